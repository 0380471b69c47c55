/*
 * fc3.h -- C ABI of the MI355X (gfx950) third-order force-constant library.
 *
 * The drop-in boundary for ForceConstants.jl's third-order pipeline.  Every entry
 * point takes plain pointers and sizes; no HIP or torch types cross it.  A Julia
 * `ccall` layer (or Python ctypes) binds these symbols and keeps the reference's
 * signatures second_order / third_order / mass_weight_third_order! / mcc3.
 *
 * Conventions
 *  - Tensors are column-major (Julia layout), 0-based: Psi[i + n*j + n*n*k], n = 3N.
 *  - Positions are host arrays of 3*N doubles, atom-major (x0,y0,z0,x1,...), Angstrom.
 *  - Output tensors are device pointers (from fc3_device_alloc or any HIP allocation
 *    on the handle's device) unless the name says host.
 *  - Every call returns an fc3_status; on failure fc3_last_error() holds a message
 *    (thread-local).  Assertion failures of the reference return the reference's
 *    message text so the binding can raise the same AssertionError.
 *  - A handle owns one HIP stream and a workspace; calls on one handle are
 *    serialised on its stream and a handle must not be used from two threads at
 *    once.  Distinct handles are independent.
 *  - Entry points that produce device results return once the work is enqueued;
 *    fc3_synchronize(h) waits for it.  Entry points with host outputs return after
 *    the data has arrived.
 */
#ifndef FC3_H
#define FC3_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FC3_OK = 0,
    FC3_ERR_ARGUMENT = 1,      /* shape / null pointer / size mismatch (AssertionError) */
    FC3_ERR_CUTOFF = 2,        /* r_cut >= box (third_order_analytical.jl:7) */
    FC3_ERR_BLOCK_SIZE = 3,    /* n % block_size != 0 or block_size <= 0 (mcc.jl:91-92) */
    FC3_ERR_OUT_OF_MEMORY = 4, /* device allocation failed */
    FC3_ERR_HIP = 5,           /* HIP runtime error */
    FC3_ERR_UNSUPPORTED = 6,   /* option not supported (e.g. D != 3) */
    FC3_ERR_CAPACITY = 7       /* neighbour list capacity exceeded */
} fc3_status;

typedef enum {
    FC3_PREC_FP32 = 0,    /* plain FP32 on FP32 MFMA (the reference's Float32 path) */
    FC3_PREC_BF16X3 = 1   /* BF16 MFMA, hi/lo split, 3 products, FP32 accumulate;
                             mcc3 with n % 4 != 0 runs the FP32 form instead */
} fc3_precision;

typedef struct fc3_context* fc3_handle;

/* ---- library / errors ------------------------------------------------------ */
int fc3_version(void);                       /* major*10000 + minor*100 + patch */
const char* fc3_last_error(void);            /* thread-local, never NULL */
const char* fc3_status_string(int status);

/* ---- handle ---------------------------------------------------------------- */
int fc3_create(fc3_handle* out, int device);
int fc3_destroy(fc3_handle h);
int fc3_synchronize(fc3_handle h);
/* Frees the handle's workspace (grown on demand by the calls that need scratch and
 * otherwise kept until fc3_destroy) after the stream has finished.  Not a reference
 * interface: returns device memory between large calls.                          */
int fc3_release_workspace(fc3_handle h);
int fc3_get_device(fc3_handle h, int* device);

/* ---- device memory (owned by the caller) ----------------------------------- */
int fc3_device_alloc(fc3_handle h, void** dptr, size_t bytes);
int fc3_device_free(fc3_handle h, void* dptr);
int fc3_memcpy_h2d(fc3_handle h, void* dst, const void* src, size_t bytes);  /* synchronous */
int fc3_memcpy_d2h(fc3_handle h, void* dst, const void* src, size_t bytes);  /* synchronous */
int fc3_memset_zero(fc3_handle h, void* dst, size_t bytes);                 /* async */
int fc3_device_memory(fc3_handle h, size_t* free_bytes, size_t* total_bytes);

/* ---- timing on the handle's stream (for measurements) ---------------------- */
int fc3_timer_start(fc3_handle h);
int fc3_timer_stop(fc3_handle h, float* ms);   /* waits for the stream */

/* ---- per-kernel timing ------------------------------------------------------ */
/* Kernel families timed when profiling is enabled on a handle. */
typedef enum {
    FC3_KERNEL_IFC3_WRITE = 0,   /* dense IFC3 writer (ASR + tol + mass weight fused) */
    FC3_KERNEL_GEMM_MODE3 = 1,   /* mcc3 mode-3 GEMM (M = n^2) */
    FC3_KERNEL_GEMM_MODE12 = 2,  /* mcc3 per-slab mode-1 GEMMs (and mode 2 on the FP32 kernels) */
    FC3_KERNEL_PAIR_BLOCKS = 3,  /* pair / triplet block evaluation */
    FC3_KERNEL_NEIGHBORS = 4,    /* neighbour list construction */
    FC3_KERNEL_GEMM_MODE2 = 5,   /* mcc3 per-slab mode-2 GEMMs on the BF16x3 kernels */
    FC3_KERNEL_COUNT = 6
} fc3_kernel_kind;
/* Enable (1) or disable (0) event timing of each kernel launch on this handle;
 * enabling clears previous records.  Not a reference interface: measurement aid. */
int fc3_profile_enable(fc3_handle h, int enable);
/* Total device time (ms) and launch count of one kernel family since enabling;
 * waits for the handle's stream. */
int fc3_profile_query(fc3_handle h, int kind, double* total_ms, int* launches);

/* ---- second order ----------------------------------------------------------- */
/* second_order(sys, LJ(sigma, epsilon, pot_r_cut), AnalyticalCalculator(tol, r_cut))
 * replaces src/ifc.jl:3-18 -> src/analytical/second_order_analytical.jl:3-42,
 * src/asr.jl:3-17, src/helper_funcs.jl:20-28.  ifc2: device, n x n FP64.        */
int fc3_second_order_lj(fc3_handle h, const double* positions, int n_atoms,
                        const double box[3], double sigma, double epsilon,
                        double pot_r_cut, double r_cut, double tol, double* ifc2);

/* second_order(sys, StillingerWeberSilicon(), AutoDiffCalculator(tol, r_cut))
 * replaces src/ifc.jl:3-18 -> src/autodiff/second_order_AD.jl:39-111.
 * sw_params: 11 doubles {epsilon, sigma, a, lambda, gamma, cos_theta0, A, B, p, q,
 *            cos_theta0_generated}: the SW_Params of SW.jl:3-14 plus the cos(theta0)
 *            that the reference's generated derivative blocks bake in (-0.333333333333,
 *            generated_derivatives/SW/H3_exec_ijk.jl:39); the runtime-built iik block
 *            uses cos_theta0.  Pass cos_theta0 twice for a self-consistent potential. */
int fc3_second_order_sw(fc3_handle h, const double* positions, int n_atoms,
                        const double box[3], const double* sw_params, double r_cut,
                        double tol, double* ifc2);

/* dynamical_matrix(sys, LJ, AnalyticalCalculator) then get_modes(D, num_rigid) on the
 * device (src/dynmat.jl:5-33, 123-131): D = IFC2 / sqrt(m_I m_J), eigen(Hermitian(D))
 * by rocSOLVER dsyevd in FP64 (upper triangle, eigenvalues ascending), then the
 * num_rigid eigenvalues of smallest magnitude set to zero (vectors untouched).
 *   masses: host, n_atoms;  freqs_sq: host, n (the eigenvalues, zeroed as above);
 *   phi64 / phi32: device n x n eigenvectors (columns = modes) in FP64 / FP32, or NULL.
 * Eigenvectors of a degenerate eigenvalue span the same space as the reference's but
 * may be another basis of it (as between any two LAPACK builds).                   */
int fc3_modes_lj(fc3_handle h, const double* positions, int n_atoms, const double box[3],
                 double sigma, double epsilon, double pot_r_cut, double r_cut, double tol,
                 const double* masses, int num_rigid_translation, double* freqs_sq,
                 double* phi64, float* phi32);

/* The same for StillingerWeberSilicon with the AutoDiffCalculator (IFC2 as
 * fc3_second_order_sw, sw_params as there).                                        */
int fc3_modes_sw(fc3_handle h, const double* positions, int n_atoms, const double box[3],
                 const double* sw_params, double r_cut, double tol, const double* masses,
                 int num_rigid_translation, double* freqs_sq, double* phi64, float* phi32);

/* ---- third order ------------------------------------------------------------ */
/* third_order(sys, LJ, AnalyticalCalculator) replaces src/ifc.jl:20-35 ->
 * src/analytical/third_order_analytical.jl:4-69, src/asr.jl:19-40,
 * src/helper_funcs.jl:20-28, fused with mass_weight_third_order!
 * (third_order_analytical.jl:185-214) and the Float32 conversion the reference's
 * callers do before mcc3 (README.md:69-74).
 *   psi64: device n^3 FP64, the third_order result (ASR + tolerance), or NULL.
 *   psi32: device n^3 FP32, the same values divided by sqrt(m_i m_j m_k) when
 *          masses != NULL (unweighted otherwise), or NULL.
 *   masses: host, n_atoms doubles, or NULL.                                      */
int fc3_third_order_lj(fc3_handle h, const double* positions, int n_atoms,
                       const double box[3], double sigma, double epsilon,
                       double pot_r_cut, double r_cut, double tol,
                       const double* masses, double* psi64, float* psi32);

/* Allocation-free variant of fc3_third_order_lj: no counterpart in the reference.  For
 * callers that cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, checks and results (bit for bit) as fc3_third_order_lj,
 * but the scratch (positions, neighbour lists, pair and self blocks) lives in a static
 * 64 MiB device array per GPU instead of the handle's workspace, and the call makes no
 * device allocation.  Systems whose scratch exceeds it (296 bytes per neighbour slot,
 * N * cap slots for N atoms with at most cap neighbours each, plus 252 bytes per atom)
 * return FC3_ERR_CAPACITY before any launch that writes the outputs.
 * All calls on one GPU share that array: calls that overlap on one GPU race.  It waits
 * for the neighbour count on the host, as fc3_third_order_lj does.                 */
int fc3_third_order_lj_noalloc(fc3_handle h, const double* positions, int n_atoms,
                               const double box[3], double sigma, double epsilon,
                               double pot_r_cut, double r_cut, double tol,
                               const double* masses, double* psi64, float* psi32);

/* Fused third_order + mass weighting + mcc3 for LJ without the dense tensor: the
 * K3 = mcc3(Psi32, phi) that fc3_third_order_lj (psi32, masses) followed by fc3_mcc3
 * (FC3_PREC_BF16X3) gives, with the mode-1 product taken over Psi's nonzeros straight
 * from the pair blocks (FP32; the values are exactly those psi32 would hold), mode 2
 * over each atom's neighbourhood columns (the only nonzero ones) and mode 3 on the
 * BF16x3 kernels.  No counterpart in the reference; it follows the sparse mode-1
 * intent of src/todo/old_mcc.jl:17-65.  Neither the FP64 nor the FP32 n^3 Psi is
 * formed; the workspace holds 12 n^2 Kc bytes of mode-1 result, Kc = 3 (max
 * neighbours + 1) rounded up to 16.
 *   phi: device n x n FP32;  k3: device n^3 FP32, 16-byte aligned;  masses: host or NULL.
 * Needs 3N % 4 == 0 and 3N >= 16 (FC3_ERR_UNSUPPORTED otherwise).                  */
int fc3_third_order_mcc3_lj(fc3_handle h, const double* positions, int n_atoms,
                            const double box[3], double sigma, double epsilon,
                            double pot_r_cut, double r_cut, double tol, const double* masses,
                            const float* phi, float* k3);

/* third_order(sys, StillingerWeberSilicon(), AutoDiffCalculator) replaces
 * src/ifc.jl:20-35 -> src/autodiff/third_order_AD.jl:71-224 (+ ASR!, apply_tols!),
 * with the same fused outputs as fc3_third_order_lj.                            */
int fc3_third_order_sw(fc3_handle h, const double* positions, int n_atoms,
                       const double box[3], const double* sw_params, double r_cut,
                       double tol, const double* masses, double* psi64, float* psi32);

/* Fused third_order + mass weighting + mcc3 for SW silicon without the dense tensor: the
 * K3 = mcc3(Psi32, phi) that fc3_third_order_sw (psi32, masses) followed by fc3_mcc3
 * (FC3_PREC_BF16X3) gives, with the mode-1 product taken over the nonzero 3x3x3 blocks of
 * the assembly's reduced block table (FP32; the values are exactly those psi32 would
 * hold), mode 2 over each atom's column atoms (those sharing a triplet with it) and
 * mode 3 on the BF16x3 kernels.  No counterpart in the reference; it follows the sparse
 * mode-1 intent of src/todo/old_mcc.jl:17-65, as fc3_third_order_mcc3_lj does for LJ.
 * Neither n^3 Psi is formed.  Arguments as fc3_third_order_sw, with phi (device n x n
 * FP32) and k3 (device n^3 FP32, 16-byte aligned) as fc3_third_order_mcc3_lj.
 * Needs 3N % 4 == 0 and 3N >= 16 (FC3_ERR_UNSUPPORTED otherwise).                  */
int fc3_third_order_mcc3_sw(fc3_handle h, const double* positions, int n_atoms,
                            const double box[3], const double* sw_params, double r_cut,
                            double tol, const double* masses, const float* phi, float* k3);

/* mass_weight_third_order!(Psi, masses) replaces
 * src/analytical/third_order_analytical.jl:185-214: psi64 (device, n^3, n = dim *
 * n_atoms) is divided in place by sqrt(m_i m_j m_k) where nonzero.              */
int fc3_mass_weight_third_order(fc3_handle h, double* psi64, int n_atoms, int dim,
                                const double* masses);

/* ---- checks of finished force constants (check_asr / check_symmetry) ----------- */
/* asr_satisfied(ifc, N_atoms, D, tol) replaces src/asr.jl:43-74 (called by
 * second_order and third_order with check_asr = true, src/ifc.jl:14,33).  ifc: device
 * FP64, n^order, n = dim * n_atoms, column-major.  For every slice (order 3: the third
 * index; order 2: the matrix), row D(i-1)+a and component b, s = sum(ifc[row, b:D:end])
 * summed in FP64 in ascending column order; the rule fails where s > tol (signed, as in
 * the reference).  *n_violations = the number of failing (slice, row, b); the first in
 * the reference's loop order (slice, atom, a, b) is returned as the key
 * (slice * n + row) * dim + b (0-based; -1 if none) with its sum.  The tensor is read
 * once.  first_violation and first_sum may be NULL.                                 */
int fc3_asr_satisfied(fc3_handle h, const double* ifc, int order, int n_atoms, int dim,
                      double tol, long long* n_violations, long long* first_violation,
                      double* first_sum);

/* issymmetric(Phi) of second_order(...; check_symmetry = true), src/ifc.jl:15: the
 * number of pairs r < c with Phi[r, c] != Phi[c, r] (exact comparison; NaN counts),
 * Phi a device FP64 n x n matrix.                                                   */
int fc3_issymmetric2(fc3_handle h, const double* ifc2, int n, long long* n_asymmetric);

/* ---- text output (save_second_order / save_third_order, fmt = :txt) ----------- */
/* save_third_order(Psi, N_atoms, D, outpath; fmt = :txt) replaces src/io.jl:39-62: the
 * device tensor psi (FP64, n^3, n = dim * n_atoms) written to the file `path` as the
 * header "#N atoms, Sum Ψ = s, Units: units" (s = round(sum(Psi), sigdigits = 7)) and
 * one line "i a j b k c value" (1-based) per nonzero, atoms i, j, k outermost and
 * components a, b, c innermost, values printed as Julia prints a Float64 (shortest
 * decimal that reads back exactly).  The nonzeros are found and ordered on the
 * device; only they cross to the host.  units: UTF-8 text, or NULL for none.      */
int fc3_save_third_order_txt(fc3_handle h, const double* psi, int n_atoms, int dim,
                             const char* units, const char* path);
/* save_second_order(Phi, N_atoms, D, outpath; fmt = :txt) replaces src/io.jl:9-28, the
 * same for the device matrix ifc2 (FP64, n^2): "#N atoms, Sum Φ = s, Units: units"
 * and lines "i a j b value".                                                         */
int fc3_save_second_order_txt(fc3_handle h, const double* ifc2, int n_atoms, int dim,
                              const char* units, const char* path);
/* x as the text files print it (Julia's print of a Float64), NUL-terminated into buf
 * (len >= 33 always suffices).  Host only.                                          */
int fc3_format_float(double x, char* buf, int len);

/* ---- modal coupling constants ----------------------------------------------- */
/* mcc3(Psi::CuArray{Float32,3}, phi::CuArray{Float32,2}) replaces src/mcc.jl:29-51
 * (and mcc3_tens_opt, src/mcc.jl:11-20):
 *   k3[n,m,l] = sum_ijk psi[i,j,k] phi[i,n] phi[j,m] phi[k,l]
 * psi, phi, k3: device FP32; k3 must not alias psi.                              */
int fc3_mcc3(fc3_handle h, const float* psi, const float* phi, int n, int precision,
             float* k3);

/* mcc3!(Psi, phi) replaces src/mcc.jl:60-80: the result overwrites psi (device),
 * using a bounded workspace (at most about 4 GiB, O(n^2) slabs plus O(n) rows of
 * Psi) instead of the reference's two extra n^3 tensors.                         */
int fc3_mcc3_inplace(fc3_handle h, float* psi, const float* phi, int n, int precision);

/* Symmetric-unique K3: the part the TODO at src/mcc.jl:116-117 asks for ("can only
 * calculate unique part and rotate to enforce symmetry").  k3u (device, n(n+1)(n+2)/6
 * floats) receives K3[a,b,l] for a <= b <= l at k3u[l(l+1)(l+2)/6 + b(b+1)/2 + a] (slab
 * by slab, the upper triangle column by column), each entry bit for bit the one fc3_mcc3
 * writes at (a,b,l).  With BF16X3 (n % 4 == 0) only the GEMM tiles that hold such
 * entries are computed in modes 1 and 2 (about 1/2 and 1/6 of their work); other
 * shapes compute the full K3 and pack it.  For a symmetric Psi, K3 is symmetric and
 * these entries determine it.  The handle's workspace grows by n^3 floats (the full
 * pipeline's intermediate lives there; fc3_release_workspace returns it).  The tile
 * lists are uploaded first, which waits for work already queued on the stream.     */
int fc3_mcc3_unique(fc3_handle h, const float* psi, const float* phi, int n, int precision,
                    float* k3u);

/* mcc3(Psi, phi, block_size) replaces src/mcc.jl:89-120: same result as fc3_mcc3,
 * delivered into a HOST array k3_host (n^3 floats) in slabs of block_size, with the
 * reference's assertions (n % block_size == 0, block_size > 0).                 */
int fc3_mcc3_blocked(fc3_handle h, const float* psi, const float* phi, int n,
                     int block_size, int precision, float* k3_host);

/* ---- batched snapshots (no counterpart in the reference, which handles one
 *      system per call: a loop of third_order (src/ifc.jl:20-35) and mcc3
 *      (src/mcc.jl:29-51) over displaced snapshots) ---------------------------- */
typedef struct {
    const double* positions;   /* host, 3 * n_atoms doubles, Angstrom */
    int n_atoms;
    double box[3];             /* cubic/orthorhombic box, Angstrom */
    const double* masses;      /* host, n_atoms (mass-weighted Psi), or NULL */
    float* phi;                /* device, n x n FP32 modes (n = 3 n_atoms): input, or
                                  output with FC3_BATCH_DEVICE_MODES */
    float* k3;                 /* device, n^3 FP32 output */
    double* freqs_sq;          /* host, n: with FC3_BATCH_DEVICE_MODES the eigenvalues of
                                  get_modes (3 rigid modes zeroed), or NULL */
} fc3_snapshot;

enum {
    FC3_BATCH_AUTOTUNE = 1,    /* time the configurations of sizes not seen before */
    FC3_BATCH_NO_GRAPH = 2,    /* issue the kernels directly instead of one graph launch */
    FC3_BATCH_DEVICE_MODES = 4 /* phi = get_modes(dynamical_matrix(sys, LJ, calc)) of every
                                  snapshot (src/dynmat.jl:5-33,123-131, num_rigid_translation
                                  = 3) on the device before the contraction: IFC2, mass
                                  weighting and rocSOLVER dsyevd, batched over same-size
                                  snapshots; needs the masses.  Synchronises the stream. */
};

/* For every snapshot: Psi = third_order(LJ, AnalyticalCalculator) mass weighted to FP32
 * (as fc3_third_order_lj's psi32), then k3 = mcc3(Psi, phi) at `precision`: either
 * through the dense Psi (one shared buffer of the largest size) or, for BF16X3 where
 * it applies and the tuner picks it, by the fused path of fc3_third_order_mcc3_lj
 * (no Psi).  All snapshots run from one captured graph (cached for an unchanged plan:
 * same sizes, pointers and configurations).  Returns once the work is enqueued.    */
int fc3_batch_third_order_mcc3_lj(fc3_handle h, const fc3_snapshot* snaps, int count,
                                  double sigma, double epsilon, double pot_r_cut,
                                  double r_cut, double tol, int precision, int flags);
/* Nodes of the last captured batch graph. */
int fc3_batch_graph_nodes(fc3_handle h, size_t* nodes);
/* Lanes (parallel branches) of the last captured batch graph: snapshots whose K3
 * outputs overlap share a lane and run in submission order; lanes run concurrently,
 * each with its own Psi and mcc3 scratch, as many as free device memory allows. */
int fc3_batch_graph_lanes(fc3_handle h, int* lanes);
/* Tuning table: the chosen configuration of size n (variant 2 = dense Psi + FP32
 * kernels, 3 = dense Psi + pre-split BF16x3 kernels, 4 = the fused path; tiles =
 * workgroups per GEMM launch) and the time of one snapshot (assembly included);
 * save/load as text lines "n precision variant tiles ms".                         */
int fc3_autotune_query(fc3_handle h, int n, int precision, int* variant, long* tiles,
                       float* ms);
int fc3_autotune_save(fc3_handle h, const char* path);
int fc3_autotune_load(fc3_handle h, const char* path);

/* ---- adjoints (no counterpart in the reference; its rrule targets) ----------- */
/* Reverse mode of mcc3 (src/mcc.jl:15-17) for a cotangent dk3 = dL/dK3 (device n^3):
 *   dpsi = dL/dPsi = dk3 x1 phi x2 phi x3 phi            (device n^3, or NULL)
 *   dphi = dL/dphi, the three mode terms                  (device n x n, or NULL; needs psi)
 * psi, phi, dk3 device FP32.  With dphi the work is 20 n^4 flops (18 without dpsi) in
 * two n^3 FP32 buffers of workspace: for FC3_PREC_BF16X3 with n % 4 == 0, n >= 16 and
 * 16-byte aligned tensors the seven mode products and the three n x n^2 x n
 * contractions run on the BF16x3 kernels, with about two more n^3 of split scratch;
 * otherwise they are FP32 GEMMs.  dpsi alone runs the mcc3 kernels
 * at `precision`.                                                                    */
int fc3_mcc3_backward(fc3_handle h, const float* psi, const float* phi, const float* dk3,
                      int n, int precision, float* dpsi, float* dphi);

/* The scalar of the differentiation workload (no counterpart in the reference, which has
 * no adjoints): L = sum_e w_e k3[e]^2 over `count` device FP32 values, with weights
 * w_e ~ U[0, 1) from Philox4x32-10 keyed by `seed` (element e: word e % 4 of the block
 * for counter (e / 4 low, e / 4 high, 0, 0), key (seed low, seed high);
 * w_e = (word >> 8) 2^-24), so that any host can regenerate them.
 *   loss: L on the host (FP64 terms, a fixed summation order), or NULL;
 *   dk3:  the cotangent dL/dk3 = 2 w k3 (device FP32; may be k3 itself), or NULL.
 * k3 and dk3 are 16-byte aligned.                                                     */
int fc3_weighted_square_loss(fc3_handle h, const float* k3, long long count,
                             unsigned long long seed, float* dk3, double* loss);

/* Reverse mode of third_order(sys, LJ(sigma, epsilon, pot_r_cut), AnalyticalCalculator)
 * (src/ifc.jl:20-35 -> third_order_analytical.jl:4-69, asr.jl:19-40,
 * helper_funcs.jl:20-28) with respect to the potential parameters:
 *   grad[0] = dL/dsigma, grad[1] = dL/depsilon   (host)
 * for the cotangent of exactly one output of fc3_third_order_lj:
 *   dpsi32: of the FP32 psi32 (mass weighted when masses != NULL), device n^3, or
 *   dpsi64: of the FP64 psi64 (masses must be NULL), device n^3.
 * Pair blocks are recomputed from the positions; entries zeroed by the tolerance
 * contribute nothing.                                                              */
int fc3_third_order_lj_param_backward(fc3_handle h, const double* positions, int n_atoms,
                                      const double box[3], double sigma, double epsilon,
                                      double pot_r_cut, double r_cut, double tol,
                                      const double* masses, const float* dpsi32,
                                      const double* dpsi64, double grad[2]);

/* Reverse mode of third_order(sys, StillingerWeberSilicon(), AutoDiffCalculator)
 * (src/ifc.jl:20-35 -> src/autodiff/third_order_AD.jl:71-224, asr.jl:19-40,
 * helper_funcs.jl:20-28) with respect to the SW parameters (SW.jl:3-14):
 *   grad = {dL/dA, dL/dB, dL/dlambda, dL/dgamma}   (host)
 * for the cotangent dpsi32 (of the FP32 psi32 of fc3_third_order_sw, mass weighted
 * when masses != NULL) or dpsi64 (of the FP64 psi64, masses NULL).  sw_params as in
 * fc3_third_order_sw.  Record blocks are recomputed; when tol > 0 the forward FP64
 * tensor is assembled into the workspace to apply the tolerance mask (8 n^3 bytes).  */
int fc3_third_order_sw_param_backward(fc3_handle h, const double* positions, int n_atoms,
                                      const double box[3], const double* sw_params,
                                      double r_cut, double tol, const double* masses,
                                      const float* dpsi32, const double* dpsi64,
                                      double grad[4]);

#ifdef __cplusplus
}
#endif

#endif /* FC3_H */
