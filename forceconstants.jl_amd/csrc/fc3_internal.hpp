// Internal declarations shared by the C-ABI layer and the HIP kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fc3.h"

namespace fc3 {

// Thread-local error message; set_error returns the status for `return set_error(...)`.
int set_error(int status, const std::string& msg);
const char* last_error();

#define FC3_HIP(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::fc3::set_error(e_ == hipErrorOutOfMemory ? FC3_ERR_OUT_OF_MEMORY  \
                                                              : FC3_ERR_HIP,           \
                                    std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define FC3_TRY(expr)                    \
    do {                                 \
        int s_ = (expr);                 \
        if (s_ != FC3_OK) return s_;     \
    } while (0)

// Per-handle kernel timing: when enabled, launch sites bracket their kernels with
// HIP events on the launch stream (ProfScope) and fc3_profile_query sums them.
struct ProfRecord {
    int kind;
    hipEvent_t start, stop;
};
struct Profiler {
    bool enabled = false;
    std::vector<ProfRecord> records;
    std::vector<hipEvent_t> pool;
};
}  // namespace fc3

struct fc3_context {
    fc3::Profiler prof;
    void* blas = nullptr;   // rocBLAS handle, created on first use (adjoint.hip)
    void* batch = nullptr;  // batched-snapshot state: workspace, graph, tuning (batch.cpp)
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    // Workspace: reserved (size computed per call), grown with a stream sync.
    void* ws = nullptr;
    size_t ws_bytes = 0;
    // Copy stream and its events (created on first use by fc3_mcc3_blocked): device ->
    // host copies of one block overlap the computation of the next.
    hipStream_t aux = nullptr;
    hipEvent_t aux_ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace fc3 {

// The handle whose API call is running on this thread (set by the C-ABI layer).
void set_current(fc3_context* h);
// A launch site of one kernel kind: a roctx range named for the kind around the enqueue
// (host side, seen by rocprofv3 --marker-trace), and, when the handle's profiling is on,
// HIP events around its kernels on the launch stream.
struct ProfScope {
    fc3_context* h = nullptr;
    hipStream_t s = nullptr;
    int rec = -1;
    ProfScope(int kind, hipStream_t stream);
    ~ProfScope();
};
// A roctx range around one C-ABI call (its enqueue on the host), named for the function.
struct TraceRange {
    explicit TraceRange(const char* name);
    ~TraceRange();
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};
#define FC3_TRACE_RANGE() ::fc3::TraceRange fc3_trace_range_(__func__)

// Allocation-free entry points (noalloc.hip): size of the static device scratch, and the
// once-per-device lookup of its address (fc3_create calls it with the device current).
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
int noalloc_init_device(int device);

// Ensure the handle's workspace holds at least `bytes`.
int ensure_workspace(fc3_context* h, size_t bytes);

// Bump allocator over a workspace region.
struct Carve {
    char* p;
    size_t off = 0, cap;
    Carve(void* base, size_t c) : p(static_cast<char*>(base)), cap(c) {}
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T* r = reinterpret_cast<T*>(p + off);
        off += count * sizeof(T);
        return r;
    }
    bool ok() const { return off <= cap; }
    static size_t need(size_t bytes) { return ((bytes + 255) & ~size_t(255)); }
};

// ---- neighbour lists (neighbors.hip) ---------------------------------------------
// Full neighbour lists of every atom, sorted by atom index, minimum-image rule of
// nearest_mirror! with the reference's strict cutoff test dist^2 < r_cut^2.
struct NeighborList {
    int n_atoms = 0;
    int cap = 0;          // slots per atom
    int* count = nullptr; // [N]
    int* nbr = nullptr;   // [N*cap]
    int* rev = nullptr;   // [N*cap]  slot of atom a in the list of nbr[a][s]
};

// Counts neighbours and returns the maximum count (synchronises the stream).
int neighbor_count(hipStream_t s, const double* d_pos, int N, const double box[3],
                   double r_cut, int* d_count, int* d_maxcount, int* h_max);
int neighbor_fill(hipStream_t s, const double* d_pos, int N, const double box[3],
                  double r_cut, NeighborList& nl);

// ---- LJ kernels (lj.hip) ---------------------------------------------------------
int lj_pair_blocks(hipStream_t s, const double* d_pos, const double box[3], double sigma,
                   double epsilon, const NeighborList& nl, double* E3, double* S3,
                   double* E2, double* S2);
int lj_write_ifc2(hipStream_t s, const NeighborList& nl, const double* E2, const double* S2,
                  double tol, double* ifc2);
// sigma^12 and sigma^6 as the reference's Float64^Int evaluates them.
void lj_sigma_powers(double sigma, double* s12, double* s6);

// Generic dense third-order writer over per-pair blocks (shared by LJ and SW):
// see the kernel comment in ifc3_writer.hip.
struct Ifc3Blocks {
    const double* E = nullptr;   // pair/direct blocks per (atom, slot), 27 doubles
    const double* S = nullptr;   // self blocks per atom, 27 doubles
};
int write_ifc3_pairwise(hipStream_t s, const NeighborList& nl, const Ifc3Blocks& b,
                        double tol, const double* d_masses, double* psi64, float* psi32);

// Psi[ii, jj, kk] /= sqrt((m[ii / D] m[jj / D]) m[kk / D]) where nonzero, n = D N.
int mass_weight3(hipStream_t s, double* psi, int N, int D, const double* d_masses);

// Sparse mode-1 product fused to the pair assembly (no dense Psi): for every atom K
// and its three slabs kk = 3K + g, U_K(g n + p, t) = sum_i phiT[p + n i] Psi32[i, jj_t, kk]
// over the columns jj_t = 3 J_t + b (t = 3 s + b; J_0 = K, J_s = nbr[K][s - 1]), the
// only columns of those slabs that hold nonzeros; Psi32 exactly the FP32 writer's
// values.  Split (bf16 hi/lo), K-blocked over t with Kc = mcc3_gather_k(cap) (zero
// past the count): hi at U[K 2 (3n) Kc + (t / 16) 16 (3n) + 16 (g n + p) + t % 16], lo at
// + 3n Kc.  Dynamic LDS of lj_mode1_sparse_lds(cap) bytes when that is at most 64 KB;
// else the coefficients go to global memory first, lj_mode1_global_bytes(N, cap) of
// `scratch` (free again on return; FC3_ERR_UNSUPPORTED without it).
size_t lj_mode1_sparse_lds(int cap);
size_t lj_mode1_global_bytes(int N, int cap);
int lj_mode1_gathered(hipStream_t s, const NeighborList& nl, const Ifc3Blocks& b, double tol,
                      const double* d_masses, const float* phiT, int Kc, __bf16* U,
                      void* scratch, size_t scratch_bytes);

// ---- Stillinger-Weber (sw.hip) ------------------------------------------------------
// Workspace for one SW assembly with neighbour capacity `cap` (upper bound).
size_t sw_workspace_bytes(int N, int cap, bool third);
// The reduced IFC3 blocks of one SW assembly: U blocks sorted by key (C N + B) N + A,
// block p's values uval[27 p + x1 + 3 x2 + 9 x3] after ASR! (before the tolerance).
struct SwBlockTable {
    const unsigned long long* ukey = nullptr;
    const double* uval = nullptr;
    int U = 0;
};
// IFC3 (psi64/psi32, either may be null) when third, else IFC2 into ifc2.  params: the
// 10 SW parameters; c0_fd / c0_iik: cos(theta0) of the generated / runtime-built blocks.
// With table non-null: the IFC3 blocks only, left in ws and described by *table (no
// dense tensor is written).
int sw_assemble(hipStream_t s, const double* d_pos, const double box[3], const double* params,
                double c0_fd, double c0_iik, const NeighborList& nl, void* ws, size_t ws_bytes,
                double tol, const double* d_masses, double* psi64, float* psi32, double* ifc2,
                SwBlockTable* table = nullptr);
// dL/d(A, B, lambda, gamma) into d_out[0..3] (device) for the cotangent dpsi32 (of the
// mass-weighted FP32 Psi, d_masses as in the forward call) or dpsi64 (of the FP64 Psi);
// fwd: the forward assembly's block table, read for the tolerance mask when tol > 0.
size_t sw_param_grad_bytes(int N, int cap);
int sw_param_grad(hipStream_t s, const double* d_pos, const double box[3], const double* params,
                  double c0_fd, double c0_iik, const NeighborList& nl, double tol,
                  const SwBlockTable* fwd, const double* d_masses, const float* dpsi32,
                  const double* dpsi64, void* ws, double* d_out);

// ---- fused SW third_order + mcc3 (sw_fused.hip) --------------------------------------
// Columns of each slab atom K from an SW block table, in the LJ fused path's convention:
// slot 0 = K, slots 1 .. colcnt[K] = the other atoms J with a block (., J, K), ascending, at
// collist[K cap2 + s - 1]; seg[K 2 (cap2 + 1) + 2 s .. + 1] = slot s's block range.
struct SwColumns {
    int* colcnt = nullptr;
    int* collist = nullptr;
    int* seg = nullptr;
    int cap2 = 0;     // most column atoms J != K of any slab
    int maxblk = 0;   // most blocks of any slab
};
// Workspace of sw_fused_columns for at most cap2_ub column atoms per slab.
size_t sw_fused_list_bytes(int N, int cap2_ub);
// Builds the columns (one host read-back of the maxima); FC3_ERR_UNSUPPORTED past cap2_ub.
int sw_fused_columns(hipStream_t s, const SwBlockTable& T, int N, int cap2_ub, void* ws,
                     SwColumns& out);
// The mode-1 product over the table's nonzeros (the FP32 writer's values: tolerance, mass
// weighting), in mcc3_from_gathered's layout with Kc = mcc3_gather_k(cols.cap2).  A slab
// whose coefficients exceed 64 KB of LDS (sw_mode1_lds(cols.maxblk)) takes them from
// global scratch instead: sw_mode1_global_bytes(T.U) of `scratch`, free again on return.
size_t sw_mode1_lds(int maxblk);
size_t sw_mode1_global_bytes(int U);
int sw_mode1_gathered(hipStream_t s, const SwBlockTable& T, int N, const SwColumns& cols,
                      double tol, const double* d_masses, const float* phiT, int Kc, __bf16* U,
                      void* scratch, size_t scratch_bytes);

// ---- mcc3 (mcc3.hip) --------------------------------------------------------------
// Kernel family and launch sizing of one mcc3: variant 2 = FP32-operand kernels (any n),
// 3 = pre-split BF16x3 kernels (BF16X3 with n % 4 == 0, n >= 16), 0 = default (3 when it
// applies); tiles = target workgroups per GEMM launch (0 = default).
struct Mcc3Config {
    int variant = 0;
    long tiles = 0;
};
bool mcc3_variant_applies(int n, int precision, int variant);
size_t mcc3_workspace_bytes(int n, int precision, bool inplace, const Mcc3Config* cfg = nullptr);
int mcc3_run(hipStream_t s, const float* psi, const float* phi, int n, int precision,
             float* k3, void* ws, size_t ws_bytes, const Mcc3Config* cfg = nullptr);
// Symmetric-unique K3 (fc3_mcc3_unique): the n(n+1)(n+2)/6 entries K3[a, b, l], a <= b <= l,
// at k3u[l(l+1)(l+2)/6 + b(b+1)/2 + a], each bitwise the full mcc3's entry; k3s is n^3
// floats of device scratch (the full pipeline's U and slab outputs live there).
long mcc3_unique_count(int n);
size_t mcc3_unique_workspace_bytes(int n, int precision);
int mcc3_unique_run(hipStream_t s, const float* psi, const float* phi, int n, int precision,
                    float* k3s, float* k3u, void* ws, size_t ws_bytes);
int mcc3_run_inplace(hipStream_t s, float* psi, const float* phi, int n, int precision,
                     void* ws, size_t ws_bytes);
// Mode products of the mcc3 adjoint on the BF16X3 kernels, out of place, FP32 in and out.
// M is n x n column-major, M(k, c) = M[k + n c]; X, Y are n^3 (Y must not alias X).
//   mode3:     Y[i + n j + n^2 c] = sum_k X[i + n j + n^2 k] M(k, c)
//   modes31_t: Y[j + n b + n^2 c] = sum_{i,k} X[i + n j + n^2 k] M(i, b) M(k, c)
//              (the (b, j) slab is stored transposed); x_swapped: X holds that tensor
//              with its first two indices swapped, X[j + n i + n^2 k]
//   modes12:   Y[b + n c + n^2 l] = sum_{i,j} X[i + n j + n^2 l] M(i, b) M(j, c)
// They apply when mcc3_adjoint_v3_applies (n % 4 == 0, n >= 16, 16-byte aligned tensors).
bool mcc3_adjoint_v3_applies(int n, const void* a, const void* b, const void* c);
size_t mcc3_adjoint_v3_bytes(int n);
int mcc3_mode3(hipStream_t s, const float* X, const float* M, int n, float* Y, void* ws,
               size_t ws_bytes);
int mcc3_modes31_t(hipStream_t s, const float* X, const float* M, int n, float* Y, void* ws,
                   size_t ws_bytes, bool x_swapped = false);
int mcc3_modes12(hipStream_t s, const float* X, const float* M, int n, float* Y, void* ws,
                 size_t ws_bytes);
// The adjoint's n x n^2 x n contractions on the BF16X3 kernel with split K: partials
// P_z[a + n b] = sum over chunk z of e < n^2 of A(a, e) B(e, b) at P + z n^2,
// z < *zc <= min(64, n)
// (the caller sums them).  form 0: A(a, e) = X[a + n e], B(e, b) = H[b + n e];
// form 1: A(a, e) = X[e + n^2 a], B(e, b) = H[e + n^2 b].  Applies for n % 4 == 0,
// 16 <= n (aligned tensors).
// Modes 2 and 3 of mcc3 on the BF16X3 kernels from the gathered mode-1 result U of
// lj_mode1_gathered: mode 2 as one batched product per atom over its neighbourhood's
// columns (gathered phi rows), mode 3 in place; K3 FP32 in k3 (n % 4 == 0, n >= 16).
// phiT: phi transposed (phiT[q + n i] = phi[i + n q]), for the row gather.
int mcc3_gather_k(int cap);
// direct: room for the FP32 intermediate that mode 3 reads without a split pass (n^3
// floats more); mcc3_from_gathered takes that path when ws_bytes allows it.
size_t mcc3_from_gathered_bytes(int n, int cap, bool direct);
int mcc3_from_gathered(hipStream_t s, const __bf16* U, const float* phi, const float* phiT, int n,
                       const int* count, const int* nbr, int cap, float* k3, void* ws,
                       size_t ws_bytes);
// Bounded-memory mcc3 (fc3_mcc3_blocked): K3 slabs [l0, l0 + cnt), cnt <= bs, into `out`
// (n^2 cnt floats, device) from the whole Psi, bitwise equal to the same slabs of
// mcc3_run.  v3 = mcc3_blocked_v3(n, precision, psi) selects the BF16X3 kernels (else
// FP32); the workspace (mcc3_blocked_bytes) holds phi's prepared copy and bounded scratch.
bool mcc3_blocked_v3(int n, int precision, const void* psi);
size_t mcc3_blocked_bytes(int n, bool v3, int bs);
int mcc3_blocked_slabs(hipStream_t s, const float* psi, const float* phi, int n, bool v3, int bs,
                       int l0, int cnt, float* out, void* ws, size_t ws_bytes);
// out[j + n i] = in[i + n j] (device, n x n)
int transpose_square(hipStream_t s, const float* in, int n, float* out);
// P_z = partial sums over K chunks z < *zc of the n x n x n^2 contraction
// sum_e A(a, e) B(b, e) (BF16X3, split K), A from X and B from H, with the K index e:
//   form 0: A(a, e) = X[a + n e], B(b, e) = H[b + n e]   (e over the last two indices)
//   form 1: A(a, e) = X[e + n^2 a], B likewise from H    (e over the first two)
//   form 2: A(a, e) = X[i + n a + n^2 k] for e = i + n k (e over the first and last),
//           B as form 0
bool mcc3_contract_applies(int n);
size_t mcc3_contract_bytes(int n);
int mcc3_contract(hipStream_t s, int form, const float* X, const float* H, int n, float* P,
                  int* zc, void* ws, size_t ws_bytes);

// ---- dynamical matrix and modes (modes.hip) -------------------------------------------
// A (device n x n FP64 IFC2) -> D = A / sqrt(m_I m_J) -> eigen(Hermitian(D)) in place:
// eigenvalues ascending into W (device) and h_w (host), eigenvectors as A's columns.
// E: n doubles, info: one int (device scratch).  Synchronises the stream.
// solver_ws: rocSOLVER's scratch (dsyevd_workspace_bytes), attached to the rocBLAS handle
// for this call only.
int dsyevd_workspace_bytes(fc3_context* h, int n, size_t* bytes);
int dynmat_eigen(fc3_context* h, double* A, int n, const double* d_masses, double* W, double* E,
                 int* info, double* h_w, void* solver_ws, size_t solver_ws_bytes);
int to_f32(hipStream_t s, const double* a, long cnt, float* b);
// D = A / sqrt(m_I m_J) in place (A: n x n, masses per atom on the device)
int dynmat_scale(hipStream_t s, double* A, int n, const double* d_masses);
// G same-size problems A + g n^2 at once (rocSOLVER dsyevd_strided_batched): W + g n
// eigenvalues ascending, eigenvectors in place, info[g]; no host synchronisation.
int dsyevd_batched_workspace_bytes(fc3_context* h, int n, int G, size_t* bytes);
int eigen_strided_batched(fc3_context* h, double* A, int n, int G, double* W, double* E,
                          int* info, void* solver_ws, size_t solver_ws_bytes);

// ---- adjoints (adjoint.hip) ---------------------------------------------------------
void release_blas(fc3_context* h);
// the context's rocBLAS handle (created on first use) on the context's stream
int blas_handle(fc3_context* h, void** out);
void release_batch(fc3_context* h);
size_t mcc3_backward_workspace_bytes(int n, int precision, bool need_dphi);
// L = sum_e w_e k3[e]^2 into d_loss (device) and, if dk3, the cotangent 2 w_e k3[e];
// w from Philox4x32-10 keyed by seed (adjoint.hip); k3 and dk3 16-byte aligned.
size_t weighted_square_loss_bytes();
int weighted_square_loss(hipStream_t s, const float* k3, long count, unsigned long long seed,
                         float* dk3, void* ws, double* d_loss);
// dpsi (n^3) and/or dphi (n x n), device FP32; uses h's workspace (sized by the above).
int mcc3_backward(fc3_context* h, const float* psi, const float* phi, const float* G, int n,
                  int precision, float* dpsi, float* dphi);
// dL/dsigma, dL/depsilon of <dpsi, Psi> into d_out[0..1] (device); partial: 2 doubles per
// 256 (atom, slot) pairs.  dpsi32 is the cotangent of the mass-weighted FP32 Psi
// (d_masses as in the forward call), dpsi64 that of the unweighted FP64 Psi.
int lj_param_grad(hipStream_t s, const double* d_pos, const double box[3], double sigma,
                  double epsilon, const NeighborList& nl, const double* E3, const double* S3,
                  double tol, const double* d_masses, const float* dpsi32, const double* dpsi64,
                  double* partial, double* d_out);

}  // namespace fc3
