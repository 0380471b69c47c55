// Dense third-order writer for pair potentials, fused with ASR!, apply_tols!, mass
// weighting and the FP32 conversion.
//
// What the reference produces for a pair potential (third_order_analytical.jl:4-69
// then asr.jl:19-40 and helper_funcs.jl:20-28), written column by column.  Column
// (jj, kk) = Psi[:, 3J+b, 3K+g] is nonzero only when
//   J != K and K is a neighbour of J:  row atom J  -> E_JK[a + 3b + 9g]
//                                      row atom K  -> -E_JK[b + 3g + 9a]  (the ASR! entry
//                                                     (K,J,K) = -(K,J,J), its only term)
//   J == K:                            row atom K  -> S_K[a + 3b + 9g]    (ASR! self entry)
//                                      row atom I in nbr(K) -> E_KI[b + 3g + 9a]
// where E_AB[u + 3v + 9w] is the value for two components (u, v) of atom A and one
// component w of atom B (lj.hip), and S the ASR! self blocks.
//
// Each workgroup owns `cpw` consecutive columns of one slab kk (a contiguous
// ncols*n range of Psi), first resolves each column's nonzero rows into LDS, then
// streams the whole range with 16-byte stores (8-byte for FP32), zeros included:
// HBM sees exactly 8 n^3 (+ 4 n^3) bytes of writes and nothing else.
#include <cstdlib>
#include "bf16_split.hpp"
#include "fc3_internal.hpp"

namespace fc3 {
namespace {

constexpr int kMaxCols = 256;

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int find_slot(const int* __restrict__ list, int cnt, int key)
{
    int lo = 0, hi = cnt - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = list[mid];
        if (v == key) return mid;
        if (v < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

__device__ __forceinline__ double tol_cut(double v, double tol) { return fabs(v) < tol ? 0.0 : v; }

template <int VEC, bool W64, bool W32, bool MW>
__global__ void __launch_bounds__(256)
k_write_ifc3_pairwise(const int* __restrict__ count, const int* __restrict__ nbr,
                      const int* __restrict__ rev, int N, int cap, const double* __restrict__ E,
                      const double* __restrict__ S, double tol, const double* __restrict__ masses,
                      double* __restrict__ psi64, float* __restrict__ psi32, int cpw)
{
    const long n = 3L * N;
    const int kk = blockIdx.y;
    const int K = kk / 3, g = kk - 3 * (kk / 3);
    const long col0 = (long)blockIdx.x * cpw;
    const int ncols = (int)((n - col0) < cpw ? (n - col0) : cpw);

    __shared__ int s_type[kMaxCols];     // 0 zero column, 1 pair column, 2 self column
    __shared__ int s_J[kMaxCols];
    __shared__ double s_v[kMaxCols][6];  // rows J (0..2) and K (3..5), after tol
    __shared__ float s_f[kMaxCols][6];   // same, mass weighted, FP32

    const int cntK = count[K];
    const int* nbrK = nbr + (size_t)K * cap;
    for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
        const long jj = col0 + c;
        const int J = (int)(jj / 3), b = (int)(jj - 3L * J);
        int type = 0;
        if (J == K) {
            type = 2;
        } else {
            const int s = find_slot(nbrK, cntK, J);
            const int rs = s >= 0 ? rev[(size_t)K * cap + s] : -1;
            if (rs >= 0) {
                type = 1;
                const double* e = E + ((size_t)J * cap + rs) * 27;
                double mJ = 0.0, mK = 0.0;
                if (MW) { mJ = masses[J]; mK = masses[K]; }
                for (int a = 0; a < 3; ++a) {
                    const double vJ = tol_cut(e[a + 3 * b + 9 * g], tol);
                    const double vK = tol_cut(-e[b + 3 * g + 9 * a], tol);
                    s_v[c][a] = vJ;
                    s_v[c][3 + a] = vK;
                    if (MW) {
                        s_f[c][a] = (float)(vJ != 0.0 ? vJ / sqrt((mJ * mJ) * mK) : vJ);
                        s_f[c][3 + a] = (float)(vK != 0.0 ? vK / sqrt((mK * mJ) * mK) : vK);
                    } else {
                        s_f[c][a] = (float)vJ;
                        s_f[c][3 + a] = (float)vK;
                    }
                }
            }
        }
        s_type[c] = type;
        s_J[c] = J;
    }
    __syncthreads();

    const long base = (long)kk * n * n + col0 * n;
    const long total = (long)ncols * n;  // elements in this workgroup's range
    const double mK = MW ? masses[K] : 0.0;

    // element value at (column c, row r)
    auto value = [&](int c, int r, double& v64, float& v32) {
        const int type = s_type[c];
        v64 = 0.0;
        v32 = 0.0f;
        if (type == 0) return;
        const int I = r / 3, a = r - 3 * (r / 3);
        if (type == 1) {
            const int J = s_J[c];
            if (I == J) { v64 = s_v[c][a]; v32 = s_f[c][a]; }
            else if (I == K) { v64 = s_v[c][3 + a]; v32 = s_f[c][3 + a]; }
            return;
        }
        // self column: J == K, b = component of the column
        const long jj = col0 + c;
        const int b = (int)(jj - 3L * K);
        double v;
        if (I == K) {
            v = S[(size_t)K * 27 + a + 3 * b + 9 * g];
        } else {
            const int s = find_slot(nbrK, cntK, I);
            if (s < 0) return;
            v = E[((size_t)K * cap + s) * 27 + b + 3 * g + 9 * a];
        }
        v = tol_cut(v, tol);
        v64 = v;
        if (MW) {
            const double mI = masses[I];
            v32 = (float)(v != 0.0 ? v / sqrt((mI * mK) * mK) : v);
        } else {
            v32 = (float)v;
        }
    };

    if (VEC == 2) {
        const long pairs = total >> 1;
        long e0 = 2L * threadIdx.x;
        int c = (int)(e0 / n);
        int r = (int)(e0 - (long)c * n);
        const int step = 2 * blockDim.x;
        for (long q = threadIdx.x; q < pairs; q += blockDim.x) {
            double a0, a1;
            float f0, f1;
            value(c, r, a0, f0);
            value(c, r + 1, a1, f1);  // n even: both rows in column c
            if (W64) {
                dvec2 v2 = {a0, a1};
                __builtin_nontemporal_store(v2, reinterpret_cast<dvec2*>(psi64 + base) + q);
            }
            if (W32) {
                fvec2 f2 = {f0, f1};
                __builtin_nontemporal_store(f2, reinterpret_cast<fvec2*>(psi32 + base) + q);
            }
            r += step;
            while (r >= n) { r -= (int)n; ++c; }
        }
    } else {
        long e = threadIdx.x;
        int c = (int)(e / n);
        int r = (int)(e - (long)c * n);
        for (; e < total; e += blockDim.x) {
            double a0;
            float f0;
            value(c, r, a0, f0);
            if (W64) psi64[base + e] = a0;
            if (W32) psi32[base + e] = f0;
            r += blockDim.x;
            while (r >= n) { r -= (int)n; ++c; }
        }
    }
}

template <int VEC, bool W64, bool W32, bool MW>
void launch_writer(hipStream_t s, const NeighborList& nl, const Ifc3Blocks& b, double tol,
                   const double* m, double* p64, float* p32, int cpw)
{
    const long n = 3L * nl.n_atoms;
    dim3 grid((unsigned)((n + cpw - 1) / cpw), (unsigned)n);
    hipLaunchKernelGGL((k_write_ifc3_pairwise<VEC, W64, W32, MW>), grid, dim3(256), 0, s,
                       nl.count, nl.nbr, nl.rev, nl.n_atoms, nl.cap, b.E, b.S, tol, m, p64, p32,
                       cpw);
}

template <int VEC>
void dispatch_writer(hipStream_t s, const NeighborList& nl, const Ifc3Blocks& b, double tol,
                     const double* m, double* p64, float* p32, int cpw)
{
    const bool w64 = p64 != nullptr, w32 = p32 != nullptr, mw = m != nullptr;
    if (w64 && w32) {
        if (mw) launch_writer<VEC, true, true, true>(s, nl, b, tol, m, p64, p32, cpw);
        else launch_writer<VEC, true, true, false>(s, nl, b, tol, m, p64, p32, cpw);
    } else if (w64) {
        launch_writer<VEC, true, false, false>(s, nl, b, tol, m, p64, p32, cpw);
    } else {
        if (mw) launch_writer<VEC, false, true, true>(s, nl, b, tol, m, p64, p32, cpw);
        else launch_writer<VEC, false, true, false>(s, nl, b, tol, m, p64, p32, cpw);
    }
}

// In-place mass weighting of a dense FP64 tensor: only nonzero entries are
// rewritten (third_order_analytical.jl:195-210).
__global__ void __launch_bounds__(256)
k_mass_weight3(double* __restrict__ psi, int N, int D, const double* __restrict__ m)
{
    const long n = (long)D * N;
    const int kk = blockIdx.y;
    const double mk = m[kk / D];
    const long j0 = (long)blockIdx.x;  // one column per block row-chunk
    double* col = psi + (long)kk * n * n + j0 * n;
    const double mj = m[j0 / D];
    for (long i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = col[i];
        if (v != 0.0) col[i] = v / sqrt((m[i / D] * mj) * mk);
    }
}

// Sparse mode-1 product fused to the assembly (no dense Psi).  Slab kk = 3K + g holds
// nonzeros only in the columns jj_t = 3 J_t + b of K's neighbourhood (t = 3 s + b; J_0 =
// K, the self column; J_s = nbr[K][s - 1], pair columns):
//   U_K(g n + p, t) = sum_i phiT[p + n i] Psi32[i, jj_t, kk]
// over the nonzero rows i of column (jj_t, kk), with Psi32 the value the FP32 writer
// stores there (tolerance, mass weighting and rounding as above):
//   pair column (s >= 1, J = J_s, when K is in J's list at slot rs): rows 3J + a with
//     E[J][rs][a + 3b + 9g], then rows 3K + a with -E[J][rs][b + 3g + 9a];
//   self column (s = 0): rows 3K + a with S[K][a + 3b + 9g], then for every neighbour
//     slot s' the rows 3 I + a (I = nbr[K][s']) with E[K][s'][b + 3g + 9a].
// Each sum is an fmaf chain from zero in that row order.  U goes out split into bf16
// hi/lo, K-blocked over t: hi at U[K 2 M Kc + (t / 16) 16 M + 16 (g n + p) + t % 16] with
// M = 3n, lo at + M Kc; columns t past the count are zero.
// Workgroup per (256 consecutive p, atom K = blockIdx.y); thread p computes every column
// and slab of its row.  The coefficients of K's neighbourhood (81 per slot: the pair
// column's J and K rows and the slot's share of the self column, tolerance, mass weight
// and FP32 rounding applied) go to LDS once per workgroup; a neighbour's three phi
// values then serve its 3 columns x 3 slabs, and every U row is written once, as 32-byte
// runs per (16 columns, slab) that are contiguous across the workgroup's lanes.
// Dynamic LDS: lj_mode1_sparse_lds(cap) bytes; with LDS = false the same coefficients
// come from k_lj_mode1_coeffs's global copy instead (for neighbourhoods past 64 KB).
//
// Atom K's coefficients (per slot s' < count[K]: cJ[27], cK[27], cS[27] (floats), then
// cSelf[27], then the slots' first rows 3 J (ints)) into coef / cSelf / jrow, by threads
// tid, tid + stride, ...
template <bool MW>
__device__ __forceinline__ void lj_stage_coeffs(int K, int tid, int stride,
                                                const int* __restrict__ count,
                                                const int* __restrict__ nbr,
                                                const int* __restrict__ rev, int cap,
                                                const double* __restrict__ E,
                                                const double* __restrict__ S, double tol,
                                                const double* __restrict__ masses, float* coef,
                                                float* cSelf, int* jrow)
{
    const int cntK = count[K];
    const int* nbrK = nbr + (size_t)K * cap;
    const double mK = MW ? masses[K] : 0.0;
    for (int q = tid; q < 27 * cntK + 27; q += stride) {
        if (q >= 27 * cntK) {
            const int e = q - 27 * cntK;
            const double v = tol_cut(S[(size_t)K * 27 + e], tol);
            cSelf[e] = MW ? (float)(v != 0.0 ? v / sqrt((mK * mK) * mK) : v) : (float)v;
            continue;
        }
        const int sl = q / 27, e = q - 27 * sl;   // e = a + 3 b + 9 g
        const int a = e % 3, b = (e / 3) % 3, g = e / 9;
        const int J = nbrK[sl];
        const int rs = rev[(size_t)K * cap + sl];   // slot of K in J's list, or < 0
        const double mJ = MW ? masses[J] : 0.0;
        float fj = 0.0f, fk = 0.0f;
        if (rs >= 0) {
            const double* ej = E + ((size_t)J * cap + rs) * 27;
            const double vJ = tol_cut(ej[a + 3 * b + 9 * g], tol);
            const double vK = tol_cut(-ej[b + 3 * g + 9 * a], tol);
            fj = MW ? (float)(vJ != 0.0 ? vJ / sqrt((mJ * mJ) * mK) : vJ) : (float)vJ;
            fk = MW ? (float)(vK != 0.0 ? vK / sqrt((mK * mJ) * mK) : vK) : (float)vK;
        }
        const double vS = tol_cut(E[((size_t)K * cap + sl) * 27 + b + 3 * g + 9 * a], tol);
        const float fs = MW ? (float)(vS != 0.0 ? vS / sqrt((mJ * mK) * mK) : vS) : (float)vS;
        coef[81 * sl + e] = fj;
        coef[81 * sl + 27 + e] = fk;
        coef[81 * sl + 54 + e] = fs;
        if (e == 0) jrow[sl] = 3 * J;
    }
}

// Floats of one atom's coefficient record in the global copy: 81 cap + 27, then cap ints
// (the LDS layout), padded to 4 bytes x 4.
__host__ __device__ inline long lj_coeff_stride(int cap) { return (81L * cap + 27 + cap + 3) / 4 * 4; }

// Workgroup per atom K: its coefficients into gc + K lj_coeff_stride(cap).
template <bool MW>
__global__ void __launch_bounds__(256)
k_lj_mode1_coeffs(const int* __restrict__ count, const int* __restrict__ nbr,
                  const int* __restrict__ rev, int cap, const double* __restrict__ E,
                  const double* __restrict__ S, double tol, const double* __restrict__ masses,
                  float* __restrict__ gc)
{
    const int K = blockIdx.x;
    float* coef = gc + (size_t)K * lj_coeff_stride(cap);
    lj_stage_coeffs<MW>(K, threadIdx.x, blockDim.x, count, nbr, rev, cap, E, S, tol, masses,
                        coef, coef + 81 * cap, reinterpret_cast<int*>(coef + 81 * cap + 27));
}

template <bool MW, bool LDS>
__global__ void __launch_bounds__(256)
k_lj_mode1_gathered(const int* __restrict__ count, const int* __restrict__ nbr,
                    const int* __restrict__ rev, int N, int cap, const double* __restrict__ E,
                    const double* __restrict__ S, double tol, const double* __restrict__ masses,
                    const float* __restrict__ phiT, int Kc, const float* __restrict__ gc,
                    __bf16* __restrict__ U)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char dyn[];
    const int n = 3 * N;
    const int K = blockIdx.y;
    const int cntK = count[K];
    const float* coef;
    const float* cSelf;
    const int* jrow;
    if constexpr (LDS) {
        float* lc = reinterpret_cast<float*>(dyn);
        lj_stage_coeffs<MW>(K, threadIdx.x, blockDim.x, count, nbr, rev, cap, E, S, tol, masses,
                            lc, lc + 81 * cap, reinterpret_cast<int*>(lc + 81 * cap + 27));
        __syncthreads();
        coef = lc;
    } else {
        coef = gc + (size_t)K * lj_coeff_stride(cap);
    }
    cSelf = coef + 81 * cap;
    jrow = reinterpret_cast<const int*>(coef + 81 * cap + 27);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;   // no barrier follows
    const float* ph = phiT + p;
    float fK[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) fK[a] = ph[(long)n * (3 * K + a)];
    // the self column's 9 values (b, g): K's rows, then every neighbour's rows
    float sv[9];
#pragma unroll
    for (int bg = 0; bg < 9; ++bg) {
        float x = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) x = fmaf(fK[a], cSelf[a + 3 * bg], x);
        sv[bg] = x;
    }
    for (int sl = 0; sl < cntK; ++sl) {
        const float* fI = ph + (long)n * jrow[sl];
        const float f0 = fI[0], f1 = fI[n], f2 = fI[2L * n];
        const float* cs = coef + 81 * sl + 54;
#pragma unroll
        for (int bg = 0; bg < 9; ++bg) {
            float x = sv[bg];
            x = fmaf(f0, cs[3 * bg], x);
            x = fmaf(f1, cs[3 * bg + 1], x);
            x = fmaf(f2, cs[3 * bg + 2], x);
            sv[bg] = x;
        }
    }
    const long M = 3L * n;
    __bf16* Uk = U + (long)K * 2 * M * Kc + 16L * p;
    for (int tb = 0; tb < Kc / 16; ++tb) {
        float acc[3][16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int t = 16 * tb + c, sl = t / 3, b = t - 3 * sl;
            if (sl == 0) {
#pragma unroll
                for (int g = 0; g < 3; ++g) acc[g][c] = sv[b + 3 * g];
            } else if (sl - 1 < cntK) {
                const float* fJ = ph + (long)n * jrow[sl - 1];
                const float j0 = fJ[0], j1 = fJ[n], j2 = fJ[2L * n];
                const float* cj = coef + 81 * (sl - 1) + 3 * b;
#pragma unroll
                for (int g = 0; g < 3; ++g) {
                    const float* cg = cj + 9 * g;
                    float x = fmaf(j0, cg[0], 0.0f);
                    x = fmaf(j1, cg[1], x);
                    x = fmaf(j2, cg[2], x);
                    x = fmaf(fK[0], cg[27], x);
                    x = fmaf(fK[1], cg[28], x);
                    x = fmaf(fK[2], cg[29], x);
                    acc[g][c] = x;
                }
            } else {
#pragma unroll
                for (int g = 0; g < 3; ++g) acc[g][c] = 0.0f;
            }
        }
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            bf16x8 h0, h1, l0, l1;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                __bf16 hh, ll;
                split_bf16(acc[g][c], hh, ll);
                h0[c] = hh;
                l0[c] = ll;
                split_bf16(acc[g][8 + c], hh, ll);
                h1[c] = hh;
                l1[c] = ll;
            }
            __bf16* Uh = Uk + (long)tb * 16 * M + 16L * g * n;
            bf16x8* dh = reinterpret_cast<bf16x8*>(Uh);
            bf16x8* dl = reinterpret_cast<bf16x8*>(Uh + M * Kc);
            dh[0] = h0;
            dh[1] = h1;
            dl[0] = l0;
            dl[1] = l1;
        }
    }
}

}  // namespace

int write_ifc3_pairwise(hipStream_t s, const NeighborList& nl, const Ifc3Blocks& b, double tol,
                        const double* d_masses, double* psi64, float* psi32)
{
    if (!psi64 && !psi32) return FC3_OK;
    const long n = 3L * nl.n_atoms;
    ProfScope prof(FC3_KERNEL_IFC3_WRITE, s);
    long cpw = 65536 / n;
    if (cpw < 1) cpw = 1;
    if (cpw > kMaxCols) cpw = kMaxCols;
    if (n % 2 == 0)
        dispatch_writer<2>(s, nl, b, tol, d_masses, psi64, psi32, (int)cpw);
    else
        dispatch_writer<1>(s, nl, b, tol, d_masses, psi64, psi32, (int)cpw);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

size_t lj_mode1_sparse_lds(int cap)
{
    return (size_t)(81 * cap + 27) * sizeof(float) + (size_t)cap * sizeof(int);
}

size_t lj_mode1_global_bytes(int N, int cap)
{
    return Carve::need(sizeof(float) * (size_t)N * lj_coeff_stride(cap));
}

int lj_mode1_gathered(hipStream_t s, const NeighborList& nl, const Ifc3Blocks& b, double tol,
                      const double* d_masses, const float* phiT, int Kc, __bf16* U,
                      void* scratch, size_t scratch_bytes)
{
    const int n = 3 * nl.n_atoms;
    if (Kc % 16 != 0 || Kc < 3 * (nl.cap + 1))
        return set_error(FC3_ERR_ARGUMENT, "fused mode 1: gathered K too small");
    // FC3_LJ_FUSED_GLOBAL=1 takes the global-coefficient kernel for any neighbourhood (the
    // two kernels multiply the same values in the same order)
    const size_t lds = lj_mode1_sparse_lds(nl.cap);
    const bool use_lds = lds <= 64 * 1024 && !(scratch && std::getenv("FC3_LJ_FUSED_GLOBAL"));
    float* gc = nullptr;
    if (!use_lds) {
        if (scratch_bytes < lj_mode1_global_bytes(nl.n_atoms, nl.cap))
            return set_error(scratch ? FC3_ERR_OUT_OF_MEMORY : FC3_ERR_UNSUPPORTED,
                             "fused mode 1: no scratch for the neighbourhoods' coefficients");
        gc = static_cast<float*>(scratch);
    }
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)nl.n_atoms);
    ProfScope prof(FC3_KERNEL_IFC3_WRITE, s);
    if (!use_lds) {
        if (d_masses)
            hipLaunchKernelGGL(k_lj_mode1_coeffs<true>, dim3((unsigned)nl.n_atoms), dim3(256), 0,
                               s, nl.count, nl.nbr, nl.rev, nl.cap, b.E, b.S, tol, d_masses, gc);
        else
            hipLaunchKernelGGL(k_lj_mode1_coeffs<false>, dim3((unsigned)nl.n_atoms), dim3(256), 0,
                               s, nl.count, nl.nbr, nl.rev, nl.cap, b.E, b.S, tol, d_masses, gc);
        FC3_HIP(hipGetLastError());
    }
#define FC3_LJ_M1(MW, L)                                                                       \
    hipLaunchKernelGGL((k_lj_mode1_gathered<MW, L>), grid, dim3(256), L ? lds : 0, s, nl.count, \
                       nl.nbr, nl.rev, nl.n_atoms, nl.cap, b.E, b.S, tol, d_masses, phiT, Kc,   \
                       gc, U)
    if (use_lds) {
        if (d_masses) FC3_LJ_M1(true, true); else FC3_LJ_M1(false, true);
    } else {
        if (d_masses) FC3_LJ_M1(true, false); else FC3_LJ_M1(false, false);
    }
#undef FC3_LJ_M1
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

int mass_weight3(hipStream_t s, double* psi, int N, int D, const double* d_masses)
{
    const long n = (long)D * N;
    hipLaunchKernelGGL(k_mass_weight3, dim3((unsigned)n, (unsigned)n), dim3(256), 0, s, psi, N, D,
                       d_masses);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

}  // namespace fc3
