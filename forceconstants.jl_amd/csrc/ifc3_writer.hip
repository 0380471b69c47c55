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
// K, the self column; J_s = nbr[K][s - 1], pair columns).  For the 16 such columns
// t = 16 blockIdx.x + c,
//   U_K(g n + p, t) = sum_i phiT[p + n i] Psi32[i, jj_t, kk]
// over the nonzero rows i of column (jj_t, kk), with Psi32 the value the FP32 writer
// stores there (tolerance, mass weighting and rounding as above).  U goes out split
// into bf16 hi/lo, K-blocked over t: hi at U[K 2 M Kc + blockIdx.x 16 M + 16 (g n + p) +
// c] with M = 3n, lo at + M Kc; columns t past the count are zero.  Pair columns have 6
// entries and the 3 self columns (all in the first block) 3 + 3 count[K] each, so
// 16 * 6 + 3 * (3 + 3 cap) entries bound a block (dynamic LDS: rows, then values).
template <bool MW>
__global__ void __launch_bounds__(256)
k_lj_mode1_gathered(const int* __restrict__ count, const int* __restrict__ nbr,
                    const int* __restrict__ rev, int N, int cap, const double* __restrict__ E,
                    const double* __restrict__ S, double tol, const double* __restrict__ masses,
                    const float* __restrict__ phiT, int Kc, __bf16* __restrict__ U)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char dyn[];
    const int n = 3 * N;
    const int kk = blockIdx.y;
    const int K = kk / 3, g = kk - 3 * (kk / 3);
    const int tb = blockIdx.x;
    __shared__ int s_cnt[16], s_off[17];
    int* e_row = reinterpret_cast<int*>(dyn);
    float* e_val = reinterpret_cast<float*>(dyn) + (16 * 6 + 3 * (3 + 3 * cap));
    const int cntK = count[K];
    const int* nbrK = nbr + (size_t)K * cap;
    const double mK = MW ? masses[K] : 0.0;
    // 1. entries per column (thread c < 16 owns column c)
    const int t = threadIdx.x;
    int J = -1, b = 0, type = 0, rs = -1;
    if (t < 16) {
        const int col = 16 * tb + t, sl = col / 3;
        b = col - 3 * sl;
        if (sl == 0) {
            J = K;
            type = 2;
        } else if (sl - 1 < cntK) {
            J = nbrK[sl - 1];
            rs = rev[(size_t)K * cap + sl - 1];   // slot of K in J's list
            if (rs >= 0) type = 1;
        }
        s_cnt[t] = type == 1 ? 6 : type == 2 ? 3 + 3 * cntK : 0;
    }
    __syncthreads();
    if (t == 0) {
        int o = 0;
        for (int c = 0; c < 16; ++c) { s_off[c] = o; o += s_cnt[c]; }
        s_off[16] = o;
    }
    __syncthreads();
    if (t < 16 && type != 0) {
        int* row = e_row + s_off[t];
        float* val = e_val + s_off[t];
        if (type == 1) {
            const double* e = E + ((size_t)J * cap + rs) * 27;
            const double mJ = MW ? masses[J] : 0.0;
            for (int a = 0; a < 3; ++a) {
                const double vJ = tol_cut(e[a + 3 * b + 9 * g], tol);
                const double vK = tol_cut(-e[b + 3 * g + 9 * a], tol);
                row[a] = 3 * J + a;
                row[3 + a] = 3 * K + a;
                if (MW) {
                    val[a] = (float)(vJ != 0.0 ? vJ / sqrt((mJ * mJ) * mK) : vJ);
                    val[3 + a] = (float)(vK != 0.0 ? vK / sqrt((mK * mJ) * mK) : vK);
                } else {
                    val[a] = (float)vJ;
                    val[3 + a] = (float)vK;
                }
            }
        } else {
            // self column: rows of K (ASR! self block) and of every neighbour I of K
            for (int a = 0; a < 3; ++a) {
                const double v = tol_cut(S[(size_t)K * 27 + a + 3 * b + 9 * g], tol);
                row[a] = 3 * K + a;
                val[a] = MW ? (float)(v != 0.0 ? v / sqrt((mK * mK) * mK) : v) : (float)v;
            }
            for (int sl = 0; sl < cntK; ++sl) {
                const int I = nbrK[sl];
                const double mI = MW ? masses[I] : 0.0;
                for (int a = 0; a < 3; ++a) {
                    const double v =
                        tol_cut(E[((size_t)K * cap + sl) * 27 + b + 3 * g + 9 * a], tol);
                    row[3 + 3 * sl + a] = 3 * I + a;
                    val[3 + 3 * sl + a] =
                        MW ? (float)(v != 0.0 ? v / sqrt((mI * mK) * mK) : v) : (float)v;
                }
            }
        }
    }
    __syncthreads();
    // 2. thread p: the 16 columns' sums, split and stored as two 32-byte runs
    const long M = 3L * n;
    __bf16* Uh = U + (long)K * 2 * M * Kc + (long)tb * 16 * M + 16L * g * n;
    for (int p = t; p < n; p += blockDim.x) {
        float acc[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            float a = 0.0f;
            for (int e = s_off[c]; e < s_off[c + 1]; ++e)
                a = fmaf(phiT[p + (long)n * e_row[e]], e_val[e], a);
            acc[c] = a;
        }
        bf16x8 h0, h1, l0, l1;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            __bf16 hh, ll;
            split_bf16(acc[c], hh, ll);
            h0[c] = hh;
            l0[c] = ll;
            split_bf16(acc[8 + c], hh, ll);
            h1[c] = hh;
            l1[c] = ll;
        }
        bf16x8* dh = reinterpret_cast<bf16x8*>(Uh + 16L * p);
        bf16x8* dl = reinterpret_cast<bf16x8*>(Uh + M * Kc + 16L * p);
        dh[0] = h0;
        dh[1] = h1;
        dl[0] = l0;
        dl[1] = l1;
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
    return (size_t)(16 * 6 + 3 * (3 + 3 * cap)) * (sizeof(int) + sizeof(float));
}

int lj_mode1_gathered(hipStream_t s, const NeighborList& nl, const Ifc3Blocks& b, double tol,
                      const double* d_masses, const float* phiT, int Kc, __bf16* U)
{
    const int n = 3 * nl.n_atoms;
    const size_t lds = lj_mode1_sparse_lds(nl.cap);
    if (lds > 64 * 1024)
        return set_error(FC3_ERR_UNSUPPORTED, "fused mode 1: too many neighbours per atom");
    if (Kc % 16 != 0 || Kc < 3 * (nl.cap + 1))
        return set_error(FC3_ERR_ARGUMENT, "fused mode 1: gathered K too small");
    dim3 grid((unsigned)(Kc / 16), (unsigned)n);
    ProfScope prof(FC3_KERNEL_IFC3_WRITE, s);
    if (d_masses)
        hipLaunchKernelGGL(k_lj_mode1_gathered<true>, grid, dim3(256), lds, s, nl.count, nl.nbr,
                           nl.rev, nl.n_atoms, nl.cap, b.E, b.S, tol, d_masses, phiT, Kc, U);
    else
        hipLaunchKernelGGL(k_lj_mode1_gathered<false>, grid, dim3(256), lds, s, nl.count, nl.nbr,
                           nl.rev, nl.n_atoms, nl.cap, b.E, b.S, tol, d_masses, phiT, Kc, U);
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
