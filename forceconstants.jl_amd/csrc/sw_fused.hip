// Fused Stillinger-Weber third_order + mcc3 without the dense Psi: the mode-1 product over
// the nonzeros of the reduced block table of one SW assembly, in the layout the gathered
// modes 2 and 3 read (mcc3_from_gathered).
//
// The table holds U blocks (I, J, K) sorted by key (K N + J) N + I; block p's values are
// uval[27 p + a + 3 b + 9 g] for the entry (3I + a, 3J + b, 3K + g) after ASR! and before
// the tolerance (sw.hip).  A block exists only for atoms of one triplet, so the columns of
// slab atom K are the atoms J within two neighbour hops of K.
//
// Columns of slab atom K, in the convention of the LJ fused path: slot 0 is J = K, slots
// 1 .. colcnt[K] the other atoms J with a block (., J, K), ascending, at
// collist[K cap2 + s - 1]; column t = 3 s + b.  seg[K][2 s], seg[K][2 s + 1] bound slot s's
// blocks in the table (rows I ascending).
#include <cstdlib>

#include "bf16_split.hpp"
#include "fc3_internal.hpp"

namespace fc3 {
namespace {

typedef unsigned long long u64;

__device__ __forceinline__ int lower_bound_key(const u64* a, int n, u64 key)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int key_j(const u64* ukey, int p, int N)
{
    return (int)((ukey[p] / (u64)N) % (u64)N);
}

// Wave per slab atom K (4 per workgroup): the number of column atoms J != K and of blocks
// of slab K, with their maxima over K in maxes[0], maxes[1] (integer atomics: the result
// does not depend on the order).  The wave walks the slab's blocks 64 at a time; a block
// opens a column atom when its J differs from the previous block's.
__global__ void __launch_bounds__(256) k_sw_col_counts(const u64* __restrict__ ukey, int U, int N,
                                                       int* __restrict__ colcnt,
                                                       int* __restrict__ maxes)
{
    const int K = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (K >= N) return;   // wave-uniform
    const u64 NN = (u64)N * N;
    const int lo = lower_bound_key(ukey, U, (u64)K * NN);
    const int hi = lower_bound_key(ukey, U, (u64)(K + 1) * NN);
    int cnt = 0;
    for (int base = lo; base < hi; base += 64) {
        const int p = base + lane;
        bool open = false;
        if (p < hi) {
            const int J = key_j(ukey, p, N);
            open = J != K && (p == lo || key_j(ukey, p - 1, N) != J);
        }
        cnt += __popcll(__ballot(open));
    }
    if (lane == 0) {
        colcnt[K] = cnt;
        atomicMax(&maxes[0], cnt);
        atomicMax(&maxes[1], hi - lo);
    }
}

// Wave per slab atom K: the column list (stride cap2) and the per-slot block ranges
// (2 (cap2 + 1) ints per K; slot 0 = J = K, slots past colcnt[K] are empty [lo, lo)).
// Every int of the slab's range table is written exactly once.
__global__ void __launch_bounds__(256) k_sw_col_lists(const u64* __restrict__ ukey, int U, int N,
                                                      int cap2, const int* __restrict__ colcnt,
                                                      int* __restrict__ collist,
                                                      int* __restrict__ seg)
{
    const int K = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (K >= N) return;   // wave-uniform
    const u64 NN = (u64)N * N;
    const int lo = lower_bound_key(ukey, U, (u64)K * NN);
    const int hi = lower_bound_key(ukey, U, (u64)(K + 1) * NN);
    const int cnt = colcnt[K];
    int* sg = seg + (size_t)K * 2 * (cap2 + 1);
    for (int s = cnt + 1 + lane; s <= cap2; s += 64) {   // empty slots
        sg[2 * s] = lo;
        sg[2 * s + 1] = lo;
    }
    const unsigned long long below = (1ull << lane) - 1ull;
    int slots = 0;
    bool self = false;
    for (int base = lo; base < hi; base += 64) {
        const int p = base + lane;
        int J = -1;
        bool open = false;
        if (p < hi) {
            J = key_j(ukey, p, N);
            open = p == lo || key_j(ukey, p - 1, N) != J;
        }
        const unsigned long long other = __ballot(open && J != K);
        self = self || __ballot(open && J == K) != 0;
        if (open) {
            // the column atom's blocks end where the next J's begin
            const int q = lower_bound_key(ukey + p, hi - p, ((u64)K * N + J + 1) * (u64)N) + p;
            if (J == K) {
                sg[0] = p;
                sg[1] = q;
            } else {
                const int s = slots + __popcll(other & below) + 1;
                collist[(size_t)K * cap2 + s - 1] = J;
                sg[2 * s] = p;
                sg[2 * s + 1] = q;
            }
        }
        slots += __popcll(other);
    }
    if (!self && lane == 0) {   // an atom without blocks of its own
        sg[0] = lo;
        sg[1] = lo;
    }
}

// Block p's coefficient e (= a + 3b + 9g) as the FP32 writer stores it: the tolerance,
// then mass weighting v / sqrt((m_I m_J) m_K), rounded to FP32; *I = its row atom.
template <bool MW>
__device__ __forceinline__ float sw_coeff(const u64* __restrict__ ukey,
                                          const double* __restrict__ uval, int p, int e, int N,
                                          double tol, const double* __restrict__ masses, int* I)
{
    const u64 key = ukey[p];
    *I = (int)(key % (u64)N);
    double v = uval[(size_t)p * 27 + e];
    if (fabs(v) < tol) v = 0.0;
    float f = (float)v;
    if (MW && v != 0.0) {
        const int J = (int)((key / (u64)N) % (u64)N), K = (int)(key / ((u64)N * N));
        f = (float)(v / sqrt(masses[*I] * masses[J] * masses[K]));
    }
    return f;
}

// Every block's coefficients into global memory (gcf[27 p + e], grow3[p] = 3 I), for slabs
// whose blocks do not fit in LDS.  Thread per coefficient.
template <bool MW>
__global__ void __launch_bounds__(256) k_sw_coeffs(const u64* __restrict__ ukey,
                                                   const double* __restrict__ uval, int U, int N,
                                                   double tol, const double* __restrict__ masses,
                                                   float* __restrict__ gcf,
                                                   int* __restrict__ grow3)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 27L * U) return;
    const int p = (int)(q / 27), e = (int)(q - 27L * p);
    int I;
    gcf[q] = sw_coeff<MW>(ukey, uval, p, e, N, tol, masses, &I);
    if (e == 0) grow3[p] = 3 * I;
}

// U_K(g n + p, t) = sum over the blocks (I, J_s, K) of slot s = t / 3 (b = t % 3), rows
// I ascending, of sum_a phiT[p + n (3I + a)] Psi32[3I + a, 3J_s + b, 3K + g], an fmaf chain
// from zero, with Psi32 the value the FP32 writer stores (sw_coeff).  Split into bf16 hi /
// lo, K-blocked over t: hi at U[K 2 M Kc + (t / 16) 16 M + 16 (g n + p) + t % 16] with
// M = 3n, lo at + M Kc; columns past the count are zero.  Workgroup per (256 rows p, atom
// K = blockIdx.y).  LDS: the slab's coefficients go to LDS once (dynamic: maxblk (27
// floats + 1 int)); else they are read from k_sw_coeffs's gcf / grow3, the same values.
template <bool MW, bool LDS>
__global__ void __launch_bounds__(256)
k_sw_mode1_gathered(const u64* __restrict__ ukey, const double* __restrict__ uval, int U, int N,
                    const int* __restrict__ colcnt, const int* __restrict__ seg, int cap2,
                    double tol, const double* __restrict__ masses, const float* __restrict__ phiT,
                    int Kc, const float* __restrict__ gcf, const int* __restrict__ grow3,
                    __bf16* __restrict__ Ug)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char dyn[];
    const int n = 3 * N;
    const int K = blockIdx.y;
    const int cnt = colcnt[K];
    const int* sg = seg + (size_t)K * 2 * (cap2 + 1);
    // the slab's blocks: [lo, hi), the union of the slots' ranges (contiguous: the table is
    // sorted by J within K, and every J of the slab is a slot)
    int lo = sg[0], hi = sg[1];
    for (int s = 1; s <= cnt; ++s) {
        lo = min(lo, sg[2 * s]);
        hi = max(hi, sg[2 * s + 1]);
    }
    const float* cf;
    const int* row3;
    if constexpr (LDS) {
        const int nb = hi - lo;
        float* lcf = reinterpret_cast<float*>(dyn);
        int* lrow3 = reinterpret_cast<int*>(lcf + 27 * (size_t)nb);
        for (int q = threadIdx.x; q < 27 * nb; q += blockDim.x) {
            const int blk = q / 27, e = q - 27 * blk;
            int I;
            lcf[q] = sw_coeff<MW>(ukey, uval, lo + blk, e, N, tol, masses, &I);
            if (e == 0) lrow3[blk] = 3 * I;
        }
        __syncthreads();
        cf = lcf;
        row3 = lrow3;
    } else {
        cf = gcf + 27 * (size_t)lo;
        row3 = grow3 + lo;
    }
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;   // no barrier follows
    const float* ph = phiT + p;
    const long M = 3L * n;
    __bf16* Uk = Ug + (long)K * 2 * M * Kc + 16L * p;
    for (int tb = 0; tb < Kc / 16; ++tb) {
        float acc[3][16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int t = 16 * tb + c, s = t / 3, b = t - 3 * s;
            float x0 = 0.0f, x1 = 0.0f, x2 = 0.0f;
            if (s <= cnt) {
                const int q0 = sg[2 * s] - lo, q1 = sg[2 * s + 1] - lo;
                for (int q = q0; q < q1; ++q) {
                    const float* fI = ph + (long)n * row3[q];
                    const float f0 = fI[0], f1 = fI[n], f2 = fI[2L * n];
                    const float* cb = cf + 27 * q + 3 * b;
                    x0 = fmaf(f0, cb[0], x0);
                    x0 = fmaf(f1, cb[1], x0);
                    x0 = fmaf(f2, cb[2], x0);
                    x1 = fmaf(f0, cb[9], x1);
                    x1 = fmaf(f1, cb[10], x1);
                    x1 = fmaf(f2, cb[11], x1);
                    x2 = fmaf(f0, cb[18], x2);
                    x2 = fmaf(f1, cb[19], x2);
                    x2 = fmaf(f2, cb[20], x2);
                }
            }
            acc[0][c] = x0;
            acc[1][c] = x1;
            acc[2][c] = x2;
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

__global__ void k_zero_ints(int* __restrict__ p, int cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = 0;
}

inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

size_t sw_fused_list_bytes(int N, int cap2_ub)
{
    return Carve::need(sizeof(int) * (size_t)N) + Carve::need(sizeof(int) * (size_t)N * cap2_ub) +
           Carve::need(sizeof(int) * (size_t)N * 2 * (cap2_ub + 1)) + Carve::need(sizeof(int) * 2);
}

int sw_fused_columns(hipStream_t s, const SwBlockTable& T, int N, int cap2_ub, void* ws,
                     SwColumns& out)
{
    Carve cv(ws, sw_fused_list_bytes(N, cap2_ub));
    out.colcnt = cv.take<int>(N);
    int* list_space = cv.take<int>((size_t)N * cap2_ub);
    int* seg_space = cv.take<int>((size_t)N * 2 * (cap2_ub + 1));
    int* maxes = cv.take<int>(2);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "fused SW columns workspace");
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, maxes, 2);
    hipLaunchKernelGGL(k_sw_col_counts, dim3(nblocks(N, 4)), dim3(256), 0, s, T.ukey, T.U, N,
                       out.colcnt, maxes);
    FC3_HIP(hipGetLastError());
    int hm[2] = {0, 0};
    FC3_HIP(hipMemcpyAsync(hm, maxes, sizeof(hm), hipMemcpyDeviceToHost, s));
    FC3_HIP(hipStreamSynchronize(s));
    if (hm[0] > cap2_ub)
        return set_error(FC3_ERR_UNSUPPORTED, "fused SW: a slab has more column atoms than bounded");
    out.cap2 = hm[0] > 0 ? hm[0] : 1;
    out.maxblk = hm[1];
    out.collist = list_space;
    out.seg = seg_space;
    hipLaunchKernelGGL(k_sw_col_lists, dim3(nblocks(N, 4)), dim3(256), 0, s, T.ukey, T.U, N,
                       out.cap2, out.colcnt, out.collist, out.seg);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

size_t sw_mode1_lds(int maxblk) { return (size_t)maxblk * (27 * sizeof(float) + sizeof(int)); }

size_t sw_mode1_global_bytes(int U)
{
    return Carve::need(sizeof(float) * 27 * (size_t)U) + Carve::need(sizeof(int) * (size_t)U);
}

int sw_mode1_gathered(hipStream_t s, const SwBlockTable& T, int N, const SwColumns& cols,
                      double tol, const double* d_masses, const float* phiT, int Kc, __bf16* U,
                      void* scratch, size_t scratch_bytes)
{
    const int n = 3 * N;
    if (Kc % 16 != 0 || Kc < 3 * (cols.cap2 + 1))
        return set_error(FC3_ERR_ARGUMENT, "fused SW mode 1: gathered K too small");
    // FC3_SW_FUSED_GLOBAL=1 takes the global-coefficient kernel for any slab size (the two
    // kernels multiply the same values in the same order)
    const size_t lds = sw_mode1_lds(cols.maxblk);
    const bool use_lds = lds <= 64 * 1024 && !std::getenv("FC3_SW_FUSED_GLOBAL");
    float* gcf = nullptr;
    int* grow3 = nullptr;
    if (!use_lds) {
        Carve cv(scratch, scratch_bytes);
        gcf = cv.take<float>(27 * (size_t)T.U);
        grow3 = cv.take<int>((size_t)T.U);
        if (!cv.ok())
            return set_error(FC3_ERR_OUT_OF_MEMORY, "fused SW mode 1: coefficient scratch");
    }
    dim3 grid(nblocks(n, 256), (unsigned)N);
    ProfScope prof(FC3_KERNEL_IFC3_WRITE, s);
    if (!use_lds && T.U > 0) {
        const unsigned gq = nblocks(27L * T.U, 256);
        if (d_masses)
            hipLaunchKernelGGL(k_sw_coeffs<true>, dim3(gq), dim3(256), 0, s, T.ukey, T.uval, T.U,
                               N, tol, d_masses, gcf, grow3);
        else
            hipLaunchKernelGGL(k_sw_coeffs<false>, dim3(gq), dim3(256), 0, s, T.ukey, T.uval, T.U,
                               N, tol, d_masses, gcf, grow3);
        FC3_HIP(hipGetLastError());
    }
#define FC3_SW_M1(MW, L)                                                                       \
    hipLaunchKernelGGL((k_sw_mode1_gathered<MW, L>), grid, dim3(256), L ? lds : 0, s, T.ukey,   \
                       T.uval, T.U, N, cols.colcnt, cols.seg, cols.cap2, tol, d_masses, phiT,   \
                       Kc, gcf, grow3, U)
    if (use_lds) {
        if (d_masses) FC3_SW_M1(true, true); else FC3_SW_M1(false, true);
    } else {
        if (d_masses) FC3_SW_M1(true, false); else FC3_SW_M1(false, false);
    }
#undef FC3_SW_M1
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

}  // namespace fc3
