// mcc3: K3[n,m,l] = sum_ijk Psi[i,j,k] phi[i,n] phi[j,m] phi[k,l]   (src/mcc.jl:15-17)
//
// Evaluated as three mode products on MFMA, ordered so that no n^3 temporary is
// needed beyond the output:
//   mode 3:     U[(i,j), l] = sum_k Psi[(i,j), k] phi[k, l]      one GEMM, M = n^2, written into K3
//   modes 1+2:  K3[:,:,l]  = phi^T U_l phi                       two n x n x n GEMMs per slab l,
//               S_l = U_l^T phi (scratch), K3_l = S_l^T phi      in place on slab l of K3
// Every GEMM has phi as its B operand, so phi is prepared once (zero-padded, and
// split into bf16 hi/lo for BF16X3); the A operand is FP32 and is split in the
// kernel while it is staged into LDS.
//
// Precisions
//   FP32:    v_mfma_f32_32x32x2_f32, FP32 operands and accumulation.
//   BF16X3:  a*b ~= ah*bh + ah*bl + al*bh with ah = bf16_rne(a), al = bf16_rne(a - ah),
//            on v_mfma_f32_32x32x16_bf16, FP32 accumulation.  Intermediates (U, S)
//            are FP32 and are split again when they are read.
#include "fc3_internal.hpp"

namespace fc3 {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BK = 32;                // K padding granularity of the prepared phi
constexpr int kXcd = 8;               // XCDs on MI355X

__device__ __forceinline__ void split_bf16(float x, __bf16& h, __bf16& l)
{
    h = static_cast<__bf16>(x);
    const float hf = static_cast<float>(h);
    l = static_cast<__bf16>(x - hf);
}

// ---------------------------------------------------------------------------------
// GEMM v2: 256 x 256 macro tile, 4 waves (2 x 2), each wave a 128 x 128 tile of 4 x 4
// MFMA 32x32 blocks; BK = 16; LDS double-buffered with a register stage in between,
// so global loads of step k+2 and the LDS stores of step k+1 overlap the MFMAs of
// step k and every step has one barrier.  The per-slab products of modes 1 and 2 are
// issued as one tall GEMM (rows of consecutive slabs stacked), so no tile is spent
// on a slab edge; the C layout C(m, c) = C[(m / R) * SS + m % R + c * CS] covers both
// the stacked-slab outputs and the single mode-3 output.
namespace v2 {

constexpr int BM = 256, BN = 256, BK = 16;
constexpr int NT = 256;
constexpr int PH = BK + 8;   // bf16 pitch (48 B): conflict-free b128 fragment reads
constexpr int PF = BK + 4;   // fp32 pitch (80 B)

struct G2 {
    const float* A;      // A(m, k) = A[m * sam + k * sak]
    long sam, sak;
    const void* Bhi;     // padded B(k, n) = B[n * ldb + k]: bf16 hi (or fp32)
    const void* Blo;     // bf16 lo
    int ldb;
    float* C;
    long R, SS, CS;      // C(m, c) = C[(m / R) * SS + m % R + c * CS]
    int M, N, K;
    int tiles_m, tiles_n;
    long per_xcd;
};

// A tile (BM x BK fp32) global -> registers, 16 values per thread.
//   AMC (A is M-contiguous): rows 4 mq + i, k 4 kq + kk       -> v[4 i + kk]
//   otherwise (K-contiguous): row t, k 4 q + kk                -> v[4 q + kk]
// With VEC the loads are 16-byte and branch-free: out-of-range rows / k are
// clamped to a valid address and zeroed afterwards (M % 4 == 0, K % 4 == 0).
template <bool AMC, bool VEC>
__device__ __forceinline__ void load_a2(const G2& g, int m0, int k0, float (&v)[16])
{
    const int t = threadIdx.x;
    if constexpr (AMC) {
        const int mq = t & 63, kq = t >> 6;
        const int mr = m0 + 4 * mq, kc = k0 + 4 * kq;
        if constexpr (VEC) {
            const bool vm = mr < g.M;
            const int mcl = vm ? mr : g.M - 4;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int k = kc + kk;
                const bool ok = vm && (k < g.K);
                const int kcl = k < g.K ? k : g.K - 1;
                const float4 x = *reinterpret_cast<const float4*>(g.A + (long)kcl * g.sak + mcl);
                v[0 + kk] = ok ? x.x : 0.0f;
                v[4 + kk] = ok ? x.y : 0.0f;
                v[8 + kk] = ok ? x.z : 0.0f;
                v[12 + kk] = ok ? x.w : 0.0f;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int m = mr + i, k = kc + kk;
                    v[4 * i + kk] = (m < g.M && k < g.K) ? g.A[(long)k * g.sak + (long)m * g.sam] : 0.0f;
                }
        }
    } else {
        const int m = m0 + t;
        if constexpr (VEC) {
            const bool vm = m < g.M;
            const long rowoff = (long)(vm ? m : g.M - 1) * g.sam;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int kc = k0 + 4 * q;
                const bool ok = vm && (kc < g.K);
                const float4 x = *reinterpret_cast<const float4*>(g.A + rowoff + (kc < g.K ? kc : 0));
                v[4 * q + 0] = ok ? x.x : 0.0f;
                v[4 * q + 1] = ok ? x.y : 0.0f;
                v[4 * q + 2] = ok ? x.z : 0.0f;
                v[4 * q + 3] = ok ? x.w : 0.0f;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int k = k0 + 4 * q + kk;
                    v[4 * q + kk] = (m < g.M && k < g.K) ? g.A[(long)m * g.sam + (long)k * g.sak] : 0.0f;
                }
        }
    }
}

// B tile of the prepared phi: 4 chunks of 16 bytes per thread, held in named
// registers (an indexed array here would be placed in scratch memory, which makes
// every load wait).  Always in bounds: phi is zero-padded to the tile grid and k0
// past the end is clamped by the caller.
struct BRegs {
    uint4 r0, r1, r2, r3;
};

template <int PREC>
__device__ __forceinline__ uint4 load_b_chunk(const G2& g, int n0, int k0, int idx)
{
    if constexpr (PREC == FC3_PREC_BF16X3) {
        const int arr = idx >> 9, col = (idx & 511) >> 1, half = idx & 1;
        const __bf16* base = static_cast<const __bf16*>(arr ? g.Blo : g.Bhi);
        return *reinterpret_cast<const uint4*>(base + (long)(n0 + col) * g.ldb + k0 + 8 * half);
    } else {
        const int col = idx >> 2, q = idx & 3;
        const float* base = static_cast<const float*>(g.Bhi);
        return *reinterpret_cast<const uint4*>(base + (long)(n0 + col) * g.ldb + k0 + 4 * q);
    }
}

template <int PREC>
__device__ __forceinline__ void load_b2(const G2& g, int n0, int k0, BRegs& b)
{
    const int t = threadIdx.x;
    b.r0 = load_b_chunk<PREC>(g, n0, k0, t);
    b.r1 = load_b_chunk<PREC>(g, n0, k0, t + NT);
    b.r2 = load_b_chunk<PREC>(g, n0, k0, t + 2 * NT);
    b.r3 = load_b_chunk<PREC>(g, n0, k0, t + 3 * NT);
}

template <bool AMC>
__device__ __forceinline__ void store_split(__bf16 (&sAh)[BM][PH], __bf16 (&sAl)[BM][PH],
                                            const float (&v)[16])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        bf16x4 h, l;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            __bf16 hh, ll;
            split_bf16(v[4 * r + kk], hh, ll);
            h[kk] = hh;
            l[kk] = ll;
        }
        if constexpr (AMC) {
            const int mq = t & 63, kq = t >> 6;
            *reinterpret_cast<bf16x4*>(&sAh[4 * mq + r][4 * kq]) = h;
            *reinterpret_cast<bf16x4*>(&sAl[4 * mq + r][4 * kq]) = l;
        } else {
            *reinterpret_cast<bf16x4*>(&sAh[t][4 * r]) = h;
            *reinterpret_cast<bf16x4*>(&sAl[t][4 * r]) = l;
        }
    }
}

template <int PREC, bool AMC, bool VEC>
__global__ void __launch_bounds__(NT, 1) k_gemm2(G2 g)
{
    const long wg = blockIdx.x;
    const long tile = (wg % kXcd) * g.per_xcd + wg / kXcd;
    if (tile >= (long)g.tiles_m * g.tiles_n) return;
    const int tm = (int)(tile / g.tiles_n), tn = (int)(tile - (long)tm * g.tiles_n);
    const int m0 = tm * BM, n0 = tn * BN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm0 = (wave & 1) * 128, wn0 = (wave >> 1) * 128;

    f32x16 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    // K is walked in pairs of steps so that the two register stages alternate
    // without copies (a copy at the loop back-edge would wait on the loads in
    // flight).  Steps past K read zero-padded phi (padded to a multiple of 2 BK)
    // and A zeroed by its bounds check, so they contribute 0.  Prefetches beyond
    // the last step are clamped to it and never consumed.
    const int nk = 2 * ((g.K + 2 * BK - 1) / (2 * BK));
    const int klast = (nk - 1) * BK;
    float av0[16], av1[16];
    BRegs bv0, bv1;
#define FC3_KN(x) ((x) < klast ? (x) : klast)

    if constexpr (PREC == FC3_PREC_BF16X3) {
        __shared__ __attribute__((aligned(16))) __bf16 sAh[2][BM][PH];
        __shared__ __attribute__((aligned(16))) __bf16 sAl[2][BM][PH];
        __shared__ __attribute__((aligned(16))) __bf16 sBh[2][BN][PH];
        __shared__ __attribute__((aligned(16))) __bf16 sBl[2][BN][PH];
        // chunk idx = t + 256 it: array idx / 512 (hi, lo), column (idx % 512) / 2, half idx % 2
#define FC3_STORE_B(st, bv)                                                                   \
        {                                                                                     \
            const int col = (t & 511) >> 1, half = t & 1;                                     \
            *reinterpret_cast<uint4*>(&sBh[st][col][8 * half]) = bv.r0;                       \
            *reinterpret_cast<uint4*>(&sBh[st][col + 128][8 * half]) = bv.r1;                 \
            *reinterpret_cast<uint4*>(&sBl[st][col][8 * half]) = bv.r2;                       \
            *reinterpret_cast<uint4*>(&sBl[st][col + 128][8 * half]) = bv.r3;                 \
        }
        const int kof = (lane >> 5) * 8;
        // one K step on LDS stage `cur`: 48 MFMAs; meanwhile the registers (avs, bvs)
        // of the next step go to the other stage and the step after next is fetched
        // into (avs, bvs).
#define FC3_STEP(cur, avs, bvs, knext)                                                        \
        {                                                                                     \
            bf16x8 ah[4], al[4], bh[4], bl[4];                                                \
            _Pragma("unroll") for (int i = 0; i < 4; ++i) {                                   \
                const int r = wm0 + i * 32 + (lane & 31);                                     \
                ah[i] = *reinterpret_cast<const bf16x8*>(&sAh[cur][r][kof]);                  \
                al[i] = *reinterpret_cast<const bf16x8*>(&sAl[cur][r][kof]);                  \
            }                                                                                 \
            _Pragma("unroll") for (int j = 0; j < 4; ++j) {                                   \
                const int c = wn0 + j * 32 + (lane & 31);                                     \
                bh[j] = *reinterpret_cast<const bf16x8*>(&sBh[cur][c][kof]);                  \
                bl[j] = *reinterpret_cast<const bf16x8*>(&sBl[cur][c][kof]);                  \
            }                                                                                 \
            _Pragma("unroll") for (int i = 0; i < 4; ++i)                                     \
                _Pragma("unroll") for (int j = 0; j < 4; ++j)                                 \
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0); \
            _Pragma("unroll") for (int i = 0; i < 4; ++i)                                     \
                _Pragma("unroll") for (int j = 0; j < 4; ++j)                                 \
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0); \
            store_split<AMC>(sAh[(cur) ^ 1], sAl[(cur) ^ 1], avs);                           \
            FC3_STORE_B((cur) ^ 1, bvs)                                                       \
            load_a2<AMC, VEC>(g, m0, FC3_KN(knext), avs);                                     \
            load_b2<PREC>(g, n0, FC3_KN(knext), bvs);                                         \
            _Pragma("unroll") for (int i = 0; i < 4; ++i)                                     \
                _Pragma("unroll") for (int j = 0; j < 4; ++j)                                 \
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0); \
            __syncthreads();                                                                  \
        }
        load_a2<AMC, VEC>(g, m0, 0, av0);
        load_b2<PREC>(g, n0, 0, bv0);
        store_split<AMC>(sAh[0], sAl[0], av0);
        FC3_STORE_B(0, bv0)
        load_a2<AMC, VEC>(g, m0, FC3_KN(BK), av1);
        load_b2<PREC>(g, n0, FC3_KN(BK), bv1);
        load_a2<AMC, VEC>(g, m0, FC3_KN(2 * BK), av0);
        load_b2<PREC>(g, n0, FC3_KN(2 * BK), bv0);
        __syncthreads();
        for (int kt = 0; kt < nk; kt += 2) {
            // even step: stage 0; stage 1 <- (av1, bv1); fetch kt+3 into (av1, bv1)
            FC3_STEP(0, av1, bv1, (kt + 3) * BK)
            // odd step: stage 1; stage 0 <- (av0, bv0); fetch kt+4 into (av0, bv0)
            FC3_STEP(1, av0, bv0, (kt + 4) * BK)
        }
#undef FC3_STEP
#undef FC3_STORE_B
    } else {
        __shared__ __attribute__((aligned(16))) float sA[2][BM][PF];
        __shared__ __attribute__((aligned(16))) float sB[2][BN][PF];
#define FC3_STORE_AF(st, avs, bvs)                                                            \
        _Pragma("unroll") for (int r = 0; r < 4; ++r) {                                       \
            const float4 x = make_float4(avs[4 * r], avs[4 * r + 1], avs[4 * r + 2], avs[4 * r + 3]); \
            if (AMC) *reinterpret_cast<float4*>(&sA[st][4 * (t & 63) + r][4 * (t >> 6)]) = x; \
            else *reinterpret_cast<float4*>(&sA[st][t][4 * r]) = x;                           \
        }                                                                                     \
        *reinterpret_cast<uint4*>(&sB[st][t >> 2][4 * (t & 3)]) = bvs.r0;                     \
        *reinterpret_cast<uint4*>(&sB[st][(t >> 2) + 64][4 * (t & 3)]) = bvs.r1;              \
        *reinterpret_cast<uint4*>(&sB[st][(t >> 2) + 128][4 * (t & 3)]) = bvs.r2;             \
        *reinterpret_cast<uint4*>(&sB[st][(t >> 2) + 192][4 * (t & 3)]) = bvs.r3;
        // MFMA step s, lane group h = lane / 32 uses k = 8 h + s (8 contiguous floats)
        const int kof = (lane >> 5) * 8;
#define FC3_STEPF(cur, avs, bvs, knext)                                                       \
        {                                                                                     \
            float fa[4][8], fb[4][8];                                                         \
            _Pragma("unroll") for (int i = 0; i < 4; ++i) {                                   \
                const int r = wm0 + i * 32 + (lane & 31);                                     \
                const float4 x0 = *reinterpret_cast<const float4*>(&sA[cur][r][kof]);         \
                const float4 x1 = *reinterpret_cast<const float4*>(&sA[cur][r][kof + 4]);     \
                fa[i][0] = x0.x; fa[i][1] = x0.y; fa[i][2] = x0.z; fa[i][3] = x0.w;           \
                fa[i][4] = x1.x; fa[i][5] = x1.y; fa[i][6] = x1.z; fa[i][7] = x1.w;           \
            }                                                                                 \
            _Pragma("unroll") for (int j = 0; j < 4; ++j) {                                   \
                const int c = wn0 + j * 32 + (lane & 31);                                     \
                const float4 x0 = *reinterpret_cast<const float4*>(&sB[cur][c][kof]);         \
                const float4 x1 = *reinterpret_cast<const float4*>(&sB[cur][c][kof + 4]);     \
                fb[j][0] = x0.x; fb[j][1] = x0.y; fb[j][2] = x0.z; fb[j][3] = x0.w;           \
                fb[j][4] = x1.x; fb[j][5] = x1.y; fb[j][6] = x1.z; fb[j][7] = x1.w;           \
            }                                                                                 \
            _Pragma("unroll") for (int s2 = 0; s2 < 4; ++s2)                                  \
                _Pragma("unroll") for (int i = 0; i < 4; ++i)                                 \
                    _Pragma("unroll") for (int j = 0; j < 4; ++j)                             \
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i][s2], fb[j][s2], acc[i][j], 0, 0, 0); \
            FC3_STORE_AF((cur) ^ 1, avs, bvs)                                                 \
            load_a2<AMC, VEC>(g, m0, FC3_KN(knext), avs);                                     \
            load_b2<PREC>(g, n0, FC3_KN(knext), bvs);                                         \
            _Pragma("unroll") for (int s2 = 4; s2 < 8; ++s2)                                  \
                _Pragma("unroll") for (int i = 0; i < 4; ++i)                                 \
                    _Pragma("unroll") for (int j = 0; j < 4; ++j)                             \
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i][s2], fb[j][s2], acc[i][j], 0, 0, 0); \
            __syncthreads();                                                                  \
        }
        load_a2<AMC, VEC>(g, m0, 0, av0);
        load_b2<PREC>(g, n0, 0, bv0);
        FC3_STORE_AF(0, av0, bv0)
        load_a2<AMC, VEC>(g, m0, FC3_KN(BK), av1);
        load_b2<PREC>(g, n0, FC3_KN(BK), bv1);
        load_a2<AMC, VEC>(g, m0, FC3_KN(2 * BK), av0);
        load_b2<PREC>(g, n0, FC3_KN(2 * BK), bv0);
        __syncthreads();
        for (int kt = 0; kt < nk; kt += 2) {
            FC3_STEPF(0, av1, bv1, (kt + 3) * BK)
            FC3_STEPF(1, av0, bv0, (kt + 4) * BK)
        }
#undef FC3_STEPF
#undef FC3_STORE_AF
    }
#undef FC3_KN

    // epilogue: register r of lane l holds row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + wn0 + j * 32 + (lane & 31);
            if (col >= g.N) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = m0 + wm0 + i * 32 + q * 8 + (lane >> 5) * 4;
                if (VEC && row + 3 < g.M) {
                    const long sl = row / g.R;
                    float* dst = g.C + sl * g.SS + (row - sl * g.R) + (long)col * g.CS;
                    *reinterpret_cast<float4*>(dst) =
                        make_float4(acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int m = row + r;
                        if (m < g.M) {
                            const long sl = m / g.R;
                            g.C[sl * g.SS + (m - sl * g.R) + (long)col * g.CS] = acc[i][j][4 * q + r];
                        }
                    }
                }
            }
        }
}

}  // namespace v2

// phi (n x n, col-major, ld n) -> zero-padded (Kpad x Npad, ld Kpad): fp32, or bf16 hi/lo.
__global__ void k_prep_phi(const float* __restrict__ phi, int n, int Kpad, int Npad, int prec,
                           float* __restrict__ pf, __bf16* __restrict__ ph, __bf16* __restrict__ pl)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)Kpad * Npad) return;
    const int c = (int)(t / Kpad), k = (int)(t - (long)c * Kpad);
    const float x = (k < n && c < n) ? phi[k + (long)c * n] : 0.0f;
    if (prec == FC3_PREC_FP32) {
        pf[t] = x;
    } else {
        __bf16 h, l;
        split_bf16(x, h, l);
        ph[t] = h;
        pl[t] = l;
    }
}

struct PhiPrep {
    const void* hi;
    const void* lo;
    int ldb;
};

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

size_t phi_bytes(int n, int prec)
{
    const size_t e = (size_t)round_up(n, BK) * round_up(n, v2::BN);
    return prec == FC3_PREC_FP32 ? Carve::need(e * 4) : 2 * Carve::need(e * 2);
}

int prep_phi(hipStream_t s, const float* phi, int n, int prec, Carve& cv, PhiPrep& out)
{
    const int Kpad = round_up(n, BK), Npad = round_up(n, v2::BN);
    const long e = (long)Kpad * Npad;
    float* pf = nullptr;
    __bf16 *ph = nullptr, *pl = nullptr;
    if (prec == FC3_PREC_FP32) {
        pf = cv.take<float>(e);
        out.hi = pf;
        out.lo = nullptr;
    } else {
        ph = cv.take<__bf16>(e);
        pl = cv.take<__bf16>(e);
        out.hi = ph;
        out.lo = pl;
    }
    out.ldb = Kpad;
    hipLaunchKernelGGL(k_prep_phi, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, s, phi, n,
                       Kpad, Npad, prec, pf, ph, pl);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

// C(m, c) = sum_k A(m, k) phi(k, c) with A(m, k) = A[m sam + k sak] and
// C(m, c) = C[(m / R) SS + m % R + c CS]  (v2 kernel).
int gemm2(hipStream_t s, int kind, int prec, const float* A, long sam, long sak,
          const PhiPrep& B, float* C, long R, long SS, long CS, long M, int N, int K)
{
    if (M >= (1L << 31)) return set_error(FC3_ERR_UNSUPPORTED, "mcc3: GEMM M too large");
    v2::G2 g;
    g.A = A; g.sam = sam; g.sak = sak;
    g.Bhi = B.hi; g.Blo = B.lo; g.ldb = B.ldb;
    g.C = C; g.R = R; g.SS = SS; g.CS = CS;
    g.M = (int)M; g.N = N; g.K = K;
    g.tiles_m = (int)((M + v2::BM - 1) / v2::BM);
    g.tiles_n = (N + v2::BN - 1) / v2::BN;
    const long tiles = (long)g.tiles_m * g.tiles_n;
    g.per_xcd = (tiles + kXcd - 1) / kXcd;
    const long grid = g.per_xcd * kXcd;
    if (grid > 0x7fffffffL) return set_error(FC3_ERR_UNSUPPORTED, "mcc3: grid too large");
    const bool amc = (sam == 1);
    const bool vec = ((amc ? sak : sam) % 4 == 0) && (R % 4 == 0) && (SS % 4 == 0) &&
                     (CS % 4 == 0) && (M % 4 == 0) && (K % 4 == 0) && (M >= 4) &&
                     ((uintptr_t)A % 16 == 0) && ((uintptr_t)C % 16 == 0);
    dim3 gd((unsigned)grid), bd(v2::NT);
    ProfScope prof(kind, s);
#define FC3_LAUNCH2(P, AM, V) hipLaunchKernelGGL((v2::k_gemm2<P, AM, V>), gd, bd, 0, s, g)
    if (prec == FC3_PREC_BF16X3) {
        if (amc) { if (vec) FC3_LAUNCH2(FC3_PREC_BF16X3, true, true); else FC3_LAUNCH2(FC3_PREC_BF16X3, true, false); }
        else     { if (vec) FC3_LAUNCH2(FC3_PREC_BF16X3, false, true); else FC3_LAUNCH2(FC3_PREC_BF16X3, false, false); }
    } else {
        if (amc) { if (vec) FC3_LAUNCH2(FC3_PREC_FP32, true, true); else FC3_LAUNCH2(FC3_PREC_FP32, true, false); }
        else     { if (vec) FC3_LAUNCH2(FC3_PREC_FP32, false, true); else FC3_LAUNCH2(FC3_PREC_FP32, false, false); }
    }
#undef FC3_LAUNCH2
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

// Slabs per mode-1/2 launch: enough workgroups to fill the chip many times over,
// with the slab scratch bounded to 1 GiB.
int slab_batch(int n)
{
    const long n2 = (long)n * n;
    const long tiles_n = (n + v2::BN - 1) / v2::BN;
    long b = (8192L * v2::BM + tiles_n * n - 1) / (tiles_n * n);
    const long cap = (1L << 30) / (4 * n2);
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    if (b > n) b = n;
    return (int)b;
}

// Modes 1 and 2 on slabs [l0, l0+cnt) of T (each n x n, stride n^2), in place:
// S_l = T_l^T phi, then T_l = S_l^T phi, each as one GEMM over the stacked slabs.
int modes12(hipStream_t s, int prec, float* T, int n, int l0, int cnt, const PhiPrep& B,
            float* S)
{
    const long n2 = (long)n * n;
    // A(m = l n + j, k = i) = T_l[i + n j] = T[m n + k];  C(m, n') -> S_l[j + n n']
    FC3_TRY(gemm2(s, FC3_KERNEL_GEMM_MODE12, prec, T + l0 * n2, n, 1, B, S, n, n2, n,
                  (long)cnt * n, n, n));
    // A(m = l n + n', k = j) = S_l[j + n n'];  C(m, m') -> T_l[n' + n m']
    FC3_TRY(gemm2(s, FC3_KERNEL_GEMM_MODE12, prec, S, n, 1, B, T + l0 * n2, n, n2, n,
                  (long)cnt * n, n, n));
    return FC3_OK;
}

}  // namespace

size_t mcc3_workspace_bytes(int n, int precision, bool inplace)
{
    const long n2 = (long)n * n;
    size_t b = phi_bytes(n, precision) + Carve::need(sizeof(float) * n2 * slab_batch(n)) + 4096;
    if (inplace) {
        // mode-3 row-block scratch: R rows x n
        const long R = 1L << 16;
        b += Carve::need(sizeof(float) * (size_t)(R < n2 ? R : n2) * n);
    }
    return b;
}

int mcc3_run(hipStream_t s, const float* psi, const float* phi, int n, int precision, float* k3,
             void* ws, size_t ws_bytes)
{
    Carve cv(ws, ws_bytes);
    PhiPrep B;
    FC3_TRY(prep_phi(s, phi, n, precision, cv, B));
    const int nb = slab_batch(n);
    const long n2 = (long)n * n;
    float* S = cv.take<float>((size_t)n2 * nb);
    // mode 3: U[(i,j), l] = sum_k Psi[(i,j), k] phi[k, l] -> K3 buffer
    FC3_TRY(gemm2(s, FC3_KERNEL_GEMM_MODE3, precision, psi, 1, n2, B, k3, n2, 0, n2, n2, n, n));
    for (int l0 = 0; l0 < n; l0 += nb) {
        const int cnt = (n - l0) < nb ? (n - l0) : nb;
        FC3_TRY(modes12(s, precision, k3, n, l0, cnt, B, S));
    }
    return FC3_OK;
}

int mcc3_run_inplace(hipStream_t s, float* psi, const float* phi, int n, int precision, void* ws,
                     size_t ws_bytes)
{
    Carve cv(ws, ws_bytes);
    PhiPrep B;
    FC3_TRY(prep_phi(s, phi, n, precision, cv, B));
    const int nb = slab_batch(n);
    const long n2 = (long)n * n;
    float* S = cv.take<float>((size_t)n2 * nb);
    for (int l0 = 0; l0 < n; l0 += nb) {
        const int cnt = (n - l0) < nb ? (n - l0) : nb;
        FC3_TRY(modes12(s, precision, psi, n, l0, cnt, B, S));
    }
    // mode 3 in place by row blocks through scratch
    const long R0 = 1L << 16;
    const long R = R0 < n2 ? R0 : n2;
    float* W = cv.take<float>((size_t)R * n);
    for (long r0 = 0; r0 < n2; r0 += R) {
        const long rows = (n2 - r0) < R ? (n2 - r0) : R;
        FC3_TRY(gemm2(s, FC3_KERNEL_GEMM_MODE3, precision, psi + r0, 1, n2, B, W, rows, 0, rows,
                      rows, n, n));
        FC3_HIP(hipMemcpy2DAsync(psi + r0, sizeof(float) * n2, W, sizeof(float) * rows,
                                 sizeof(float) * rows, n, hipMemcpyDeviceToDevice, s));
    }
    return FC3_OK;
}

}  // namespace fc3
