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

constexpr int BK = 96;                // K padding granularity of the prepared phi (6 x 16)
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
constexpr int PF = BK + 4;   // fp32 row pitch (80 B): conflict-free fragment reads

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

// BF16 operand tiles in LDS: each row (an A row or a B column) holds BK = 16 bf16,
// two 16-byte chunks, unpadded.  Blocks of 4 rows fill the 32 banks once; inside a
// block, row r sits in slot (r & 3) ^ (block & 3) and its chunks are swapped on odd
// blocks.  This makes the MFMA fragment reads (8 consecutive rows, one chunk each),
// the row-wise stores (8 rows x one chunk) and the transposing stores of an
// M-contiguous A (4 rows that are 4 apart x 4 k-quads) all bank-conflict-free.
// Returns the element offset of chunk c of row r.
__device__ __forceinline__ int swz(int r, int c)
{
    const int b = r >> 2;
    return 64 * b + 16 * ((r & 3) ^ (b & 3)) + 8 * (c ^ (b & 1));
}

// ---- global -> registers -------------------------------------------------------
// A tile (BM x BK fp32), 16 values per thread:
//   AMC (M-contiguous): thread (mq = t >> 2, kq = t & 3) holds rows 4 mq + i and
//                       k = 4 kq + kk as v[4 i + kk];
//   K-contiguous:       thread t holds row t, k = 4 q + kk as v[4 q + kk].
// FAST: the tile is inside M and K needs no tail, so no checks.  Otherwise the
// loads are clamped to valid addresses and zeroed (M % 4 == 0, K % 4 == 0 when
// VEC) or fully scalar.
template <bool AMC, bool VEC, bool FAST>
__device__ __forceinline__ void load_a2(const G2& g, int m0, int k0, float (&v)[16])
{
    const int t = threadIdx.x;
    if constexpr (AMC) {
        const int mq = t >> 2, kq = t & 3;
        const int mr = m0 + 4 * mq, kc = k0 + 4 * kq;
        if constexpr (FAST) {
            const float* p = g.A + (long)kc * g.sak + mr;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const float4 x = *reinterpret_cast<const float4*>(p + (long)kk * g.sak);
                v[0 + kk] = x.x; v[4 + kk] = x.y; v[8 + kk] = x.z; v[12 + kk] = x.w;
            }
        } else if constexpr (VEC) {
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
        if constexpr (FAST) {
            const float* p = g.A + (long)m * g.sam + k0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 x = *reinterpret_cast<const float4*>(p + 4 * q);
                v[4 * q + 0] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
            }
        } else if constexpr (VEC) {
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

// B tile of the prepared phi: 4 chunks of 16 bytes per thread in named registers
// (an indexed array would live in scratch memory and make every load wait).
// Chunk idx = t + 256 it: BF16X3 -> array idx / 512 (hi, lo), column (idx % 512) / 2,
// chunk idx % 2;  FP32 -> column idx / 4, quarter idx % 4.  Always in bounds (phi
// is zero-padded to the tile grid; k0 past the end is clamped by the caller).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct BRegs {
    u32x4 r0, r1, r2, r3;
};

template <int PREC>
__device__ __forceinline__ void load_b2(const G2& g, int n0, int k0, BRegs& b)
{
    const int t = threadIdx.x;
    if constexpr (PREC == FC3_PREC_BF16X3) {
        const int col = t >> 1, c = t & 1;
        const __bf16* ph = static_cast<const __bf16*>(g.Bhi) + (long)(n0 + col) * g.ldb + k0 + 8 * c;
        const __bf16* pl = static_cast<const __bf16*>(g.Blo) + (long)(n0 + col) * g.ldb + k0 + 8 * c;
        const long o = 128L * g.ldb;
        b.r0 = *reinterpret_cast<const u32x4*>(ph);
        b.r1 = *reinterpret_cast<const u32x4*>(ph + o);
        b.r2 = *reinterpret_cast<const u32x4*>(pl);
        b.r3 = *reinterpret_cast<const u32x4*>(pl + o);
    } else {
        const int col = t >> 2, q = t & 3;
        const float* p = static_cast<const float*>(g.Bhi) + (long)(n0 + col) * g.ldb + k0 + 4 * q;
        const long o = 64L * g.ldb;
        b.r0 = *reinterpret_cast<const u32x4*>(p);
        b.r1 = *reinterpret_cast<const u32x4*>(p + o);
        b.r2 = *reinterpret_cast<const u32x4*>(p + 2 * o);
        b.r3 = *reinterpret_cast<const u32x4*>(p + 3 * o);
    }
}

// ---- registers -> LDS (BF16X3) -------------------------------------------------
template <bool AMC>
__device__ __forceinline__ void store_a_split(__bf16* sAh, __bf16* sAl, const float (&v)[16])
{
    const int t = threadIdx.x;
    if constexpr (AMC) {
        const int mq = t >> 2, kq = t & 3;
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
            const int off = swz(4 * mq + r, kq >> 1) + 4 * (kq & 1);
            *reinterpret_cast<bf16x4*>(sAh + off) = h;
            *reinterpret_cast<bf16x4*>(sAl + off) = l;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            bf16x8 h, l;
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) {
                __bf16 hh, ll;
                split_bf16(v[8 * c + kk], hh, ll);
                h[kk] = hh;
                l[kk] = ll;
            }
            const int off = swz(t, c);
            *reinterpret_cast<bf16x8*>(sAh + off) = h;
            *reinterpret_cast<bf16x8*>(sAl + off) = l;
        }
    }
}

__device__ __forceinline__ void store_b_bf16(__bf16* sBh, __bf16* sBl, const BRegs& b)
{
    const int t = threadIdx.x;
    const int col = t >> 1, c = t & 1;
    *reinterpret_cast<u32x4*>(sBh + swz(col, c)) = b.r0;
    *reinterpret_cast<u32x4*>(sBh + swz(col + 128, c)) = b.r1;
    *reinterpret_cast<u32x4*>(sBl + swz(col, c)) = b.r2;
    *reinterpret_cast<u32x4*>(sBl + swz(col + 128, c)) = b.r3;
}

constexpr int TILE_E = BM * BK;   // bf16 elements of one operand tile (BM == BN)

// The K loop, BF16X3.  Two LDS stages and a ring of three register stages for
// the global prefetch, so a global load has three K steps to arrive.  In step k:
//   read fragments of step k from stage k % 2
//   MFMA (al bh), (ah bl)
//   store registers (data of step k+1, loaded in step k-2) -> stage (k+1) % 2
//   global load of step k+3 into the registers just freed
//   MFMA (ah bh); barrier
// The loop is unrolled by 6 so that stage and register indices are static; K is
// padded to a multiple of 6 BK (zero-padded phi, A zeroed by its bounds check),
// and prefetches past the end are clamped to the last step and never consumed.
constexpr int KSTEP_MULT = 6;

template <bool AMC, bool VEC, bool FAST>
__device__ __forceinline__ void kloop_bf16(const G2& g, int m0, int n0, __bf16 (*s)[4][TILE_E],
                                           f32x16 (&acc)[4][4])
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm0 = (wave & 1) * 128, wn0 = (wave >> 1) * 128;
    const int nk = KSTEP_MULT * ((g.K + KSTEP_MULT * BK - 1) / (KSTEP_MULT * BK));
    const int klast = (nk - 1) * BK;
    float av0[16], av1[16], av2[16];
    BRegs bv0, bv1, bv2;
    int offa[4], offb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        offa[i] = swz(wm0 + i * 32 + (lane & 31), lane >> 5);
        offb[i] = swz(wn0 + i * 32 + (lane & 31), lane >> 5);
    }
#define FC3_KN(x) ((x) < klast ? (x) : klast)
    // Issue order inside one step (one wave per SIMD issues at most one instruction
    // per slot, so MFMAs must be spread between the other instructions): the first
    // fragments, then each remaining fragment read, split VALU, LDS store and global
    // load placed between MFMAs.
#define FC3_SGB(mask, n) __builtin_amdgcn_sched_group_barrier(mask, n, 0);
#define FC3_INTERLEAVE()                                                                      \
    FC3_SGB(0x100, 6)                                                                         \
    _Pragma("unroll") for (int q = 0; q < 10; ++q) { FC3_SGB(0x008, 1) FC3_SGB(0x100, 1) }     \
    _Pragma("unroll") for (int q = 0; q < 12; ++q) {                                          \
        FC3_SGB(0x008, 1) FC3_SGB(0x002, 4) FC3_SGB(0x200, 1) }                               \
    _Pragma("unroll") for (int q = 0; q < 8; ++q) {                                           \
        FC3_SGB(0x008, 1) FC3_SGB(0x020, 1) FC3_SGB(0x002, 3) }                               \
    _Pragma("unroll") for (int q = 0; q < 18; ++q) { FC3_SGB(0x008, 1) FC3_SGB(0x002, 2) }
#define FC3_MMA(X, Y)                                                                         \
    _Pragma("unroll") for (int i = 0; i < 4; ++i)                                             \
        _Pragma("unroll") for (int j = 0; j < 4; ++j)                                         \
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(X[i], Y[j], acc[i][j], 0, 0, 0);
#define FC3_STEP(cur, avs, bvs, knext)                                                        \
    {                                                                                         \
        bf16x8 ah[4], al[4], bh[4], bl[4];                                                    \
        _Pragma("unroll") for (int i = 0; i < 4; ++i) {                                       \
            ah[i] = *reinterpret_cast<const bf16x8*>(&s[cur][0][offa[i]]);                    \
            al[i] = *reinterpret_cast<const bf16x8*>(&s[cur][1][offa[i]]);                    \
            bh[i] = *reinterpret_cast<const bf16x8*>(&s[cur][2][offb[i]]);                    \
            bl[i] = *reinterpret_cast<const bf16x8*>(&s[cur][3][offb[i]]);                    \
        }                                                                                     \
        FC3_MMA(al, bh)                                                                       \
        FC3_MMA(ah, bl)                                                                       \
        store_a_split<AMC>(s[(cur) ^ 1][0], s[(cur) ^ 1][1], avs);                           \
        store_b_bf16(s[(cur) ^ 1][2], s[(cur) ^ 1][3], bvs);                                 \
        load_a2<AMC, VEC, FAST>(g, m0, FC3_KN(knext), avs);                                  \
        load_b2<FC3_PREC_BF16X3>(g, n0, FC3_KN(knext), bvs);                                  \
        FC3_MMA(ah, bh)                                                                       \
        FC3_INTERLEAVE()                                                                      \
        __syncthreads();                                                                      \
    }
    // prologue: stage 0 <- step 0; registers: ring slot r holds step 1 + ((k - 1 - r) mod 3)
    load_a2<AMC, VEC, FAST>(g, m0, 0, av0);
    load_b2<FC3_PREC_BF16X3>(g, n0, 0, bv0);
    load_a2<AMC, VEC, FAST>(g, m0, FC3_KN(BK), av1);
    load_b2<FC3_PREC_BF16X3>(g, n0, FC3_KN(BK), bv1);
    load_a2<AMC, VEC, FAST>(g, m0, FC3_KN(2 * BK), av2);
    load_b2<FC3_PREC_BF16X3>(g, n0, FC3_KN(2 * BK), bv2);
    store_a_split<AMC>(s[0][0], s[0][1], av0);
    store_b_bf16(s[0][2], s[0][3], bv0);
    load_a2<AMC, VEC, FAST>(g, m0, FC3_KN(3 * BK), av0);
    load_b2<FC3_PREC_BF16X3>(g, n0, FC3_KN(3 * BK), bv0);
    __syncthreads();
    // now av1 = step 1, av2 = step 2, av0 = step 3
    for (int kt = 0; kt < nk; kt += KSTEP_MULT) {
        FC3_STEP(0, av1, bv1, (kt + 4) * BK)   // k = kt    : stage 1 <- k+1, fetch k+4
        FC3_STEP(1, av2, bv2, (kt + 5) * BK)   // k = kt + 1: stage 0 <- k+1 (= kt+2)
        FC3_STEP(0, av0, bv0, (kt + 6) * BK)   // k = kt + 2
        FC3_STEP(1, av1, bv1, (kt + 7) * BK)   // k = kt + 3
        FC3_STEP(0, av2, bv2, (kt + 8) * BK)   // k = kt + 4
        FC3_STEP(1, av0, bv0, (kt + 9) * BK)   // k = kt + 5
    }
#undef FC3_STEP
#undef FC3_MMA
#undef FC3_INTERLEAVE
#undef FC3_SGB
#undef FC3_KN
}

// The K loop, FP32 (same pipeline; padded row-major LDS tiles).
template <bool AMC, bool VEC, bool FAST>
__device__ __forceinline__ void kloop_fp32(const G2& g, int m0, int n0, float (*sA)[BM][PF],
                                           float (*sB)[BN][PF], f32x16 (&acc)[4][4])
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm0 = (wave & 1) * 128, wn0 = (wave >> 1) * 128;
    const int nk = 2 * ((g.K + 2 * BK - 1) / (2 * BK));
    const int klast = (nk - 1) * BK;
    float av0[16], av1[16];
    BRegs bv0, bv1;
    // MFMA step s, lane group h = lane / 32 uses k = 8 h + s (8 contiguous floats)
    const int kof = (lane >> 5) * 8;
#define FC3_KN(x) ((x) < klast ? (x) : klast)
#define FC3_STORE(st, avs, bvs)                                                               \
    {                                                                                         \
        _Pragma("unroll") for (int r = 0; r < 4; ++r) {                                       \
            const float4 x = make_float4(avs[4 * r], avs[4 * r + 1], avs[4 * r + 2], avs[4 * r + 3]); \
            if (AMC) *reinterpret_cast<float4*>(&sA[st][4 * (t >> 2) + r][4 * (t & 3)]) = x;  \
            else *reinterpret_cast<float4*>(&sA[st][t][4 * r]) = x;                           \
        }                                                                                     \
        *reinterpret_cast<u32x4*>(&sB[st][t >> 2][4 * (t & 3)]) = bvs.r0;                     \
        *reinterpret_cast<u32x4*>(&sB[st][(t >> 2) + 64][4 * (t & 3)]) = bvs.r1;              \
        *reinterpret_cast<u32x4*>(&sB[st][(t >> 2) + 128][4 * (t & 3)]) = bvs.r2;             \
        *reinterpret_cast<u32x4*>(&sB[st][(t >> 2) + 192][4 * (t & 3)]) = bvs.r3;             \
    }
#define FC3_STEPF(cur, avs, bvs, knext)                                                       \
    {                                                                                         \
        float fa[4][8], fb[4][8];                                                             \
        _Pragma("unroll") for (int i = 0; i < 4; ++i) {                                       \
            const int r = wm0 + i * 32 + (lane & 31);                                         \
            const float4 x0 = *reinterpret_cast<const float4*>(&sA[cur][r][kof]);             \
            const float4 x1 = *reinterpret_cast<const float4*>(&sA[cur][r][kof + 4]);         \
            fa[i][0] = x0.x; fa[i][1] = x0.y; fa[i][2] = x0.z; fa[i][3] = x0.w;               \
            fa[i][4] = x1.x; fa[i][5] = x1.y; fa[i][6] = x1.z; fa[i][7] = x1.w;               \
            const int c = wn0 + i * 32 + (lane & 31);                                         \
            const float4 y0 = *reinterpret_cast<const float4*>(&sB[cur][c][kof]);             \
            const float4 y1 = *reinterpret_cast<const float4*>(&sB[cur][c][kof + 4]);         \
            fb[i][0] = y0.x; fb[i][1] = y0.y; fb[i][2] = y0.z; fb[i][3] = y0.w;               \
            fb[i][4] = y1.x; fb[i][5] = y1.y; fb[i][6] = y1.z; fb[i][7] = y1.w;               \
        }                                                                                     \
        _Pragma("unroll") for (int s2 = 0; s2 < 4; ++s2)                                      \
            _Pragma("unroll") for (int i = 0; i < 4; ++i)                                     \
                _Pragma("unroll") for (int j = 0; j < 4; ++j)                                 \
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i][s2], fb[j][s2], acc[i][j], 0, 0, 0); \
        FC3_STORE((cur) ^ 1, avs, bvs)                                                        \
        load_a2<AMC, VEC, FAST>(g, m0, FC3_KN(knext), avs);                                  \
        load_b2<FC3_PREC_FP32>(g, n0, FC3_KN(knext), bvs);                                    \
        _Pragma("unroll") for (int s2 = 4; s2 < 8; ++s2)                                      \
            _Pragma("unroll") for (int i = 0; i < 4; ++i)                                     \
                _Pragma("unroll") for (int j = 0; j < 4; ++j)                                 \
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i][s2], fb[j][s2], acc[i][j], 0, 0, 0); \
        __syncthreads();                                                                      \
    }
    load_a2<AMC, VEC, FAST>(g, m0, 0, av0);
    load_b2<FC3_PREC_FP32>(g, n0, 0, bv0);
    FC3_STORE(0, av0, bv0)
    load_a2<AMC, VEC, FAST>(g, m0, FC3_KN(BK), av1);
    load_b2<FC3_PREC_FP32>(g, n0, FC3_KN(BK), bv1);
    load_a2<AMC, VEC, FAST>(g, m0, FC3_KN(2 * BK), av0);
    load_b2<FC3_PREC_FP32>(g, n0, FC3_KN(2 * BK), bv0);
    __syncthreads();
    for (int kt = 0; kt < nk; kt += 2) {
        FC3_STEPF(0, av1, bv1, (kt + 3) * BK)
        FC3_STEPF(1, av0, bv0, (kt + 4) * BK)
    }
#undef FC3_STEPF
#undef FC3_STORE
#undef FC3_KN
}

// 256 x 256 macro tile, 4 waves (2 x 2), each a 128 x 128 tile of 4 x 4 MFMA 32x32
// blocks; XCD-aware tile order.  Interior tiles (inside M, K a multiple of 2 BK)
// take the unchecked load path.
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

    if constexpr (PREC == FC3_PREC_BF16X3) {
        const bool fast = VEC && (m0 + BM <= g.M) && (g.K % (KSTEP_MULT * BK) == 0);
        __shared__ __attribute__((aligned(16))) __bf16 s[2][4][TILE_E];
        if (fast) kloop_bf16<AMC, VEC, VEC>(g, m0, n0, s, acc);
        else kloop_bf16<AMC, VEC, false>(g, m0, n0, s, acc);
    } else {
        const bool fast = VEC && (m0 + BM <= g.M) && (g.K % (2 * BK) == 0);
        __shared__ __attribute__((aligned(16))) float sA[2][BM][PF];
        __shared__ __attribute__((aligned(16))) float sB[2][BN][PF];
        if (fast) kloop_fp32<AMC, VEC, VEC>(g, m0, n0, sA, sB, acc);
        else kloop_fp32<AMC, VEC, false>(g, m0, n0, sA, sB, acc);
    }

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
