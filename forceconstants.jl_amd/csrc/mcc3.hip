// mcc3: K3[n,m,l] = sum_ijk Psi[i,j,k] phi[i,n] phi[j,m] phi[k,l]   (src/mcc.jl:15-17)
//
// Evaluated as three mode products on MFMA, ordered so that no n^3 temporary is
// needed beyond the output:
//   mode 3:     U[(i,j), l] = sum_k Psi[(i,j), k] phi[k, l]      one GEMM over M = n^2 rows
//   modes 1+2:  K3[:,:,l]  = phi^T U_l phi                       two n x n x n GEMMs per slab l,
//               S_l = U_l^T phi (scratch), K3_l = S_l^T phi      stacked over batches of slabs
// Every GEMM has phi as its B operand, so phi is prepared once.
//
// Kernel families
//   v2 (FP32):   v_mfma_f32_32x32x2_f32, FP32 operands read straight from Psi / K3 and
//                staged through padded LDS; U lives in the K3 buffer.  Also runs BF16X3
//                requests when n % 4 != 0 (FP32 is the more accurate of the two).
//   v3 (BF16X3): a*b ~= ah*bh + ah*bl + al*bh with ah = bf16_rne(a), al = bf16_rne(a - ah),
//                on v_mfma_f32_32x32x16_bf16 with FP32 accumulation.  Operands are split
//                once into K-blocked hi/lo planes (each GEMM writes its output already
//                split for the next one) and copied global -> LDS without registers.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "bf16_split.hpp"
#include "fc3_internal.hpp"

namespace fc3 {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BK = 96;                // K padding granularity of the prepared phi (6 x 16)
constexpr int kXcd = 8;               // XCDs on MI355X

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
    const float* B;      // zero-padded B(k, n) = B[n * ldb + k]
    int ldb;
    float* C;
    long R, SS, CS;      // C(m, c) = C[(m / R) * SS + m % R + c * CS]
    int M, N, K;
    int tiles_m, tiles_n;
    long per_xcd;
};

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

__device__ __forceinline__ void load_b2(const G2& g, int n0, int k0, BRegs& b)
{
    const int t = threadIdx.x;
    const int col = t >> 2, q = t & 3;
    const float* p = g.B + (long)(n0 + col) * g.ldb + k0 + 4 * q;
    const long o = 64L * g.ldb;
    b.r0 = *reinterpret_cast<const u32x4*>(p);
    b.r1 = *reinterpret_cast<const u32x4*>(p + o);
    b.r2 = *reinterpret_cast<const u32x4*>(p + 2 * o);
    b.r3 = *reinterpret_cast<const u32x4*>(p + 3 * o);
}

constexpr int TILE_E = BM * BK;   // bf16 elements of one operand tile (BM == BN)

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
        load_b2(g, n0, FC3_KN(knext), bvs);                                    \
        _Pragma("unroll") for (int s2 = 4; s2 < 8; ++s2)                                      \
            _Pragma("unroll") for (int i = 0; i < 4; ++i)                                     \
                _Pragma("unroll") for (int j = 0; j < 4; ++j)                                 \
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i][s2], fb[j][s2], acc[i][j], 0, 0, 0); \
        __syncthreads();                                                                      \
    }
    load_a2<AMC, VEC, FAST>(g, m0, 0, av0);
    load_b2(g, n0, 0, bv0);
    FC3_STORE(0, av0, bv0)
    load_a2<AMC, VEC, FAST>(g, m0, FC3_KN(BK), av1);
    load_b2(g, n0, FC3_KN(BK), bv1);
    load_a2<AMC, VEC, FAST>(g, m0, FC3_KN(2 * BK), av0);
    load_b2(g, n0, FC3_KN(2 * BK), bv0);
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
template <bool AMC, bool VEC>
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

    const bool fast = VEC && (m0 + BM <= g.M) && (g.K % (2 * BK) == 0);
    __shared__ __attribute__((aligned(16))) float sA[2][BM][PF];
    __shared__ __attribute__((aligned(16))) float sB[2][BN][PF];
    if (fast) kloop_fp32<AMC, VEC, VEC>(g, m0, n0, sA, sB, acc);
    else kloop_fp32<AMC, VEC, false>(g, m0, n0, sA, sB, acc);

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

// ---------------------------------------------------------------------------------
// GEMM v3 (BF16X3 only): both operands arrive already split into bf16 hi/lo planes in
// a K-blocked layout: K is cut into blocks of BK = 16, and inside a block the rows
// follow each other at 32 bytes (row r, k-block kb at kb * KS + 16 r).  One K step of
// a 256-row tile is then 8 KB of contiguous memory, so the K loop is global -> LDS
// copies (global_load_lds, 16 bytes per lane, each instruction one contiguous KB) with
// no registers, VALU or LDS stores in between, then fragment reads and MFMAs.  Four
// LDS stages in a ring; step k multiplies fragments read during step k-1, reads stage
// k+1, and refills its own stage with step k+4.  The epilogue writes C as FP32, or
// split again into bf16 hi/lo planes in the K-blocked layout the next mode product reads.
namespace v3 {

using v2::BM;
using v2::BN;
using v2::BK;
using v2::NT;
using v2::TILE_E;
constexpr int NSTAGE = 4;

// Operand and output addressing (element offsets; lo planes at + a_lo, b_lo, c_lo):
//   A(m, k) = A[(m / a_R) a_SS + (k / 16) a_KS + 16 (m % a_R) + k % 16]
//   B(k, c) = B[(k / 16) b_KS + 16 c + k % 16]
//   C(m, c) = C[(m / c_R) c_SS + ((m % c_R) / 16) c_KS + (m % c_R) % 16 + c c_CS]
// (c_KS = 16 makes C plain with rows contiguous inside each group of c_R rows.)
struct G3 {
    const __bf16* A;
    long a_lo, a_SS, a_KS;
    const __bf16* B;
    long b_lo, b_KS;
    void* C;
    long c_SS, c_KS, c_CS, c_lo;
    int a_R, c_R;
    int M, N, K;         // K % BK == 0, M % 4 == 0, c_R % 4 == 0
    int tiles_m, tiles_n;
    long per_xcd;
    // grid y (nz workgroup rows, z = blockIdx.y), either
    //   split K (zk = 1): z takes K blocks [z kc, min((z + 1) kc, K / BK)), or
    //   a batch (zk = 0): z multiplies A + z a_ZS by B + z b_ZS over all of K;
    // either way it writes C + z c_ZS.  An unsplit product: zk = 1, kc = K / BK, nz = 1.
    int kc, zk, nz;
    long a_ZS, b_ZS, c_ZS;
    // sparse K (k_gemm3<..., true>, unsplit products only): row tile tm multiplies only
    // the K blocks kb with kmask[tm kmask_ld + kb] != 0, in ascending order; the others
    // hold only zeros in A, so the result is the same
    const unsigned char* kmask;
    int kmask_ld;
    int kmask_div;       // row tile tm reads mask row tm / kmask_div
    // optional outer level of the A row map: A row m -> (m / a_R2) a_SS2 + row m % a_R2
    // of the two-level map above (a_R2 = 2^30 when unused)
    int a_R2;
    long a_SS2;
    // epilogue (256 x 256 tiles): 0 = each lane stores its accumulators directly; 1 = C is
    // contiguous along rows (c_KS == 16) and 2 = along columns within 16-row groups
    // (c_CS == 16): the tile goes out through LDS so that consecutive lanes store
    // consecutive addresses
    int epi;
    // optional explicit tile list: the launch computes only the tlist_n tiles
    // tlist[t] = (tm << 8) | tn, in the list's order (instead of every tile in the grouped
    // order); a tile's result does not depend on which other tiles a launch holds
    const int* tlist;
    int tlist_n;
    // sparse K: a row tile without K blocks writes nothing (its output is all zero, and
    // the caller guarantees that no consumer reads it)
    int skip_empty;
    // first-round phase offsets: the first 256 workgroups (in dispatch order, one per CU)
    // wait ((index / 8) % kStaggerGroups) x `stagger` ticks of the 100 MHz wall clock
    // before their first tile, so that the CUs do not all reach their 256 KB epilogue
    // stores at the same moment (0 = no wait)
    int stagger;
};
constexpr int kStaggerGroups = 4;
// Short-K launches (K <= 512, at least 4096 workgroups) start staggered by 8 us steps:
// measured on M = 373248, N = 2592 random operands, K = 240 ran 8-12% faster, K = 480
// 1%, while K = 2592 lost 0.5% (its epilogue is 4% of a tile, and the first round's
// wait is not won back).
constexpr int kStaggerTicks = 800;

// row tiles walked together, column by column, in the grouped tile order
constexpr int kGM = 3;

// K blocks a sparse-K tile can list (K <= 65536)
constexpr int kMaxKBlocks = 4096;

// Workgroup barrier for the copy ring: s_waitcnt vmcnt(N) lgkmcnt(0) (gfx9 encoding,
// expcnt unconstrained), then s_barrier.  After it, every wave's copies older than its
// last N are in LDS and every wave's fragment reads issued before it have returned.
// The signal fences keep the compiler from moving memory accesses across; unlike
// __syncthreads() there is no release fence, which would wait for all copies in flight.
template <int N>
__device__ __forceinline__ void ring_barrier()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_sched_barrier(0);   // nothing of the step is scheduled past the wait
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt((N & 0xf) | ((N >> 4) << 14) | (0x7 << 4));
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// s_waitcnt vmcnt(0), other counters unconstrained
__device__ __forceinline__ void wait_vmcnt0() { __builtin_amdgcn_s_waitcnt((0x7 << 4) | (0xf << 8)); }

// Operand tile layout in LDS: row r (an A row or a B column) holds BK = 16 bf16 as two
// 16-byte chunks c = 0, 1; chunk (r, c) sits at 16-byte slot
//   32 (r >> 4) + 16 c + (r & 15).
// A 64-lane ds_read_b128 is served 16 lanes per cycle from 64 four-byte banks; the
// fragment reads give each group of 16 lanes 16 consecutive rows of one chunk, which
// land in 16 distinct slots modulo 16, i.e. on all 64 banks once.  The copies write
// 64 consecutive slots per instruction.
__device__ __forceinline__ int lds_chunk(int r, int c) { return 32 * (r >> 4) + 16 * c + (r & 15); }

__device__ __forceinline__ void lds_copy16(const __bf16* g, __bf16* lds)
{
    __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(g), reinterpret_cast<void*>(lds),
                                     16, 0, 0);
}

// Tile shape: a BM x BN macro tile over 4 waves arranged WM x WN, each wave TM x TN MFMA
// 32x32 blocks: 256 x 256 as 2 x 2 waves of 4 x 4 blocks.  (A 256 x 288 shape, 4 x 1
// waves of 2 x 9 blocks, would fit n = 2592 = 9 x 288 without an almost empty last
// column tile, but ran 20% slower: every wave then reads all B fragments.)
// Narrow shapes 1-3 (256 x 32, 64, 128 as 4 x 1 waves of 2 x TN blocks) take the last
// N % 256 columns as the last tile column of the launch, so that no 256-wide tile is
// spent on them.
template <int SHAPE>
struct Shape;
template <>
struct Shape<0> {
    static constexpr int WM = 2, WN = 2, TM = 4, TN = 4;
};
template <>
struct Shape<1> {
    static constexpr int WM = 4, WN = 1, TM = 2, TN = 1;
};
template <>
struct Shape<2> {
    static constexpr int WM = 4, WN = 1, TM = 2, TN = 2;
};
template <>
struct Shape<3> {
    static constexpr int WM = 4, WN = 1, TM = 2, TN = 4;
};

// One tile of shape SHAPE at (m0, n0) on the four LDS stage objects of the kernel
// (passed in, so that after inlining every access still names its own object).
// Four C values of rows row..row + 3 (one 16-row group, row % 4 == 0) and column col,
// at the address of the layout in g, as FP32 or split into bf16 hi/lo planes.
template <bool SPLIT_OUT>
__device__ __forceinline__ void store4(const G3& g, int kz, int row, int col, float x0, float x1,
                                       float x2, float x3)
{
    const unsigned sl = (unsigned)row / (unsigned)g.c_R;
    const int rr = row - (int)sl * g.c_R;   // rr % 4 == 0: the 4 rows share a K block
    const long off = (long)sl * g.c_SS + (long)(rr >> 4) * g.c_KS + (rr & 15) + (long)col * g.c_CS;
    if constexpr (SPLIT_OUT) {
        __bf16* C = static_cast<__bf16*>(g.C) + kz * g.c_ZS;
        bf16x4 h, l;
        __bf16 hh, ll;
        split_bf16(x0, hh, ll); h[0] = hh; l[0] = ll;
        split_bf16(x1, hh, ll); h[1] = hh; l[1] = ll;
        split_bf16(x2, hh, ll); h[2] = hh; l[2] = ll;
        split_bf16(x3, hh, ll); h[3] = hh; l[3] = ll;
        *reinterpret_cast<bf16x4*>(C + off) = h;
        *reinterpret_cast<bf16x4*>(C + off + g.c_lo) = l;
    } else {
        float* C = static_cast<float*>(g.C) + kz * g.c_ZS;
        *reinterpret_cast<float4*>(C + off) = make_float4(x0, x1, x2, x3);
    }
}

// LDS visibility of this wave's stores to its other lanes (and the order of the wave's
// LDS accesses around it)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The 128 x 128 wave tile of a 256 x 256 tile, written through the wave's own LDS stage
// buffer `st` (>= 18 KB; the stages are free once every wave's copies have landed and
// its fragment reads are done, which the barrier at the top ensures).  Each element goes
// to the address the direct epilogue would give it (store4); only the lane that stores
// it changes, so that a wave's store covers contiguous memory:
//   epi 1 (rows contiguous): 32-column strips; lane = (column pair, 4-row chunk of 128),
//          one instruction = 2 columns x 128 rows;
//   epi 2 (columns contiguous in 16-row groups): 32-row strips; lane = (column, 4-row
//          chunk of a 16-row group), one instruction = 16 columns x 16 rows.
template <bool SPLIT_OUT>
__device__ __forceinline__ void staged_epilogue(const G3& g, int m0, int n0, int wm0, int wn0,
                                                int kz, const f32x16 (&acc)[4][4], float* st)
{
    const int lane = threadIdx.x & 63;
    __syncthreads();
    if (g.epi == 1) {
        constexpr int PR = 132;   // [column][row] pitch in floats: 128 rows + 4
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = 32 * i + 8 * q + 4 * (lane >> 5), c = lane & 31;
                    *reinterpret_cast<float4*>(st + c * PR + r) =
                        make_float4(acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2],
                                    acc[i][j][4 * q + 3]);
                }
            wave_lds_sync();
#pragma unroll 1
            for (int t = 0; t < 16; ++t) {
                const int c = 2 * t + (lane >> 5), r = 4 * (lane & 31);
                const float4 v = *reinterpret_cast<const float4*>(st + c * PR + r);
                const int row = m0 + wm0 + r, col = n0 + wn0 + 32 * j + c;
                if (row < g.M && col < g.N) store4<SPLIT_OUT>(g, kz, row, col, v.x, v.y, v.z, v.w);
            }
            wave_lds_sync();
        }
    } else {
        constexpr int PC = 36;    // [column][row] pitch in floats: 32 rows + 4
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = 8 * q + 4 * (lane >> 5), c = 32 * j + (lane & 31);
                    *reinterpret_cast<float4*>(st + c * PC + r) =
                        make_float4(acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2],
                                    acc[i][j][4 * q + 3]);
                }
            wave_lds_sync();
#pragma unroll 1
            for (int t = 0; t < 16; ++t) {
                const int c = 16 * (t & 7) + (lane >> 2), r = 16 * (t >> 3) + 4 * (lane & 3);
                const float4 v = *reinterpret_cast<const float4*>(st + c * PC + r);
                const int row = m0 + wm0 + 32 * i + r, col = n0 + wn0 + c;
                if (row < g.M && col < g.N) store4<SPLIT_OUT>(g, kz, row, col, v.x, v.y, v.z, v.w);
            }
            wave_lds_sync();
        }
    }
}

// SPK: the K steps are the nk_list K blocks klist[0..nk_list) (sparse K), not 0..K/16.
template <bool SPLIT_OUT, int SHAPE, bool SPK = false>
__device__ __forceinline__ void gemm3_tile(const G3& g, const int m0, const int n0, __bf16* s0,
                                           __bf16* s1, __bf16* s2, __bf16* s3,
                                           const unsigned short* klist = nullptr, int nk_list = 0)
{
    using SH = Shape<SHAPE>;
    constexpr int TBM = SH::WM * SH::TM * 32, TBN = SH::WN * SH::TN * 32;
    static_assert(TBM == BM, "A tile is 256 rows");
    constexpr int AE = TBM * BK, BE = TBN * BK;        // bf16 elements per plane
    constexpr int CA = TBM / 32, CB = TBN / 32;        // 64-chunk copies per plane
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm0 = (wave % SH::WM) * (SH::TM * 32), wn0 = (wave / SH::WM) * (SH::TN * 32);

    // Copy plan: wave w fills plane w (0 A hi, 1 A lo, 2 B hi, 3 B lo) of a stage with
    // CA or CB copies of 64 chunks; chunk P = 64 q + lane of the plane (layout:
    // lds_chunk) holds row r = 16 (P >> 5) + (P & 15), k-half c = (P >> 4) & 1.  Copy q
    // covers rows 32 q .. 32 q + 31, one contiguous KB of the K block (one slab permitting).
    constexpr int CMAX = CA > CB ? CA : CB;
    // wave-uniform (an SGPR value), so that branches on it are scalar branches
    const bool is_a = __builtin_amdgcn_readfirstlane(wave) < 2;
    const int ncopy = is_a ? CA : CB;
    const __bf16* src[CMAX];
#pragma unroll
    for (int q = 0; q < CMAX; ++q) {
        const int P = 64 * q + lane;
        const int r = 16 * (P >> 5) + (P & 15);
        const int c = (P >> 4) & 1;
        if (is_a) {
            int m = m0 + (q < CA ? r : 0);
            if (m >= g.M) m = g.M - 1;        // rows past M feed only discarded rows of C
            const unsigned g2 = (unsigned)m / (unsigned)g.a_R2;
            const int m2 = m - (int)g2 * g.a_R2;
            const unsigned sl = (unsigned)m2 / (unsigned)g.a_R;
            const long arow = (long)g2 * g.a_SS2 + (long)sl * g.a_SS + 16L * (m2 - (int)sl * g.a_R);
            src[q] = g.A + arow + (wave ? g.a_lo : 0) + 8 * c;
        } else {
            src[q] = g.B + (wave == 3 ? g.b_lo : 0) + 16L * (n0 + (q < CB ? r : 0)) + 8 * c;
        }
    }
    const long kstride = is_a ? g.a_KS : g.b_KS;   // elements per K block
    const int kz = blockIdx.y, kbeg = g.zk ? kz * g.kc : 0;
    // >= 1: the host launches no empty chunk; a sparse-K tile may have none (zero output)
    const int nk = SPK ? nk_list : g.zk ? min(g.kc, g.K / BK - kbeg) : g.K / BK;
    const long zoff = kbeg * kstride + (long)kz * (is_a ? g.a_ZS : g.b_ZS);
#pragma unroll
    for (int q = 0; q < CMAX; ++q) src[q] += zoff;
    const int wofs = wave == 0 ? 0 : wave == 1 ? AE : wave == 2 ? 2 * AE : 2 * AE + BE;
    // this wave's copies of step k into stage array `st`; steps past the end copy the
    // last step again (into a stage whose fragments are read but not used), so that
    // every step has the same number of copies in flight
    auto issue = [&](__bf16* st, int k) {
        const int kc = k < nk ? k : nk - 1;
        const long kk = (SPK ? (long)klist[kc] : (long)kc) * kstride;
#pragma unroll
        for (int q = 0; q < CMAX; ++q)
            if (q < ncopy) lds_copy16(src[q] + kk, st + wofs + 512 * q);
    };

    f32x16 acc[SH::TM][SH::TN];
#pragma unroll
    for (int i = 0; i < SH::TM; ++i)
#pragma unroll
        for (int j = 0; j < SH::TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    int offa[SH::TM], offb[SH::TN];
#pragma unroll
    for (int i = 0; i < SH::TM; ++i) offa[i] = 8 * lds_chunk(wm0 + i * 32 + (lane & 31), lane >> 5);
#pragma unroll
    for (int j = 0; j < SH::TN; ++j) offb[j] = 8 * lds_chunk(wn0 + j * 32 + (lane & 31), lane >> 5);

#define FC3_MMA3(X, Y)                                                                      \
    _Pragma("unroll") for (int i = 0; i < SH::TM; ++i)                                      \
        _Pragma("unroll") for (int j = 0; j < SH::TN; ++j)                                  \
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(X[i], Y[j], acc[i][j], 0, 0, 0);
#define FC3_RING_BARRIER(D)                                                                 \
    {                                                                                       \
        if constexpr (CA == CB) ring_barrier<(D) * CA>();                                   \
        else if (is_a) ring_barrier<(D) * CA>();                                            \
        else ring_barrier<(D) * CB>();                                                      \
    }
    // Fragments of one stage (hi and lo of this wave's A rows and B columns), in two
    // register sets: step k multiplies set k % 2 while set (k + 1) % 2 is read from the
    // next stage, so no MFMA waits for LDS.
    bf16x8 ah0[SH::TM], al0[SH::TM], bh0[SH::TN], bl0[SH::TN];
    bf16x8 ah1[SH::TM], al1[SH::TM], bh1[SH::TN], bl1[SH::TN];
#define FC3_READ(ST, AH, AL, BH, BL)                                                        \
    {                                                                                       \
        _Pragma("unroll") for (int i = 0; i < SH::TM; ++i) {                                \
            AH[i] = *reinterpret_cast<const bf16x8*>(ST + offa[i]);                         \
            AL[i] = *reinterpret_cast<const bf16x8*>(ST + AE + offa[i]);                    \
        }                                                                                   \
        _Pragma("unroll") for (int j = 0; j < SH::TN; ++j) {                                \
            BH[j] = *reinterpret_cast<const bf16x8*>(ST + 2 * AE + offb[j]);                \
            BL[j] = *reinterpret_cast<const bf16x8*>(ST + 2 * AE + BE + offb[j]);           \
        }                                                                                   \
    }
    // Scheduling hint for one step: the fragment reads go out two per MFMA from the top of
    // the step (left alone, the scheduler sinks them to the barrier, where their latency
    // is exposed), then the copies one per two MFMAs (when every wave issues the same
    // number, so that they are not behind a branch), then the remaining MFMAs.
#define FC3_ORDER()                                                                         \
    {                                                                                       \
        constexpr int NREAD = 2 * (SH::TM + SH::TN), NMMA = 3 * SH::TM * SH::TN;            \
        constexpr int NCP = CA == CB ? CA : 0;                                              \
        static_assert(NREAD / 2 + 2 * NCP <= NMMA, "step schedule");                        \
        _Pragma("unroll") for (int r = 0; r < NREAD / 2; ++r) {                             \
            __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);                              \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                              \
        }                                                                                   \
        _Pragma("unroll") for (int r = 0; r < NCP; ++r) {                                   \
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                              \
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);                              \
        }                                                                                   \
        __builtin_amdgcn_sched_group_barrier(0x008, NMMA - NREAD / 2 - 2 * NCP, 0);         \
    }
    // Step K (stage K % 4, fragments already in registers): read stage K + 1 into the
    // other register set, multiply, and refill stage K % 4 (whose reads every wave
    // finished before the barrier that ended step K - 1) with step K + 4.  The closing
    // barrier waits for the fragment reads and leaves the copies of steps K + 3 and
    // K + 4 in flight, so stage K + 2 has landed for the next step's reads.
#define FC3_STEP4(K, SNX, SCUR, AH, AL, BH, BL, NAH, NAL, NBH, NBL)                         \
    {                                                                                       \
        FC3_READ(SNX, NAH, NAL, NBH, NBL)                                                   \
        FC3_MMA3(AL, BH)                                                                    \
        issue(SCUR, (K) + 4);                                                               \
        FC3_MMA3(AH, BL)                                                                    \
        FC3_MMA3(AH, BH)                                                                    \
        FC3_ORDER()                                                                         \
        FC3_RING_BARRIER(2)                                                                 \
    }
#define FC3_EVEN(K, SNX, SCUR) FC3_STEP4(K, SNX, SCUR, ah0, al0, bh0, bl0, ah1, al1, bh1, bl1)
#define FC3_ODD(K, SNX, SCUR) FC3_STEP4(K, SNX, SCUR, ah1, al1, bh1, bl1, ah0, al0, bh0, bl0)

    // prologue: steps 0..3 in flight; once step 0 has landed its fragments are read, and
    // the second barrier retires those reads (stage 0 may be refilled) with step 1 landed
    // (a sparse-K tile without K blocks skips the loop: its output is zero)
    if (!SPK || nk > 0) {
        issue(s0, 0);
        issue(s1, 1);
        issue(s2, 2);
        issue(s3, 3);
        FC3_RING_BARRIER(3)
        FC3_READ(s0, ah0, al0, bh0, bl0)
        FC3_RING_BARRIER(2)
        int k = 0;
        for (; k + 4 <= nk; k += 4) {
            FC3_EVEN(k, s1, s0)
            FC3_ODD(k + 1, s2, s1)
            FC3_EVEN(k + 2, s3, s2)
            FC3_ODD(k + 3, s0, s3)
        }
        if (k < nk) FC3_EVEN(k, s1, s0)
        if (k + 1 < nk) FC3_ODD(k + 1, s2, s1)
        if (k + 2 < nk) FC3_EVEN(k + 2, s3, s2)
#undef FC3_EVEN
#undef FC3_ODD
#undef FC3_STEP4
#undef FC3_ORDER
#undef FC3_READ
#undef FC3_RING_BARRIER
#undef FC3_MMA3
        wait_vmcnt0();   // no copy may still be landing when the workgroup's LDS is released

    }
    // epilogue: register r of lane l holds row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
    if constexpr (SHAPE == 0) {
        if (g.epi != 0) {
            staged_epilogue<SPLIT_OUT>(g, m0, n0, wm0, wn0, kz, acc,
                                       reinterpret_cast<float*>(wave == 0   ? s0
                                                                : wave == 1 ? s1
                                                                : wave == 2 ? s2
                                                                            : s3));
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < SH::TM; ++i)
#pragma unroll
        for (int j = 0; j < SH::TN; ++j) {
            const int col = n0 + wn0 + j * 32 + (lane & 31);
            if (col >= g.N) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = m0 + wm0 + i * 32 + q * 8 + (lane >> 5) * 4;
                if (row >= g.M) continue;   // M % 4 == 0: the 4 rows are all valid
                const unsigned sl = (unsigned)row / (unsigned)g.c_R;
                const int rr = row - (int)sl * g.c_R;   // rr % 4 == 0: the 4 rows share a K block
                const long off = (long)sl * g.c_SS + (long)(rr >> 4) * g.c_KS + (rr & 15) +
                                 (long)col * g.c_CS;
                const float x0 = acc[i][j][4 * q], x1 = acc[i][j][4 * q + 1];
                const float x2 = acc[i][j][4 * q + 2], x3 = acc[i][j][4 * q + 3];
                if constexpr (SPLIT_OUT) {
                    __bf16* C = static_cast<__bf16*>(g.C) + kz * g.c_ZS;
                    bf16x4 h, l;
                    __bf16 hh, ll;
                    split_bf16(x0, hh, ll); h[0] = hh; l[0] = ll;
                    split_bf16(x1, hh, ll); h[1] = hh; l[1] = ll;
                    split_bf16(x2, hh, ll); h[2] = hh; l[2] = ll;
                    split_bf16(x3, hh, ll); h[3] = hh; l[3] = ll;
                    *reinterpret_cast<bf16x4*>(C + off) = h;
                    *reinterpret_cast<bf16x4*>(C + off + g.c_lo) = l;
                } else {
                    float* C = static_cast<float*>(g.C) + kz * g.c_ZS;
                    *reinterpret_cast<float4*>(C + off) = make_float4(x0, x1, x2, x3);
                }
            }
        }
}

// One tile with an FP32 A operand split in registers (k_gemm3_af32):
// A(m, k) = Af[m + k a_KS], FP32 and column-major, so rows are contiguous for each k.
// The rows k < K must all exist and be finite, since the B rows past the contraction
// length are zero.
// Copies: a K block's 16 k-rows of the 256-row tile are 16 contiguous KB runs. Wave 0
// copies k-rows 0-7 and wave 1 k-rows 8-15, into the bytes that gemm3_tile's bf16 hi and
// lo planes take: k-row k at float 256 k of the stage. B is copied as in gemm3_tile.
// A fragments: each lane's fragment of block i is row wm0 + 32 i + lane % 32, k-rows
// 8 (lane / 32) .. + 7.
//   - It is read as 8 floats during the step before the one that uses it.
//   - It is split into bf16 hi / lo (split_bf16, as the split-transpose kernels write
//     them) at the top of its own step, after the barrier that retired the reads.
// The MFMAs therefore see the operands of the pre-split path, and C is the same bit for
// bit.
template <int SHAPE>
__device__ __forceinline__ void gemm3_tile_af32(const G3& g, const int m0, const int n0, __bf16* s0,
                                                __bf16* s1, __bf16* s2, __bf16* s3)
{
    using SH = Shape<SHAPE>;
    constexpr int TBM = SH::WM * SH::TM * 32, TBN = SH::WN * SH::TN * 32;
    static_assert(TBM == BM, "A tile is 256 rows");
    constexpr int AE = TBM * BK, BE = TBN * BK;        // bf16 elements per plane
    constexpr int CA = TBM / 32, CB = TBN / 32;        // copies per wave: A k-rows, B chunks
    static_assert(CA == 8, "8 k-rows per A wave");
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm0 = (wave % SH::WM) * (SH::TM * 32), wn0 = (wave / SH::WM) * (SH::TN * 32);
    constexpr int CMAX = CA > CB ? CA : CB;
    const bool is_a = __builtin_amdgcn_readfirstlane(wave) < 2;
    const int ncopy = is_a ? CA : CB;
    const __bf16* src[CMAX];
#pragma unroll
    for (int q = 0; q < CMAX; ++q) {
        if (is_a) {
            int m = m0 + 4 * lane;
            if (m > g.M - 4) m = g.M - 4;   // rows past M feed only discarded rows of C
            src[q] = reinterpret_cast<const __bf16*>(reinterpret_cast<const float*>(g.A) + m +
                                                     (long)(8 * wave + q) * g.a_KS);
        } else {
            const int P = 64 * q + lane;
            const int r = 16 * (P >> 5) + (P & 15);
            const int c = (P >> 4) & 1;
            src[q] = g.B + (wave == 3 ? g.b_lo : 0) + 16L * (n0 + (q < CB ? r : 0)) + 8 * c;
        }
    }
    // bf16 elements per K block: A 16 k-rows of floats, B as in gemm3_tile
    const long kstride = is_a ? 32L * g.a_KS : g.b_KS;
    const int nk = g.K / BK;
    const int wofs = wave == 0 ? 0 : wave == 1 ? AE : wave == 2 ? 2 * AE : 2 * AE + BE;
    auto issue = [&](__bf16* st, int k) {
        const long kk = (long)(k < nk ? k : nk - 1) * kstride;
#pragma unroll
        for (int q = 0; q < CMAX; ++q)
            if (q < ncopy) lds_copy16(src[q] + kk, st + wofs + 512 * q);
    };

    f32x16 acc[SH::TM][SH::TN];
#pragma unroll
    for (int i = 0; i < SH::TM; ++i)
#pragma unroll
        for (int j = 0; j < SH::TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    int offb[SH::TN];
#pragma unroll
    for (int j = 0; j < SH::TN; ++j) offb[j] = 8 * lds_chunk(wn0 + j * 32 + (lane & 31), lane >> 5);
    // float offset of this lane's A fragment row in a stage (k-row 8 (lane / 32) + u at
    // + 256 u)
    const int offa = 256 * 8 * (lane >> 5) + wm0 + (lane & 31);

    float xa[SH::TM][8];   // raw A fragments of the next stage
    bf16x8 ah[SH::TM], al[SH::TM];
    bf16x8 bh0[SH::TN], bl0[SH::TN], bh1[SH::TN], bl1[SH::TN];
#define FC3_MMA3(X, Y)                                                                      \
    _Pragma("unroll") for (int i = 0; i < SH::TM; ++i)                                      \
        _Pragma("unroll") for (int j = 0; j < SH::TN; ++j)                                  \
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(X[i], Y[j], acc[i][j], 0, 0, 0);
#define FC3_RING_BARRIER(D)                                                                 \
    {                                                                                       \
        if constexpr (CA == CB) ring_barrier<(D) * CA>();                                   \
        else if (is_a) ring_barrier<(D) * CA>();                                            \
        else ring_barrier<(D) * CB>();                                                      \
    }
#define FC3_READ_NEXT(ST, BH, BL)                                                           \
    {                                                                                       \
        const float* sf_ = reinterpret_cast<const float*>(ST) + offa;                       \
        _Pragma("unroll") for (int i = 0; i < SH::TM; ++i)                                  \
            _Pragma("unroll") for (int u = 0; u < 8; ++u) xa[i][u] = sf_[256 * u + 32 * i]; \
        _Pragma("unroll") for (int j = 0; j < SH::TN; ++j) {                                \
            BH[j] = *reinterpret_cast<const bf16x8*>(ST + 2 * AE + offb[j]);                \
            BL[j] = *reinterpret_cast<const bf16x8*>(ST + 2 * AE + BE + offb[j]);           \
        }                                                                                   \
    }
    // split_bf16 on pairs (one packed conversion per 2 values; round to nearest even as
    // the scalar conversion)
#define FC3_SPLIT()                                                                         \
    {                                                                                       \
        typedef float f2v_ __attribute__((ext_vector_type(2)));                             \
        typedef __bf16 b2v_ __attribute__((ext_vector_type(2)));                            \
        _Pragma("unroll") for (int i = 0; i < SH::TM; ++i)                                  \
            _Pragma("unroll") for (int u = 0; u < 8; u += 2) {                              \
                const f2v_ x_ = {xa[i][u], xa[i][u + 1]};                                   \
                const b2v_ h_ = __builtin_convertvector(x_, b2v_);                          \
                const f2v_ hf_ = __builtin_convertvector(h_, f2v_);                         \
                const f2v_ d_ = {x_[0] - hf_[0], x_[1] - hf_[1]};                           \
                const b2v_ l_ = __builtin_convertvector(d_, b2v_);                          \
                ah[i][u] = h_[0];                                                           \
                ah[i][u + 1] = h_[1];                                                       \
                al[i][u] = l_[0];                                                           \
                al[i][u + 1] = l_[1];                                                       \
            }                                                                               \
    }
    // Scheduling hint for one step: block 0's split first, then each block's first four
    // MFMAs (lo x B hi) with the next block's split, the reads of the next stage two per
    // MFMA, the copies one per two MFMAs, then the rest.
#define FC3_ORDER_AF32()                                                                    \
    {                                                                                       \
        constexpr int NV = 24, NMMA = 3 * SH::TM * SH::TN;                                  \
        constexpr int NRD = SH::TM * 4 + 2 * SH::TN;                                        \
        constexpr int NCP = CA == CB ? CA : 0;                                              \
        __builtin_amdgcn_sched_group_barrier(0x002, NV, 0);                                 \
        _Pragma("unroll") for (int i = 1; i < SH::TM; ++i) {                                \
            __builtin_amdgcn_sched_group_barrier(0x008, SH::TN, 0);                         \
            __builtin_amdgcn_sched_group_barrier(0x002, NV, 0);                             \
        }                                                                                   \
        __builtin_amdgcn_sched_group_barrier(0x008, SH::TN, 0);                             \
        _Pragma("unroll") for (int r = 0; r < NRD / 2; ++r) {                               \
            __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);                              \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                              \
        }                                                                                   \
        _Pragma("unroll") for (int r = 0; r < NCP; ++r) {                                   \
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                              \
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);                              \
        }                                                                                   \
        __builtin_amdgcn_sched_group_barrier(0x008, NMMA - SH::TM * SH::TN - NRD / 2 - 2 * NCP, 0); \
    }
    // Step K (stage K % 4 read into xa / B set during step K - 1): split it, read stage
    // K + 1, multiply, refill stage K % 4 with step K + 4; the barrier leaves the copies
    // of steps K + 3 and K + 4 in flight.
#define FC3_STEP4(K, SNX, SCUR, BH, BL, NBH, NBL)                                           \
    {                                                                                       \
        FC3_SPLIT()                                                                         \
        FC3_READ_NEXT(SNX, NBH, NBL)                                                        \
        FC3_MMA3(al, BH)                                                                    \
        issue(SCUR, (K) + 4);                                                               \
        FC3_MMA3(ah, BL)                                                                    \
        FC3_MMA3(ah, BH)                                                                    \
        FC3_ORDER_AF32()                                                                    \
        FC3_RING_BARRIER(2)                                                                 \
    }
#define FC3_EVEN(K, SNX, SCUR) FC3_STEP4(K, SNX, SCUR, bh0, bl0, bh1, bl1)
#define FC3_ODD(K, SNX, SCUR) FC3_STEP4(K, SNX, SCUR, bh1, bl1, bh0, bl0)
    {
        issue(s0, 0);
        issue(s1, 1);
        issue(s2, 2);
        issue(s3, 3);
        FC3_RING_BARRIER(3)
        FC3_READ_NEXT(s0, bh0, bl0)
        FC3_RING_BARRIER(2)
        int k = 0;
        for (; k + 4 <= nk; k += 4) {
            FC3_EVEN(k, s1, s0)
            FC3_ODD(k + 1, s2, s1)
            FC3_EVEN(k + 2, s3, s2)
            FC3_ODD(k + 3, s0, s3)
        }
        if (k < nk) FC3_EVEN(k, s1, s0)
        if (k + 1 < nk) FC3_ODD(k + 1, s2, s1)
        if (k + 2 < nk) FC3_EVEN(k + 2, s3, s2)
#undef FC3_EVEN
#undef FC3_ODD
#undef FC3_STEP4
#undef FC3_SPLIT
#undef FC3_ORDER_AF32
#undef FC3_READ_NEXT
#undef FC3_RING_BARRIER
#undef FC3_MMA3
        wait_vmcnt0();   // no copy may still be landing when the workgroup's LDS is released
    }
    if constexpr (SHAPE == 0) {
        if (g.epi != 0) {
            staged_epilogue<false>(g, m0, n0, wm0, wn0, 0, acc,
                                   reinterpret_cast<float*>(wave == 0   ? s0
                                                            : wave == 1 ? s1
                                                            : wave == 2 ? s2
                                                                        : s3));
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < SH::TM; ++i)
#pragma unroll
        for (int j = 0; j < SH::TN; ++j) {
            const int col = n0 + wn0 + j * 32 + (lane & 31);
            if (col >= g.N) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = m0 + wm0 + i * 32 + q * 8 + (lane >> 5) * 4;
                if (row >= g.M) continue;
                store4<false>(g, 0, row, col, acc[i][j][4 * q], acc[i][j][4 * q + 1],
                              acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]);
            }
        }
}

// Elements of one LDS stage of the widest (256 x 256) tile: A hi, A lo, B hi, B lo.
constexpr int kStageE = 2 * BM * BK + 2 * 256 * BK;

// The GEMM over g's M x N in 256 x 256 tiles; with NARROW = 1..3 the last tile column is
// a 256 x 32/64/128 tile of that shape (the N % 256 columns), in the same launch, so its
// A panel is read while the neighbouring 256-wide tiles of its row group hold it in the
// caches.
template <bool SPLIT_OUT, int NARROW, bool SPK = false>
__global__ void __launch_bounds__(NT, 1) k_gemm3(G3 g)
{
    // Tile order: each XCD takes a contiguous range of tiles (workgroups are dealt to the
    // 8 XCDs round robin), walked in groups of GM row tiles column by column, so that the
    // tiles in flight on one XCD share A and B panels.  Measured at n = 2592 (11 tile
    // columns, 32 CUs per XCD), random operands: GM = 3 670 ms per mcc3, 2 675, 1 678,
    // 4 678, 8 686, 16 732, 32 760; the LJ workload, same GPU: GM = 3 625 ms, 8 630 ms.
    const long wg = blockIdx.x;
    const long tile = (wg % kXcd) * g.per_xcd + wg / kXcd;
    if (g.stagger > 0) {
        const long lin = (long)blockIdx.y * gridDim.x + wg;   // dispatch order
        if (lin < 256) {
            const long wait = ((lin / kXcd) % kStaggerGroups) * g.stagger;   // bounded
            const long t0 = wall_clock64();
            while (wall_clock64() - t0 < wait) __builtin_amdgcn_s_sleep(2);
        }
    }
    int tm, tn;
    if (g.tlist) {
        if (tile >= g.tlist_n) return;
        const int v = g.tlist[tile];
        tm = v >> 8;
        tn = v & 255;
    } else {
        if (tile >= (long)g.tiles_m * g.tiles_n) return;
        const long grp = tile / (kGM * (long)g.tiles_n), first = grp * kGM;
        const int gsz = (int)((g.tiles_m - first) < kGM ? (g.tiles_m - first) : kGM);
        const long w = tile - grp * kGM * (long)g.tiles_n;
        tm = (int)(first + w % gsz);
        tn = (int)(w / gsz);
    }
    const int m0 = tm * BM, n0 = tn * 256;
    // One LDS object per stage: the compiler then sees that the fragment reads of one
    // stage do not alias the copies in flight into the others, and adds no wait.
    __shared__ __attribute__((aligned(16))) __bf16 s0[kStageE];
    __shared__ __attribute__((aligned(16))) __bf16 s1[kStageE];
    __shared__ __attribute__((aligned(16))) __bf16 s2[kStageE];
    __shared__ __attribute__((aligned(16))) __bf16 s3[kStageE];
    if constexpr (SPK) {
        // this row tile's nonzero K blocks, in ascending order (wave 0 compacts the mask)
        __shared__ unsigned short klist[kMaxKBlocks];
        __shared__ int kcount;
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            // unsplit products: the row tile's mask row; a batch (zk = 0): product z's row
            const long mrow = g.zk ? (long)(tm / g.kmask_div) : (long)blockIdx.y;
            const unsigned char* mk = g.kmask + mrow * g.kmask_ld;
            int c = 0;
            for (int b0 = 0; b0 < g.kmask_ld; b0 += 64) {
                const int kb = b0 + lane;
                const bool f = kb < g.kmask_ld && mk[kb] != 0;
                const unsigned long long m = __ballot(f);
                if (f) klist[c + __popcll(m & ((1ull << lane) - 1ull))] = (unsigned short)kb;
                c += __popcll(m);
            }
            if (lane == 0) kcount = c;
        }
        __syncthreads();
        const int nkl = kcount;
        if (nkl == 0 && g.skip_empty) return;   // uniform: every thread read kcount
        if constexpr (NARROW != 0) {
            if (tn == g.tiles_n - 1) {
                gemm3_tile<SPLIT_OUT, NARROW, true>(g, m0, n0, s0, s1, s2, s3, klist, nkl);
                return;
            }
        }
        gemm3_tile<SPLIT_OUT, 0, true>(g, m0, n0, s0, s1, s2, s3, klist, nkl);
    } else {
        if constexpr (NARROW != 0) {
            if (tn == g.tiles_n - 1) {
                gemm3_tile<SPLIT_OUT, NARROW>(g, m0, n0, s0, s1, s2, s3);
                return;
            }
        }
        gemm3_tile<SPLIT_OUT, 0>(g, m0, n0, s0, s1, s2, s3);
    }
}

// C = A B with A FP32 column-major (A(m, k) = Af[m + k a_KS]) split in registers; FP32
// C; unsplit K, no batch.  Tile order, stagger and narrow last tile column as k_gemm3.
template <int NARROW>
__global__ void __launch_bounds__(NT, 1) k_gemm3_af32(G3 g)
{
    const long wg = blockIdx.x;
    const long tile = (wg % kXcd) * g.per_xcd + wg / kXcd;
    if (g.stagger > 0 && wg < 256) {
        const long wait = ((wg / kXcd) % kStaggerGroups) * g.stagger;   // bounded
        const long t0 = wall_clock64();
        while (wall_clock64() - t0 < wait) __builtin_amdgcn_s_sleep(2);
    }
    if (tile >= (long)g.tiles_m * g.tiles_n) return;
    const long grp = tile / (kGM * (long)g.tiles_n), first = grp * kGM;
    const int gsz = (int)((g.tiles_m - first) < kGM ? (g.tiles_m - first) : kGM);
    const long w = tile - grp * kGM * (long)g.tiles_n;
    const int tm = (int)(first + w % gsz);
    const int tn = (int)(w / gsz);
    const int m0 = tm * BM, n0 = tn * 256;
    __shared__ __attribute__((aligned(16))) __bf16 s0[kStageE];
    __shared__ __attribute__((aligned(16))) __bf16 s1[kStageE];
    __shared__ __attribute__((aligned(16))) __bf16 s2[kStageE];
    __shared__ __attribute__((aligned(16))) __bf16 s3[kStageE];
    if constexpr (NARROW != 0) {
        if (tn == g.tiles_n - 1) {
            gemm3_tile_af32<NARROW>(g, m0, n0, s0, s1, s2, s3);
            return;
        }
    }
    gemm3_tile_af32<0>(g, m0, n0, s0, s1, s2, s3);
}

// Psi rows [r0, r0 + rows) (M-contiguous: Psi[r + n2 k], k < K) -> the K-blocked A
// operand with K padded to Kp (a multiple of 16, zeros for k >= K):
// X[(k / 16) 16 rows + 16 (r - r0) + k % 16] (hi), lo at + rows Kp.  Block per (K block,
// 256 rows) on a 1-D grid, rows fastest: thread r loads its row's 16 values (each load
// 1 KB contiguous over the block) and stores them as two 32-byte runs per plane, so the
// block writes 8 KB contiguous per plane and needs no LDS.
// With kmask, the block also records whether any of its values is nonzero, at
// kmask[tile nkb + kb] (the mode-3 GEMM's row tile and K block), and an all-zero block is
// not written at all: the GEMM skips it.
__global__ void __launch_bounds__(256) k_split_transpose(const float* __restrict__ psi, long n2,
                                                          int K, int Kp, long r0, long rows,
                                                          __bf16* __restrict__ X,
                                                          unsigned char* __restrict__ kmask,
                                                          int nkb)
{
    // 1-D grid, row tile fastest: neighbouring blocks read neighbouring KB of the same
    // k rows and write neighbouring 8 KB of the same K block
    const long ntiles = (rows + 255) / 256;
    const long tile = blockIdx.x % ntiles;
    const int kb = (int)(blockIdx.x / ntiles);
    const long r = tile * 256 + threadIdx.x;
    const bool valid = r < rows;
    const int k0 = 16 * kb;
    bf16x8 h0, h1, l0, l1;
    bool nz = false;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int k = k0 + j;
        const float x = (valid && k < K) ? psi[r0 + r + n2 * k] : 0.0f;
        nz |= x != 0.0f;
        __bf16 hh, ll;
        split_bf16(x, hh, ll);
        if (j < 8) { h0[j] = hh; l0[j] = ll; } else { h1[j - 8] = hh; l1[j - 8] = ll; }
    }
    if (kmask) {   // every thread of the block reaches this barrier
        const int any = __syncthreads_or(nz ? 1 : 0);
        if (threadIdx.x == 0) kmask[tile * nkb + kb] = any ? 1 : 0;
        if (!any) return;
    }
    if (!valid) return;
    const long off = (long)kb * 16 * rows + 16 * r;
    bf16x8* dh = reinterpret_cast<bf16x8*>(X + off);
    bf16x8* dl = reinterpret_cast<bf16x8*>(X + rows * Kp + off);
    dh[0] = h0;
    dh[1] = h1;
    dl[0] = l0;
    dl[1] = l1;
}

// The mode-3 operand of a tensor T held with its first two indices swapped: rows
// r = i + n j, j in [j0, j0 + u), of T[i + n j + n^2 k] = G[j + n i + n^2 k], split and
// K-blocked over k exactly as k_split_transpose on T (rows = u n, r0 = n j0: X[(k / 16)
// 16 rows + 16 r + k % 16], lo at + rows Kp, zeros for k >= n), without forming T.
// Block per (32 j, 16 i, K block) on a 1-D grid, j tiles fastest (neighbouring blocks
// read neighbouring 128-byte runs of the same rows of G), through an LDS tile
// [16 k][16 i][32 j]; each j's 16 rows go out as 512 contiguous bytes per plane.
__global__ void __launch_bounds__(256) k_split_transpose_swapped(const float* __restrict__ G,
                                                                  long n2, int n, int Kp, long j0,
                                                                  long u,
                                                                  __bf16* __restrict__ X)
{
    __shared__ float t[16][16][33];
    const long ni = (n + 15) / 16, nj = (u + 31) / 32;
    long b = blockIdx.x;
    const long jt = b % nj;
    b /= nj;
    const int it = (int)(b % ni);
    const int kb = (int)(b / ni);
    const int i0 = 16 * it;
    const long jl0 = 32 * jt;
    const int lj = threadIdx.x & 31, grp = threadIdx.x >> 5;
    // all 32 loads of a thread in flight before the first LDS store
    float x[32];
    const bool jin = jl0 + lj < u;
    const float* g0 = G + (j0 + jl0 + lj) + n2 * (16L * kb);
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        const int q = grp + 8 * c, kk = q >> 4, ii = q & 15;
        const bool in = jin && 16 * kb + kk < n && i0 + ii < n;
        x[c] = in ? g0[(long)n * (i0 + ii) + n2 * kk] : 0.0f;
    }
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        const int q = grp + 8 * c;
        t[q >> 4][q & 15][lj] = x[c];
    }
    __syncthreads();
    const long rows = u * n;
#pragma unroll
    for (int q = threadIdx.x; q < 1024; q += 256) {
        const int h = q & 1, ii = (q >> 1) & 15, jj = q >> 5;
        const int i = i0 + ii;
        const long jl = jl0 + jj;
        if (i >= n || jl >= u) continue;
        bf16x8 hi, lo;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            __bf16 a, c;
            split_bf16(t[8 * h + v][ii][jj], a, c);
            hi[v] = a;
            lo[v] = c;
        }
        const long off = (long)kb * 16 * rows + 16 * (i + (long)n * jl) + 8 * h;
        *reinterpret_cast<bf16x8*>(X + off) = hi;
        *reinterpret_cast<bf16x8*>(X + rows * Kp + off) = lo;
    }
}

// Mode-3 operand in the row order of U's K-blocked layout: Psi rows (i, j), j in [j0, j0 +
// u) (r0 = n j0), as X rows m = (i / 16) 16 u + 16 (j - j0) + i % 16 for i < Kp (rows
// i >= n, the K padding of U, are zero), split and K-blocked over k like
// k_split_transpose, with rows = u Kp.  A 256-row GEMM tile is then 16 i x 16 j (for
// u >= 16): its output column is one contiguous run of U.  Block (K block, row tile) on a
// 1-D grid, tiles fastest; kmask and the skipped all-zero blocks as k_split_transpose.
__global__ void __launch_bounds__(256) k_split_transpose_u(const float* __restrict__ psi, long n2,
                                                            int n, int Kp, long r0, int u,
                                                            __bf16* __restrict__ X,
                                                            unsigned char* __restrict__ kmask,
                                                            int nkb)
{
    const long rows = (long)u * Kp;
    const long ntiles = (rows + 255) / 256;   // 1-D grid, row tile fastest
    const long tile = blockIdx.x % ntiles;
    const int kb = (int)(blockIdx.x / ntiles);
    const long m = tile * 256 + threadIdx.x;
    const long ib = m / (16L * u), rem = m - ib * 16L * u;
    const int jj = (int)(rem >> 4), i = (int)(ib * 16 + (rem & 15));
    const bool inb = m < rows, valid = inb && i < n;
    const int k0 = 16 * kb;
    bf16x8 h0, h1, l0, l1;
    bool nz = false;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int k = k0 + t;
        const float x = (valid && k < n) ? psi[r0 + i + (long)n * jj + n2 * k] : 0.0f;
        nz |= x != 0.0f;
        __bf16 hh, ll;
        split_bf16(x, hh, ll);
        if (t < 8) { h0[t] = hh; l0[t] = ll; } else { h1[t - 8] = hh; l1[t - 8] = ll; }
    }
    if (kmask) {   // every thread of the block reaches this barrier
        const int any = __syncthreads_or(nz ? 1 : 0);
        if (threadIdx.x == 0) kmask[tile * nkb + kb] = any ? 1 : 0;
        if (!any) return;
    }
    if (!inb) return;
    const long off = (long)kb * 16 * rows + 16 * m;
    bf16x8* dh = reinterpret_cast<bf16x8*>(X + off);
    bf16x8* dl = reinterpret_cast<bf16x8*>(X + rows * Kp + off);
    dh[0] = h0;
    dh[1] = h1;
    dl[0] = l0;
    dl[1] = l1;
}

// k_split_transpose_u over two neighbouring i blocks per workgroup (512 threads: 32 i x
// 16 j), so that each load of a k covers whole 128-byte lines of Psi (32 consecutive i)
// instead of half lines that another workgroup fetches again.  Same X and kmask as
// k_split_transpose_u.  Block (K block, (i-block pair, 16-j chunk)) on a 1-D grid;
// u % 16 == 0.
__global__ void __launch_bounds__(512) k_split_transpose_u2(const float* __restrict__ psi, long n2,
                                                             int n, int Kp, long r0, int u,
                                                             __bf16* __restrict__ X,
                                                             unsigned char* __restrict__ kmask,
                                                             int nkb)
{
    __shared__ int any2[2];
    const long rows = (long)u * Kp;
    const int chunks = u / 16;
    const long nblk = (long)((Kp / 16 + 1) / 2) * chunks;   // 1-D grid, row block fastest
    const int by = (int)(blockIdx.x % nblk);
    const int kb = (int)(blockIdx.x / nblk);
    const int ib2 = by / chunks, jc = by - ib2 * chunks;
    const int il = threadIdx.x & 31, jl = threadIdx.x >> 5, half = il >> 4;
    const int i = 32 * ib2 + il, jj = 16 * jc + jl;
    const long ib = 2L * ib2 + half;
    const long m = ib * 16L * u + 16L * jj + (il & 15);
    const bool inb = m < rows, valid = inb && i < n;
    if (threadIdx.x < 2) any2[threadIdx.x] = 0;
    __syncthreads();
    const int k0 = 16 * kb;
    bf16x8 h0, h1, l0, l1;
    bool nz = false;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int k = k0 + t;
        const float x = (valid && k < n) ? psi[r0 + i + (long)n * jj + n2 * k] : 0.0f;
        nz |= x != 0.0f;
        __bf16 hh, ll;
        split_bf16(x, hh, ll);
        if (t < 8) { h0[t] = hh; l0[t] = ll; } else { h1[t - 8] = hh; l1[t - 8] = ll; }
    }
    bool keep = true;
    if (kmask) {   // every thread of the block reaches these barriers
        if (nz) atomicOr(&any2[half], 1);
        __syncthreads();
        const long tile = ib * (u / 16) + jc;   // the 256-row tile of rows (ib, jc)
        if (threadIdx.x == 0 || threadIdx.x == 16) {
            if (ib * 16 < Kp) kmask[tile * nkb + kb] = any2[half] ? 1 : 0;
        }
        keep = any2[half] != 0;
    }
    if (!keep || !inb) return;
    const long off = (long)kb * 16 * rows + 16 * m;
    bf16x8* dh = reinterpret_cast<bf16x8*>(X + off);
    bf16x8* dl = reinterpret_cast<bf16x8*>(X + rows * Kp + off);
    dh[0] = h0;
    dh[1] = h1;
    dl[0] = l0;
    dl[1] = l1;
}

// bytes [0, cnt) of p set to v
__global__ void k_fill_u8(unsigned char* __restrict__ p, long cnt, unsigned char v)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (long)gridDim.x * blockDim.x)
        p[e] = v;
}

// ---- gathered mode 1 (n % 16 == 0) ---------------------------------------------------
// Psi's (i, j) pattern, Psi viewed as an n^2 x n matrix (rows r = i + n j contiguous):
// pat[r] = 1 when some Psi[r + n^2 k] != 0.  Thread per 4 consecutive rows, k looped in
// K blocks of 16 (n % 16 == 0): every load of a wave is 1 KB contiguous, Psi is read
// once, and a block's 16 loads go out before the first is tested.  Also jk[jb nkb + kb]
// = 1 when some Psi[i + n j + n^2 k] != 0 with j in column group jb = j / 16 and k in K
// block kb (jk zeroed before; nkb = n / 16): one ballot per wave and block, at most two
// groups per wave (a wave's 256 rows lie within 16 n).
__global__ void __launch_bounds__(256) k_psi_pattern(const float* __restrict__ psi, long n2, int n,
                                                     unsigned char* __restrict__ pat,
                                                     unsigned char* __restrict__ jk, int nkb)
{
    const long r = 4 * ((long)blockIdx.x * 256 + threadIdx.x);
    const bool live = r < n2;   // no early return: the whole wave takes part in the ballots
    const int lane = threadIdx.x & 63;
    // the bit patterns: a value is nonzero when its bits without the sign are (-0.0f == 0)
    using v2::u32x4;
    const unsigned mag = 0x7fffffffu;
    const u32x4* p = reinterpret_cast<const u32x4*>(psi + (live ? r : 0));
    const long step = n2 / 4;
    const int jb = live ? (int)(r / n / 16) : -1;
    const int jb_first = __shfl(jb, 0);   // lane 0 is live whenever any lane is
    u32x4 any = {0u, 0u, 0u, 0u};
    for (int kb = 0; kb < nkb; ++kb) {
        u32x4 v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
            v[u] = live ? __builtin_nontemporal_load(p + step * (16L * kb + u)) : u32x4{0u, 0u, 0u, 0u};
        u32x4 blk = v[0];
#pragma unroll
        for (int u = 1; u < 16; ++u) blk |= v[u];
        any |= blk;
        const bool nz = ((blk.x | blk.y | blk.z | blk.w) & mag) != 0;
        const unsigned long long m0 = __ballot(nz && jb == jb_first);
        const unsigned long long m1 = __ballot(nz && jb != jb_first);
        if (lane == 0 && m0) jk[(long)jb_first * nkb + kb] = 1;
        if (m1 && lane == __ffsll((long long)m1) - 1) jk[(long)jb * nkb + kb] = 1;
    }
    if (!live) return;
    const bool a = (any.x & mag) != 0, b = (any.y & mag) != 0, c = (any.z & mag) != 0,
               d = (any.w & mag) != 0;
    uchar4 o;
    o.x = a; o.y = b; o.z = c; o.w = d;
    *reinterpret_cast<uchar4*>(pat + r) = o;
}

// For each group jb of 16 columns j: the rows i (ascending) with pat[i + n j] set for some
// j of the group, list[jb n + t], their number cnt[jb], and the mode-1 mask of the group,
// T[jb nkb + kb] = 1 for the K blocks kb < ceil(cnt / 16) of the gathered rows.  Block per
// group; the 4 waves scan i in chunks of 256 and compact with ballots.
__global__ void __launch_bounds__(256) k_group_lists(const unsigned char* __restrict__ pat, int n,
                                                     int* __restrict__ list, int* __restrict__ cnt,
                                                     int nkb, unsigned char* __restrict__ T)
{
    __shared__ int wcount[4];
    __shared__ int base;
    const int jb = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    for (int i0 = 0; i0 < n; i0 += 256) {
        const int i = i0 + threadIdx.x;
        bool f = false;
        if (i < n) {
            const unsigned char* q = pat + i + (long)n * 16 * jb;
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) f |= q[(long)n * jj] != 0;
        }
        const unsigned long long m = __ballot(f);
        if (lane == 0) wcount[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; ++w) off += wcount[w];
        if (f) list[(long)jb * n + off + __popcll(m & ((1ull << lane) - 1ull))] = i;
        __syncthreads();
        if (threadIdx.x == 0) base += wcount[0] + wcount[1] + wcount[2] + wcount[3];
        __syncthreads();
    }
    const int c = base;
    if (threadIdx.x == 0) cnt[jb] = c;
    for (int kb = threadIdx.x; kb < nkb; kb += 256) T[(long)jb * nkb + kb] = kb * 16 < c ? 1 : 0;
}

// The mode-1 B operand of each group: B_jb(t, c) = phi[list[jb n + t], c] for t < cnt[jb],
// zero for the rest of the last K block and for columns c >= n; split, K-blocked, hi at
// Bg[jb 2 n Npad + (t / 16) 16 Npad + 16 c + t % 16], lo at + n Npad.  Thread (column c,
// 16-row block tb, group jb); blocks past the gathered rows are not written (never read).
__global__ void __launch_bounds__(256) k_gather_phi_list(const float* __restrict__ phi, int n,
                                                         const int* __restrict__ list,
                                                         const int* __restrict__ cnt, int Npad,
                                                         __bf16* __restrict__ Bg)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int tb = blockIdx.y, jb = blockIdx.z;
    const int L = cnt[jb];
    if (c >= Npad || 16 * tb >= L) return;
    bf16x8 h0, h1, l0, l1;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int t = 16 * tb + u;
        const float x = (t < L && c < n) ? phi[list[(long)jb * n + t] + (long)n * c] : 0.0f;
        __bf16 hh, ll;
        split_bf16(x, hh, ll);
        if (u < 8) { h0[u] = hh; l0[u] = ll; } else { h1[u - 8] = hh; l1[u - 8] = ll; }
    }
    const long off = (long)jb * 2 * n * Npad + (long)tb * 16 * Npad + 16L * c;
    bf16x8* dh = reinterpret_cast<bf16x8*>(Bg + off);
    bf16x8* dl = reinterpret_cast<bf16x8*>(Bg + off + (long)n * Npad);
    dh[0] = h0;
    dh[1] = h1;
    dl[0] = l0;
    dl[1] = l1;
}

// Mode-3 operand of the gathered layout for the groups [jb0, jb0 + G): X rows
// m = (jb - jb0) 16 n + 256 tb + 16 jj + t % 16 hold Psi row (i, j) with i = list[jb n + t]
// (t = 16 tb + t % 16; zero when t >= cnt[jb]) and j = 16 jb + jj, split and K-blocked over
// k like k_split_transpose_u (rows = 16 G n).  A 256-row tile is one (jb, tb): a tile past
// the group's rows only records an empty mask (its GEMM tile then writes nothing, and mode 1
// never reads those rows); so does a K block that jk marks zero for the whole group.
// kmask[tile nkb + kb] = any nonzero in the tile's K block kb.
// Block (K block, tile) on a 1-D grid, tiles fastest; thread (jj, t % 16).
__global__ void __launch_bounds__(256) k_split_gather(const float* __restrict__ psi, long n2, int n,
                                                      int jb0, const int* __restrict__ list,
                                                      const int* __restrict__ cnt,
                                                      const unsigned char* __restrict__ jk,
                                                      __bf16* __restrict__ X, long rows,
                                                      unsigned char* __restrict__ kmask, int nkb)
{
    const long ntiles = rows / 256;                // 1-D grid, tile fastest
    const long tile = blockIdx.x % ntiles;
    const int kb = (int)(blockIdx.x / ntiles);
    const int tpg = n / 16;                        // tiles per group
    const int jb = jb0 + (int)(tile / tpg), tb = (int)(tile % tpg);
    const int L = cnt[jb];
    // uniform over the block: rows past the group's list, or a K block in which the whole
    // group is zero (k_psi_pattern's jk): nothing to read
    if (16 * tb >= L || !jk[(long)jb * nkb + kb]) {
        if (threadIdx.x == 0) kmask[tile * nkb + kb] = 0;
        return;
    }
    const int tl = threadIdx.x & 15, jj = threadIdx.x >> 4;
    const int t = 16 * tb + tl, j = 16 * jb + jj;
    const int i = t < L ? list[(long)jb * n + t] : -1;
    const int k0 = 16 * kb;
    bf16x8 h0, h1, l0, l1;
    bool nz = false;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int k = k0 + u;
        const float x = (i >= 0 && k < n) ? psi[i + (long)n * j + n2 * k] : 0.0f;
        nz |= x != 0.0f;
        __bf16 hh, ll;
        split_bf16(x, hh, ll);
        if (u < 8) { h0[u] = hh; l0[u] = ll; } else { h1[u - 8] = hh; l1[u - 8] = ll; }
    }
    const int any = __syncthreads_or(nz ? 1 : 0);
    if (threadIdx.x == 0) kmask[tile * nkb + kb] = any ? 1 : 0;
    if (!any) return;
    const long m = tile * 256 + threadIdx.x;       // 16 jj + t % 16
    const long off = (long)kb * 16 * rows + 16 * m;
    bf16x8* dh = reinterpret_cast<bf16x8*>(X + off);
    bf16x8* dl = reinterpret_cast<bf16x8*>(X + rows * 16 * nkb + off);
    dh[0] = h0;
    dh[1] = h1;
    dl[0] = l0;
    dl[1] = l1;
}

// Per slab z = blockIdx.z: a rows x K matrix with K contiguous, src[z src_zs + e + ld r]
// (K % 4 == 0, ld % 4 == 0) -> the K-blocked operand Y[z y_zs + (e / 16) 16 rows + 16 r +
// e % 16] (hi), lo at + lo_off; past K nothing is written (a last partial K block keeps
// whatever its padding holds).  FOLD: K runs over kn-long segments kfs apart, element
// (r, e) at src[z src_zs + e % kn + kfs (e / kn) + ld r] (kn % 4 == 0, so no float4
// crosses a segment); for a tensor X[i + n r + n^2 k] with kn = n, kfs = n^2 the K index
// is e = i + n k.
// Block per tile of 128 rows x 64 e (4 K blocks), staged through LDS: the loads are float4
// runs along each row, and each K block's 128 rows go out as contiguous 16-byte pieces
// (a wave stores 1 KB runs), where a thread per 4 values would scatter 8-byte stores
// 16 rows apart.
constexpr int kSrRows = 128, kSrE = 64;
template <bool FOLD>
__global__ void __launch_bounds__(256) k_split_staged(const float* __restrict__ src, long src_zs,
                                                       long ld, long K, long rows,
                                                       __bf16* __restrict__ Y, long y_zs,
                                                       long lo_off, long kn, long kfs)
{
    __shared__ float t[kSrRows][kSrE + 4];
    const long e0 = (long)blockIdx.x * kSrE, r0 = (long)blockIdx.y * kSrRows;
    src += (long)blockIdx.z * src_zs;
    Y += (long)blockIdx.z * y_zs;
    const int tid = threadIdx.x;
#pragma unroll
    for (int it = 0; it < kSrRows * kSrE / 4 / 256; ++it) {
        const int idx = it * 256 + tid, rr = idx / (kSrE / 4), c4 = idx % (kSrE / 4);
        const long r = r0 + rr, e = e0 + 4 * c4;
        if (r < rows && e < K) {   // K % 4 == 0: the 4 values are all inside
            const long ea = FOLD ? e % kn + kfs * (e / kn) : e;
            *reinterpret_cast<float4*>(&t[rr][4 * c4]) =
                *reinterpret_cast<const float4*>(src + ea + ld * r);
        }
    }
    __syncthreads();
    // item = (K block kb of the tile, row rr, half h): 8 values -> 16 bytes per plane
#pragma unroll
    for (int it = 0; it < kSrRows * kSrE / 8 / 256; ++it) {
        const int idx = it * 256 + tid, h = idx & 1, rr = (idx >> 1) % kSrRows,
                  kb = (idx >> 1) / kSrRows;
        const long r = r0 + rr, e = e0 + 16 * kb + 8 * h;
        if (r >= rows || e >= K) continue;
        bf16x8 hi, lo;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            __bf16 a, b;
            split_bf16(t[rr][16 * kb + 8 * h + u], a, b);
            hi[u] = a;
            lo[u] = b;
        }
        const long off = (e >> 4) * 16 * rows + 16 * r + (e & 15);
        if (e + 8 <= K) {
            *reinterpret_cast<bf16x8*>(Y + off) = hi;
            *reinterpret_cast<bf16x8*>(Y + off + lo_off) = lo;
        } else {
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e + u < K) {
                    Y[off + u] = hi[u];
                    Y[off + lo_off + u] = lo[u];
                }
        }
    }
}

// Zeroes the last K block (block_elems bf16 at block_off) of both planes of `slabs` slabs
// of a K-blocked operand: the rows' K padding, which no producer writes.
__global__ void __launch_bounds__(256) k_zero_kblock(__bf16* __restrict__ base, long slabs,
                                                      long slab_stride, long plane_off,
                                                      long block_off, long block_elems)
{
    const long per = 2 * block_elems;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < slabs * per;
         t += (long)gridDim.x * blockDim.x) {
        const long sl = t / per, r = t - sl * per;
        const long p = r / block_elems, e = r - p * block_elems;
        base[sl * slab_stride + p * plane_off + block_off + e] = __bf16(0.0f);
    }
}

}  // namespace v3

// phi (n x n, col-major, ld n) -> zero-padded FP32 (Kpad x Npad, ld Kpad).
__global__ void k_prep_phi(const float* __restrict__ phi, int n, int Kpad, int Npad,
                           float* __restrict__ pf)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)Kpad * Npad) return;
    const int c = (int)(t / Kpad), k = (int)(t - (long)c * Kpad);
    pf[t] = (k < n && c < n) ? phi[k + (long)c * n] : 0.0f;
}

struct PhiPrep {
    const float* b;
    int ldb;
};

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// extra: zero columns past the padded width, for B operands that start at a column
// offset (a tile reads up to 256 columns from its first one)
size_t phi_bytes(int n, int extra = 0)
{
    return Carve::need((size_t)round_up(n, BK) * (round_up(n, v2::BN) + extra) * 4);
}

int prep_phi(hipStream_t s, const float* phi, int n, Carve& cv, PhiPrep& out, int extra = 0)
{
    const int Kpad = round_up(n, BK), Npad = round_up(n, v2::BN) + extra;
    const long e = (long)Kpad * Npad;
    float* pf = cv.take<float>(e);
    out.b = pf;
    out.ldb = Kpad;
    hipLaunchKernelGGL(k_prep_phi, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, s, phi, n,
                       Kpad, Npad, pf);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

// C(m, c) = sum_k A(m, k) phi(k, c) with A(m, k) = A[m sam + k sak] and
// C(m, c) = C[(m / R) SS + m % R + c CS]  (v2 kernel, FP32).
int gemm2(hipStream_t s, int kind, const float* A, long sam, long sak, const PhiPrep& B,
          float* C, long R, long SS, long CS, long M, int N, int K)
{
    if (M >= (1L << 31)) return set_error(FC3_ERR_UNSUPPORTED, "mcc3: GEMM M too large");
    v2::G2 g;
    g.A = A; g.sam = sam; g.sak = sak;
    g.B = B.b; g.ldb = B.ldb;
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
#define FC3_LAUNCH2(AM, V) hipLaunchKernelGGL((v2::k_gemm2<AM, V>), gd, bd, 0, s, g)
    if (amc) { if (vec) FC3_LAUNCH2(true, true); else FC3_LAUNCH2(true, false); }
    else     { if (vec) FC3_LAUNCH2(false, true); else FC3_LAUNCH2(false, false); }
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
int modes12(hipStream_t s, float* T, int n, int l0, int cnt, const PhiPrep& B,
            float* S)
{
    const long n2 = (long)n * n;
    // A(m = l n + j, k = i) = T_l[i + n j] = T[m n + k];  C(m, n') -> S_l[j + n n']
    FC3_TRY(gemm2(s, FC3_KERNEL_GEMM_MODE12, T + l0 * n2, n, 1, B, S, n, n2, n,
                  (long)cnt * n, n, n));
    // A(m = l n + n', k = j) = S_l[j + n n'];  C(m, m') -> T_l[n' + n m']
    FC3_TRY(gemm2(s, FC3_KERNEL_GEMM_MODE12, S, n, 1, B, T + l0 * n2, n, n2, n,
                  (long)cnt * n, n, n));
    return FC3_OK;
}


// ---------------------------------------------------------------------------------
// BF16X3 with pre-split operands (v3 kernels), used when n % 4 == 0.  The contraction
// length is padded to Kp = n rounded up to 16 with zeros.  Every operand is in the
// K-blocked layout of v3::G3; a slab of a mode-1/2 operand holds P = n Kp bf16 per plane:
//   phi:    B(k, c) = phi[k + n c], Kp rows, columns zero-padded to a multiple of BN.
//   mode 3: Psi row blocks (rows (i, j) = i + n j) are split-transposed into X, and the
//           GEMM writes U_l(i, j) = U[(i, j), l] split, already as the mode-1 A operand
//           A(l n + j, i): slab l hi at bf16 offset l 2P, lo at +P.  For n % 16 == 0
//           (P = n^2) that is exactly the byte range of K3 slab l and U lives there;
//           otherwise U has its own scratch.
//   mode 1: S_l = U_l^T phi, written split as the mode-2 A operand (same layout).
//   mode 2: K3_l = S_l^T phi, FP32 (over the bytes of U_l when U lives in K3).
bool v3_applies(int n, int prec) { return prec == FC3_PREC_BF16X3 && n % 4 == 0 && n >= 16; }
bool aligned16(const void* p) { return (uintptr_t)p % 16 == 0; }
inline int padk(int n) { return round_up(n, v3::BK); }

struct Split3 {
    const __bf16* b;
    long lo, ks;
};

// phi -> K-blocked B operand: hi at P[(k / 16) 16 Npad + 16 c + k % 16], lo at + Kp Npad;
// zero for k >= n or c >= n.
__global__ void k_prep_phi3(const float* __restrict__ phi, int n, int Kp, int Npad,
                            __bf16* __restrict__ P)
{
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long e = (long)Kp * Npad;
    if (o >= e) return;
    const long kb = o / (16L * Npad), rem = o - kb * 16L * Npad;
    const int c = (int)(rem >> 4), k = (int)(16 * kb + (rem & 15));
    const float x = (c < n && k < n) ? phi[k + (long)c * n] : 0.0f;
    __bf16 h, l;
    split_bf16(x, h, l);
    P[o] = h;
    P[e + o] = l;
}

size_t phi3_bytes(int n, int extra = 0)
{
    return Carve::need((size_t)2 * 2 * (round_up(n, v3::BN) + extra) * padk(n));
}

int prep_phi3(hipStream_t s, const float* phi, int n, Carve& cv, Split3& out, int extra = 0)
{
    const int Npad = round_up(n, v3::BN) + extra, Kp = padk(n);
    const long e = (long)Kp * Npad;
    __bf16* p = cv.take<__bf16>(2 * e);
    out.b = p;
    out.lo = e;
    out.ks = 16L * Npad;
    hipLaunchKernelGGL(k_prep_phi3, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, s, phi, n, Kp,
                       Npad, p);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

// Zero the K padding of `slabs` slabs of a K-blocked mode-1/2 operand (no-op if n % 16 == 0).
int zero_kpad(hipStream_t s, __bf16* base, int n, long slabs)
{
    const int Kp = padk(n);
    if (Kp == n || slabs <= 0) return FC3_OK;
    const long P = (long)n * Kp;
    const long blk = 16L * n;
    const long tot = slabs * 2 * blk;
    const unsigned grid = (unsigned)std::min<long>((tot + 255) / 256, 65536);
    hipLaunchKernelGGL(v3::k_zero_kblock, dim3(grid), dim3(256), 0, s, base, slabs, 2 * P, P,
                       (long)(Kp / 16 - 1) * blk, blk);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

// k_split_staged over `slabs` slabs (see there); K % 4 == 0, ld % 4 == 0, src 16-byte aligned.
// kn > 0: K folded over kn-long segments kfs apart (k_split_staged<true>)
int split_staged(hipStream_t s, const float* src, long src_zs, long ld, long K, long rows,
                 long slabs, __bf16* Y, long y_zs, long lo_off, long kn = 0, long kfs = 0)
{
    if (slabs < 1 || slabs > 65535 || (rows + v3::kSrRows - 1) / v3::kSrRows > 65535 ||
        (K + v3::kSrE - 1) / v3::kSrE > 0x7fffffffL)
        return set_error(FC3_ERR_UNSUPPORTED, "mcc3: split grid out of range");
    if (kn != 0 && (kn < 0 || kn % 4 != 0))
        return set_error(FC3_ERR_ARGUMENT, "mcc3: folded split segment must be a multiple of 4");
    const dim3 grid((unsigned)((K + v3::kSrE - 1) / v3::kSrE),
                    (unsigned)((rows + v3::kSrRows - 1) / v3::kSrRows), (unsigned)slabs);
    if (kn)
        hipLaunchKernelGGL(v3::k_split_staged<true>, grid, dim3(256), 0, s, src, src_zs, ld, K,
                           rows, Y, y_zs, lo_off, kn, kfs);
    else
        hipLaunchKernelGGL(v3::k_split_staged<false>, grid, dim3(256), 0, s, src, src_zs, ld, K,
                           rows, Y, y_zs, lo_off, K, 0L);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

// cnt slabs T_l (n x n, T_l[i + n j] at src[l n^2 + i + n j]) -> the K-blocked A operand
// of the mode-1 product A(l n + j, i) = T_l[i + n j]: hi at
// Y[l 2P + (i / 16) 16 n + 16 j + i % 16], lo at + P, with P = n Kp (Kp = n rounded up
// to 16; the padding k >= n is not written).  n % 4 == 0.
int split_slabs(hipStream_t s, const float* src, int n, int cnt, __bf16* Y)
{
    const long n2 = (long)n * n, P = (long)n * padk(n);
    return split_staged(s, src, n2, n, n, n, cnt, Y, 2 * P, P);
}

// Output addressing of v3::G3 (see there); lo plane at + lo for split outputs.
struct OutMap {
    void* C;
    long R, SS, KS, CS, lo;
};

// One k_gemm3 launch over g's M x N: 256-wide tile columns, the last one narrow (NARROW =
// 1..3: 32, 64, 128 columns) when it holds N % 256 <= 128 columns.
template <bool SPLIT, int NARROW, bool SPK = false>
int launch3(hipStream_t s, int kind, v3::G3 g)
{
    const int rem = g.N % 256;
    g.tiles_m = (int)((g.M + v3::BM - 1) / v3::BM);
    g.tiles_n = NARROW ? (g.N - rem) / 256 + 1 : (g.N + 255) / 256;
    const long tiles = g.tlist ? (long)g.tlist_n : (long)g.tiles_m * g.tiles_n;
    if (tiles == 0) return FC3_OK;
    g.per_xcd = (tiles + kXcd - 1) / kXcd;
    const long grid = g.per_xcd * kXcd;
    if (grid > 0x7fffffffL) return set_error(FC3_ERR_UNSUPPORTED, "mcc3: grid too large");
    if (g.nz < 1 || g.nz > 65535) return set_error(FC3_ERR_UNSUPPORTED, "mcc3: grid y out of range");
    const long steps = g.zk ? g.kc : g.K / v3::BK;   // K steps of one tile
    g.stagger = (!SPK && steps <= 32 && grid * g.nz >= 4096) ? v3::kStaggerTicks : 0;
    ProfScope prof(kind, s);
    hipLaunchKernelGGL((v3::k_gemm3<SPLIT, NARROW, SPK>), dim3((unsigned)grid, (unsigned)g.nz),
                       dim3(v3::NT), 0, s, g);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

template <bool SPLIT, bool SPK = false>
int launch3_cols(hipStream_t s, int kind, const v3::G3& g)
{
    // the N % 256 last columns: a narrow last tile column for <= 128 of them, else (or
    // none) all columns in 256-wide tiles
    const int rem = g.N % 256;
    const int shape = rem == 0 ? 0 : rem <= 32 ? 1 : rem <= 64 ? 2 : rem <= 128 ? 3 : 0;
    if (shape == 1) return launch3<SPLIT, 1, SPK>(s, kind, g);
    if (shape == 2) return launch3<SPLIT, 2, SPK>(s, kind, g);
    if (shape == 3) return launch3<SPLIT, 3, SPK>(s, kind, g);
    return launch3<SPLIT, 0, SPK>(s, kind, g);
}

template <int NARROW>
int launch3_af32(hipStream_t s, int kind, v3::G3 g)
{
    const int rem = g.N % 256;
    g.tiles_m = (int)((g.M + v3::BM - 1) / v3::BM);
    g.tiles_n = NARROW ? (g.N - rem) / 256 + 1 : (g.N + 255) / 256;
    const long tiles = (long)g.tiles_m * g.tiles_n;
    if (tiles == 0) return FC3_OK;
    g.per_xcd = (tiles + kXcd - 1) / kXcd;
    const long grid = g.per_xcd * kXcd;
    if (grid > 0x7fffffffL) return set_error(FC3_ERR_UNSUPPORTED, "mcc3: grid too large");
    g.stagger = (g.K / v3::BK <= 32 && grid >= 4096) ? v3::kStaggerTicks : 0;
    ProfScope prof(kind, s);
    hipLaunchKernelGGL((v3::k_gemm3_af32<NARROW>), dim3((unsigned)grid), dim3(v3::NT), 0, s, g);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

// C(m, c) = sum_k A(m, k) phi(k, c) for an FP32 column-major A(m, k) = A[m + k lda] with
// K rows (zero B rows past the contraction length; the A rows there finite), split into
// bf16 hi / lo inside the GEMM: C equals, bit for bit, split-transposing A and running
// gemm3.  FP32 C through o; M % 4 == 0, M >= 4, lda % 4 == 0, A 16-byte aligned.
int gemm3_af32(hipStream_t s, int kind, const float* A, long lda, const Split3& B,
               const OutMap& o, long M, int N, int K)
{
    if (M >= (1L << 31) || o.R >= (1L << 31))
        return set_error(FC3_ERR_UNSUPPORTED, "mcc3: GEMM M too large");
    const bool ok = K % v3::BK == 0 && K > 0 && M % 4 == 0 && M >= 4 && lda % 4 == 0 &&
                    lda >= M && B.lo % 8 == 0 && B.ks % 8 == 0 && o.R % 4 == 0 &&
                    o.SS % 4 == 0 && o.KS % 4 == 0 && o.CS % 4 == 0 && aligned16(A) &&
                    aligned16(B.b) && aligned16(o.C);
    if (!ok) return set_error(FC3_ERR_UNSUPPORTED, "mcc3: FP32-operand GEMM layout");
    v3::G3 g{};
    g.A = reinterpret_cast<const __bf16*>(A);
    g.a_KS = lda;
    g.B = B.b; g.b_lo = B.lo; g.b_KS = B.ks;
    g.C = o.C; g.c_SS = o.SS; g.c_KS = o.KS; g.c_CS = o.CS; g.c_lo = o.lo; g.c_R = (int)o.R;
    g.M = (int)M; g.N = N; g.K = K;
    g.zk = 1; g.kc = K / v3::BK; g.nz = 1;
    g.a_R2 = 1 << 30;
    g.epi = o.KS == 16 ? 1 : o.CS == 16 ? 2 : 0;
    const int rem = N % 256;
    const int shape = rem == 0 ? 0 : rem <= 32 ? 1 : rem <= 64 ? 2 : rem <= 128 ? 3 : 0;
    if (shape == 1) return launch3_af32<1>(s, kind, g);
    if (shape == 2) return launch3_af32<2>(s, kind, g);
    if (shape == 3) return launch3_af32<3>(s, kind, g);
    return launch3_af32<0>(s, kind, g);
}

// C(m, c) = sum_k A(m, k) phi(k, c); A(m, k) = A[(m / a_R) a_SS + (k / 16) a_KS + 16 (m % a_R)
// + k % 16], lo at + a_lo.
int gemm3(hipStream_t s, int kind, const __bf16* A, long a_lo, long a_R, long a_SS, long a_KS,
          const Split3& B, const OutMap& o, bool split_out, long M, int N, int K, int kc = 0,
          long c_ZS = 0, int batch = 0, long a_ZS = 0, long b_ZS = 0,
          const unsigned char* kmask = nullptr, int kmask_div = 1, long a_R2 = 0, long a_SS2 = 0,
          const int* tlist = nullptr, int tlist_n = 0, bool skip_empty = false)
{
    if (M >= (1L << 31) || a_R >= (1L << 31) || o.R >= (1L << 31))
        return set_error(FC3_ERR_UNSUPPORTED, "mcc3: GEMM M too large");
    const bool ok = K % v3::BK == 0 && M % 4 == 0 && M > 0 && a_lo % 8 == 0 && a_SS % 8 == 0 &&
                    a_KS % 8 == 0 && B.lo % 8 == 0 && B.ks % 8 == 0 && o.R % 4 == 0 &&
                    o.SS % 4 == 0 && o.KS % 4 == 0 && o.CS % 4 == 0 && o.lo % 4 == 0 &&
                    aligned16(A) && aligned16(B.b) && aligned16(o.C) && a_ZS % 8 == 0 &&
                    a_SS2 % 8 == 0 &&
                    b_ZS % 8 == 0 && c_ZS % 4 == 0;
    if (!ok) return set_error(FC3_ERR_UNSUPPORTED, "mcc3: v3 GEMM operand layout");
    v3::G3 g;
    g.A = A; g.a_lo = a_lo; g.a_SS = a_SS; g.a_KS = a_KS; g.a_R = (int)a_R;
    g.B = B.b; g.b_lo = B.lo; g.b_KS = B.ks;
    g.C = o.C; g.c_SS = o.SS; g.c_KS = o.KS; g.c_CS = o.CS; g.c_lo = o.lo; g.c_R = (int)o.R;
    g.M = (int)M; g.N = N; g.K = K;
    if (batch > 0) {   // batch of products over z < batch
        g.zk = 0; g.kc = K / v3::BK; g.nz = batch;
    } else {           // K split into chunks of kc blocks (kc = 0: unsplit)
        g.zk = 1; g.kc = kc > 0 ? kc : K / v3::BK;
        g.nz = (K / v3::BK + g.kc - 1) / g.kc;
    }
    g.a_ZS = a_ZS; g.b_ZS = b_ZS; g.c_ZS = c_ZS;
    g.kmask = kmask;
    g.kmask_ld = K / v3::BK;
    g.kmask_div = kmask_div > 0 ? kmask_div : 1;
    if (a_R2 >= (1L << 30)) return set_error(FC3_ERR_UNSUPPORTED, "mcc3: A row groups too large");
    g.a_R2 = a_R2 > 0 ? (int)a_R2 : (1 << 30);
    g.a_SS2 = a_R2 > 0 ? a_SS2 : 0;
    g.epi = o.KS == 16 ? 1 : o.CS == 16 ? 2 : 0;
    g.tlist = tlist;
    g.tlist_n = tlist ? tlist_n : 0;
    g.skip_empty = (kmask && skip_empty) ? 1 : 0;
    g.stagger = 0;   // set per launch shape in launch3
    if (tlist && ((batch == 0 && g.nz != 1) || (N + 255) / 256 > 256))
        return set_error(FC3_ERR_UNSUPPORTED, "mcc3: tile list shape");
    if (kmask) {   // sparse K: split output; unsplit, or a batch with one mask row per product
        if (!split_out || (batch == 0 && g.nz != 1) || g.kmask_ld > v3::kMaxKBlocks)
            return set_error(FC3_ERR_UNSUPPORTED, "mcc3: sparse-K GEMM shape");
        return launch3_cols<true, true>(s, kind, g);
    }
    return split_out ? launch3_cols<true>(s, kind, g) : launch3_cols<false>(s, kind, g);
}

// Launch sizing for the pre-split path.  A launch needs about 16k workgroups (64 per
// CU) to run at full rate: smaller ones lose 5-15% to ramp-up and tail.  Batches are
// equalised so that the last one is not a small remainder.
constexpr long kTilesPerLaunch = 16384;   // default; the batch autotuner may choose others
constexpr size_t kScratchSlabs = size_t(4) << 30, kScratchRows = size_t(8) << 30;
constexpr size_t kScratchSlabsInplace = size_t(1) << 30, kScratchRowsInplace = size_t(2) << 30;

// Slabs per mode-1/2 launch; each slab takes 4 n Kp bytes of split scratch.
int slab_batch3(int n, size_t cap_bytes, long tiles_per_launch)
{
    const long tiles_n = (n + v3::BN - 1) / v3::BN;
    long b = (tiles_per_launch * v3::BM + tiles_n * n - 1) / (tiles_n * n);
    const long cap = (long)(cap_bytes / (4 * (size_t)n * padk(n)));
    b = std::max(1L, std::min({b, cap, (long)n}));
    const long batches = (n + b - 1) / b;
    b = (n + batches - 1) / batches;
    // a multiple of 16 when n is one: every batch can then take the sparse mode 1
    if (n % 16 == 0 && b >= 16) b = b / 16 * 16;
    return (int)b;
}

// Mode-3 row block: whole columns j of Psi(i, j, :) (a multiple of n rows); the split
// copy X takes 4 Kp bytes per row.
long mode3_rows(int n, size_t cap_bytes, long tiles_per_launch)
{
    const long tiles_n = (n + v3::BN - 1) / v3::BN;
    const long target = (tiles_per_launch + tiles_n - 1) / tiles_n * v3::BM;
    long u = (target + n - 1) / n;                                   // columns j per block
    const long cap = (long)(cap_bytes / (4 * (size_t)padk(n))) / n;
    u = std::max(1L, std::min({u, cap, (long)n}));
    const long blocks = (n + u - 1) / u;
    u = (n + blocks - 1) / blocks;
    // whole groups of 16 columns when n % 16 == 0: every 256-row tile of the mode-3 GEMM
    // is then one 16 i x 16 j block of U, the granularity of the mode-1 block mask
    if (n % 16 == 0 && u >= 16) u = u / 16 * 16;
    return u * (long)n;
}

// Psi rows [r0, r0 + rows) -> X (split, K-blocked, K padded) -> C(m, l) for those rows.
// With kmask (mask_bytes(rows, n) bytes of scratch; split output only) the split records
// which 256-row x 16-K blocks hold a nonzero and the GEMM multiplies only those: a force-
// constant tensor is mostly zero blocks (LJ FCC 6x6x6: 3.5% of them are not), and the
// skipped products are all exactly zero, so U is the same bit for bit.
size_t mask_bytes(long rows, int n) { return Carve::need((size_t)((rows + 255) / 256) * (padk(n) / 16)); }

// swapped: psi holds the tensor with its first two indices swapped (r0, rows multiples
// of n; no kmask)
int mode3_block(hipStream_t s, const float* psi, int n, long r0, long rows, const Split3& B,
                __bf16* X, const OutMap& o, bool split_out, int ncols = 0,
                unsigned char* kmask = nullptr, bool swapped = false)
{
    const long n2 = (long)n * n;
    const int Kp = padk(n);
    if (kmask && (swapped || !split_out || Kp / 16 > v3::kMaxKBlocks)) kmask = nullptr;
    if (swapped) {
        if (r0 % n != 0 || rows % n != 0)
            return set_error(FC3_ERR_ARGUMENT, "mcc3: swapped mode-3 rows must be whole slabs");
        const long u = rows / n;
        const dim3 gt((unsigned)((Kp / 16) * ((u + 31) / 32) * ((n + 15) / 16)));
        hipLaunchKernelGGL(v3::k_split_transpose_swapped, gt, dim3(256), 0, s, psi, n2, n, Kp,
                           r0 / n, u, X);
    } else {
        const dim3 gt((unsigned)((Kp / 16) * ((rows + 255) / 256)));
        hipLaunchKernelGGL(v3::k_split_transpose, gt, dim3(256), 0, s, psi, n2, n, Kp, r0, rows,
                           X, kmask, Kp / 16);
    }
    FC3_HIP(hipGetLastError());
    return gemm3(s, FC3_KERNEL_GEMM_MODE3, X, rows * Kp, rows, 0, 16 * rows, B, o, split_out,
                 rows, ncols > 0 ? ncols : n, Kp, 0, 0, 0, 0, 0, kmask);
}

// Mode 3 of the out-of-place forms, Psi rows (i, j) with j in [j0, j0 + u) -> U's slabs
// l < ncols: the rows go through k_split_transpose_u, so that a GEMM tile is 16 i x 16 j
// and its output columns are contiguous runs of U (written through the staged epilogue);
// U's K-padding rows i in [n, Kp) receive zeros.  X: u Kp rows (x_rows), kmask:
// mask_bytes(u Kp, n) bytes or null (then every block is multiplied).
long x_rows(long R, int n) { return R / n * padk(n); }

int mode3_block_u(hipStream_t s, const float* psi, int n, long j0, int u, const Split3& B,
                  __bf16* X, __bf16* U, long P, int ncols, unsigned char* kmask)
{
    const long n2 = (long)n * n;
    const int Kp = padk(n);
    if (kmask && Kp / 16 > v3::kMaxKBlocks) kmask = nullptr;
    const long rows = (long)u * Kp;
    if (u % 16 == 0) {   // whole 128-byte lines of Psi per load
        const long pairs = (Kp / 16 + 1) / 2;
        const dim3 gt((unsigned)((Kp / 16) * pairs * (u / 16)));
        hipLaunchKernelGGL(v3::k_split_transpose_u2, gt, dim3(512), 0, s, psi, n2, n, Kp, n * j0,
                           u, X, kmask, Kp / 16);
    } else {
        const dim3 gt((unsigned)((Kp / 16) * ((rows + 255) / 256)));
        hipLaunchKernelGGL(v3::k_split_transpose_u, gt, dim3(256), 0, s, psi, n2, n, Kp, n * j0,
                           u, X, kmask, Kp / 16);
    }
    FC3_HIP(hipGetLastError());
    // C(m, c), m = (i / 16) 16 u + 16 (j - j0) + i % 16 -> U_c(i, j) = U[c 2P + (i / 16) 16 n +
    // 16 j + i % 16]
    const OutMap oU{U + 16 * j0, 16L * u, 16L * n, 16, 2 * P, P};
    return gemm3(s, FC3_KERNEL_GEMM_MODE3, X, rows * Kp, rows, 0, 16 * rows, B, oU, true, rows,
                 ncols, Kp, 0, 0, 0, 0, 0, kmask);
}

// Modes 3 and 1 skip the all-zero K blocks of their operands unless FC3_MCC3_DENSE is set
// to a nonzero value (then every block is multiplied: the reference for checking that the
// skipping changes nothing).
bool mode3_skips_zero_blocks()
{
    const char* e = std::getenv("FC3_MCC3_DENSE");
    return !(e && *e && std::atoi(e) != 0);
}

// Modes 1 and 2 on slabs [l0, l0 + cnt): U holds them as the mode-1 A operand (slab
// stride 2P bf16 with P = n Kp, lo at +P, K-blocked); S is scratch of the same layout; T
// receives the FP32 slabs K3_l[b + n c].  tl1 / tl2 (with counts): explicit tile lists of
// the mode-1 / mode-2 GEMMs (null: all tiles).
int modes12_v3(hipStream_t s, const __bf16* U, int n, int cnt, const Split3& B, __bf16* S,
               float* T, const int* tl1 = nullptr, int ntl1 = 0, const int* tl2 = nullptr,
               int ntl2 = 0)
{
    const long n2 = (long)n * n;
    const int Kp = padk(n);
    const long P = (long)n * Kp;
    // A(l n + j, i) = U_l(i, j);  C(l n + j, b) = S_l(j, b) -> mode-2 A(l n + b, j)
    const OutMap oS{S, n, 2 * P, 16L * n, 16, P};
    FC3_TRY(gemm3(s, FC3_KERNEL_GEMM_MODE12, U, P, n, 2 * P, 16L * n, B, oS, true, (long)cnt * n,
                  n, Kp, 0, 0, 0, 0, 0, nullptr, 1, 0, 0, tl1, ntl1));
    // A(l n + b, j) = S_l(j, b);  C(l n + b, c) = K3_l[b + n c]
    const OutMap oT{T, n, n2, 16, n, 0};
    return gemm3(s, FC3_KERNEL_GEMM_MODE2, S, P, n, 2 * P, 16L * n, B, oT, false, (long)cnt * n,
                 n, Kp, 0, 0, 0, 0, 0, nullptr, 1, 0, 0, tl2, ntl2);
}

// Gathered mode 1 (n % 16 == 0, zero blocks skipped): Psi's (i, j) pattern, for each group
// of 16 columns j the list of rows i it touches, the groups' mode-1 masks and gathered phi.
bool gathered_applies(int n) { return n % 16 == 0 && mode3_skips_zero_blocks(); }
size_t gathered_bytes(int n)
{
    const long G = n / 16, Npad = round_up(n, v3::BN);
    return Carve::need((size_t)n * n) + Carve::need(sizeof(int) * (size_t)G * n) +
           Carve::need(sizeof(int) * (size_t)G) + 2 * Carve::need((size_t)G * G) +
           Carve::need(2 * sizeof(__bf16) * (size_t)G * n * Npad);
}

size_t mcc3_v3_workspace_bytes(int n, bool inplace, long tiles)
{
    const long P = (long)n * padk(n);
    const int nb = slab_batch3(n, inplace ? kScratchSlabsInplace : kScratchSlabs, tiles);
    const long R = mode3_rows(n, inplace ? kScratchRowsInplace : kScratchRows, tiles);
    const size_t slabs = Carve::need(2 * sizeof(__bf16) * P * nb);
    const size_t u = (!inplace && padk(n) != n) ? Carve::need(2 * sizeof(__bf16) * P * n) : 0;
    return phi3_bytes(n) + slabs * (inplace ? 2 : 1) + u +
           Carve::need(2 * sizeof(__bf16) * (size_t)x_rows(R, n) * padk(n)) +
           mask_bytes(x_rows(R, n), n) + (n % 16 == 0 ? gathered_bytes(n) : 0) + 4096;
}

// ---- symmetric-unique K3 (the "calculate only the unique part" TODO of src/mcc.jl:116-117)
// The entries K3[a, b, l] with a <= b <= l, packed as out[l(l+1)(l+2)/6 + b(b+1)/2 + a]
// (slab by slab, each slab's upper triangle column by column).  The BF16X3 path runs the
// full pipeline's mode 3 and computes in modes 1 and 2 only the GEMM tiles that hold a
// needed entry: mode 1 (S_l(j, b) = sum_i U_l(i, j) phi(i, b)) the column tiles with
// some b <= l, mode 2 (K3_l(b, c)) the tiles with some b <= c <= l, about 1/2 and 1/6 of
// their work.  Each needed entry comes out of the same tile, rows and K loop as in the full
// mcc3, so it is the full K3's entry bit for bit.
__global__ void k_pack_unique(const float* __restrict__ T, int n, int l0,
                              float* __restrict__ out)
{
    const int b = blockIdx.x;
    const long l = (long)l0 + blockIdx.y;
    if (b > l) return;
    const float* src = T + (size_t)blockIdx.y * n * n + (size_t)b * n;
    float* dst = out + l * (l + 1) * (l + 2) / 6 + (long)b * (b + 1) / 2;
    for (int a = threadIdx.x; a <= b; a += blockDim.x) dst[a] = src[a];
}

// Tiles (tm << 8 | tn) of a GEMM over M x N (256 x 256 tiles, the last column up to N - 1)
// for which need(tm, row range, column range) holds, in the kernel's grouped order.
template <class F>
void push_tiles(std::vector<int>& v, long M, int N, F need)
{
    const int tiles_m = (int)((M + 255) / 256), tiles_n = (N + 255) / 256;
    for (int first = 0; first < tiles_m; first += v3::kGM) {
        const int gsz = std::min(v3::kGM, tiles_m - first);
        for (int tn = 0; tn < tiles_n; ++tn)
            for (int tm = first; tm < first + gsz; ++tm)
                if (need(256L * tm, std::min(256L * tm + 255, M - 1), 256 * tn,
                         std::min(256 * tn + 255, N - 1)))
                    v.push_back((tm << 8) | tn);
    }
}

// Mode-1 and mode-2 tile lists of the slab batch [l0, l0 + cnt) for the unique entries.
// gathered: mode 1 is a batch of products over the column groups, each with rows
// m = 16 l + j % 16 (the same tiles for every group); else one product with rows
// m = l n + j.  Mode 2 always has rows m = l n + b.
void unique_tiles(int n, int l0, int cnt, bool gathered, std::vector<int>& t1,
                  std::vector<int>& t2)
{
    const long M = (long)cnt * n;
    push_tiles(t1, gathered ? 16L * cnt : M, n, [&](long r0, long r1, int c0, int) {
        const long lmax = gathered ? std::min(r1 / 16, (long)cnt - 1)
                                   : std::min(r1 / n, (long)cnt - 1);
        return c0 <= l0 + lmax;   // some column b <= some slab l of the tile
    });
    push_tiles(t2, M, n, [&](long r0, long r1, int c0, int c1) {
        for (long ll = r0 / n; ll <= std::min(r1 / n, (long)cnt - 1); ++ll) {
            const long blo = std::max(r0 - ll * n, 0L);
            if (std::max(blo, (long)c0) <= std::min((long)l0 + ll, (long)c1)) return true;
        }
        return false;   // no b <= c <= l in the tile
    });
}

// Bytes of the tile lists of a whole unique mcc3 (upper bound).
size_t unique_list_bytes(int n, int nb)
{
    const long tiles_n = (n + 255) / 256;
    const long batches = (n + nb - 1) / nb;
    return Carve::need(sizeof(int) * 2 * ((long)n * n / 256 + batches) * tiles_n);
}

int mcc3_v3(hipStream_t s, const float* psi, const float* phi, int n, float* k3, void* ws,
            size_t ws_bytes, long tiles, float* k3u = nullptr)
{
    Carve cv(ws, ws_bytes);
    Split3 B;
    FC3_TRY(prep_phi3(s, phi, n, cv, B));
    const long n2 = (long)n * n;
    const int Kp = padk(n);
    const long P = (long)n * Kp;
    const int nb = slab_batch3(n, kScratchSlabs, tiles);
    const long R = mode3_rows(n, kScratchRows, tiles);
    __bf16* S = cv.take<__bf16>((size_t)2 * P * nb);
    __bf16* X = cv.take<__bf16>((size_t)2 * x_rows(R, n) * Kp);
    // U lives in the K3 buffer when its slabs have K3's size, else in scratch
    __bf16* U = Kp == n ? reinterpret_cast<__bf16*>(k3) : cv.take<__bf16>((size_t)2 * P * n);
    const bool skip = mode3_skips_zero_blocks();
    unsigned char* mask = skip ? cv.take<unsigned char>(mask_bytes(x_rows(R, n), n)) : nullptr;
    // gathered mode 1 (n % 16 == 0): pattern, row lists and masks of the column groups, and
    // the groups' gathered phi
    const bool gath = skip && n % 16 == 0;
    const int NG = n / 16, Npad = round_up(n, v3::BN);
    unsigned char* pat = gath ? cv.take<unsigned char>((size_t)n * n) : nullptr;
    int* glist = gath ? cv.take<int>((size_t)NG * n) : nullptr;
    int* gcnt = gath ? cv.take<int>(NG) : nullptr;
    unsigned char* gmask = gath ? cv.take<unsigned char>((size_t)NG * NG) : nullptr;
    unsigned char* jk = gath ? cv.take<unsigned char>((size_t)NG * NG) : nullptr;
    __bf16* Bg = gath ? cv.take<__bf16>((size_t)2 * NG * n * Npad) : nullptr;
    // unique entries only: every batch's mode-1 and mode-2 tile lists, uploaded once
    const size_t list_cap = k3u ? unique_list_bytes(n, nb) / sizeof(int) : 0;
    int* dl = k3u ? cv.take<int>(list_cap) : nullptr;
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 workspace too small");
    std::vector<int> lists;
    std::vector<long> off;   // per batch: mode-1 start, mode-2 start; then the end
    if (k3u) {
        for (int l0 = 0; l0 < n; l0 += nb) {
            const int cnt = (n - l0) < nb ? (n - l0) : nb;
            std::vector<int> t1, t2;
            unique_tiles(n, l0, cnt, gath, t1, t2);
            off.push_back((long)lists.size());
            lists.insert(lists.end(), t1.begin(), t1.end());
            off.push_back((long)lists.size());
            lists.insert(lists.end(), t2.begin(), t2.end());
        }
        off.push_back((long)lists.size());
        if (lists.size() > list_cap) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 tile lists");
        FC3_HIP(hipMemcpyAsync(dl, lists.data(), sizeof(int) * lists.size(),
                               hipMemcpyHostToDevice, s));
        // the host lists must outlive their copy (this waits for work already queued)
        FC3_HIP(hipStreamSynchronize(s));
    }
    FC3_TRY(zero_kpad(s, S, n, nb));
    if (gath) {
        // Psi's (i, j) pattern (one read of Psi), the groups' row lists and gathered phi
        hipLaunchKernelGGL(v3::k_fill_u8, dim3((unsigned)((NG * NG + 255) / 256)), dim3(256), 0, s,
                           jk, (long)NG * NG, (unsigned char)0);
        hipLaunchKernelGGL(v3::k_psi_pattern, dim3((unsigned)((n2 / 4 + 255) / 256)), dim3(256), 0,
                           s, psi, n2, n, pat, jk, NG);
        hipLaunchKernelGGL(v3::k_group_lists, dim3((unsigned)NG), dim3(256), 0, s, pat, n, glist,
                           gcnt, NG, gmask);
        hipLaunchKernelGGL(v3::k_gather_phi_list,
                           dim3((unsigned)(Npad / 256), (unsigned)NG, (unsigned)NG), dim3(256), 0,
                           s, phi, n, glist, gcnt, Npad, Bg);
        FC3_HIP(hipGetLastError());
        // mode 3 by blocks of whole column groups: U_l(t, j) = sum_k Psi[list(t), j, k] phi[k, l]
        // in U's gathered layout U[l 2P + 16 n jb + 256 (t / 16) + 16 (j % 16) + t % 16];
        // tiles past a group's rows are neither split nor multiplied nor written
        const int G = (int)(R / n / 16);
        for (int jb0 = 0; jb0 < NG; jb0 += G) {
            const int g = std::min(G, NG - jb0);
            const long rows = 16L * g * n;
            hipLaunchKernelGGL(v3::k_split_gather, dim3((unsigned)((Kp / 16) * (rows / 256))),
                               dim3(256), 0, s, psi, n2, n, jb0, glist, gcnt, jk, X, rows,
                               mask, Kp / 16);
            FC3_HIP(hipGetLastError());
            const OutMap oU{U + 16L * n * jb0, rows, 0, 16, 2 * P, P};
            FC3_TRY(gemm3(s, FC3_KERNEL_GEMM_MODE3, X, rows * Kp, rows, 0, 16 * rows, B, oU, true,
                          rows, n, Kp, 0, 0, 0, 0, 0, mask, 1, 0, 0, nullptr, 0, true));
        }
    } else {
        // mode 3 by blocks of whole columns j (R is a multiple of n); it also writes U's
        // K padding (zeros)
        for (long r0 = 0; r0 < n2; r0 += R) {
            const long rows = (n2 - r0) < R ? (n2 - r0) : R;
            FC3_TRY(mode3_block_u(s, psi, n, r0 / n, (int)(rows / n), B, X, U, P, n, mask));
        }
    }
    int bi = 0;
    for (int l0 = 0; l0 < n; l0 += nb, bi += 2) {
        const int cnt = (n - l0) < nb ? (n - l0) : nb;
        const int* t1 = k3u ? dl + off[bi] : nullptr;
        const int* t2 = k3u ? dl + off[bi + 1] : nullptr;
        const int n1 = k3u ? (int)(off[bi + 1] - off[bi]) : 0;
        const int n2t = k3u ? (int)(off[bi + 2] - off[bi + 1]) : 0;
        if (gath) {
            // mode 1, one product per column group jb (grid y): S_l(j, b) = sum_t U_l(t, j)
            // Bg_jb(t, b) over the group's ceil(cnt / 16) K blocks, written as the mode-2
            // operand S[l 2P + 16 n jb + 16 b + j % 16]
            const Split3 Bj{Bg, (long)n * Npad, 16L * Npad};
            const OutMap oS{S, 16, 2 * P, 16, 16, P};
            FC3_TRY(gemm3(s, FC3_KERNEL_GEMM_MODE12, U + 2 * P * l0, P, 16, 2 * P, 256, Bj, oS,
                          true, 16L * cnt, n, Kp, 0, 16L * n, NG, 16L * n, 2L * n * Npad, gmask,
                          1, 0, 0, t1, n1));
            // mode 2: A(l n + b, j) = S_l(j, b);  C(l n + b, c) = K3_l[b + n c]
            const OutMap oT{k3 + n2 * l0, n, n2, 16, n, 0};
            FC3_TRY(gemm3(s, FC3_KERNEL_GEMM_MODE2, S, P, n, 2 * P, 16L * n, B, oT, false,
                          (long)cnt * n, n, Kp, 0, 0, 0, 0, 0, nullptr, 1, 0, 0, t2, n2t));
        } else {
            FC3_TRY(modes12_v3(s, U + 2 * P * l0, n, cnt, B, S, k3 + n2 * l0, t1, n1, t2, n2t));
        }
        if (k3u) {
            hipLaunchKernelGGL(k_pack_unique, dim3((unsigned)n, (unsigned)cnt), dim3(256), 0, s,
                               k3 + n2 * l0, n, l0, k3u);
            FC3_HIP(hipGetLastError());
        }
    }
    return FC3_OK;
}

int mcc3_v3_inplace(hipStream_t s, float* psi, const float* phi, int n, void* ws, size_t ws_bytes)
{
    Carve cv(ws, ws_bytes);
    Split3 B;
    FC3_TRY(prep_phi3(s, phi, n, cv, B));
    const long n2 = (long)n * n;
    const int Kp = padk(n);
    const long P = (long)n * Kp;
    const int nb = slab_batch3(n, kScratchSlabsInplace, kTilesPerLaunch);
    const long R = mode3_rows(n, kScratchRowsInplace, kTilesPerLaunch);
    __bf16* Y = cv.take<__bf16>((size_t)2 * P * nb);
    __bf16* S = cv.take<__bf16>((size_t)2 * P * nb);
    __bf16* X = cv.take<__bf16>((size_t)2 * R * Kp);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 workspace too small");
    FC3_TRY(zero_kpad(s, Y, n, nb));
    FC3_TRY(zero_kpad(s, S, n, nb));
    // modes 1 and 2 on the slabs of Psi (they commute with mode 3)
    for (int l0 = 0; l0 < n; l0 += nb) {
        const int cnt = (n - l0) < nb ? (n - l0) : nb;
        FC3_TRY(split_slabs(s, psi + n2 * l0, n, cnt, Y));
        FC3_TRY(modes12_v3(s, Y, n, cnt, B, S, psi + n2 * l0));
    }
    // mode 3 by row blocks: each block is copied (split) to X before the GEMM overwrites it
    for (long r0 = 0; r0 < n2; r0 += R) {
        const long rows = (n2 - r0) < R ? (n2 - r0) : R;
        const OutMap oP{psi + r0, rows, 0, 16, n2, 0};
        FC3_TRY(mode3_block(s, psi, n, r0, rows, B, X, oP, false));
    }
    return FC3_OK;
}

// Scratch of the adjoint's mode products: the split M, the mode-1 operand U of
// modes31_t (n slabs), mode-3 row blocks, and two slab batches for modes12.
struct AdjScratch {
    long P, R;
    int nb;
};
AdjScratch adj_scratch(int n)
{
    AdjScratch a;
    a.P = (long)n * padk(n);
    a.nb = slab_batch3(n, kScratchSlabsInplace, kTilesPerLaunch);
    a.R = mode3_rows(n, kScratchRowsInplace, kTilesPerLaunch);
    return a;
}

}  // namespace

long mcc3_unique_count(int n) { return (long)n * (n + 1) * (n + 2) / 6; }

size_t mcc3_unique_workspace_bytes(int n, int precision)
{
    size_t b = mcc3_workspace_bytes(n, precision, false);
    if (v3_applies(n, precision))
        b += unique_list_bytes(n, slab_batch3(n, kScratchSlabs, kTilesPerLaunch));
    return b;
}

int mcc3_unique_run(hipStream_t s, const float* psi, const float* phi, int n, int precision,
                    float* k3s, float* k3u, void* ws, size_t ws_bytes)
{
    if (v3_applies(n, precision) && aligned16(psi) && aligned16(k3s))
        return mcc3_v3(s, psi, phi, n, k3s, ws, ws_bytes, kTilesPerLaunch, k3u);
    // other shapes and FP32: the full K3 into the scratch, then its unique entries
    FC3_TRY(mcc3_run(s, psi, phi, n, precision, k3s, ws, ws_bytes));
    for (int l0 = 0; l0 < n; l0 += 65535) {
        const int cnt = std::min(65535, n - l0);
        hipLaunchKernelGGL(k_pack_unique, dim3((unsigned)n, (unsigned)cnt), dim3(256), 0, s,
                           k3s + (size_t)n * n * l0, n, l0, k3u);
        FC3_HIP(hipGetLastError());
    }
    return FC3_OK;
}


bool mcc3_adjoint_v3_applies(int n, const void* a, const void* b, const void* c)
{
    return v3_applies(n, FC3_PREC_BF16X3) && aligned16(a) && aligned16(b) && aligned16(c);
}

size_t mcc3_adjoint_v3_bytes(int n)
{
    const AdjScratch a = adj_scratch(n);
    return phi3_bytes(n) + Carve::need(2 * sizeof(__bf16) * (size_t)a.P * n) +
           Carve::need(2 * sizeof(__bf16) * (size_t)a.R * padk(n)) +
           2 * Carve::need(2 * sizeof(__bf16) * (size_t)a.P * a.nb) + 4096;
}

int mcc3_mode3(hipStream_t s, const float* X, const float* M, int n, float* Y, void* ws,
               size_t ws_bytes)
{
    Carve cv(ws, ws_bytes);
    Split3 B;
    FC3_TRY(prep_phi3(s, M, n, cv, B));
    const AdjScratch a = adj_scratch(n);
    const long n2 = (long)n * n;
    __bf16* Xs = cv.take<__bf16>((size_t)2 * a.R * padk(n));
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 adjoint workspace too small");
    for (long r0 = 0; r0 < n2; r0 += a.R) {
        const long rows = (n2 - r0) < a.R ? (n2 - r0) : a.R;
        const OutMap oY{Y + r0, rows, 0, 16, n2, 0};
        FC3_TRY(mode3_block(s, X, n, r0, rows, B, Xs, oY, false));
    }
    return FC3_OK;
}

int mcc3_modes31_t(hipStream_t s, const float* X, const float* M, int n, float* Y, void* ws,
                   size_t ws_bytes, bool x_swapped)
{
    Carve cv(ws, ws_bytes);
    Split3 B;
    FC3_TRY(prep_phi3(s, M, n, cv, B));
    const AdjScratch a = adj_scratch(n);
    const long n2 = (long)n * n, P = a.P;
    __bf16* U = cv.take<__bf16>((size_t)2 * P * n);
    __bf16* Xs = cv.take<__bf16>((size_t)2 * a.R * padk(n));
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 adjoint workspace too small");
    FC3_TRY(zero_kpad(s, U, n, n));
    // mode 3 into U, the mode-1 operand: A(c n + j, i) = U_c(i, j) (as in the forward)
    for (long r0 = 0; r0 < n2; r0 += a.R) {
        const long rows = (n2 - r0) < a.R ? (n2 - r0) : a.R;
        const OutMap oU{U + 16 * (r0 / n), n, 16, 16L * n, 2 * P, P};
        FC3_TRY(mode3_block(s, X, n, r0, rows, B, Xs, oU, true, 0, nullptr, x_swapped));
    }
    // mode 1: C(c n + j, b) = Y[j + n b + n^2 c]
    const int nb = a.nb;
    for (int l0 = 0; l0 < n; l0 += nb) {
        const int cnt = (n - l0) < nb ? (n - l0) : nb;
        const OutMap oY{Y + n2 * l0, n, n2, 16, n, 0};
        FC3_TRY(gemm3(s, FC3_KERNEL_GEMM_MODE12, U + 2 * P * l0, P, n, 2 * P, 16L * n, B, oY,
                      false, (long)cnt * n, n, padk(n)));
    }
    return FC3_OK;
}

int mcc3_modes12(hipStream_t s, const float* X, const float* M, int n, float* Y, void* ws,
                 size_t ws_bytes)
{
    Carve cv(ws, ws_bytes);
    Split3 B;
    FC3_TRY(prep_phi3(s, M, n, cv, B));
    const AdjScratch a = adj_scratch(n);
    const long n2 = (long)n * n, P = a.P;
    const int nb = a.nb;
    __bf16* Ys = cv.take<__bf16>((size_t)2 * P * nb);
    __bf16* S = cv.take<__bf16>((size_t)2 * P * nb);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 adjoint workspace too small");
    FC3_TRY(zero_kpad(s, Ys, n, nb));
    FC3_TRY(zero_kpad(s, S, n, nb));
    for (int l0 = 0; l0 < n; l0 += nb) {
        const int cnt = (n - l0) < nb ? (n - l0) : nb;
        FC3_TRY(split_slabs(s, X + n2 * l0, n, cnt, Ys));
        FC3_TRY(modes12_v3(s, Ys, n, cnt, B, S, Y + n2 * l0));
    }
    return FC3_OK;
}

// Gathered phi rows for the fused path: for atom K, B_K(t, q) = phi[row_K(t), q] with
// row_K(3 s + b) = 3 K + b (s = 0) or 3 nbr[K][s - 1] + b (s <= count[K]), zero past
// the count or past n; split, K-blocked, hi at Bg[K 2 Kc Npad + (t / 16) 16 Npad + 16 q +
// t % 16], lo at + Kc Npad.  Thread (q, t block, K) writes 16 t as two 32-byte runs.
__global__ void __launch_bounds__(256) k_gather_phi_nbr(const float* __restrict__ phiT, int n,
                                                         const int* __restrict__ count,
                                                         const int* __restrict__ nbr, int cap,
                                                         int Kc, int Npad, __bf16* __restrict__ Bg)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Npad) return;
    const int tb = blockIdx.y, K = blockIdx.z;
    const int cnt = count[K];
    bf16x8 h0, h1, l0, l1;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int t = 16 * tb + c, sl = t / 3, b = t - 3 * (t / 3);
        int row = -1;
        if (sl == 0) row = 3 * K + b;
        else if (sl - 1 < cnt) row = 3 * nbr[(size_t)K * cap + sl - 1] + b;
        const float x = (row >= 0 && q < n) ? phiT[q + (long)n * row] : 0.0f;
        __bf16 hh, ll;
        split_bf16(x, hh, ll);
        if (c < 8) { h0[c] = hh; l0[c] = ll; } else { h1[c - 8] = hh; l1[c - 8] = ll; }
    }
    const long base = (long)K * 2 * Kc * Npad + (long)tb * 16 * Npad + 16L * q;
    bf16x8* dh = reinterpret_cast<bf16x8*>(Bg + base);
    bf16x8* dl = reinterpret_cast<bf16x8*>(Bg + base + (long)Kc * Npad);
    dh[0] = h0;
    dh[1] = h1;
    dl[0] = l0;
    dl[1] = l1;
}

int mcc3_gather_k(int cap) { return round_up(3 * (cap + 1), 16); }

size_t mcc3_from_gathered_bytes(int n, int cap, bool direct)
{
    const int N = n / 3, Kc = mcc3_gather_k(cap), Npad = round_up(n, v3::BN);
    const size_t base = phi3_bytes(n) + Carve::need(2 * sizeof(__bf16) * (size_t)Kc * Npad * N) +
                        4096;
    if (direct)   // the n^2 x Kp FP32 intermediate, read by mode 3 directly
        return base + Carve::need(sizeof(float) * (size_t)n * n * padk(n));
    const long R = mode3_rows(n, kScratchRowsInplace, kTilesPerLaunch);
    return base + Carve::need(2 * sizeof(__bf16) * (size_t)R * padk(n));
}

__global__ void k_fill_f32(float* __restrict__ p, long cnt, float v)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < cnt;
         i += (long)gridDim.x * blockDim.x)
        p[i] = v;
}

int mcc3_from_gathered(hipStream_t s, const __bf16* U, const float* phi, const float* phiT, int n,
                       const int* count, const int* nbr, int cap, float* k3, void* ws,
                       size_t ws_bytes)
{
    if (!v3_applies(n, FC3_PREC_BF16X3) || !aligned16(U) || !aligned16(k3) || n % 3 != 0)
        return set_error(FC3_ERR_UNSUPPORTED, "fused mcc3: needs n % 4 == 0, n >= 16");
    const int N = n / 3, Kc = mcc3_gather_k(cap), Npad = round_up(n, v3::BN);
    const int Kp = padk(n);
    // direct: mode 2 writes the intermediate to an n^2 x Kp FP32 scratch that mode 3 reads
    // as its FP32 A operand (no split pass); else (a smaller workspace) mode 2 writes it
    // into k3 and mode 3 runs in place by row blocks, each split into X first.  Both give
    // the same K3 bit for bit.
    const bool direct = ws_bytes >= mcc3_from_gathered_bytes(n, cap, true) &&
                        !std::getenv("FC3_FUSED_SPLIT");
    Carve cv(ws, ws_bytes);
    Split3 B;
    FC3_TRY(prep_phi3(s, phi, n, cv, B));
    const long n2 = (long)n * n;
    const long R = mode3_rows(n, kScratchRowsInplace, kTilesPerLaunch);
    __bf16* Bg = cv.take<__bf16>((size_t)2 * Kc * Npad * N);
    float* V = direct ? cv.take<float>((size_t)n2 * Kp) : nullptr;
    __bf16* X = direct ? nullptr : cv.take<__bf16>((size_t)2 * R * Kp);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "fused mcc3 workspace too small");
    hipLaunchKernelGGL(k_gather_phi_nbr, dim3((unsigned)((Npad + 255) / 256), (unsigned)(Kc / 16),
                                              (unsigned)N),
                       dim3(256), 0, s, phiT, n, count, nbr, cap, Kc, Npad, Bg);
    FC3_HIP(hipGetLastError());
    // mode 2, one product per atom K over its 3 slabs: C(g n + p, q) = K3[p + n q + n^2 (3K + g)]
    // = sum_t U_K(g n + p, t) B_K(t, q)
    const Split3 BK{Bg, (long)Kc * Npad, 16L * Npad};
    const OutMap oT{direct ? V : k3, n, n2, 16, n, 0};
    FC3_TRY(gemm3(s, FC3_KERNEL_GEMM_MODE12, U, 3L * n * Kc, 3L * n, 0, 16L * 3 * n, BK, oT, false,
                  3L * n, n, Kc, 0, 3 * n2, N, 2L * 3 * n * Kc, 2L * Kc * Npad));
    if (direct) {
        // mode 3 from the FP32 intermediate: K3(m, l) = sum_k V[m + n^2 k] phi(k, l); the
        // K padding slabs of V are zero (phi's padded rows are zero too)
        if (Kp > n) {
            const long pad = n2 * (Kp - n);
            hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)std::min<long>((pad + 255) / 256, 65536)),
                               dim3(256), 0, s, V + n2 * n, pad, 0.0f);
            FC3_HIP(hipGetLastError());
        }
        const OutMap oK{k3, n2, 0, 16, n2, 0};
        return gemm3_af32(s, FC3_KERNEL_GEMM_MODE3, V, n2, B, oK, n2, n, Kp);
    }
    // mode 3 in place by row blocks (each block split into X before the GEMM overwrites it)
    for (long r0 = 0; r0 < n2; r0 += R) {
        const long rows = (n2 - r0) < R ? (n2 - r0) : R;
        const OutMap oK{k3 + r0, rows, 0, 16, n2, 0};
        FC3_TRY(mode3_block(s, k3, n, r0, rows, B, X, oK, false));
    }
    return FC3_OK;
}

bool mcc3_contract_applies(int n)
{
    return v3_applies(n, FC3_PREC_BF16X3);
}

// one split operand: 2 n^3 bf16, plus 16 Npad of slack that the column tiles past n read
static size_t contract_operand_bytes(int n)
{
    return Carve::need(sizeof(__bf16) * (2 * (size_t)n * n * n + 16 * (size_t)round_up(n, v3::BN)));
}

size_t mcc3_contract_bytes(int n) { return 2 * contract_operand_bytes(n) + 4096; }

int mcc3_contract(hipStream_t s, int form, const float* X, const float* H, int n, float* P,
                  int* zc, void* ws, size_t ws_bytes)
{
    Carve cv(ws, ws_bytes);
    const long n2 = (long)n * n, n3 = n2 * n;
    __bf16* SA = cv.take<__bf16>(contract_operand_bytes(n) / sizeof(__bf16));
    __bf16* SB = cv.take<__bf16>(contract_operand_bytes(n) / sizeof(__bf16));
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 adjoint workspace too small");
    for (int t = 0; t < 2; ++t) {
        const float* src = t == 0 ? X : H;
        __bf16* dst = t == 0 ? SA : SB;
        if (form == 0 || (form == 2 && t == 1)) {
            // rows r < n, K index e < n^2 at stride n: src[r + n e]
            const dim3 gt((unsigned)((n2 / 16) * ((n + 255) / 256)));
            hipLaunchKernelGGL(v3::k_split_transpose, gt, dim3(256), 0, s, src, (long)n, (int)n2,
                               (int)n2, 0L, (long)n, dst, (unsigned char*)nullptr, 0);
        } else if (form == 2) {
            // rows r < n, K index e = i + n k: src[i + n r + n^2 k] (form 0 of the tensor
            // with its first two indices swapped, without forming it)
            FC3_TRY(split_staged(s, src, 0, n, n2, n, 1, dst, 0, (long)n * n2, n, n2));
        } else {
            // rows r < n, K index e < n^2 contiguous: src[e + n^2 r]
            FC3_TRY(split_staged(s, src, 0, n2, n2, n, 1, dst, 0, (long)n * n2));
        }
        FC3_HIP(hipGetLastError());
    }
    // K split into at most min(64, n) chunks; partial z at P + z n^2 as P_z[a + n b]
    const int nkb = (int)(n2 / v3::BK);
    const int maxz = n < 64 ? n : 64;
    const int kc = (nkb + maxz - 1) / maxz;
    *zc = (nkb + kc - 1) / kc;
    const Split3 B{SB, n3, 16L * n};
    const OutMap oP{P, n, 0, 16, n, 0};
    return gemm3(s, FC3_KERNEL_GEMM_MODE12, SA, n3, n, 0, 16L * n, B, oP, false, n, n, (int)n2,
                 kc, n2);
}

size_t mcc3_workspace_bytes(int n, int precision, bool inplace, const Mcc3Config* cfg)
{
    const long n2 = (long)n * n;
    const long tiles = cfg && cfg->tiles > 0 ? cfg->tiles : kTilesPerLaunch;
    size_t b = phi_bytes(n) + Carve::need(sizeof(float) * n2 * slab_batch(n)) + 4096;
    if (inplace) {
        // mode-3 row-block scratch: R rows x n
        const long R = 1L << 16;
        b += Carve::need(sizeof(float) * (size_t)(R < n2 ? R : n2) * n);
    }
    // the pre-split path is taken only for 16-byte aligned tensors; size for either path
    if (v3_applies(n, precision)) b = std::max(b, mcc3_v3_workspace_bytes(n, inplace, tiles));
    return b;
}

bool mcc3_variant_applies(int n, int precision, int variant)
{
    return variant == 2 || (variant == 3 && v3_applies(n, precision));
}

int mcc3_run(hipStream_t s, const float* psi, const float* phi, int n, int precision, float* k3,
             void* ws, size_t ws_bytes, const Mcc3Config* cfg)
{
    const int variant = cfg && cfg->variant ? cfg->variant : 3;
    const long tiles = cfg && cfg->tiles > 0 ? cfg->tiles : kTilesPerLaunch;
    if (variant == 3 && v3_applies(n, precision) && aligned16(psi) && aligned16(k3))
        return mcc3_v3(s, psi, phi, n, k3, ws, ws_bytes, tiles);
    Carve cv(ws, ws_bytes);
    PhiPrep B;
    FC3_TRY(prep_phi(s, phi, n, cv, B));
    const int nb = slab_batch(n);
    const long n2 = (long)n * n;
    float* S = cv.take<float>((size_t)n2 * nb);
    // mode 3: U[(i,j), l] = sum_k Psi[(i,j), k] phi[k, l] -> K3 buffer
    FC3_TRY(gemm2(s, FC3_KERNEL_GEMM_MODE3, psi, 1, n2, B, k3, n2, 0, n2, n2, n, n));
    for (int l0 = 0; l0 < n; l0 += nb) {
        const int cnt = (n - l0) < nb ? (n - l0) : nb;
        FC3_TRY(modes12(s, k3, n, l0, cnt, B, S));
    }
    return FC3_OK;
}

int mcc3_run_inplace(hipStream_t s, float* psi, const float* phi, int n, int precision, void* ws,
                     size_t ws_bytes)
{
    if (v3_applies(n, precision) && aligned16(psi))
        return mcc3_v3_inplace(s, psi, phi, n, ws, ws_bytes);   // default launch sizing
    Carve cv(ws, ws_bytes);
    PhiPrep B;
    FC3_TRY(prep_phi(s, phi, n, cv, B));
    const int nb = slab_batch(n);
    const long n2 = (long)n * n;
    float* S = cv.take<float>((size_t)n2 * nb);
    for (int l0 = 0; l0 < n; l0 += nb) {
        const int cnt = (n - l0) < nb ? (n - l0) : nb;
        FC3_TRY(modes12(s, psi, n, l0, cnt, B, S));
    }
    // mode 3 in place by row blocks through scratch
    const long R0 = 1L << 16;
    const long R = R0 < n2 ? R0 : n2;
    float* W = cv.take<float>((size_t)R * n);
    for (long r0 = 0; r0 < n2; r0 += R) {
        const long rows = (n2 - r0) < R ? (n2 - r0) : R;
        FC3_TRY(gemm2(s, FC3_KERNEL_GEMM_MODE3, psi + r0, 1, n2, B, W, rows, 0, rows,
                      rows, n, n));
        FC3_HIP(hipMemcpy2DAsync(psi + r0, sizeof(float) * n2, W, sizeof(float) * rows,
                                 sizeof(float) * rows, n, hipMemcpyDeviceToDevice, s));
    }
    return FC3_OK;
}

// ---- bounded device memory: mcc3(Psi, phi, block_size) (src/mcc.jl:89-120) ----------------
// K3 slabs [l0, l0 + cnt) from the whole Psi: mode 3 against phi's columns l0 .. l0 + cnt
// only, then modes 1 and 2 on those slabs.  The device holds the block's output slabs and
// bounded scratch (512 MiB of slab scratch, 2 GiB of split Psi rows); Psi is read, and for
// BF16X3 split, once per block.  A GEMM tile's result depends only on its own rows and
// columns and the full K loop, never on which other tiles a launch holds, so every slab
// is bitwise the slab the unblocked mcc3_run computes.
namespace {
constexpr size_t kBlkSlabScratch = size_t(512) << 20, kBlkRowScratch = size_t(2) << 30;
constexpr int kBlkExtraCols = 256;   // a tile reads up to 256 phi columns from its first

long blk_rows(int n)
{
    const size_t Kp = padk(n);
    long u = (long)(kBlkRowScratch / (4 * Kp)) / n;   // whole columns j per row block
    u = std::max(1L, std::min(u, (long)n));
    const long blocks = (n + u - 1) / u;
    return (n + blocks - 1) / blocks * (long)n;
}
}  // namespace

bool mcc3_blocked_v3(int n, int precision, const void* psi)
{
    return v3_applies(n, precision) && aligned16(psi);
}

size_t mcc3_blocked_bytes(int n, bool v3, int bs)
{
    const size_t n2 = (size_t)n * n;
    if (v3) {
        const size_t Kp = padk(n), P = (size_t)n * Kp;
        const size_t nb = std::min<size_t>(bs, slab_batch3(n, kBlkSlabScratch, kTilesPerLaunch));
        return phi3_bytes(n, kBlkExtraCols) + Carve::need(4 * P * nb) +
               Carve::need(4 * (size_t)x_rows(blk_rows(n), n) * Kp) +
               mask_bytes(x_rows(blk_rows(n), n), n) +
               ((int)Kp != n ? Carve::need(4 * P * bs) : 0) + 4096;
    }
    const size_t nb = std::min<size_t>(bs, slab_batch(n));
    return phi_bytes(n, kBlkExtraCols) + Carve::need(4 * n2 * nb) + 4096;
}

int mcc3_blocked_slabs(hipStream_t s, const float* psi, const float* phi, int n, bool v3, int bs,
                       int l0, int cnt, float* out, void* ws, size_t ws_bytes)
{
    if (cnt < 1 || cnt > bs || l0 < 0 || l0 + cnt > n)
        return set_error(FC3_ERR_ARGUMENT, "mcc3 blocked: slab range out of bounds");
    const long n2 = (long)n * n;
    Carve cv(ws, ws_bytes);
    if (v3) {
        if (!aligned16(out)) return set_error(FC3_ERR_ARGUMENT, "mcc3 blocked: unaligned block");
        Split3 B;
        FC3_TRY(prep_phi3(s, phi, n, cv, B, kBlkExtraCols));
        const int Kp = padk(n);
        const long P = (long)n * Kp;
        const int nb = (int)std::min<long>(bs, slab_batch3(n, kBlkSlabScratch, kTilesPerLaunch));
        const long R = blk_rows(n);
        __bf16* S = cv.take<__bf16>((size_t)2 * P * nb);
        __bf16* X = cv.take<__bf16>((size_t)2 * x_rows(R, n) * Kp);
        // U lives in the output block when its slabs have K3's size, else in scratch
        __bf16* U = Kp == n ? reinterpret_cast<__bf16*>(out) : cv.take<__bf16>((size_t)2 * P * bs);
        unsigned char* mask =
            mode3_skips_zero_blocks() ? cv.take<unsigned char>(mask_bytes(x_rows(R, n), n)) : nullptr;
        if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 blocked workspace too small");
        FC3_TRY(zero_kpad(s, S, n, nb));
        const Split3 Bc{B.b + 16L * l0, B.lo, B.ks};   // B(k, c) = phi[k, l0 + c]
        for (long r0 = 0; r0 < n2; r0 += R) {
            const long rows = (n2 - r0) < R ? (n2 - r0) : R;
            FC3_TRY(mode3_block_u(s, psi, n, r0 / n, (int)(rows / n), Bc, X, U, P, cnt, mask));
        }
        for (int c0 = 0; c0 < cnt; c0 += nb) {
            const int c = (cnt - c0) < nb ? (cnt - c0) : nb;
            FC3_TRY(modes12_v3(s, U + 2 * P * c0, n, c, B, S, out + n2 * c0));
        }
        return FC3_OK;
    }
    PhiPrep B;
    FC3_TRY(prep_phi(s, phi, n, cv, B, kBlkExtraCols));
    const int nb = std::min(bs, slab_batch(n));
    float* S = cv.take<float>((size_t)n2 * nb);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 blocked workspace too small");
    const PhiPrep Bc{B.b + (long)l0 * B.ldb, B.ldb};    // B(k, c) = phi[k, l0 + c]
    // mode 3: U[(i, j), c] = sum_k Psi[(i, j), k] phi[k, l0 + c] -> the output block
    FC3_TRY(gemm2(s, FC3_KERNEL_GEMM_MODE3, psi, 1, n2, Bc, out, n2, 0, n2, n2, cnt, n));
    for (int c0 = 0; c0 < cnt; c0 += nb) {
        const int c = (cnt - c0) < nb ? (cnt - c0) : nb;
        FC3_TRY(modes12(s, out, n, c0, c, B, S));
    }
    return FC3_OK;
}

}  // namespace fc3
