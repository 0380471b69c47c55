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

constexpr int BM = 128, BN = 128, BK = 32;
constexpr int NT = 256;               // 4 waves: 2 (m) x 2 (n), 64 x 64 each
constexpr int PADH = 8;               // bf16 row padding (16 bytes)
constexpr int PADF = 4;               // fp32 row padding (16 bytes)
constexpr int kXcd = 8;               // XCDs on MI355X

struct GemmArgs {
    const float* A;   // element (m, k) of batch b at A[b*sab + m*sam + k*sak]
    long sam, sak, sab;
    const void* Bhi;  // padded B (Kpad x Npad col-major, ld = ldb): bf16 hi or fp32
    const void* Blo;  // bf16 lo (BF16X3 only)
    int ldb;
    float* C;         // element (m, n) of batch b at C[b*scb + m + n*scn]
    long scn, scb;
    int M, N, K, batch;
    int tiles_m, tiles_n;
    long per_xcd;     // ceil(tiles_m * tiles_n * batch / kXcd)
};

__device__ __forceinline__ void split_bf16(float x, __bf16& h, __bf16& l)
{
    h = static_cast<__bf16>(x);
    const float hf = static_cast<float>(h);
    l = static_cast<__bf16>(x - hf);
}

// A tile (BM x BK) global -> registers.  Each thread holds a 4 (rows) x 4 (k) block:
// rows 4*mq .. 4*mq+3, k 4*kq .. 4*kq+3 of the tile, v[row][k].
template <bool AMC, bool VEC>
__device__ __forceinline__ void load_a(const GemmArgs& g, const float* __restrict__ A, int m0,
                                       int k0, float v[4][4])
{
    const int t = threadIdx.x;
    const int mq = t & 31, kq = t >> 5;
    const int mr = m0 + 4 * mq, kc = k0 + 4 * kq;
    if constexpr (AMC) {
        // M contiguous: one float4 along m per k
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int k = kc + kk;
            if (VEC && k < g.K && mr + 3 < g.M) {
                const float4 x = *reinterpret_cast<const float4*>(A + (long)k * g.sak + mr);
                v[0][kk] = x.x; v[1][kk] = x.y; v[2][kk] = x.z; v[3][kk] = x.w;
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    v[i][kk] = (k < g.K && mr + i < g.M) ? A[(long)k * g.sak + (long)(mr + i) * g.sam]
                                                         : 0.0f;
            }
        }
    } else {
        // K contiguous: one float4 along k per row
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = mr + i;
            if (VEC && m < g.M && kc + 3 < g.K) {
                const float4 x = *reinterpret_cast<const float4*>(A + (long)m * g.sam + kc);
                v[i][0] = x.x; v[i][1] = x.y; v[i][2] = x.z; v[i][3] = x.w;
            } else {
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
                    v[i][kk] = (m < g.M && kc + kk < g.K)
                                   ? A[(long)m * g.sam + (long)(kc + kk) * g.sak]
                                   : 0.0f;
            }
        }
    }
}

template <int PREC, bool AMC, bool VEC>
__global__ void __launch_bounds__(NT) k_gemm(GemmArgs g)
{
    // XCD-aware tile order: consecutive workgroup ids land on different XCDs, so
    // tile t is assigned to workgroup (t % per_xcd) * 8 + t / per_xcd, giving each
    // XCD a contiguous run of tiles that share A rows (one M-tile, all N-tiles).
    const long wg = blockIdx.x;
    const long tile = (wg % kXcd) * g.per_xcd + wg / kXcd;
    const long tiles_per_batch = (long)g.tiles_m * g.tiles_n;
    if (tile >= tiles_per_batch * g.batch) return;
    const int b = (int)(tile / tiles_per_batch);
    const long tb = tile - (long)b * tiles_per_batch;
    const int tm = (int)(tb / g.tiles_n), tn = (int)(tb - (long)tm * g.tiles_n);
    const int m0 = tm * BM, n0 = tn * BN;

    const float* __restrict__ A = g.A + (long)b * g.sab;
    float* __restrict__ C = g.C + (long)b * g.scb;

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 64;

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int nk = (g.K + BK - 1) / BK;
    float av[4][4];

    if constexpr (PREC == FC3_PREC_BF16X3) {
        __shared__ __attribute__((aligned(16))) __bf16 sAh[BM][BK + PADH];
        __shared__ __attribute__((aligned(16))) __bf16 sAl[BM][BK + PADH];
        __shared__ __attribute__((aligned(16))) __bf16 sBh[BN][BK + PADH];
        __shared__ __attribute__((aligned(16))) __bf16 sBl[BN][BK + PADH];
        const __bf16* __restrict__ Bh = static_cast<const __bf16*>(g.Bhi);
        const __bf16* __restrict__ Bl = static_cast<const __bf16*>(g.Blo);
        uint4 bv[4];
        auto load_b = [&](int k0) {
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                const int idx = t + NT * it;
                const int arr = idx >> 9, rem = idx & 511;
                const int col = rem >> 2, q = rem & 3;
                const __bf16* src = (arr ? Bl : Bh) + (long)(n0 + col) * g.ldb + k0 + 8 * q;
                bv[it] = *reinterpret_cast<const uint4*>(src);
            }
        };
        load_a<AMC, VEC>(g, A, m0, 0, av);
        load_b(0);
        for (int kt = 0; kt < nk; ++kt) {
            __syncthreads();
            {
                const int mq = t & 31, kq = t >> 5;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    bf16x4 h, l;
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        __bf16 hh, ll;
                        split_bf16(av[i][kk], hh, ll);
                        h[kk] = hh;
                        l[kk] = ll;
                    }
                    *reinterpret_cast<bf16x4*>(&sAh[4 * mq + i][4 * kq]) = h;
                    *reinterpret_cast<bf16x4*>(&sAl[4 * mq + i][4 * kq]) = l;
                }
#pragma unroll
                for (int it = 0; it < 4; ++it) {
                    const int idx = t + NT * it;
                    const int arr = idx >> 9, rem = idx & 511;
                    const int col = rem >> 2, q = rem & 3;
                    __bf16* dst = arr ? &sBl[col][8 * q] : &sBh[col][8 * q];
                    *reinterpret_cast<uint4*>(dst) = bv[it];
                }
            }
            __syncthreads();
            if (kt + 1 < nk) {
                load_a<AMC, VEC>(g, A, m0, (kt + 1) * BK, av);
                load_b((kt + 1) * BK);
            }
#pragma unroll
            for (int s = 0; s < BK / 16; ++s) {
                const int kof = s * 16 + (lane >> 5) * 8;
                bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const int r = wm0 + i * 32 + (lane & 31);
                    ah[i] = *reinterpret_cast<const bf16x8*>(&sAh[r][kof]);
                    al[i] = *reinterpret_cast<const bf16x8*>(&sAl[r][kof]);
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int c = wn0 + j * 32 + (lane & 31);
                    bh[j] = *reinterpret_cast<const bf16x8*>(&sBh[c][kof]);
                    bl[j] = *reinterpret_cast<const bf16x8*>(&sBl[c][kof]);
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                    }
            }
        }
    } else {
        __shared__ __attribute__((aligned(16))) float sA[BM][BK + PADF];
        __shared__ __attribute__((aligned(16))) float sB[BN][BK + PADF];
        const float* __restrict__ Bf = static_cast<const float*>(g.Bhi);
        float4 bv[4];
        auto load_b = [&](int k0) {
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                const int idx = t + NT * it;
                const int col = idx >> 3, q = idx & 7;
                bv[it] = *reinterpret_cast<const float4*>(Bf + (long)(n0 + col) * g.ldb + k0 + 4 * q);
            }
        };
        load_a<AMC, VEC>(g, A, m0, 0, av);
        load_b(0);
        for (int kt = 0; kt < nk; ++kt) {
            __syncthreads();
            {
                const int mq = t & 31, kq = t >> 5;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    *reinterpret_cast<float4*>(&sA[4 * mq + i][4 * kq]) =
                        make_float4(av[i][0], av[i][1], av[i][2], av[i][3]);
#pragma unroll
                for (int it = 0; it < 4; ++it) {
                    const int idx = t + NT * it;
                    const int col = idx >> 3, q = idx & 7;
                    *reinterpret_cast<float4*>(&sB[col][4 * q]) = bv[it];
                }
            }
            __syncthreads();
            if (kt + 1 < nk) {
                load_a<AMC, VEC>(g, A, m0, (kt + 1) * BK, av);
                load_b((kt + 1) * BK);
            }
            // k is permuted inside the tile: MFMA step s, lane group h = lane/32 uses
            // k = 16 h + s, so each lane reads 16 contiguous floats.
            const int kof = (lane >> 5) * 16;
            float fa[2][16], fb[2][16];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int r = wm0 + i * 32 + (lane & 31);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 x = *reinterpret_cast<const float4*>(&sA[r][kof + 4 * q]);
                    fa[i][4 * q] = x.x; fa[i][4 * q + 1] = x.y; fa[i][4 * q + 2] = x.z; fa[i][4 * q + 3] = x.w;
                }
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int c = wn0 + j * 32 + (lane & 31);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 x = *reinterpret_cast<const float4*>(&sB[c][kof + 4 * q]);
                    fb[j][4 * q] = x.x; fb[j][4 * q + 1] = x.y; fb[j][4 * q + 2] = x.z; fb[j][4 * q + 3] = x.w;
                }
            }
#pragma unroll
            for (int s = 0; s < 16; ++s)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i][s], fb[j][s], acc[i][j], 0, 0, 0);
        }
    }

    // Epilogue: 32x32 accumulator layout: register r of lane l holds row
    // 8*(r/4) + 4*(l/32) + r%4, column l%32.
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn0 + j * 32 + (lane & 31);
            if (col >= g.N) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = m0 + wm0 + i * 32 + q * 8 + (lane >> 5) * 4;
                float* dst = C + row + (long)col * g.scn;
                if (VEC && row + 3 < g.M) {
                    *reinterpret_cast<float4*>(dst) =
                        make_float4(acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2],
                                    acc[i][j][4 * q + 3]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (row + r < g.M) dst[r] = acc[i][j][4 * q + r];
                }
            }
        }
}

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
    const size_t e = (size_t)round_up(n, BK) * round_up(n, BN);
    return prec == FC3_PREC_FP32 ? Carve::need(e * 4) : 2 * Carve::need(e * 2);
}

int prep_phi(hipStream_t s, const float* phi, int n, int prec, Carve& cv, PhiPrep& out)
{
    const int Kpad = round_up(n, BK), Npad = round_up(n, BN);
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

// C[b] (M x N) = A[b] (M x K) * phi_padded (K x N), batched.
int gemm(hipStream_t s, int kind, int prec, const float* A, long sam, long sak, long sab, const PhiPrep& B,
         float* C, long scn, long scb, int M, int N, int K, int batch)
{
    GemmArgs g;
    g.A = A; g.sam = sam; g.sak = sak; g.sab = sab;
    g.Bhi = B.hi; g.Blo = B.lo; g.ldb = B.ldb;
    g.C = C; g.scn = scn; g.scb = scb;
    g.M = M; g.N = N; g.K = K; g.batch = batch;
    g.tiles_m = (M + BM - 1) / BM;
    g.tiles_n = (N + BN - 1) / BN;
    const long tiles = (long)g.tiles_m * g.tiles_n * batch;
    g.per_xcd = (tiles + kXcd - 1) / kXcd;
    const long grid = g.per_xcd * kXcd;
    const bool amc = (sam == 1);
    // 16-byte vector access needs every stride and base aligned to 4 floats.
    const bool vec = ((amc ? sak : sam) % 4 == 0) && (sab % 4 == 0) && (scn % 4 == 0) &&
                     (scb % 4 == 0) && ((uintptr_t)A % 16 == 0) && ((uintptr_t)C % 16 == 0);
    if (grid > 0x7fffffffL) return set_error(FC3_ERR_UNSUPPORTED, "mcc3: grid too large");
    dim3 gd((unsigned)grid), bd(NT);
    ProfScope prof(kind, s);
#define FC3_LAUNCH(P, AM, V) hipLaunchKernelGGL((k_gemm<P, AM, V>), gd, bd, 0, s, g)
    if (prec == FC3_PREC_BF16X3) {
        if (amc) { if (vec) FC3_LAUNCH(FC3_PREC_BF16X3, true, true); else FC3_LAUNCH(FC3_PREC_BF16X3, true, false); }
        else     { if (vec) FC3_LAUNCH(FC3_PREC_BF16X3, false, true); else FC3_LAUNCH(FC3_PREC_BF16X3, false, false); }
    } else {
        if (amc) { if (vec) FC3_LAUNCH(FC3_PREC_FP32, true, true); else FC3_LAUNCH(FC3_PREC_FP32, true, false); }
        else     { if (vec) FC3_LAUNCH(FC3_PREC_FP32, false, true); else FC3_LAUNCH(FC3_PREC_FP32, false, false); }
    }
#undef FC3_LAUNCH
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

// Slabs per batched mode-1/2 launch: enough tiles to fill the chip several times.
int slab_batch(int n)
{
    const long tiles = (long)((n + BM - 1) / BM) * ((n + BN - 1) / BN);
    long b = (2048 + tiles - 1) / tiles;
    if (b < 1) b = 1;
    if (b > n) b = n;
    return (int)b;
}

// Modes 1 and 2 on slabs [l0, l0+cnt) of T (each n x n, stride n^2), in place.
int modes12(hipStream_t s, int prec, float* T, int n, int l0, int cnt, const PhiPrep& B,
            float* S)
{
    const long n2 = (long)n * n;
    // S_l[j, n'] = sum_i T_l[i, j] phi[i, n']   (A[m=j, k=i] = T_l[i + n j])
    FC3_TRY(gemm(s, FC3_KERNEL_GEMM_MODE12, prec, T + l0 * n2, n, 1, n2, B, S, n, n2, n, n, n, cnt));
    // T_l[n', m'] = sum_j S_l[j, n'] phi[j, m']  (A[m=n', k=j] = S_l[j + n n'])
    FC3_TRY(gemm(s, FC3_KERNEL_GEMM_MODE12, prec, S, n, 1, n2, B, T + l0 * n2, n, n2, n, n, n, cnt));
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
    FC3_TRY(gemm(s, FC3_KERNEL_GEMM_MODE3, precision, psi, 1, n2, 0, B, k3, n2, 0, (int)n2, n, n, 1));
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
        FC3_TRY(gemm(s, FC3_KERNEL_GEMM_MODE3, precision, psi + r0, 1, n2, 0, B, W, rows, 0, (int)rows, n, n, 1));
        FC3_HIP(hipMemcpy2DAsync(psi + r0, sizeof(float) * n2, W, sizeof(float) * rows,
                                 sizeof(float) * rows, n, hipMemcpyDeviceToDevice, s));
    }
    return FC3_OK;
}

}  // namespace fc3
