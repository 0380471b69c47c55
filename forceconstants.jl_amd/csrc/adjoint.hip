// Hand-written adjoints (reverse mode) of the two hot-path operations.
//
// mcc3 (src/mcc.jl:15-17): K3[n,m,l] = sum_ijk Psi[i,j,k] phi[i,n] phi[j,m] phi[k,l].
// For a cotangent G = dL/dK3:
//   dL/dPsi = G x1 phi x2 phi x3 phi         (the same contraction with phi transposed)
//   dL/dphi = T1 + T2 + T3, one term per mode, each a contraction of Psi with G after the
//             other two modes of G are mapped back by phi:
//     T1[a,b] = sum_jk Psi[a,j,k] H23[b,j,k],  H23 = G x2 phi x3 phi
//     T2[a,b] = sum_ik Psi[i,a,k] H13[i,b,k],  H13 = G x1 phi x3 phi
//     T3[a,b] = sum_ij Psi[i,j,a] H12[i,j,b],  H12 = G x1 phi x2 phi
//   (x_q phi maps index q of G to index i with weight phi[i, .]).  G x3 phi is shared by
//   H23 and H13; H12 also gives dL/dPsi = H12 x3 phi.  18 n^4 flops, two n^3 buffers.
//   For BF16X3 (n % 4 == 0) the mode products and the n x n^2 x n contractions run on
//   the BF16x3 mcc3 kernel (the contractions with split K); otherwise they are FP32
//   GEMMs (rocBLAS, column-major like the tensors).  dL/dPsi alone runs on the
//   library's mcc3 kernels.
//
// LJ parameters (third_order_analytical.jl:4-69, asr.jl:19-40): every entry of Psi is
// +-E_JK[t] of one pair block or a self block S_J = -sum_K E_JK, so dL/dtheta =
// sum over pair blocks of dE_JK[t]/dtheta times the cotangent gathered from the (at most
// four) positions that hold E_JK[t]: (3J+u, 3J+v, 3K+w) with +1, (3K+w, 3J+u, 3K+v) with
// -1 (ASR entry of row atom K), (3K+w, 3J+u, 3J+v) with +1 (self column of J), and the
// self entry (3J+u, 3J+v, 3J+w) with -1, for t = u + 3v + 9w.  apply_tols! zeroes an
// entry (and its derivative) when |value| < tol.  The blocks are recomputed from the
// positions, so no dense derivative tensor is stored.  E is linear in epsilon and in
// (sigma^12, sigma^6), so dE/dsigma is E evaluated with (12 sigma^11, 6 sigma^5).
#include <rocblas/rocblas.h>

#include "fc3_internal.hpp"
#include "geom.hpp"

namespace fc3 {
namespace {

#define FC3_BLAS(expr)                                                                     \
    do {                                                                                   \
        const rocblas_status st_ = (expr);                                                 \
        if (st_ != rocblas_status_success)                                                 \
            return ::fc3::set_error(FC3_ERR_HIP, std::string("rocBLAS: ") +                \
                                                     rocblas_status_to_string(st_));       \
    } while (0)

// out[j + n i] = in[i + n j]
__global__ void k_transpose(const float* __restrict__ in, int n, float* __restrict__ out)
{
    __shared__ float t[32][33];
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    for (int y = threadIdx.y; y < 32; y += 8) {
        const int i = i0 + threadIdx.x, j = j0 + y;
        if (i < n && j < n) t[y][threadIdx.x] = in[i + (long)n * j];
    }
    __syncthreads();
    for (int y = threadIdx.y; y < 32; y += 8) {
        const int j = j0 + threadIdx.x, i = i0 + y;
        if (i < n && j < n) out[j + (long)n * i] = t[threadIdx.x][y];
    }
}

// out[e] (+)= sum_k T[e + n2 k], k ascending, for e < n2 (accumulate or overwrite)
__global__ void k_sum_slabs(const float* __restrict__ T, long n2, int cnt, float* __restrict__ out,
                            int accumulate = 1)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n2) return;
    float acc = 0.0f;
    for (int k = 0; k < cnt; ++k) acc += T[e + n2 * k];
    out[e] = accumulate ? out[e] + acc : acc;
}

// One thread per (atom J, neighbour slot s): accumulates sum_t dE_JK[t]/dsigma * W[t]
// and sum_t E_JK[t]/epsilon * W[t] into per-block partials (deterministic order).
template <typename DT>
__global__ void __launch_bounds__(256)
k_lj_param_grad(const double* __restrict__ pos, double Lx, double Ly, double Lz, double s12,
                double s6, double sigma, double eps, const int* __restrict__ count,
                const int* __restrict__ nbr, int N, int cap, const double* __restrict__ E3,
                const double* __restrict__ S3, double tol, const double* __restrict__ masses,
                const DT* __restrict__ dpsi, double* __restrict__ partial)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double gs = 0.0, ge = 0.0;
    if (t < N * cap) {
        const int J = t / cap, s = t - J * cap;
        if (s < count[J]) {
            const int K = nbr[t];
            const double L[3] = {Lx, Ly, Lz};
            double r[3] = {pos[3 * J] - pos[3 * K], pos[3 * J + 1] - pos[3 * K + 1],
                           pos[3 * J + 2] - pos[3 * K + 2]};
            geom::mirror(r, L);
            const double dist = sqrt(geom::sumsq(r));
            // d/dsigma of (s12, s6) = (12 s12 / sigma, 6 s6 / sigma); E is linear in both
            const geom::LJRadial fs =
                geom::lj_radial(12.0 * s12 / sigma, 6.0 * s6 / sigma, eps, dist, true);
            const long n = 3L * N;
            const long n2 = n * n;
            double cJJK = 1.0, cKJK = 1.0, cKJJ = 1.0, cJJJ = 1.0;
            if (masses) {
                const double mJ = masses[J], mK = masses[K];
                cJJK = 1.0 / sqrt((mJ * mJ) * mK);
                cKJK = 1.0 / sqrt((mK * mJ) * mK);
                cKJJ = 1.0 / sqrt((mK * mJ) * mJ);
                cJJJ = 1.0 / sqrt((mJ * mJ) * mJ);
            }
            const double* e = E3 + (size_t)t * 27;
            const double* sj = S3 + (size_t)J * 27;
            for (int w = 0; w < 3; ++w)
                for (int v = 0; v < 3; ++v)
                    for (int u = 0; u < 3; ++u) {
                        const int q = u + 3 * v + 9 * w;
                        double W = 0.0;
                        if (fabs(e[q]) >= tol) {
                            W += cJJK * (double)dpsi[(3L * J + u) + n * (3L * J + v) + n2 * (3L * K + w)];
                            W -= cKJK * (double)dpsi[(3L * K + w) + n * (3L * J + u) + n2 * (3L * K + v)];
                            W += cKJJ * (double)dpsi[(3L * K + w) + n * (3L * J + u) + n2 * (3L * J + v)];
                        }
                        if (fabs(sj[q]) >= tol)
                            W -= cJJJ * (double)dpsi[(3L * J + u) + n * (3L * J + v) + n2 * (3L * J + w)];
                        if (W == 0.0) continue;
                        const int hi = u > v ? u : v, lo = u > v ? v : u;
                        const double ds = (J < K) ? -geom::phi3(fs, dist, r, hi, lo, w)
                                                  : -geom::phi3(fs, dist, r, w, hi, lo);
                        gs += ds * W;
                        ge += (e[q] / eps) * W;
                    }
        }
    }
    // block reduction in a fixed order
    __shared__ double rs[256], re[256];
    rs[threadIdx.x] = gs;
    re[threadIdx.x] = ge;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            rs[threadIdx.x] += rs[threadIdx.x + h];
            re[threadIdx.x] += re[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = rs[0];
        partial[2 * blockIdx.x + 1] = re[0];
    }
}

// ---- weighted square loss L = sum_e w_e K3_e^2 (the differentiation workload) -------
// w_e ~ U[0, 1) from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC11): element e takes
// word e % 4 of the block for counter (e / 4 low word, e / 4 high word, 0, 0) under key
// (seed low word, seed high word); w_e = (word >> 8) 2^-24, exact in FP32.  Every term
// w K3^2 is formed in FP64; thread sums run over a fixed grid stride and the block sums
// combine in a fixed tree, so L is the same on every run.  The cotangent 2 w K3 is one
// FP32 rounding of the exact (2 w) K3.
__device__ __forceinline__ void philox4x32_10(unsigned c[4], unsigned k0, unsigned k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;   // Weyl key increments
            k1 += 0xBB67AE85u;
        }
        const unsigned long long p0 = 0xD2511F53ull * c[0];   // round multipliers
        const unsigned long long p1 = 0xCD9E8D57ull * c[2];
        const unsigned n0 = (unsigned)(p1 >> 32) ^ c[1] ^ k0;
        const unsigned n2 = (unsigned)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (unsigned)p1;
        c[2] = n2;
        c[3] = (unsigned)p0;
    }
}

constexpr int kLossBlocks = 2048;

__global__ void __launch_bounds__(256)
k_weighted_square_loss(const float* __restrict__ k3, long count, unsigned long long seed,
                       float* __restrict__ dk3, double* __restrict__ partial)
{
    double acc = 0.0;
    const long nq = (count + 3) / 4, nfull = count / 4, S = (long)gridDim.x * blockDim.x;
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // four grid-stride steps at a time with their loads issued first (one 16-byte load in
    // flight per thread leaves HBM half idle); the same q order per thread, so the same
    // sums as the plain loop below
    for (; q + 3 * S < nfull; q += 4 * S) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = reinterpret_cast<const float4*>(k3)[q + u * S];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long qq = q + u * S;
            unsigned c[4] = {(unsigned)qq, (unsigned)((unsigned long long)qq >> 32), 0u, 0u};
            philox4x32_10(c, (unsigned)seed, (unsigned)(seed >> 32));
            const float x[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
            float g[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float w = (float)(c[j] >> 8) * 5.9604644775390625e-8f;   // 2^-24
                acc += (double)w * (double)x[j] * (double)x[j];
                g[j] = (2.0f * w) * x[j];
            }
            if (dk3) reinterpret_cast<float4*>(dk3)[qq] = make_float4(g[0], g[1], g[2], g[3]);
        }
    }
    for (; q < nq; q += S) {
        unsigned c[4] = {(unsigned)q, (unsigned)((unsigned long long)q >> 32), 0u, 0u};
        philox4x32_10(c, (unsigned)seed, (unsigned)(seed >> 32));
        const long e0 = 4 * q;
        float x[4];
        if (e0 + 3 < count) {
            const float4 v = reinterpret_cast<const float4*>(k3)[q];   // k3 16-byte aligned
            x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
        } else {
            for (int j = 0; j < 4; ++j) x[j] = e0 + j < count ? k3[e0 + j] : 0.0f;
        }
        float g[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float w = (float)(c[j] >> 8) * 5.9604644775390625e-8f;   // 2^-24
            if (e0 + j < count) acc += (double)w * (double)x[j] * (double)x[j];
            g[j] = (2.0f * w) * x[j];
        }
        if (dk3) {
            if (e0 + 3 < count) {
                reinterpret_cast<float4*>(dk3)[q] = make_float4(g[0], g[1], g[2], g[3]);
            } else {
                for (int j = 0; j < 4; ++j)
                    if (e0 + j < count) dk3[e0 + j] = g[j];
            }
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_sum_in_order(const double* __restrict__ partial, int cnt, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0;
    for (int i = 0; i < cnt; ++i) a += partial[i];
    out[0] = a;
}

__global__ void k_sum_partials(const double* __restrict__ partial, int cnt, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int i = 0; i < cnt; ++i) {
        a += partial[2 * i];
        b += partial[2 * i + 1];
    }
    out[0] = a;
    out[1] = b;
}

}  // namespace

int transpose_square(hipStream_t s, const float* in, int n, float* out)
{
    hipLaunchKernelGGL(k_transpose, dim3((n + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, s, in, n,
                       out);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

void release_blas(fc3_context* h)
{
    if (h && h->blas) {
        rocblas_destroy_handle(static_cast<rocblas_handle>(h->blas));
        h->blas = nullptr;
    }
}

static int get_blas(fc3_context* h, rocblas_handle* out)
{
    if (!h->blas) {
        rocblas_handle b = nullptr;
        FC3_BLAS(rocblas_create_handle(&b));
        h->blas = b;
        // no kernels that sum with atomics (rocBLAS's default, stated here): the FP32
        // pullback and the eigensolver return the same bits on every run
        FC3_BLAS(rocblas_set_atomics_mode(b, rocblas_atomics_not_allowed));
    }
    rocblas_handle b = static_cast<rocblas_handle>(h->blas);
    FC3_BLAS(rocblas_set_stream(b, h->stream));
    FC3_BLAS(rocblas_set_pointer_mode(b, rocblas_pointer_mode_host));
    *out = b;
    return FC3_OK;
}

int blas_handle(fc3_context* h, void** out)
{
    rocblas_handle b = nullptr;
    FC3_TRY(get_blas(h, &b));
    *out = b;
    return FC3_OK;
}

size_t weighted_square_loss_bytes() { return Carve::need(sizeof(double) * (kLossBlocks + 1)); }

int weighted_square_loss(hipStream_t s, const float* k3, long count, unsigned long long seed,
                         float* dk3, void* ws, double* d_loss)
{
    Carve cv(ws, weighted_square_loss_bytes());
    double* partial = cv.take<double>(kLossBlocks + 1);
    const long nq = (count + 3) / 4;
    const long blocks_needed = (nq + 255) / 256;
    const int grid = (int)(blocks_needed < kLossBlocks ? (blocks_needed > 0 ? blocks_needed : 1)
                                                       : kLossBlocks);
    hipLaunchKernelGGL(k_weighted_square_loss, dim3(grid), dim3(256), 0, s, k3, count, seed, dk3,
                       partial);
    hipLaunchKernelGGL(k_sum_in_order, dim3(1), dim3(64), 0, s, partial, grid, d_loss);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

size_t mcc3_backward_workspace_bytes(int n, int precision, bool need_dphi)
{
    const size_t n3 = (size_t)n * n * n;
    if (need_dphi) {
        size_t b = 2 * Carve::need(sizeof(float) * n3) + 4096;
        if (precision == FC3_PREC_BF16X3 && mcc3_adjoint_v3_applies(n, nullptr, nullptr, nullptr))
            b += Carve::need(sizeof(float) * (size_t)n * n) +
                 std::max(mcc3_adjoint_v3_bytes(n),
                          mcc3_contract_applies(n) ? mcc3_contract_bytes(n) : size_t(0));
        return b;
    }
    return Carve::need(sizeof(float) * (size_t)n * n) + mcc3_workspace_bytes(n, precision, false) +
           4096;
}

int mcc3_backward(fc3_context* h, const float* psi, const float* phi, const float* G, int n,
                  int precision, float* dpsi, float* dphi)
{
    hipStream_t s = h->stream;
    const long n2 = (long)n * n;
    Carve cv(h->ws, h->ws_bytes);
    if (!dphi) {
        // dL/dPsi alone: mcc3 of G with phi^T on the library's kernels (either precision)
        float* phiT = cv.take<float>((size_t)n2);
        hipLaunchKernelGGL(k_transpose, dim3((n + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, s, phi,
                           n, phiT);
        FC3_HIP(hipGetLastError());
        char* rest = cv.take<char>(0);
        if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 backward workspace");
        return mcc3_run(s, G, phiT, n, precision, dpsi, rest, cv.cap - cv.off);
    }
    rocblas_handle b;
    FC3_TRY(get_blas(h, &b));
    float* T1 = cv.take<float>((size_t)n2 * n);
    float* T2 = cv.take<float>((size_t)n2 * n);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 backward workspace");
    const float one = 1.0f, zero = 0.0f;
    const rocblas_operation N_ = rocblas_operation_none, T_ = rocblas_operation_transpose;
    const bool v3 = precision == FC3_PREC_BF16X3 && mcc3_adjoint_v3_applies(n, psi, G, dpsi ? dpsi : G);
    const bool cv3 = v3 && mcc3_contract_applies(n);
    float* phiT = nullptr;
    char* rest = nullptr;
    size_t rest_bytes = 0;
    const unsigned sum_grid = (unsigned)((n2 + 255) / 256);
    if (v3) {
        // The mode products on the BF16x3 kernels (M = phi^T maps an index back); the
        // three n x n^2 x n contractions with Psi on the BF16x3 kernel with split K, or
        // as FP32 GEMMs where that does not apply.
        phiT = cv.take<float>((size_t)n2);
        rest = cv.take<char>(0);
        if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 backward workspace");
        rest_bytes = cv.cap - cv.off;
        hipLaunchKernelGGL(k_transpose, dim3((n + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, s, phi,
                           n, phiT);
        FC3_HIP(hipGetLastError());
        int zc = 0;
        // H23 = G x2 phi x3 phi: modes31_t of the slab-transposed G (G's first two
        // indices swapped) writes H23 in its natural layout; its mode-3 split reads G in
        // that order directly
        FC3_TRY(mcc3_modes31_t(s, G, phiT, n, T2, rest, rest_bytes, true));
        // dphi = Psi_(1) * H23_(1)^T
        if (cv3) {
            FC3_TRY(mcc3_contract(s, 0, psi, T2, n, T1, &zc, rest, rest_bytes));
            hipLaunchKernelGGL(k_sum_slabs, dim3(sum_grid), dim3(256), 0, s, T1, n2, zc, dphi, 0);
            FC3_HIP(hipGetLastError());
        } else {
            FC3_BLAS(rocblas_sgemm(b, N_, T_, n, n, (int)n2, &one, psi, n, T2, n, &zero, dphi, n));
        }
        // Y = H13 with its first two indices swapped (Y_k = H13_k^T):
        // dphi(a, b) += sum_{i,k} Psi(i, a, k) Y(b, i, k), the first form with both
        // slab-transposed; the contraction's split reads Psi in that order directly
        FC3_TRY(mcc3_modes31_t(s, G, phiT, n, T2, rest, rest_bytes));
        if (cv3) {
            FC3_TRY(mcc3_contract(s, 2, psi, T2, n, T1, &zc, rest, rest_bytes));
        } else {
            FC3_BLAS(rocblas_sgemm_strided_batched(b, T_, T_, n, n, n, &one, psi, n, n2, T2, n,
                                                   n2, &zero, T1, n, n2, n));
            zc = n;
        }
        hipLaunchKernelGGL(k_sum_slabs, dim3(sum_grid), dim3(256), 0, s, T1, n2, zc, dphi, 1);
        FC3_HIP(hipGetLastError());
        // T2 = H12 = G x1 phi x2 phi; dphi += Psi_(3)^T H12_(3)
        FC3_TRY(mcc3_modes12(s, G, phiT, n, T2, rest, rest_bytes));
        if (cv3) {
            FC3_TRY(mcc3_contract(s, 1, psi, T2, n, T1, &zc, rest, rest_bytes));
            hipLaunchKernelGGL(k_sum_slabs, dim3(sum_grid), dim3(256), 0, s, T1, n2, zc, dphi, 1);
            FC3_HIP(hipGetLastError());
        }
    } else {
        // 1. T1 = G x3 phi:  (n^2 x n) = G (n^2 x n) * phi^T
        FC3_BLAS(rocblas_sgemm(b, N_, T_, (int)n2, n, n, &one, G, (int)n2, phi, n, &zero, T1,
                               (int)n2));
        // 2. T2 = H23, slab k: T1_k * phi^T
        FC3_BLAS(rocblas_sgemm_strided_batched(b, N_, T_, n, n, n, &one, T1, n, n2, phi, n, 0,
                                               &zero, T2, n, n2, n));
        // 3. dphi = Psi_(1) * H23_(1)^T   (contraction over n^2)
        FC3_BLAS(rocblas_sgemm(b, N_, T_, n, n, (int)n2, &one, psi, n, T2, n, &zero, dphi, n));
        // 4. T2 = H13, slab k: phi * T1_k
        FC3_BLAS(rocblas_sgemm_strided_batched(b, N_, N_, n, n, n, &one, phi, n, 0, T1, n, n2,
                                               &zero, T2, n, n2, n));
        // 5. dphi += sum_k Psi_k^T H13_k  (per-slab products into T1, then a slab sum)
        FC3_BLAS(rocblas_sgemm_strided_batched(b, T_, N_, n, n, n, &one, psi, n, n2, T2, n, n2,
                                               &zero, T1, n, n2, n));
        hipLaunchKernelGGL(k_sum_slabs, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, T1,
                           n2, n, dphi);
        FC3_HIP(hipGetLastError());
        // 6. T2 = H12, slab l: (phi * G_l) * phi^T
        FC3_BLAS(rocblas_sgemm_strided_batched(b, N_, N_, n, n, n, &one, phi, n, 0, G, n, n2,
                                               &zero, T1, n, n2, n));
        FC3_BLAS(rocblas_sgemm_strided_batched(b, N_, T_, n, n, n, &one, T1, n, n2, phi, n, 0,
                                               &zero, T2, n, n2, n));
    }
    // 7. dphi += Psi_(3)^T H12_(3)   (contraction over n^2).  As one GEMM this is only
    //    (n/256)^2 output tiles over a K of n^2, which rocBLAS runs at a small fraction of
    //    the MFMA rate; instead K is cut into S chunks whose n x n products go to slabs of
    //    T1 (free after step 6), summed in chunk order.
    if (!cv3) {
        // S + 1 slabs must fit T1's n slabs
        const int S = n <= 2 ? 1 : (n - 1 < 64 ? n - 1 : 64);
        const long L = n2 / S, rem = n2 - (long)S * L;
        FC3_BLAS(rocblas_sgemm_strided_batched(b, T_, N_, n, n, (int)L, &one, psi, (int)n2, L, T2,
                                               (int)n2, L, &zero, T1, n, n2, S));
        if (rem)
            FC3_BLAS(rocblas_sgemm(b, T_, N_, n, n, (int)rem, &one, psi + (long)S * L, (int)n2,
                                   T2 + (long)S * L, (int)n2, &zero, T1 + (long)S * n2, n));
        hipLaunchKernelGGL(k_sum_slabs, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, T1, n2,
                           S + (rem ? 1 : 0), dphi);
        FC3_HIP(hipGetLastError());
    }
    // 8. dpsi = H12 x3 phi
    if (dpsi && v3) return mcc3_mode3(s, T2, phiT, n, dpsi, rest, rest_bytes);
    if (dpsi)
        FC3_BLAS(rocblas_sgemm(b, N_, T_, (int)n2, n, n, &one, T2, (int)n2, phi, n, &zero, dpsi,
                               (int)n2));
    return FC3_OK;
}

int lj_param_grad(hipStream_t s, const double* d_pos, const double box[3], double sigma,
                  double epsilon, const NeighborList& nl, const double* E3, const double* S3,
                  double tol, const double* d_masses, const float* dpsi32, const double* dpsi64,
                  double* partial, double* d_out)
{
    const int N = nl.n_atoms;
    const long tot = (long)N * nl.cap;
    const int blocks = (int)((tot + 255) / 256);
    // sigma^12, sigma^6 exactly as the forward pass evaluates them
    double s12 = 0.0, s6 = 0.0;
    lj_sigma_powers(sigma, &s12, &s6);
    if (dpsi32)
        hipLaunchKernelGGL(k_lj_param_grad<float>, dim3(blocks), dim3(256), 0, s, d_pos, box[0],
                           box[1], box[2], s12, s6, sigma, epsilon, nl.count, nl.nbr, N, nl.cap,
                           E3, S3, tol, d_masses, dpsi32, partial);
    else
        hipLaunchKernelGGL(k_lj_param_grad<double>, dim3(blocks), dim3(256), 0, s, d_pos, box[0],
                           box[1], box[2], s12, s6, sigma, epsilon, nl.count, nl.nbr, N, nl.cap,
                           E3, S3, tol, nullptr, dpsi64, partial);
    FC3_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, partial, blocks, d_out);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

}  // namespace fc3
