// Stillinger-Weber force constants on gfx950 (FP64).
//
// The reference's AutoDiffCalculator loops (third_order_AD.jl:71-170,
// second_order_AD.jl:39-111) add, for every centre i and neighbour slot j (all j != i in
// range), a pair term when j > i and a triplet term for every k > j in range, each as a
// few permuted 3x3x3 (3x3) blocks added into the dense tensor.  Here:
//   1. one thread per (centre i, slot j) evaluates its pair term and triplet terms in
//      closed form (sw_closed.hpp) and emits one record (block key, 27 or 9 values) per
//      permuted block the reference would add, numbered in the reference's loop order;
//   2. a stable radix sort on the key groups the records of each output block while
//      keeping that order, so the segmented sum adds contributions in the same
//      sequence as the reference's `.+=`;
//   3. ASR! is applied on the block list (a second sort by (first atom, second atom)
//      gives each row of blocks), then a column writer streams the dense tensor with
//      apply_tols!, mass weighting and the FP32 copy fused.
// Block keys order a block (A, B, C) by (C, B, A) so that the blocks of one output
// column (B, C) are contiguous and sorted by row atom.
#include <hipcub/hipcub.hpp>

#include <vector>

#include "fc3_internal.hpp"
#include "geom.hpp"
#include "sw_closed.hpp"

namespace fc3 {
namespace {

typedef unsigned long long u64;

__device__ __forceinline__ int first_above(const int* list, int cnt, int key)
{
    int lo = 0, hi = cnt;  // first slot with list[slot] > key
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (list[mid] > key) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__global__ void k_sw_counts(const int* __restrict__ count, const int* __restrict__ nbr, int N,
                            int cap, int* __restrict__ rec3, int* __restrict__ rec2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int d = count[i];
    const int f = first_above(nbr + (size_t)i * cap, d, i);
    const int npair = d - f;
    const int ntrip = d * (d - 1) / 2;
    rec3[i] = 6 * npair + 24 * ntrip;
    rec2[i] = 2 * npair + 6 * ntrip;
}

// permutedims semantics: dst[x1,x2,x3] = src[j] with j[perm[d]] = x_d (perm 1-based)
__device__ __forceinline__ void permute27(const double* src, int p1, int p2, int p3, double sgn,
                                          double* dst)
{
    const int pm[3] = {p1 - 1, p2 - 1, p3 - 1};
    for (int x3 = 0; x3 < 3; ++x3)
        for (int x2 = 0; x2 < 3; ++x2)
            for (int x1 = 0; x1 < 3; ++x1) {
                int j[3];
                j[pm[0]] = x1;
                j[pm[1]] = x2;
                j[pm[2]] = x3;
                dst[x1 + 3 * x2 + 9 * x3] = sgn * src[j[0] + 3 * j[1] + 9 * j[2]];
            }
}

struct EmitCtx {
    const double* pos;
    const int* count;
    const int* nbr;
    int N, cap;
    double L[3];
    sw::Params P;
    double c0_fd, c0_iik;
    const int* off3;   // per-atom record offsets (IFC3), or null
    const int* off2;   // per-atom record offsets (IFC2), or null
    u64* key3;
    double* val3;
    u64* key2;
    double* val2;
};

__device__ __forceinline__ u64 key3of(int A, int B, int C, int N)
{
    return ((u64)C * N + B) * N + A;
}

// One thread per (centre i, slot s, slot u) with u >= s.  Centre i's records lie at
// off3[i] (off2[i]) in slot order: slot s's pair term (6 IFC3 / 2 IFC2 records, when nb[s]
// > i), then its triplets (s, tt) for tt = s + 1 .. d - 1 (24 / 6 records each).  Thread
// (i, s, s) writes the pair term and thread (i, s, u > s) the triplet (s, u), each at its
// place in that order, so the records are those of a sequential loop over the slots.
__global__ void __launch_bounds__(64) k_sw_emit(EmitCtx c)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)c.N * c.cap * c.cap) return;
    const int i = (int)(t / ((long)c.cap * c.cap));
    const int su = (int)(t - (long)i * c.cap * c.cap);
    const int s = su / c.cap, u = su - s * c.cap;
    const int d = c.count[i];
    if (s >= d || u >= d || u < s) return;
    const int* nb = c.nbr + (size_t)i * c.cap;
    const int N = c.N;
    const int f = first_above(nb, d, i);
    const int j = nb[s];
    // records before this thread's: the pair terms of slots < s (and of s itself for a
    // triplet thread), the triplets of slots < s, and slot s's triplets before u
    const int pairs_before = (s > f ? s - f : 0) + (u > s && j > i ? 1 : 0);
    const int trips_before = s * (d - 1) - s * (s - 1) / 2 + (u > s ? u - s - 1 : 0);
    long r3 = c.off3 ? c.off3[i] + 6L * pairs_before + 24L * trips_before : 0;
    long r2 = c.off2 ? c.off2[i] + 2L * pairs_before + 6L * trips_before : 0;
    const double ri[3] = {c.pos[3 * i], c.pos[3 * i + 1], c.pos[3 * i + 2]};
    double rij[3] = {ri[0] - c.pos[3 * j], ri[1] - c.pos[3 * j + 1], ri[2] - c.pos[3 * j + 2]};
    geom::mirror(rij, c.L);

    double blk[27];
    if (u == s && j > i) {
        // pair term: (i,i,j),(i,j,i),(j,i,i) get -H3, (j,j,i),(j,i,j),(i,j,j) get +H3;
        // IFC2 (i,j) and (j,i) get -H2 (third_order_AD.jl:103-111, second_order_AD.jl:66-70)
        double h2[9];
        sw::pair_blocks(rij, c.P, c.off2 ? h2 : nullptr, c.off3 ? blk : nullptr);
        if (c.off3) {
            const int ks[6][3] = {{i, i, j}, {i, j, i}, {j, i, i}, {j, j, i}, {j, i, j}, {i, j, j}};
            for (int q = 0; q < 6; ++q) {
                const double sg = q < 3 ? -1.0 : 1.0;
                c.key3[r3] = key3of(ks[q][0], ks[q][1], ks[q][2], N);
                for (int e = 0; e < 27; ++e) c.val3[r3 * 27 + e] = sg * blk[e];
                ++r3;
            }
        }
        if (c.off2) {
            c.key2[r2] = (u64)j * N + i;
            for (int e = 0; e < 9; ++e) c.val2[r2 * 9 + e] = -h2[e];
            ++r2;
            c.key2[r2] = (u64)i * N + j;
            for (int e = 0; e < 9; ++e) c.val2[r2 * 9 + e] = -h2[e];
            ++r2;
        }
    }
    if (u > s) {
        const int k = nb[u];
        double rik[3] = {ri[0] - c.pos[3 * k], ri[1] - c.pos[3 * k + 1], ri[2] - c.pos[3 * k + 2]};
        geom::mirror(rik, c.L);
        // the reference evaluates its blocks at r_arr = [r_i; r_i - r_ij; r_i - r_ik]
        double u[3], v[3];
        for (int q = 0; q < 3; ++q) {
            const double nj = ri[q] - rij[q], nk = ri[q] - rik[q];
            u[q] = ri[q] - nj;
            v[q] = ri[q] - nk;
        }
        const double p = (u[0] * u[0] + u[1] * u[1]) + u[2] * u[2];
        const double q2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
        const double sdot = (u[0] * v[0] + u[1] * v[1]) + u[2] * v[2];
        const double le = c.P.lambda * c.P.eps, gs = c.P.gamma * c.P.sigma, rc = c.P.a * c.P.sigma;
        double Hfd[4][4][4];
        sw::h_partials(p, q2, sdot, c.c0_fd, le, gs, rc, Hfd);
        sw::TForms tf;
        sw::t_forms(u, v, tf);
        if (c.off3) {
            double Hik[4][4][4];
            sw::h_partials(p, q2, sdot, c.c0_iik, le, gs, rc, Hik);
            // placements of each block as the reference's set_* helpers add it
            struct Pl { int A, B, C, p1, p2, p3; };
            auto emit = [&](const Pl& pl) {
                c.key3[r3] = key3of(pl.A, pl.B, pl.C, N);
                permute27(blk, pl.p1, pl.p2, pl.p3, 1.0, c.val3 + r3 * 27);
                ++r3;
            };
            // iij -> set_third_order_terms!(i, j): (i,i,j) id, (i,j,i) (1,3,2), (j,i,i) (3,2,1)
            sw::block3(Hfd, tf, 0, 0, 1, blk);
            emit({i, i, j, 1, 2, 3}); emit({i, j, i, 1, 3, 2}); emit({j, i, i, 3, 2, 1});
            // ijj -> set_third_order_terms_alt!(j, i): (i,j,j) id, (j,i,j) (2,1,3), (j,j,i) (3,2,1)
            sw::block3(Hfd, tf, 0, 1, 1, blk);
            emit({i, j, j, 1, 2, 3}); emit({j, i, j, 2, 1, 3}); emit({j, j, i, 3, 2, 1});
            // iik (runtime-built block, cos(theta0) = -1/3) -> set_third_order_terms!(i, k)
            sw::block3(Hik, tf, 0, 0, 2, blk);
            emit({i, i, k, 1, 2, 3}); emit({i, k, i, 1, 3, 2}); emit({k, i, i, 3, 2, 1});
            // ikk -> set_third_order_terms_alt!(k, i)
            sw::block3(Hfd, tf, 0, 2, 2, blk);
            emit({i, k, k, 1, 2, 3}); emit({k, i, k, 2, 1, 3}); emit({k, k, i, 3, 2, 1});
            // jjk -> set_third_order_terms!(j, k)
            sw::block3(Hfd, tf, 1, 1, 2, blk);
            emit({j, j, k, 1, 2, 3}); emit({j, k, j, 1, 3, 2}); emit({k, j, j, 3, 2, 1});
            // jkk -> set_third_order_terms_alt!(k, j)
            sw::block3(Hfd, tf, 1, 2, 2, blk);
            emit({j, k, k, 1, 2, 3}); emit({k, j, k, 2, 1, 3}); emit({k, k, j, 3, 2, 1});
            // ijk -> the six-permutation placement (third_order_AD.jl:212-224)
            sw::block3(Hfd, tf, 0, 1, 2, blk);
            emit({i, j, k, 1, 2, 3}); emit({i, k, j, 1, 3, 2}); emit({j, i, k, 2, 1, 3});
            emit({j, k, i, 2, 3, 1}); emit({k, i, j, 3, 1, 2}); emit({k, j, i, 3, 2, 1});
        }
        if (c.off2) {
            const int pr[3][2] = {{i, j}, {i, k}, {j, k}};
            const int gr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
            double b2[9];
            for (int w = 0; w < 3; ++w) {
                sw::block2(Hfd, tf, gr[w][0], gr[w][1], b2);
                const int A = pr[w][0], B = pr[w][1];
                c.key2[r2] = (u64)B * N + A;                       // (A, B) += block
                for (int e = 0; e < 9; ++e) c.val2[r2 * 9 + e] = b2[e];
                ++r2;
                c.key2[r2] = (u64)A * N + B;                       // (B, A) += block^T
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) c.val2[r2 * 9 + a + 3 * b] = b2[b + 3 * a];
                ++r2;
            }
        }
    }
}

// self records (zero blocks) so that ASR! has a slot for every (A,A,A) / (A,A)
__global__ void k_sw_self(const int* __restrict__ count, int N, long base3, u64* key3,
                          double* val3, long base2, u64* key2, double* val2)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    if (key3) {
        key3[base3 + a] = key3of(a, a, a, N);
        for (int e = 0; e < 27; ++e) val3[(base3 + a) * 27 + e] = 0.0;
    }
    if (key2) {
        key2[base2 + a] = (u64)a * N + a;
        for (int e = 0; e < 9; ++e) val2[(base2 + a) * 9 + e] = 0.0;
    }
}

__global__ void k_iota(int* x, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = (int)t;
}

__global__ void k_flags(const u64* __restrict__ key, long n, int* __restrict__ flag)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    flag[p] = (p == 0 || key[p] != key[p - 1]) ? 1 : 0;
}

__global__ void k_segstart(const int* __restrict__ flag, const int* __restrict__ pos, long n,
                           int* __restrict__ seg)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (flag[p]) seg[pos[p]] = (int)p;
    if (p == n - 1) seg[pos[p] + flag[p]] = (int)n;
}

// Segmented sum in record order: U[u] = sum of the records of segment u.
template <int B>
__global__ void k_reduce(const u64* __restrict__ skey, const int* __restrict__ sidx,
                         const int* __restrict__ seg, int U, const double* __restrict__ val,
                         u64* __restrict__ ukey, double* __restrict__ uval)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)U * B) return;
    const int u = (int)(t / B), e = (int)(t - (long)u * B);
    double acc = 0.0;
    for (int p = seg[u]; p < seg[u + 1]; ++p) acc += val[(long)sidx[p] * B + e];
    uval[t] = acc;
    if (e == 0) ukey[u] = skey[seg[u]];
}

// key2 of a unique block for the row-major view: IFC3 (A,B,C) -> (A N + B) N + C;
// IFC2 (A,B) -> A N + B.
template <int B>
__global__ void k_rowkey(const u64* __restrict__ ukey, int U, int N, u64* __restrict__ k2)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= U) return;
    const u64 k = ukey[u];
    if (B == 27) {
        const u64 A = k % N, Bm = (k / N) % N, C = k / ((u64)N * N);
        k2[u] = (A * N + Bm) * N + C;
    } else {
        const u64 A = k % N, Bm = k / N;
        k2[u] = A * N + Bm;
    }
}

__device__ __forceinline__ int lower_bound_u64(const u64* a, int n, u64 key)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// ASR!: every block whose first and last atoms agree gets -(sum over the other
// blocks of its row, ascending last atom).
template <int B>
__global__ void k_asr(const u64* __restrict__ ukey, int U, int N, const u64* __restrict__ k2s,
                      const int* __restrict__ k2idx, double* __restrict__ uval)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)U * B) return;
    const int u = (int)(t / B), e = (int)(t - (long)u * B);
    const u64 k = ukey[u];
    if (B == 27) {
        const u64 A = k % N, Bm = (k / N) % N, C = k / ((u64)N * N);
        if (A != C) return;
        const u64 lo = (A * N + Bm) * N;
        int p = lower_bound_u64(k2s, U, lo);
        double acc = 0.0;
        for (; p < U && k2s[p] < lo + N; ++p) {
            const u64 Cp = k2s[p] - lo;
            if (Cp == A) continue;
            acc += uval[(long)k2idx[p] * 27 + e];
        }
        uval[t] = -acc;
    } else {
        const u64 A = k % N, Bm = k / N;
        if (A != Bm) return;
        const u64 lo = A * N;
        int p = lower_bound_u64(k2s, U, lo);
        double acc = 0.0;
        for (; p < U && k2s[p] < lo + N; ++p) {
            const u64 Bp = k2s[p] - lo;
            if (Bp == A) continue;
            acc += uval[(long)k2idx[p] * 9 + e];
        }
        uval[t] = -acc;
    }
}

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));
constexpr int kMaxCols = 256;

// Dense IFC3 writer over a block list sorted by (C, B, A).
template <int VEC, bool W64, bool W32, bool MW>
__global__ void __launch_bounds__(256)
k_write_blocks3(const u64* __restrict__ ukey, const double* __restrict__ uval, int U, int N,
                double tol, const double* __restrict__ masses, double* __restrict__ psi64,
                float* __restrict__ psi32, int cpw)
{
    const long n = 3L * N;
    const int kk = blockIdx.y;
    const int K = kk / 3, g = kk - 3 * K;
    const long col0 = (long)blockIdx.x * cpw;
    const int ncols = (int)((n - col0) < cpw ? (n - col0) : cpw);
    __shared__ int s_lo[kMaxCols], s_hi[kMaxCols];
    for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
        const long jj = col0 + c;
        const int J = (int)(jj / 3);
        const u64 base = ((u64)K * N + J) * N;
        const int lo = lower_bound_u64(ukey, U, base);
        const int hi = lower_bound_u64(ukey, U, base + N);
        s_lo[c] = lo;
        s_hi[c] = hi;
    }
    __syncthreads();
    const long base = (long)kk * n * n + col0 * n;
    const long total = (long)ncols * n;
    const double mK = MW ? masses[K] : 0.0;
    auto value = [&](int c, int r, double& v64, float& v32) {
        v64 = 0.0;
        v32 = 0.0f;
        const int lo = s_lo[c], hi = s_hi[c];
        if (lo >= hi) return;
        const int I = r / 3, a = r - 3 * (r / 3);
        const long jj = col0 + c;
        const int J = (int)(jj / 3), b = (int)(jj - 3L * J);
        const u64 want = ((u64)K * N + J) * N + I;
        const int p = lo + lower_bound_u64(ukey + lo, hi - lo, want);
        if (p >= hi || ukey[p] != want) return;
        double v = uval[(long)p * 27 + a + 3 * b + 9 * g];
        v = fabs(v) < tol ? 0.0 : v;
        v64 = v;
        if (MW) {
            const double m = masses[I] * masses[J];
            v32 = (float)(v != 0.0 ? v / sqrt(m * mK) : v);
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
            value(c, r + 1, a1, f1);
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

__global__ void __launch_bounds__(256)
k_write_blocks2(const u64* __restrict__ ukey, const double* __restrict__ uval, int U, int N,
                double tol, double* __restrict__ out)
{
    const long n = 3L * N;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const long j = t / n, i = t - j * n;
    const int I = (int)(i / 3), a = (int)(i - 3L * I);
    const int J = (int)(j / 3), b = (int)(j - 3L * J);
    const u64 want = (u64)J * N + I;
    const int p = lower_bound_u64(ukey, U, want);
    double v = 0.0;
    if (p < U && ukey[p] == want) v = uval[(long)p * 9 + a + 3 * b];
    out[t] = fabs(v) < tol ? 0.0 : v;
}

inline unsigned blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int key_bits(u64 maxkey)
{
    int b = 1;
    while (b < 64 && (maxkey >> b) != 0) ++b;
    return b;
}

// Sort records, reduce segments, apply ASR!.  Returns unique blocks (ukey, uval, U).
template <int B>
int reduce_blocks(hipStream_t s, long R, int N, u64* key, double* val, Carve& cv, u64*& ukey,
                  double*& uval, int& U)
{
    int* idx = cv.take<int>(R);
    int* sidx = cv.take<int>(R);
    u64* skey = cv.take<u64>(R);
    int* flag = cv.take<int>(R);
    int* pos = cv.take<int>(R);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "SW workspace too small (records)");
    hipLaunchKernelGGL(k_iota, dim3(blocks(R, 256)), dim3(256), 0, s, idx, R);
    const u64 maxkey = (B == 27) ? (u64)N * N * N : (u64)N * N;
    const int bits = key_bits(maxkey);
    size_t tb = 0;
    FC3_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, skey, idx, sidx, (int)R, 0, bits, s));
    void* tmp = cv.take<char>(tb);
    size_t tb2 = 0;
    FC3_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, flag, pos, (int)R, s));
    void* tmp2 = cv.take<char>(tb2);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "SW workspace too small (sort)");
    FC3_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, skey, idx, sidx, (int)R, 0, bits, s));
    hipLaunchKernelGGL(k_flags, dim3(blocks(R, 256)), dim3(256), 0, s, skey, R, flag);
    FC3_HIP(hipcub::DeviceScan::ExclusiveSum(tmp2, tb2, flag, pos, (int)R, s));
    int last[2];
    FC3_HIP(hipMemcpyAsync(&last[0], pos + R - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    FC3_HIP(hipMemcpyAsync(&last[1], flag + R - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    FC3_HIP(hipStreamSynchronize(s));
    U = last[0] + last[1];
    int* seg = cv.take<int>(U + 1);
    ukey = cv.take<u64>(U);
    uval = cv.take<double>((size_t)U * B);
    // row view for ASR!
    u64* k2 = cv.take<u64>(U);
    u64* k2s = cv.take<u64>(U);
    int* uidx = cv.take<int>(U);
    int* uidxs = cv.take<int>(U);
    size_t tb3 = 0;
    FC3_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, k2, k2s, uidx, uidxs, U, 0, bits, s));
    void* tmp3 = cv.take<char>(tb3);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "SW workspace too small (blocks)");
    hipLaunchKernelGGL(k_segstart, dim3(blocks(R, 256)), dim3(256), 0, s, flag, pos, R, seg);
    hipLaunchKernelGGL((k_reduce<B>), dim3(blocks((long)U * B, 256)), dim3(256), 0, s, skey, sidx,
                       seg, U, val, ukey, uval);
    hipLaunchKernelGGL((k_rowkey<B>), dim3(blocks(U, 256)), dim3(256), 0, s, ukey, U, N, k2);
    hipLaunchKernelGGL(k_iota, dim3(blocks(U, 256)), dim3(256), 0, s, uidx, (long)U);
    FC3_HIP(hipcub::DeviceRadixSort::SortPairs(tmp3, tb3, k2, k2s, uidx, uidxs, U, 0, bits, s));
    hipLaunchKernelGGL((k_asr<B>), dim3(blocks((long)U * B, 256)), dim3(256), 0, s, ukey, U, N,
                       k2s, uidxs, uval);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

// LDS-staged writer for even n <= kLdsColElems (the same values as k_write_blocks3).
// The workgroup's cpw columns go in chunks of C = max(1, kChunkElems / n) columns, held in
// a dynamic LDS buffer of C n doubles (at most 48 KB):
//   - the chunk's column buffer in LDS is zeroed;
//   - every block of its columns scatters its 3 rows (after the tolerance) into it;
//   - the chunk streams out as 16-byte (FP64) and 8-byte (FP32) non-temporal stores, with
//     the mass weighting applied on the way.
// No element searches the block keys: a column's block range comes from two searches.
constexpr int kLdsColElems = 6144;   // largest LDS column buffer (48 KB): n <= 6144
// elements per chunk; measured at 648 DoF: 3072 (4 columns, 21 KB) 0.55 ms, 1296 0.71 ms,
// 5832 0.67 ms
constexpr int kChunkElems = 3072;

template <bool W64, bool W32, bool MW>
__global__ void __launch_bounds__(256)
k_write_blocks3_lds(const u64* __restrict__ ukey, const double* __restrict__ uval, int U, int N,
                    double tol, const double* __restrict__ masses, double* __restrict__ psi64,
                    float* __restrict__ psi32, int cpw)
{
    const int n = 3 * N;
    const int kk = blockIdx.y;
    const int K = kk / 3, g = kk - 3 * K;
    const long col0 = (long)blockIdx.x * cpw;
    const int ncols = (int)((n - col0) < cpw ? (n - col0) : cpw);
    __shared__ int s_lo[kMaxCols], s_hi[kMaxCols];
    extern __shared__ double buf[];   // C n doubles
    for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
        const int J = (int)((col0 + c) / 3);
        const u64 base = ((u64)K * N + J) * N;
        s_lo[c] = lower_bound_u64(ukey, U, base);
        s_hi[c] = lower_bound_u64(ukey, U, base + N);
    }
    const int C = n < kChunkElems ? kChunkElems / n : 1;
    const double mK = MW ? masses[K] : 0.0;
    for (int c0 = 0; c0 < ncols; c0 += C) {
        const int cc = (ncols - c0) < C ? (ncols - c0) : C;
        const int elems = cc * n;
        __syncthreads();   // ranges ready; the previous chunk's stream has read buf
        for (int e = threadIdx.x; e < elems; e += blockDim.x) buf[e] = 0.0;
        __syncthreads();
        for (int c = 0; c < cc; ++c) {
            const int lo = s_lo[c0 + c], hi = s_hi[c0 + c];
            const int b = (int)((col0 + c0 + c) % 3);
            for (int p = lo + (int)threadIdx.x; p < hi; p += blockDim.x) {
                const int I = (int)(ukey[p] % (u64)N);
                const double* v = uval + (long)p * 27 + 3 * b + 9 * g;
#pragma unroll
                for (int a = 0; a < 3; ++a) buf[c * n + 3 * I + a] = fabs(v[a]) < tol ? 0.0 : v[a];
            }
        }
        __syncthreads();
        const long gbase = (long)kk * n * n + (col0 + c0) * n;
        for (int q = threadIdx.x; 2 * q < elems; q += blockDim.x) {
            const int e = 2 * q;
            const double a0 = buf[e], a1 = buf[e + 1];
            if (W64) {
                dvec2 v2 = {a0, a1};
                __builtin_nontemporal_store(v2, reinterpret_cast<dvec2*>(psi64 + gbase) + q);
            }
            if (W32) {
                float f0 = (float)a0, f1 = (float)a1;
                if (MW) {
                    const int c = e / n, r = e - c * n;   // n even: e + 1 is in column c too
                    const double mJ = masses[(int)((col0 + c0 + c) / 3)];
                    if (a0 != 0.0) f0 = (float)(a0 / sqrt(masses[r / 3] * mJ * mK));
                    if (a1 != 0.0) f1 = (float)(a1 / sqrt(masses[(r + 1) / 3] * mJ * mK));
                }
                fvec2 f2 = {f0, f1};
                __builtin_nontemporal_store(f2, reinterpret_cast<fvec2*>(psi32 + gbase) + q);
            }
        }
    }
}

// ---- parameter adjoint ----------------------------------------------------------
// Cotangent of one placed record value at (3A+x1, 3B+x2, 3C+x3): the reference's ASR!
// overwrites every (A, B, A) block, so a record with C == A contributes nothing, and a
// record with C != A reaches its own position and, with a minus sign, the ASR entry
// (3A+x1, 3B+x2, 3A+x3).  apply_tols! zeroes entries (and their derivatives) with
// |value| < tol.  The forward values come from the reduced block table of the forward
// assembly (ukey sorted, uval after ASR!, the values the dense writer stores), so no
// dense forward tensor is kept: block (A, B, C) is found by binary search on its key.
template <typename DT>
struct GradCtx {
    const DT* dpsi;
    const u64* ukey;         // forward block keys (mask), or null when tol == 0
    const double* uval;
    int U, N;
    double tol;
    const double* masses;    // mass factors for a cotangent of the mass-weighted Psi, or null
    long n;
};

// The forward block (A, B, C)'s 27 values, or null when the block is absent (all zero).
template <typename DT>
__device__ const double* fwd_block(const GradCtx<DT>& g, int A, int B, int C)
{
    const u64 want = key3of(A, B, C, g.N);
    const int p = lower_bound_u64(g.ukey, g.U, want);
    return (p < g.U && g.ukey[p] == want) ? g.uval + (long)p * 27 : nullptr;
}

// Cotangent at (3A+x1, 3B+x2, 3C+x3); fb: that block's forward values (tol > 0 only).
template <typename DT>
__device__ double pos_cot(const GradCtx<DT>& g, const double* fb, int A, int B, int C, int x1,
                          int x2, int x3)
{
    if (g.ukey && !(fb && fabs(fb[x1 + 3 * x2 + 9 * x3]) >= g.tol)) return 0.0;
    const long p = (3L * A + x1) + g.n * (3L * B + x2) + g.n * g.n * (3L * C + x3);
    double w = (double)g.dpsi[p];
    if (g.masses) w /= sqrt((g.masses[A] * g.masses[B]) * g.masses[C]);
    return w;
}

// Adds to W[src index] the cotangent of the placement (A, B, C) with permutation perm
// (permutedims semantics of permute27) and sign sgn.
template <typename DT>
__device__ void place_cot(const GradCtx<DT>& g, int A, int B, int C, int p1, int p2, int p3,
                          double sgn, double* W)
{
    if (C == A) return;
    const int pm[3] = {p1 - 1, p2 - 1, p3 - 1};
    const double* fc = g.ukey ? fwd_block(g, A, B, C) : nullptr;
    const double* fa = g.ukey ? fwd_block(g, A, B, A) : nullptr;
    for (int x3 = 0; x3 < 3; ++x3)
        for (int x2 = 0; x2 < 3; ++x2)
            for (int x1 = 0; x1 < 3; ++x1) {
                int jj[3];
                jj[pm[0]] = x1;
                jj[pm[1]] = x2;
                jj[pm[2]] = x3;
                const double w = pos_cot(g, fc, A, B, C, x1, x2, x3) -
                                 pos_cot(g, fa, A, B, A, x1, x2, x3);
                W[jj[0] + 3 * jj[1] + 9 * jj[2]] += sgn * w;
            }
}

// One thread per (centre i, slot s), the record loop of k_sw_emit: accumulates
// dL/d(A, B, lambda, gamma) into per-block partials (4 doubles per block).
template <typename DT>
__global__ void __launch_bounds__(64) k_sw_param_grad(EmitCtx c, GradCtx<DT> g,
                                                      double* __restrict__ partial)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double gA = 0.0, gB = 0.0, gl = 0.0, gg = 0.0;
    if (t < c.N * c.cap) {
        const int i = t / c.cap, s = t - i * c.cap;
        const int d = c.count[i];
        if (s < d) {
            const int* nb = c.nbr + (size_t)i * c.cap;
            const int j = nb[s];
            const double ri[3] = {c.pos[3 * i], c.pos[3 * i + 1], c.pos[3 * i + 2]};
            double rij[3] = {ri[0] - c.pos[3 * j], ri[1] - c.pos[3 * j + 1],
                             ri[2] - c.pos[3 * j + 2]};
            geom::mirror(rij, c.L);
            double W[27];
            if (j > i) {
                for (int e = 0; e < 27; ++e) W[e] = 0.0;
                const int ks[6][3] = {{i, i, j}, {i, j, i}, {j, i, i}, {j, j, i}, {j, i, j}, {i, j, j}};
                for (int q = 0; q < 6; ++q)
                    place_cot(g, ks[q][0], ks[q][1], ks[q][2], 1, 2, 3, q < 3 ? -1.0 : 1.0, W);
                double blk[27], b1[27], b0[27];
                sw::pair_blocks(rij, c.P, nullptr, blk);
                sw::Params P1 = c.P, P0 = c.P;
                P1.B = 1.0;
                P0.B = 0.0;
                sw::pair_blocks(rij, P1, nullptr, b1);   // phi2 is affine in B
                sw::pair_blocks(rij, P0, nullptr, b0);
                for (int e = 0; e < 27; ++e) {
                    gA += (blk[e] / c.P.A) * W[e];
                    gB += (b1[e] - b0[e]) * W[e];
                }
            }
            for (int tt = s + 1; tt < d; ++tt) {
                const int k = nb[tt];
                double rik[3] = {ri[0] - c.pos[3 * k], ri[1] - c.pos[3 * k + 1],
                                 ri[2] - c.pos[3 * k + 2]};
                geom::mirror(rik, c.L);
                double u[3], v[3];
                for (int q = 0; q < 3; ++q) {
                    const double nj = ri[q] - rij[q], nk = ri[q] - rik[q];
                    u[q] = ri[q] - nj;
                    v[q] = ri[q] - nk;
                }
                const double p = (u[0] * u[0] + u[1] * u[1]) + u[2] * u[2];
                const double q2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
                const double sdot = (u[0] * v[0] + u[1] * v[1]) + u[2] * v[2];
                const double le = c.P.lambda * c.P.eps, gs = c.P.gamma * c.P.sigma,
                             rc = c.P.a * c.P.sigma;
                double Hfd[4][4][4], dHfd[4][4][4], Hik[4][4][4], dHik[4][4][4];
                sw::h_partials(p, q2, sdot, c.c0_fd, le, gs, rc, Hfd, dHfd, c.P.gamma);
                sw::h_partials(p, q2, sdot, c.c0_iik, le, gs, rc, Hik, dHik, c.P.gamma);
                sw::TForms tf;
                sw::t_forms(u, v, tf);
                struct Pl { int A, B, C, p1, p2, p3; };
                // the seven blocks of k_sw_emit with their placements
                auto block = [&](const double (*H)[4][4], const double (*dH)[4][4], int G1, int G2,
                                 int G3, const Pl* pl, int np) {
                    for (int e = 0; e < 27; ++e) W[e] = 0.0;
                    for (int q = 0; q < np; ++q)
                        place_cot(g, pl[q].A, pl[q].B, pl[q].C, pl[q].p1, pl[q].p2, pl[q].p3, 1.0, W);
                    double b[27], db[27];
                    sw::block3(H, tf, G1, G2, G3, b);
                    sw::block3(dH, tf, G1, G2, G3, db);
                    for (int e = 0; e < 27; ++e) {
                        gl += (b[e] / c.P.lambda) * W[e];
                        gg += db[e] * W[e];
                    }
                };
                const Pl iij[3] = {{i, i, j, 1, 2, 3}, {i, j, i, 1, 3, 2}, {j, i, i, 3, 2, 1}};
                const Pl ijj[3] = {{i, j, j, 1, 2, 3}, {j, i, j, 2, 1, 3}, {j, j, i, 3, 2, 1}};
                const Pl iik[3] = {{i, i, k, 1, 2, 3}, {i, k, i, 1, 3, 2}, {k, i, i, 3, 2, 1}};
                const Pl ikk[3] = {{i, k, k, 1, 2, 3}, {k, i, k, 2, 1, 3}, {k, k, i, 3, 2, 1}};
                const Pl jjk[3] = {{j, j, k, 1, 2, 3}, {j, k, j, 1, 3, 2}, {k, j, j, 3, 2, 1}};
                const Pl jkk[3] = {{j, k, k, 1, 2, 3}, {k, j, k, 2, 1, 3}, {k, k, j, 3, 2, 1}};
                const Pl ijk[6] = {{i, j, k, 1, 2, 3}, {i, k, j, 1, 3, 2}, {j, i, k, 2, 1, 3},
                                   {j, k, i, 2, 3, 1}, {k, i, j, 3, 1, 2}, {k, j, i, 3, 2, 1}};
                block(Hfd, dHfd, 0, 0, 1, iij, 3);
                block(Hfd, dHfd, 0, 1, 1, ijj, 3);
                block(Hik, dHik, 0, 0, 2, iik, 3);
                block(Hfd, dHfd, 0, 2, 2, ikk, 3);
                block(Hfd, dHfd, 1, 1, 2, jjk, 3);
                block(Hfd, dHfd, 1, 2, 2, jkk, 3);
                block(Hfd, dHfd, 0, 1, 2, ijk, 6);
            }
        }
    }
    __shared__ double red[4][64];
    red[0][threadIdx.x] = gA;
    red[1][threadIdx.x] = gB;
    red[2][threadIdx.x] = gl;
    red[3][threadIdx.x] = gg;
    __syncthreads();
    for (int h = 32; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) partial[4 * blockIdx.x + q] = red[q][0];
}

__global__ void k_sum_partials4(const double* __restrict__ partial, int cnt, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < cnt; ++i)
        for (int q = 0; q < 4; ++q) a[q] += partial[4 * i + q];
    for (int q = 0; q < 4; ++q) out[q] = a[q];
}

}  // namespace

size_t sw_workspace_bytes(int N, int cap, bool third)
{
    const long d = cap;
    const long R = (long)N * (third ? (6 * d + 12 * d * (d - 1)) : (2 * d + 3 * d * (d - 1))) + N;
    const int B = third ? 27 : 9;
    // records (key, values), sort buffers (idx x2, skey, flag, pos), unique copies,
    // row view, radix-sort temporaries (bounded generously)
    size_t b = 0;
    b += Carve::need(sizeof(u64) * R) + Carve::need(sizeof(double) * R * B);
    b += 4 * Carve::need(sizeof(int) * R) + Carve::need(sizeof(u64) * R);
    b += Carve::need(sizeof(int) * (R + 1)) + Carve::need(sizeof(u64) * R) +
         Carve::need(sizeof(double) * R * B);
    b += 2 * Carve::need(sizeof(u64) * R) + 2 * Carve::need(sizeof(int) * R);
    b += 3 * Carve::need(sizeof(u64) * R * 2 + (64 << 10));
    b += 4 * Carve::need(sizeof(int) * (N + 1)) + (1 << 20);
    return b;
}

int sw_assemble(hipStream_t s, const double* d_pos, const double box[3], const double* params,
                double c0_fd, double c0_iik, const NeighborList& nl, void* ws, size_t ws_bytes,
                double tol, const double* d_masses, double* psi64, float* psi32, double* ifc2,
                SwBlockTable* table)
{
    const int N = nl.n_atoms;
    const bool third = (psi64 || psi32 || table);
    Carve cv(ws, ws_bytes);
    int* rec3 = cv.take<int>(N + 1);
    int* rec2 = cv.take<int>(N + 1);
    int* off3 = cv.take<int>(N + 1);
    int* off2 = cv.take<int>(N + 1);
    hipLaunchKernelGGL(k_sw_counts, dim3(blocks(N, 256)), dim3(256), 0, s, nl.count, nl.nbr, N,
                       nl.cap, rec3, rec2);
    FC3_HIP(hipGetLastError());
    std::vector<int> h3(N), h2(N);
    FC3_HIP(hipMemcpyAsync(h3.data(), rec3, sizeof(int) * N, hipMemcpyDeviceToHost, s));
    FC3_HIP(hipMemcpyAsync(h2.data(), rec2, sizeof(int) * N, hipMemcpyDeviceToHost, s));
    FC3_HIP(hipStreamSynchronize(s));
    std::vector<int> o3(N), o2(N);
    long R3 = 0, R2 = 0;
    for (int i = 0; i < N; ++i) {
        o3[i] = (int)R3;
        o2[i] = (int)R2;
        R3 += h3[i];
        R2 += h2[i];
    }
    if (R3 + N >= (1L << 31)) return set_error(FC3_ERR_UNSUPPORTED, "SW: too many records");
    FC3_HIP(hipMemcpyAsync(off3, o3.data(), sizeof(int) * N, hipMemcpyHostToDevice, s));
    FC3_HIP(hipMemcpyAsync(off2, o2.data(), sizeof(int) * N, hipMemcpyHostToDevice, s));
    const long RT3 = R3 + N, RT2 = R2 + N;
    u64* key3 = nullptr;
    double* val3 = nullptr;
    u64* key2 = nullptr;
    double* val2 = nullptr;
    if (third) {
        key3 = cv.take<u64>(RT3);
        val3 = cv.take<double>((size_t)RT3 * 27);
    } else {
        key2 = cv.take<u64>(RT2);
        val2 = cv.take<double>((size_t)RT2 * 9);
    }
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "SW workspace too small");
    EmitCtx c;
    c.pos = d_pos;
    c.count = nl.count;
    c.nbr = nl.nbr;
    c.N = N;
    c.cap = nl.cap;
    c.L[0] = box[0]; c.L[1] = box[1]; c.L[2] = box[2];
    c.P = sw::Params{params[0], params[1], params[2], params[3], params[4],
                     params[5], params[6], params[7], params[8], params[9]};
    c.c0_fd = c0_fd;
    c.c0_iik = c0_iik;
    c.off3 = third ? off3 : nullptr;
    c.off2 = third ? nullptr : off2;
    c.key3 = key3; c.val3 = val3; c.key2 = key2; c.val2 = val2;
    {
        ProfScope prof(FC3_KERNEL_PAIR_BLOCKS, s);
        hipLaunchKernelGGL(k_sw_emit, dim3(blocks((long)N * nl.cap * nl.cap, 64)), dim3(64), 0, s,
                           c);
        hipLaunchKernelGGL(k_sw_self, dim3(blocks(N, 256)), dim3(256), 0, s, nl.count, N, R3,
                           key3, val3, R2, key2, val2);
        FC3_HIP(hipGetLastError());
    }
    u64* ukey = nullptr;
    double* uval = nullptr;
    int U = 0;
    const long n = 3L * N;
    if (third) {
        FC3_TRY(reduce_blocks<27>(s, RT3, N, key3, val3, cv, ukey, uval, U));
        if (table) {   // the reduced blocks only (they stay in ws), no dense tensor
            table->ukey = ukey;
            table->uval = uval;
            table->U = U;
            return FC3_OK;
        }
        long cpw = 65536 / n;
        if (cpw < 1) cpw = 1;
        if (cpw > kMaxCols) cpw = kMaxCols;
        dim3 grid(blocks(n, (int)cpw), (unsigned)n);
        ProfScope prof(FC3_KERNEL_IFC3_WRITE, s);
        const bool mw = d_masses != nullptr;
#define LW(V, A, Bq, M) hipLaunchKernelGGL((k_write_blocks3<V, A, Bq, M>), grid, dim3(256), 0, s, \
                                           ukey, uval, U, N, tol, d_masses, psi64, psi32, (int)cpw)
#define LL(A, Bq, M) hipLaunchKernelGGL((k_write_blocks3_lds<A, Bq, M>), grid, dim3(256), lds, s, \
                                        ukey, uval, U, N, tol, d_masses, psi64, psi32, (int)cpw)
        if (n % 2 == 0 && n <= kLdsColElems) {
            // one LDS chunk of columns per workgroup: many short workgroups, a short tail
            cpw = std::min<long>(n < kChunkElems ? kChunkElems / n : 1, kMaxCols);
            grid = dim3(blocks(n, (int)cpw), (unsigned)n);
            const size_t lds = sizeof(double) * (size_t)cpw * n;
            if (psi64 && psi32) { if (mw) LL(true, true, true); else LL(true, true, false); }
            else if (psi64) LL(true, false, false);
            else { if (mw) LL(false, true, true); else LL(false, true, false); }
        } else if (n % 2 == 0) {
            if (psi64 && psi32) { if (mw) LW(2, true, true, true); else LW(2, true, true, false); }
            else if (psi64) LW(2, true, false, false);
            else { if (mw) LW(2, false, true, true); else LW(2, false, true, false); }
        } else {
            if (psi64 && psi32) { if (mw) LW(1, true, true, true); else LW(1, true, true, false); }
            else if (psi64) LW(1, true, false, false);
            else { if (mw) LW(1, false, true, true); else LW(1, false, true, false); }
        }
#undef LW
#undef LL
        FC3_HIP(hipGetLastError());
    } else {
        FC3_TRY(reduce_blocks<9>(s, RT2, N, key2, val2, cv, ukey, uval, U));
        hipLaunchKernelGGL(k_write_blocks2, dim3(blocks(n * n, 256)), dim3(256), 0, s, ukey, uval, U,
                           N, tol, ifc2);
        FC3_HIP(hipGetLastError());
    }
    return FC3_OK;
}

}  // namespace fc3

namespace fc3 {

size_t sw_param_grad_bytes(int N, int cap)
{
    const long blocks_ = ((long)N * cap + 63) / 64;
    return Carve::need(sizeof(double) * (4 * blocks_ + 4));
}

int sw_param_grad(hipStream_t s, const double* d_pos, const double box[3], const double* params,
                  double c0_fd, double c0_iik, const NeighborList& nl, double tol,
                  const SwBlockTable* fwd, const double* d_masses, const float* dpsi32,
                  const double* dpsi64, void* ws, double* d_out)
{
    const int N = nl.n_atoms;
    const long nblk = ((long)N * nl.cap + 63) / 64;
    Carve cv(ws, sw_param_grad_bytes(N, nl.cap));
    double* partial = cv.take<double>(4 * nblk);
    EmitCtx c;
    c.pos = d_pos;
    c.count = nl.count;
    c.nbr = nl.nbr;
    c.N = N;
    c.cap = nl.cap;
    c.L[0] = box[0]; c.L[1] = box[1]; c.L[2] = box[2];
    c.P = sw::Params{params[0], params[1], params[2], params[3], params[4],
                     params[5], params[6], params[7], params[8], params[9]};
    c.c0_fd = c0_fd;
    c.c0_iik = c0_iik;
    c.off3 = nullptr; c.off2 = nullptr;
    c.key3 = nullptr; c.val3 = nullptr; c.key2 = nullptr; c.val2 = nullptr;
    const long n = 3L * N;
    if (dpsi32) {
        GradCtx<float> g{dpsi32, tol > 0.0 ? fwd->ukey : nullptr, fwd ? fwd->uval : nullptr,
                         fwd ? fwd->U : 0, N, tol, d_masses, n};
        hipLaunchKernelGGL(k_sw_param_grad<float>, dim3((unsigned)nblk), dim3(64), 0, s, c, g, partial);
    } else {
        GradCtx<double> g{dpsi64, tol > 0.0 ? fwd->ukey : nullptr, fwd ? fwd->uval : nullptr,
                          fwd ? fwd->U : 0, N, tol, nullptr, n};
        hipLaunchKernelGGL(k_sw_param_grad<double>, dim3((unsigned)nblk), dim3(64), 0, s, c, g, partial);
    }
    FC3_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials4, dim3(1), dim3(64), 0, s, partial, (int)nblk, d_out);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

}  // namespace fc3
