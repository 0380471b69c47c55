// Text output of dense force constants: save_second_order / save_third_order with
// fmt = :txt (src/io.jl:9-62).
//
// The reference walks the whole dense tensor on the host in the order atoms i, j(, k),
// then components a, b(, c), and prints one line "i a j b [k c] value" (1-based) per
// nonzero, after a header "#N atoms, Sum X = round(sum, sigdigits = 7), Units: u".
// Here the tensor stays on the device:
//   k_nnz_count   one pass over the tensor (HBM-bound): nonzeros and value sums per
//                 workgroup chunk (fixed chunks, fixed reduction order: deterministic);
//   k_nnz_compact second pass: every nonzero of a chunk written at the chunk's offset
//                 with the key of its line in the reference's print order;
//   radix sort    of (key, value) pairs (hipCUB), so that only the nonzeros (16 bytes
//                 each) cross PCIe, already in print order;
// and the host formats the lines.  Values print as Julia prints a Float64 (the shortest
// decimal that reads back to the same double, plain or exponent notation by Julia's
// rule), so the file reads back to the exact tensor.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fc3_internal.hpp"

namespace fc3 {
namespace {

constexpr int kChunks = 8192;   // workgroups per pass: 32 per CU
constexpr int kNT = 256;

// print-order key of the element at column-major linear index e of an n^dims tensor,
// n = D N: ((i N + j) N + k) D^3 + (a D + b) D + c for dims = 3 (i, a from the first
// index ii = D i + a, etc.), (i N + j) D^2 + a D + b for dims = 2
__device__ __forceinline__ unsigned long long print_key(long e, int n, int N, int D, int dims)
{
    const long ii = e % n, rest = e / n;
    const long jj = rest % n, kk = rest / n;
    const long i = ii / D, a = ii % D, j = jj / D, b = jj % D;
    if (dims == 2) return (unsigned long long)(((i * N + j) * D + a) * D + b);
    const long k = kk / D, c = kk % D;
    return (unsigned long long)((((i * N + j) * N + k) * D + a) * D + b) * D + c;
}

// chunk c covers elements [c L, min((c + 1) L, E)); L is a multiple of 2 kNT
__global__ void __launch_bounds__(kNT) k_nnz_count(const double* __restrict__ x, long E, long L,
                                                   long long* __restrict__ cnt,
                                                   double* __restrict__ sum)
{
    __shared__ long long sc[kNT];
    __shared__ double ss[kNT];
    const long b = (long)blockIdx.x * L, end = min(b + L, E);
    long long c = 0;
    double s = 0.0;
    for (long e = b + threadIdx.x; e < end; e += kNT) {
        const double v = x[e];
        if (v != 0.0) {
            ++c;
            s += v;
        }
    }
    sc[threadIdx.x] = c;
    ss[threadIdx.x] = s;
    __syncthreads();
    for (int w = kNT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sc[threadIdx.x] += sc[threadIdx.x + w];
            ss[threadIdx.x] += ss[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cnt[blockIdx.x] = sc[0];
        sum[blockIdx.x] = ss[0];
    }
}

// nonzeros of chunk blockIdx.x, in memory order, at off[blockIdx.x]
__global__ void __launch_bounds__(kNT) k_nnz_compact(const double* __restrict__ x, long E, long L,
                                                     int n, int N, int D, int dims,
                                                     const long long* __restrict__ off,
                                                     unsigned long long* __restrict__ keys,
                                                     double* __restrict__ vals)
{
    __shared__ int wsum[kNT / 64];
    const long b = (long)blockIdx.x * L, end = min(b + L, E);
    long long o = off[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (long t0 = b; t0 < end; t0 += kNT) {
        const long e = t0 + threadIdx.x;
        const double v = e < end ? x[e] : 0.0;
        const bool nz = v != 0.0;
        const unsigned long long m = __ballot(nz);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int wofs = 0, tot = 0;
        for (int w = 0; w < kNT / 64; ++w) {
            if (w < wave) wofs += wsum[w];
            tot += wsum[w];
        }
        if (nz) {
            keys[o + wofs + before] = print_key(e, n, N, D, dims);
            vals[o + wofs + before] = v;
        }
        o += tot;
        __syncthreads();   // wsum is rewritten by the next tile
    }
}

// ---- host: Julia's printing of a Float64 and of round(x, sigdigits = 7) ------------

// Julia's mod for floats: the result has the sign of the divisor (y > 0 here)
double jl_mod(double a, double y)
{
    double r = std::fmod(a, y);
    if (r < 0) r += y;
    return r;
}

}  // namespace

// show(io, x::Float64) as Julia's Base.Ryu.writeshortest with print's defaults: the
// shortest digit string d_1..d_m that reads back to x, with pt = (decimal exponent of
// d_1) + 1.  Plain notation when -4 < pt <= 6 (and, for pt >= m, when x is the integer
// those digits spell up to 0.05), with ".0" after integers; otherwise d_1.d_2..d_m e exp
// (exp = pt - 1, no padding; "d_1.0" when m = 1).  Returns the length written to out
// (at most 32 chars, not terminated).
int julia_float_chars(double x, char* out)
{
    char* p = out;
    if (std::isnan(x)) { std::memcpy(out, "NaN", 3); return 3; }
    if (std::signbit(x)) *p++ = '-';
    if (std::isinf(x)) { std::memcpy(p, "Inf", 3); return (int)(p - out) + 3; }
    if (x == 0.0) { std::memcpy(p, "0.0", 3); return (int)(p - out) + 3; }
    char sci[40];
    const auto r = std::to_chars(sci, sci + sizeof sci, std::fabs(x), std::chars_format::scientific);
    // sci: d[.ddd]e(+|-)XX
    char dig[24];
    int m = 0;
    const char* q = sci;
    for (; q < r.ptr && *q != 'e'; ++q)
        if (*q != '.') dig[m++] = *q;
    int e10 = 0;   // the exponent: the characters after 'e' up to r.ptr (no terminator)
    std::from_chars(q + 1 + (q[1] == '+'), r.ptr, e10);
    const int pt = e10 + 1;
    bool plain = -4 < pt && pt <= 6;
    if (plain && pt >= m && std::fabs(jl_mod(x + 0.05, std::pow(10.0, pt - m)) - 0.05) > 0.05)
        plain = false;
    if (plain) {
        if (pt <= 0) {
            *p++ = '0';
            *p++ = '.';
            for (int z = 0; z < -pt; ++z) *p++ = '0';
            for (int d = 0; d < m; ++d) *p++ = dig[d];
        } else if (pt < m) {
            for (int d = 0; d < pt; ++d) *p++ = dig[d];
            *p++ = '.';
            for (int d = pt; d < m; ++d) *p++ = dig[d];
        } else {
            for (int d = 0; d < m; ++d) *p++ = dig[d];
            for (int z = 0; z < pt - m; ++z) *p++ = '0';
            *p++ = '.';
            *p++ = '0';
        }
    } else {
        *p++ = dig[0];
        *p++ = '.';
        if (m == 1) *p++ = '0';
        for (int d = 1; d < m; ++d) *p++ = dig[d];
        *p++ = 'e';
        const auto er = std::to_chars(p, p + 8, pt - 1);
        p = er.ptr;
    }
    return (int)(p - out);
}

// round(x, sigdigits = 7) (base/floatfuncs.jl): h = floor(log10|x|) + 1 digits before the
// point; x * 10^(7 - h) rounded to an integer (ties to even), scaled back
double julia_round_sig7(double x)
{
    if (x == 0.0 || !std::isfinite(x)) return x;
    const int h = (int)std::floor(std::log10(std::fabs(x))) + 1;
    const int d = 7 - h;
    double r;
    if (d >= 0) {
        const double sc = std::pow(10.0, d);
        r = std::nearbyint(x * sc) / sc;
    } else {
        const double isc = std::pow(10.0, -d);
        r = std::nearbyint(x / isc) * isc;
    }
    return std::isfinite(r) ? r : x;
}

// The nonzeros of the device tensor x (n^dims FP64, n = D N) written as the reference's
// text format to `path`.
int save_txt(fc3_context* h, const double* x, int N, int D, int dims, const char* symbol,
             const char* units, const char* path)
{
    const long n = (long)D * N;
    const long E = dims == 2 ? n * n : n * n * n;
    long L = (E + kChunks - 1) / kChunks;
    L = (L + 2 * kNT - 1) / (2 * kNT) * (2 * kNT);
    const int chunks = (int)((E + L - 1) / L);
    hipStream_t s = h->stream;
    // pass 1: counts and sums per chunk
    FC3_TRY(ensure_workspace(h, Carve::need(sizeof(long long) * chunks) +
                                    Carve::need(sizeof(double) * chunks) + 1024));
    std::vector<long long> cnt(chunks);
    std::vector<double> sums(chunks);
    {
        Carve cv(h->ws, h->ws_bytes);
        long long* dc = cv.take<long long>(chunks);
        double* ds = cv.take<double>(chunks);
        hipLaunchKernelGGL(k_nnz_count, dim3(chunks), dim3(kNT), 0, s, x, E, L, dc, ds);
        FC3_HIP(hipGetLastError());
        FC3_HIP(hipMemcpyAsync(cnt.data(), dc, sizeof(long long) * chunks, hipMemcpyDeviceToHost, s));
        FC3_HIP(hipMemcpyAsync(sums.data(), ds, sizeof(double) * chunks, hipMemcpyDeviceToHost, s));
        FC3_HIP(hipStreamSynchronize(s));
    }
    std::vector<long long> off(chunks);
    long long nnz = 0;
    double total = 0.0;
    for (int c = 0; c < chunks; ++c) {
        off[c] = nnz;
        nnz += cnt[c];
        total += sums[c];
    }
    // pass 2: compaction and sort
    int bits = 1;
    while (bits < 64 && (1ull << bits) < (unsigned long long)E) ++bits;
    std::vector<unsigned long long> keys((size_t)nnz);
    std::vector<double> vals((size_t)nnz);
    if (nnz > 0) {
        size_t tmp_bytes = 0;
        FC3_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (unsigned long long*)nullptr,
                                                   (unsigned long long*)nullptr, (double*)nullptr,
                                                   (double*)nullptr, (int)std::min<long long>(nnz, 1LL << 30),
                                                   0, bits, s));
        if (nnz >= (1LL << 31))
            return set_error(FC3_ERR_UNSUPPORTED, "text output: more than 2^31 nonzeros");
        const size_t need = Carve::need(sizeof(long long) * chunks) +
                            2 * Carve::need(sizeof(unsigned long long) * nnz) +
                            2 * Carve::need(sizeof(double) * nnz) + Carve::need(tmp_bytes) + 2048;
        FC3_TRY(ensure_workspace(h, need));
        Carve cv(h->ws, h->ws_bytes);
        long long* doff = cv.take<long long>(chunks);
        unsigned long long* k0 = cv.take<unsigned long long>(nnz);
        unsigned long long* k1 = cv.take<unsigned long long>(nnz);
        double* v0 = cv.take<double>(nnz);
        double* v1 = cv.take<double>(nnz);
        void* tmp = cv.take<char>(tmp_bytes);
        if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "text output workspace");
        FC3_HIP(hipMemcpyAsync(doff, off.data(), sizeof(long long) * chunks, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_nnz_compact, dim3(chunks), dim3(kNT), 0, s, x, E, L, (int)n, N, D,
                           dims, doff, k0, v0);
        FC3_HIP(hipGetLastError());
        FC3_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, v0, v1, (int)nnz, 0,
                                                   bits, s));
        FC3_HIP(hipMemcpyAsync(keys.data(), k1, sizeof(unsigned long long) * nnz,
                               hipMemcpyDeviceToHost, s));
        FC3_HIP(hipMemcpyAsync(vals.data(), v1, sizeof(double) * nnz, hipMemcpyDeviceToHost, s));
        FC3_HIP(hipStreamSynchronize(s));
    }
    // host: header and one line per nonzero
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(FC3_ERR_ARGUMENT, std::string("cannot open ") + path + " for writing");
    std::string buf;
    buf.reserve(size_t(1) << 22);
    char num[40];
    buf += "#" + std::to_string(N) + " atoms, Sum " + symbol + " = ";
    buf.append(num, julia_float_chars(julia_round_sig7(total), num));
    buf += ", Units: ";
    buf += units ? units : "";
    buf += "\n";
    const unsigned long long uD = D, uN = N;
    for (long long t = 0; t < nnz; ++t) {
        unsigned long long k = keys[t];
        unsigned idx[6];
        int ni = 0;
        if (dims == 3) {
            const unsigned c = (unsigned)(k % uD); k /= uD;
            const unsigned b = (unsigned)(k % uD); k /= uD;
            const unsigned a = (unsigned)(k % uD); k /= uD;
            const unsigned kk = (unsigned)(k % uN); k /= uN;
            const unsigned j = (unsigned)(k % uN); k /= uN;
            const unsigned i = (unsigned)k;
            idx[0] = i; idx[1] = a; idx[2] = j; idx[3] = b; idx[4] = kk; idx[5] = c;
            ni = 6;
        } else {
            const unsigned b = (unsigned)(k % uD); k /= uD;
            const unsigned a = (unsigned)(k % uD); k /= uD;
            const unsigned j = (unsigned)(k % uN); k /= uN;
            const unsigned i = (unsigned)k;
            idx[0] = i; idx[1] = a; idx[2] = j; idx[3] = b;
            ni = 4;
        }
        for (int q = 0; q < ni; ++q) {
            const auto r = std::to_chars(num, num + sizeof num, idx[q] + 1);
            buf.append(num, r.ptr);
            buf += ' ';
        }
        buf.append(num, julia_float_chars(vals[t], num));
        buf += '\n';
        if (buf.size() > (size_t(1) << 22)) {
            if (std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) {
                std::fclose(f);
                return set_error(FC3_ERR_ARGUMENT, std::string("write failed: ") + path);
            }
            buf.clear();
        }
    }
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (std::fclose(f) != 0 || !ok) return set_error(FC3_ERR_ARGUMENT, std::string("write failed: ") + path);
    return FC3_OK;
}

}  // namespace fc3

using namespace fc3;

extern "C" {

int fc3_format_float(double x, char* buf, int len)
{
    char tmp[40];
    const int m = julia_float_chars(x, tmp);
    if (!buf || len <= m) return set_error(FC3_ERR_ARGUMENT, "buffer too small");
    std::memcpy(buf, tmp, m);
    buf[m] = '\0';
    return FC3_OK;
}

int fc3_save_second_order_txt(fc3_handle h, const double* ifc2, int n_atoms, int dim,
                              const char* units, const char* path)
{
    FC3_TRACE_RANGE();
    if (!h) return set_error(FC3_ERR_ARGUMENT, "null fc3 handle");
    FC3_HIP(hipSetDevice(h->device));
    set_current(h);
    if (!ifc2 || !path) return set_error(FC3_ERR_ARGUMENT, "ifc2 and path must be non-null");
    if (n_atoms < 1 || dim < 1 || (long long)n_atoms * dim > 65535)
        return set_error(FC3_ERR_ARGUMENT, "Incorrect dimensions");
    return save_txt(h, ifc2, n_atoms, dim, 2, "\xCE\xA6", units, path);   // "Φ"
}

int fc3_save_third_order_txt(fc3_handle h, const double* psi, int n_atoms, int dim,
                             const char* units, const char* path)
{
    FC3_TRACE_RANGE();
    if (!h) return set_error(FC3_ERR_ARGUMENT, "null fc3 handle");
    FC3_HIP(hipSetDevice(h->device));
    set_current(h);
    if (!psi || !path) return set_error(FC3_ERR_ARGUMENT, "psi and path must be non-null");
    if (n_atoms < 1 || dim < 1 || (long long)n_atoms * dim > 65535)
        return set_error(FC3_ERR_ARGUMENT, "Incorrect dimensions");
    return save_txt(h, psi, n_atoms, dim, 3, "\xCE\xA8", units, path);   // "Ψ"
}

}  // extern "C"
