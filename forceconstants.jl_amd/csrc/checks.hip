// Checks the reference runs on finished force constants (check_asr / check_symmetry of
// second_order and third_order, src/ifc.jl:14-15,33):
//
//   asr_satisfied(ifc2, N, D, tol)   src/asr.jl:43-59: for every atom i and components
//       a, b, s = sum(ifc2[D(i-1)+a, b:D:end]); the rule fails where s > tol (signed: a
//       negative sum never fails);
//   asr_satisfied(ifc3, N, D, tol)   src/asr.jl:62-74: the same for every slice
//       ifc3[:, :, D(k-1)+c];
//   issymmetric(ifc2)                Phi == Phi^T entry by entry (a NaN fails).
//
// One thread per (row, b, slice): the thread sums its row's n / D entries at stride D
// columns in FP64, in ascending column order.  Neighbouring threads own neighbouring rows,
// so every load of a wave is one contiguous 512-byte run: the pass reads the tensor once
// at HBM rate (8 n^3 bytes for IFC3).  Violations are counted, and the first one in the
// reference's loop order (slice, then atom, then a, then b) is kept with its sum, for the
// caller's message.
#include <cmath>
#include <string>

#include "fc3_internal.hpp"

namespace fc3 {
namespace {

constexpr int kNT = 256;

unsigned blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct AsrResult {
    unsigned long long count;
    unsigned long long first;   // smallest loop-order key of a violation (~0 = none)
    double first_sum;
};

__global__ void k_asr_init(AsrResult* r)
{
    if (threadIdx.x == 0) {
        r->count = 0;
        r->first = ~0ull;
        r->first_sum = 0.0;
    }
}

// grid (ceil(n / kNT), slices, D): x = row r = D i + a, y = slice, z = b
__global__ void __launch_bounds__(kNT) k_asr_rows(const double* __restrict__ A, int n, int D,
                                                  double tol, AsrResult* __restrict__ out)
{
    const int r = blockIdx.x * kNT + threadIdx.x;
    const long slice = blockIdx.y;
    const int b = blockIdx.z;
    bool bad = false;
    unsigned long long key = ~0ull;
    double s = 0.0;
    if (r < n) {
        const double* p = A + (size_t)slice * n * n + (size_t)b * n + r;
        const size_t step = (size_t)D * n;
        const int cnt = n / D;
        for (int j = 0; j < cnt; ++j) s += p[j * step];
        bad = s > tol;
        // reference loop order: slice (k, c), atom i, a, b
        key = ((unsigned long long)slice * n + r) * D + b;
    }
    const unsigned long long nbad = __popcll(__ballot(bad));
    if (nbad == 0) return;
    if ((threadIdx.x & 63) == 0) atomicAdd(&out->count, nbad);
    if (bad) atomicMin(&out->first, key);
    // the thread that holds the first violation (known only after every wave has posted)
    // is found by the host from `first`; its sum is written by a second, tiny launch
}

__global__ void k_asr_first_sum(const double* __restrict__ A, int n, int D, AsrResult* r)
{
    if (threadIdx.x != 0 || r->first == ~0ull) return;
    const unsigned long long key = r->first;
    const int b = (int)(key % D);
    const unsigned long long rs = key / D;
    const int row = (int)(rs % n);
    const long slice = (long)(rs / n);
    const double* p = A + (size_t)slice * n * n + (size_t)b * n + row;
    double s = 0.0;
    for (int j = 0; j < n / D; ++j) s += p[(size_t)j * D * n];
    r->first_sum = s;
}

__global__ void __launch_bounds__(kNT) k_sym2(const double* __restrict__ A, int n,
                                              unsigned long long* __restrict__ count)
{
    const long e = (long)blockIdx.x * kNT + threadIdx.x;
    bool bad = false;
    if (e < (long)n * n) {
        const int r = (int)(e % n), c = (int)(e / n);
        if (r < c) bad = !(A[e] == A[(size_t)r * n + c]);
    }
    const unsigned long long nbad = __popcll(__ballot(bad));
    if (nbad && (threadIdx.x & 63) == 0) atomicAdd(count, nbad);
}

__global__ void k_zero_u64(unsigned long long* p)
{
    if (threadIdx.x == 0) *p = 0;
}

}  // namespace
}  // namespace fc3

using namespace fc3;

extern "C" {

int fc3_asr_satisfied(fc3_handle h, const double* ifc, int order, int n_atoms, int dim, double tol,
                      long long* n_violations, long long* first_violation, double* first_sum)
{
    FC3_TRACE_RANGE();
    if (!h) return set_error(FC3_ERR_ARGUMENT, "null fc3 handle");
    FC3_HIP(hipSetDevice(h->device));
    set_current(h);
    if (!ifc || !n_violations) return set_error(FC3_ERR_ARGUMENT, "ifc and n_violations must be non-null");
    if (order != 2 && order != 3) return set_error(FC3_ERR_ARGUMENT, "order must be 2 or 3");
    if (n_atoms < 1 || dim < 1 || (long long)n_atoms * dim > 65535)
        return set_error(FC3_ERR_ARGUMENT, "Incorrect dimensions");
    if (std::isnan(tol)) return set_error(FC3_ERR_ARGUMENT, "tol must not be NaN");
    const int n = n_atoms * dim;
    FC3_TRY(ensure_workspace(h, sizeof(AsrResult)));
    AsrResult* d = static_cast<AsrResult*>(h->ws);
    hipStream_t s = h->stream;
    hipLaunchKernelGGL(k_asr_init, dim3(1), dim3(64), 0, s, d);
    const unsigned slices = order == 3 ? (unsigned)n : 1u;
    hipLaunchKernelGGL(k_asr_rows, dim3(blocks(n, kNT), slices, (unsigned)dim), dim3(kNT), 0, s,
                       ifc, n, dim, tol, d);
    hipLaunchKernelGGL(k_asr_first_sum, dim3(1), dim3(64), 0, s, ifc, n, dim, d);
    FC3_HIP(hipGetLastError());
    AsrResult hr;
    FC3_HIP(hipMemcpyAsync(&hr, d, sizeof(hr), hipMemcpyDeviceToHost, s));
    FC3_HIP(hipStreamSynchronize(s));
    *n_violations = (long long)hr.count;
    if (first_violation) *first_violation = hr.count ? (long long)hr.first : -1;
    if (first_sum) *first_sum = hr.count ? hr.first_sum : 0.0;
    return FC3_OK;
}

int fc3_issymmetric2(fc3_handle h, const double* ifc2, int n, long long* n_asymmetric)
{
    FC3_TRACE_RANGE();
    if (!h) return set_error(FC3_ERR_ARGUMENT, "null fc3 handle");
    FC3_HIP(hipSetDevice(h->device));
    set_current(h);
    if (!ifc2 || !n_asymmetric) return set_error(FC3_ERR_ARGUMENT, "ifc2 and n_asymmetric must be non-null");
    if (n < 1 || n > 65535) return set_error(FC3_ERR_ARGUMENT, "Incorrect dimensions");
    FC3_TRY(ensure_workspace(h, sizeof(unsigned long long)));
    auto* d = static_cast<unsigned long long*>(h->ws);
    hipStream_t s = h->stream;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, d);
    hipLaunchKernelGGL(k_sym2, dim3(blocks((long)n * n, kNT)), dim3(kNT), 0, s, ifc2, n, d);
    FC3_HIP(hipGetLastError());
    unsigned long long c = 0;
    FC3_HIP(hipMemcpyAsync(&c, d, sizeof(c), hipMemcpyDeviceToHost, s));
    FC3_HIP(hipStreamSynchronize(s));
    *n_asymmetric = (long long)c;
    return FC3_OK;
}

}  // extern "C"
