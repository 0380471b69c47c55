// Device-side dynamical matrix and normal modes (src/dynmat.jl:5-33, 123-131):
//   D[i, j] = Phi2[i, j] / sqrt(m_I m_J)          (I, J the atoms of rows i, j)
//   eigen(Hermitian(D)): eigenvalues ascending, eigenvectors as columns, from the upper
//   triangle as Julia's Hermitian wrapper reads it.
// The eigensolver is rocSOLVER's divide-and-conquer dsyevd in FP64 on the library's
// stream; the caller zeroes the rigid-translation eigenvalues as get_modes does.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "fc3_internal.hpp"

namespace fc3 {
namespace {

__global__ void k_dynmat(double* __restrict__ A, int n, const double* __restrict__ m)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = (int)(e % n), j = (int)(e / n);
    A[e] /= sqrt(m[i / 3] * m[j / 3]);
}

__global__ void k_to_f32(const double* __restrict__ a, long cnt, float* __restrict__ b)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt) b[e] = (float)a[e];
}

}  // namespace

int dynmat_eigen(fc3_context* h, double* A, int n, const double* d_masses, double* W, double* E,
                 int* info, double* h_w)
{
    hipStream_t s = h->stream;
    const long nn = (long)n * n;
    hipLaunchKernelGGL(k_dynmat, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, A, n, d_masses);
    FC3_HIP(hipGetLastError());
    void* hb = nullptr;
    FC3_TRY(blas_handle(h, &hb));
    const rocblas_status st = rocsolver_dsyevd(static_cast<rocblas_handle>(hb), rocblas_evect_original,
                                               rocblas_fill_upper, n, A, n, W, E, info);
    if (st != rocblas_status_success)
        return set_error(FC3_ERR_HIP, std::string("rocSOLVER dsyevd: ") + rocblas_status_to_string(st));
    int hinfo = 0;
    FC3_HIP(hipMemcpyAsync(h_w, W, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    FC3_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
    FC3_HIP(hipStreamSynchronize(s));
    if (hinfo != 0)
        return set_error(FC3_ERR_HIP, "rocSOLVER dsyevd did not converge (info " +
                                          std::to_string(hinfo) + ")");
    return FC3_OK;
}

int to_f32(hipStream_t s, const double* a, long cnt, float* b)
{
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, a, cnt, b);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

}  // namespace fc3
