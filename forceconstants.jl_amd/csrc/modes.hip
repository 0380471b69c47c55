// Device-side dynamical matrix and normal modes (src/dynmat.jl:5-33, 123-131):
//   D[i, j] = Phi2[i, j] / sqrt(m_I m_J)          (I, J the atoms of rows i, j)
//   eigen(Hermitian(D)): eigenvalues ascending, eigenvectors as columns, from the upper
//   triangle as Julia's Hermitian wrapper reads it.
// The eigensolver is rocSOLVER's divide-and-conquer dsyevd in FP64 on the library's
// stream; the caller zeroes the rigid-translation eigenvalues as get_modes does.
// rocSOLVER's device scratch is carved from the library's workspace (sized by a rocBLAS
// size query) and handed to the rocBLAS handle for the call only, so the handle never
// allocates or frees device memory of its own between the library's calls.
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

int dsyevd_workspace_bytes(fc3_context* h, int n, size_t* bytes)
{
    void* hb = nullptr;
    FC3_TRY(blas_handle(h, &hb));
    rocblas_handle b = static_cast<rocblas_handle>(hb);
    size_t sz = 0;
    if (rocblas_start_device_memory_size_query(b) != rocblas_status_success)
        return set_error(FC3_ERR_HIP, "rocBLAS device memory size query could not start");
    // in query mode the call only reports its scratch size (no launch, no access)
    (void)rocsolver_dsyevd(b, rocblas_evect_original, rocblas_fill_upper, n, nullptr, n, nullptr,
                           nullptr, nullptr);
    if (rocblas_stop_device_memory_size_query(b, &sz) != rocblas_status_success)
        return set_error(FC3_ERR_HIP, "rocBLAS device memory size query failed");
    *bytes = sz;
    return FC3_OK;
}

int dynmat_eigen(fc3_context* h, double* A, int n, const double* d_masses, double* W, double* E,
                 int* info, double* h_w, void* solver_ws, size_t solver_ws_bytes)
{
    hipStream_t s = h->stream;
    const long nn = (long)n * n;
    hipLaunchKernelGGL(k_dynmat, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, A, n, d_masses);
    FC3_HIP(hipGetLastError());
    void* hb = nullptr;
    FC3_TRY(blas_handle(h, &hb));
    rocblas_handle b = static_cast<rocblas_handle>(hb);
    if (solver_ws_bytes &&
        rocblas_set_workspace(b, solver_ws, solver_ws_bytes) != rocblas_status_success)
        return set_error(FC3_ERR_HIP, "rocBLAS: setting the eigensolver workspace failed");
    const rocblas_status st = rocsolver_dsyevd(b, rocblas_evect_original, rocblas_fill_upper, n, A,
                                               n, W, E, info);
    // back to rocBLAS-managed memory, which the handle holds none of now: the workspace
    // above belongs to the library and may be reallocated by its next call
    const rocblas_status st2 = solver_ws_bytes ? rocblas_set_workspace(b, nullptr, 0)
                                               : rocblas_status_success;
    if (st != rocblas_status_success)
        return set_error(FC3_ERR_HIP, std::string("rocSOLVER dsyevd: ") + rocblas_status_to_string(st));
    if (st2 != rocblas_status_success)
        return set_error(FC3_ERR_HIP, "rocBLAS: releasing the eigensolver workspace failed");
    int hinfo = 0;
    FC3_HIP(hipMemcpyAsync(h_w, W, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    FC3_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
    FC3_HIP(hipStreamSynchronize(s));
    if (hinfo != 0)
        return set_error(FC3_ERR_HIP, "rocSOLVER dsyevd did not converge (info " +
                                          std::to_string(hinfo) + ")");
    return FC3_OK;
}

int dynmat_scale(hipStream_t s, double* A, int n, const double* d_masses)
{
    const long nn = (long)n * n;
    hipLaunchKernelGGL(k_dynmat, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, A, n, d_masses);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

int dsyevd_batched_workspace_bytes(fc3_context* h, int n, int G, size_t* bytes)
{
    void* hb = nullptr;
    FC3_TRY(blas_handle(h, &hb));
    rocblas_handle b = static_cast<rocblas_handle>(hb);
    size_t sz = 0;
    if (rocblas_start_device_memory_size_query(b) != rocblas_status_success)
        return set_error(FC3_ERR_HIP, "rocBLAS device memory size query could not start");
    (void)rocsolver_dsyevd_strided_batched(b, rocblas_evect_original, rocblas_fill_upper, n,
                                           nullptr, n, (rocblas_stride)n * n, nullptr, n, nullptr,
                                           n, nullptr, G);
    if (rocblas_stop_device_memory_size_query(b, &sz) != rocblas_status_success)
        return set_error(FC3_ERR_HIP, "rocBLAS device memory size query failed");
    *bytes = sz;
    return FC3_OK;
}

// G matrices A + g n^2 (upper triangles) -> eigenvalues W + g n (ascending) and
// eigenvectors in place, one rocSOLVER call on the library's stream; info: G ints.
// No host synchronisation: the caller reads W and info after the stream.
int eigen_strided_batched(fc3_context* h, double* A, int n, int G, double* W, double* E,
                          int* info, void* solver_ws, size_t solver_ws_bytes)
{
    void* hb = nullptr;
    FC3_TRY(blas_handle(h, &hb));
    rocblas_handle b = static_cast<rocblas_handle>(hb);
    if (solver_ws_bytes &&
        rocblas_set_workspace(b, solver_ws, solver_ws_bytes) != rocblas_status_success)
        return set_error(FC3_ERR_HIP, "rocBLAS: setting the eigensolver workspace failed");
    const rocblas_status st = rocsolver_dsyevd_strided_batched(
        b, rocblas_evect_original, rocblas_fill_upper, n, A, n, (rocblas_stride)n * n, W, n, E, n,
        info, G);
    const rocblas_status st2 = solver_ws_bytes ? rocblas_set_workspace(b, nullptr, 0)
                                               : rocblas_status_success;
    if (st != rocblas_status_success)
        return set_error(FC3_ERR_HIP, std::string("rocSOLVER dsyevd (batched): ") +
                                          rocblas_status_to_string(st));
    if (st2 != rocblas_status_success)
        return set_error(FC3_ERR_HIP, "rocBLAS: releasing the eigensolver workspace failed");
    return FC3_OK;
}

int to_f32(hipStream_t s, const double* a, long cnt, float* b)
{
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, a, cnt, b);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

}  // namespace fc3
