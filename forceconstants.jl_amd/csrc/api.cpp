#include <algorithm>
// C-ABI entry points (include/fc3.h): argument validation with the reference's
// assertion messages, workspace management and dispatch to the HIP kernels.
#include <cmath>
#include <cstring>
#include <string>
#include <functional>
#include <vector>

#include "fc3_internal.hpp"

namespace fc3 {

namespace {
thread_local std::string g_err = "";
}

int set_error(int status, const std::string& msg)
{
    g_err = msg;
    return status;
}

const char* last_error() { return g_err.c_str(); }

int ensure_workspace(fc3_context* h, size_t bytes)
{
    if (bytes <= h->ws_bytes) return FC3_OK;
    if (h->ws) {
        FC3_HIP(hipStreamSynchronize(h->stream));
        FC3_HIP(hipFree(h->ws));
        h->ws = nullptr;
        h->ws_bytes = 0;
    }
    const size_t want = (bytes + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
    hipError_t e = hipMalloc(&h->ws, want);
    if (e != hipSuccess) {
        h->ws = nullptr;
        (void)hipGetLastError();   // a caller may retry smaller: leave no error behind
        return set_error(FC3_ERR_OUT_OF_MEMORY,
                         "workspace allocation of " + std::to_string(want) + " bytes failed: " +
                             hipGetErrorString(e));
    }
    h->ws_bytes = want;
    return FC3_OK;
}

}  // namespace fc3

using namespace fc3;

namespace {

int check_handle(fc3_handle h)
{
    if (!h) return set_error(FC3_ERR_ARGUMENT, "null fc3 handle");
    FC3_HIP(hipSetDevice(h->device));
    set_current(h);
    return FC3_OK;
}

int check_system(const double* pos, int n_atoms, const double* box)
{
    if (!pos || !box) return set_error(FC3_ERR_ARGUMENT, "positions and box must be non-null");
    if (n_atoms < 1) return set_error(FC3_ERR_ARGUMENT, "n_atoms must be >= 1");
    if ((long long)n_atoms * 3 > 65535)
        return set_error(FC3_ERR_UNSUPPORTED, "3*n_atoms must be <= 65535");
    for (int c = 0; c < 3; ++c)
        if (!(box[c] > 0.0) || !std::isfinite(box[c]))
            return set_error(FC3_ERR_ARGUMENT, "box sizes must be positive and finite");
    return FC3_OK;
}

// Neighbour list + position upload into the handle workspace.  `extra` bytes of
// workspace are reserved after the neighbour data for the caller.
struct Prepared {
    double* pos = nullptr;
    double* masses = nullptr;
    NeighborList nl;
    char* extra = nullptr;
    size_t extra_bytes = 0;
};

int prepare(fc3_handle h, const double* pos, int N, const double* box, double r_cut,
            const double* masses, const std::function<size_t(int)>& extra_bytes, Prepared& P)
{
    // phase 1: counts (small workspace)
    const size_t head = Carve::need(sizeof(double) * 3 * N) + Carve::need(sizeof(double) * N) +
                        Carve::need(sizeof(int) * N) + Carve::need(sizeof(int));
    FC3_TRY(ensure_workspace(h, head));
    {
        Carve cv(h->ws, h->ws_bytes);
        P.pos = cv.take<double>(3 * (size_t)N);
        P.masses = cv.take<double>(N);
        int* count = cv.take<int>(N);
        int* maxc = cv.take<int>(1);
        FC3_HIP(hipMemcpyAsync(P.pos, pos, sizeof(double) * 3 * N, hipMemcpyHostToDevice,
                               h->stream));
        int hmax = 0;
        FC3_TRY(neighbor_count(h->stream, P.pos, N, box, r_cut, count, maxc, &hmax));
        P.nl.cap = hmax > 0 ? hmax : 1;
    }
    // phase 2: full workspace
    const int cap = P.nl.cap;
    const size_t slots = (size_t)N * cap;
    const size_t extra = extra_bytes(cap);
    const size_t total = head + 2 * Carve::need(sizeof(int) * slots) + Carve::need(extra) + 1024;
    FC3_TRY(ensure_workspace(h, total));  // may move the workspace: re-upload below
    Carve cv(h->ws, h->ws_bytes);
    P.pos = cv.take<double>(3 * (size_t)N);
    P.masses = cv.take<double>(N);
    P.nl.n_atoms = N;
    P.nl.count = cv.take<int>(N);
    cv.take<int>(1);
    P.nl.nbr = cv.take<int>(slots);
    P.nl.rev = cv.take<int>(slots);
    P.extra = cv.take<char>(extra);
    P.extra_bytes = extra;
    FC3_HIP(hipMemcpyAsync(P.pos, pos, sizeof(double) * 3 * N, hipMemcpyHostToDevice, h->stream));
    if (masses)
        FC3_HIP(hipMemcpyAsync(P.masses, masses, sizeof(double) * N, hipMemcpyHostToDevice,
                               h->stream));
    // counts again (cheap) so that they live in the final workspace
    int* maxc = reinterpret_cast<int*>(reinterpret_cast<char*>(P.nl.count) +
                                       Carve::need(sizeof(int) * N));
    int hmax = 0;
    FC3_TRY(neighbor_count(h->stream, P.pos, N, box, r_cut, P.nl.count, maxc, &hmax));
    if (hmax > cap) return set_error(FC3_ERR_CAPACITY, "neighbour count changed between passes");
    FC3_TRY(neighbor_fill(h->stream, P.pos, N, box, r_cut, P.nl));
    return FC3_OK;
}

}  // namespace

extern "C" {

int fc3_version(void) { return 100; }

const char* fc3_last_error(void) { return fc3::last_error(); }

const char* fc3_status_string(int status)
{
    switch (status) {
        case FC3_OK: return "ok";
        case FC3_ERR_ARGUMENT: return "invalid argument";
        case FC3_ERR_CUTOFF: return "cutoff larger than box";
        case FC3_ERR_BLOCK_SIZE: return "invalid block_size";
        case FC3_ERR_OUT_OF_MEMORY: return "out of device memory";
        case FC3_ERR_HIP: return "HIP runtime error";
        case FC3_ERR_UNSUPPORTED: return "unsupported";
        case FC3_ERR_CAPACITY: return "capacity exceeded";
        default: return "unknown status";
    }
}

int fc3_create(fc3_handle* out, int device)
{
    if (!out) return set_error(FC3_ERR_ARGUMENT, "out must be non-null");
    *out = nullptr;
    int ndev = 0;
    FC3_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return set_error(FC3_ERR_ARGUMENT, "device " + std::to_string(device) +
                                               " out of range (" + std::to_string(ndev) +
                                               " devices)");
    FC3_HIP(hipSetDevice(device));
    auto* h = new fc3_context();
    h->device = device;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&h->ev_start);
    if (e == hipSuccess) e = hipEventCreate(&h->ev_stop);
    if (e != hipSuccess) {
        delete h;
        return set_error(FC3_ERR_HIP, std::string("fc3_create: ") + hipGetErrorString(e));
    }
    // the static scratch of the allocation-free entry points, looked up on this device
    if (const int st = noalloc_init_device(device); st != FC3_OK) {
        fc3_destroy(h);
        return st;
    }
    *out = h;
    return FC3_OK;
}

int fc3_destroy(fc3_handle h)
{
    if (!h) return FC3_OK;
    hipSetDevice(h->device);
    hipStreamSynchronize(h->stream);
    release_blas(h);
    release_batch(h);
    if (h->aux) hipStreamSynchronize(h->aux);
    if (h->ws) hipFree(h->ws);
    for (auto e : h->aux_ev)
        if (e) hipEventDestroy(e);
    if (h->aux) hipStreamDestroy(h->aux);
    for (auto& r : h->prof.records) {
        hipEventDestroy(r.start);
        hipEventDestroy(r.stop);
    }
    for (auto e : h->prof.pool) hipEventDestroy(e);
    if (h->ev_start) hipEventDestroy(h->ev_start);
    if (h->ev_stop) hipEventDestroy(h->ev_stop);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return FC3_OK;
}

int fc3_release_workspace(fc3_handle h)
{
    FC3_TRY(check_handle(h));
    FC3_HIP(hipStreamSynchronize(h->stream));
    if (h->ws) FC3_HIP(hipFree(h->ws));
    h->ws = nullptr;
    h->ws_bytes = 0;
    return FC3_OK;
}

int fc3_synchronize(fc3_handle h)
{
    FC3_TRY(check_handle(h));
    FC3_HIP(hipStreamSynchronize(h->stream));
    return FC3_OK;
}

int fc3_get_device(fc3_handle h, int* device)
{
    FC3_TRY(check_handle(h));
    if (!device) return set_error(FC3_ERR_ARGUMENT, "device must be non-null");
    *device = h->device;
    return FC3_OK;
}

int fc3_device_alloc(fc3_handle h, void** dptr, size_t bytes)
{
    FC3_TRY(check_handle(h));
    if (!dptr) return set_error(FC3_ERR_ARGUMENT, "dptr must be non-null");
    *dptr = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess) {
        *dptr = nullptr;
        return set_error(FC3_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(bytes) +
                                                    " bytes failed: " + hipGetErrorString(e));
    }
    return FC3_OK;
}

int fc3_device_free(fc3_handle h, void* dptr)
{
    FC3_TRY(check_handle(h));
    if (!dptr) return FC3_OK;
    FC3_HIP(hipStreamSynchronize(h->stream));
    FC3_HIP(hipFree(dptr));
    return FC3_OK;
}

int fc3_memcpy_h2d(fc3_handle h, void* dst, const void* src, size_t bytes)
{
    FC3_TRY(check_handle(h));
    if (bytes == 0) return FC3_OK;
    if (!dst || !src) return set_error(FC3_ERR_ARGUMENT, "null pointer in fc3_memcpy_h2d");
    FC3_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    FC3_HIP(hipStreamSynchronize(h->stream));
    return FC3_OK;
}

int fc3_memcpy_d2h(fc3_handle h, void* dst, const void* src, size_t bytes)
{
    FC3_TRY(check_handle(h));
    if (bytes == 0) return FC3_OK;
    if (!dst || !src) return set_error(FC3_ERR_ARGUMENT, "null pointer in fc3_memcpy_d2h");
    FC3_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    FC3_HIP(hipStreamSynchronize(h->stream));
    return FC3_OK;
}

int fc3_memset_zero(fc3_handle h, void* dst, size_t bytes)
{
    FC3_TRY(check_handle(h));
    if (bytes == 0) return FC3_OK;
    if (!dst) return set_error(FC3_ERR_ARGUMENT, "null pointer in fc3_memset_zero");
    FC3_HIP(hipMemsetAsync(dst, 0, bytes, h->stream));
    return FC3_OK;
}

int fc3_device_memory(fc3_handle h, size_t* free_bytes, size_t* total_bytes)
{
    FC3_TRY(check_handle(h));
    size_t f = 0, t = 0;
    FC3_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return FC3_OK;
}

int fc3_timer_start(fc3_handle h)
{
    FC3_TRY(check_handle(h));
    FC3_HIP(hipEventRecord(h->ev_start, h->stream));
    return FC3_OK;
}

int fc3_timer_stop(fc3_handle h, float* ms)
{
    FC3_TRY(check_handle(h));
    FC3_HIP(hipEventRecord(h->ev_stop, h->stream));
    FC3_HIP(hipEventSynchronize(h->ev_stop));
    float t = 0.0f;
    FC3_HIP(hipEventElapsedTime(&t, h->ev_start, h->ev_stop));
    if (ms) *ms = t;
    return FC3_OK;
}

int fc3_second_order_lj(fc3_handle h, const double* positions, int n_atoms, const double box[3],
                        double sigma, double epsilon, double pot_r_cut, double r_cut, double tol,
                        double* ifc2)
{
    FC3_TRACE_RANGE();
    (void)pot_r_cut;
    FC3_TRY(check_handle(h));
    FC3_TRY(check_system(positions, n_atoms, box));
    if (!ifc2) return set_error(FC3_ERR_ARGUMENT, "ifc2 must be non-null");
    const int N = n_atoms;
    Prepared P;
    const size_t per_slot = Carve::need(sizeof(double) * 27) + Carve::need(sizeof(double) * 9);
    FC3_TRY(prepare(h, positions, N, box, r_cut, nullptr, [&](int cap) {
        return sizeof(double) * ((size_t)N * cap * (27 + 9) + (size_t)N * (27 + 9)) + 4096;
    }, P));
    (void)per_slot;
    const size_t slots = (size_t)N * P.nl.cap;
    double* E3 = reinterpret_cast<double*>(P.extra);
    double* E2 = E3 + slots * 27;
    double* S2 = E2 + slots * 9;
    FC3_TRY(lj_pair_blocks(h->stream, P.pos, box, sigma, epsilon, P.nl, E3, nullptr, E2, S2));
    FC3_TRY(lj_write_ifc2(h->stream, P.nl, E2, S2, tol, ifc2));
    return FC3_OK;
}

int fc3_third_order_lj(fc3_handle h, const double* positions, int n_atoms, const double box[3],
                       double sigma, double epsilon, double pot_r_cut, double r_cut, double tol,
                       const double* masses, double* psi64, float* psi32)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    FC3_TRY(check_system(positions, n_atoms, box));
    for (int c = 0; c < 3; ++c)
        if (!(pot_r_cut < box[c])) return set_error(FC3_ERR_CUTOFF, "Cutoff larger than L/2");
    if (!psi64 && !psi32) return set_error(FC3_ERR_ARGUMENT, "psi64 and psi32 are both null");
    if (masses)
        for (int i = 0; i < n_atoms; ++i)
            if (!(masses[i] > 0.0))
                return set_error(FC3_ERR_ARGUMENT, "masses must be positive");
    const int N = n_atoms;
    Prepared P;
    FC3_TRY(prepare(h, positions, N, box, r_cut, masses, [&](int cap) {
        return sizeof(double) * ((size_t)N * cap * (27 + 9) + (size_t)N * (27 + 9)) + 4096;
    }, P));
    const size_t slots = (size_t)N * P.nl.cap;
    double* E3 = reinterpret_cast<double*>(P.extra);
    double* E2 = E3 + slots * 27;
    double* S3 = E2 + slots * 9;
    FC3_TRY(lj_pair_blocks(h->stream, P.pos, box, sigma, epsilon, P.nl, E3, S3, E2, nullptr));
    Ifc3Blocks b;
    b.E = E3;
    b.S = S3;
    FC3_TRY(write_ifc3_pairwise(h->stream, P.nl, b, tol, masses ? P.masses : nullptr, psi64,
                                psi32));
    return FC3_OK;
}

int fc3_third_order_mcc3_lj(fc3_handle h, const double* positions, int n_atoms,
                            const double box[3], double sigma, double epsilon, double pot_r_cut,
                            double r_cut, double tol, const double* masses, const float* phi,
                            float* k3)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    FC3_TRY(check_system(positions, n_atoms, box));
    for (int c = 0; c < 3; ++c)
        if (!(pot_r_cut < box[c])) return set_error(FC3_ERR_CUTOFF, "Cutoff larger than L/2");
    if (!phi || !k3) return set_error(FC3_ERR_ARGUMENT, "phi, k3 must be non-null");
    if (masses)
        for (int i = 0; i < n_atoms; ++i)
            if (!(masses[i] > 0.0))
                return set_error(FC3_ERR_ARGUMENT, "masses must be positive");
    const int N = n_atoms;
    const long n = 3L * N;
    if (n % 4 != 0 || n < 16 || n * n >= (1L << 31) || (uintptr_t)k3 % 16 != 0)
        return set_error(FC3_ERR_UNSUPPORTED,
                         "fused third_order + mcc3: needs 3N % 4 == 0, 3N >= 16, aligned k3");
    // U: the gathered mode-1 result, 2 (3n) Kc bf16 per atom
    auto u_bytes = [&](int cap) {
        return Carve::need(sizeof(uint16_t) * 2 * 3 * (size_t)n * mcc3_gather_k(cap) * N);
    };
    // with room for the FP32 intermediate (mode 3 without a split pass) when the device
    // has it, else with the bounded split scratch (the same K3)
    // (neighbourhoods past 64 KB of LDS: room for their coefficients in global memory,
    // which mcc3_from_gathered's scratch holds until it starts)
    auto need = [&](int cap, bool direct) {
        return Carve::need(sizeof(double) * ((size_t)N * cap * (27 + 9) + (size_t)N * (27 + 9))) +
               Carve::need(sizeof(float) * (size_t)n * n) + u_bytes(cap) +
               std::max(mcc3_from_gathered_bytes((int)n, cap, direct),
                        lj_mode1_sparse_lds(cap) > 64 * 1024 ? lj_mode1_global_bytes(N, cap)
                                                             : size_t(0)) +
               4096;
    };
    Prepared P;
    int rc = prepare(h, positions, N, box, r_cut, masses,
                     [&](int cap) { return need(cap, true); }, P);
    if (rc == FC3_ERR_OUT_OF_MEMORY)
        rc = prepare(h, positions, N, box, r_cut, masses,
                     [&](int cap) { return need(cap, false); }, P);
    FC3_TRY(rc);
    const int Kc = mcc3_gather_k(P.nl.cap);
    Carve cv(P.extra, P.extra_bytes);
    const size_t slots = (size_t)N * P.nl.cap;
    double* E3 = cv.take<double>(slots * 27 + slots * 9 + (size_t)N * (27 + 9));
    double* E2 = E3 + slots * 27;
    double* S3 = E2 + slots * 9;
    float* phiT = cv.take<float>((size_t)n * n);
    __bf16* U = cv.take<__bf16>(2 * 3 * (size_t)n * Kc * N);
    char* rest = cv.take<char>(0);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "fused third_order + mcc3 workspace");
    FC3_TRY(lj_pair_blocks(h->stream, P.pos, box, sigma, epsilon, P.nl, E3, S3, E2, nullptr));
    Ifc3Blocks b;
    b.E = E3;
    b.S = S3;
    FC3_TRY(transpose_square(h->stream, phi, (int)n, phiT));
    FC3_TRY(lj_mode1_gathered(h->stream, P.nl, b, tol, masses ? P.masses : nullptr, phiT, Kc, U,
                              rest, cv.cap - cv.off));
    return mcc3_from_gathered(h->stream, U, phi, phiT, (int)n, P.nl.count, P.nl.nbr, P.nl.cap,
                              k3, rest, cv.cap - cv.off);
}

int fc3_mass_weight_third_order(fc3_handle h, double* psi64, int n_atoms, int dim,
                                const double* masses)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    // @assert D ∈ [1,2,3] (third_order_analytical.jl:190)
    if (dim < 1 || dim > 3) return set_error(FC3_ERR_ARGUMENT, "D ∈ [1, 2, 3]");
    if (!psi64 || !masses || n_atoms < 1)
        return set_error(FC3_ERR_ARGUMENT, "psi64/masses must be non-null, n_atoms >= 1");
    if ((long long)n_atoms * dim > 65535)
        return set_error(FC3_ERR_UNSUPPORTED, "D*n_atoms must be <= 65535");
    FC3_TRY(ensure_workspace(h, sizeof(double) * n_atoms + 256));
    double* dm = static_cast<double*>(h->ws);
    FC3_HIP(hipMemcpyAsync(dm, masses, sizeof(double) * n_atoms, hipMemcpyHostToDevice,
                           h->stream));
    FC3_TRY(mass_weight3(h->stream, psi64, n_atoms, dim, dm));
    return FC3_OK;
}

int fc3_second_order_sw(fc3_handle h, const double* positions, int n_atoms, const double box[3],
                        const double* sw_params, double r_cut, double tol, double* ifc2)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    FC3_TRY(check_system(positions, n_atoms, box));
    if (!sw_params || !ifc2) return set_error(FC3_ERR_ARGUMENT, "sw_params, ifc2 must be non-null");
    if (!(r_cut <= sw_params[2] * sw_params[1]))
        return set_error(FC3_ERR_ARGUMENT,
                         "For SW silicon force constant cutoff must be less than potential cutoff");
    const int N = n_atoms;
    Prepared P;
    FC3_TRY(prepare(h, positions, N, box, r_cut, nullptr,
                    [&](int cap) { return sw_workspace_bytes(N, cap, false); }, P));
    return sw_assemble(h->stream, P.pos, box, sw_params, sw_params[10], sw_params[5], P.nl,
                       P.extra, P.extra_bytes, tol, nullptr, nullptr, nullptr, ifc2);
}

int fc3_modes_lj(fc3_handle h, const double* positions, int n_atoms, const double box[3],
                 double sigma, double epsilon, double pot_r_cut, double r_cut, double tol,
                 const double* masses, int num_rigid_translation, double* freqs_sq,
                 double* phi64, float* phi32)
{
    FC3_TRACE_RANGE();
    (void)pot_r_cut;
    FC3_TRY(check_handle(h));
    FC3_TRY(check_system(positions, n_atoms, box));
    if (!masses || !freqs_sq) return set_error(FC3_ERR_ARGUMENT, "masses, freqs_sq must be non-null");
    if (num_rigid_translation < 0 || num_rigid_translation > 3 * n_atoms)
        return set_error(FC3_ERR_ARGUMENT, "num_rigid_translation out of range");
    for (int i = 0; i < n_atoms; ++i)
        if (!(masses[i] > 0.0)) return set_error(FC3_ERR_ARGUMENT, "masses must be positive");
    const int N = n_atoms, n = 3 * N;
    size_t sv = 0;
    FC3_TRY(dsyevd_workspace_bytes(h, n, &sv));
    Prepared P;
    FC3_TRY(prepare(h, positions, N, box, r_cut, masses, [&](int cap) {
        return Carve::need(sizeof(double) * ((size_t)N * cap * (27 + 9) + (size_t)N * (27 + 9))) +
               Carve::need(sizeof(double) * (size_t)n * n) + 2 * Carve::need(sizeof(double) * n) +
               Carve::need(sizeof(int)) + Carve::need(sv) + 4096;
    }, P));
    Carve cv(P.extra, P.extra_bytes);
    const size_t slots = (size_t)N * P.nl.cap;
    double* E3 = cv.take<double>(slots * 27 + slots * 9 + (size_t)N * (27 + 9));
    double* E2 = E3 + slots * 27;
    double* S2 = E2 + slots * 9;
    double* A = cv.take<double>((size_t)n * n);
    double* W = cv.take<double>(n);
    double* E = cv.take<double>(n);
    int* info = cv.take<int>(1);
    char* svw = cv.take<char>(sv);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "modes workspace");
    FC3_TRY(lj_pair_blocks(h->stream, P.pos, box, sigma, epsilon, P.nl, E3, nullptr, E2, S2));
    FC3_TRY(lj_write_ifc2(h->stream, P.nl, E2, S2, tol, A));
    FC3_TRY(dynmat_eigen(h, A, n, P.masses, W, E, info, freqs_sq, svw, sv));
    // get_modes (dynmat.jl:123-131): the num_rigid_translation eigenvalues of smallest
    // magnitude (stable order, as sortperm) are set to zero; the vectors are untouched
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return std::fabs(freqs_sq[a]) < std::fabs(freqs_sq[b]); });
    for (int r = 0; r < num_rigid_translation; ++r) freqs_sq[idx[r]] = 0.0;
    if (phi64)
        FC3_HIP(hipMemcpyAsync(phi64, A, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice,
                               h->stream));
    if (phi32) FC3_TRY(to_f32(h->stream, A, (long)n * n, phi32));
    return FC3_OK;
}

int fc3_modes_sw(fc3_handle h, const double* positions, int n_atoms, const double box[3],
                 const double* sw_params, double r_cut, double tol, const double* masses,
                 int num_rigid_translation, double* freqs_sq, double* phi64, float* phi32)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    FC3_TRY(check_system(positions, n_atoms, box));
    if (!sw_params || !masses || !freqs_sq)
        return set_error(FC3_ERR_ARGUMENT, "sw_params, masses, freqs_sq must be non-null");
    if (!(r_cut <= sw_params[2] * sw_params[1]))
        return set_error(FC3_ERR_ARGUMENT,
                         "For SW silicon force constant cutoff must be less than potential cutoff");
    if (num_rigid_translation < 0 || num_rigid_translation > 3 * n_atoms)
        return set_error(FC3_ERR_ARGUMENT, "num_rigid_translation out of range");
    for (int i = 0; i < n_atoms; ++i)
        if (!(masses[i] > 0.0)) return set_error(FC3_ERR_ARGUMENT, "masses must be positive");
    const int N = n_atoms, n = 3 * N;
    size_t sv = 0;
    FC3_TRY(dsyevd_workspace_bytes(h, n, &sv));
    Prepared P;
    FC3_TRY(prepare(h, positions, N, box, r_cut, masses, [&](int cap) {
        return Carve::need(sw_workspace_bytes(N, cap, false)) +
               Carve::need(sizeof(double) * (size_t)n * n) + 2 * Carve::need(sizeof(double) * n) +
               Carve::need(sizeof(int)) + Carve::need(sv) + 4096;
    }, P));
    Carve cv(P.extra, P.extra_bytes);
    const size_t swb = sw_workspace_bytes(N, P.nl.cap, false);
    char* sw_ws = cv.take<char>(swb);
    double* A = cv.take<double>((size_t)n * n);
    double* W = cv.take<double>(n);
    double* E = cv.take<double>(n);
    int* info = cv.take<int>(1);
    char* svw = cv.take<char>(sv);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "modes workspace");
    FC3_TRY(sw_assemble(h->stream, P.pos, box, sw_params, sw_params[10], sw_params[5], P.nl, sw_ws,
                        swb, tol, nullptr, nullptr, nullptr, A));
    FC3_TRY(dynmat_eigen(h, A, n, P.masses, W, E, info, freqs_sq, svw, sv));
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return std::fabs(freqs_sq[a]) < std::fabs(freqs_sq[b]); });
    for (int r = 0; r < num_rigid_translation; ++r) freqs_sq[idx[r]] = 0.0;
    if (phi64)
        FC3_HIP(hipMemcpyAsync(phi64, A, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice,
                               h->stream));
    if (phi32) FC3_TRY(to_f32(h->stream, A, (long)n * n, phi32));
    return FC3_OK;
}

int fc3_third_order_sw(fc3_handle h, const double* positions, int n_atoms, const double box[3],
                       const double* sw_params, double r_cut, double tol, const double* masses,
                       double* psi64, float* psi32)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    FC3_TRY(check_system(positions, n_atoms, box));
    if (!sw_params) return set_error(FC3_ERR_ARGUMENT, "sw_params must be non-null");
    if (!psi64 && !psi32) return set_error(FC3_ERR_ARGUMENT, "psi64 and psi32 are both null");
    if (!(r_cut <= sw_params[2] * sw_params[1]))
        return set_error(FC3_ERR_ARGUMENT,
                         "For AutoDiff SW silicon force constant cutoff must be less than potential cutoff");
    if (masses)
        for (int i = 0; i < n_atoms; ++i)
            if (!(masses[i] > 0.0)) return set_error(FC3_ERR_ARGUMENT, "masses must be positive");
    const int N = n_atoms;
    Prepared P;
    FC3_TRY(prepare(h, positions, N, box, r_cut, masses,
                    [&](int cap) { return sw_workspace_bytes(N, cap, true); }, P));
    return sw_assemble(h->stream, P.pos, box, sw_params, sw_params[10], sw_params[5], P.nl,
                       P.extra, P.extra_bytes, tol, masses ? P.masses : nullptr, psi64, psi32,
                       nullptr);
}

int fc3_third_order_mcc3_sw(fc3_handle h, const double* positions, int n_atoms,
                            const double box[3], const double* sw_params, double r_cut,
                            double tol, const double* masses, const float* phi, float* k3)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    FC3_TRY(check_system(positions, n_atoms, box));
    if (!sw_params) return set_error(FC3_ERR_ARGUMENT, "sw_params must be non-null");
    if (!phi || !k3) return set_error(FC3_ERR_ARGUMENT, "phi, k3 must be non-null");
    if (!(r_cut <= sw_params[2] * sw_params[1]))
        return set_error(FC3_ERR_ARGUMENT,
                         "For AutoDiff SW silicon force constant cutoff must be less than potential cutoff");
    if (masses)
        for (int i = 0; i < n_atoms; ++i)
            if (!(masses[i] > 0.0)) return set_error(FC3_ERR_ARGUMENT, "masses must be positive");
    const int N = n_atoms;
    const long n = 3L * N;
    if (n % 4 != 0 || n < 16 || n * n >= (1L << 31) || (uintptr_t)k3 % 16 != 0)
        return set_error(FC3_ERR_UNSUPPORTED,
                         "fused third_order + mcc3: needs 3N % 4 == 0, 3N >= 16, aligned k3");
    // A block joins atoms of one triplet, so a slab atom's column atoms lie within two
    // neighbour hops: at most cap + cap (cap - 1) = cap^2 of them besides itself.
    auto cap2_bound = [&](int cap) { return (int)std::min<long>((long)N - 1 > 0 ? N - 1 : 1,
                                                                (long)cap * cap); };
    auto need = [&](int cap, bool direct) {
        const int c2 = cap2_bound(cap);
        return Carve::need(sw_workspace_bytes(N, cap, true)) +
               Carve::need(sizeof(float) * (size_t)n * n) +
               Carve::need(sizeof(uint16_t) * 2 * 3 * (size_t)n * mcc3_gather_k(c2) * N) +
               Carve::need(sw_fused_list_bytes(N, c2)) +
               mcc3_from_gathered_bytes((int)n, c2, direct) + 4096;
    };
    Prepared P;
    int rc = prepare(h, positions, N, box, r_cut, masses,
                     [&](int cap) { return need(cap, true); }, P);
    if (rc == FC3_ERR_OUT_OF_MEMORY)
        rc = prepare(h, positions, N, box, r_cut, masses,
                     [&](int cap) { return need(cap, false); }, P);
    FC3_TRY(rc);
    const int c2 = cap2_bound(P.nl.cap);
    Carve cv(P.extra, P.extra_bytes);
    char* sws = cv.take<char>(sw_workspace_bytes(N, P.nl.cap, true));
    float* phiT = cv.take<float>((size_t)n * n);
    __bf16* U = cv.take<__bf16>(2 * 3 * (size_t)n * mcc3_gather_k(c2) * N);
    char* lists = cv.take<char>(sw_fused_list_bytes(N, c2));
    char* rest = cv.take<char>(0);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "fused SW third_order + mcc3 workspace");
    SwBlockTable T;
    FC3_TRY(sw_assemble(h->stream, P.pos, box, sw_params, sw_params[10], sw_params[5], P.nl, sws,
                        sw_workspace_bytes(N, P.nl.cap, true), tol, masses ? P.masses : nullptr,
                        nullptr, nullptr, nullptr, &T));
    SwColumns cols;
    FC3_TRY(sw_fused_columns(h->stream, T, N, c2, lists, cols));
    const int Kc = mcc3_gather_k(cols.cap2);
    FC3_TRY(transpose_square(h->stream, phi, (int)n, phiT));
    // `rest` is free until mcc3_from_gathered: it holds mode 1's coefficients when a slab's
    // do not fit in LDS
    FC3_TRY(sw_mode1_gathered(h->stream, T, N, cols, tol, masses ? P.masses : nullptr, phiT, Kc,
                              U, rest, cv.cap - cv.off));
    return mcc3_from_gathered(h->stream, U, phi, phiT, (int)n, cols.colcnt, cols.collist,
                              cols.cap2, k3, rest, cv.cap - cv.off);
}

int fc3_mcc3(fc3_handle h, const float* psi, const float* phi, int n, int precision, float* k3)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    if (!psi || !phi || !k3) return set_error(FC3_ERR_ARGUMENT, "psi, phi, k3 must be non-null");
    if (n < 1) return set_error(FC3_ERR_ARGUMENT, "n must be >= 1");
    if (precision != FC3_PREC_FP32 && precision != FC3_PREC_BF16X3)
        return set_error(FC3_ERR_ARGUMENT, "unknown precision");
    if ((long long)n * n >= (1LL << 31))
        return set_error(FC3_ERR_UNSUPPORTED, "mcc3: n^2 must be < 2^31");
    if ((const void*)psi == (const void*)k3)
        return set_error(FC3_ERR_ARGUMENT, "k3 must not alias psi (use fc3_mcc3_inplace)");
    const size_t ws = mcc3_workspace_bytes(n, precision, false);
    FC3_TRY(ensure_workspace(h, ws));
    return mcc3_run(h->stream, psi, phi, n, precision, k3, h->ws, h->ws_bytes);
}

int fc3_mcc3_inplace(fc3_handle h, float* psi, const float* phi, int n, int precision)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    if (!psi || !phi) return set_error(FC3_ERR_ARGUMENT, "psi, phi must be non-null");
    if (n < 1) return set_error(FC3_ERR_ARGUMENT, "n must be >= 1");
    if (precision != FC3_PREC_FP32 && precision != FC3_PREC_BF16X3)
        return set_error(FC3_ERR_ARGUMENT, "unknown precision");
    if ((long long)n * n >= (1LL << 31))
        return set_error(FC3_ERR_UNSUPPORTED, "mcc3: n^2 must be < 2^31");
    const size_t ws = mcc3_workspace_bytes(n, precision, true);
    FC3_TRY(ensure_workspace(h, ws));
    return mcc3_run_inplace(h->stream, psi, phi, n, precision, h->ws, h->ws_bytes);
}

int fc3_mcc3_unique(fc3_handle h, const float* psi, const float* phi, int n, int precision,
                    float* k3u)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    if (!psi || !phi || !k3u) return set_error(FC3_ERR_ARGUMENT, "psi, phi, k3u must be non-null");
    if (n < 1) return set_error(FC3_ERR_ARGUMENT, "n must be >= 1");
    if (precision != FC3_PREC_FP32 && precision != FC3_PREC_BF16X3)
        return set_error(FC3_ERR_ARGUMENT, "unknown precision");
    if ((long long)n * n >= (1LL << 31))
        return set_error(FC3_ERR_UNSUPPORTED, "mcc3: n^2 must be < 2^31");
    // the full pipeline's U and slab outputs take n^3 floats of workspace ahead of the
    // mcc3 scratch
    const size_t sb = Carve::need(sizeof(float) * (size_t)n * n * n);
    FC3_TRY(ensure_workspace(h, sb + mcc3_unique_workspace_bytes(n, precision)));
    char* ws = static_cast<char*>(h->ws);
    return mcc3_unique_run(h->stream, psi, phi, n, precision, reinterpret_cast<float*>(ws), k3u,
                           ws + sb, h->ws_bytes - sb);
}

int fc3_mcc3_blocked(fc3_handle h, const float* psi, const float* phi, int n, int block_size,
                     int precision, float* k3_host)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    if (!psi || !phi || !k3_host)
        return set_error(FC3_ERR_ARGUMENT, "psi, phi, k3_host must be non-null");
    // @assert size(phi)[1] % block_size == 0 && block_size > 0  (mcc.jl:91-92)
    if (block_size <= 0 || n % block_size != 0)
        return set_error(FC3_ERR_BLOCK_SIZE, "size(phi)[1] % block_size == 0 && block_size > 0");
    if (precision != FC3_PREC_FP32 && precision != FC3_PREC_BF16X3)
        return set_error(FC3_ERR_ARGUMENT, "unknown precision");
    if ((long long)n * n >= (1LL << 31))
        return set_error(FC3_ERR_UNSUPPORTED, "mcc3: n^2 must be < 2^31");
    // The reference's block_size bounds device memory (mcc.jl:85).  Here the device holds
    // two blocks of bs output slabs (one being computed, one being copied to the host)
    // plus bounded scratch, never the whole K3.  bs is a multiple of block_size of about
    // 4 GiB per block (fewer passes over Psi), block_size itself when that is larger.
    const size_t slab = sizeof(float) * (size_t)n * n;
    const long bmem = std::max<long>(1, (long)((size_t(4) << 30) / slab));
    long bsl = block_size >= bmem ? block_size : (bmem / block_size) * block_size;
    const int bs = (int)std::min<long>(bsl, n);
    const bool v3 = mcc3_blocked_v3(n, precision, psi);
    const size_t blk = Carve::need(slab * bs);
    const size_t scratch = mcc3_blocked_bytes(n, v3, bs);
    FC3_TRY(ensure_workspace(h, 2 * blk + scratch + 1024));
    if (!h->aux) {
        FC3_HIP(hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking));
        for (auto& e : h->aux_ev) FC3_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    Carve cv(h->ws, h->ws_bytes);
    float* dbuf[2] = {cv.take<float>((size_t)n * n * bs), cv.take<float>((size_t)n * n * bs)};
    char* sw = cv.take<char>(scratch);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "mcc3 blocked workspace");
    hipEvent_t* computed = h->aux_ev;       // block in dbuf[b] computed
    hipEvent_t* copied = h->aux_ev + 2;     // dbuf[b] copied to the host
    const int nblk = (n + bs - 1) / bs;
    int st = FC3_OK;
    hipError_t e = hipSuccess;
    for (int t = 0; t <= nblk && st == FC3_OK && e == hipSuccess; ++t) {
        if (t < nblk) {   // compute block t into dbuf[t % 2] once its previous copy is done
            const int b = t % 2, l0 = t * bs, cnt = std::min(bs, n - l0);
            if (t >= 2) e = hipStreamWaitEvent(h->stream, copied[b], 0);
            if (e == hipSuccess)
                st = mcc3_blocked_slabs(h->stream, psi, phi, n, v3, bs, l0, cnt, dbuf[b], sw,
                                        scratch);
            if (st == FC3_OK && e == hipSuccess) e = hipEventRecord(computed[b], h->stream);
        }
        if (t >= 1 && st == FC3_OK && e == hipSuccess) {   // copy block t - 1 out
            const int b = (t - 1) % 2, l0 = (t - 1) * bs, cnt = std::min(bs, n - l0);
            e = hipStreamWaitEvent(h->aux, computed[b], 0);
            if (e == hipSuccess)
                e = hipMemcpyAsync(k3_host + (size_t)n * n * l0, dbuf[b], slab * cnt,
                                   hipMemcpyDeviceToHost, h->aux);
            if (e == hipSuccess) e = hipEventRecord(copied[b], h->aux);
        }
    }
    const hipError_t e1 = hipStreamSynchronize(h->stream);
    const hipError_t e2 = hipStreamSynchronize(h->aux);
    if (st != FC3_OK) return st;
    for (hipError_t x : {e, e1, e2})
        if (x != hipSuccess)
            return set_error(x == hipErrorOutOfMemory ? FC3_ERR_OUT_OF_MEMORY : FC3_ERR_HIP,
                             std::string("mcc3 blocked: ") + hipGetErrorString(x));
    return FC3_OK;
}

int fc3_mcc3_backward(fc3_handle h, const float* psi, const float* phi, const float* dk3, int n,
                      int precision, float* dpsi, float* dphi)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    if (!phi || !dk3) return set_error(FC3_ERR_ARGUMENT, "phi, dk3 must be non-null");
    if (!dpsi && !dphi) return set_error(FC3_ERR_ARGUMENT, "dpsi and dphi are both null");
    if (dphi && !psi) return set_error(FC3_ERR_ARGUMENT, "psi must be non-null for dphi");
    if (n < 1) return set_error(FC3_ERR_ARGUMENT, "n must be >= 1");
    if (precision != FC3_PREC_FP32 && precision != FC3_PREC_BF16X3)
        return set_error(FC3_ERR_ARGUMENT, "unknown precision");
    if ((long long)n * n >= (1LL << 31))
        return set_error(FC3_ERR_UNSUPPORTED, "mcc3: n^2 must be < 2^31");
    if (dpsi && ((const void*)dpsi == (const void*)dk3 || (const void*)dpsi == (const void*)psi))
        return set_error(FC3_ERR_ARGUMENT, "dpsi must not alias psi or dk3");
    FC3_TRY(ensure_workspace(h, mcc3_backward_workspace_bytes(n, precision, dphi != nullptr)));
    return mcc3_backward(h, psi, phi, dk3, n, precision, dpsi, dphi);
}

int fc3_weighted_square_loss(fc3_handle h, const float* k3, long long count,
                             unsigned long long seed, float* dk3, double* loss)
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    if (!k3 || count < 0) return set_error(FC3_ERR_ARGUMENT, "k3 must be non-null, count >= 0");
    if (!dk3 && !loss) return set_error(FC3_ERR_ARGUMENT, "dk3 and loss are both null");
    if (reinterpret_cast<uintptr_t>(k3) % 16 != 0 || reinterpret_cast<uintptr_t>(dk3) % 16 != 0)
        return set_error(FC3_ERR_ARGUMENT, "k3 and dk3 must be 16-byte aligned");
    if (dk3 && (const void*)dk3 != (const void*)k3) {
        const uintptr_t a = reinterpret_cast<uintptr_t>(k3), b = reinterpret_cast<uintptr_t>(dk3);
        const uintptr_t bytes = sizeof(float) * (uintptr_t)count;
        if (a < b + bytes && b < a + bytes)
            return set_error(FC3_ERR_ARGUMENT, "dk3 must be k3 itself or not overlap it");
    }
    FC3_TRY(ensure_workspace(h, weighted_square_loss_bytes() + Carve::need(sizeof(double))));
    Carve cv(h->ws, h->ws_bytes);
    double* d_loss = cv.take<double>(1);
    char* rest = cv.take<char>(weighted_square_loss_bytes());
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "loss workspace");
    FC3_TRY(weighted_square_loss(h->stream, k3, (long)count, seed, dk3, rest, d_loss));
    if (loss) {
        FC3_HIP(hipMemcpyAsync(loss, d_loss, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        FC3_HIP(hipStreamSynchronize(h->stream));
    }
    return FC3_OK;
}

int fc3_third_order_lj_param_backward(fc3_handle h, const double* positions, int n_atoms,
                                      const double box[3], double sigma, double epsilon,
                                      double pot_r_cut, double r_cut, double tol,
                                      const double* masses, const float* dpsi32,
                                      const double* dpsi64, double grad[2])
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    FC3_TRY(check_system(positions, n_atoms, box));
    for (int c = 0; c < 3; ++c)
        if (!(pot_r_cut < box[c])) return set_error(FC3_ERR_CUTOFF, "Cutoff larger than L/2");
    if (!grad) return set_error(FC3_ERR_ARGUMENT, "grad must be non-null");
    if ((dpsi32 == nullptr) == (dpsi64 == nullptr))
        return set_error(FC3_ERR_ARGUMENT, "exactly one of dpsi32, dpsi64 must be non-null");
    if (dpsi64 && masses)
        return set_error(FC3_ERR_ARGUMENT, "dpsi64 is the cotangent of the unweighted Psi: masses must be null");
    if (!(sigma > 0.0)) return set_error(FC3_ERR_ARGUMENT, "sigma must be positive");
    if (masses)
        for (int i = 0; i < n_atoms; ++i)
            if (!(masses[i] > 0.0)) return set_error(FC3_ERR_ARGUMENT, "masses must be positive");
    const int N = n_atoms;
    Prepared P;
    FC3_TRY(prepare(h, positions, N, box, r_cut, masses, [&](int cap) {
        const size_t blocks = ((size_t)N * cap + 255) / 256;
        return sizeof(double) * ((size_t)N * cap * (27 + 9) + (size_t)N * 27 + 2 * blocks + 2) +
               4096;
    }, P));
    const size_t slots = (size_t)N * P.nl.cap;
    double* E3 = reinterpret_cast<double*>(P.extra);
    double* E2 = E3 + slots * 27;
    double* S3 = E2 + slots * 9;
    double* out = S3 + (size_t)N * 27;
    double* partial = out + 2;
    FC3_TRY(lj_pair_blocks(h->stream, P.pos, box, sigma, epsilon, P.nl, E3, S3, E2, nullptr));
    FC3_TRY(lj_param_grad(h->stream, P.pos, box, sigma, epsilon, P.nl, E3, S3, tol,
                          masses ? P.masses : nullptr, dpsi32, dpsi64, partial, out));
    FC3_HIP(hipMemcpyAsync(grad, out, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    FC3_HIP(hipStreamSynchronize(h->stream));
    return FC3_OK;
}

int fc3_third_order_sw_param_backward(fc3_handle h, const double* positions, int n_atoms,
                                      const double box[3], const double* sw_params, double r_cut,
                                      double tol, const double* masses, const float* dpsi32,
                                      const double* dpsi64, double grad[4])
{
    FC3_TRACE_RANGE();
    FC3_TRY(check_handle(h));
    FC3_TRY(check_system(positions, n_atoms, box));
    if (!sw_params || !grad) return set_error(FC3_ERR_ARGUMENT, "sw_params, grad must be non-null");
    if (!(r_cut <= sw_params[2] * sw_params[1]))
        return set_error(FC3_ERR_ARGUMENT,
                         "For SW silicon force constant cutoff must be less than potential cutoff");
    if ((dpsi32 == nullptr) == (dpsi64 == nullptr))
        return set_error(FC3_ERR_ARGUMENT, "exactly one of dpsi32, dpsi64 must be non-null");
    if (dpsi64 && masses)
        return set_error(FC3_ERR_ARGUMENT, "dpsi64 is the cotangent of the unweighted Psi: masses must be null");
    if (!(sw_params[6] != 0.0) || !(sw_params[3] != 0.0) || !(sw_params[4] != 0.0))
        return set_error(FC3_ERR_ARGUMENT, "A, lambda and gamma must be nonzero");
    if (masses)
        for (int i = 0; i < n_atoms; ++i)
            if (!(masses[i] > 0.0)) return set_error(FC3_ERR_ARGUMENT, "masses must be positive");
    const int N = n_atoms;
    const long n = 3L * N;
    (void)n;
    // with tol > 0 the tolerance mask needs the forward values: the assembly's reduced
    // block table (MBs), recomputed here, not a dense 8 n^3-byte forward tensor
    Prepared P;
    FC3_TRY(prepare(h, positions, N, box, r_cut, masses, [&](int cap) {
        return sw_param_grad_bytes(N, cap) + (tol > 0.0 ? sw_workspace_bytes(N, cap, true) : 0) +
               Carve::need(4 * sizeof(double)) + 4096;
    }, P));
    Carve cv(P.extra, P.extra_bytes);
    double* out = cv.take<double>(4);
    char* gws = cv.take<char>(sw_param_grad_bytes(N, P.nl.cap));
    char* rest = cv.take<char>(0);
    if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "SW pullback workspace too small");
    const size_t rest_bytes = P.extra_bytes - (size_t)(rest - P.extra);
    SwBlockTable fwd;
    if (tol > 0.0)
        FC3_TRY(sw_assemble(h->stream, P.pos, box, sw_params, sw_params[10], sw_params[5], P.nl,
                            rest, rest_bytes, tol, nullptr, nullptr, nullptr, nullptr, &fwd));
    FC3_TRY(sw_param_grad(h->stream, P.pos, box, sw_params, sw_params[10], sw_params[5], P.nl, tol,
                          tol > 0.0 ? &fwd : nullptr, masses ? P.masses : nullptr, dpsi32, dpsi64,
                          gws, out));
    FC3_HIP(hipMemcpyAsync(grad, out, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    FC3_HIP(hipStreamSynchronize(h->stream));
    return FC3_OK;
}

}  // extern "C"
