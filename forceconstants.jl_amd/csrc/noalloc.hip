// Allocation-free variant of third_order for 12-6 LJ (fc3_third_order_lj_noalloc).
//
// fc3_third_order_lj keeps its scratch (uploaded positions and masses, neighbour counts,
// sorted neighbour lists with reverse slots, the 27- and 9-value pair blocks and the self
// blocks) in the handle's workspace, which it grows with hipMalloc.  Callers that cannot
// allocate (serving loops, harnesses that forbid it) use this variant instead: the same
// kernels, with that scratch carved from `scratch`, a static device array of
// kNoallocScratchBytes (64 MiB) that every GPU the library runs on holds until the
// process ends.  The caller passes the outputs, as for fc3_third_order_lj, and the
// results are identical.
//
// The array's name is a host-side stand-in, not a device address: its address on each
// device is taken once, with that device current, by fc3_create (noalloc_init_device)
// and kept per device; kernels receive that address.  Every call on one GPU shares that
// GPU's array and leaves its contents behind: everything read is written first in the
// same call (the maximum-count cell is cleared by neighbor_count), and calls that
// overlap on one GPU race.
#include <cmath>
#include <mutex>
#include <string>

#include "fc3_internal.hpp"

static __device__ __align__(256) unsigned char scratch[fc3::kNoallocScratchBytes];

namespace fc3 {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};   // address of `scratch` on each device

}  // namespace

int noalloc_init_device(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return set_error(FC3_ERR_UNSUPPORTED, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device]) return FC3_OK;
    void* p = nullptr;
    FC3_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = p;
    return FC3_OK;
}

namespace {

void* scratch_of(int device)
{
    std::lock_guard<std::mutex> lock(g_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}

int too_big(size_t need, int n_atoms, int cap)
{
    return set_error(FC3_ERR_CAPACITY,
                     "fc3_third_order_lj_noalloc: " + std::to_string(n_atoms) + " atoms with " +
                         std::to_string(cap) + " neighbours need " + std::to_string(need) +
                         " bytes of scratch; the static scratch holds " +
                         std::to_string(kNoallocScratchBytes));
}

}  // namespace
}  // namespace fc3

using namespace fc3;

extern "C" int fc3_third_order_lj_noalloc(fc3_handle h, const double* positions, int n_atoms,
                                          const double box[3], double sigma, double epsilon,
                                          double pot_r_cut, double r_cut, double tol,
                                          const double* masses, double* psi64, float* psi32)
{
    if (!h) return set_error(FC3_ERR_ARGUMENT, "null fc3 handle");
    FC3_HIP(hipSetDevice(h->device));
    set_current(nullptr);   // no per-kernel timing events: they would be created here
    if (!positions || !box) return set_error(FC3_ERR_ARGUMENT, "positions and box must be non-null");
    if (n_atoms < 1) return set_error(FC3_ERR_ARGUMENT, "n_atoms must be >= 1");
    if ((long long)n_atoms * 3 > 65535)
        return set_error(FC3_ERR_UNSUPPORTED, "3*n_atoms must be <= 65535");
    for (int c = 0; c < 3; ++c) {
        if (!(box[c] > 0.0) || !std::isfinite(box[c]))
            return set_error(FC3_ERR_ARGUMENT, "box sizes must be positive and finite");
        if (!(pot_r_cut < box[c])) return set_error(FC3_ERR_CUTOFF, "Cutoff larger than L/2");
    }
    if (!psi64 && !psi32) return set_error(FC3_ERR_ARGUMENT, "psi64 and psi32 are both null");
    if (masses)
        for (int i = 0; i < n_atoms; ++i)
            if (!(masses[i] > 0.0)) return set_error(FC3_ERR_ARGUMENT, "masses must be positive");
    void* base = scratch_of(h->device);
    if (!base) return set_error(FC3_ERR_HIP, "static scratch not initialised on this device");
    const int N = n_atoms;
    hipStream_t s = h->stream;

    // phase 1: positions, masses, counts and the maximum-count cell
    Carve cv(base, kNoallocScratchBytes);
    double* pos = cv.take<double>(3 * (size_t)N);
    double* dm = cv.take<double>(N);
    int* count = cv.take<int>(N);
    int* maxc = cv.take<int>(1);
    if (!cv.ok()) return too_big(cv.off, N, 0);
    FC3_HIP(hipMemcpyAsync(pos, positions, sizeof(double) * 3 * N, hipMemcpyHostToDevice, s));
    if (masses)
        FC3_HIP(hipMemcpyAsync(dm, masses, sizeof(double) * N, hipMemcpyHostToDevice, s));
    int hmax = 0;
    FC3_TRY(neighbor_count(s, pos, N, box, r_cut, count, maxc, &hmax));
    const int cap = hmax > 0 ? hmax : 1;

    // phase 2: lists and blocks, sized by the counted capacity, checked before any launch
    const size_t slots = (size_t)N * cap;
    NeighborList nl;
    nl.n_atoms = N;
    nl.cap = cap;
    nl.count = count;
    nl.nbr = cv.take<int>(slots);
    nl.rev = cv.take<int>(slots);
    double* E3 = cv.take<double>(slots * 27);
    double* E2 = cv.take<double>(slots * 9);
    double* S3 = cv.take<double>((size_t)N * 27);
    if (!cv.ok()) return too_big(cv.off, N, cap);
    FC3_TRY(neighbor_fill(s, pos, N, box, r_cut, nl));
    FC3_TRY(lj_pair_blocks(s, pos, box, sigma, epsilon, nl, E3, S3, E2, nullptr));
    Ifc3Blocks b;
    b.E = E3;
    b.S = S3;
    return write_ifc3_pairwise(s, nl, b, tol, masses ? dm : nullptr, psi64, psi32);
}
