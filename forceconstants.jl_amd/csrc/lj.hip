// 12-6 Lennard-Jones force constants (analytical calculator) on gfx950.
//
// Stage 1 (k_lj_pair_blocks): one thread per (atom a, neighbour slot s) computes the
//   3x3 and 3x3x3 pair blocks of that ordered pair; one thread per atom then forms
//   the self blocks that ASR! assigns (asr.jl:3-40) by summing its row of blocks in
//   ascending neighbour order, the order of the reference's strided sum.
// Stage 2 writes the dense tensors by output column (k_lj_write_ifc2 and
//   ifc3_writer.hip): every element is stored exactly once, zeros included, with
//   16-byte coalesced stores, so no memset pass and no scattered writes exist.
//
// Which value the reference leaves at each position.  third_order! for LJ
// (third_order_analytical.jl:14-62) writes, for every pair i<j and every (a,b,c),
// one value to six permuted positions, so the last iteration that reaches a
// position wins.  For a position holding two components (u, v) of atom A and one
// component w of atom B the winner is:
//     A < B :  -phi3(r, max(u,v), min(u,v), w)
//     A > B :  -phi3(r, w, max(u,v), min(u,v))     with r = mirror(r_A - r_B)
// independent of where the components sit among the three indices.  (The values
// differ only in the rounding of the triple product r_x r_y r_z.)  E3[a][s] holds
// this value for A = a, B = nbr[a][s] at [u + 3v + 9w].
#include "fc3_internal.hpp"
#include "geom.hpp"

namespace fc3 {
namespace {

__global__ void __launch_bounds__(256)
k_lj_pair_blocks(const double* __restrict__ pos, double Lx, double Ly, double Lz, double s12,
                 double s6, double eps, const int* __restrict__ count,
                 const int* __restrict__ nbr, int N, int cap, double* __restrict__ E3,
                 double* __restrict__ E2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * cap) return;
    const int a = t / cap, s = t - a * cap;
    if (s >= count[a]) return;
    const int b = nbr[t];
    const double L[3] = {Lx, Ly, Lz};
    double r[3] = {pos[3 * a] - pos[3 * b], pos[3 * a + 1] - pos[3 * b + 1],
                   pos[3 * a + 2] - pos[3 * b + 2]};
    geom::mirror(r, L);
    const double dist = sqrt(geom::sumsq(r));
    const geom::LJRadial f = geom::lj_radial(s12, s6, eps, dist, true);
    double* e3 = E3 + (size_t)t * 27;
    for (int w = 0; w < 3; ++w)
        for (int v = 0; v < 3; ++v)
            for (int u = 0; u < 3; ++u) {
                const int hi = u > v ? u : v, lo = u > v ? v : u;
                const double val = (a < b) ? -geom::phi3(f, dist, r, hi, lo, w)
                                           : -geom::phi3(f, dist, r, w, hi, lo);
                e3[u + 3 * v + 9 * w] = val;
            }
    double* e2 = E2 + (size_t)t * 9;
    for (int v = 0; v < 3; ++v)
        for (int u = 0; u < 3; ++u) e2[u + 3 * v] = -geom::phi2(f, dist, r, u, v);
}

// Self blocks: S[a] = -(sum over the row), ascending neighbour index (ASR!).
template <int B>
__global__ void k_self_blocks(const int* __restrict__ count, int N, int cap,
                              const double* __restrict__ E, double* __restrict__ S)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * B) return;
    const int a = t / B, e = t - a * B;
    double acc = 0.0;
    const int c = count[a];
    for (int s = 0; s < c; ++s) acc += E[((size_t)a * cap + s) * B + e];
    S[(size_t)a * B + e] = -acc;
}

__device__ __forceinline__ int find_slot(const int* __restrict__ list, int cnt, int key)
{
    int lo = 0, hi = cnt - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = list[mid];
        if (v == key) return mid;
        if (v < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

// IFC2 writer: one thread per element, column-major.
__global__ void __launch_bounds__(256)
k_lj_write_ifc2(const int* __restrict__ count, const int* __restrict__ nbr, int N, int cap,
                const double* __restrict__ E2, const double* __restrict__ S2, double tol,
                double* __restrict__ out)
{
    const long n = 3L * N;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const long j = t / n, i = t - j * n;
    const int I = (int)(i / 3), a = (int)(i - 3L * I);
    const int J = (int)(j / 3), b = (int)(j - 3L * J);
    double v = 0.0;
    if (I == J) {
        v = S2[(size_t)I * 9 + a + 3 * b];
    } else {
        const int s = find_slot(nbr + (size_t)I * cap, count[I], J);
        if (s >= 0) v = E2[((size_t)I * cap + s) * 9 + a + 3 * b];
    }
    out[t] = (fabs(v) < tol) ? 0.0 : v;
}

}  // namespace

void lj_sigma_powers(double sigma, double* s12, double* s6)
{
    // Julia's Float64^Int: compensated power by squaring (same algorithm as geom::ipow)
    auto hpow = [](double x, int n) {
        double yh = 1.0, yl = 0.0, xh = x, xl = 0.0;
        while (n > 1) {
            if (n & 1) {
                const double ph = yh * xh;
                double pl = __builtin_fma(yh, xh, -ph);
                pl = pl + (yh * xl + yl * xh);
                yh = ph; yl = pl;
            }
            const double sh = xh * xh;
            double sl = __builtin_fma(xh, xh, -sh);
            sl = sl + 2.0 * xh * xl;
            xh = sh; xl = sl;
            n >>= 1;
        }
        const double ph = yh * xh;
        double pl = __builtin_fma(yh, xh, -ph);
        pl = pl + (yh * xl + yl * xh);
        return ph + pl;
    };
    *s12 = hpow(sigma, 12);
    *s6 = hpow(sigma, 6);
}

int lj_pair_blocks(hipStream_t s, const double* d_pos, const double box[3], double sigma,
                   double epsilon, const NeighborList& nl, double* E3, double* S3, double* E2,
                   double* S2)
{
    // sigma^12 and sigma^6 as Julia evaluates them (compensated Float64^Int); the
    // host computes them with the same algorithm as the device for bit agreement.
    const int N = nl.n_atoms;
    const long tot = (long)N * nl.cap;
    double s12 = 0.0, s6 = 0.0;
    lj_sigma_powers(sigma, &s12, &s6);
    ProfScope prof(FC3_KERNEL_PAIR_BLOCKS, s);
    hipLaunchKernelGGL(k_lj_pair_blocks, dim3((tot + 255) / 256), dim3(256), 0, s, d_pos, box[0],
                       box[1], box[2], s12, s6, epsilon, nl.count, nl.nbr, N, nl.cap, E3, E2);
    FC3_HIP(hipGetLastError());
    if (S3) {
        hipLaunchKernelGGL(k_self_blocks<27>, dim3((N * 27 + 255) / 256), dim3(256), 0, s,
                           nl.count, N, nl.cap, E3, S3);
        FC3_HIP(hipGetLastError());
    }
    if (S2) {
        hipLaunchKernelGGL(k_self_blocks<9>, dim3((N * 9 + 255) / 256), dim3(256), 0, s, nl.count,
                           N, nl.cap, E2, S2);
        FC3_HIP(hipGetLastError());
    }
    return FC3_OK;
}

int lj_write_ifc2(hipStream_t s, const NeighborList& nl, const double* E2, const double* S2,
                  double tol, double* ifc2)
{
    const long n = 3L * nl.n_atoms;
    const long tot = n * n;
    hipLaunchKernelGGL(k_lj_write_ifc2, dim3((tot + 255) / 256), dim3(256), 0, s, nl.count,
                       nl.nbr, nl.n_atoms, nl.cap, E2, S2, tol, ifc2);
    FC3_HIP(hipGetLastError());
    return FC3_OK;
}

}  // namespace fc3
