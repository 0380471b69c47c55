/*
 * fc3_oracle.c -- CPU restatement of ForceConstants.jl's third-order pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg,
 * never by the product path.  It follows the reference's loop structure literally
 * (dense column-major arrays, write order, strided ASR sums) so that it reproduces
 * the reference's values, not just the mathematics.
 *
 * Layout: column-major (Julia), 0-based: Psi[i + n*j + n*n*k], n = 3N.
 *
 * Reference files followed (paths relative to the ForceConstants.jl root):
 *   src/helper_funcs.jl:7-28             nearest_mirror!, apply_tols!
 *   src/interactions/LennardJones.jl:28-52  LJ radial derivatives
 *   src/analytical/second_order_analytical.jl:3-51  second_order!, phi2
 *   src/analytical/third_order_analytical.jl:4-69,185-224  third_order!, mass weighting, phi3
 *   src/asr.jl:3-40                      ASR! (2nd and 3rd order)
 *   src/interactions/SW.jl:60-83         SW pair and three-body energy terms
 *   src/autodiff/second_order_AD.jl:39-111  SW second_order! loop + scatter
 *   src/autodiff/third_order_AD.jl:71-224   SW third_order! loop + permuted scatter
 *   src/autodiff/autodiff_helper.jl:85-108  SW block definitions (d3h / dr_x dr_y dr_z)
 *
 * Parity status of each function is stated in its comment; every function here is
 * pinned by a test in tests/test_oracle_*.py (golden values of the reference test
 * suite, sympy/mpmath derivatives, finite differences or invariants).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <omp.h>

#define D3 3

/* ------------------------------------------------------------------------- */
/* Integer power, x^n for n >= 1, evaluated by binary exponentiation carried in
 * double-double (Dekker products through fma) and rounded once at the end.  The
 * reference evaluates sigma^12, r^13 ... with Julia's Float64^Int, which is a
 * compensated power-by-squaring that is correctly rounded in all but rare cases;
 * n == 2 and n == 3 are plain products (Julia's literal_pow).                    */
static double ipow(double x, int n)
{
    if (n == 2) return x * x;
    if (n == 3) return x * x * x;
    double yh = 1.0, yl = 0.0, xh = x, xl = 0.0;
    while (n > 1) {
        if (n & 1) {
            double ph = yh * xh;
            double pl = fma(yh, xh, -ph);
            pl = pl + (yh * xl + yl * xh);
            yh = ph; yl = pl;
        }
        double sh = xh * xh;
        double sl = fma(xh, xh, -sh);
        sl = sl + 2.0 * xh * xl;
        xh = sh; xl = sl;
        n >>= 1;
    }
    double ph = yh * xh;
    double pl = fma(yh, xh, -ph);
    pl = pl + (yh * xl + yl * xh);
    return ph + pl;
}

double fco_ipow(double x, int n) { return ipow(x, n); }

/* threads the threaded loops use (OpenMP's default team size) */
int fco_num_threads(void) { return omp_get_max_threads(); }

/* nearest_mirror! (helper_funcs.jl:7-18): one shift per component, strict bounds. */
static void mirror(double r[3], const double L[3])
{
    for (int c = 0; c < 3; ++c) {
        if (r[c] > L[c] / 2) r[c] -= L[c];
        else if (r[c] < -L[c] / 2) r[c] += L[c];
    }
}

/* sum(x -> x^2, r) : ((x*x + y*y) + z*z) */
static double sumsq(const double r[3]) { return (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]; }

/* ------------------------------------------------------------------------- */
/* 12-6 Lennard-Jones radial derivatives, LennardJones.jl:38-52, same operation
 * order (Julia n-ary * folds left; 14(x) is 14*x).                             */
static double lj_d1(double s, double e, double r)
{
    double s12 = ipow(s, 12), s6 = ipow(s, 6);
    return (-1.0 * e) * (48.0 * (s12 / ipow(r, 13)) - 24.0 * (s6 / ipow(r, 7)));
}
static double lj_d2(double s, double e, double r)
{
    double s12 = ipow(s, 12), s6 = ipow(s, 6);
    return (4.0 * e) * (156.0 * (s12 / ipow(r, 14)) - 42.0 * (s6 / ipow(r, 8)));
}
static double lj_d3(double s, double e, double r)
{
    double s12 = ipow(s, 12), s6 = ipow(s, 6);
    return (-4.0 * e) * (156.0 * (14.0 * (s12 / ipow(r, 15))) - 42.0 * (8.0 * (s6 / ipow(r, 9))));
}

/* phi2, second_order_analytical.jl:44-51 */
static double lj_phi2(double d1, double d2, double dist, const double r[3], int a, int b)
{
    double t = ((r[a] * r[b]) / (dist * dist)) * (d2 - (d1 / dist));
    if (a == b) return t + (d1 / dist);
    return t;
}

/* phi3, third_order_analytical.jl:216-224; the product r[a]*r[b]*r[c] is taken in
 * argument order, which is what makes the reference's last-write-wins visible.  */
static double lj_phi3(double d1, double d2, double d3, double dist, const double r[3],
                      int a, int b, int c)
{
    double t1 = (((r[a] * r[b]) * r[c]) / ((dist * dist) * dist)) *
                ((d3 - ((3.0 * d2) / dist)) + ((3.0 * d1) / (dist * dist)));
    double dl = ((b == c ? r[a] : 0.0) + (c == a ? r[b] : 0.0)) + (a == b ? r[c] : 0.0);
    double t2 = (dl / (dist * dist)) * (d2 - (d1 / dist));
    return t1 + t2;
}

/* apply_tols!, helper_funcs.jl:20-28 (elementwise, so threads change nothing) */
static void apply_tols(double* a, size_t len, double tol)
{
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < len; ++i)
        if (fabs(a[i]) < tol) a[i] = 0.0;
}

/* ASR!(::Matrix), asr.jl:3-17: diagonal block entry = -(sum of its row at stride 3),
 * summed in ascending atom order (the entry itself is still zero when summed).   */
static void asr2(double* f, int N)
{
    const long n = 3L * N;
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                long ii = 3L * i + a, jj = 3L * i + b;
                double s = 0.0;
                for (int j = 0; j < N; ++j) s += f[ii + n * (3L * j + b)];
                f[ii + n * jj] = -s;
            }
}

/* ASR!(::Array{T,3}), asr.jl:19-40: every (i,j,i) entry is zeroed and replaced by
 * -(sum over k of Psi[i a, j b, k c]) in ascending k.  Threaded over i like the
 * reference's @tasks loop: atom i's iteration reads and writes only rows 3i..3i+2,
 * and each sum keeps its ascending order, so the result does not depend on threads. */
static void asr3(double* p, int N)
{
    const long n = 3L * N, n2 = n * n;
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int a = 0; a < 3; ++a) {
                long ii = 3L * i + a;
                for (int b = 0; b < 3; ++b) {
                    long jj = 3L * j + b;
                    for (int c = 0; c < 3; ++c) {
                        long kk = 3L * i + c;
                        p[ii + n * jj + n2 * kk] = 0.0;
                        double s = 0.0;
                        for (int k = 0; k < N; ++k) s += p[ii + n * jj + n2 * (3L * k + c)];
                        p[ii + n * jj + n2 * kk] = -s;
                    }
                }
            }
}

/* ------------------------------------------------------------------------- */
/* second_order (ifc.jl:3-18) with the analytical LJ calculator.
 * Returns 0, or 2 when the cutoff check fails.                                  */
int fco_lj_ifc2(const double* pos, int N, const double* L, double sigma, double eps,
                double pot_rcut, double rcut, double tol, double* out)
{
    (void)pot_rcut;
    const long n = 3L * N;
    memset(out, 0, sizeof(double) * n * n);
    const double rc2 = rcut * rcut;
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) {
            double r[3] = {pos[3 * i] - pos[3 * j], pos[3 * i + 1] - pos[3 * j + 1],
                           pos[3 * i + 2] - pos[3 * j + 2]};
            mirror(r, L);
            double d2s = sumsq(r);
            if (d2s < rc2) {
                double dist = sqrt(d2s);
                double d1 = lj_d1(sigma, eps, dist), d2 = lj_d2(sigma, eps, dist);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        long ii = 3L * i + a, jj = 3L * j + b;
                        double v = -lj_phi2(d1, d2, dist, r, a, b);
                        out[ii + n * jj] = v;
                        out[jj + n * ii] = v;
                    }
            }
        }
    asr2(out, N);
    apply_tols(out, (size_t)(n * n), tol);
    return 0;
}

/* third_order! with the analytical LJ calculator (third_order_analytical.jl:4-69)
 * followed by ASR! and apply_tols! (ifc.jl:20-35).  The six permuted stores of
 * each (a,b,c) iteration are replayed in the reference's order.  The pair loop is
 * threaded over i with dynamic scheduling, as the reference's @tasks loop
 * (:14-15): pair (i<j) stores only into the permutations of the atom multisets
 * {i,i,j} and {i,j,j}, which no other pair shares, so no two iterations store to
 * one entry and the result does not depend on the thread count.                 */
int fco_lj_ifc3(const double* pos, int N, const double* L, double sigma, double eps,
                double pot_rcut, double rcut, double tol, double* out)
{
    for (int c = 0; c < 3; ++c)
        if (!(pot_rcut < L[c])) return 2;
    const long n = 3L * N, n2 = n * n;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) memset(out + n2 * k, 0, sizeof(double) * n2);
    const double rc2 = rcut * rcut;
#define P3(x, y, z) out[(x) + n * (y) + n2 * (z)]
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) {
            for (int fam = 0; fam < 2; ++fam) {
                double r[3];
                if (fam == 0)
                    for (int c = 0; c < 3; ++c) r[c] = pos[3 * i + c] - pos[3 * j + c];
                else
                    for (int c = 0; c < 3; ++c) r[c] = pos[3 * j + c] - pos[3 * i + c];
                mirror(r, L);
                double d2s = sumsq(r);
                if (!(d2s < rc2)) continue;
                double dist = sqrt(d2s);
                double d1 = lj_d1(sigma, eps, dist), d2 = lj_d2(sigma, eps, dist),
                       d3 = lj_d3(sigma, eps, dist);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        for (int c = 0; c < 3; ++c) {
                            double v = -lj_phi3(d1, d2, d3, dist, r, a, b, c);
                            long ii, jj, kk;
                            if (fam == 0) { ii = 3L * i + a; jj = 3L * i + b; kk = 3L * j + c; }
                            else          { ii = 3L * i + a; jj = 3L * j + b; kk = 3L * j + c; }
                            P3(ii, jj, kk) = v; P3(ii, kk, jj) = v;
                            P3(jj, kk, ii) = v; P3(jj, ii, kk) = v;
                            P3(kk, ii, jj) = v; P3(kk, jj, ii) = v;
                        }
            }
        }
#undef P3
    asr3(out, N);
    apply_tols(out, (size_t)(n2 * n), tol);
    return 0;
}

/* The value the reference leaves at a position that only pair (p<q) writes
 * directly, obtained by replaying that pair's stores and keeping the last one
 * that lands on (x,y,z).  Returns 0.0 when the pair is out of range or does not
 * touch the position.                                                          */
static double lj_direct(const double* pos, const double* L, double sigma, double eps,
                        double rc2, int p, int q, long x, long y, long z)
{
    double val = 0.0;
    for (int fam = 0; fam < 2; ++fam) {
        double r[3];
        if (fam == 0)
            for (int c = 0; c < 3; ++c) r[c] = pos[3 * p + c] - pos[3 * q + c];
        else
            for (int c = 0; c < 3; ++c) r[c] = pos[3 * q + c] - pos[3 * p + c];
        mirror(r, L);
        double d2s = sumsq(r);
        if (!(d2s < rc2)) continue;
        double dist = sqrt(d2s);
        double d1 = lj_d1(sigma, eps, dist), d2 = lj_d2(sigma, eps, dist),
               d3 = lj_d3(sigma, eps, dist);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    long ii, jj, kk;
                    if (fam == 0) { ii = 3L * p + a; jj = 3L * p + b; kk = 3L * q + c; }
                    else          { ii = 3L * p + a; jj = 3L * q + b; kk = 3L * q + c; }
                    long w[6][3] = {{ii, jj, kk}, {ii, kk, jj}, {jj, kk, ii},
                                    {jj, ii, kk}, {kk, ii, jj}, {kk, jj, ii}};
                    for (int s = 0; s < 6; ++s)
                        if (w[s][0] == x && w[s][1] == y && w[s][2] == z)
                            val = -lj_phi3(d1, d2, d3, dist, r, a, b, c);
                }
    }
    return val;
}

/* Value of a single entry of the LJ third_order result before apply_tols!, for
 * systems too large for a dense CPU tensor.  Same semantics as fco_lj_ifc3.      */
static double lj_entry(const double* pos, int N, const double* L, double sigma, double eps,
                       double rc2, long x, long y, long z)
{
    int I = (int)(x / 3), J = (int)(y / 3), K = (int)(z / 3);
    if (I == K) {
        /* ASR entry: -(sum_k Psi[x, y, 3k + c]) with the self term excluded */
        int c = (int)(z % 3);
        double s = 0.0;
        for (int k = 0; k < N; ++k) {
            if (k == I) continue;
            s += lj_entry(pos, N, L, sigma, eps, rc2, x, y, 3L * k + c);
        }
        return -s;
    }
    /* first != third: only a pair of two distinct atoms can write it */
    int A = I, B = K;            /* the two atoms; J equals one of them */
    if (!(J == A || J == B)) return 0.0;
    int p = A < B ? A : B, q = A < B ? B : A;
    return lj_direct(pos, L, sigma, eps, rc2, p, q, x, y, z);
}

/* Columns Psi[:, jj[c], kk[c]] of the LJ third_order result (ASR + tols). */
int fco_lj_ifc3_columns(const double* pos, int N, const double* L, double sigma, double eps,
                        double pot_rcut, double rcut, double tol, int ncols,
                        const int64_t* jj, const int64_t* kk, double* out)
{
    for (int c = 0; c < 3; ++c)
        if (!(pot_rcut < L[c])) return 2;
    const long n = 3L * N;
    const double rc2 = rcut * rcut;
    for (int c = 0; c < ncols; ++c)
        for (long x = 0; x < n; ++x) {
            double v = lj_entry(pos, N, L, sigma, eps, rc2, x, jj[c], kk[c]);
            out[x + n * c] = (fabs(v) < tol) ? 0.0 : v;
        }
    return 0;
}

/* mass_weight_third_order!, third_order_analytical.jl:185-214 (D = 3). */
/* mass_weight_third_order! (third_order_analytical.jl:185-214) for D = n / N in 1:3:
 * ii = D*i + a, jj = D*j + b, kk = D*k + c; nonzero entries divided by
 * sqrt(m_i m_j m_k), the product taken left to right as the reference writes it. */
void fco_mass_weight3_d(double* psi, int N, int D, const double* m)
{
    const long n = (long)D * N, n2 = n * n;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k)
                for (int a = 0; a < D; ++a)
                    for (int b = 0; b < D; ++b)
                        for (int c = 0; c < D; ++c) {
                            long idx = ((long)D * i + a) + n * ((long)D * j + b) +
                                       n2 * ((long)D * k + c);
                            if (psi[idx] != 0) psi[idx] /= sqrt((m[i] * m[j]) * m[k]);
                        }
}

void fco_mass_weight3(double* psi, int N, const double* m) { fco_mass_weight3_d(psi, N, 3, m); }

/* ========================================================================= */
/* Stillinger-Weber.  Derivative blocks are computed by order-3 hyper-dual
 * arithmetic (Fike & Alonso's hyper-dual numbers extended to three nilpotent
 * directions): evaluating h at x + e1*u + e2*v + e3*w gives d3h(u,v,w) in the
 * e1e2e3 coefficient, exactly up to rounding.  This is an automatic derivative,
 * as the reference's AutoDiffCalculator is, and shares no code with the HIP
 * closed forms.                                                              */
typedef struct { double v[8]; } hd3;   /* 1, e1, e2, e12, e3, e13, e23, e123 */

static hd3 hd_const(double x) { hd3 r; memset(&r, 0, sizeof r); r.v[0] = x; return r; }
static hd3 hd_add(hd3 a, hd3 b) { for (int i = 0; i < 8; ++i) a.v[i] += b.v[i]; return a; }
static hd3 hd_sub(hd3 a, hd3 b) { for (int i = 0; i < 8; ++i) a.v[i] -= b.v[i]; return a; }
static hd3 hd_scale(hd3 a, double s) { for (int i = 0; i < 8; ++i) a.v[i] *= s; return a; }
static hd3 hd_mul(hd3 a, hd3 b)
{
    /* basis index bitmask: bit0 = e1, bit1 = e2, bit2 = e3; e_i^2 = 0 */
    hd3 r; memset(&r, 0, sizeof r);
    for (int i = 0; i < 8; ++i) {
        if (a.v[i] == 0.0) continue;
        for (int j = 0; j < 8; ++j) {
            if (i & j) continue;
            r.v[i | j] += a.v[i] * b.v[j];
        }
    }
    return r;
}
/* f(a) for a scalar function with derivatives f0..f3 at a.v[0] */
static hd3 hd_apply(hd3 a, double f0, double f1, double f2, double f3)
{
    hd3 d = a; d.v[0] = 0.0;
    hd3 d2 = hd_mul(d, d), d3 = hd_mul(d2, d);
    hd3 r = hd_const(f0);
    r = hd_add(r, hd_scale(d, f1));
    r = hd_add(r, hd_scale(d2, f2 / 2.0));
    r = hd_add(r, hd_scale(d3, f3 / 6.0));
    return r;
}
static hd3 hd_sqrt(hd3 a)
{
    double x = a.v[0], s = sqrt(x);
    return hd_apply(a, s, 0.5 / s, -0.25 / (s * x), 0.375 / (s * x * x));
}
static hd3 hd_inv(hd3 a)
{
    double x = a.v[0], i1 = 1.0 / x;
    return hd_apply(a, i1, -i1 * i1, 2.0 * i1 * i1 * i1, -6.0 * i1 * i1 * i1 * i1);
}
static hd3 hd_exp(hd3 a)
{
    double e = exp(a.v[0]);
    return hd_apply(a, e, e, e, e);
}
static hd3 hd_div(hd3 a, hd3 b) { return hd_mul(a, hd_inv(b)); }

/* SW parameter vector layout shared with the C-ABI:
 *   0 epsilon, 1 sigma, 2 a, 3 lambda, 4 gamma, 5 cos_theta0, 6 A, 7 B, 8 p, 9 q */
typedef struct {
    double eps, sigma, a, lambda, gamma, c0, A, B, p, q;
} swp_t;

static swp_t swp_load(const double* v)
{
    swp_t s = {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9]};
    return s;
}

/* h(ri, rj, rk), SW.jl:68-83 with r_ij = ri - rj, r_ik = ri - rk.  c0 is passed
 * separately because the reference's derivative blocks bake two different values. */
static hd3 sw_h(const hd3 x[9], const swp_t* P, double c0)
{
    hd3 rij[3], rik[3];
    for (int c = 0; c < 3; ++c) { rij[c] = hd_sub(x[c], x[3 + c]); rik[c] = hd_sub(x[c], x[6 + c]); }
    hd3 sij = hd_add(hd_add(hd_mul(rij[0], rij[0]), hd_mul(rij[1], rij[1])), hd_mul(rij[2], rij[2]));
    hd3 sik = hd_add(hd_add(hd_mul(rik[0], rik[0]), hd_mul(rik[1], rik[1])), hd_mul(rik[2], rik[2]));
    hd3 dij = hd_sqrt(sij), dik = hd_sqrt(sik);
    hd3 dot = hd_add(hd_add(hd_mul(rij[0], rik[0]), hd_mul(rij[1], rik[1])), hd_mul(rij[2], rik[2]));
    hd3 cs = hd_div(dot, hd_mul(dij, dik));
    hd3 dc = hd_sub(cs, hd_const(c0));
    double le = P->eps * P->lambda, gs = P->sigma * P->gamma, rc = P->a * P->sigma;
    hd3 e1 = hd_exp(hd_scale(hd_inv(hd_sub(dij, hd_const(rc))), gs));
    hd3 e2 = hd_exp(hd_scale(hd_inv(hd_sub(dik, hd_const(rc))), gs));
    return hd_mul(hd_mul(hd_scale(hd_mul(dc, dc), le), e1), e2);
}

/* phi2(r) of SW.jl:76-78 as a function of the 3-vector r (pair term) */
static hd3 sw_pair(const hd3 r[3], const swp_t* P)
{
    hd3 s = hd_add(hd_add(hd_mul(r[0], r[0]), hd_mul(r[1], r[1])), hd_mul(r[2], r[2]));
    hd3 d = hd_sqrt(s);
    hd3 sr = hd_scale(hd_inv(d), P->sigma);             /* sigma / r */
    /* (sigma/r)^p for p = 4, and (sigma/r)^q for q = 0 */
    hd3 srp = hd_const(1.0);
    for (int t = 0; t < (int)P->p; ++t) srp = hd_mul(srp, sr);
    hd3 srq = hd_const(1.0);
    for (int t = 0; t < (int)P->q; ++t) srq = hd_mul(srq, sr);
    hd3 poly = hd_sub(hd_scale(srp, P->B), srq);
    hd3 ex = hd_exp(hd_scale(hd_inv(hd_sub(d, hd_const(P->a * P->sigma))), P->sigma));
    return hd_scale(hd_mul(poly, ex), P->A * P->eps);
}

static void seed9(hd3 x[9], const double* r9, int u, int v, int w)
{
    for (int t = 0; t < 9; ++t) x[t] = hd_const(r9[t]);
    x[u].v[1] += 1.0;
    x[v].v[2] += 1.0;
    x[w].v[4] += 1.0;
}

/* d3h / (d x_{g0*3+a} d x_{g1*3+b} d x_{g2*3+c}), groups 0 = i, 1 = j, 2 = k */
static void sw_block3(const double* r9, const swp_t* P, double c0, int g0, int g1, int g2,
                      double* out27)
{
    hd3 x[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) {
                seed9(x, r9, 3 * g0 + a, 3 * g1 + b, 3 * g2 + c);
                out27[a + 3 * b + 9 * c] = sw_h(x, P, c0).v[7];
            }
}

/* d2h / (d x_{g0*3+a} d x_{g1*3+b}) */
static void sw_block2(const double* r9, const swp_t* P, double c0, int g0, int g1, double* out9)
{
    hd3 x[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            for (int t = 0; t < 9; ++t) x[t] = hd_const(r9[t]);
            x[3 * g0 + a].v[1] += 1.0;
            x[3 * g1 + b].v[2] += 1.0;
            out9[a + 3 * b] = sw_h(x, P, c0).v[3];
        }
}

/* Pair blocks: d2 phi2 / dr dr and d3 phi2 / dr dr dr w.r.t. r = r_ij. */
static void sw_pair_blocks(const double r[3], const swp_t* P, double* h2, double* h3)
{
    hd3 x[3];
    if (h2)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                for (int t = 0; t < 3; ++t) x[t] = hd_const(r[t]);
                x[a].v[1] += 1.0; x[b].v[2] += 1.0;
                h2[a + 3 * b] = sw_pair(x, P).v[3];
            }
    if (h3)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    for (int t = 0; t < 3; ++t) x[t] = hd_const(r[t]);
                    x[a].v[1] += 1.0; x[b].v[2] += 1.0; x[c].v[4] += 1.0;
                    h3[a + 3 * b + 9 * c] = sw_pair(x, P).v[7];
                }
}

/* Exported for fixture checks.  out = 7 blocks x 27 in the order
 * iij, ijj, iik, ikk, jjk, jkk, ijk; c0_fd is the constant of the FastDifferentiation
 * blocks, c0_iik the one of the runtime-built iik block.                         */
void fco_sw_blocks3(const double* r9, const double* swv, double c0_fd, double c0_iik,
                    double* out)
{
    swp_t P = swp_load(swv);
    sw_block3(r9, &P, c0_fd, 0, 0, 1, out + 0 * 27);    /* iij */
    sw_block3(r9, &P, c0_fd, 0, 1, 1, out + 1 * 27);    /* ijj */
    sw_block3(r9, &P, c0_iik, 0, 0, 2, out + 2 * 27);   /* iik */
    sw_block3(r9, &P, c0_fd, 0, 2, 2, out + 3 * 27);    /* ikk */
    sw_block3(r9, &P, c0_fd, 1, 1, 2, out + 4 * 27);    /* jjk */
    sw_block3(r9, &P, c0_fd, 1, 2, 2, out + 5 * 27);    /* jkk */
    sw_block3(r9, &P, c0_fd, 0, 1, 2, out + 6 * 27);    /* ijk */
}

/* out = 3 blocks x 9 in the order ij, ik, jk */
void fco_sw_blocks2(const double* r9, const double* swv, double c0_fd, double* out)
{
    swp_t P = swp_load(swv);
    sw_block2(r9, &P, c0_fd, 0, 1, out + 0);
    sw_block2(r9, &P, c0_fd, 0, 2, out + 9);
    sw_block2(r9, &P, c0_fd, 1, 2, out + 18);
}

void fco_sw_pair_blocks(const double* r3, const double* swv, double* h2, double* h3)
{
    swp_t P = swp_load(swv);
    sw_pair_blocks(r3, &P, h2, h3);
}

/* SW energy (finitediff_helper.jl:24-64 with SW.jl:60-83): used by finite-
 * difference checks of the derivative blocks.                                   */
double fco_sw_energy(const double* pos, int N, const double* L, const double* swv, double rcut)
{
    swp_t P = swp_load(swv);
    const double rc2 = rcut * rcut;
    double E = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            double rij[3];
            for (int c = 0; c < 3; ++c) rij[c] = pos[3 * i + c] - pos[3 * j + c];
            mirror(rij, L);
            double dij2 = sumsq(rij);
            if (!(dij2 < rc2)) continue;
            double dij = sqrt(dij2);
            if (j > i) {
                double sr = P.sigma / dij;
                E += P.A * P.eps * (P.B * pow(sr, P.p) - pow(sr, P.q)) *
                     exp(P.sigma / (dij - P.a * P.sigma));
            }
            for (int k = j + 1; k < N; ++k) {
                if (k == i) continue;
                double rik[3];
                for (int c = 0; c < 3; ++c) rik[c] = pos[3 * i + c] - pos[3 * k + c];
                mirror(rik, L);
                double dik2 = sumsq(rik);
                if (!(dik2 < rc2)) continue;
                double dik = sqrt(dik2);
                double cs = (rij[0] * rik[0] + rij[1] * rik[1] + rij[2] * rik[2]) / (dij * dik);
                E += P.lambda * P.eps * (cs - P.c0) * (cs - P.c0) *
                     exp(P.gamma * P.sigma / (dij - P.a * P.sigma)) *
                     exp(P.gamma * P.sigma / (dik - P.a * P.sigma));
            }
        }
    return E;
}

/* permutedims of a 3x3x3 block stored a + 3b + 9c: out[x1,x2,x3] = in[x_{perm}] */
static double blk(const double* B, int a, int b, int c) { return B[a + 3 * b + 9 * c]; }

/* arr[r1,r2,r3] .+= block (no permutation), third_order_AD.jl:180-188 */
static void add_plain(double* p, long n, int A, int Bm, int C, const double* B, double sgn)
{
    const long n2 = n * n;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c)
                p[(3L * A + a) + n * (3L * Bm + b) + n2 * (3L * C + c)] += sgn * blk(B, a, b, c);
}

/* out[x,y,z] = B[perm applied]: for permutedims!(dst, B, (p1,p2,p3)),
 * dst[x1,x2,x3] = B[x_{q}] with q the inverse; implemented directly.           */
static void add_perm(double* p, long n, int A, int Bm, int C, const double* B,
                     int p1, int p2, int p3)
{
    /* dst[i1,i2,i3] = src[j] where j[p_d] = i_d, i.e. dims of dst are dims (p1,p2,p3) of src */
    const long n2 = n * n;
    int perm[3] = {p1 - 1, p2 - 1, p3 - 1};
    for (int i1 = 0; i1 < 3; ++i1)
        for (int i2 = 0; i2 < 3; ++i2)
            for (int i3 = 0; i3 < 3; ++i3) {
                int idx[3] = {i1, i2, i3};
                int src[3];
                src[perm[0]] = idx[0]; src[perm[1]] = idx[1]; src[perm[2]] = idx[2];
                p[(3L * A + i1) + n * (3L * Bm + i2) + n2 * (3L * C + i3)] +=
                    blk(B, src[0], src[1], src[2]);
            }
}

/* set_third_order_terms!(arr, r1, r2, block, block2): aab, :191-199 */
static void set_aab(double* p, long n, int a, int b, const double* B)
{
    add_perm(p, n, a, a, b, B, 1, 2, 3);
    add_perm(p, n, a, b, a, B, 1, 3, 2);
    add_perm(p, n, b, a, a, B, 3, 2, 1);
}
/* set_third_order_terms_alt!(arr, r1, r2, block, block2): abb, :202-210 */
static void set_abb(double* p, long n, int r1, int r2, const double* B)
{
    add_perm(p, n, r2, r1, r1, B, 1, 2, 3);
    add_perm(p, n, r1, r2, r1, B, 2, 1, 3);
    add_perm(p, n, r1, r1, r2, B, 3, 2, 1);
}
/* 6-permutation version, :212-224 */
static void set_abc(double* p, long n, int a, int b, int c, const double* B)
{
    add_perm(p, n, a, b, c, B, 1, 2, 3);
    add_perm(p, n, a, c, b, B, 1, 3, 2);
    add_perm(p, n, b, a, c, B, 2, 1, 3);
    add_perm(p, n, b, c, a, B, 2, 3, 1);
    add_perm(p, n, c, a, b, B, 3, 1, 2);
    add_perm(p, n, c, b, a, B, 3, 2, 1);
}

/* Block providers of the SW assembly: the blocks themselves, or (g_sw_dparam >= 0, set
 * by fco_sw_ifc3_dparam below) their derivatives with respect to one parameter. */
static int g_sw_dparam = -1;
static void sw_block3_dp(const double* r9, const swp_t* P, double c0, int g0, int g1, int g2,
                         double* out27, int which);
static void sw_pair_block3_dp(const double r[3], const swp_t* P, double* h3, int which);
static void sw_b3(const double* r9, const swp_t* P, double c0, int g0, int g1, int g2, double* out)
{
    if (g_sw_dparam >= 0) sw_block3_dp(r9, P, c0, g0, g1, g2, out, g_sw_dparam);
    else sw_block3(r9, P, c0, g0, g1, g2, out);
}
static void sw_pb3(const double r[3], const swp_t* P, double* h3)
{
    if (g_sw_dparam >= 0) sw_pair_block3_dp(r, P, h3, g_sw_dparam);
    else sw_pair_blocks(r, P, NULL, h3);
}

/* SW third_order! with the AutoDiffCalculator (third_order_AD.jl:71-170) + ASR! +
 * apply_tols!.  c0_fd / c0_iik: cos(theta0) as baked into the FastDifferentiation
 * blocks and as used by the runtime-built iik block.                            */
int fco_sw_ifc3(const double* pos, int N, const double* L, const double* swv, double rcut,
                double tol, double c0_fd, double c0_iik, double* out)
{
    swp_t P = swp_load(swv);
    if (!(rcut <= P.a * P.sigma)) return 2;
    const long n = 3L * N, n2 = n * n;
    memset(out, 0, sizeof(double) * n2 * n);
    const double rc2 = rcut * rcut;
    double B[27], r9[9];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            if (i == j) continue;
            double rij[3];
            for (int c = 0; c < 3; ++c) rij[c] = pos[3 * i + c] - pos[3 * j + c];
            mirror(rij, L);
            if (!(sumsq(rij) < rc2)) continue;
            if (j > i) {
                sw_pb3(rij, &P, B);
                add_plain(out, n, i, i, j, B, -1.0);
                add_plain(out, n, i, j, i, B, -1.0);
                add_plain(out, n, j, i, i, B, -1.0);
                add_plain(out, n, j, j, i, B, 1.0);
                add_plain(out, n, j, i, j, B, 1.0);
                add_plain(out, n, i, j, j, B, 1.0);
            }
            for (int k = j + 1; k < N; ++k) {
                if (k == i) continue;
                double rik[3];
                for (int c = 0; c < 3; ++c) rik[c] = pos[3 * i + c] - pos[3 * k + c];
                mirror(rik, L);
                if (!(sumsq(rik) < rc2)) continue;
                for (int c = 0; c < 3; ++c) {
                    r9[c] = pos[3 * i + c];
                    r9[3 + c] = pos[3 * i + c] - rij[c];
                    r9[6 + c] = pos[3 * i + c] - rik[c];
                }
                sw_b3(r9, &P, c0_fd, 0, 0, 1, B);  set_aab(out, n, i, j, B);     /* iij */
                sw_b3(r9, &P, c0_fd, 0, 1, 1, B);  set_abb(out, n, j, i, B);     /* ijj */
                sw_b3(r9, &P, c0_iik, 0, 0, 2, B); set_aab(out, n, i, k, B);     /* iik */
                sw_b3(r9, &P, c0_fd, 0, 2, 2, B);  set_abb(out, n, k, i, B);     /* ikk */
                sw_b3(r9, &P, c0_fd, 1, 1, 2, B);  set_aab(out, n, j, k, B);     /* jjk */
                sw_b3(r9, &P, c0_fd, 1, 2, 2, B);  set_abb(out, n, k, j, B);     /* jkk */
                sw_b3(r9, &P, c0_fd, 0, 1, 2, B);  set_abc(out, n, i, j, k, B);  /* ijk */
            }
        }
    asr3(out, N);
    apply_tols(out, (size_t)(n2 * n), tol);
    return 0;
}

/* SW second_order! with the AutoDiffCalculator (second_order_AD.jl:39-111) + ASR! +
 * apply_tols!.                                                                  */
int fco_sw_ifc2(const double* pos, int N, const double* L, const double* swv, double rcut,
                double tol, double c0_fd, double* out)
{
    swp_t P = swp_load(swv);
    if (!(rcut <= P.a * P.sigma)) return 2;
    const long n = 3L * N;
    memset(out, 0, sizeof(double) * n * n);
    const double rc2 = rcut * rcut;
    double H[9], H3[27], r9[9];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            if (i == j) continue;
            double rij[3];
            for (int c = 0; c < 3; ++c) rij[c] = pos[3 * i + c] - pos[3 * j + c];
            mirror(rij, L);
            if (!(sumsq(rij) < rc2)) continue;
            if (j > i) {
                sw_pair_blocks(rij, &P, H, NULL);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        out[(3L * i + a) + n * (3L * j + b)] += -H[a + 3 * b];
                        out[(3L * j + a) + n * (3L * i + b)] += -H[a + 3 * b];
                    }
            }
            for (int k = j + 1; k < N; ++k) {
                if (k == i) continue;
                double rik[3];
                for (int c = 0; c < 3; ++c) rik[c] = pos[3 * i + c] - pos[3 * k + c];
                mirror(rik, L);
                if (!(sumsq(rik) < rc2)) continue;
                for (int c = 0; c < 3; ++c) {
                    r9[c] = pos[3 * i + c];
                    r9[3 + c] = pos[3 * i + c] - rij[c];
                    r9[6 + c] = pos[3 * i + c] - rik[c];
                }
                int pr[3][2] = {{i, j}, {i, k}, {j, k}};
                int gr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
                for (int t = 0; t < 3; ++t) {
                    sw_block2(r9, &P, c0_fd, gr[t][0], gr[t][1], H3);
                    int A = pr[t][0], Bq = pr[t][1];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            out[(3L * A + a) + n * (3L * Bq + b)] += H3[a + 3 * b];
                            out[(3L * Bq + a) + n * (3L * A + b)] += H3[b + 3 * a];
                        }
                }
            }
        }
    asr2(out, N);
    apply_tols(out, (size_t)(n * n), tol);
    return 0;
}

/* ========================================================================= */
/* Sparse (COO) form of the LJ third_order result for systems too large for a
 * dense host tensor.  Same semantics as fco_lj_ifc3: for each pair p<q the six
 * permuted stores of the reference are replayed into a local 6x6x6 tensor over
 * the components of {p, q}; ASR! entries (a,b,a) take -(a,b,b); self blocks
 * (a,a,a) are -(sum over k ascending of (a,a,k)).  Entries with |v| < tol are
 * dropped.  Returns the number of entries written (<= cap), or -1 on cutoff
 * failure, -2 if cap is too small.                                             */
static void lj_pair_local(const double* pos, const double* L, double sigma, double eps,
                          double rc2, int p, int q, double loc[216], int* inrange)
{
    memset(loc, 0, sizeof(double) * 216);
    *inrange = 0;
    for (int fam = 0; fam < 2; ++fam) {
        double r[3];
        if (fam == 0)
            for (int c = 0; c < 3; ++c) r[c] = pos[3 * p + c] - pos[3 * q + c];
        else
            for (int c = 0; c < 3; ++c) r[c] = pos[3 * q + c] - pos[3 * p + c];
        mirror(r, L);
        double d2s = sumsq(r);
        if (!(d2s < rc2)) continue;
        *inrange = 1;
        double dist = sqrt(d2s);
        double d1 = lj_d1(sigma, eps, dist), d2 = lj_d2(sigma, eps, dist),
               d3 = lj_d3(sigma, eps, dist);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    double v = -lj_phi3(d1, d2, d3, dist, r, a, b, c);
                    /* local component index: atom p -> 0..2, atom q -> 3..5 */
                    int ii, jj, kk;
                    if (fam == 0) { ii = a; jj = b; kk = 3 + c; }
                    else          { ii = a; jj = 3 + b; kk = 3 + c; }
#define LOC(x, y, z) loc[(x) + 6 * (y) + 36 * (z)]
                    LOC(ii, jj, kk) = v; LOC(ii, kk, jj) = v;
                    LOC(jj, kk, ii) = v; LOC(jj, ii, kk) = v;
                    LOC(kk, ii, jj) = v; LOC(kk, jj, ii) = v;
#undef LOC
                }
    }
}

long fco_lj_ifc3_coo(const double* pos, int N, const double* L, double sigma, double eps,
                     double pot_rcut, double rcut, double tol, long cap, int32_t* oi,
                     int32_t* oj, int32_t* ok, double* ov)
{
    for (int c = 0; c < 3; ++c)
        if (!(pot_rcut < L[c])) return -1;
    const double rc2 = rcut * rcut;
    long m = 0;
    double loc[216];
    double* self = (double*)calloc((size_t)N * 27, sizeof(double));
#define EMIT(I, J, K, V)                                            \
    do {                                                            \
        double v_ = (V);                                            \
        if (!(fabs(v_) < tol)) {                                    \
            if (m >= cap) { free(self); return -2; }                \
            oi[m] = (I); oj[m] = (J); ok[m] = (K); ov[m] = v_; ++m; \
        }                                                           \
    } while (0)
    for (int p = 0; p < N; ++p)
        for (int q = p + 1; q < N; ++q) {
            int inr;
            lj_pair_local(pos, L, sigma, eps, rc2, p, q, loc, &inr);
            if (!inr) continue;
            int at[2] = {p, q};
            for (int x = 0; x < 6; ++x)
                for (int y = 0; y < 6; ++y)
                    for (int z = 0; z < 6; ++z) {
                        int A = at[x / 3], B = at[y / 3], C = at[z / 3];
                        if (A == B && B == C) continue;          /* self: below */
                        double v;
                        if (A == C) {
                            /* ASR! entry (A,B,A) = -(A,B,B) */
                            v = -loc[x + 6 * y + 36 * ((y / 3) * 3 + z % 3)];
                        } else {
                            v = loc[x + 6 * y + 36 * z];
                        }
                        EMIT(3 * A + x % 3, 3 * B + y % 3, 3 * C + z % 3, v);
                    }
        }
    /* self blocks: -(sum over k ascending of Psi[a a, a b, k c]) */
    for (int a = 0; a < N; ++a) {
        double acc[27] = {0};
        for (int k = 0; k < N; ++k) {
            if (k == a) continue;
            int p = a < k ? a : k, q = a < k ? k : a, inr;
            lj_pair_local(pos, L, sigma, eps, rc2, p, q, loc, &inr);
            if (!inr) continue;
            int la = (a == p) ? 0 : 3, lk = (k == p) ? 0 : 3;
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y)
                    for (int z = 0; z < 3; ++z)
                        acc[x + 3 * y + 9 * z] += loc[(la + x) + 6 * (la + y) + 36 * (lk + z)];
        }
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y)
                for (int z = 0; z < 3; ++z)
                    EMIT(3 * a + x, 3 * a + y, 3 * a + z, -acc[x + 3 * y + 9 * z]);
    }
#undef EMIT
    free(self);
    return m;
}

/* ========================================================================= */
/* Forward-mode parameter derivatives of the third-order tensors: the same loops
 * as fco_lj_ifc3 / fco_sw_ifc3, instantiated on dual numbers in one potential
 * parameter, so that every entry carries d Psi / d p beside its value.  ASR! is
 * linear and is applied to the derivatives as to the values; apply_tols! is not
 * (an entry it zeroes has derivative zero: the caller masks with the values).
 * The reference has no parameter derivatives (no ForwardDiff in Project.toml);
 * these are the checker of the GPU parameter pullbacks.                        */

/* ---- LJ: dual numbers a + b eps, eps^2 = 0 ---- */
typedef struct { double v, d; } du;
static du du_c(double x) { du r = {x, 0.0}; return r; }
static du du_add(du a, du b) { du r = {a.v + b.v, a.d + b.d}; return r; }
static du du_sub(du a, du b) { du r = {a.v - b.v, a.d - b.d}; return r; }
static du du_mul(du a, du b) { du r = {a.v * b.v, a.v * b.d + a.d * b.v}; return r; }
static du du_scale(du a, double s) { du r = {a.v * s, a.d * s}; return r; }
static du du_divd(du a, double s) { du r = {a.v / s, a.d / s}; return r; }
static du du_pow(du x, int n)
{
    du r = du_c(1.0);
    for (int t = 0; t < n; ++t) r = du_mul(r, x);
    return r;
}
/* LennardJones.jl:38-52 with sigma, epsilon dual */
static du lj_d1_du(du s, du e, double r)
{
    du s12 = du_pow(s, 12), s6 = du_pow(s, 6);
    return du_mul(du_scale(e, -1.0), du_sub(du_scale(du_divd(s12, ipow(r, 13)), 48.0),
                                            du_scale(du_divd(s6, ipow(r, 7)), 24.0)));
}
static du lj_d2_du(du s, du e, double r)
{
    du s12 = du_pow(s, 12), s6 = du_pow(s, 6);
    return du_mul(du_scale(e, 4.0), du_sub(du_scale(du_divd(s12, ipow(r, 14)), 156.0),
                                           du_scale(du_divd(s6, ipow(r, 8)), 42.0)));
}
static du lj_d3_du(du s, du e, double r)
{
    du s12 = du_pow(s, 12), s6 = du_pow(s, 6);
    return du_mul(du_scale(e, -4.0), du_sub(du_scale(du_divd(s12, ipow(r, 15)), 156.0 * 14.0),
                                            du_scale(du_divd(s6, ipow(r, 9)), 42.0 * 8.0)));
}
/* third_order_analytical.jl:216-224 on duals */
static du lj_phi3_du(du d1, du d2, du d3, double dist, const double r[3], int a, int b, int c)
{
    const double g = ((r[a] * r[b]) * r[c]) / ((dist * dist) * dist);
    du t1 = du_scale(du_add(du_sub(d3, du_divd(du_scale(d2, 3.0), dist)),
                            du_divd(du_scale(d1, 3.0), dist * dist)), g);
    const double dl = ((b == c ? r[a] : 0.0) + (c == a ? r[b] : 0.0)) + (a == b ? r[c] : 0.0);
    du t2 = du_scale(du_sub(d2, du_divd(d1, dist)), dl / (dist * dist));
    return du_add(t1, t2);
}

/* d Psi / d sigma (which = 0) or d Psi / d epsilon (which = 1) of fco_lj_ifc3, before
 * the tolerance pass.  Returns 0, 2 on the cutoff check, 1 on a bad `which`.    */
int fco_lj_ifc3_dparam(const double* pos, int N, const double* L, double sigma, double eps,
                       double pot_rcut, double rcut, int which, double* out)
{
    for (int c = 0; c < 3; ++c)
        if (!(pot_rcut < L[c])) return 2;
    if (which != 0 && which != 1) return 1;
    const long n = 3L * N, n2 = n * n;
    memset(out, 0, sizeof(double) * n2 * n);
    const double rc2 = rcut * rcut;
    const du s = {sigma, which == 0 ? 1.0 : 0.0}, e = {eps, which == 1 ? 1.0 : 0.0};
#define P3(x, y, z) out[(x) + n * (y) + n2 * (z)]
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j)
            for (int fam = 0; fam < 2; ++fam) {
                double r[3];
                if (fam == 0)
                    for (int c = 0; c < 3; ++c) r[c] = pos[3 * i + c] - pos[3 * j + c];
                else
                    for (int c = 0; c < 3; ++c) r[c] = pos[3 * j + c] - pos[3 * i + c];
                mirror(r, L);
                double d2s = sumsq(r);
                if (!(d2s < rc2)) continue;
                double dist = sqrt(d2s);
                du d1 = lj_d1_du(s, e, dist), d2 = lj_d2_du(s, e, dist), d3 = lj_d3_du(s, e, dist);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        for (int c = 0; c < 3; ++c) {
                            double v = -lj_phi3_du(d1, d2, d3, dist, r, a, b, c).d;
                            long ii, jj, kk;
                            if (fam == 0) { ii = 3L * i + a; jj = 3L * i + b; kk = 3L * j + c; }
                            else          { ii = 3L * i + a; jj = 3L * j + b; kk = 3L * j + c; }
                            P3(ii, jj, kk) = v; P3(ii, kk, jj) = v;
                            P3(jj, kk, ii) = v; P3(jj, ii, kk) = v;
                            P3(kk, ii, jj) = v; P3(kk, jj, ii) = v;
                        }
            }
#undef P3
    asr3(out, N);
    return 0;
}

/* ---- SW: order-3 hyper-dual numbers in the coordinates (e1, e2, e3) times a dual
 * direction e4 in one parameter: 16 coefficients, basis index = bitmask of e1..e4,
 * e_i^2 = 0.  The e1 e2 e3 e4 coefficient of a block entry is d/dp of the third
 * derivative.                                                                   */
typedef struct { double v[16]; } hq;
static hq hq_c(double x) { hq r; memset(&r, 0, sizeof r); r.v[0] = x; return r; }
static hq hq_add(hq a, hq b) { for (int i = 0; i < 16; ++i) a.v[i] += b.v[i]; return a; }
static hq hq_sub(hq a, hq b) { for (int i = 0; i < 16; ++i) a.v[i] -= b.v[i]; return a; }
static hq hq_scale(hq a, double s) { for (int i = 0; i < 16; ++i) a.v[i] *= s; return a; }
static hq hq_mul(hq a, hq b)
{
    hq r; memset(&r, 0, sizeof r);
    for (int i = 0; i < 16; ++i) {
        if (a.v[i] == 0.0) continue;
        for (int j = 0; j < 16; ++j) {
            if (i & j) continue;
            r.v[i | j] += a.v[i] * b.v[j];
        }
    }
    return r;
}
/* f(a) from f and its first four derivatives at a.v[0] (the nilpotent part has at
 * most four distinct factors, so the Taylor series ends at the fourth power)   */
static hq hq_apply(hq a, double f0, double f1, double f2, double f3, double f4)
{
    hq d = a; d.v[0] = 0.0;
    hq d2 = hq_mul(d, d), d3 = hq_mul(d2, d), d4 = hq_mul(d3, d);
    hq r = hq_c(f0);
    r = hq_add(r, hq_scale(d, f1));
    r = hq_add(r, hq_scale(d2, f2 / 2.0));
    r = hq_add(r, hq_scale(d3, f3 / 6.0));
    r = hq_add(r, hq_scale(d4, f4 / 24.0));
    return r;
}
static hq hq_sqrt(hq a)
{
    double x = a.v[0], s = sqrt(x);
    return hq_apply(a, s, 0.5 / s, -0.25 / (s * x), 0.375 / (s * x * x),
                    -0.9375 / (s * x * x * x));
}
static hq hq_inv(hq a)
{
    double x = a.v[0], i1 = 1.0 / x, i2 = i1 * i1;
    return hq_apply(a, i1, -i2, 2.0 * i2 * i1, -6.0 * i2 * i2, 24.0 * i2 * i2 * i1);
}
static hq hq_exp(hq a)
{
    double e = exp(a.v[0]);
    return hq_apply(a, e, e, e, e, e);
}
static hq hq_div(hq a, hq b) { return hq_mul(a, hq_inv(b)); }

/* parameters with parameter `which` (index in the swp_t vector layout) seeded on e4 */
typedef struct { hq eps, sigma, a, lambda, gamma, A, B; double p, q; } swq_t;
static swq_t swq_seed(const swp_t* P, int which)
{
    swq_t q;
    q.eps = hq_c(P->eps); q.sigma = hq_c(P->sigma); q.a = hq_c(P->a);
    q.lambda = hq_c(P->lambda); q.gamma = hq_c(P->gamma); q.A = hq_c(P->A); q.B = hq_c(P->B);
    q.p = P->p; q.q = P->q;
    hq* slot[8] = {&q.eps, &q.sigma, &q.a, &q.lambda, &q.gamma, NULL, &q.A, &q.B};
    if (which >= 0 && which < 8 && slot[which]) slot[which]->v[8] = 1.0;
    return q;
}
/* sw_h and sw_pair above, on hq with hq parameters */
static hq sw_h_q(const hq x[9], const swq_t* P, double c0)
{
    hq rij[3], rik[3];
    for (int c = 0; c < 3; ++c) { rij[c] = hq_sub(x[c], x[3 + c]); rik[c] = hq_sub(x[c], x[6 + c]); }
    hq sij = hq_add(hq_add(hq_mul(rij[0], rij[0]), hq_mul(rij[1], rij[1])), hq_mul(rij[2], rij[2]));
    hq sik = hq_add(hq_add(hq_mul(rik[0], rik[0]), hq_mul(rik[1], rik[1])), hq_mul(rik[2], rik[2]));
    hq dij = hq_sqrt(sij), dik = hq_sqrt(sik);
    hq dot = hq_add(hq_add(hq_mul(rij[0], rik[0]), hq_mul(rij[1], rik[1])), hq_mul(rij[2], rik[2]));
    hq cs = hq_div(dot, hq_mul(dij, dik));
    hq dc = hq_sub(cs, hq_c(c0));
    hq le = hq_mul(P->eps, P->lambda), gs = hq_mul(P->sigma, P->gamma), rc = hq_mul(P->a, P->sigma);
    hq e1 = hq_exp(hq_mul(hq_inv(hq_sub(dij, rc)), gs));
    hq e2 = hq_exp(hq_mul(hq_inv(hq_sub(dik, rc)), gs));
    return hq_mul(hq_mul(hq_mul(hq_mul(dc, dc), le), e1), e2);
}
static hq sw_pair_q(const hq r[3], const swq_t* P)
{
    hq s = hq_add(hq_add(hq_mul(r[0], r[0]), hq_mul(r[1], r[1])), hq_mul(r[2], r[2]));
    hq d = hq_sqrt(s);
    hq sr = hq_mul(hq_inv(d), P->sigma);
    hq srp = hq_c(1.0);
    for (int t = 0; t < (int)P->p; ++t) srp = hq_mul(srp, sr);
    hq srq = hq_c(1.0);
    for (int t = 0; t < (int)P->q; ++t) srq = hq_mul(srq, sr);
    hq poly = hq_sub(hq_mul(srp, P->B), srq);
    hq ex = hq_exp(hq_mul(hq_inv(hq_sub(d, hq_mul(P->a, P->sigma))), P->sigma));
    return hq_mul(hq_mul(poly, ex), hq_mul(P->A, P->eps));
}
static void sw_block3_dp(const double* r9, const swp_t* P, double c0, int g0, int g1, int g2,
                         double* out27, int which)
{
    swq_t Q = swq_seed(P, which);
    hq x[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) {
                for (int t = 0; t < 9; ++t) x[t] = hq_c(r9[t]);
                x[3 * g0 + a].v[1] += 1.0;
                x[3 * g1 + b].v[2] += 1.0;
                x[3 * g2 + c].v[4] += 1.0;
                out27[a + 3 * b + 9 * c] = sw_h_q(x, &Q, c0).v[15];
            }
}
static void sw_pair_block3_dp(const double r[3], const swp_t* P, double* h3, int which)
{
    swq_t Q = swq_seed(P, which);
    hq x[3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) {
                for (int t = 0; t < 3; ++t) x[t] = hq_c(r[t]);
                x[a].v[1] += 1.0; x[b].v[2] += 1.0; x[c].v[4] += 1.0;
                h3[a + 3 * b + 9 * c] = sw_pair_q(x, &Q).v[15];
            }
}

/* d Psi / d p of fco_sw_ifc3 before the tolerance pass, p = swv[which] (0 epsilon,
 * 1 sigma, 2 a, 3 lambda, 4 gamma, 6 A, 7 B).  Returns 0, 2 on the cutoff check, 1 on
 * a bad `which`.  Not reentrant (the block provider is switched for the call).    */
int fco_sw_ifc3_dparam(const double* pos, int N, const double* L, const double* swv,
                       double rcut, double c0_fd, double c0_iik, int which, double* out)
{
    if (which < 0 || which > 7 || which == 5) return 1;
    g_sw_dparam = which;
    const int rc = fco_sw_ifc3(pos, N, L, swv, rcut, 0.0, c0_fd, c0_iik, out);
    g_sw_dparam = -1;
    return rc;
}
