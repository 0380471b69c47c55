// Stillinger-Weber derivative blocks in closed form (FP64, device code).
//
// Three-body term (SW.jl:60-83), with u = r_ij = r_i - r_j and v = r_ik = r_i - r_k:
//   h = le * (C - c0)^2 * e(p) * e(q),   p = u.u, q = v.v, s = u.v, C = s / sqrt(p q),
//   e(x) = exp(gs / (sqrt(x) - rc)),     le = lambda*eps, gs = gamma*sigma, rc = a*sigma.
// h depends on the positions only through the three quadratic forms t = (p, q, s), so
// with H(p, q, s) the chain rule closes after two terms (t'' is constant, t''' = 0):
//   d3h/dx dy dz = sum_abc H_abc t_a,x t_b,y t_c,z
//                + sum_ab  H_ab (t_a,xy t_b,z + t_a,xz t_b,y + t_a,yz t_b,x)
//   d2h/dx dy    = sum_ab  H_ab t_a,x t_b,y + sum_a H_a t_a,xy
// For a derivative w.r.t. atom G in {i, j, k}, component alpha:
//   dp = 2 su(G) u_alpha, dq = 2 sv(G) v_alpha, ds = su(G) v_alpha + sv(G) u_alpha,
//   su = (1, -1, 0), sv = (1, 0, -1), and the second derivatives of t are
//   2 su su' delta, 2 sv sv' delta, (su sv' + sv su') delta.
// The partials of H in (p, q, s) factor: C = s * p^-1/2 * q^-1/2 is a product of
// one-variable functions, F = (C - c0)^2 follows by Faa di Bruno, and H = le F e(p) e(q)
// by the Leibniz rule.
//
// Pair term phi2(r) = A eps (B (sigma/r)^p - (sigma/r)^q) exp(sigma/(r - rc)) (SW.jl:76-78):
// the radial derivatives phi2', phi2'', phi2''' by the product rule, then the usual
// radial tensors (the same structure as the LJ phi3 of third_order_analytical.jl:216-224).
#pragma once

#include <hip/hip_runtime.h>

namespace fc3 {
namespace sw {

struct Params {
    double eps, sigma, a, lambda, gamma, c0, A, B, p, q;
};

// Derivatives of x -> exp(gs / (sqrt(x) - rc)) with respect to x, orders 0..3; with de,
// also their derivatives with respect to gamma (gs = gamma sigma: the exponent and all
// its x-derivatives are linear in gs, so d/dgamma multiplies each by 1/gamma).
__device__ __forceinline__ void exp_radial_sq(double x, double gs, double rc, double e[4],
                                              double* de = nullptr, double gamma = 1.0)
{
    const double rho = sqrt(x);
    const double d = rho - rc;
    const double ps1 = -gs / (d * d), ps2 = 2.0 * gs / (d * d * d), ps3 = -6.0 * gs / (d * d * d * d);
    const double r1 = 0.5 / rho, r2 = -0.25 / (rho * x), r3 = 0.375 / (rho * x * x);
    const double f1 = ps1 * r1;
    const double f2 = ps2 * r1 * r1 + ps1 * r2;
    const double f3 = ps3 * r1 * r1 * r1 + 3.0 * ps2 * r1 * r2 + ps1 * r3;
    const double e0 = exp(gs / d);
    e[0] = e0;
    e[1] = e0 * f1;
    e[2] = e0 * (f2 + f1 * f1);
    e[3] = e0 * (f3 + 3.0 * f1 * f2 + f1 * f1 * f1);
    if (de) {
        const double g0 = gs / d, ig = 1.0 / gamma;
        de[0] = e0 * g0 * ig;
        de[1] = e0 * f1 * (g0 + 1.0) * ig;
        de[2] = e0 * (g0 * (f2 + f1 * f1) + f2 + 2.0 * f1 * f1) * ig;
        de[3] = e0 * (g0 * (f3 + 3.0 * f1 * f2 + f1 * f1 * f1) + f3 + 6.0 * f1 * f2 +
                      3.0 * f1 * f1 * f1) * ig;
    }
}

// Partials of H(p, q, s) = le (C - c0)^2 e(p) e(q): Hd[i][j][k] = d^(i+j+k) H / dp^i dq^j ds^k
// for i + j + k <= 3; with dHg, also their derivatives with respect to gamma.
__device__ inline void h_partials(double p, double q, double s, double c0, double le, double gs,
                                  double rc, double Hd[4][4][4],
                                  double (*dHg)[4][4] = nullptr, double gamma = 1.0)
{
    const double a = sqrt(p), b = sqrt(q);
    // pi(p) = p^-1/2 and kappa(q) = q^-1/2 with derivatives 0..3
    const double pi[4] = {1.0 / a, -0.5 / (a * p), 0.75 / (a * p * p), -1.875 / (a * p * p * p)};
    const double ka[4] = {1.0 / b, -0.5 / (b * q), 0.75 / (b * q * q), -1.875 / (b * q * q * q)};
    const double sg[4] = {s, 1.0, 0.0, 0.0};
    // C partials: Cd(i, j, k) = sg[k] pi[i] ka[j]
    auto Cd = [&](int i, int j, int k) -> double {
        if (i > 3 || j > 3 || k > 3) return 0.0;
        return sg[k] * pi[i] * ka[j];
    };
    const double dc = Cd(0, 0, 0) - c0;
    // F partial for a multi-index (i, j, k), i + j + k <= 3, via Faa di Bruno on (.)^2
    auto Fd = [&](int i, int j, int k) -> double {
        const int ord = i + j + k;
        if (ord == 0) return dc * dc;
        int lab[3], nl = 0;  // the variables being differentiated, as labels 0 p, 1 q, 2 s
        for (int t = 0; t < i; ++t) lab[nl++] = 0;
        for (int t = 0; t < j; ++t) lab[nl++] = 1;
        for (int t = 0; t < k; ++t) lab[nl++] = 2;
        auto C1 = [&](int x) { return Cd(x == 0, x == 1, x == 2); };
        auto C2 = [&](int x, int y) {
            return Cd((x == 0) + (y == 0), (x == 1) + (y == 1), (x == 2) + (y == 2));
        };
        if (ord == 1) return 2.0 * dc * C1(lab[0]);
        if (ord == 2) return 2.0 * (C1(lab[0]) * C1(lab[1]) + dc * C2(lab[0], lab[1]));
        const int x = lab[0], y = lab[1], z = lab[2];
        const double Cxyz = Cd(i, j, k);
        return 2.0 * (C2(x, y) * C1(z) + C2(x, z) * C1(y) + C2(y, z) * C1(x) + dc * Cxyz);
    };
    double ep[4], eq[4], dep[4], deq[4];
    exp_radial_sq(p, gs, rc, ep, dHg ? dep : nullptr, gamma);
    exp_radial_sq(q, gs, rc, eq, dHg ? deq : nullptr, gamma);
    const double binom[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 1, 0}, {1, 3, 3, 1}};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) {
                if (i + j + k > 3) {
                    Hd[i][j][k] = 0.0;
                    if (dHg) dHg[i][j][k] = 0.0;
                    continue;
                }
                double acc = 0.0, dacc = 0.0;
                for (int i1 = 0; i1 <= i; ++i1)
                    for (int j1 = 0; j1 <= j; ++j1) {
                        const double w = binom[i][i1] * binom[j][j1] * Fd(i - i1, j - j1, k);
                        acc += w * ep[i1] * eq[j1];
                        if (dHg) dacc += w * (dep[i1] * eq[j1] + ep[i1] * deq[j1]);
                    }
                Hd[i][j][k] = le * acc;
                if (dHg) dHg[i][j][k] = le * dacc;
            }
}

// First / second derivative coefficients of t = (p, q, s) for groups G = 0 (i), 1 (j), 2 (k).
struct TForms {
    double T[3][3][3];  // T[a][G][alpha]
    double S[3][3][3];  // S[a][G1][G2]  (times delta_{alpha beta})
};

__device__ inline void t_forms(const double u[3], const double v[3], TForms& tf)
{
    const double su[3] = {1.0, -1.0, 0.0}, sv[3] = {1.0, 0.0, -1.0};
    for (int G = 0; G < 3; ++G)
        for (int al = 0; al < 3; ++al) {
            tf.T[0][G][al] = 2.0 * su[G] * u[al];
            tf.T[1][G][al] = 2.0 * sv[G] * v[al];
            tf.T[2][G][al] = su[G] * v[al] + sv[G] * u[al];
        }
    for (int G1 = 0; G1 < 3; ++G1)
        for (int G2 = 0; G2 < 3; ++G2) {
            tf.S[0][G1][G2] = 2.0 * su[G1] * su[G2];
            tf.S[1][G1][G2] = 2.0 * sv[G1] * sv[G2];
            tf.S[2][G1][G2] = su[G1] * sv[G2] + sv[G1] * su[G2];
        }
}

__device__ __forceinline__ double H2(const double Hd[4][4][4], int a, int b)
{
    return Hd[(a == 0) + (b == 0)][(a == 1) + (b == 1)][(a == 2) + (b == 2)];
}
__device__ __forceinline__ double H3(const double Hd[4][4][4], int a, int b, int c)
{
    return Hd[(a == 0) + (b == 0) + (c == 0)][(a == 1) + (b == 1) + (c == 1)]
             [(a == 2) + (b == 2) + (c == 2)];
}

// out[al + 3 be + 9 ga] = d3h / d r_{G1,al} d r_{G2,be} d r_{G3,ga}
__device__ inline void block3(const double Hd[4][4][4], const TForms& tf, int G1, int G2, int G3,
                              double* out)
{
    double h3[3][3][3], h2[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            h2[a][b] = H2(Hd, a, b);
            for (int c = 0; c < 3; ++c) h3[a][b][c] = H3(Hd, a, b, c);
        }
    // X[b][c][al] = sum_a h3[a][b][c] T[a][G1][al]
    double X[3][3][3];
    for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c)
            for (int al = 0; al < 3; ++al) {
                double acc = 0.0;
                for (int a = 0; a < 3; ++a) acc += h3[a][b][c] * tf.T[a][G1][al];
                X[b][c][al] = acc;
            }
    // Y[c][al][be] = sum_b X[b][c][al] T[b][G2][be]
    double Y[3][3][3];
    for (int c = 0; c < 3; ++c)
        for (int al = 0; al < 3; ++al)
            for (int be = 0; be < 3; ++be) {
                double acc = 0.0;
                for (int b = 0; b < 3; ++b) acc += X[b][c][al] * tf.T[b][G2][be];
                Y[c][al][be] = acc;
            }
    // second-order terms: W12[b] = sum_a h2[a][b] S[a][G1][G2] etc.
    double W12[3], W13[3], W23[3];
    for (int b = 0; b < 3; ++b) {
        double x12 = 0.0, x13 = 0.0, x23 = 0.0;
        for (int a = 0; a < 3; ++a) {
            x12 += h2[a][b] * tf.S[a][G1][G2];
            x13 += h2[a][b] * tf.S[a][G1][G3];
            x23 += h2[a][b] * tf.S[a][G2][G3];
        }
        W12[b] = x12; W13[b] = x13; W23[b] = x23;
    }
    for (int ga = 0; ga < 3; ++ga)
        for (int be = 0; be < 3; ++be)
            for (int al = 0; al < 3; ++al) {
                double acc = 0.0;
                for (int c = 0; c < 3; ++c) acc += Y[c][al][be] * tf.T[c][G3][ga];
                double t2 = 0.0;
                for (int b = 0; b < 3; ++b) {
                    if (al == be) t2 += W12[b] * tf.T[b][G3][ga];
                    if (al == ga) t2 += W13[b] * tf.T[b][G2][be];
                    if (be == ga) t2 += W23[b] * tf.T[b][G1][al];
                }
                out[al + 3 * be + 9 * ga] = acc + t2;
            }
}

// out[al + 3 be] = d2h / d r_{G1,al} d r_{G2,be}
__device__ inline void block2(const double Hd[4][4][4], const TForms& tf, int G1, int G2,
                              double* out)
{
    for (int be = 0; be < 3; ++be)
        for (int al = 0; al < 3; ++al) {
            double acc = 0.0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    acc += H2(Hd, a, b) * tf.T[a][G1][al] * tf.T[b][G2][be];
            if (al == be) {
                const double h1[3] = {Hd[1][0][0], Hd[0][1][0], Hd[0][0][1]};
                for (int a = 0; a < 3; ++a) acc += h1[a] * tf.S[a][G1][G2];
            }
            out[al + 3 * be] = acc;
        }
}

// Radial derivatives of the pair term, orders 1..3.
__device__ inline void pair_radial(double r, const Params& P, double& d1, double& d2, double& d3)
{
    const double x = P.sigma / r;
    const double xp = pow(x, P.p), xq = pow(x, P.q);
    const double Ae = P.A * P.eps;
    const double f0 = Ae * (P.B * xp - xq);
    const double f1 = -Ae * (P.B * P.p * xp - P.q * xq) / r;
    const double f2 = Ae * (P.B * P.p * (P.p + 1.0) * xp - P.q * (P.q + 1.0) * xq) / (r * r);
    const double f3 = -Ae * (P.B * P.p * (P.p + 1.0) * (P.p + 2.0) * xp -
                             P.q * (P.q + 1.0) * (P.q + 2.0) * xq) / (r * r * r);
    const double rc = P.a * P.sigma;
    const double d = r - rc;
    const double s1 = -P.sigma / (d * d), s2 = 2.0 * P.sigma / (d * d * d),
                 s3 = -6.0 * P.sigma / (d * d * d * d);
    const double g0 = exp(P.sigma / d);
    const double g1 = g0 * s1, g2 = g0 * (s2 + s1 * s1), g3 = g0 * (s3 + 3.0 * s1 * s2 + s1 * s1 * s1);
    d1 = f1 * g0 + f0 * g1;
    d2 = f2 * g0 + 2.0 * f1 * g1 + f0 * g2;
    d3 = f3 * g0 + 3.0 * f2 * g1 + 3.0 * f1 * g2 + f0 * g3;
}

// Hessian (9) and third-derivative tensor (27) of phi2(|r|) w.r.t. the vector r.
__device__ inline void pair_blocks(const double r[3], const Params& P, double* h2, double* h3)
{
    const double dist = sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
    double d1, d2, d3;
    pair_radial(dist, P, d1, d2, d3);
    const double A3 = (d3 - 3.0 * d2 / dist + 3.0 * d1 / (dist * dist)) / (dist * dist * dist);
    const double B2 = (d2 - d1 / dist) / (dist * dist);
    if (h3)
        for (int c = 0; c < 3; ++c)
            for (int b = 0; b < 3; ++b)
                for (int a = 0; a < 3; ++a) {
                    const double dl = (b == c ? r[a] : 0.0) + (a == c ? r[b] : 0.0) +
                                      (a == b ? r[c] : 0.0);
                    h3[a + 3 * b + 9 * c] = r[a] * r[b] * r[c] * A3 + dl * B2;
                }
    if (h2)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a)
                h2[a + 3 * b] = r[a] * r[b] * B2 + (a == b ? d1 / dist : 0.0);
}

}  // namespace sw
}  // namespace fc3
