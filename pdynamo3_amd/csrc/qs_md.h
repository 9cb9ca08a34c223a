// Gaussian integrals by the McMurchie-Davidson scheme (L. E. McMurchie and E. R. Davidson, J. Comput. Phys. 26 (1978) 218;
// the recurrences as in T. Helgaker, P. Jorgensen, J. Olsen, Molecular Electronic-Structure Theory, ch. 9), for shells of
// angular momentum l <= 2 (s, p, d).  Host and device code: the GPU kernels of the SCF wrapper call these functions, and
// a host build of the same header is checked against the compiled reference on the CPU.
//
// Primitive Cartesian Gaussians x_A^i y_A^j z_A^k exp(-a r_A^2).  For a pair (a at A, b at B): p = a + b,
// P = (a A + b B) / p, mu = a b / p.  Hermite expansion coefficients per direction (E[i][j][t], t <= i + j):
//   E[0][0][0] = exp(-mu X_AB^2)
//   E[i+1][j][t] = E[i][j][t-1] / (2p) + X_PA E[i][j][t] + (t+1) E[i][j][t+1]
//   E[i][j+1][t] = E[i][j][t-1] / (2p) + X_PB E[i][j][t] + (t+1) E[i][j][t+1]
// Overlap S = E^x_0 E^y_0 E^z_0 (pi/p)^(3/2); kinetic energy from one-dimensional overlaps,
//   T_ij = -2 b^2 S_i,j+2 + b (2j+1) S_ij - j (j-1) / 2 S_i,j-2.
// Coulomb-type integrals through the Hermite integrals R_tuv(alpha, X, Y, Z):
//   R^n_000 = (-2 alpha)^n F_n(alpha |XYZ|^2),  R^n_t+1,u,v = t R^n+1_t-1,u,v + X R^n+1_t,u,v  (and so for u, v),
// with F_n the Boys function:
//   nuclear attraction  (a|1/r_C|b) = (2 pi / p) sum_tuv E_tuv R_tuv(p, P - C)
//   three-centre        (ab|c) = 2 pi^(5/2) / (p g sqrt(p + g)) sum_tuv E^ab_tuv sum_t'u'v' (-1)^(t'+u'+v') E^c_t'u'v'
//                                R_t+t',u+u',v+v'(p g / (p + g), P - C)
//   two-centre          (c|d) as the three-centre form with the pair replaced by the single Gaussian c.
// A single Gaussian's Hermite coefficients follow from the pair recurrence with b = 0 (E[0][0][0] = 1, X_PA = 0).
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define QS_HD __host__ __device__ inline
#else
#define QS_HD inline
#endif

namespace qs {

constexpr int kMaxL = 2;                       // largest angular momentum of one shell
constexpr int kMaxR = 3 * kMaxL;               // largest t + u + v of a Hermite integral (three d shells)
constexpr int kBoysOrders = kMaxR + 1;         // F_0 .. F_6 are needed
constexpr int kBoysTaylor = 6;                 // Taylor terms beyond the tabulated order
constexpr int kBoysTable = kBoysOrders + kBoysTaylor + 1;   // tabulated orders F_0 .. F_13
constexpr double kBoysStep = 0.05;             // table spacing in T
constexpr int kBoysPoints = 801;               // T = 0 .. 40
constexpr double kBoysTMax = kBoysStep * (kBoysPoints - 1);

QS_HD int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian powers of component c of angular momentum l in the reference's order (for z = 0..l, for y = 0..l-z, x = l-y-z).
QS_HD void cart_powers(int l, int c, int &x, int &y, int &z) {
    int n = 0;
    for (int zz = 0; zz <= l; zz++)
        for (int yy = 0; yy <= l - zz; yy++) {
            if (n == c) { x = l - yy - zz; y = yy; z = zz; return; }
            n++;
        }
    x = y = z = 0;
}

// Boys function F_0 .. F_m (m <= kMaxR) at T from the table F[k][n] = F_n(k step), n < kBoysTable: a Taylor expansion of
// F_m about the nearest tabulated T (|dT| <= step / 2, kBoysTaylor + 1 terms) and downward recursion
// F_n = (2 T F_n+1 + exp(-T)) / (2n + 1); beyond the table, F_0 = sqrt(pi / T) / 2 and upward recursion
// F_n+1 = ((2n + 1) F_n - exp(-T)) / (2T), which is stable for large T.
QS_HD void boys(int m, double T, const double *__restrict__ table, double *F) {
    if (T < kBoysTMax) {
        const int k = static_cast<int>(T / kBoysStep + 0.5);
        const double d = T - k * kBoysStep;
        const double *row = table + k * kBoysTable;
        double s = 0.0, term = 1.0;
        for (int j = 0; j <= kBoysTaylor; j++) {
            s += row[m + j] * term;
            term *= -d / (j + 1);
        }
        F[m] = s;
        const double e = exp(-T);
        for (int n = m - 1; n >= 0; n--) F[n] = (2.0 * T * F[n + 1] + e) / (2 * n + 1);
    } else {
        const double e = exp(-T);
        F[0] = 0.5 * sqrt(M_PI / T);
        for (int n = 0; n < m; n++) F[n + 1] = ((2 * n + 1) * F[n] - e) / (2.0 * T);
    }
}

// One-dimensional Hermite coefficients E[i][j][t] of a pair, i <= la, j <= lb (lb <= kMaxL + 2 for the kinetic energy).
// Layout E[(i * 5 + j) * 7 + t].
constexpr int kEJ = 5, kET = 7;
QS_HD void hermite_1d(int la, int lb, double p, double XPA, double XPB, double e00, double *E) {
    const double h = 0.5 / p;
    for (int k = 0; k < (kMaxL + 1) * kEJ * kET; k++) E[k] = 0.0;
    E[0] = e00;
    for (int i = 0; i < la; i++)
        for (int t = 0; t <= i + 1; t++) {
            const double em = t > 0 ? E[(i * kEJ) * kET + t - 1] : 0.0;
            const double e0 = t <= i ? E[(i * kEJ) * kET + t] : 0.0;
            const double ep = t + 1 <= i ? E[(i * kEJ) * kET + t + 1] : 0.0;
            E[((i + 1) * kEJ) * kET + t] = h * em + XPA * e0 + (t + 1) * ep;
        }
    for (int i = 0; i <= la; i++)
        for (int j = 0; j < lb; j++)
            for (int t = 0; t <= i + j + 1; t++) {
                const double *r = E + (i * kEJ + j) * kET;
                const double em = t > 0 ? r[t - 1] : 0.0;
                const double e0 = t <= i + j ? r[t] : 0.0;
                const double ep = t + 1 <= i + j ? r[t + 1] : 0.0;
                E[(i * kEJ + j + 1) * kET + t] = h * em + XPB * e0 + (t + 1) * ep;
            }
}

// Hermite coefficients of a single Gaussian x^i exp(-g x^2), i <= kMaxL: G[i * 3 + t].
QS_HD void hermite_single(int l, double g, double *G) {
    const double h = 0.5 / g;
    for (int k = 0; k < 9; k++) G[k] = 0.0;
    G[0] = 1.0;
    for (int i = 0; i < l; i++)
        for (int t = 0; t <= i + 1; t++) {
            const double em = t > 0 ? G[i * 3 + t - 1] : 0.0;
            const double ep = t + 1 <= i ? G[i * 3 + t + 1] : 0.0;
            G[(i + 1) * 3 + t] = h * em + (t + 1) * ep;
        }
}

// Hermite integrals R_tuv = R^0_tuv for t + u + v <= L (L <= kMaxR), R[(t * 7 + u) * 7 + v].  The levels R^n are formed
// in place from n = L down to 0: the entries of total order s at level n take orders s - 1 and s - 2 of level n + 1,
// which still hold level n + 1 while s runs downwards.
constexpr int kRD = kMaxR + 1;
QS_HD int ridx(int t, int u, int v) { return (t * kRD + u) * kRD + v; }
QS_HD void hermite_r(int L, double alpha, double X, double Y, double Z, const double *__restrict__ table, double *R) {
    double F[kBoysOrders];
    boys(L, alpha * (X * X + Y * Y + Z * Z), table, F);
    // Level n = L: only R^L_000.
    double fac = 1.0;
    for (int n = 0; n < L; n++) fac *= -2.0 * alpha;
    R[0] = fac * F[L];
    for (int n = L - 1; n >= 0; n--) {
        fac /= -2.0 * alpha;
        for (int s = L - n; s >= 1; s--)
            for (int t = 0; t <= s; t++)
                for (int u = 0; u <= s - t; u++) {
                    const int v = s - t - u;
                    double r;
                    if (t > 0) r = (t > 1 ? (t - 1) * R[ridx(t - 2, u, v)] : 0.0) + X * R[ridx(t - 1, u, v)];
                    else if (u > 0) r = (u > 1 ? (u - 1) * R[ridx(t, u - 2, v)] : 0.0) + Y * R[ridx(t, u - 1, v)];
                    else r = (v > 1 ? (v - 1) * R[ridx(t, u, v - 2)] : 0.0) + Z * R[ridx(t, u, v - 1)];
                    R[ridx(t, u, v)] = r;
                }
        R[0] = fac * F[n];
    }
}

}  // namespace qs
