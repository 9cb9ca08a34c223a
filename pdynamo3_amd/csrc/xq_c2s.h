// Host-side helpers shared by the quadrature's basis set-up and the SCF wrapper's integrals: factorials and the
// normalised Cartesian -> real solid harmonic transformation of Schlegel & Frisch (Int. J. Quantum Chem. 54 (1995) 83,
// eq. 15) in the reference's function orders (Cartesian: for z = 0..l, for y = 0..l-z, x = l-y-z; spherical: m = 0, 1c,
// 1s, 2c, 2s, ...; GaussianBasis.c:13-27).
#pragma once

#include <cmath>

namespace xq {

namespace {

// Cartesian powers of component c of angular momentum l (l <= 3) in the reference order.
inline void c2s_cart(int l, int c, int &x, int &y, int &z) {
    int n = 0;
    for (int zz = 0; zz <= l; zz++)
        for (int yy = 0; yy <= l - zz; yy++) {
            if (n == c) { x = l - yy - zz; y = yy; z = zz; return; }
            n++;
        }
    x = y = z = 0;
}

double odd_fact(int n) {
    double f = 1.0;
    for (int i = 3; i <= n; i += 2) f *= i;
    return f;
}
double fact(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; i++) f *= i;
    return f;
}
double binom(int n, int k) { return (k < 0 || k > n) ? 0.0 : fact(n) / (fact(k) * fact(n - k)); }

// c2s block for one l (normalised Cartesian -> normalised real solid harmonics), written at (r0, c0) of a row-major matrix
// with ns columns.  Spherical order m = 0, 1c, 1s, 2c, 2s, ...
void c2s_block(int l, double *out, int ns, int r0, int c0) {
    for (int c = 0; c < (l + 1) * (l + 2) / 2; c++) {
        int x, y, z;
        c2s_cart(l, c, x, y, z);
        for (int m = 0; m <= l; m++) {
            const int jj = l - m - z;
            if (jj < 0 || (jj % 2) != 0) continue;
            const int j = jj / 2;
            double s0 = 0.0, sc = 0.0, ss = 0.0;
            for (int i = 0; i <= (l - m) / 2; i++) {
                if (j > i) continue;
                const double t = binom(l, i) * binom(i, j) * ((i % 2) ? -1.0 : 1.0) * fact(2 * l - 2 * i) / fact(l - m - 2 * i);
                for (int k = 0; k <= j; k++) {
                    const int p = x - 2 * k;  // power of x taken from (x + i y)^m
                    if (p < 0 || p > m) continue;
                    const double u = t * binom(j, k) * binom(m, p);
                    const int q = m - p;      // power of (i y)
                    const double sign = ((q / 2) % 2) ? -1.0 : 1.0;
                    if (m == 0) s0 += u;
                    else if ((q % 2) == 0) sc += std::sqrt(2.0) * sign * u;
                    else ss += std::sqrt(2.0) * sign * u;
                }
            }
            const double norm = std::sqrt(fact(2 * x) * fact(2 * y) * fact(2 * z) * fact(l) / (fact(2 * l) * fact(x) * fact(y) * fact(z))) *
                                std::sqrt(fact(l - m) / fact(l + m)) / (std::pow(2.0, l) * fact(l));
            if (m == 0) out[(r0 + c) * ns + c0] = norm * s0;
            else {
                out[(r0 + c) * ns + c0 + 2 * m - 1] = norm * sc;
                out[(r0 + c) * ns + c0 + 2 * m] = norm * ss;
            }
        }
    }
}

}  // namespace

}  // namespace xq
