// The integral work items of the SCF wrapper (one shell pair, fit shell pair or shell pair x fit shell each), as host and
// device functions: the GPU kernels of qs_integrals.hip call them one item per thread, and a host build of this header
// checks the same code against the compiled reference on the CPU.  The mathematics is in qs_md.h; the Boys table and the
// c2s transformation come in as arguments.
#pragma once

#include "qs_md.h"

namespace qs {

struct QPair { int x, y; };

struct QShell {                 // one angular momentum of one contracted shell (an SP shell gives two)
    double x, y, z;
    int l, nPrim, prim;         // primitives: exponents and coefficients from offset prim
    int f0, nf;                 // first function and function count (2l+1 spherical or Cartesian)
    int atom, pad;
};

struct PairEntry {              // a significant shell pair (A >= B) and its first row of electron-fit integrals
    int a, b;
    long long row;
};

constexpr double kUnderFlow = 1.0e-12;    // electron-fit integrals dropped by the reference's storage
// Nuclei are Gaussian charge distributions of width 1e-4 bohr, exponent 4 / (w^2 pi), as the reference's electron-nuclear
// integrals take them by default (GaussianNucleus.h: _DefaultNuclearWidth, _DefaultNuclearWidthE).
constexpr double kNuclearWidth = 1.0e-4;
constexpr double kNuclearExponent = 4.0 / (kNuclearWidth * kNuclearWidth * M_PI);

// Cartesian factor of component c of angular momentum l: sqrt((2l-1)!! / ((2x-1)!! (2y-1)!! (2z-1)!!)).
QS_HD double cart_factor(int l, int c) {
    int x, y, z;
    cart_powers(l, c, x, y, z);
    auto of = [](int m) { double f = 1.0; for (int i = 3; i <= m; i += 2) f *= i; return f; };
    return sqrt(of(2 * l - 1) / (of(2 * x - 1) * of(2 * y - 1) * of(2 * z - 1)));
}

// Transform a Cartesian block in place along one index: [outer][ncart(l)][inner] -> [outer][2l+1][inner]; c2s holds
// the transformation of l = 0..2 as [l][6][5].  Used only when the basis is spherical.
QS_HD void to_spherical(int l, const double *c2s, double *blk, int outer, int inner, double *tmp) {
    const int nc = ncart(l), ns = 2 * l + 1;
    for (int o = 0; o < outer; o++)
        for (int i = 0; i < inner; i++) {
            for (int s = 0; s < ns; s++) {
                double v = 0.0;
                for (int c = 0; c < nc; c++) v += c2s[(l * 6 + c) * 5 + s] * blk[(o * nc + c) * inner + i];
                tmp[s] = v;
            }
            for (int s = 0; s < ns; s++) blk[(o * ns + s) * inner + i] = tmp[s];
        }
}

QS_HD long long packed_index(int i, int j) {
    return i >= j ? static_cast<long long>(i) * (i + 1) / 2 + j : static_cast<long long>(j) * (j + 1) / 2 + i;
}

// Overlap, kinetic and electron-nuclear blocks of shell pair k (A >= B), written into the packed matrices (mu >= nu).
QS_HD void one_electron_pair(int k, const QPair *__restrict__ pairs, const QShell *__restrict__ shells,
                             const double *__restrict__ exps, const double *__restrict__ coefs, int nAtoms,
                             const double *__restrict__ atoms, int spherical, const double *__restrict__ boysTable,
                             const double *__restrict__ c2s, double *__restrict__ S, double *__restrict__ T,
                             double *__restrict__ V) {
    const QShell A = shells[pairs[k].x], B = shells[pairs[k].y];
    const int na = ncart(A.l), nb = ncart(B.l);
    double s[36], t[36], v[36], tmp[6];
    for (int i = 0; i < 36; i++) { s[i] = 0.0; t[i] = 0.0; v[i] = 0.0; }
    double Ex[(kMaxL + 1) * kEJ * kET], Ey[(kMaxL + 1) * kEJ * kET], Ez[(kMaxL + 1) * kEJ * kET];
    double R[kRD * kRD * kRD];
    const double ABx = A.x - B.x, ABy = A.y - B.y, ABz = A.z - B.z;
    for (int qa = 0; qa < A.nPrim; qa++)
        for (int qb = 0; qb < B.nPrim; qb++) {
            const double a = exps[A.prim + qa], b = exps[B.prim + qb], p = a + b, mu = a * b / p;
            const double cab = coefs[A.prim + qa] * coefs[B.prim + qb];
            const double Px = (a * A.x + b * B.x) / p, Py = (a * A.y + b * B.y) / p, Pz = (a * A.z + b * B.z) / p;
            hermite_1d(A.l, B.l + 2, p, Px - A.x, Px - B.x, exp(-mu * ABx * ABx), Ex);
            hermite_1d(A.l, B.l + 2, p, Py - A.y, Py - B.y, exp(-mu * ABy * ABy), Ey);
            hermite_1d(A.l, B.l + 2, p, Pz - A.z, Pz - B.z, exp(-mu * ABz * ABz), Ez);
            const double sq = sqrt(M_PI / p);
            auto S1 = [&](const double *E, int i, int j) { return j < 0 ? 0.0 : E[(i * kEJ + j) * kET] * sq; };
            for (int ia = 0; ia < na; ia++) {
                int xa, ya, za;
                cart_powers(A.l, ia, xa, ya, za);
                for (int ib = 0; ib < nb; ib++) {
                    int xb, yb, zb;
                    cart_powers(B.l, ib, xb, yb, zb);
                    const double sx = S1(Ex, xa, xb), sy = S1(Ey, ya, yb), sz = S1(Ez, za, zb);
                    auto T1 = [&](const double *E, int i, int j) {
                        return -2.0 * b * b * S1(E, i, j + 2) + b * (2 * j + 1) * S1(E, i, j) -
                               0.5 * j * (j - 1) * S1(E, i, j - 2);
                    };
                    const double c = cab * cart_factor(A.l, ia) * cart_factor(B.l, ib);
                    s[ia * nb + ib] += c * sx * sy * sz;
                    t[ia * nb + ib] += c * (T1(Ex, xa, xb) * sy * sz + sx * T1(Ey, ya, yb) * sz + sx * sy * T1(Ez, za, zb));
                }
            }
            // Nuclear attraction of Gaussian nuclei Z (z/pi)^(3/2) exp(-z r^2) (z = kNuclearExponent):
            // -Z_C (2 pi / p) sqrt(z / (p + z)) sum_tuv E_tuv R_tuv(p z / (p + z), P - C).
            const int L = A.l + B.l;
            const double zN = kNuclearExponent, aN = p * zN / (p + zN), fN = 2.0 * M_PI / p * sqrt(zN / (p + zN));
            for (int C = 0; C < nAtoms; C++) {
                const double Zc = atoms[4 * C + 3];
                hermite_r(L, aN, Px - atoms[4 * C], Py - atoms[4 * C + 1], Pz - atoms[4 * C + 2], boysTable, R);
                const double f = -Zc * fN * cab;
                for (int ia = 0; ia < na; ia++) {
                    int xa, ya, za;
                    cart_powers(A.l, ia, xa, ya, za);
                    for (int ib = 0; ib < nb; ib++) {
                        int xb, yb, zb;
                        cart_powers(B.l, ib, xb, yb, zb);
                        double acc = 0.0;
                        for (int tt = 0; tt <= xa + xb; tt++)
                            for (int uu = 0; uu <= ya + yb; uu++)
                                for (int vv = 0; vv <= za + zb; vv++)
                                    acc += Ex[(xa * kEJ + xb) * kET + tt] * Ey[(ya * kEJ + yb) * kET + uu] *
                                           Ez[(za * kEJ + zb) * kET + vv] * R[ridx(tt, uu, vv)];
                        v[ia * nb + ib] += f * cart_factor(A.l, ia) * cart_factor(B.l, ib) * acc;
                    }
                }
            }
        }
    int ma = na, mb = nb;
    if (spherical) {
        double *blks[3] = {s, t, v};
        for (double *blk : blks) {
            to_spherical(A.l, c2s, blk, 1, nb, tmp);          // rows: A's functions
            to_spherical(B.l, c2s, blk, 2 * A.l + 1, 1, tmp); // columns: B's functions
        }
        ma = 2 * A.l + 1;
        mb = 2 * B.l + 1;
    }
    for (int i = 0; i < ma; i++)
        for (int j = 0; j < mb; j++) {
            const int mu = A.f0 + i, nu = B.f0 + j;
            if (mu < nu) continue;                     // the diagonal pair's upper triangle
            const long long o = packed_index(mu, nu);
            S[o] = s[i * mb + j];
            T[o] = t[i * mb + j];
            V[o] = v[i * mb + j];
        }
}

// (C|D) of fit shell pair k (C >= D) into the packed fit matrix; the diagonal pair also writes its functions' integrals
// into row nFit.
QS_HD void fit_matrix_pair(int k, const QPair *__restrict__ pairs, const QShell *__restrict__ shells,
                           const double *__restrict__ exps, const double *__restrict__ coefs, int nFit, int spherical,
                           const double *__restrict__ boysTable, const double *__restrict__ c2s, double *__restrict__ M) {
    const QShell C = shells[pairs[k].x], D = shells[pairs[k].y];
    const int nc = ncart(C.l), nd = ncart(D.l);
    double blk[36], tmp[6], Gc[3][9], Gd[3][9], R[kRD * kRD * kRD];
    for (int i = 0; i < 36; i++) blk[i] = 0.0;
    for (int qc = 0; qc < C.nPrim; qc++)
        for (int qd = 0; qd < D.nPrim; qd++) {
            const double g = exps[C.prim + qc], d = exps[D.prim + qd];
            const double alpha = g * d / (g + d);
            const double pref = 2.0 * pow(M_PI, 2.5) / (g * d * sqrt(g + d)) * coefs[C.prim + qc] * coefs[D.prim + qd];
            for (int dim = 0; dim < 3; dim++) { hermite_single(C.l, g, Gc[dim]); hermite_single(D.l, d, Gd[dim]); }
            hermite_r(C.l + D.l, alpha, C.x - D.x, C.y - D.y, C.z - D.z, boysTable, R);
            for (int ic = 0; ic < nc; ic++) {
                int xc, yc, zc;
                cart_powers(C.l, ic, xc, yc, zc);
                for (int id = 0; id < nd; id++) {
                    int xd, yd, zd;
                    cart_powers(D.l, id, xd, yd, zd);
                    double acc = 0.0;
                    for (int t1 = 0; t1 <= xc; t1++)
                        for (int u1 = 0; u1 <= yc; u1++)
                            for (int v1 = 0; v1 <= zc; v1++) {
                                const double e1 = Gc[0][xc * 3 + t1] * Gc[1][yc * 3 + u1] * Gc[2][zc * 3 + v1];
                                if (e1 == 0.0) continue;
                                for (int t2 = 0; t2 <= xd; t2++)
                                    for (int u2 = 0; u2 <= yd; u2++)
                                        for (int v2 = 0; v2 <= zd; v2++) {
                                            const double e2 = Gd[0][xd * 3 + t2] * Gd[1][yd * 3 + u2] * Gd[2][zd * 3 + v2];
                                            const double sg = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
                                            acc += e1 * e2 * sg * R[ridx(t1 + t2, u1 + u2, v1 + v2)];
                                        }
                            }
                    blk[ic * nd + id] += pref * cart_factor(C.l, ic) * cart_factor(D.l, id) * acc;
                }
            }
        }
    int mc = nc, md = nd;
    if (spherical) {
        to_spherical(C.l, c2s, blk, 1, nd, tmp);
        to_spherical(D.l, c2s, blk, 2 * C.l + 1, 1, tmp);
        mc = 2 * C.l + 1;
        md = 2 * D.l + 1;
    }
    for (int i = 0; i < mc; i++)
        for (int j = 0; j < md; j++) {
            const int p = C.f0 + i, q = D.f0 + j;
            if (p < q) continue;
            M[packed_index(p, q)] = blk[i * md + j];
        }
    // The fit functions' integrals (int chi_P dr) into row nFit, computed by the diagonal pair's thread.
    if (pairs[k].x == pairs[k].y) {
        double ov[6];
        for (int ic = 0; ic < nc; ic++) {
            int x, y, z;
            cart_powers(C.l, ic, x, y, z);
            double acc = 0.0;
            for (int q = 0; q < C.nPrim; q++) {
                const double g = exps[C.prim + q];
                double G[9];
                hermite_single(C.l, g, G);
                acc += coefs[C.prim + q] * pow(M_PI / g, 1.5) * G[x * 3] * G[y * 3] * G[z * 3];
            }
            ov[ic] = acc * cart_factor(C.l, ic);
        }
        if (spherical) {
            to_spherical(C.l, c2s, ov, 1, 1, tmp);
            mc = 2 * C.l + 1;
        }
        for (int i = 0; i < mc; i++) M[packed_index(nFit, C.f0 + i)] = ov[i];
    }
}

// (mu nu|P) of item k = (significant orbital shell pair k / nFitShells, fit shell k % nFitShells) into its rows of B
// ([row][fit function]); magnitudes <= kUnderFlow are stored as zero.
QS_HD void fit_integrals_item(long long k, int nFitShells, const PairEntry *__restrict__ pairs,
                              const QShell *__restrict__ oShells, const double *__restrict__ oExps,
                              const double *__restrict__ oCoefs, int oSpherical, const QShell *__restrict__ fShells,
                              const double *__restrict__ fExps, const double *__restrict__ fCoefs, int fSpherical, int nFit,
                              const double *__restrict__ boysTable, const double *__restrict__ c2s,
                              double *__restrict__ B) {
    const PairEntry pe = pairs[k / nFitShells];
    const QShell A = oShells[pe.a], Bs = oShells[pe.b], C = fShells[k % nFitShells];
    const int na = ncart(A.l), nb = ncart(Bs.l), nc = ncart(C.l);
    double blk[216], tmp[6];
    for (int i = 0; i < 216; i++) blk[i] = 0.0;
    double Ex[(kMaxL + 1) * kEJ * kET], Ey[(kMaxL + 1) * kEJ * kET], Ez[(kMaxL + 1) * kEJ * kET];
    double G[9], R[kRD * kRD * kRD];
    // Cartesian powers of the three shells' components; their normalisation factors multiply the finished block.
    int pa[6][3], pb[6][3], pc[6][3];
    for (int i = 0; i < na; i++) cart_powers(A.l, i, pa[i][0], pa[i][1], pa[i][2]);
    for (int i = 0; i < nb; i++) cart_powers(Bs.l, i, pb[i][0], pb[i][1], pb[i][2]);
    for (int i = 0; i < nc; i++) cart_powers(C.l, i, pc[i][0], pc[i][1], pc[i][2]);
    const double ABx = A.x - Bs.x, ABy = A.y - Bs.y, ABz = A.z - Bs.z;
    for (int qa = 0; qa < A.nPrim; qa++)
        for (int qb = 0; qb < Bs.nPrim; qb++) {
            const double a = oExps[A.prim + qa], b = oExps[Bs.prim + qb], p = a + b, mu = a * b / p;
            const double cab = oCoefs[A.prim + qa] * oCoefs[Bs.prim + qb];
            const double Px = (a * A.x + b * Bs.x) / p, Py = (a * A.y + b * Bs.y) / p, Pz = (a * A.z + b * Bs.z) / p;
            hermite_1d(A.l, Bs.l, p, Px - A.x, Px - Bs.x, exp(-mu * ABx * ABx), Ex);
            hermite_1d(A.l, Bs.l, p, Py - A.y, Py - Bs.y, exp(-mu * ABy * ABy), Ey);
            hermite_1d(A.l, Bs.l, p, Pz - A.z, Pz - Bs.z, exp(-mu * ABz * ABz), Ez);
            for (int qc = 0; qc < C.nPrim; qc++) {
                const double g = fExps[C.prim + qc];
                const double alpha = p * g / (p + g);
                // 2 pi^(5/2) / (p g sqrt(p + g)); pi^(5/2) = 17.493418327624862.
                const double pref = 2.0 * 17.493418327624862 / (p * g * sqrt(p + g)) * cab * fCoefs[C.prim + qc];
                hermite_single(C.l, g, G);   // the same in x, y and z
                hermite_r(A.l + Bs.l + C.l, alpha, Px - C.x, Py - C.y, Pz - C.z, boysTable, R);
                for (int ic = 0; ic < nc; ic++) {
                    const int xc = pc[ic][0], yc = pc[ic][1], zc = pc[ic][2];
                    // W_tuv = sum_t'u'v' (-1)^(t'+u'+v') E^c_t'u'v' R_t+t',u+u',v+v' for t + u + v <= la + lb.
                    double W[5 * 5 * 5];
                    const int Lab = A.l + Bs.l;
                    for (int t = 0; t <= Lab; t++)
                        for (int u = 0; u <= Lab - t; u++)
                            for (int v = 0; v <= Lab - t - u; v++) {
                                double acc = 0.0;
                                for (int t2 = 0; t2 <= xc; t2++)
                                    for (int u2 = 0; u2 <= yc; u2++)
                                        for (int v2 = 0; v2 <= zc; v2++) {
                                            const double e = G[xc * 3 + t2] * G[yc * 3 + u2] * G[zc * 3 + v2];
                                            const double sg = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
                                            acc += sg * e * R[ridx(t + t2, u + u2, v + v2)];
                                        }
                                W[(t * 5 + u) * 5 + v] = acc;
                            }
                    for (int ia = 0; ia < na; ia++) {
                        const int xa = pa[ia][0], ya = pa[ia][1], za = pa[ia][2];
                        for (int ib = 0; ib < nb; ib++) {
                            const int xb = pb[ib][0], yb = pb[ib][1], zb = pb[ib][2];
                            double acc = 0.0;
                            for (int t = 0; t <= xa + xb; t++)
                                for (int u = 0; u <= ya + yb; u++)
                                    for (int v = 0; v <= za + zb; v++)
                                        acc += Ex[(xa * kEJ + xb) * kET + t] * Ey[(ya * kEJ + yb) * kET + u] *
                                               Ez[(za * kEJ + zb) * kET + v] * W[(t * 5 + u) * 5 + v];
                            blk[(ia * nb + ib) * nc + ic] += pref * acc;
                        }
                    }
                }
            }
        }
    for (int ia = 0; ia < na; ia++)
        for (int ib = 0; ib < nb; ib++) {
            const double fab = cart_factor(A.l, ia) * cart_factor(Bs.l, ib);
            for (int ic = 0; ic < nc; ic++) blk[(ia * nb + ib) * nc + ic] *= fab * cart_factor(C.l, ic);
        }
    int ma = na, mb = nb, mc = nc;
    if (oSpherical) {
        to_spherical(A.l, c2s, blk, 1, nb * nc, tmp);
        to_spherical(Bs.l, c2s, blk, 2 * A.l + 1, nc, tmp);
        ma = 2 * A.l + 1;
        mb = 2 * Bs.l + 1;
    }
    if (fSpherical) {
        to_spherical(C.l, c2s, blk, ma * mb, 1, tmp);
        mc = 2 * C.l + 1;
    }
    for (int i = 0; i < ma; i++)
        for (int j = 0; j < mb; j++) {
            double *row = B + (pe.row + i * mb + j) * nFit + C.f0;
            for (int c = 0; c < mc; c++) {
                const double x = blk[(i * mb + j) * mc + c];
                row[c] = fabs(x) > kUnderFlow ? x : 0.0;
            }
        }
}

}  // namespace qs
