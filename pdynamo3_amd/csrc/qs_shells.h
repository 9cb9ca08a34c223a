// Host-side set-up of the SCF wrapper's integrals: shells from the element specification, significant shell pairs, the
// Boys table and the c2s table.  Shared by the library (qs_integrals.hip) and the host build that checks the integral
// code on the CPU.
#pragma once

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "qs_kernels.h"
#include "xq_c2s.h"

namespace qs {

constexpr double kPairScreen = 1.0e-15;   // shell pairs whose largest primitive prefactor is below this are dropped

struct QBasis {
    int nAtoms = 0, nFunctions = 0, spherical = 1;
    std::vector<int> Z;
    std::vector<double> xyz;
    std::vector<QShell> shells;
    std::vector<double> exps, coefs;
};

// F_n(T) by its series e^-T sum_i (2T)^i / ((2n+1)(2n+3)...(2n+2i+1)), all terms positive, in long double.
inline double boys_series(int n, double T) {
    long double term = 1.0L / (2 * n + 1), sum = term;
    for (int i = 1; i < 2000; i++) {
        term *= 2.0L * T / (2 * n + 2 * i + 1);
        sum += term;
        if (term < sum * 1e-22L) break;
    }
    return static_cast<double>(sum * expl(-static_cast<long double>(T)));
}

// The Boys table F[k][n] = F_n(k kBoysStep), n < kBoysTable, and the c2s transformations of l = 0..2 as [l][6][5].
inline void boys_table(std::vector<double> &t) {
    t.assign(static_cast<size_t>(kBoysPoints) * kBoysTable, 0.0);
    for (int k = 0; k < kBoysPoints; k++)
        for (int n = 0; n < kBoysTable; n++) t[static_cast<size_t>(k) * kBoysTable + n] = boys_series(n, k * kBoysStep);
}

inline void c2s_table(double *t) {
    for (int k = 0; k < 90; k++) t[k] = 0.0;
    for (int l = 0; l <= 2; l++) {
        std::vector<double> m(static_cast<size_t>(ncart(l)) * (2 * l + 1), 0.0);
        xq::c2s_block(l, m.data(), 2 * l + 1, 0, 0);
        for (int c = 0; c < ncart(l); c++)
            for (int sI = 0; sI < 2 * l + 1; sI++) t[(l * 6 + c) * 5 + sI] = m[static_cast<size_t>(c) * (2 * l + 1) + sI];
    }
}

// Shells of every atom from the element specification (the format of xq_basis_set), one QShell per angular momentum.
inline std::string build_shells(QBasis *b, int nSpec, const double *spec) {
    struct Element { int z, start, nShells; };
    std::vector<Element> elements;
    for (int p = 0; p < nSpec;) {
        if (p + 2 > nSpec) return "truncated basis specification";
        Element e{static_cast<int>(spec[p]), p + 2, static_cast<int>(spec[p + 1])};
        p += 2;
        for (int s = 0; s < e.nShells; s++) {
            if (p + 3 > nSpec) return "truncated basis specification";
            const int lLow = static_cast<int>(spec[p]), lHigh = static_cast<int>(spec[p + 1]), np = static_cast<int>(spec[p + 2]);
            if (lLow < 0 || lHigh < lLow || lHigh > kMaxL || np < 1 || np > 64)
                return "unsupported shell for the SCF integrals (angular momenta 0 to 2)";
            p += 3 + np * (1 + lHigh - lLow + 1);
        }
        if (p > nSpec) return "truncated basis specification";
        elements.push_back(e);
    }
    int f = 0;
    for (int a = 0; a < b->nAtoms; a++) {
        const Element *el = nullptr;
        for (const Element &e : elements) if (e.z == b->Z[a]) el = &e;
        if (el == nullptr) return "no basis for an atom's element";
        int p = el->start;
        for (int s = 0; s < el->nShells; s++) {
            const int lLow = static_cast<int>(spec[p]), lHigh = static_cast<int>(spec[p + 1]), np = static_cast<int>(spec[p + 2]);
            p += 3;
            const int rowLen = 1 + lHigh - lLow + 1;
            for (int l = lLow; l <= lHigh; l++) {
                QShell sh{};
                sh.x = b->xyz[3 * a]; sh.y = b->xyz[3 * a + 1]; sh.z = b->xyz[3 * a + 2];
                sh.l = l;
                sh.nPrim = np;
                sh.prim = static_cast<int>(b->exps.size());
                sh.f0 = f;
                sh.nf = b->spherical ? 2 * l + 1 : ncart(l);
                sh.atom = a;
                for (int q = 0; q < np; q++) {
                    const double alpha = spec[p + q * rowLen];
                    const double cin = spec[p + q * rowLen + 1 + (l - lLow)];
                    // Unnormalise (GaussianBasis.c:728-751).
                    const double e2 = 2.0 * alpha;
                    double fac = std::pow(M_PI, 1.5) / (e2 * std::sqrt(e2));
                    for (int i = 1; i <= l; i++) fac *= static_cast<double>(2 * i - 1) / (2.0 * e2);
                    b->exps.push_back(alpha);
                    b->coefs.push_back(cin / std::sqrt(fac));
                }
                f += sh.nf;
                b->shells.push_back(sh);
            }
            p += np * rowLen;
        }
    }
    b->nFunctions = f;
    return "";
}


// Shell pairs A >= B whose largest primitive prefactor |c_a c_b| exp(-mu R^2) (pi/p)^(3/2) reaches the screen.
inline std::vector<QPair> shell_pairs(const QBasis *b, bool screen) {
    std::vector<QPair> out;
    for (int A = 0; A < static_cast<int>(b->shells.size()); A++)
        for (int B = 0; B <= A; B++) {
            const QShell &sa = b->shells[A], &sb = b->shells[B];
            const double dx = sa.x - sb.x, dy = sa.y - sb.y, dz = sa.z - sb.z, r2 = dx * dx + dy * dy + dz * dz;
            double m = 0.0;
            for (int qa = 0; qa < sa.nPrim; qa++)
                for (int qb = 0; qb < sb.nPrim; qb++) {
                    const double a = b->exps[sa.prim + qa], c = b->exps[sb.prim + qb], p = a + c;
                    m = std::max(m, std::fabs(b->coefs[sa.prim + qa] * b->coefs[sb.prim + qb]) * std::exp(-a * c / p * r2) *
                                        std::pow(M_PI / p, 1.5) * std::max(1.0, 1.0 / p));
                }
            if (!screen || m > kPairScreen) out.push_back(QPair{A, B});
        }
    return out;
}

}  // namespace qs
