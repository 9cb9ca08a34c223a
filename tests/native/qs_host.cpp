// Host build of the SCF wrapper's integral work items (pdynamo3_amd/csrc/qs_kernels.h, qs_shells.h), for checking the
// integral code against the compiled reference on the CPU.  Test infrastructure only: the product runs the same items on
// the GPU (qs_integrals.hip); nothing in the package loads this library.
#include <cstring>
#include <vector>

#include "../../pdynamo3_amd/csrc/qs_shells.h"

using namespace qs;

namespace {

struct Tables {
    std::vector<double> boys;
    double c2s[90];
    Tables() { boys_table(boys); c2s_table(c2s); }
};
const Tables &tables() { static Tables t; return t; }

int make(QBasis &b, int nAtoms, const int *Z, const double *xyz, int nSpec, const double *spec, int spherical) {
    b.nAtoms = nAtoms;
    b.spherical = spherical ? 1 : 0;
    b.Z.assign(Z, Z + nAtoms);
    b.xyz.assign(xyz, xyz + 3 * nAtoms);
    return build_shells(&b, nSpec, spec).empty() ? 0 : 3;
}

}  // namespace

extern "C" {

int qsh_boys(int m, double T, double *F) {
    boys(m, T, tables().boys.data(), F);
    return 0;
}

int qsh_one_electron(int nAtoms, const int *Z, const double *xyz, int nSpec, const double *spec, int spherical, double *S,
                     double *T, double *V) {
    QBasis b;
    if (make(b, nAtoms, Z, xyz, nSpec, spec, spherical)) return 3;
    std::vector<QPair> pairs = shell_pairs(&b, false);
    std::vector<double> atoms(4 * nAtoms);
    for (int a = 0; a < nAtoms; a++) {
        for (int c = 0; c < 3; c++) atoms[4 * a + c] = xyz[3 * a + c];
        atoms[4 * a + 3] = Z[a];
    }
    for (int k = 0; k < static_cast<int>(pairs.size()); k++)
        one_electron_pair(k, pairs.data(), b.shells.data(), b.exps.data(), b.coefs.data(), nAtoms, atoms.data(), b.spherical,
                          tables().boys.data(), tables().c2s, S, T, V);
    return 0;
}

int qsh_fit_matrix(int nAtoms, const int *Z, const double *xyz, int nSpec, const double *spec, int spherical, double *M) {
    QBasis b;
    if (make(b, nAtoms, Z, xyz, nSpec, spec, spherical)) return 3;
    std::vector<QPair> pairs = shell_pairs(&b, false);
    for (int k = 0; k < static_cast<int>(pairs.size()); k++)
        fit_matrix_pair(k, pairs.data(), b.shells.data(), b.exps.data(), b.coefs.data(), b.nFunctions, b.spherical,
                        tables().boys.data(), tables().c2s, M);
    return 0;
}

// Dense (mu nu|P) [packed orbital pair][fit function] over the significant shell pairs (others left at zero).
int qsh_fit_integrals(int nAtoms, const int *Z, const double *xyz, int nOrb, const double *orbSpec, int orbSpherical,
                      int nFitSpec, const double *fitSpec, int fitSpherical, double *dense) {
    QBasis o, f;
    if (make(o, nAtoms, Z, xyz, nOrb, orbSpec, orbSpherical) || make(f, nAtoms, Z, xyz, nFitSpec, fitSpec, fitSpherical))
        return 3;
    std::vector<QPair> sp = shell_pairs(&o, true);
    std::vector<PairEntry> pairs(sp.size());
    long long rows = 0;
    for (size_t k = 0; k < sp.size(); k++) {
        pairs[k] = PairEntry{sp[k].x, sp[k].y, rows};
        rows += static_cast<long long>(o.shells[sp[k].x].nf) * o.shells[sp[k].y].nf;
    }
    const int nFit = f.nFunctions, nShellsF = static_cast<int>(f.shells.size());
    std::vector<double> B(static_cast<size_t>(rows) * nFit, 0.0);
    for (long long k = 0; k < static_cast<long long>(sp.size()) * nShellsF; k++)
        fit_integrals_item(k, nShellsF, pairs.data(), o.shells.data(), o.exps.data(), o.coefs.data(), o.spherical,
                           f.shells.data(), f.exps.data(), f.coefs.data(), f.spherical, nFit, tables().boys.data(),
                           tables().c2s, B.data());
    for (size_t k = 0; k < sp.size(); k++) {
        const QShell &A = o.shells[sp[k].x], &Bs = o.shells[sp[k].y];
        for (int i = 0; i < A.nf; i++)
            for (int j = 0; j < Bs.nf; j++) {
                const int mu = A.f0 + i, nu = Bs.f0 + j;
                if (mu < nu) continue;
                std::memcpy(dense + packed_index(mu, nu) * nFit, B.data() + (pairs[k].row + i * Bs.nf + j) * nFit,
                            nFit * sizeof(double));
            }
    }
    return 0;
}

}  // extern "C"
