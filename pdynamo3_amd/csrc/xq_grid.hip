// Grid construction on the device: Euler-Maclaurin radial shells x pruned Lebedev-Laikov spheres around every atom, Becke
// partition weights in FP64, the |w| > 1e-30 cull and per-atom blocks of at most 128 points.
//
// Follows DFTGrid_Construct (pMolecule/QCModel/__extensions__/csource/DFTGrid.c:105-263) with DFTGrid_Atom_Parameters
// (:474-523), DFTGrid_Radial_Points (:562-575), the Lebedev size <-> degree maps (pScientific/Quadrature/__extensions__/
// csource/Lebedev.c:62-193), the octahedral orbit expansion order (Lebedev.c:1565-1719) and the Becke weight
// (DFTGridWeights.c:43-80, 220-258).
//
// Device layout: every radial shell (atom, radius) is one work-group; each thread computes the Becke weight of one angular
// point in FP64.  Candidate points of one atom are contiguous and ordered (radial shell, angular index) exactly as the
// reference orders them, so the order-preserving cull reproduces the reference's point and block order.
//
// Becke weight: P_A / sum_i P_i with P_i = prod_{j != i} s(nu_ij).  P_A (the point's own atom) is formed in full.  For
// i != A the product is formed factor by factor, starting with the factors from atom A and from the atom nearest the
// point, then the others nearest atom to A first (the factors that can make it vanish come early), and stops once it falls below 2^-64 max(|P_A|, |P_near|), P_near
// the product of the atom nearest the point: all later factors lie in [0, 1], so the dropped part of sum_i P_i is below
// nAtoms * 2^-64 of max(P_A, P_near) <= the sum.  This turns the reference's O(nAtoms^2) pair loop per point into
// O(nAtoms) for points that sit in their own atom's cell or in the cell of the atom nearest them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <hip/hip_runtime.h>

#include "lebedev_tables.h"
#include "xq_internal.h"

// No fused multiply-add in this file: the cull |w| > 1e-30 depends on Becke cell factors that cancel to exactly zero near
// nuclei, so point coordinates, distances and the cell polynomial are rounded operation by operation as in the reference.
#pragma clang fp contract(off)

namespace xq {

namespace {

// Bragg radii in Angstrom by atomic number (DFTGrid.c:29-40).
const double kBragg[120] = {
    0.75, 0.35, 0.35, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 0.50, 1.80, 1.50, 1.25, 1.10, 1.00, 1.00, 1.00, 1.00, 2.20,
    1.80, 1.60, 1.40, 1.35, 1.40, 1.40, 1.40, 1.35, 1.35, 1.35, 1.35, 1.30, 1.25, 1.15, 1.15, 1.15, 1.15, 2.35, 2.00, 1.80,
    1.55, 1.45, 1.45, 1.35, 1.30, 1.35, 1.40, 1.60, 1.55, 1.55, 1.45, 1.45, 1.40, 1.40, 1.40, 2.60, 2.15, 1.95, 1.85, 1.85,
    1.85, 1.85, 1.85, 1.85, 1.80, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.55, 1.45, 1.35, 1.35, 1.30, 1.35, 1.35, 1.35,
    1.50, 1.90, 1.80, 1.60, 1.90, 1.90, 1.90, 2.60, 2.15, 1.95, 1.80, 1.80, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75,
    1.75, 1.75, 1.75, 1.75, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55};
const double kAngstromToBohr = 1.0e-10 / 5.291772083e-11;   // Units.h:35, Constants.h:15
const double kBFTol[5] = {1.0e-8, 1.0e-9, 1.0e-10, 1.0e-12, 1.0e-15};    // DFTGrid.c:43
const double kRhoTol[5] = {1.0e-13, 1.0e-14, 1.0e-15, 1.0e-17, 1.0e-20}; // DFTGrid.c:46
constexpr int kMinL = 9;                 // DFTGrid.c:101
constexpr double kRadialCutFactor = 0.2; // DFTGrid.c:102
constexpr double kWeightTol = 1.0e-30;   // DFTGrid.c:103

void atom_parameters(int accuracy, int z, int *nR, int *lMax, double *rMax) {
    switch (accuracy) {
        case 0: *nR = z < 11 ? 21 : z < 19 ? 42 : z < 37 ? 75 : 84; *rMax = 20.0; *lMax = 23; break;
        case 1: *nR = z < 11 ? 35 : z < 19 ? 70 : z < 37 ? 95 : 104; *rMax = 25.0; *lMax = 35; break;
        case 2: *nR = z < 11 ? 49 : z < 19 ? 88 : z < 37 ? 112 : 123; *rMax = 30.0; *lMax = z < 11 ? 35 : 41; break;
        case 3:
            *nR = z < 11 ? 70 : z < 19 ? 123 : z < 37 ? 130 : 155; *rMax = 35.0;
            *lMax = z < 11 ? 41 : z < 19 ? 47 : z < 89 ? 53 : 59;
            break;
        default: *nR = z < 11 ? 100 : z < 19 ? 125 : z < 37 ? 160 : 205; *rMax = 35.0; *lMax = 65; break;
    }
}

double bragg_radius(int z) {
    if (z < 0) z = 0;
    if (z >= 120) z = 119;
    return kAngstromToBohr * kBragg[z];
}

// Number of points of the smallest rule integrating degree l (Lebedev.c:134-193).
int lebedev_size_for_l(int l) {
    static const int small[16] = {6, 6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350};
    static const int large[22] = {0, 0, 0, 0, 0, 434, 590, 770, 974, 1202, 1454, 1730, 2030, 2354, 2702, 3074, 3470,
                                  3890, 4334, 4802, 5294, 5810};
    if (l < 3) return 6;
    if (l <= 31) return small[l / 2];
    if (l <= 131) return large[l / 6];
    return 0;
}

// Degree of the smallest rule with at least n points (Lebedev.c:62-98).
int lebedev_l_for_size(int n) {
    static const int ls[XQ_LEBEDEV_NRULES] = {3, 5, 7, 9, 11, 13, 15, 17, 19, 21, 23, 25, 27, 29, 31, 35,
                                              41, 47, 53, 59, 65, 71, 77, 83, 89, 95, 101, 107, 113, 119, 125, 131};
    for (int k = 0; k < XQ_LEBEDEV_NRULES; k++)
        if (n <= xq_lebedev_npoints[k]) return ls[k];
    return 0;
}

// Expand rule k into unit vectors and weights, in the orbit order of Lebedev.c:1565-1719.  Each template row names, for x, y
// and z, the generator slot (1 = a, 2 = b, 3 = c, 0 = zero) and its sign.
int lebedev_expand(int k, double *out /* x, y, z, w per point */) {
    static const signed char T1[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    static const signed char T2[12][3] = {{0, 1, 1}, {0, 1, -1}, {0, -1, 1}, {0, -1, -1}, {1, 0, 1}, {1, 0, -1},
                                          {-1, 0, 1}, {-1, 0, -1}, {1, 1, 0}, {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}};
    static const signed char T3[8][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
                                         {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
    static const signed char T4[24][3] = {{1, 1, 2}, {1, 1, -2}, {1, -1, 2}, {1, -1, -2}, {-1, 1, 2}, {-1, 1, -2},
                                          {-1, -1, 2}, {-1, -1, -2}, {1, 2, 1}, {1, -2, 1}, {1, 2, -1}, {1, -2, -1},
                                          {-1, 2, 1}, {-1, -2, 1}, {-1, 2, -1}, {-1, -2, -1}, {2, 1, 1}, {-2, 1, 1},
                                          {2, 1, -1}, {-2, 1, -1}, {2, -1, 1}, {-2, -1, 1}, {2, -1, -1}, {-2, -1, -1}};
    static const signed char T5[24][3] = {{1, 2, 0}, {1, -2, 0}, {-1, 2, 0}, {-1, -2, 0}, {2, 1, 0}, {2, -1, 0},
                                          {-2, 1, 0}, {-2, -1, 0}, {1, 0, 2}, {1, 0, -2}, {-1, 0, 2}, {-1, 0, -2},
                                          {2, 0, 1}, {2, 0, -1}, {-2, 0, 1}, {-2, 0, -1}, {0, 1, 2}, {0, 1, -2},
                                          {0, -1, 2}, {0, -1, -2}, {0, 2, 1}, {0, 2, -1}, {0, -2, 1}, {0, -2, -1}};
    static const signed char T6[48][3] = {
        {1, 2, 3}, {1, 2, -3}, {1, -2, 3}, {1, -2, -3}, {-1, 2, 3}, {-1, 2, -3}, {-1, -2, 3}, {-1, -2, -3},
        {1, 3, 2}, {1, 3, -2}, {1, -3, 2}, {1, -3, -2}, {-1, 3, 2}, {-1, 3, -2}, {-1, -3, 2}, {-1, -3, -2},
        {2, 1, 3}, {2, 1, -3}, {2, -1, 3}, {2, -1, -3}, {-2, 1, 3}, {-2, 1, -3}, {-2, -1, 3}, {-2, -1, -3},
        {2, 3, 1}, {2, 3, -1}, {2, -3, 1}, {2, -3, -1}, {-2, 3, 1}, {-2, 3, -1}, {-2, -3, 1}, {-2, -3, -1},
        {3, 1, 2}, {3, 1, -2}, {3, -1, 2}, {3, -1, -2}, {-3, 1, 2}, {-3, 1, -2}, {-3, -1, 2}, {-3, -1, -2},
        {3, 2, 1}, {3, 2, -1}, {3, -2, 1}, {3, -2, -1}, {-3, 2, 1}, {-3, 2, -1}, {-3, -2, 1}, {-3, -2, -1}};
    int m = 0;
    for (int o = xq_lebedev_first[k]; o < xq_lebedev_first[k + 1]; o++) {
        const xq_lebedev_orbit &orb = xq_lebedev_orbits[o];
        double g[4] = {0.0, 0.0, 0.0, 0.0};
        const signed char(*T)[3] = nullptr;
        int count = 0;
        switch (orb.type) {
            case 1: g[1] = 1.0; T = T1; count = 6; break;
            case 2: g[1] = 1.0 / std::sqrt(2.0); T = T2; count = 12; break;
            case 3: g[1] = 1.0 / std::sqrt(3.0); T = T3; count = 8; break;
            case 4: g[1] = orb.a; g[2] = std::sqrt(1.0 - 2.0 * orb.a * orb.a); T = T4; count = 24; break;
            case 5: g[1] = orb.a; g[2] = std::sqrt(1.0 - orb.a * orb.a); T = T5; count = 24; break;
            case 6: g[1] = orb.a; g[2] = orb.b; g[3] = std::sqrt(1.0 - orb.a * orb.a - orb.b * orb.b); T = T6; count = 48; break;
        }
        for (int n = 0; n < count; n++, m++) {
            for (int c = 0; c < 3; c++) {
                int s = T[n][c];
                out[4 * m + c] = s == 0 ? 0.0 : (s > 0 ? g[s] : -g[-s]);
            }
            out[4 * m + 3] = orb.v;
        }
    }
    return m;
}

// Becke cell-function constants (DFTGridWeights.c:21-35).
__constant__ double c_beckeC[11] = {1.000000000000000e+00, -3.333333333333333e+00, 9.000000000000000e+00,
                                    -17.142857142857142e+00, 23.333333333333332e+00, -22.909090909090910e+00,
                                    16.153846153846153e+00, -8.000000000000000e+00, 2.647058823529412e+00,
                                    -0.526315789473684e+00, 0.047619047619048e+00};
constexpr double kBeckeA = 1.850069046020527e+00;

struct RadialShell {
    int atom;
    int rule;        // first unit vector of the Lebedev rule in the device rule array
    int nAng;
    int out;         // first candidate point of the shell
    double r;
    double wr;       // 4 pi w_r
};

__device__ __forceinline__ double becke_poly(double nu) {
    double nu2 = nu * nu, term = nu, sum = 0.0;
#pragma unroll
    for (int t = 0; t <= 10; t++) {
        sum += c_beckeC[t] * term;
        term *= nu2;
    }
    return sum;
}

// Factor of P_k from the pair (k, m), k != m, given the pair's 1/R and size adjustment (stored once for the ordered pair
// (max, min)): s(nu) for k > m, 1 - s(nu) for k < m, with mu = (r_max - r_min) / R.
__device__ __forceinline__ double becke_factor_of(bool kHigh, double rk, double rm, double invR, double a) {
    const double rhi = kHigh ? rk : rm, rlo = kHigh ? rm : rk;
    const double mu = (rhi - rlo) * invR;
    const double nu = mu + a * (1.0 - mu * mu);
    const double s = kBeckeA * becke_poly(nu);
    return kHigh ? (0.5 - s) : (0.5 + s);
}

__device__ __forceinline__ int pair_index(int k, int m) {
    const int hi = k > m ? k : m, lo = k > m ? m : k;
    return (hi * (hi - 1)) / 2 + lo;
}

// Candidate points of one radial shell per work-group (thread = angular point) and their Becke weights
// w = 4 pi w_r w_ang P_A / sum_i P_i (DFTGrid.c:158-211, DFTGridWeights.c:220-258).  P_A and the product of the atom
// nearest the point come first; when even P_A / (P_A + P_near) leaves the weight below half the cull tolerance (or
// P_A = 0), the point is culled (as the reference culls it) without the other products.  The atom coordinates, the pair
// data of the shell's atom A (row A of 1/R and a) and the atoms in order of distance from A are staged in LDS.  P_A is
// computed in full; each other P_i takes its factors with A and the nearest atom, then the others nearest A first, and
// stops once it falls below 2^-64 of max(|P_A|, |P_near|) (both <= the sum), where it no longer changes the sum beyond
// rounding.  Each factor is the reference's cell function of the same mu and nu, rounded as the reference rounds it (so factors that vanish there vanish here); the products are taken in another
// order, so weights agree with the reference's to rounding and the same points pass the cull.
__global__ void __launch_bounds__(256) candidate_points_kernel(const RadialShell *__restrict__ shells,
                                                               const double *__restrict__ rules,
                                                               const double *__restrict__ xyz, int nAtoms,
                                                               const double *__restrict__ invR,
                                                               const double *__restrict__ aij,
                                                               const int *__restrict__ order,
                                                               double *__restrict__ cx, double *__restrict__ cy,
                                                               double *__restrict__ cz, double *__restrict__ cw,
                                                               int *__restrict__ keep) {
    extern __shared__ double lds[];
    double *sx = lds, *sy = lds + nAtoms, *sz = lds + 2 * nAtoms, *sInv = lds + 3 * nAtoms, *sA = lds + 4 * nAtoms;
    int *sOrd = reinterpret_cast<int *>(lds + 5 * nAtoms);
    const RadialShell sh = shells[blockIdx.x];
    const int A = sh.atom;
    for (int j = threadIdx.x; j < nAtoms; j += blockDim.x) {
        sOrd[j] = order[static_cast<size_t>(A) * nAtoms + j];
        sx[j] = xyz[3 * j];
        sy[j] = xyz[3 * j + 1];
        sz[j] = xyz[3 * j + 2];
        const int ij = j == A ? 0 : pair_index(A, j);
        sInv[j] = j == A ? 0.0 : invR[ij];
        sA[j] = j == A ? 0.0 : aij[ij];
    }
    __syncthreads();
    auto dist = [&](int i, double px, double py, double pz) {
        const double dx = sx[i] - px, dy = sy[i] - py, dz = sz[i] - pz;
        return sqrt(dx * dx + dy * dy + dz * dz);
    };
    const double ax = sx[A], ay = sy[A], az = sz[A];
    for (int k = threadIdx.x; k < sh.nAng; k += blockDim.x) {
        const double *u = rules + 4 * (sh.rule + k);
        const double px = u[0] * sh.r + ax, py = u[1] * sh.r + ay, pz = u[2] * sh.r + az;
        const double rA = dist(A, px, py, pz);
        // P_A in full.
        double PA = 1.0;
        for (int j = 0; j < nAtoms; j++) {
            if (j == A) continue;
            PA *= becke_factor_of(A > j, rA, dist(j, px, py, pz), sInv[j], sA[j]);
        }
        // Products below 2^-64 of max(|P_A|, |P_near|) <= sum_i P_i no longer change the sum beyond rounding.
        double floor = fabs(PA) * 0x1p-64;
        const double wru = sh.wr * u[3];
        const int o = sh.out + k;
        cx[o] = px;
        cy[o] = py;
        cz[o] = pz;
        // The atom nearest the point (other than A) gets its full product first: sum_i P_i >= P_A + P_near, so a point
        // whose weight cannot exceed half the cull tolerance even then is culled without the other products.
        int near = -1;
        double dNear = 0.0;
        for (int j = 0; j < nAtoms; j++) {
            if (j == A) continue;
            const double dx = sx[j] - px, dy = sy[j] - py, dz = sz[j] - pz, d2 = dx * dx + dy * dy + dz * dz;
            if (near < 0 || d2 < dNear) { near = j; dNear = d2; }
        }
        double sum = PA, rn = 0.0;
        if (near >= 0) {
            rn = dist(near, px, py, pz);
            double Pn = becke_factor_of(near > A, rn, rA, sInv[near], sA[near]);
            for (int j = 0; j < nAtoms; j++) {
                if (j == near || j == A) continue;
                const int ij = pair_index(near, j);
                Pn *= becke_factor_of(near > j, rn, dist(j, px, py, pz), invR[ij], aij[ij]);
            }
            sum += Pn;
            floor = fmax(floor, fabs(Pn) * 0x1p-64);
            if (PA == 0.0 || (sum > 0.0 && fabs(wru) * (fabs(PA) / sum) < 0.5 * kWeightTol)) {
                cw[o] = 0.0;
                keep[o] = 0;
                continue;
            }
        }
        for (int i = 0; i < nAtoms; i++) {
            if (i == A || i == near) continue;
            const double ri = dist(i, px, py, pz);
            double Pi = becke_factor_of(i > A, ri, rA, sInv[i], sA[i]);
            if (near >= 0 && fabs(Pi) >= floor) {   // then the atom nearest the point
                const int in = pair_index(i, near);
                Pi *= becke_factor_of(i > near, ri, rn, invR[in], aij[in]);
            }
            for (int jj = 0; jj < nAtoms && fabs(Pi) >= floor; jj++) {
                const int j = sOrd[jj];
                if (j == i || j == A || j == near) continue;
                const int ij = pair_index(i, j);
                Pi *= becke_factor_of(i > j, ri, dist(j, px, py, pz), invR[ij], aij[ij]);
            }
            if (fabs(Pi) >= floor) sum += Pi;
        }
        const double w = wru * (PA / sum);
        cw[o] = w;
        keep[o] = fabs(w) > kWeightTol ? 1 : 0;
    }
}

// Order-preserving compaction of the kept candidates: dst = exclusive prefix sum of keep.
__global__ void compact_points_kernel(int n, const int *__restrict__ keep, const int *__restrict__ dst,
                                      const double *__restrict__ cx, const double *__restrict__ cy,
                                      const double *__restrict__ cz, const double *__restrict__ cw,
                                      double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
                                      double *__restrict__ w) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && keep[i]) {
        int d = dst[i];
        x[d] = cx[i];
        y[d] = cy[i];
        z[d] = cz[i];
        w[d] = cw[i];
    }
}

// Exclusive prefix sum of keep in three passes over tiles of kScanTile entries: tile sums, a scan of the tile sums in
// one work-group, then each tile's own scan offset by its tile's prefix.
constexpr int kScanTile = 1024;

__device__ __forceinline__ int block_exclusive_scan_256(int v, int *sh, int *total) {
    // Work-group of 256: inclusive wave scans, then the four wave totals.
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < wv; k++) base += sh[k];
    if (total) *total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return base + x - v;
}

__global__ void __launch_bounds__(256) scan_tile_sums_kernel(int n, const int *__restrict__ in, int *__restrict__ sums) {
    __shared__ int sh[4];
    const int base = blockIdx.x * kScanTile + 4 * threadIdx.x;
    int v = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) v += base + k < n ? in[base + k] : 0;
    int total = 0;
    (void)block_exclusive_scan_256(v, sh, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(256) scan_sums_kernel(int nTiles, int *__restrict__ sums, int *__restrict__ total) {
    __shared__ int sh[4];
    int run = 0;
    for (int base = 0; base < nTiles; base += 256) {
        const int i = base + threadIdx.x;
        const int v = i < nTiles ? sums[i] : 0;
        int chunk = 0;
        const int ex = block_exclusive_scan_256(v, sh, &chunk);
        if (i < nTiles) sums[i] = run + ex;
        run += chunk;
    }
    if (threadIdx.x == 0) *total = run;
}

__global__ void __launch_bounds__(256) scan_apply_kernel(int n, const int *__restrict__ in, const int *__restrict__ sums,
                                                         int *__restrict__ out) {
    __shared__ int sh[4];
    const int base = blockIdx.x * kScanTile + 4 * threadIdx.x;
    int e[4], v = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) { e[k] = base + k < n ? in[base + k] : 0; v += e[k]; }
    int run = sums[blockIdx.x] + block_exclusive_scan_256(v, sh, nullptr);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k < n) out[base + k] = run;
        run += e[k];
    }
}

// Becke pair data (DFTGridWeights.c:43-80): 1/R_ij and the size adjustment a_ij for i > j, packed as pair_index.
void becke_pairs(const xq_grid *g, std::vector<double> &invR, std::vector<double> &aij) {
    const int nA = g->nAtoms;
    const size_t nPairs = nA > 1 ? static_cast<size_t>(nA) * (nA - 1) / 2 : 1;
    invR.assign(nPairs, 0.0);
    aij.assign(nPairs, 0.0);
    for (int i = 0, ij = 0; i < nA; i++) {
        for (int j = 0; j < i; j++, ij++) {
            double dx = g->xyz[3 * i] - g->xyz[3 * j], dy = g->xyz[3 * i + 1] - g->xyz[3 * j + 1],
                   dz = g->xyz[3 * i + 2] - g->xyz[3 * j + 2];
            double chi = bragg_radius(g->Z[i]) / bragg_radius(g->Z[j]);
            double u = (chi - 1.0) / (chi + 1.0), a = u / (u * u - 1.0);
            invR[ij] = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
            aij[ij] = a > 0.5 ? 0.5 : (a < -0.5 ? -0.5 : a);
        }
    }
}

// Largest number of atoms per lane of the weight-derivative kernel: 64 x 16 = 1024 atoms (more atoms per lane spill
// registers).
constexpr int kMaxLaneAtoms = 16;

// Screen of the weight-derivative terms.  The term of the ordered pair (k, m) is ifac = (ds_km/dmu) (A_k / s_km) fac_k, and
// moves an atom or the point by at most |ifac| 2 / R_km.  |ds/dmu| <= 1.6875 x 2 = 3.375 (|ds/dnu| peaks at nu = 0 with
// (3/2)^3 / 2, |dnu/dmu| = |1 - 2 a mu| <= 2); |A_k / s_km| <= Q_k, the product of k's factors without its smallest one;
// |fac_k| = |eXC w| / sum_i A_i for k != g.  Atoms with Q_k < 2^-80 sum_i A_i (k != g) are outside the set K: the terms of
// pairs (k, m) with k outside K and of pairs (m, k) with k outside K are each below 3.375 2^-80 |eXC w| 2 / R_km, and
// are left out (below 1e-15 hartree/bohr in all for any molecule this kernel takes); every other pair term is computed.
// sum_i A_i is taken over the atoms whose bound passes half the screen at max(A_g, A_near), A_near the product of the
// atom nearest the point (K lies among them); each atom left out of it has A_i below 2^-81 max(A_g, A_near).
constexpr double kWeightScreen = 0x1p-80;
constexpr int kNearAtoms = 24;           // atoms whose factors bound a cell product (the point's atom and its nearest)

// Derivative polynomial of the cell function's polynomial: d_t = (2 t + 1) c_t, so that d/dnu sum_t c_t nu^(2t+1) =
// sum_t d_t nu^(2t).
__constant__ double c_beckeD[11] = {1.0 * 1.000000000000000e+00, 3.0 * -3.333333333333333e+00,
                                    5.0 * 9.000000000000000e+00, 7.0 * -17.142857142857142e+00,
                                    9.0 * 23.333333333333332e+00, 11.0 * -22.909090909090910e+00,
                                    13.0 * 16.153846153846153e+00, 15.0 * -8.000000000000000e+00,
                                    17.0 * 2.647058823529412e+00, 19.0 * -0.526315789473684e+00,
                                    21.0 * 0.047619047619048e+00};

// Atom k's cell-factor product with the factor s of one partner left out, from the product of its non-zero factors (of
// either sign: a factor can round to a tiny negative value) and the count of its zero factors.  invS = 1 / s (used only
// when no factor vanished, so s != 0).
__device__ __forceinline__ double excluded_product(double prodNZ, int zeros, double s, double invS) {
    if (zeros == 0) return prodNZ * invS;
    return (zeros == 1 && s == 0.0) ? prodNZ : 0.0;
}

// XC gradient weight terms (DFTGridWeights_Derivatives, DFTGridWeights.c:98-215).  One wave per point, WPG points per
// round of a work-group; lane l owns atoms l, l + 64, ... (at most LA of them) and accumulates their gradient in
// registers.  Per point (grid atom g, weight w, energy density eXC = eps rho): A_i = prod_j s_ij; p = A_g / sum_i A_i;
// ew = eXC w / (p sum_i A_i); dA_i/dmu_ij = ds_ij prod_{k != j} s_ik; for every ordered pair (i, j) the reference adds
// ifac = dA_i/dmu_ij fac_i, fac_i = ew (1 - p) for i = g and -ew p otherwise, times the derivatives of mu_ij with respect
// to atom i, atom j and the point (which moves with atom g).  The pair loop runs over m in K (wave-uniform) and the
// lane's atoms k, evaluating each pair once: the terms of (k, m) and (m, k) that move k, the point term of (k, m), and
// for k outside K the terms of (m, k) that move m (summed over the wave) and the point; the reference's O(nAtoms^3)
// update of dA/dmu becomes O(nAtoms^2) per point, and the screen (kWeightScreen) leaves O(|K| nAtoms): the atoms whose
// cell product is not negligible at the point.  Both factors of a pair come from one polynomial (s_km and s_mk = 1 - s_km share nu), and
// one division gives both 1 / s_km and 1 / s_mk.
// LDS per wave: the distances R, their inverses and the non-zero products (3 nAtoms doubles), and the zero counts (nAtoms
// bytes, after the doubles of all WPG waves).  WPG waves per work-group (4, or 2 for molecules whose per-wave LDS would
// not fit four times in 64 KB); the waves share no LDS.
template <int LA, int WPG>
__global__ void __launch_bounds__(64 * WPG) weight_gradient_kernel(int nPoints,
                                                              const double *__restrict__ px, const double *__restrict__ py,
                                                              const double *__restrict__ pz, const double *__restrict__ pw,
                                                              const int *__restrict__ pointAtom,
                                                              const double *__restrict__ eXC, const double *__restrict__ xyz,
                                                              int nAtoms, const double *__restrict__ invR,
                                                              const double *__restrict__ aij,
                                                              const int *__restrict__ nearAtoms, int nNear,
                                                              double *__restrict__ grad) {
    extern __shared__ double lds[];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    double *R = lds + static_cast<size_t>(wv) * 3 * nAtoms, *iR = R + nAtoms, *prodNZ = iR + nAtoms;
    unsigned char *zeros = reinterpret_cast<unsigned char *>(lds + static_cast<size_t>(3 * WPG) * nAtoms) + wv * nAtoms;
    double G[LA][3];
#pragma unroll
    for (int a = 0; a < LA; a++) G[a][0] = G[a][1] = G[a][2] = 0.0;
    // Work-group g takes the rounds g, g + gridDim.x, ... of WPG consecutive points (a point's cost depends on the atoms
    // whose cells matter there, so contiguous ranges would differ in cost).
    for (int base = blockIdx.x * WPG; base < nPoints; base += gridDim.x * WPG) {
        const int p = base + wv;
        const bool active = p < nPoints && eXC[p] != 0.0;
        const double rx = active ? px[p] : 0.0, ry = active ? py[p] : 0.0, rz = active ? pz[p] : 0.0;
        if (active)
            for (int i = lane; i < nAtoms; i += 64) {
                const double dx = xyz[3 * i] - rx, dy = xyz[3 * i + 1] - ry, dz = xyz[3 * i + 2] - rz;
                R[i] = sqrt(dx * dx + dy * dy + dz * dz);
                iR[i] = 1.0 / R[i];
            }
        __syncthreads();
        // Products of the cell factors, their zero counts, and the screened set K (flag 0x80 of the count).  A_g and the
        // product of the atom nearest the point are lower bounds of sum_i A_i; the partial product of atom i over the
        // atoms nearest g bounds every product of i with one factor left out (|s| <= 1), so an atom whose bound is below
        // the screen at the larger of the two is outside K.  The
        // remaining candidates take their full products with the factors spread over the wave.  The others keep a zero
        // product and a count of 2, which leaves their terms out, as the screen leaves out those of the atoms outside K.
        double sumA = 0.0;
        int g = 0;
        unsigned long long maskK[LA];   // the atoms of K among atoms 64 a .. 64 a + 63 (the same in every lane)
#pragma unroll
        for (int a = 0; a < LA; a++) maskK[a] = 0ull;
        if (active) {
            g = pointAtom[p];
            auto full_product = [&](int i, double &prod, int &nz, double &minS) {
                prod = 1.0;
                minS = 1.0;
                nz = 0;
                for (int j = lane; j < nAtoms; j += 64) {
                    if (j == i) continue;
                    const int ij = pair_index(i, j);
                    const double s = becke_factor_of(i > j, R[i], R[j], invR[ij], aij[ij]);
                    if (s == 0.0) nz++;
                    else {
                        prod *= s;
                        minS = fmin(minS, fabs(s));
                    }
                }
                // Butterfly reductions: every lane ends with the same bits.
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    prod *= __shfl_xor(prod, o, 64);
                    minS = fmin(minS, __shfl_xor(minS, o, 64));
                    nz += __shfl_xor(nz, o, 64);
                }
            };
            double pg, mg;
            int zg;
            full_product(g, pg, zg, mg);
            // The atom nearest the point other than g (the wave's arg-min of R, ties to the lower index) and its product:
            // sum_i A_i >= max(A_g, A_near), which bounds the screen for points deep in another atom's cell too.
            double dn = 0.0;
            int near = -1;
            for (int j = lane; j < nAtoms; j += 64)
                if (j != g && (near < 0 || R[j] < dn)) { dn = R[j]; near = j; }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const double od = __shfl_xor(dn, o, 64);
                const int on = __shfl_xor(near, o, 64);
                if (on >= 0 && (near < 0 || od < dn || (od == dn && on < near))) { dn = od; near = on; }
            }
            double pn = 0.0, mn = 1.0;
            int zn = 2;
            if (near >= 0) full_product(near, pn, zn, mn);
            const double bound = 0x1p-81 * fmax(zg == 0 ? fabs(pg) : 0.0, zn == 0 ? fabs(pn) : 0.0);   // half the screen
            const int *nearList = nearAtoms + static_cast<size_t>(g) * nNear;
            bool cand[LA];
            double Q[LA];
#pragma unroll
            for (int a = 0; a < LA; a++) {
                const int i = lane + 64 * a;
                cand[a] = false;
                Q[a] = 0.0;
                if (i >= nAtoms) continue;
                if (i != g) {
                    double pp = 1.0, mp = 1.0;
                    int zp = 0;
                    for (int q = 0; q < nNear; q++) {
                        const int j = nearList[q];
                        if (j == i) continue;
                        const int ij = pair_index(i, j);
                        const double s = fabs(becke_factor_of(i > j, R[i], R[j], invR[ij], aij[ij]));
                        if (s == 0.0) zp++;
                        else {
                            pp *= s;
                            mp = fmin(mp, s);
                        }
                    }
                    const double qp = zp == 0 ? pp / mp : (zp == 1 ? pp : 0.0);
                    cand[a] = !(qp < bound);
                }
                prodNZ[i] = 0.0;
                zeros[i] = 2;
            }
#pragma unroll
            for (int a = 0; a < LA; a++) {
                unsigned long long mask = __ballot(cand[a] || lane + 64 * a == g || lane + 64 * a == near);
                while (mask != 0ull) {
                    const int bit = __ffsll(static_cast<long long>(mask)) - 1;
                    mask &= mask - 1ull;
                    const int c = 64 * a + bit;
                    double pc = pg, mc = mg;
                    int zc = zg;
                    if (c == near) { pc = pn; mc = mn; zc = zn; }
                    else if (c != g) full_product(c, pc, zc, mc);
                    if (zc == 0) sumA += pc;
                    if (lane == bit) {
                        cand[a] = true;
                        prodNZ[c] = pc;
                        zeros[c] = static_cast<unsigned char>(zc < 2 ? zc : 2);
                        Q[a] = zc == 0 ? fabs(pc) / mc : (zc == 1 ? fabs(pc) : 0.0);
                    }
                }
            }
#pragma unroll
            for (int a = 0; a < LA; a++) {
                const int i = lane + 64 * a;
                const bool inK = i < nAtoms && cand[a] && (i == g || !(Q[a] < kWeightScreen * sumA));
                if (inK) zeros[i] |= 0x80;
                maskK[a] = __ballot(inK);
            }
        }
        __syncthreads();
        if (active) {
            const double Ag = (zeros[g] & 0x7f) == 0 ? prodNZ[g] : 0.0;
            const double pW = Ag / sumA;
            const double ew = eXC[p] * pw[p] / (pW * sumA);
            double gs[3] = {0.0, 0.0, 0.0};
            // The lane's atoms k (their data held in registers across the pair loop).
            double kR[LA], kP[LA], kf[LA];
            int kNZ[LA];
            bool kIn[LA];
#pragma unroll
            for (int a = 0; a < LA; a++) {
                const int k = lane + 64 * a < nAtoms ? lane + 64 * a : 0;
                kR[a] = R[k];
                kP[a] = prodNZ[k];
                kNZ[a] = zeros[k] & 0x7f;
                kIn[a] = (zeros[k] & 0x80) != 0;
                kf[a] = k == g ? ew * (1.0 - pW) : -ew * pW;
            }
            // Every pair (k, m) with m in K and k a lane atom, each evaluated once: the terms of (k, m) and (m, k) that
            // move k and the point term of (k, m); for k outside K also the terms of (m, k) that move m and the point
            // (the terms of (k, m) that move m are below the screen), summed over the wave for m.
            for (int a2 = 0; a2 < LA; a2++)
            for (unsigned long long mk = maskK[a2]; mk != 0ull; mk &= mk - 1ull) {
                const int m = 64 * a2 + __ffsll(static_cast<long long>(mk)) - 1;
                const int nzm = zeros[m] & 0x7f;
                const double xm = xyz[3 * m], ym = xyz[3 * m + 1], zmc = xyz[3 * m + 2];
                const double Rm = R[m], pm = prodNZ[m], iRm = iR[m];
                const double emx = (xm - rx) * iRm, emy = (ym - ry) * iRm, emz = (zmc - rz) * iRm;
                const double fm = m == g ? ew * (1.0 - pW) : -ew * pW;
                double hx = 0.0, hy = 0.0, hz = 0.0;
#pragma unroll
                for (int a = 0; a < LA; a++) {
                    const int k = lane + 64 * a;
                    if (k >= nAtoms || k == m) continue;
                    const int km = pair_index(k, m);
                    const double r = invR[km], am = aij[km];
                    const bool kHigh = k > m;
                    // mu and nu of the pair stored as (max, min), as the cell factors take them.
                    const double muP = (kHigh ? kR[a] - Rm : Rm - kR[a]) * r;
                    const double nu = muP + am * (1.0 - muP * muP), nu2 = nu * nu;
                    const double S = kBeckeA * becke_poly(nu);
                    const double skm = kHigh ? 0.5 - S : 0.5 + S, smk = kHigh ? 0.5 + S : 0.5 - S;
                    double dsum = c_beckeD[10];
#pragma unroll
                    for (int tt = 9; tt >= 0; tt--) dsum = dsum * nu2 + c_beckeD[tt];
                    dsum *= -kBeckeA * (1.0 - 2.0 * am * muP);
                    const double sProd = skm * smk;
                    double invKM, invMK;
                    if (fabs(sProd) >= 0x1p-1000) {
                        const double inv = 1.0 / sProd;
                        invKM = smk * inv;
                        invMK = skm * inv;
                    } else {
                        invKM = 1.0 / skm;
                        invMK = 1.0 / smk;
                    }
                    const double ifKM = dsum * excluded_product(kP[a], kNZ[a], skm, invKM) * kf[a];
                    const double ifMK = dsum * excluded_product(pm, nzm, smk, invMK) * fm;
                    const double mu = kHigh ? muP : -muP, mr2 = mu * r * r;
                    const double xk = xyz[3 * k], yk = xyz[3 * k + 1], zk = xyz[3 * k + 2], iRk = iR[k];
                    const double kex = (xk - rx) * iRk, key = (yk - ry) * iRk, kez = (zk - rz) * iRk;
                    const double dx = xk - xm, dy = yk - ym, dz = zk - zmc;
                    // Pair (k, m) moves atom k through d mu / d r_k and the point through -(d mu / d r_k + d mu / d r_m);
                    // pair (m, k) moves atom k through its second atom: both along r (e_k) - mu r^2 (r_k - r_m).
                    const double d = ifKM - ifMK;
                    G[a][0] += d * (kex * r - dx * mr2);
                    G[a][1] += d * (key * r - dy * mr2);
                    G[a][2] += d * (kez * r - dz * mr2);
                    double wr = ifKM * r;
                    if (!kIn[a]) {
                        // Pair (m, k) moves atom m along r (e_m) - (-mu) r^2 (r_m - r_k) = r (e_m) - mu r^2 (r_k - r_m).
                        hx += ifMK * (emx * r - dx * mr2);
                        hy += ifMK * (emy * r - dy * mr2);
                        hz += ifMK * (emz * r - dz * mr2);
                        wr -= ifMK * r;
                    }
                    gs[0] += wr * (emx - kex);
                    gs[1] += wr * (emy - key);
                    gs[2] += wr * (emz - kez);
                }
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    hx += __shfl_xor(hx, o, 64);
                    hy += __shfl_xor(hy, o, 64);
                    hz += __shfl_xor(hz, o, 64);
                }
                if (lane == (m & 63)) {
#pragma unroll
                    for (int a = 0; a < LA; a++)
                        if (a == (m >> 6)) { G[a][0] += hx; G[a][1] += hy; G[a][2] += hz; }
                }
            }
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) gs[d] += __shfl_xor(gs[d], o, 64);
            if (lane == (g & 63)) {
#pragma unroll
                for (int a = 0; a < LA; a++)
                    if (a == (g >> 6)) { G[a][0] += gs[0]; G[a][1] += gs[1]; G[a][2] += gs[2]; }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < LA; a++) {
        const int k = lane + 64 * a;
        if (k < nAtoms) {
            atomicAdd(&grad[3 * k], G[a][0]);
            atomicAdd(&grad[3 * k + 1], G[a][1]);
            atomicAdd(&grad[3 * k + 2], G[a][2]);
        }
    }
}

}  // namespace

int build_grid(xq_grid *g) {
    const int nA = g->nAtoms;
    g->bfTol = kBFTol[g->accuracy];
    g->rhoTol = kRhoTol[g->accuracy];
    // Lebedev rules used by this grid, expanded once.
    std::vector<int> ruleFirst(XQ_LEBEDEV_NRULES, -1);
    std::vector<double> rules;
    auto rule_of = [&](int l) -> int {
        int n = lebedev_size_for_l(l), k = 0;
        while (k < XQ_LEBEDEV_NRULES && xq_lebedev_npoints[k] != n) k++;
        if (ruleFirst[k] < 0) {
            ruleFirst[k] = static_cast<int>(rules.size() / 4);
            rules.resize(rules.size() + 4 * static_cast<size_t>(n));
            lebedev_expand(k, rules.data() + 4 * static_cast<size_t>(ruleFirst[k]));
        }
        return k;
    };
    // Radial shells, per atom in order (DFTGrid.c:138-211).
    std::vector<RadialShell> shells;
    std::vector<int> atomFirst(nA + 1, 0);
    int nCand = 0;
    for (int a = 0; a < nA; a++) {
        int nR, lMax;
        double rMax;
        atom_parameters(g->accuracy, g->Z[a], &nR, &lMax, &rMax);
        const int nAngMax = lebedev_size_for_l(lMax), nAngMin = lebedev_size_for_l(kMinL);
        const double range = bragg_radius(g->Z[a]), rCut = kRadialCutFactor * range;
        const double d = 1.0 + static_cast<double>(nR);
        atomFirst[a] = nCand;
        for (int i = 0; i < nR; i++) {
            const double x = static_cast<double>(i + 1) / d;
            const double r = range * std::pow(x / (1.0 - x), 2.0);
            const double wr = 2.0 * std::pow(range, 3.0) * std::pow(x, 5.0) / std::pow(1.0 - x, 7.0) / d;
            if (r > rMax) break;
            int l;
            if (r > rCut) l = lMax;
            else {
                int n = static_cast<int>(std::ceil(static_cast<double>(nAngMax) * r / rCut));
                l = n < nAngMin ? kMinL : lebedev_l_for_size(n);
            }
            const int k = rule_of(l);
            RadialShell s;
            s.atom = a;
            s.rule = ruleFirst[k];
            s.nAng = xq_lebedev_npoints[k];
            s.out = nCand;
            s.r = r;
            s.wr = 4.0 * M_PI * wr;
            shells.push_back(s);
            nCand += s.nAng;
        }
    }
    atomFirst[nA] = nCand;
    // Becke pair data (DFTGridWeights.c:43-80): 1/R_ij and the size adjustment a_ij, for i > j.
    const size_t nPairs = nA > 1 ? static_cast<size_t>(nA) * (nA - 1) / 2 : 1;
    std::vector<double> invR, aij;
    becke_pairs(g, invR, aij);
    // Device work.
    RadialShell *d_shells = nullptr;
    double *d_rules = nullptr, *d_xyz = nullptr, *d_invR = nullptr, *d_aij = nullptr;
    int *d_order = nullptr;
    double *cx = nullptr, *cy = nullptr, *cz = nullptr, *cw = nullptr;
    int *keep = nullptr, *dst = nullptr, *total = nullptr, *tileSums = nullptr;
    // LDS of the candidate kernel: atom coordinates, the shell atom's pair row and its distance order (44 bytes per
    // atom).
    const size_t lds = (5 * sizeof(double) + sizeof(int)) * static_cast<size_t>(nA);
    if (lds > 64 * 1024) {
        set_error("grid construction supports at most 1489 atoms");
        return kInvalidArgument;
    }
    // Per atom A, all atoms in order of distance from A (A first).
    std::vector<int> order(static_cast<size_t>(nA) * nA);
    {
        std::vector<double> d2(nA);
        for (int a = 0; a < nA; a++) {
            int *row = order.data() + static_cast<size_t>(a) * nA;
            for (int j = 0; j < nA; j++) {
                const double dx = g->xyz[3 * a] - g->xyz[3 * j], dy = g->xyz[3 * a + 1] - g->xyz[3 * j + 1],
                             dz = g->xyz[3 * a + 2] - g->xyz[3 * j + 2];
                d2[j] = dx * dx + dy * dy + dz * dz;
                row[j] = j;
            }
            std::stable_sort(row, row + nA, [&](int p, int q) { return d2[p] < d2[q]; });
        }
    }
    int status = kOK;
    auto cleanup = [&]() {
        // The kernels that use these arrays must be done before the arrays are reused.
        (void)hipStreamSynchronize(g->stream);
        dev_free(g, &d_shells); dev_free(g, &d_rules); dev_free(g, &d_xyz); dev_free(g, &d_invR); dev_free(g, &d_aij);
        dev_free(g, &d_order);
        dev_free(g, &cx); dev_free(g, &cy); dev_free(g, &cz); dev_free(g, &cw); dev_free(g, &keep); dev_free(g, &dst); dev_free(g, &total);
        dev_free(g, &tileSums);
    };
    do {
        if ((status = dev_alloc(g, &d_shells, shells.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_rules, rules.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_xyz, g->xyz.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_invR, nPairs)) != kOK) break;
        if ((status = dev_alloc(g, &d_aij, nPairs)) != kOK) break;
        if ((status = dev_alloc(g, &d_order, order.size())) != kOK) break;
        if ((status = dev_alloc(g, &cx, nCand)) != kOK) break;
        if ((status = dev_alloc(g, &cy, nCand)) != kOK) break;
        if ((status = dev_alloc(g, &cz, nCand)) != kOK) break;
        if ((status = dev_alloc(g, &cw, nCand)) != kOK) break;
        if ((status = dev_alloc(g, &keep, nCand)) != kOK) break;
        if ((status = dev_alloc(g, &dst, nCand)) != kOK) break;
        if ((status = dev_alloc(g, &total, 1)) != kOK) break;
        if ((status = dev_alloc(g, &tileSums, static_cast<size_t>((nCand + kScanTile - 1) / kScanTile))) != kOK) break;
        hipStream_t s = g->stream;
        if ((status = check(hipMemcpyAsync(d_shells, shells.data(), shells.size() * sizeof(RadialShell),
                                           hipMemcpyHostToDevice, s), "copy shells")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_rules, rules.data(), rules.size() * sizeof(double), hipMemcpyHostToDevice, s),
                            "copy rules")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_xyz, g->xyz.data(), g->xyz.size() * sizeof(double), hipMemcpyHostToDevice, s),
                            "copy xyz")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_invR, invR.data(), nPairs * sizeof(double), hipMemcpyHostToDevice, s),
                            "copy invR")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_aij, aij.data(), nPairs * sizeof(double), hipMemcpyHostToDevice, s),
                            "copy aij")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, s),
                            "copy order")) != kOK) break;
        if (!shells.empty()) {
            hipLaunchKernelGGL(candidate_points_kernel, dim3(shells.size()), dim3(256), lds, s, d_shells, d_rules, d_xyz, nA,
                               d_invR, d_aij, d_order, cx, cy, cz, cw, keep);
            if ((status = check(hipGetLastError(), "candidate_points_kernel")) != kOK) break;
            const int nTiles = (nCand + kScanTile - 1) / kScanTile;
            hipLaunchKernelGGL(scan_tile_sums_kernel, dim3(nTiles), dim3(256), 0, s, nCand, keep, tileSums);
            hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(256), 0, s, nTiles, tileSums, total);
            hipLaunchKernelGGL(scan_apply_kernel, dim3(nTiles), dim3(256), 0, s, nCand, keep, tileSums, dst);
            if ((status = check(hipGetLastError(), "scan kernels")) != kOK) break;
        }
        // Per-atom kept counts from the scan at the atom boundaries.
        std::vector<int> scanned(nCand > 0 ? nCand : 1, 0), kept(nCand > 0 ? nCand : 1, 0);
        int nKept = 0;
        if (nCand > 0) {
            // Wait for the kernels first: a pageable copy that waits for its stream holds the runtime's staging path.
            if ((status = check(hipStreamSynchronize(s), "grid sync")) != kOK) break;
            if ((status = check(hipMemcpyAsync(scanned.data(), dst, nCand * sizeof(int), hipMemcpyDeviceToHost, s),
                                "copy scan")) != kOK) break;
            if ((status = check(hipMemcpyAsync(kept.data(), keep, nCand * sizeof(int), hipMemcpyDeviceToHost, s),
                                "copy keep")) != kOK) break;
            if ((status = check(hipMemcpyAsync(&nKept, total, sizeof(int), hipMemcpyDeviceToHost, s), "copy total")) != kOK)
                break;
            if ((status = check(hipStreamSynchronize(s), "grid sync")) != kOK) break;
        }
        g->nPoints = nKept;
        if ((status = dev_alloc(g, &g->d_x, nKept)) != kOK) break;
        if ((status = dev_alloc(g, &g->d_y, nKept)) != kOK) break;
        if ((status = dev_alloc(g, &g->d_z, nKept)) != kOK) break;
        if ((status = dev_alloc(g, &g->d_w, nKept)) != kOK) break;
        if (nCand > 0) {
            hipLaunchKernelGGL(compact_points_kernel, dim3((nCand + 255) / 256), dim3(256), 0, s, nCand, keep, dst, cx, cy,
                               cz, cw, g->d_x, g->d_y, g->d_z, g->d_w);
            if ((status = check(hipGetLastError(), "compact_points_kernel")) != kOK) break;
        }
        // Blocks of at most kBlockSize consecutive points per atom.
        g->blocks.clear();
        for (int a = 0; a < nA; a++) {
            const int c0 = atomFirst[a], c1 = atomFirst[a + 1];
            if (c1 <= c0) continue;
            const int p0 = scanned[c0];
            const int p1 = scanned[c1 - 1] + kept[c1 - 1];
            for (int p = p0; p < p1; p += kBlockSize) {
                DevBlock b;
                b.atom = a;
                b.pStart = p;
                b.nPoints = (p1 - p) < kBlockSize ? (p1 - p) : kBlockSize;
                b.nKept = 0;
                b.vOffset = 0;
                b.iOffset = 0;
                b.pad = 0;
                g->blocks.push_back(b);
            }
        }
        g->nBlocks = static_cast<int>(g->blocks.size());
        if ((status = dev_alloc(g, &g->d_blocks, g->blocks.size())) != kOK) break;
        if (!g->blocks.empty() &&
            (status = check(hipMemcpyAsync(g->d_blocks, g->blocks.data(), g->blocks.size() * sizeof(DevBlock),
                                           hipMemcpyHostToDevice, s), "copy blocks")) != kOK) break;
        status = check(hipStreamSynchronize(s), "grid sync");
    } while (false);
    cleanup();   // only the kept points and the block table stay resident (and counted on the handle)
    return status;
}

// Estimated point count without the cull (DFTGrid_EstimatedPoints, DFTGrid.c:301-356).
long long estimated_points(int accuracy, int nAtoms, const int *Z) {
    long long p = 0;
    for (int a = 0; a < nAtoms; a++) {
        int nR, lMax;
        double rMax;
        atom_parameters(accuracy, Z[a], &nR, &lMax, &rMax);
        const int nAngMax = lebedev_size_for_l(lMax), nAngMin = lebedev_size_for_l(kMinL);
        const double range = bragg_radius(Z[a]), rCut = kRadialCutFactor * range, d = 1.0 + nR;
        for (int i = 0; i < nR; i++) {
            const double x = static_cast<double>(i + 1) / d;
            const double r = range * std::pow(x / (1.0 - x), 2.0);
            if (r > rMax) break;
            int l;
            if (r > rCut) l = lMax;
            else {
                int n = static_cast<int>(std::ceil(static_cast<double>(nAngMax) * r / rCut));
                l = n < nAngMin ? kMinL : lebedev_l_for_size(n);
            }
            p += lebedev_size_for_l(l);
        }
    }
    return p;
}


// XC gradient weight terms of every point with a non-zero energy density eXC (device arrays; grad is incremented).
int weight_gradients(xq_grid *g, const double *d_eXC, double *d_grad, hipStream_t s) {
    const int nA = g->nAtoms, G = g->nPoints;
    if (G <= 0 || nA < 2) return kOK;
    // Waves per work-group: four up to 640 atoms, two above (each wave needs 25 bytes of LDS per atom: 4 x 25 x 640 and
    // 2 x 25 x 1024 bytes both fit in 64 KB).  The kernel instance below takes the same count.
    const size_t perWave = (3 * sizeof(double) + 1) * static_cast<size_t>(nA);
    const int wpg = nA <= 640 ? 4 : 2;
    const size_t lds = wpg * perWave;
    if (nA > 64 * kMaxLaneAtoms || lds > 64 * 1024) {
        set_error("XC gradients support at most " + std::to_string(64 * kMaxLaneAtoms) + " atoms");
        return kInvalidArgument;
    }
    std::vector<double> invR, aij;
    becke_pairs(g, invR, aij);
    std::vector<int> pointAtom(static_cast<size_t>(G), 0);
    for (const DevBlock &b : g->blocks)
        for (int k = 0; k < b.nPoints; k++) pointAtom[static_cast<size_t>(b.pStart) + k] = b.atom;
    // Per atom: itself and the atoms nearest it (the partial products that bound the cell products of a point).
    const int nNear = std::min(kNearAtoms, nA);
    std::vector<int> nearAtoms(static_cast<size_t>(nA) * nNear), order(nA);
    for (int a = 0; a < nA; a++) {
        std::vector<double> d2(nA);
        for (int j = 0; j < nA; j++) {
            const double dx = g->xyz[3 * j] - g->xyz[3 * a], dy = g->xyz[3 * j + 1] - g->xyz[3 * a + 1],
                         dz = g->xyz[3 * j + 2] - g->xyz[3 * a + 2];
            d2[j] = j == a ? -1.0 : dx * dx + dy * dy + dz * dz;
            order[j] = j;
        }
        std::partial_sort(order.begin(), order.begin() + nNear, order.end(),
                          [&](int u, int v) { return d2[u] < d2[v] || (d2[u] == d2[v] && u < v); });
        for (int q = 0; q < nNear; q++) nearAtoms[static_cast<size_t>(a) * nNear + q] = order[q];
    }
    double *d_invR = nullptr, *d_aij = nullptr, *d_xyz = nullptr;
    int *d_pointAtom = nullptr, *d_near = nullptr;
    int status = kOK;
    do {
        if ((status = dev_alloc(g, &d_invR, invR.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_aij, aij.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_xyz, g->xyz.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_pointAtom, pointAtom.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_near, nearAtoms.size())) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_near, nearAtoms.data(), nearAtoms.size() * sizeof(int),
                                           hipMemcpyHostToDevice, s), "copy near atoms")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_invR, invR.data(), invR.size() * sizeof(double), hipMemcpyHostToDevice, s),
                            "copy invR")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_aij, aij.data(), aij.size() * sizeof(double), hipMemcpyHostToDevice, s),
                            "copy aij")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_xyz, g->xyz.data(), g->xyz.size() * sizeof(double), hipMemcpyHostToDevice, s),
                            "copy xyz")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_pointAtom, pointAtom.data(), pointAtom.size() * sizeof(int),
                                           hipMemcpyHostToDevice, s), "copy point atoms")) != kOK) break;
        const int nGroups = std::min(2048 * 4 / wpg, (G + 63) / 64);
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(nGroups), dim3(64 * wpg), lds, s, G, g->d_x, g->d_y, g->d_z,
                               g->d_w, d_pointAtom, d_eXC, d_xyz, nA, d_invR, d_aij, d_near, nNear, d_grad);
        };
        // Atoms per lane: the smallest instance that covers nAtoms (fewer registers for small systems).
        if (nA <= 64) launch(weight_gradient_kernel<1, 4>);
        else if (nA <= 128) launch(weight_gradient_kernel<2, 4>);
        else if (nA <= 192) launch(weight_gradient_kernel<3, 4>);
        else if (nA <= 256) launch(weight_gradient_kernel<4, 4>);
        else if (nA <= 384) launch(weight_gradient_kernel<6, 4>);
        else if (nA <= 640) launch(weight_gradient_kernel<10, 4>);
        else launch(weight_gradient_kernel<kMaxLaneAtoms, 2>);
        if ((status = check(hipGetLastError(), "weight_gradient_kernel")) != kOK) break;
        status = check(hipStreamSynchronize(s), "gradient sync");
    } while (false);
    (void)hipStreamSynchronize(s);
    dev_free(g, &d_invR);
    dev_free(g, &d_aij);
    dev_free(g, &d_xyz);
    dev_free(g, &d_pointAtom);
    dev_free(g, &d_near);
    return status;
}

}  // namespace xq
