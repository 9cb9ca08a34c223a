// Gaussian basis set-up and block-wise basis values and gradients on the device.
//
// Host side: the element specification is turned into per-atom shells with unnormalised primitive coefficients and the
// Cartesian factors of GaussianBasis_UnnormalizePrimitives / _NormalizePrimitiveCCBF (pMolecule/QCModel/GaussianBases/
// __extensions__/csource/GaussianBasis.c:448-501, 728-751), and the Cartesian -> real solid harmonic transformation of
// Schlegel & Frisch (Int. J. Quantum Chem. 54 (1995) 83, eq. 15), with the reference's function order (GaussianBasis.c:13-27).
//
// Device side, once per geometry ("in core", DFTIntegrator.c:209-210,491):
//   1. count kernel, one work-group per grid block: bounding sphere of the block's points, a rigorous per-shell upper bound on
//      |phi| over the sphere (shells that cannot reach bfTol are skipped), exact values of the remaining functions, and the
//      reference's filter max_p |phi_m(p)| > bfTol (GridFunctionDataBlock.c:130-185) -> ascending kept-function list.
//   2. eval kernel, one work-group per block: values and gradients of the kept functions (GaussianBasisIntegrals_f1Op1.c:
//      100-197 and GaussianBasisTransform.c:85-125), stored function-major in rows of 128 points (coalesced stores).
// The kept sets are exactly the reference's: screening removes only functions whose bound is below the tolerance.
#include <cmath>
#include <cstring>
#include <hip/hip_runtime.h>

#include "xq_internal.h"
#include "xq_c2s.h"

namespace xq {

namespace {

constexpr int kMaxShellsLDS = 16384;

// Cartesian powers in the reference order (for z = 0..l, for y = 0..l-z, x = l-y-z), for l = 0..3, concatenated.
__host__ __device__ constexpr int cart_offset(int l) { return l * (l + 1) * (l + 2) / 6; }
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

struct CartPow {
    int x[20], y[20], z[20];
};
__host__ __device__ constexpr CartPow make_cart_pow() {
    CartPow p{};
    int n = 0;
    for (int l = 0; l <= 3; l++)
        for (int z = 0; z <= l; z++)
            for (int y = 0; y <= l - z; y++) {
                p.x[n] = l - y - z;
                p.y[n] = y;
                p.z[n] = z;
                n++;
            }
    return p;
}
constexpr CartPow kCart = make_cart_pow();

// ------------------------------------------------------------------------------------------------------------------------
// Device evaluation of one shell at one point.  Template over (lLow, lHigh, spherical) so that every array index is a
// compile-time constant and stays in registers.
// ------------------------------------------------------------------------------------------------------------------------
template <int LL, int LH>
struct ShellDims {
    static constexpr int nCBF = cart_offset(LH + 1) - cart_offset(LL);
    static constexpr int nSph = (LH + 1) * (LH + 1) - LL * LL;
};

template <int LL, int LH, bool SPH, bool GRAD>
__device__ __forceinline__ void shell_eval(const xq::DevShell &sh, const double *__restrict__ coef,
                                           const double *__restrict__ trans, double dx, double dy, double dz, double *v,
                                           double *vx, double *vy, double *vz) {
    constexpr int NC = ShellDims<LL, LH>::nCBF;
    constexpr int NB = SPH ? ShellDims<LL, LH>::nSph : NC;
    constexpr int C0 = cart_offset(LL);
    const double r2 = dx * dx + dy * dy + dz * dz;
    double e0[NC], e1[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) { e0[c] = 0.0; e1[c] = 0.0; }
    const double *ex = coef + sh.coef, *cc = coef + sh.coef + sh.nPrim;
    for (int q = 0; q < sh.nPrim; q++) {
        const double a = ex[q], e = exp(-a * r2);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const double k = cc[q * NC + c];
            e0[c] += k * e;
            e1[c] += k * a * e;
        }
    }
    double X0[LH + 2], Y0[LH + 2], Z0[LH + 2];
    X0[0] = Y0[0] = Z0[0] = 1.0;
#pragma unroll
    for (int i = 1; i <= LH + 1; i++) { X0[i] = dx * X0[i - 1]; Y0[i] = dy * Y0[i - 1]; Z0[i] = dz * Z0[i - 1]; }
    double g0[NC], gx[NC], gy[NC], gz[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int i = kCart.x[C0 + c], j = kCart.y[C0 + c], k = kCart.z[C0 + c];
        g0[c] = X0[i] * Y0[j] * Z0[k] * e0[c];
        if (GRAD) {
            const double E1 = -2.0 * e1[c];
            const double xi1 = i > 0 ? i * X0[i > 0 ? i - 1 : 0] : 0.0;
            const double yj1 = j > 0 ? j * Y0[j > 0 ? j - 1 : 0] : 0.0;
            const double zk1 = k > 0 ? k * Z0[k > 0 ? k - 1 : 0] : 0.0;
            gx[c] = (xi1 * e0[c] + X0[i + 1] * E1) * Y0[j] * Z0[k];
            gy[c] = (yj1 * e0[c] + Y0[j + 1] * E1) * X0[i] * Z0[k];
            gz[c] = (zk1 * e0[c] + Z0[k + 1] * E1) * X0[i] * Y0[j];
        }
    }
    if (SPH) {
        const double *T = trans + sh.trans;
#pragma unroll
        for (int s = 0; s < NB; s++) {
            double a = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const double t = T[c * NB + s];
                a += t * g0[c];
                if (GRAD) { bx += t * gx[c]; by += t * gy[c]; bz += t * gz[c]; }
            }
            v[s] = a;
            if (GRAD) { vx[s] = bx; vy[s] = by; vz[s] = bz; }
        }
    } else {
#pragma unroll
        for (int s = 0; s < NB; s++) {
            v[s] = g0[s];
            if (GRAD) { vx[s] = gx[s]; vy[s] = gy[s]; vz[s] = gz[s]; }
        }
    }
}

// Dispatch on the shell type; calls body(nBasis, v, vx, vy, vz).  Supported: S, P, D, F, SP.
template <bool GRAD, typename Body>
__device__ __forceinline__ void with_shell(const xq::DevShell &sh, bool sph, const double *coef, const double *trans,
                                           double dx, double dy, double dz, Body body) {
    double v[16], vx[16], vy[16], vz[16];
    const int key = sh.lLow * 4 + sh.lHigh;
    if (sph) {
        switch (key) {
            case 0: shell_eval<0, 0, true, GRAD>(sh, coef, trans, dx, dy, dz, v, vx, vy, vz); body(1, v, vx, vy, vz); break;
            case 5: shell_eval<1, 1, true, GRAD>(sh, coef, trans, dx, dy, dz, v, vx, vy, vz); body(3, v, vx, vy, vz); break;
            case 1: shell_eval<0, 1, true, GRAD>(sh, coef, trans, dx, dy, dz, v, vx, vy, vz); body(4, v, vx, vy, vz); break;
            case 10: shell_eval<2, 2, true, GRAD>(sh, coef, trans, dx, dy, dz, v, vx, vy, vz); body(5, v, vx, vy, vz); break;
            case 15: shell_eval<3, 3, true, GRAD>(sh, coef, trans, dx, dy, dz, v, vx, vy, vz); body(7, v, vx, vy, vz); break;
            default: break;
        }
    } else {
        switch (key) {
            case 0: shell_eval<0, 0, false, GRAD>(sh, coef, trans, dx, dy, dz, v, vx, vy, vz); body(1, v, vx, vy, vz); break;
            case 5: shell_eval<1, 1, false, GRAD>(sh, coef, trans, dx, dy, dz, v, vx, vy, vz); body(3, v, vx, vy, vz); break;
            case 1: shell_eval<0, 1, false, GRAD>(sh, coef, trans, dx, dy, dz, v, vx, vy, vz); body(4, v, vx, vy, vz); break;
            case 10: shell_eval<2, 2, false, GRAD>(sh, coef, trans, dx, dy, dz, v, vx, vy, vz); body(6, v, vx, vy, vz); break;
            case 15: shell_eval<3, 3, false, GRAD>(sh, coef, trans, dx, dy, dz, v, vx, vy, vz); body(10, v, vx, vy, vz); break;
            default: break;
        }
    }
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Upper bound of max_{r >= d} r^L exp(-a r^2).
__device__ __forceinline__ double radial_bound(int L, double a, double d) {
    if (L == 0) return exp(-a * d * d);
    const double rs = sqrt(L / (2.0 * a)), r = d > rs ? d : rs;
    double p = 1.0;
    for (int i = 0; i < L; i++) p *= r;
    return p * exp(-a * r * r);
}

// Work-group of kBlockSize threads, one thread per point.  Writes the kept functions of block b to tmpIdx[b * nF ...] and
// their count to counts[b].
__global__ void __launch_bounds__(kBlockSize) basis_count_kernel(
    int nBlocks, const xq::DevBlock *__restrict__ blocks, const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, int nShells, const xq::DevShell *__restrict__ shells, const double *__restrict__ coef,
    const double *__restrict__ trans, int spherical, int nF, double bfTol, int *__restrict__ tmpIdx,
    int *__restrict__ counts) {
    __shared__ unsigned char cand[kMaxShellsLDS];
    __shared__ double red[4][2];
    __shared__ int nKeptS;
    const int b = blockIdx.x;
    const xq::DevBlock blk = blocks[b];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const bool valid = t < blk.nPoints;
    const double x = valid ? px[blk.pStart + t] : 0.0, y = valid ? py[blk.pStart + t] : 0.0,
                 z = valid ? pz[blk.pStart + t] : 0.0;
    // Bounding sphere: centroid and largest distance.
    double sx = wave_sum(x), sy = wave_sum(y), sz = wave_sum(z);
    if (lane == 0) { red[wave][0] = sx; red[wave][1] = sy; red[wave + 2][0] = sz; }
    __syncthreads();
    const double inv = 1.0 / blk.nPoints;
    const double cx = (red[0][0] + red[1][0]) * inv, cy = (red[0][1] + red[1][1]) * inv, cz = (red[2][0] + red[3][0]) * inv;
    __syncthreads();
    double rr = valid ? sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz)) : 0.0;
    rr = wave_max(rr);
    if (lane == 0) red[wave][0] = rr;
    __syncthreads();
    const double radius = fmax(red[0][0], red[1][0]);
    // Candidate shells.
    for (int s = t; s < nShells; s += blockDim.x) {
        const xq::DevShell &sh = shells[s];
        const double dd = sqrt((sh.x - cx) * (sh.x - cx) + (sh.y - cy) * (sh.y - cy) + (sh.z - cz) * (sh.z - cz)) - radius;
        const double d = dd > 0.0 ? dd : 0.0;
        double bound = 0.0;
        const double *ex = coef + sh.coef;
        for (int q = 0; q < sh.nPrim; q++)
            for (int L = sh.lLow; L <= sh.lHigh; L++) bound += sh.bound[q] * radial_bound(L, ex[q], d);
        cand[s] = bound * (1.0 + 1.0e-12) > bfTol ? 1 : 0;
    }
    if (t == 0) nKeptS = 0;
    __syncthreads();
    int *out = tmpIdx + static_cast<size_t>(b) * nF;
    for (int s = 0; s < nShells; s++) {
        if (!cand[s]) continue;
        const xq::DevShell sh = shells[s];
        with_shell<false>(sh, spherical != 0, coef, trans, x - sh.x, y - sh.y, z - sh.z,
                          [&](int nB, const double *v, const double *, const double *, const double *) {
                              for (int k = 0; k < nB; k++) {
                                  double m = wave_max(valid ? fabs(v[k]) : 0.0);
                                  if (lane == 0) red[wave][0] = m;
                                  __syncthreads();
                                  if (t == 0 && fmax(red[0][0], red[1][0]) > bfTol) out[nKeptS++] = sh.fStart + k;
                                  __syncthreads();
                              }
                          });
    }
    if (t == 0) counts[b] = nKeptS;
}

// Values and gradients of the kept functions of block b into rows [vOffset, vOffset + nKept) of f, fx, fy, fz (row stride
// kBlockSize; entries past the block's points are zero).  T is the storage type.
template <typename T>
__global__ void __launch_bounds__(kBlockSize) basis_eval_kernel(
    const xq::DevBlock *__restrict__ blocks, const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, const xq::DevShell *__restrict__ shells, const int *__restrict__ shellOf,
    const double *__restrict__ coef, const double *__restrict__ trans, int spherical, int nF,
    const int *__restrict__ tmpIdx, int *__restrict__ idx, T *__restrict__ f, T *__restrict__ fx, T *__restrict__ fy,
    T *__restrict__ fz) {
    const int b = blockIdx.x;
    const xq::DevBlock blk = blocks[b];
    const int t = threadIdx.x;
    const bool valid = t < blk.nPoints;
    const double x = valid ? px[blk.pStart + t] : 0.0, y = valid ? py[blk.pStart + t] : 0.0,
                 z = valid ? pz[blk.pStart + t] : 0.0;
    const int *kept = tmpIdx + static_cast<size_t>(b) * nF;
    for (int k = t; k < blk.nKept; k += blockDim.x) idx[blk.iOffset + k] = kept[k];
    int k = 0;
    while (k < blk.nKept) {
        const xq::DevShell sh = shells[shellOf[kept[k]]];
        with_shell<true>(sh, spherical != 0, coef, trans, x - sh.x, y - sh.y, z - sh.z,
                         [&](int nB, const double *v, const double *vx, const double *vy, const double *vz) {
                             for (int j = 0; j < nB; j++) {
                                 if (k < blk.nKept && kept[k] == sh.fStart + j) {
                                     const size_t o = static_cast<size_t>(blk.vOffset + k) * kBlockSize + t;
                                     f[o] = static_cast<T>(valid ? v[j] : 0.0);
                                     fx[o] = static_cast<T>(valid ? vx[j] : 0.0);
                                     fy[o] = static_cast<T>(valid ? vy[j] : 0.0);
                                     fz[o] = static_cast<T>(valid ? vz[j] : 0.0);
                                     k++;
                                 }
                             }
                         });
    }
}

// Values, gradients and second derivatives of one shell (GaussianBasisIntegrals_f1Op1.c:203-331): with e2 = 4 sum c a^2 e^{-a r^2},
// d2/dx2 phi = (i(i-1) x^{i-2} e0 + (2i+1) x^i e1 + x^{i+2} e2) y^j z^k and
// d2/dxdy phi = (i j x^{i-1} y^{j-1} e0 + (i x^{i-1} y^{j+1} + j x^{i+1} y^{j-1}) e1 + x^{i+1} y^{j+1} e2) z^k; out[0..9] = phi, x, y,
// z, xx, xy, xz, yy, yz, zz (nBasis values each, spherical = c2s^T Cartesian).
template <int LL, int LH, bool SPH>
__device__ __forceinline__ void shell_eval2(const xq::DevShell &sh, const double *__restrict__ coef,
                                            const double *__restrict__ trans, double dx, double dy, double dz,
                                            double (*out)[16]) {
    constexpr int NC = ShellDims<LL, LH>::nCBF;
    constexpr int NB = SPH ? ShellDims<LL, LH>::nSph : NC;
    constexpr int C0 = cart_offset(LL);
    const double r2 = dx * dx + dy * dy + dz * dz;
    double e0[NC], e1[NC], e2[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) { e0[c] = 0.0; e1[c] = 0.0; e2[c] = 0.0; }
    const double *ex = coef + sh.coef, *cc = coef + sh.coef + sh.nPrim;
    for (int q = 0; q < sh.nPrim; q++) {
        const double a = ex[q], e = exp(-a * r2);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const double k = cc[q * NC + c];
            e0[c] += k * e;
            e1[c] += k * a * e;
            e2[c] += k * (a * a) * e;
        }
    }
    double X0[LH + 3], Y0[LH + 3], Z0[LH + 3], X1[LH + 3], Y1[LH + 3], Z1[LH + 3], X2[LH + 3], Y2[LH + 3], Z2[LH + 3];
    X0[0] = Y0[0] = Z0[0] = 1.0;
    X1[0] = Y1[0] = Z1[0] = X2[0] = Y2[0] = Z2[0] = 0.0;
#pragma unroll
    for (int i = 1; i <= LH + 2; i++) {
        X0[i] = dx * X0[i - 1]; Y0[i] = dy * Y0[i - 1]; Z0[i] = dz * Z0[i - 1];
        X1[i] = i * X0[i - 1]; Y1[i] = i * Y0[i - 1]; Z1[i] = i * Z0[i - 1];
        X2[i] = i * X1[i - 1]; Y2[i] = i * Y1[i - 1]; Z2[i] = i * Z1[i - 1];
    }
    double g[10][NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int i = kCart.x[C0 + c], j = kCart.y[C0 + c], k = kCart.z[C0 + c];
        const double E0 = e0[c], E1 = -2.0 * e1[c], E2 = 4.0 * e2[c];
        g[0][c] = X0[i] * Y0[j] * Z0[k] * E0;
        g[1][c] = (X1[i] * E0 + X0[i + 1] * E1) * Y0[j] * Z0[k];
        g[2][c] = (Y1[j] * E0 + Y0[j + 1] * E1) * X0[i] * Z0[k];
        g[3][c] = (Z1[k] * E0 + Z0[k + 1] * E1) * X0[i] * Y0[j];
        g[4][c] = (X2[i] * E0 + (dx * X1[i] + X1[i + 1]) * E1 + X0[i + 2] * E2) * Y0[j] * Z0[k];
        g[5][c] = (X1[i] * Y1[j] * E0 + (X1[i] * Y0[j + 1] + X0[i + 1] * Y1[j]) * E1 + X0[i + 1] * Y0[j + 1] * E2) * Z0[k];
        g[6][c] = (X1[i] * Z1[k] * E0 + (X1[i] * Z0[k + 1] + X0[i + 1] * Z1[k]) * E1 + X0[i + 1] * Z0[k + 1] * E2) * Y0[j];
        g[7][c] = (Y2[j] * E0 + (dy * Y1[j] + Y1[j + 1]) * E1 + Y0[j + 2] * E2) * X0[i] * Z0[k];
        g[8][c] = (Y1[j] * Z1[k] * E0 + (Y1[j] * Z0[k + 1] + Y0[j + 1] * Z1[k]) * E1 + Y0[j + 1] * Z0[k + 1] * E2) * X0[i];
        g[9][c] = (Z2[k] * E0 + (dz * Z1[k] + Z1[k + 1]) * E1 + Z0[k + 2] * E2) * X0[i] * Y0[j];
    }
#pragma unroll
    for (int d = 0; d < 10; d++) {
        if (SPH) {
            const double *T = trans + sh.trans;
#pragma unroll
            for (int s = 0; s < NB; s++) {
                double a = 0.0;
#pragma unroll
                for (int c = 0; c < NC; c++) a += T[c * NB + s] * g[d][c];
                out[d][s] = a;
            }
        } else {
#pragma unroll
            for (int s = 0; s < NB; s++) out[d][s] = g[d][s];
        }
    }
}

// Dispatch on the shell type; calls body(nBasis, out).
template <typename Body>
__device__ __forceinline__ void with_shell2(const xq::DevShell &sh, bool sph, const double *coef, const double *trans,
                                            double dx, double dy, double dz, Body body) {
    double out[10][16];
    const int key = sh.lLow * 4 + sh.lHigh;
    if (sph) {
        switch (key) {
            case 0: shell_eval2<0, 0, true>(sh, coef, trans, dx, dy, dz, out); body(1, out); break;
            case 5: shell_eval2<1, 1, true>(sh, coef, trans, dx, dy, dz, out); body(3, out); break;
            case 1: shell_eval2<0, 1, true>(sh, coef, trans, dx, dy, dz, out); body(4, out); break;
            case 10: shell_eval2<2, 2, true>(sh, coef, trans, dx, dy, dz, out); body(5, out); break;
            case 15: shell_eval2<3, 3, true>(sh, coef, trans, dx, dy, dz, out); body(7, out); break;
            default: break;
        }
    } else {
        switch (key) {
            case 0: shell_eval2<0, 0, false>(sh, coef, trans, dx, dy, dz, out); body(1, out); break;
            case 5: shell_eval2<1, 1, false>(sh, coef, trans, dx, dy, dz, out); body(3, out); break;
            case 1: shell_eval2<0, 1, false>(sh, coef, trans, dx, dy, dz, out); body(4, out); break;
            case 10: shell_eval2<2, 2, false>(sh, coef, trans, dx, dy, dz, out); body(6, out); break;
            case 15: shell_eval2<3, 3, false>(sh, coef, trans, dx, dy, dz, out); body(10, out); break;
            default: break;
        }
    }
}

// Values and first and second derivatives of the kept functions of blocks b0 + blockIdx.x into rows
// [vOffset - rowBase, + nKept) of the ten planes out[0..9] (row stride kBlockSize; zero past a block's points).
__global__ void __launch_bounds__(kBlockSize) basis_eval2_kernel(
    int b0, long long rowBase, const xq::DevBlock *__restrict__ blocks, const double *__restrict__ px,
    const double *__restrict__ py, const double *__restrict__ pz, const xq::DevShell *__restrict__ shells,
    const int *__restrict__ shellOf, const double *__restrict__ coef, const double *__restrict__ trans, int spherical,
    const int *__restrict__ idx, xq::Planes10 out) {
    const int b = b0 + static_cast<int>(blockIdx.x);
    const xq::DevBlock blk = blocks[b];
    const int t = threadIdx.x;
    const bool valid = t < blk.nPoints;
    const double x = valid ? px[blk.pStart + t] : 0.0, y = valid ? py[blk.pStart + t] : 0.0,
                 z = valid ? pz[blk.pStart + t] : 0.0;
    const int *kept = idx + blk.iOffset;
    // Work-group (b, h) writes the kept rows h nKept / S .. (h + 1) nKept / S - 1 (S = gridDim.y); a shell that
    // straddles two ranges is evaluated by both.
    const int S = static_cast<int>(gridDim.y), h = static_cast<int>(blockIdx.y);
    const int k1 = static_cast<int>((static_cast<long long>(h) + 1) * blk.nKept / S);
    int k = static_cast<int>(static_cast<long long>(h) * blk.nKept / S);
    while (k < k1) {
        const xq::DevShell sh = shells[shellOf[kept[k]]];
        with_shell2(sh, spherical != 0, coef, trans, x - sh.x, y - sh.y, z - sh.z, [&](int nB, const double (*v)[16]) {
            for (int j = 0; j < nB; j++) {
                if (k < k1 && kept[k] == sh.fStart + j) {
                    const size_t o = static_cast<size_t>(blk.vOffset - rowBase + k) * kBlockSize + t;
#pragma unroll
                    for (int d = 0; d < 10; d++) out.p[d][o] = valid ? v[d][j] : 0.0;
                    k++;
                }
            }
        });
    }
}

}  // namespace

void free_basis(xq_grid *g) {
    free_mixed(g);
    free_f64(g);
    dev_free(g, &g->d_shells);
    dev_free(g, &g->d_coef);
    dev_free(g, &g->d_trans);
    dev_free(g, &g->d_pow);
    dev_free(g, &g->d_idx);
    dev_free(g, &g->d_shellOf);
    g->h_shellOf.clear();
    g->h_funcAtom.clear();
    void **values[4] = {&g->d_f, &g->d_fx, &g->d_fy, &g->d_fz};
    for (void **a : values) {
        g->deviceBytes -= static_cast<long long>(cache_free(*a));
        *a = nullptr;
    }
    g->hasBasis = false;
}

int set_basis(xq_grid *g, int nSpec, const double *spec, int spherical, int precision) {
    if (precision != XQ_PRECISION_FP64 && precision != XQ_PRECISION_MIXED) {
        set_error("unknown precision mode");
        return kInvalidArgument;
    }
    // Parse the element specification.
    struct Element { int z, start, nShells; };
    std::vector<Element> elements;
    for (int p = 0; p < nSpec;) {
        if (p + 2 > nSpec) { set_error("truncated basis specification"); return kInvalidArgument; }
        Element e{static_cast<int>(spec[p]), p + 2, static_cast<int>(spec[p + 1])};
        p += 2;
        for (int s = 0; s < e.nShells; s++) {
            if (p + 3 > nSpec) { set_error("truncated basis specification"); return kInvalidArgument; }
            const int lLow = static_cast<int>(spec[p]), lHigh = static_cast<int>(spec[p + 1]), np = static_cast<int>(spec[p + 2]);
            if (lLow < 0 || lHigh < lLow || lHigh > kMaxL || np < 1 || np > kMaxPrim) {
                set_error("unsupported shell in basis specification");
                return kInvalidArgument;
            }
            const int key = lLow * 4 + lHigh;
            if (!(key == 0 || key == 5 || key == 1 || key == 10 || key == 15)) {
                set_error("unsupported shell type (supported: S, P, D, F, SP)");
                return kInvalidArgument;
            }
            p += 3 + np * (1 + lHigh - lLow + 1);
        }
        if (p > nSpec) { set_error("truncated basis specification"); return kInvalidArgument; }
        elements.push_back(e);
    }
    std::vector<DevShell> shells;
    std::vector<double> coef, trans;
    std::vector<int> shellOf;
    int f = 0;
    for (int a = 0; a < g->nAtoms; a++) {
        const Element *el = nullptr;
        for (const Element &e : elements) if (e.z == g->Z[a]) el = &e;
        if (el == nullptr) { set_error("no basis for an atom's element"); return kInvalidArgument; }
        int p = el->start;
        for (int s = 0; s < el->nShells; s++) {
            DevShell sh{};
            sh.x = g->xyz[3 * a]; sh.y = g->xyz[3 * a + 1]; sh.z = g->xyz[3 * a + 2];
            sh.atom = a;
            sh.lLow = static_cast<int>(spec[p]); sh.lHigh = static_cast<int>(spec[p + 1]); sh.nPrim = static_cast<int>(spec[p + 2]);
            p += 3;
            sh.nCBF = 0; sh.nBasis = 0;
            for (int l = sh.lLow; l <= sh.lHigh; l++) { sh.nCBF += ncart(l); sh.nBasis += 2 * l + 1; }
            if (!spherical) sh.nBasis = sh.nCBF;
            sh.fStart = f;
            sh.coef = static_cast<int>(coef.size());
            sh.pow = 0;
            std::vector<double> ex(sh.nPrim), cc(static_cast<size_t>(sh.nPrim) * sh.nCBF);
            for (int q = 0; q < sh.nPrim; q++) {
                const double alpha = spec[p++];
                ex[q] = alpha;
                for (int l = sh.lLow, c = 0; l <= sh.lHigh; l++) {
                    // Unnormalise (GaussianBasis.c:728-751).
                    const double e2 = 2.0 * alpha;
                    double fac = std::pow(M_PI, 1.5) / (e2 * std::sqrt(e2));
                    for (int i = 1; i <= l; i++) fac *= static_cast<double>(2 * i - 1) / (2.0 * e2);
                    const double k = spec[p++] / std::sqrt(fac);
                    // Cartesian factors (GaussianBasis.c:448-501).
                    const int off = cart_offset(l);
                    for (int i = 0; i < ncart(l); i++, c++) {
                        const double m = std::sqrt(odd_fact(2 * l - 1) / (odd_fact(2 * kCart.x[off + i] - 1) *
                                                                         odd_fact(2 * kCart.y[off + i] - 1) *
                                                                         odd_fact(2 * kCart.z[off + i] - 1)));
                        cc[static_cast<size_t>(q) * sh.nCBF + c] = k * m;
                    }
                }
            }
            coef.insert(coef.end(), ex.begin(), ex.end());
            coef.insert(coef.end(), cc.begin(), cc.end());
            double K = 1.0;
            if (spherical) {
                sh.trans = static_cast<int>(trans.size());
                std::vector<double> T(static_cast<size_t>(sh.nCBF) * sh.nBasis, 0.0);
                for (int l = sh.lLow, r0 = 0, c0 = 0; l <= sh.lHigh; l++) {
                    c2s_block(l, T.data(), sh.nBasis, r0, c0);
                    r0 += ncart(l);
                    c0 += 2 * l + 1;
                }
                K = 0.0;
                for (int sI = 0; sI < sh.nBasis; sI++) {
                    double col = 0.0;
                    for (int c = 0; c < sh.nCBF; c++) col += std::fabs(T[static_cast<size_t>(c) * sh.nBasis + sI]);
                    K = std::max(K, col);
                }
                trans.insert(trans.end(), T.begin(), T.end());
            } else sh.trans = -1;
            for (int q = 0; q < sh.nPrim; q++) {
                double amax = 0.0;
                for (int c = 0; c < sh.nCBF; c++) amax = std::max(amax, std::fabs(cc[static_cast<size_t>(q) * sh.nCBF + c]));
                sh.bound[q] = amax * K;
            }
            for (int j = 0; j < sh.nBasis; j++) shellOf.push_back(static_cast<int>(shells.size()));
            f += sh.nBasis;
            shells.push_back(sh);
        }
    }
    if (shells.size() > static_cast<size_t>(kMaxShellsLDS)) { set_error("too many shells"); return kInvalidArgument; }
    free_basis(g);
    free_work(g);
    g->nFunctions = f;
    g->nShells = static_cast<int>(shells.size());
    g->spherical = spherical ? 1 : 0;
    g->precision = precision;
    hipStream_t s = g->stream;
    XQ_TRY(dev_alloc(g, &g->d_shells, shells.size()));
    XQ_TRY(dev_alloc(g, &g->d_coef, coef.size()));
    XQ_TRY(dev_alloc(g, &g->d_trans, trans.size()));
    XQ_CHECK(hipMemcpyAsync(g->d_shells, shells.data(), shells.size() * sizeof(DevShell), hipMemcpyHostToDevice, s));
    XQ_CHECK(hipMemcpyAsync(g->d_coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, s));
    if (!trans.empty())
        XQ_CHECK(hipMemcpyAsync(g->d_trans, trans.data(), trans.size() * sizeof(double), hipMemcpyHostToDevice, s));
    // Pass 1: kept functions per block.
    int *tmpIdx = nullptr, *counts = nullptr, *d_shellOf = nullptr;
    int status = kOK;
    do {
        if ((status = dev_alloc(g, &tmpIdx, static_cast<size_t>(g->nBlocks) * f)) != kOK) break;
        if ((status = dev_alloc(g, &counts, g->nBlocks)) != kOK) break;
        if ((status = dev_alloc(g, &d_shellOf, shellOf.size())) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_shellOf, shellOf.data(), shellOf.size() * sizeof(int), hipMemcpyHostToDevice, s),
                            "copy shellOf")) != kOK) break;
        if (g->nBlocks > 0) {
            hipLaunchKernelGGL(basis_count_kernel, dim3(g->nBlocks), dim3(kBlockSize), 0, s, g->nBlocks, g->d_blocks, g->d_x,
                               g->d_y, g->d_z, g->nShells, g->d_shells, g->d_coef, g->d_trans, g->spherical, f, g->bfTol,
                               tmpIdx, counts);
            if ((status = check(hipGetLastError(), "basis_count_kernel")) != kOK) break;
        }
        std::vector<int> h_counts(g->nBlocks > 0 ? g->nBlocks : 1, 0);
        // Wait for the kernels first: a pageable copy that waits for its stream holds the runtime's staging path.
        if ((status = check(hipStreamSynchronize(s), "basis sync")) != kOK) break;
        if (g->nBlocks > 0 &&
            (status = check(hipMemcpyAsync(h_counts.data(), counts, g->nBlocks * sizeof(int), hipMemcpyDeviceToHost, s),
                            "copy counts")) != kOK) break;
        if ((status = check(hipStreamSynchronize(s), "basis sync")) != kOK) break;
        long long rows = 0;
        double bfv = 0.0, bfs = 0.0, bfp = 0.0;
        for (int b = 0; b < g->nBlocks; b++) {
            bfs += static_cast<double>(h_counts[b]) * h_counts[b] * g->blocks[b].nPoints;
            const double padded = 32.0 * ((h_counts[b] + 31) / 32);
            bfp += padded * padded * kBlockSize;
            g->blocks[b].nKept = h_counts[b];
            g->blocks[b].vOffset = rows;
            g->blocks[b].iOffset = static_cast<int>(rows);
            rows += h_counts[b];
            bfv += static_cast<double>(h_counts[b]) * g->blocks[b].nPoints;
        }
        g->nValueRows = rows;
        g->bfValues = bfv;
        g->bfSquares = bfs;
        g->bfSquaresPadded = bfp;
        if ((status = check(hipMemcpyAsync(g->d_blocks, g->blocks.data(), g->nBlocks * sizeof(DevBlock), hipMemcpyHostToDevice, s),
                            "copy blocks")) != kOK) break;
        const size_t nVal = static_cast<size_t>(rows > 0 ? rows : 1) * kBlockSize;
        const size_t elt = sizeof(double);
        if ((status = dev_alloc(g, &g->d_idx, static_cast<size_t>(rows > 0 ? rows : 1))) != kOK) break;
        void **arrays[4] = {&g->d_f, &g->d_fx, &g->d_fy, &g->d_fz};
        for (void **a : arrays) {
            hipError_t e = cache_alloc(a, nVal * elt);
            if (e != hipSuccess) { set_error("hipMalloc failed for basis values"); status = kOutOfMemory; break; }
            g->deviceBytes += static_cast<long long>(nVal * elt);
        }
        if (status != kOK) break;
        if (g->nBlocks > 0) {
            hipLaunchKernelGGL(basis_eval_kernel<double>, dim3(g->nBlocks), dim3(kBlockSize), 0, s, g->d_blocks, g->d_x,
                               g->d_y, g->d_z, g->d_shells, d_shellOf, g->d_coef, g->d_trans, g->spherical, f, tmpIdx,
                               g->d_idx, static_cast<double *>(g->d_f), static_cast<double *>(g->d_fx),
                               static_cast<double *>(g->d_fy), static_cast<double *>(g->d_fz));
            if ((status = check(hipGetLastError(), "basis_eval_kernel")) != kOK) break;
        }
        // The operand-order planes of the chosen precision replace the FP64 rows.
        if (precision == XQ_PRECISION_MIXED)
            status = prepare_mixed(g, static_cast<double *>(g->d_f), static_cast<double *>(g->d_fx),
                                   static_cast<double *>(g->d_fy), static_cast<double *>(g->d_fz));
        else
            status = prepare_f64(g, static_cast<double *>(g->d_f), static_cast<double *>(g->d_fx),
                                 static_cast<double *>(g->d_fy), static_cast<double *>(g->d_fz));
        if (status != kOK) break;
        if ((status = check(hipStreamSynchronize(s), "basis sync")) != kOK) break;
        for (void **a : arrays) {
            g->deviceBytes -= static_cast<long long>(cache_free(*a));
            *a = nullptr;
        }
        status = check(hipStreamSynchronize(s), "basis sync");
    } while (false);
    (void)hipStreamSynchronize(s);   // the kernels that use the scratch arrays are done before they are reused
    dev_free(g, &tmpIdx);
    dev_free(g, &counts);
    dev_free(g, &d_shellOf);
    if (status != kOK) { free_basis(g); return status; }
    g->h_shellOf = shellOf;     // kept for the second-derivative evaluation of the XC gradients
    g->h_funcAtom.resize(shellOf.size());
    for (size_t fn = 0; fn < shellOf.size(); fn++) g->h_funcAtom[fn] = shells[shellOf[fn]].atom;
    g->hasBasis = true;
    return kOK;
}


// Basis values, gradients and second derivatives of blocks [b0, b1) into the planes (rows from the blocks' value-row offsets
// less rowBase), for the XC gradients.
int eval_basis2(xq_grid *g, int b0, int b1, long long rowBase, const Planes10 &out, hipStream_t s) {
    if (b1 <= b0) return kOK;
    if (g->d_shellOf == nullptr) {
        XQ_TRY(dev_alloc(g, &g->d_shellOf, g->h_shellOf.size()));
        XQ_CHECK(hipMemcpyAsync(g->d_shellOf, g->h_shellOf.data(), g->h_shellOf.size() * sizeof(int), hipMemcpyHostToDevice,
                                g->stream));
        XQ_CHECK(hipStreamSynchronize(g->stream));
    }
    // Four work-groups per block, each a quarter of its kept rows: the blocks' rows are evaluated one shell after another,
    // so a launch would otherwise last as long as its largest block.
    hipLaunchKernelGGL(basis_eval2_kernel, dim3(b1 - b0, 4), dim3(kBlockSize), 0, s, b0, rowBase, g->d_blocks, g->d_x, g->d_y,
                       g->d_z, g->d_shells, g->d_shellOf, g->d_coef, g->d_trans, g->spherical, g->d_idx, out);
    XQ_CHECK(hipGetLastError());
    return kOK;
}

}  // namespace xq
