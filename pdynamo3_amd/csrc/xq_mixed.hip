// Mixed-precision quadrature: both contractions on int8 matrix cores with exact integer accumulation (Ozaki-style
// slicing), FP64 everywhere else.
//
// Representation.  A value x in a group that shares the power of two 2^e (|x| < 2^e) is cut into signed 7-bit digits:
//     y = x 2^-e;  for k = 0..D-1: y *= 128, d_k = trunc(y), y -= d_k      (d_k in [-127, 127], |remainder| < 2^-7D)
//     x = 2^e sum_k d_k 2^(-7(k+1)) + r,   |r| < 2^(e - 7D).
// (The last digit is rounded to nearest instead, see slice().)  Every step is exact in FP64.  A product of two sliced groups is sum_{k,l} 2^(-7(k+l+2)) sum d_k d_l; the inner sums are
// int8 x int8 GEMMs with int32 accumulation (v_mfma_i32_32x32x32_i8), exact for the chunk lengths used here, and terms with
// the same k + l share one int32 accumulator.  Only the representation error r and the dropped products (k + l above the
// kept order) differ from FP64, both bounded relative to the group's power of two.
//
//   X = P_red Phi (density):     P rows (one power per row of P, per SCF step), 5 digits; Phi per (point, chunk of 32
//                                kept functions), 4 digits, stored; kept products k + l <= 4.  Error ~2^-28 of the largest
//                                term per point, ~2^-35 relative on P (P's rounding would be systematic across the grid).
//   V += Phi W^T + W Phi^T:      Phi per (function, chunk of 32 points), 4 digits, stored; W per (function, chunk of 32
//                                points), 4 digits, formed per step; kept products k + l <= 3.
// The chunk partial sums are promoted to FP64 with their powers of two and summed in FP64.  The density epilogue uses Phi
// rebuilt from its digits (28 bits) and grad Phi stored in FP32; the functional, the weights and every reduction are FP64.
//
// MFMA lane layout (32 x 32 x 32, int8): lane l supplies A[l % 32][16 (l / 32) .. + 15] and B[16 (l / 32) .. + 15][l % 32]
// (16 consecutive K bytes each) and receives D[(v % 4) + 4 (l / 32) + 8 (v / 4)][l % 32] for v = 0..15.
#include <hip/hip_runtime.h>

#include "xq_functional.h"
#include "xq_internal.h"

namespace xq {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

namespace {

constexpr int kC = 32;          // chunk length (functions or points) and MFMA tile edge
constexpr int kDPhi = 4;        // digits of Phi and W
constexpr int kDP = 5;          // digits of P

__device__ __forceinline__ v16i mfma_i8(const v4i &a, const v4i &b, const v16i &c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// Power of two bounding max|x| (frexp exponent); 0 for an all-zero group.
__device__ __forceinline__ int group_exponent(double amax) {
    int e = 0;
    if (amax > 0.0) (void)frexp(amax, &e);
    return e;
}

// Stored exponents are int8: a group whose largest magnitude is below 2^-100 (contributions below 1e-30 of unit-sized
// terms) is stored as zero digits with exponent 0.
constexpr int kMinStoredExponent = -100;

// Digits 0..D-2 truncate (the remainder keeps x's sign and stays below one unit); the last digit rounds to nearest, so the
// representation error is unbiased and at most half a unit of the last digit (clamped at +-127 in the rare case
// |y| >= 127.5, which the first D-1 truncations make possible only for the last digit).
template <int D>
__device__ __forceinline__ void slice(double x, int e, signed char *d) {
    double y = ldexp(x, -e);
#pragma unroll
    for (int k = 0; k < D - 1; k++) {
        y *= 128.0;
        const double t = trunc(y);
        d[k] = static_cast<signed char>(t);
        y -= t;
    }
    const double t = fmin(fmax(rint(y * 128.0), -127.0), 127.0);
    d[D - 1] = static_cast<signed char>(t);
}

template <int D>
__device__ __forceinline__ double unslice(const signed char *d, int e) {
    double s = 0.0, w = 1.0 / 128.0;
#pragma unroll
    for (int k = 0; k < D; k++) { s += w * d[k]; w *= 1.0 / 128.0; }
    return ldexp(s, e);
}

// ------------------------------------------------------------------------------------------------------------------------
// Set-up, once per geometry: FP64 values of one block -> the two digit layouts and FP32 gradients.
//   D1  [row][chunk q (4)][digit (4)][32 points]   exponents E1 [row][q]       (rows of the block; per function & chunk)
//   D2  [chunk c][digit (4)][point (128)][32 k]     exponents E2 [c][point]     (per point & chunk of 32 kept functions)
//   G   [3][row][128] FP32
// Offsets: rows start at the block's value-row offset R (D1, E1 at 512 R bytes and 4 R; G at 128 R floats); chunks c of
// block b start at chunk offset C_b = sum over earlier blocks of ceil(nKept / 32) (D2 at 16384 C_b bytes, E2 at 128 C_b).
// ------------------------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlockSize) mixed_prepare_kernel(
    const xq::DevBlock *__restrict__ blocks, const int *__restrict__ chunkOffset, const double *__restrict__ f,
    const double *__restrict__ fx, const double *__restrict__ fy, const double *__restrict__ fz,
    signed char *__restrict__ D1, signed char *__restrict__ E1, signed char *__restrict__ D2, signed char *__restrict__ E2,
    float *__restrict__ G, size_t rowsTotal) {
    const int b = blockIdx.x;
    const xq::DevBlock blk = blocks[b];
    const int nb = blk.nKept, t = threadIdx.x;   // t = point
    const size_t R = static_cast<size_t>(blk.vOffset);
    const size_t C = static_cast<size_t>(chunkOffset[b]);
    // Gradients in FP32.
    for (int r = 0; r < nb; r++) {
        const size_t o = (R + r) * kBlockSize + t;
        G[o] = static_cast<float>(fx[o]);
        G[rowsTotal * kBlockSize + o] = static_cast<float>(fy[o]);
        G[2 * rowsTotal * kBlockSize + o] = static_cast<float>(fz[o]);
    }
    // D1/E1: per (row, chunk of 32 points).  Thread t handles point t; the chunk max is a reduction over 32 lanes.
    const int q = t >> 5, pk = t & 31;
    for (int r = 0; r < nb; r++) {
        const double x = f[(R + r) * kBlockSize + t];
        double m = fabs(x);
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 32));
        int e = group_exponent(m);
        signed char d[kDPhi] = {0, 0, 0, 0};
        if (e < kMinStoredExponent) e = 0;
        else slice<kDPhi>(x, e, d);
        const size_t base = ((R + r) * 4 + q) * (kDPhi * kC);
#pragma unroll
        for (int k = 0; k < kDPhi; k++) D1[base + k * kC + pk] = d[k];
        if (pk == 0) E1[(R + r) * 4 + q] = static_cast<signed char>(e);
    }
    // D2/E2: per (point, chunk of 32 kept functions).
    const int nChunks = (nb + kC - 1) / kC;
    for (int c = 0; c < nChunks; c++) {
        double m = 0.0;
        for (int k = 0; k < kC; k++) {
            const int r = c * kC + k;
            if (r < nb) m = fmax(m, fabs(f[(R + r) * kBlockSize + t]));
        }
        int e = group_exponent(m);
        const bool tiny = e < kMinStoredExponent;
        if (tiny) e = 0;
        E2[(C + c) * kBlockSize + t] = static_cast<signed char>(e);
        for (int k = 0; k < kC; k++) {
            const int r = c * kC + k;
            signed char d[kDPhi] = {0, 0, 0, 0};
            if (r < nb && !tiny) slice<kDPhi>(f[(R + r) * kBlockSize + t], e, d);
#pragma unroll
            for (int j = 0; j < kDPhi; j++) D2[(((C + c) * kDPhi + j) * kBlockSize + t) * kC + k] = d[j];
        }
    }
}

// Row exponents of P (power of two above max_k |P_ik|), once per step.
__global__ void p_row_exponent_kernel(int n, const double *__restrict__ P, int *__restrict__ eP) {
    const int i = blockIdx.x;
    double m = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) m = fmax(m, fabs(P[static_cast<size_t>(i) * n + k]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    __shared__ double s[4];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) eP[i] = group_exponent(fmax(fmax(s[0], s[1]), fmax(s[2], s[3])));
}

// ------------------------------------------------------------------------------------------------------------------------
// Density kernel (per block, 256 threads = 4 waves): X = P_red Phi on int8 MFMA, rho and grad rho, the FP64 functional,
// E_xc / N_e block sums; writes rho, grad rho and (w vrho, w vsigma) per point for the Fock stage.
// Wave wv owns the 32 points 32 wv .. 32 wv + 31; the 32 x 32 P_red tile of each (row chunk I, column chunk c) is shared
// through LDS, and the next tile's P entries are fetched into registers while the current tile is multiplied.
// ------------------------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) density_kernel_mixed(
    const xq::DevBlock *__restrict__ blocks, const int *__restrict__ chunkOffset, const int *__restrict__ idxAll,
    const signed char *__restrict__ D2, const signed char *__restrict__ E2, const float *__restrict__ G, size_t rowsTotal,
    int n, const double *__restrict__ P, const int *__restrict__ eP, const double *__restrict__ w, FunctionalSet fs,
    double rhoTol, double *__restrict__ rhoOut, double *__restrict__ vrs, double *__restrict__ blockEN,
    int *__restrict__ skip) {
    __shared__ __attribute__((aligned(16))) signed char sA[kDP][kC][kC];   // P_red tile digits [digit][row][k]
    __shared__ int sPe[kC];
    __shared__ double rsum[4][kBlockSize];
    __shared__ double wsum[4][2];
    const int b = blockIdx.x;
    const xq::DevBlock blk = blocks[b];
    const int nb = blk.nKept, np = blk.nPoints;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    if (nb <= 0) {
        if (t == 0) { blockEN[2 * b] = 0.0; blockEN[2 * b + 1] = 0.0; skip[b] = 1; }
        return;
    }
    const int *idx = idxAll + blk.iOffset;
    const size_t R = static_cast<size_t>(blk.vOffset);
    const size_t C = static_cast<size_t>(chunkOffset[b]);
    const int nChunks = (nb + kC - 1) / kC;
    const size_t plane = rowsTotal * kBlockSize;
    const int p = 32 * wv + lr;          // this lane's point (its D column)
    // Tile entries this thread stages: rows r = e >> 5, columns k = e & 31 for e = t + 256 q.
    auto fetch = [&](int I, int c, double *pv) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = t + 256 * q, gr = I * kC + (e >> 5), gk = c * kC + (e & 31);
            pv[q] = (gr < nb && gk < nb) ? P[static_cast<size_t>(idx[gr]) * n + idx[gk]] : 0.0;
        }
    };
    double pr = 0.0, pgx = 0.0, pgy = 0.0, pgz = 0.0;
    double pnext[4];
    fetch(0, 0, pnext);
    for (int I = 0; I < nChunks; I++) {
        double X[16];
#pragma unroll
        for (int v = 0; v < 16; v++) X[v] = 0.0;
        __syncthreads();
        if (t < kC) {
            const int r = I * kC + t;
            sPe[t] = r < nb ? eP[idx[r]] : 0;
        }
        for (int c = 0; c < nChunks; c++) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int e = t + 256 * q, r = e >> 5, k = e & 31;
                signed char d[kDP] = {0, 0, 0, 0, 0};
                if (pnext[q] != 0.0) slice<kDP>(pnext[q], sPe[r], d);
#pragma unroll
                for (int j = 0; j < kDP; j++) sA[j][r][k] = d[j];
            }
            __syncthreads();
            // Prefetch the next tile's P entries while this one is multiplied.
            const int cn = c + 1 < nChunks ? c + 1 : 0, In = c + 1 < nChunks ? I : I + 1;
            if (In < nChunks) fetch(In, cn, pnext);
            v4i a[kDP], bq[kDPhi];
#pragma unroll
            for (int j = 0; j < kDP; j++) a[j] = *reinterpret_cast<const v4i *>(&sA[j][lr][16 * lh]);
#pragma unroll
            for (int j = 0; j < kDPhi; j++)
                bq[j] = *reinterpret_cast<const v4i *>(&D2[(((C + c) * kDPhi + j) * kBlockSize + p) * kC + 16 * lh]);
            v16i acc[5];
#pragma unroll
            for (int s = 0; s < 5; s++) acc[s] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < kDP; i++)
#pragma unroll
                for (int j = 0; j < kDPhi; j++)
                    if (i + j <= 4) acc[i + j] = mfma_i8(a[i], bq[j], acc[i + j]);
            const int ePhi = E2[(C + c) * kBlockSize + p];
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int rl = (v & 3) + 4 * lh + 8 * (v >> 2);
                double s = 0.0;
#pragma unroll
                for (int k = 4; k >= 0; k--) s = s * (1.0 / 128.0) + static_cast<double>(acc[k][v]);
                X[v] += ldexp(s, sPe[rl] + ePhi - 14);
            }
        }
        // Epilogue: rho += Phi X, grad rho += grad Phi X (Phi rebuilt from the point-scaled digits).
        const int ePhiI = E2[(C + I) * kBlockSize + p];
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int rl = (v & 3) + 4 * lh + 8 * (v >> 2), r = I * kC + rl;
            if (r < nb) {
                signed char d[kDPhi];
#pragma unroll
                for (int j = 0; j < kDPhi; j++) d[j] = D2[(((C + I) * kDPhi + j) * kBlockSize + p) * kC + rl];
                const double phi = unslice<kDPhi>(d, ePhiI);
                const size_t o = (R + r) * kBlockSize + p;
                pr += phi * X[v];
                pgx += static_cast<double>(G[o]) * X[v];
                pgy += static_cast<double>(G[plane + o]) * X[v];
                pgz += static_cast<double>(G[2 * plane + o]) * X[v];
            }
        }
    }
    pr += __shfl_xor(pr, 32, 64);
    pgx += __shfl_xor(pgx, 32, 64);
    pgy += __shfl_xor(pgy, 32, 64);
    pgz += __shfl_xor(pgz, 32, 64);
    if (lh == 0) { rsum[0][p] = pr; rsum[1][p] = pgx; rsum[2][p] = pgy; rsum[3][p] = pgz; }
    __syncthreads();
    // Functional, E and N.
    double rho = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    if (t < np) { rho = rsum[0][t]; gx = 2.0 * rsum[1][t]; gy = 2.0 * rsum[2][t]; gz = 2.0 * rsum[3][t]; }
    double mx = t < np ? fabs(rho) : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) wsum[wv][0] = mx;
    __syncthreads();
    const bool skipped = !(fmax(fmax(wsum[0][0], wsum[1][0]), fmax(wsum[2][0], wsum[3][0])) > rhoTol);
    __syncthreads();
    double e = 0.0, ne = 0.0;
    if (t < np) {
        double vr = 0.0, vs = 0.0;
        if (!skipped) {
            double eps;
            evaluate_functional(fs, rho, gx * gx + gy * gy + gz * gz, eps, vr, vs);
            const double ww = w[blk.pStart + t];
            vr *= ww;
            vs *= ww;
            e = eps * rho * ww;
            ne = rho * ww;
        }
        const size_t pp = static_cast<size_t>(blk.pStart) + t;
        rhoOut[4 * pp] = rho;
        rhoOut[4 * pp + 1] = gx;
        rhoOut[4 * pp + 2] = gy;
        rhoOut[4 * pp + 3] = gz;
        vrs[2 * pp] = vr;
        vrs[2 * pp + 1] = vs;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { e += __shfl_xor(e, o, 64); ne += __shfl_xor(ne, o, 64); }
    if (lane == 0) { wsum[wv][0] = e; wsum[wv][1] = ne; }
    __syncthreads();
    if (t == 0) {
        blockEN[2 * b] = ((wsum[0][0] + wsum[1][0]) + wsum[2][0]) + wsum[3][0];
        blockEN[2 * b + 1] = ((wsum[0][1] + wsum[1][1]) + wsum[2][1]) + wsum[3][1];
        skip[b] = skipped ? 1 : 0;
    }
}

// ------------------------------------------------------------------------------------------------------------------------
// W digits (per block, 128 threads = points): W_rp = (w vrho / 2) Phi_rp + 2 w vsigma (grad rho . grad Phi_rp), cut into
// 4 digits per (row, chunk of 32 points), in D1's layout (DW, EW).  Phi comes from D1 (28 bits), grad Phi from G.
// ------------------------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlockSize) w_digits_kernel(
    const xq::DevBlock *__restrict__ blocks, const signed char *__restrict__ D1, const signed char *__restrict__ E1,
    const float *__restrict__ G, size_t rowsTotal, const double *__restrict__ rhoIn, const double *__restrict__ vrs,
    const int *__restrict__ skip, signed char *__restrict__ DW, signed char *__restrict__ EW) {
    const int b = blockIdx.x;
    if (skip[b]) return;
    const xq::DevBlock blk = blocks[b];
    const int nb = blk.nKept, np = blk.nPoints, t = threadIdx.x;
    const size_t R = static_cast<size_t>(blk.vOffset), plane = rowsTotal * kBlockSize;
    const int q = t >> 5, pk = t & 31;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
    if (t < np) {
        const size_t pp = static_cast<size_t>(blk.pStart) + t;
        const double s2 = 2.0 * vrs[2 * pp + 1];
        c0 = 0.5 * vrs[2 * pp];
        c1 = s2 * rhoIn[4 * pp + 1];
        c2 = s2 * rhoIn[4 * pp + 2];
        c3 = s2 * rhoIn[4 * pp + 3];
    }
    for (int r = 0; r < nb; r++) {
        const size_t base = ((R + r) * 4 + q) * (kDPhi * kC);
        signed char d[kDPhi];
#pragma unroll
        for (int k = 0; k < kDPhi; k++) d[k] = D1[base + k * kC + pk];
        const double phi = unslice<kDPhi>(d, E1[(R + r) * 4 + q]);
        const size_t o = (R + r) * kBlockSize + t;
        const double x = c0 * phi + c1 * static_cast<double>(G[o]) + c2 * static_cast<double>(G[plane + o]) +
                         c3 * static_cast<double>(G[2 * plane + o]);
        double m = fabs(x);
#pragma unroll
        for (int o2 = 16; o2 > 0; o2 >>= 1) m = fmax(m, __shfl_xor(m, o2, 32));
        int e = group_exponent(m);
        signed char dw[kDPhi] = {0, 0, 0, 0};
        if (e < kMinStoredExponent) e = 0;
        else slice<kDPhi>(x, e, dw);
#pragma unroll
        for (int k = 0; k < kDPhi; k++) DW[base + k * kC + pk] = dw[k];
        if (pk == 0) EW[(R + r) * 4 + q] = static_cast<signed char>(e);
    }
}

// ------------------------------------------------------------------------------------------------------------------------
// Fock kernel (per block, 256 threads = 4 waves): V += Phi W^T + W Phi^T on lower-triangle 32 x 32 tiles (I, J), J <= I.
// The four waves share a tile and take one 32-point chunk each (q = wv).  Lane (lr, lh) loads, for rows I*32 + lr and
// J*32 + lr, the 16 points 32 q + 16 lh .. + 15 of the Phi and W digits: its A operands (rows I) and B operands (rows J).
// Partial tiles are summed over the waves in LDS and added to V with FP64 atomics.
// ------------------------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) fock_kernel_mixed(
    const xq::DevBlock *__restrict__ blocks, const int *__restrict__ idxAll, const signed char *__restrict__ D1,
    const signed char *__restrict__ E1, const signed char *__restrict__ DW, const signed char *__restrict__ EW, int n,
    const int *__restrict__ skip, double *__restrict__ V) {
    __shared__ double red[4][kC][kC + 1];
    const int b = blockIdx.x;
    if (skip[b]) return;
    const xq::DevBlock blk = blocks[b];
    const int nb = blk.nKept;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int *idx = idxAll + blk.iOffset;
    const size_t R = static_cast<size_t>(blk.vOffset);
    const int nChunks = (nb + kC - 1) / kC;
    const int q = wv;
    for (int I = 0; I < nChunks; I++) {
        for (int J = 0; J <= I; J++) {
            v4i fd[2][kDPhi], wd[2][kDPhi];
            int eF[2], eW[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int rr = (h == 0 ? I : J) * kC + lr;
                if (rr < nb) {
                    const size_t base = ((R + rr) * 4 + q) * (kDPhi * kC) + 16 * lh;
#pragma unroll
                    for (int k = 0; k < kDPhi; k++) {
                        fd[h][k] = *reinterpret_cast<const v4i *>(&D1[base + k * kC]);
                        wd[h][k] = *reinterpret_cast<const v4i *>(&DW[base + k * kC]);
                    }
                    eF[h] = E1[(R + rr) * 4 + q];
                    eW[h] = EW[(R + rr) * 4 + q];
                } else {
#pragma unroll
                    for (int k = 0; k < kDPhi; k++) { fd[h][k] = v4i{0, 0, 0, 0}; wd[h][k] = v4i{0, 0, 0, 0}; }
                    eF[h] = 0;
                    eW[h] = 0;
                }
            }
            v16i a1[4], a2[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                a1[s] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
                a2[s] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            }
#pragma unroll
            for (int i = 0; i < kDPhi; i++)
#pragma unroll
                for (int j = 0; j < kDPhi; j++)
                    if (i + j <= 3) {
                        a1[i + j] = mfma_i8(fd[0][i], wd[1][j], a1[i + j]);   // Phi_I W_J^T
                        a2[i + j] = mfma_i8(wd[0][i], fd[1][j], a2[i + j]);   // W_I Phi_J^T
                    }
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int rl = (v & 3) + 4 * lh + 8 * (v >> 2);
                const int eFI = __shfl(eF[0], rl, 64), eWI = __shfl(eW[0], rl, 64);
                double s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int k = 3; k >= 0; k--) {
                    s1 = s1 * (1.0 / 128.0) + static_cast<double>(a1[k][v]);
                    s2 = s2 * (1.0 / 128.0) + static_cast<double>(a2[k][v]);
                }
                red[wv][rl][lr] = ldexp(s1, eFI + eW[1] - 14) + ldexp(s2, eWI + eF[1] - 14);
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int el = t + 256 * u, i = el >> 5, j = el & 31;
                const int gi = I * kC + i, gj = J * kC + j;
                if (gi < nb && gj < nb && gj <= gi) {
                    const double x = ((red[0][i][j] + red[1][i][j]) + red[2][i][j]) + red[3][i][j];
                    atomicAdd(&V[static_cast<size_t>(idx[gi]) * n + idx[gj]], x);
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace

// Host side: digit storage at basis set-up and the per-step launch.
int prepare_mixed(xq_grid *g, const double *f, const double *fx, const double *fy, const double *fz) {
    std::vector<int> chunkOffset(g->nBlocks + 1, 0);
    long long chunks = 0;
    for (int b = 0; b < g->nBlocks; b++) {
        chunkOffset[b] = static_cast<int>(chunks);
        chunks += (g->blocks[b].nKept + kC - 1) / kC;
    }
    chunkOffset[g->nBlocks] = static_cast<int>(chunks);
    g->nChunks = chunks;
    const size_t rows = static_cast<size_t>(g->nValueRows > 0 ? g->nValueRows : 1);
    XQ_TRY(dev_alloc(g, &g->d_chunkOffset, chunkOffset.size()));
    XQ_TRY(dev_alloc(g, &g->d_D1, rows * 512));
    XQ_TRY(dev_alloc(g, &g->d_E1, rows * 4));
    XQ_TRY(dev_alloc(g, &g->d_D2, static_cast<size_t>(chunks > 0 ? chunks : 1) * 16384));
    XQ_TRY(dev_alloc(g, &g->d_E2, static_cast<size_t>(chunks > 0 ? chunks : 1) * 128));
    XQ_TRY(dev_alloc(g, &g->d_G, 3 * rows * kBlockSize));
    XQ_TRY(dev_alloc(g, &g->d_DW, rows * 512));
    XQ_TRY(dev_alloc(g, &g->d_EW, rows * 4));
    XQ_CHECK(hipMemcpyAsync(g->d_chunkOffset, chunkOffset.data(), chunkOffset.size() * sizeof(int), hipMemcpyHostToDevice,
                            g->stream));
    if (g->nBlocks > 0)
        hipLaunchKernelGGL(mixed_prepare_kernel, dim3(g->nBlocks), dim3(kBlockSize), 0, g->stream, g->d_blocks,
                           g->d_chunkOffset, f, fx, fy, fz, g->d_D1, g->d_E1, g->d_D2, g->d_E2, g->d_G,
                           static_cast<size_t>(g->nValueRows));
    XQ_CHECK(hipGetLastError());
    return kOK;
}

void free_mixed(xq_grid *g) {
    dev_free(&g->d_chunkOffset);
    dev_free(&g->d_D1);
    dev_free(&g->d_E1);
    dev_free(&g->d_D2);
    dev_free(&g->d_E2);
    dev_free(&g->d_G);
    dev_free(&g->d_DW);
    dev_free(&g->d_EW);
    dev_free(&g->d_eP);
}

int launch_fused_mixed(xq_grid *g, const FunctionalSet &fs) {
    if (g->d_eP == nullptr) XQ_TRY(dev_alloc(g, &g->d_eP, static_cast<size_t>(g->nFunctions)));
    hipStream_t s = g->stream;
    const size_t rows = static_cast<size_t>(g->nValueRows);
    hipLaunchKernelGGL(p_row_exponent_kernel, dim3(g->nFunctions), dim3(256), 0, s, g->nFunctions, g->d_P, g->d_eP);
    (void)hipEventRecord(g->evRho0, s);
    hipLaunchKernelGGL(density_kernel_mixed, dim3(g->nBlocks), dim3(256), 0, s, g->d_blocks, g->d_chunkOffset, g->d_idx,
                       g->d_D2, g->d_E2, g->d_G, rows, g->nFunctions, g->d_P, g->d_eP, g->d_w, fs, g->rhoTol, g->d_rho,
                       g->d_vrs, g->d_blockEN, g->d_skip);
    (void)hipEventRecord(g->evRho1, s);
    hipLaunchKernelGGL(w_digits_kernel, dim3(g->nBlocks), dim3(kBlockSize), 0, s, g->d_blocks, g->d_D1, g->d_E1, g->d_G,
                       rows, g->d_rho, g->d_vrs, g->d_skip, g->d_DW, g->d_EW);
    hipLaunchKernelGGL(fock_kernel_mixed, dim3(g->nBlocks), dim3(256), 0, s, g->d_blocks, g->d_idx, g->d_D1, g->d_E1,
                       g->d_DW, g->d_EW, g->nFunctions, g->d_skip, g->d_V);
    (void)hipEventRecord(g->evFock1, s);
    return kOK;
}

}  // namespace xq
