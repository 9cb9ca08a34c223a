// XC nuclear gradients on the device, FP64, restricted closed shell or spin unrestricted: DFTIntegrator_Integrate with a
// gradients argument and no Fock matrix (DFTIntegrator.c:159-509, as DFTGridIntegrator.Gradients calls it,
// DFTGridIntegrator.py:72-88).
//
// Per block, with the basis values, gradients and second derivatives of its kept functions (order + 1, :203-207):
//   rho, grad rho, the block skip and the functional as in the quadrature; vrho, vsigma *= w; eXC = eps rho (:332-350);
//   direct terms (DFTIntegrator_Gradients :643-747 and DFTIntegrator_GradientsSigma :752-808, factor 4 for the restricted
//   case): for each kept row i, g_i = sum_p [(2 vrho X + 4 vsigma Y) grad phi + 4 vsigma X (Hess phi . grad rho)]_ip with
//   X = P_red Phi and Y = P_red (grad Phi . grad rho); g_i is subtracted from the atom of function i and added to the
//   block's atom (UGradientContributions :909-941, the grid moving with its atom).  The reference sums per function then
//   per atom; here the terms of one atom's rows are summed per point first, so the last bits differ.
// Spin unrestricted (DFTIntegrator.c:405-487): rho_s and grad rho_s of both spins, the block skipped when both are below
//   rhoTol, the polarised functional, eXC = eps (rho_a + rho_b), and the direct terms of each spin with its own P_s and
//   vrho_s and the sigma coefficient 4 vsigma_ss grad rho_s + 2 vsigma_ab grad rho_t (t the other spin).
// The weight terms (DFTGridWeights_Derivatives) follow per point from eXC (weight_gradient_kernel, xq_grid.hip).
//
// Device layout: the basis data of a batch of consecutive blocks in ten planes of rows of 128 points (as the set-up's FP64
// values), plus X and U planes per spin; batches keep the scratch bounded.  The products P_red Phi and P_red U run on
// FP64 MFMA with P_red staged in LDS, as in the FP64 quadrature's density kernel, one work-group per row chunk of a
// block and spin; per batch: the X pass, the per-block functional (which forms U), the Z pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "xq_functional.h"
#include "xq_internal.h"

namespace xq {

namespace {

// Rows of basis data per batch (12 or 14 planes of 128 doubles per row: about 6.4 or 7.5 GB).  Four times fewer,
// larger batches than 131,072 rows took 3.6 ms off an (H2O)64 gradient call: fewer launches, each with its own tail.
constexpr long long kBatchRows = 524288;
constexpr int kC = 32;                           // rows per chunk (two 16-row MFMA tiles)
constexpr int kPanelF = 4;                       // chunks of P_red columns staged in LDS at once
constexpr int kPanelPitch = kPanelF * kC + 2;    // LDS row pitch in doubles (padded against bank conflicts)
constexpr int kIdxCache = 1024;                  // kept-function indices cached in LDS
constexpr int kGradCopies = 64;                  // copies of the gradient array that spread the atomics

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma_f64(double a, double b, const v4d &c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// P_red Bp for one block on FP64 MFMA (v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]
// and receives D[4 v + l / 16][l % 16]).  256 threads; wave wv owns points 32 wv .. + 31 (two 16-point tiles); for each
// chunk I of 32 rows, epi(I, acc) receives acc[hm][hn][v] = (P_red Bp)[32 I + 16 hm + 4 v + l / 16][32 wv + 16 hn + l % 16].
// P_red is staged in LDS by panels of kPanelF chunks of columns; Bp is the block's first row of a plane of rows of 128
// points (rows >= nb are not read).  Row chunks I0 .. I1 - 1.
template <typename Epi>
__device__ __forceinline__ void reduced_product(const double *__restrict__ P, int n, const int *idx, int nb,
                                                const double *__restrict__ Bp, double (&sPf)[kC][kPanelPitch], int I0,
                                                int I1, Epi epi) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, l16 = lane & 15, l4 = lane >> 4;
    const int nChunks = (nb + kC - 1) / kC;
    const int sr = t >> 3, s8 = t & 7;     // staging: row sr, columns 4 s8 .. + 3 of each chunk of the panel
    auto fetchB = [&](int c, double (&bq)[8][2]) {
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int k = kC * c + 4 * s + l4;
#pragma unroll
            for (int h = 0; h < 2; h++)
                bq[s][h] = k < nb ? Bp[static_cast<size_t>(k) * kBlockSize + kC * wv + 16 * h + l16] : 0.0;
        }
    };
    for (int I = I0; I < I1; I++) {
        v4d acc[2][2];
#pragma unroll
        for (int hm = 0; hm < 2; hm++)
#pragma unroll
            for (int hn = 0; hn < 2; hn++) acc[hm][hn] = v4d{0.0, 0.0, 0.0, 0.0};
        double bnext[8][2];
        fetchB(0, bnext);
        const int r = kC * I + sr;
        const size_t rowOff = r < nb ? static_cast<size_t>(idx[r]) * n : 0;
        for (int c0 = 0; c0 < nChunks; c0 += kPanelF) {
            const int nc = nChunks - c0 < kPanelF ? nChunks - c0 : kPanelF;
            double pv[kPanelF][4];
#pragma unroll
            for (int m = 0; m < kPanelF; m++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    const int k = kC * (c0 + m) + 4 * s8 + jj;
                    pv[m][jj] = (m < nc && r < nb && k < nb) ? P[rowOff + idx[k]] : 0.0;
                }
            __syncthreads();   // every wave is done with the previous panel
#pragma unroll
            for (int m = 0; m < kPanelF; m++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++) sPf[sr][kC * m + 4 * s8 + jj] = pv[m][jj];
            __syncthreads();
            for (int cl = 0; cl < nc; cl++) {
                const int c = c0 + cl;
                double bq[8][2];
#pragma unroll
                for (int s = 0; s < 8; s++) { bq[s][0] = bnext[s][0]; bq[s][1] = bnext[s][1]; }
                if (c + 1 < nChunks) fetchB(c + 1, bnext);
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    const double a0 = sPf[l16][kC * cl + 4 * s + l4], a1 = sPf[16 + l16][kC * cl + 4 * s + l4];
                    acc[0][0] = mfma_f64(a0, bq[s][0], acc[0][0]);
                    acc[0][1] = mfma_f64(a0, bq[s][1], acc[0][1]);
                    acc[1][0] = mfma_f64(a1, bq[s][0], acc[1][0]);
                    acc[1][1] = mfma_f64(a1, bq[s][1], acc[1][1]);
                }
            }
        }
        epi(I, acc);
    }
}

// XCD ranges of a launch: XCD x takes entries start[x] .. start[x + 1] - 1 of its list.
struct XcdRanges {
    int start[9];
};

// The direct terms split by row chunk, in three launches per batch of blocks.  Items {block, row chunk I, slot of the
// item's shares, batch-local block} run one work-group each; the row chunks of a block are consecutive items of one XCD
// (work-group g runs on XCD g % 8), so that they read the block's basis rows at about the same time, through that XCD's
// L2.
//   grad_rows_x_kernel: per spin, X rows of chunk I = P_s Phi and the chunk's shares of rho_s and grad rho_s per point;
//   grad_functional_kernel (one work-group per block): rho and grad rho as the sums of the shares in chunk order, the
//     block skip, the functional, eXC and the coefficients (u0, c) per point, and U_s = u0 Phi + c . grad Phi for the
//     block's rows;
//   grad_rows_z_kernel: per spin, Z rows of chunk I = P_s U_s and the per-row terms of chunk I.
// The two row-chunk kernels take spin blockIdx.y of nSpins.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) grad_rows_x_kernel(
    const int4 *__restrict__ items, XcdRanges xr, long long rowBase, const DevBlock *__restrict__ blocks,
    const int *__restrict__ idxAll, int n, const double *__restrict__ P0, const double *__restrict__ P1, Planes10 B,
    double *__restrict__ Xp0, double *__restrict__ Xp1, double *__restrict__ part, int nSpins) {
    __shared__ double sPf[kC][kPanelPitch];
    __shared__ int sIdx[kIdxCache];
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    if (slot >= xr.start[xcd + 1] - xr.start[xcd]) return;
    const int4 it = items[xr.start[xcd] + slot];
    const int b = it.x, I = it.y;
    const DevBlock blk = blocks[b];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, l16 = lane & 15, l4 = lane >> 4;
    const int nb = blk.nKept;
    const int *idxG = idxAll + blk.iOffset;
    for (int k = t; k < nb && k < kIdxCache; k += 256) sIdx[k] = idxG[k];
    __syncthreads();
    const int *idx = nb <= kIdxCache ? sIdx : idxG;
    const size_t R0 = static_cast<size_t>(blk.vOffset - rowBase) * kBlockSize;
    const double *F = B.p[0] + R0, *Fx = B.p[1] + R0, *Fy = B.p[2] + R0, *Fz = B.p[3] + R0;
    const int sp = blockIdx.y;
    {
        double *X = (sp == 0 ? Xp0 : Xp1) + R0;
        double pr[2] = {0.0, 0.0}, pgx[2] = {0.0, 0.0}, pgy[2] = {0.0, 0.0}, pgz[2] = {0.0, 0.0};
        reduced_product(sp == 0 ? P0 : P1, n, idx, nb, F, sPf, I, I + 1, [&](int, v4d (&acc)[2][2]) {
#pragma unroll
            for (int hm = 0; hm < 2; hm++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = kC * I + 16 * hm + 4 * v + l4;
                    if (i < nb) {
#pragma unroll
                        for (int hn = 0; hn < 2; hn++) {
                            const size_t o = static_cast<size_t>(i) * kBlockSize + kC * wv + 16 * hn + l16;
                            const double x = acc[hm][hn][v];
                            X[o] = x;
                            pr[hn] += F[o] * x;
                            pgx[hn] += Fx[o] * x;
                            pgy[hn] += Fy[o] * x;
                            pgz[hn] += Fz[o] * x;
                        }
                    }
                }
        });
        double *q = part + (static_cast<size_t>(it.z) * nSpins + sp) * 4 * kBlockSize;
#pragma unroll
        for (int hn = 0; hn < 2; hn++) {
#pragma unroll
            for (int o = 16; o < 64; o <<= 1) {
                pr[hn] += __shfl_xor(pr[hn], o, 64);
                pgx[hn] += __shfl_xor(pgx[hn], o, 64);
                pgy[hn] += __shfl_xor(pgy[hn], o, 64);
                pgz[hn] += __shfl_xor(pgz[hn], o, 64);
            }
            if (l4 == 0) {
                const int p = kC * wv + 16 * hn + l16;
                q[p] = pr[hn];
                q[kBlockSize + p] = pgx[hn];
                q[2 * kBlockSize + p] = pgy[hn];
                q[3 * kBlockSize + p] = pgz[hn];
            }
        }
    }
}

// One work-group of 256 threads per block of the batch (blockList: {block, slot of its first item's shares}).  coef
// receives [batch-local block][4 kSpins][point]; skipOut 1 for a block without kept functions or below rhoTol.
template <int kSpins>
__global__ void __launch_bounds__(256) grad_functional_kernel(
    const int2 *__restrict__ blockList, long long rowBase, const DevBlock *__restrict__ blocks,
    const double *__restrict__ w, FunctionalSet fs, double rhoTol, Planes10 B, const double *__restrict__ part,
    double *__restrict__ Up0, double *__restrict__ Up1, double *__restrict__ coef, int *__restrict__ skipOut,
    double *__restrict__ eXC) {
    __shared__ double sCoef[4 * kSpins][kBlockSize];
    __shared__ double smax[4];
    const int2 bl = blockList[blockIdx.x];
    const int b = bl.x;
    const DevBlock blk = blocks[b];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int nb = blk.nKept, np = blk.nPoints, T = (nb + kC - 1) / kC;
    const bool valid = t < np;           // np <= 128: threads 128 .. 255 hold no point here
    double rs[kSpins][4];
#pragma unroll
    for (int sp = 0; sp < kSpins; sp++) rs[sp][0] = rs[sp][1] = rs[sp][2] = rs[sp][3] = 0.0;
    if (valid) {
        for (int I = 0; I < T; I++)
#pragma unroll
            for (int sp = 0; sp < kSpins; sp++) {
                const double *q = part + ((static_cast<size_t>(bl.y) + I) * kSpins + sp) * 4 * kBlockSize + t;
#pragma unroll
                for (int k = 0; k < 4; k++) rs[sp][k] += q[k * kBlockSize];
            }
    }
    double mx = 0.0;
    if (valid) {
#pragma unroll
        for (int sp = 0; sp < kSpins; sp++) mx = fmax(mx, fabs(rs[sp][0]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) smax[wv] = mx;
    __syncthreads();
    const int local = static_cast<int>(blockIdx.x);
    if (nb <= 0 || !(fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3])) > rhoTol)) {   // the block is skipped (:332)
        if (valid) eXC[blk.pStart + t] = 0.0;
        if (t == 0) skipOut[local] = 1;
        return;
    }
    if (t == 0) skipOut[local] = 0;
    double cf[kSpins][4];
#pragma unroll
    for (int sp = 0; sp < kSpins; sp++) cf[sp][0] = cf[sp][1] = cf[sp][2] = cf[sp][3] = 0.0;
    if (valid) {
        const double ww = w[blk.pStart + t];
        if constexpr (kSpins == 1) {
            const double rho = rs[0][0], gx = 2.0 * rs[0][1], gy = 2.0 * rs[0][2], gz = 2.0 * rs[0][3];
            double eps, vr, vs;
            evaluate_functional(fs, rho, gx * gx + gy * gy + gz * gz, eps, vr, vs);
            vr *= ww;
            vs *= ww;
            eXC[blk.pStart + t] = eps * rho;
            cf[0][0] = 2.0 * vr;
            cf[0][1] = 4.0 * vs * gx;
            cf[0][2] = 4.0 * vs * gy;
            cf[0][3] = 4.0 * vs * gz;
        } else {
            const double ra = rs[0][0], rb = rs[1][0];
            double ga[3], gb[3];
#pragma unroll
            for (int d = 0; d < 3; d++) { ga[d] = 2.0 * rs[0][1 + d]; gb[d] = 2.0 * rs[1][1 + d]; }
            PolOut f;
            evaluate_functional_pol(fs, ra, rb, ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2],
                                    ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2],
                                    gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2], f);
            eXC[blk.pStart + t] = f.e * (ra + rb);
            const double vsaa = f.vsaa * ww, vsab = f.vsab * ww, vsbb = f.vsbb * ww;
            cf[0][0] = 2.0 * f.va * ww;
            cf[1][0] = 2.0 * f.vb * ww;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                cf[0][1 + d] = 4.0 * vsaa * ga[d] + 2.0 * vsab * gb[d];
                cf[1][1 + d] = 4.0 * vsbb * gb[d] + 2.0 * vsab * ga[d];
            }
        }
    }
    if (t < kBlockSize) {
#pragma unroll
        for (int sp = 0; sp < kSpins; sp++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                sCoef[4 * sp + k][t] = cf[sp][k];
                coef[(static_cast<size_t>(local) * 4 * kSpins + 4 * sp + k) * kBlockSize + t] = cf[sp][k];
            }
    }
    __syncthreads();
    // U_s at point t % 128 for rows t / 128, + 2, ...
    const size_t R0 = static_cast<size_t>(blk.vOffset - rowBase) * kBlockSize;
    const double *F = B.p[0] + R0, *Fx = B.p[1] + R0, *Fy = B.p[2] + R0, *Fz = B.p[3] + R0;
    const int p = t & (kBlockSize - 1);
    double *Us[2] = {Up0 + R0, kSpins == 2 ? Up1 + R0 : nullptr};
#pragma unroll
    for (int sp = 0; sp < kSpins; sp++) {
        const double c0 = sCoef[4 * sp][p], cx = sCoef[4 * sp + 1][p], cy = sCoef[4 * sp + 2][p],
                     cz = sCoef[4 * sp + 3][p];
        double *U = Us[sp];
        for (int i = t >> 7; i < nb; i += 2) {
            const size_t o = static_cast<size_t>(i) * kBlockSize + p;
            U[o] = c0 * F[o] + cx * Fx[o] + cy * Fy[o] + cz * Fz[o];
        }
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) grad_rows_z_kernel(
    const int4 *__restrict__ items, XcdRanges xr, long long rowBase, const DevBlock *__restrict__ blocks,
    const int *__restrict__ idxAll, const int *__restrict__ atomOf, int n, const double *__restrict__ P0,
    const double *__restrict__ P1, Planes10 B, const double *__restrict__ Xp0, const double *__restrict__ Xp1,
    const double *__restrict__ Up0, const double *__restrict__ Up1, const double *__restrict__ coef,
    const int *__restrict__ skip, double *__restrict__ gradCopies, int nAtoms, int nSpins) {
    __shared__ double sPf[kC][kPanelPitch];
    __shared__ int sIdx[kIdxCache];
    __shared__ double sCoef[3][kBlockSize];
    __shared__ double sRed[4][kC][3];
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    if (slot >= xr.start[xcd + 1] - xr.start[xcd]) return;
    const int4 it = items[xr.start[xcd] + slot];
    const int b = it.x, I = it.y, local = it.w;
    if (skip[local]) return;
    const DevBlock blk = blocks[b];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, l16 = lane & 15, l4 = lane >> 4;
    const int nb = blk.nKept, np = blk.nPoints;
    const int *idxG = idxAll + blk.iOffset;
    const int sp = blockIdx.y;
    for (int k = t; k < nb && k < kIdxCache; k += 256) sIdx[k] = idxG[k];
    if (t < kBlockSize) {
#pragma unroll
        for (int k = 0; k < 3; k++)
            sCoef[k][t] = coef[(static_cast<size_t>(local) * 4 * nSpins + 4 * sp + 1 + k) * kBlockSize + t];
    }
    __syncthreads();
    const int *idx = nb <= kIdxCache ? sIdx : idxG;
    const size_t R0 = static_cast<size_t>(blk.vOffset - rowBase) * kBlockSize;
    const double *Fx = B.p[1] + R0, *Fy = B.p[2] + R0, *Fz = B.p[3] + R0;
    const double *Hxx = B.p[4] + R0, *Hxy = B.p[5] + R0, *Hxz = B.p[6] + R0, *Hyy = B.p[7] + R0, *Hyz = B.p[8] + R0,
                 *Hzz = B.p[9] + R0;
    double *grad = gradCopies + static_cast<size_t>((b + sp) % kGradCopies) * 3 * nAtoms;
    double gs[3] = {0.0, 0.0, 0.0};
    {
        const double *X = (sp == 0 ? Xp0 : Xp1) + R0;
        const double *sA = sCoef[0], *sB = sCoef[1], *sC = sCoef[2];
        reduced_product(sp == 0 ? P0 : P1, n, idx, nb, (sp == 0 ? Up0 : Up1) + R0, sPf, I, I + 1, [&](int, v4d (&acc)[2][2]) {
            double c[2][4][3];
#pragma unroll
            for (int hm = 0; hm < 2; hm++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    c[hm][v][0] = c[hm][v][1] = c[hm][v][2] = 0.0;
                    const int i = kC * I + 16 * hm + 4 * v + l4;
                    if (i < nb) {
#pragma unroll
                        for (int hn = 0; hn < 2; hn++) {
                            const int p = kC * wv + 16 * hn + l16;
                            if (p < np) {
                                const size_t o = static_cast<size_t>(i) * kBlockSize + p;
                                const double z = acc[hm][hn][v], x = X[o];
                                const double ax = sA[p], ay = sB[p], az = sC[p];
                                const double hxy = Hxy[o], hxz = Hxz[o], hyz = Hyz[o];
                                c[hm][v][0] += z * Fx[o] + x * (Hxx[o] * ax + hxy * ay + hxz * az);
                                c[hm][v][1] += z * Fy[o] + x * (hxy * ax + Hyy[o] * ay + hyz * az);
                                c[hm][v][2] += z * Fz[o] + x * (hxz * ax + hyz * ay + Hzz[o] * az);
                            }
                        }
                    }
                }
            // Sum over the 16 points of each lane group, then over the four waves.
#pragma unroll
            for (int hm = 0; hm < 2; hm++)
#pragma unroll
                for (int v = 0; v < 4; v++)
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        double e = c[hm][v][d];
#pragma unroll
                        for (int o = 1; o < 16; o <<= 1) e += __shfl_xor(e, o, 64);
                        if (l16 == 0) sRed[wv][16 * hm + 4 * v + l4][d] = e;
                    }
            __syncthreads();
            if (t < kC) {
                const int i = kC * I + t;
                if (i < nb) {
                    const int atom = atomOf[idxG[i]];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double gi = ((sRed[0][t][d] + sRed[1][t][d]) + sRed[2][t][d]) + sRed[3][t][d];
                        atomicAdd(&grad[3 * atom + d], -gi);
                        gs[d] += gi;
                    }
                }
            }
        });
    }
    // The grid moves with the block's atom (UGradientContributions :909-941): it receives the sum of the row terms.
    if (wv == 0) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double e = gs[d];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
            if (lane == 0) atomicAdd(&grad[3 * blk.atom + d], e);
        }
    }
}

// grad[k] += sum of the kGradCopies copies (one thread per entry).
__global__ void __launch_bounds__(256) sum_copies_kernel(int nEntries, const double *__restrict__ copies,
                                                         double *__restrict__ grad) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nEntries) return;
    double s = 0.0;
    for (int c = 0; c < kGradCopies; c++) s += copies[static_cast<size_t>(c) * nEntries + k];
    grad[k] += s;
}

}  // namespace

namespace {

// The launch order of the batch b0 .. b1 - 1, written to order[b0 .. b1): as the density launch's (xq_mixed.hip,
// build_block_order), 8 contiguous ranges of equal estimated cost (T^2 + T + 1 for a block of T row chunks), one per
// XCD, each sorted by decreasing cost, so that the batch's longest blocks do not start last.
XcdRanges batch_order(const xq_grid *g, int b0, int b1, std::vector<int> &order) {
    std::vector<double> cost(static_cast<size_t>(b1 - b0));
    double total = 0.0;
    for (int b = b0; b < b1; b++) {
        const double T = static_cast<double>((g->blocks[b].nKept + kC - 1) / kC);
        cost[b - b0] = T * T + T + 1.0;
        total += cost[b - b0];
        order[b] = b;
    }
    XcdRanges xr;
    xr.start[0] = b0;
    int b = b0;
    double acc = 0.0;
    for (int x = 1; x < 8; x++) {
        while (b < b1 && acc + 0.5 * cost[b - b0] < total * x / 8.0) acc += cost[(b++) - b0];
        xr.start[x] = b;
    }
    xr.start[8] = b1;
    for (int x = 0; x < 8; x++)
        std::stable_sort(order.begin() + xr.start[x], order.begin() + xr.start[x + 1],
                         [&](int u, int v) { return cost[u - b0] > cost[v - b0]; });
    return xr;
}

// The direct and weight terms for one (kSpins = 1) or two (alpha, beta) dense densities already on the device.
int gradients_on_device(xq_grid *g, const FunctionalSet &fs, int spins, const double *d_P0, const double *d_P1,
                        double *grad) {
    hipStream_t s = g->stream;
    const int nA = g->nAtoms, nB = g->nBlocks;
    // Function -> atom (the reference's functionCenters).
    std::vector<int> atomOf(g->h_funcAtom);
    if (atomOf.empty()) atomOf.push_back(0);
    long long maxRows = 0;
    for (const DevBlock &b : g->blocks) maxRows = std::max<long long>(maxRows, b.nKept);
    const long long batchRows = std::max(kBatchRows, maxRows);
    std::vector<double> zeros(3 * static_cast<size_t>(nA), 0.0);
    std::vector<double> zeroCopies(zeros.size() * kGradCopies, 0.0);
    // Ten basis planes, X and U per spin.
    const int nPlanes = 10 + 2 * spins;
    double *planes[14] = {};
    double *d_grad = nullptr, *d_eXC = nullptr, *d_copies = nullptr, *d_part = nullptr, *d_coef = nullptr;
    int *d_atomOf = nullptr, *d_skip = nullptr;
    int4 *d_items = nullptr;
    int2 *d_blockList = nullptr;
    // Batches of blocks whose kept rows fit the planes, and their launch orders.
    std::vector<std::pair<int, int>> batches;
    for (int b0 = 0; b0 < nB;) {
        const long long rowBase = g->blocks[b0].vOffset;
        int b1 = b0;
        while (b1 < nB && g->blocks[b1].vOffset + g->blocks[b1].nKept - rowBase <= batchRows) b1++;
        batches.emplace_back(b0, b1);
        b0 = b1;
    }
    std::vector<int> order(static_cast<size_t>(nB > 0 ? nB : 1), 0);
    // Per batch: its items (each block of the launch order cut into its row chunks, consecutive on the block's XCD),
    // their XCD ranges, and per block {block, slot of its first item's shares} (slots in block order).
    std::vector<int4> items;
    std::vector<int2> blockList(static_cast<size_t>(nB > 0 ? nB : 1));
    std::vector<XcdRanges> ranges;
    std::vector<size_t> itemBase;
    size_t maxItems = 1, maxBlocks = 1;
    for (const auto &bt : batches) {
        const int b0 = bt.first, b1 = bt.second;
        const XcdRanges xb = batch_order(g, b0, b1, order);
        int slot = 0;
        for (int b = b0; b < b1; b++) {
            blockList[b] = make_int2(b, slot);
            slot += (g->blocks[b].nKept + kC - 1) / kC;
        }
        maxItems = std::max(maxItems, static_cast<size_t>(slot));
        maxBlocks = std::max(maxBlocks, static_cast<size_t>(b1 - b0));
        itemBase.push_back(items.size());
        XcdRanges xi;
        for (int x = 0; x < 8; x++) {
            xi.start[x] = static_cast<int>(items.size() - itemBase.back());
            for (int k = xb.start[x]; k < xb.start[x + 1]; k++) {
                const int b = order[k];
                for (int I = 0; I < (g->blocks[b].nKept + kC - 1) / kC; I++)
                    items.push_back(make_int4(b, I, blockList[b].y + I, b - b0));
            }
        }
        xi.start[8] = static_cast<int>(items.size() - itemBase.back());
        ranges.push_back(xi);
    }
    if (items.empty()) items.push_back(make_int4(0, 0, 0, 0));
    int status = kOK;
    do {
        bool ok = true;
        for (int k = 0; k < nPlanes; k++)
            if ((status = dev_alloc(g, &planes[k], static_cast<size_t>(batchRows) * kBlockSize)) != kOK) {
                ok = false;
                break;
            }
        if (!ok) break;
        if ((status = dev_alloc(g, &d_grad, zeros.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_copies, zeroCopies.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_eXC, static_cast<size_t>(g->nPoints > 0 ? g->nPoints : 1))) != kOK) break;
        if ((status = dev_alloc(g, &d_atomOf, atomOf.size())) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_grad, zeros.data(), zeros.size() * sizeof(double), hipMemcpyHostToDevice, s),
                            "zero gradients")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_copies, zeroCopies.data(), zeroCopies.size() * sizeof(double),
                                           hipMemcpyHostToDevice, s), "zero gradient copies")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_atomOf, atomOf.data(), atomOf.size() * sizeof(int), hipMemcpyHostToDevice, s),
                            "copy function atoms")) != kOK) break;
        if ((status = dev_alloc(g, &d_items, items.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_blockList, blockList.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_part, maxItems * spins * 4 * kBlockSize)) != kOK) break;
        if ((status = dev_alloc(g, &d_coef, maxBlocks * spins * 4 * kBlockSize)) != kOK) break;
        if ((status = dev_alloc(g, &d_skip, maxBlocks)) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice, s),
                            "copy items")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_blockList, blockList.data(), blockList.size() * sizeof(int2),
                                           hipMemcpyHostToDevice, s), "copy block list")) != kOK) break;
        Planes10 B;
        for (int k = 0; k < 10; k++) B.p[k] = planes[k];
        double *Xa = planes[10], *Xb = spins == 2 ? planes[11] : planes[10];
        double *Ua = planes[10 + spins], *Ub = spins == 2 ? planes[13] : Ua;
        for (size_t k = 0; k < batches.size() && status == kOK; k++) {
            const int b0 = batches[k].first, b1 = batches[k].second;
            const long long rowBase = g->blocks[b0].vOffset;
            if ((status = eval_basis2(g, b0, b1, rowBase, B, s)) != kOK) break;
            const XcdRanges &xr = ranges[k];
            int perXcd = 0;
            for (int x = 0; x < 8; x++) perXcd = std::max(perXcd, xr.start[x + 1] - xr.start[x]);
            const int4 *it = d_items + itemBase[k];
            const int2 *bl = d_blockList + b0;
            const double *P1 = spins == 2 ? d_P1 : d_P0;
            if (perXcd > 0)
                hipLaunchKernelGGL(grad_rows_x_kernel, dim3(8 * perXcd, spins), dim3(256), 0, s, it, xr, rowBase,
                                   g->d_blocks, g->d_idx, g->nFunctions, d_P0, P1, B, Xa, Xb, d_part, spins);
            if (spins == 1)
                hipLaunchKernelGGL(grad_functional_kernel<1>, dim3(b1 - b0), dim3(256), 0, s, bl, rowBase, g->d_blocks,
                                   g->d_w, fs, g->rhoTol, B, d_part, Ua, Ub, d_coef, d_skip, d_eXC);
            else
                hipLaunchKernelGGL(grad_functional_kernel<2>, dim3(b1 - b0), dim3(256), 0, s, bl, rowBase, g->d_blocks,
                                   g->d_w, fs, g->rhoTol, B, d_part, Ua, Ub, d_coef, d_skip, d_eXC);
            if (perXcd > 0)
                hipLaunchKernelGGL(grad_rows_z_kernel, dim3(8 * perXcd, spins), dim3(256), 0, s, it, xr, rowBase,
                                   g->d_blocks, g->d_idx, d_atomOf, g->nFunctions, d_P0, P1, B, Xa, Xb, Ua, Ub, d_coef,
                                   d_skip, d_copies, nA, spins);
            status = check(hipGetLastError(), "direct gradient kernels");
        }
        if (status != kOK) break;
        const int nEntries = static_cast<int>(zeros.size());
        hipLaunchKernelGGL(sum_copies_kernel, dim3((nEntries + 255) / 256), dim3(256), 0, s, nEntries, d_copies, d_grad);
        if ((status = check(hipGetLastError(), "sum_copies_kernel")) != kOK) break;
        if ((status = weight_gradients(g, d_eXC, d_grad, s)) != kOK) break;
        if ((status = check(hipStreamSynchronize(s), "gradient sync")) != kOK) break;
        if ((status = check(hipMemcpyAsync(zeros.data(), d_grad, zeros.size() * sizeof(double), hipMemcpyDeviceToHost, s),
                            "copy gradients")) != kOK) break;
        if ((status = check(hipStreamSynchronize(s), "gradient sync")) != kOK) break;
        for (size_t k = 0; k < zeros.size(); k++) grad[k] += zeros[k];
    } while (false);
    (void)hipStreamSynchronize(s);
    for (double *&p : planes) dev_free(g, &p);
    dev_free(g, &d_grad);
    dev_free(g, &d_copies);
    dev_free(g, &d_eXC);
    dev_free(g, &d_atomOf);
    dev_free(g, &d_items);
    dev_free(g, &d_blockList);
    dev_free(g, &d_part);
    dev_free(g, &d_coef);
    dev_free(g, &d_skip);
    return status;
}

}  // namespace

int gradients(xq_grid *g, int nIDs, const int *ids, const double *Ppacked, double *grad) {
    FunctionalSet fs;
    XQ_TRY(functional_set(nIDs, ids, &fs));
    XQ_TRY(upload_density(g, Ppacked, g->stream));
    return gradients_on_device(g, fs, 1, g->d_P, nullptr, grad);
}

// Unrestricted: the alpha density goes to a dense copy of its own, the beta density to the handle's P.
int gradients_uks(xq_grid *g, int nIDs, const int *ids, const double *Pa, const double *Pb, double *grad) {
    FunctionalSet fs;
    XQ_TRY(functional_set(nIDs, ids, &fs));
    hipStream_t s = g->stream;
    const size_t n = static_cast<size_t>(g->nFunctions);
    double *d_Pa = nullptr;
    int status = kOK;
    do {
        if ((status = dev_alloc(g, &d_Pa, n * n > 0 ? n * n : 1)) != kOK) break;
        if ((status = upload_density(g, Pa, s)) != kOK) break;
        if (n > 0 && (status = check(hipMemcpyAsync(d_Pa, g->d_P, n * n * sizeof(double), hipMemcpyDeviceToDevice, s),
                                     "copy alpha density")) != kOK) break;
        // upload_density refills the pinned staging buffer, which the previous copy reads when it executes.
        if ((status = check(hipStreamSynchronize(s), "density sync")) != kOK) break;
        if ((status = upload_density(g, Pb, s)) != kOK) break;
        status = gradients_on_device(g, fs, 2, d_Pa, g->d_P, grad);
    } while (false);
    (void)hipStreamSynchronize(s);
    dev_free(g, &d_Pa);
    return status;
}

}  // namespace xq
