// Fused per-block quadrature on FP64 matrix cores (v_mfma_f64_16x16x4f64): density, functional and Fock contribution of one
// grid block in one work-group, reading the block's basis values once from HBM (the second and third touches hit the
// caches).  Same mathematics as xq_integrate.hip (DFTIntegrator.c:159-509); see that file for the formulae.
//
// Work-group = 256 threads = 4 waves.  MFMA lane layout (16 x 16 x 4, FP64): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and receives D[4 v + l / 16][l % 16] for v = 0..3.
//   Phase A  X = P_red Phi in macro tiles of 64 rows x 128 points; wave w owns rows 16 w .. 16 w + 15 and all eight 16-point
//            tiles (8 accumulators).  K runs over the kept functions in chunks of 16 staged in LDS (P_red gathered from the
//            dense P, Phi rows copied).  The epilogue multiplies X by Phi and grad Phi from memory and keeps per-lane point
//            partials of rho and grad rho; one LDS reduction at the end.
//   Phase B  FP64 functional, one thread per point; per-point coefficients of W in LDS; E_xc and N_e block sums.
//   Phase C  D = Phi W^T + W Phi^T on lower-triangle macro tiles of 64 x 64; wave w owns rows 16 w .. 16 w + 15 and four
//            16-column tiles.  K = the block's points in chunks of 16, W formed while staging.  FP64 atomics into V.
#include <hip/hip_runtime.h>

#include "xq_functional.h"
#include "xq_internal.h"

namespace xq {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int kKC = 16;   // K chunk staged in LDS

__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__global__ void __launch_bounds__(256) fused_block_kernel_f64(
    const xq::DevBlock *__restrict__ blocks, const int *__restrict__ idxAll, const double *__restrict__ f,
    const double *__restrict__ fx, const double *__restrict__ fy, const double *__restrict__ fz, int n,
    const double *__restrict__ P, const double *__restrict__ w, FunctionalSet fs, double rhoTol,
    double *__restrict__ blockEN, int *__restrict__ skip, double *__restrict__ V) {
    // LDS: phase A uses As/Bs; phase C reuses the same storage for its four tiles.
    __shared__ double lds[4 * 64 * (kKC + 1)];          // 34.8 KB
    __shared__ double coef[4][kBlockSize];              // a_p, b_p (x, y, z)
    __shared__ double red[16][kBlockSize];              // point partial sums
    __shared__ double rsum[4][kBlockSize];              // rho, grad rho
    __shared__ double wsum[4][2];
    const int b = blockIdx.x;
    const xq::DevBlock blk = blocks[b];
    const int nb = blk.nKept, np = blk.nPoints;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int li = lane & 15, lk = lane >> 4;
    if (nb <= 0) {
        if (t == 0) { blockEN[2 * b] = 0.0; blockEN[2 * b + 1] = 0.0; skip[b] = 1; }
        return;
    }
    const int *idx = idxAll + blk.iOffset;
    const size_t row0 = static_cast<size_t>(blk.vOffset);
    double (*As)[kKC + 1] = reinterpret_cast<double (*)[kKC + 1]>(lds);                 // [64][17]
    double (*Bs)[kBlockSize + 1] = reinterpret_cast<double (*)[kBlockSize + 1]>(lds + 64 * (kKC + 1));  // [16][129]

    // ---------------------------------------------------------------- Phase A: X = P_red Phi, rho and grad rho partials.
    double pr[8], pgx[8], pgy[8], pgz[8];
#pragma unroll
    for (int q = 0; q < 8; q++) { pr[q] = 0.0; pgx[q] = 0.0; pgy[q] = 0.0; pgz[q] = 0.0; }
    for (int I = 0; I < nb; I += 64) {
        v4d acc[8];
#pragma unroll
        for (int q = 0; q < 8; q++) acc[q] = v4d{0.0, 0.0, 0.0, 0.0};
        for (int K = 0; K < nb; K += kKC) {
#pragma unroll
            for (int q = 0; q < 4; q++) {          // 64 x 16 of P_red
                const int e = t + 256 * q, r = e >> 4, c = e & 15;
                As[r][c] = (I + r < nb && K + c < nb) ? P[static_cast<size_t>(idx[I + r]) * n + idx[K + c]] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {          // 16 x 128 of Phi
                const int e = t + 256 * q, r = e >> 7, c = e & 127;
                Bs[r][c] = K + r < nb ? f[(row0 + K + r) * kBlockSize + c] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < kKC; kk += 4) {
                const double a = As[16 * wv + li][kk + lk];
#pragma unroll
                for (int q = 0; q < 8; q++) acc[q] = mfma16(a, Bs[kk + lk][16 * q + li], acc[q]);
            }
            __syncthreads();
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int r = I + 16 * wv + 4 * v + lk;
            if (r < nb) {
                const size_t o = (row0 + r) * kBlockSize;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int p = 16 * q + li;
                    const double xv = acc[q][v];
                    pr[q] += f[o + p] * xv;
                    pgx[q] += fx[o + p] * xv;
                    pgy[q] += fy[o + p] * xv;
                    pgz[q] += fz[o + p] * xv;
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
        for (int q = 0; q < 8; q++) red[4 * wv + lk][16 * q + li] = c == 0 ? pr[q] : c == 1 ? pgx[q] : c == 2 ? pgy[q] : pgz[q];
        __syncthreads();
        if (t < kBlockSize) {
            double s = 0.0;
            for (int g = 0; g < 16; g++) s += red[g][t];
            rsum[c][t] = s;
        }
        __syncthreads();
    }

    // ---------------------------------------------------------------- Phase B: functional, W coefficients, E and N.
    double rho = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    if (t < np) { rho = rsum[0][t]; gx = 2.0 * rsum[1][t]; gy = 2.0 * rsum[2][t]; gz = 2.0 * rsum[3][t]; }
    double m = t < np ? fabs(rho) : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (lane == 0) wsum[wv][0] = m;
    __syncthreads();
    const bool skipped = !(fmax(fmax(wsum[0][0], wsum[1][0]), fmax(wsum[2][0], wsum[3][0])) > rhoTol);
    __syncthreads();
    double e = 0.0, ne = 0.0;
    if (t < kBlockSize) {
        double vr = 0.0, vs = 0.0;
        if (!skipped && t < np) {
            double eps;
            evaluate_functional(fs, rho, gx * gx + gy * gy + gz * gz, eps, vr, vs);
            const double ww = w[blk.pStart + t];
            vr *= ww;
            vs *= ww;
            e = eps * rho * ww;
            ne = rho * ww;
        }
        coef[0][t] = 0.5 * vr;
        coef[1][t] = 2.0 * vs * gx;
        coef[2][t] = 2.0 * vs * gy;
        coef[3][t] = 2.0 * vs * gz;
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
    if (skipped) return;

    // ---------------------------------------------------------------- Phase C: V += Phi W^T + W Phi^T (lower triangle).
    double (*FI)[kKC + 1] = reinterpret_cast<double (*)[kKC + 1]>(lds);
    double (*WI)[kKC + 1] = reinterpret_cast<double (*)[kKC + 1]>(lds + 64 * (kKC + 1));
    double (*FJ)[kKC + 1] = reinterpret_cast<double (*)[kKC + 1]>(lds + 2 * 64 * (kKC + 1));
    double (*WJ)[kKC + 1] = reinterpret_cast<double (*)[kKC + 1]>(lds + 3 * 64 * (kKC + 1));
    for (int I = 0; I < nb; I += 64) {
        for (int J = 0; J <= I; J += 64) {
            v4d acc[4];
#pragma unroll
            for (int q = 0; q < 4; q++) acc[q] = v4d{0.0, 0.0, 0.0, 0.0};
            for (int K = 0; K < np; K += kKC) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int el = t + 256 * q, r = el >> 4, c = el & 15, p = K + c;
                    double fi = 0.0, wi = 0.0, fj = 0.0, wj = 0.0;
                    if (p < np) {
                        const double c0 = coef[0][p], c1 = coef[1][p], c2 = coef[2][p], c3 = coef[3][p];
                        if (I + r < nb) {
                            const size_t o = (row0 + I + r) * kBlockSize + p;
                            fi = f[o];
                            wi = c0 * fi + c1 * fx[o] + c2 * fy[o] + c3 * fz[o];
                        }
                        if (J + r < nb) {
                            const size_t o = (row0 + J + r) * kBlockSize + p;
                            fj = f[o];
                            wj = c0 * fj + c1 * fx[o] + c2 * fy[o] + c3 * fz[o];
                        }
                    }
                    FI[r][c] = fi;
                    WI[r][c] = wi;
                    FJ[r][c] = fj;
                    WJ[r][c] = wj;
                }
                __syncthreads();
#pragma unroll
                for (int kk = 0; kk < kKC; kk += 4) {
                    const double a1 = FI[16 * wv + li][kk + lk], a2 = WI[16 * wv + li][kk + lk];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        acc[q] = mfma16(a1, WJ[16 * q + li][kk + lk], acc[q]);
                        acc[q] = mfma16(a2, FJ[16 * q + li][kk + lk], acc[q]);
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = I + 16 * wv + 4 * v + lk;
                if (r >= nb) continue;
                const size_t mrow = static_cast<size_t>(idx[r]) * n;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int c = J + 16 * q + li;
                    if (c < nb && c <= r) atomicAdd(&V[mrow + idx[c]], acc[q][v]);
                }
            }
        }
    }
}

}  // namespace

void launch_fused_f64(xq_grid *g, const FunctionalSet &fs, hipStream_t s) {
    hipLaunchKernelGGL(fused_block_kernel_f64, dim3(g->nBlocks), dim3(256), 0, s, g->d_blocks, g->d_idx,
                       static_cast<const double *>(g->d_f), static_cast<const double *>(g->d_fx),
                       static_cast<const double *>(g->d_fy), static_cast<const double *>(g->d_fz), g->nFunctions, g->d_P,
                       g->d_w, fs, g->rhoTol, g->d_blockEN, g->d_skip, g->d_V);
}

}  // namespace xq
