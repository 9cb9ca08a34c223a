// XC nuclear gradients on the device, FP64, restricted closed shell with GGA functionals: DFTIntegrator_Integrate with a
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
// The weight terms (DFTGridWeights_Derivatives) follow per point from eXC (weight_gradient_kernel, xq_grid.hip).
//
// Device layout: the basis data of a batch of consecutive blocks in ten planes of rows of 128 points (as the set-up's FP64
// values), plus X and U planes; batches keep the scratch bounded.  The products P_red Phi and P_red U run on FP64 MFMA
// with P_red staged in LDS, as in the FP64 quadrature's density kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "xq_functional.h"
#include "xq_internal.h"

namespace xq {

namespace {

// Rows of basis data per batch (12 planes of 128 doubles per row: about 1.6 GB).
constexpr long long kBatchRows = 131072;
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
// points (rows >= nb are not read).
template <typename Epi>
__device__ __forceinline__ void reduced_product(const double *__restrict__ P, int n, const int *idx, int nb,
                                                const double *__restrict__ Bp, double (&sPf)[kC][kPanelPitch], Epi epi) {
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
    for (int I = 0; I < nChunks; I++) {
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

// Direct terms of the blocks of one batch (one work-group of 256 threads per block; work-group g runs on XCD g % 8 and
// each XCD takes a contiguous range of blocks, so that its L2 holds the P rows they gather).  B: the ten basis planes of
// the batch; Xp, Up: scratch planes with the same rows; eXC[point] = eps rho (zero for skipped or empty blocks); the
// gradient terms go to copy (work-group % kGradCopies) of gradCopies.
//   pass 1: X = P_red Phi, rho, grad rho, the functional (vrho, vsigma times w);
//   U = 2 vrho Phi + 4 vsigma (grad rho . grad Phi), so that P_red U = 2 vrho X + 4 vsigma Y;
//   pass 2: Z = P_red U and per row g_i = sum_p [Z grad phi + X (Hess phi . 4 vsigma grad rho)]_ip.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) grad_direct_kernel(
    int b0, int nGroups, int groupsPerXcd, long long rowBase, const DevBlock *__restrict__ blocks,
    const int *__restrict__ idxAll, const int *__restrict__ atomOf, int n, const double *__restrict__ P,
    const double *__restrict__ w, FunctionalSet fs, double rhoTol, Planes10 B, double *__restrict__ Xp,
    double *__restrict__ Up, double *__restrict__ eXC, double *__restrict__ gradCopies, int nAtoms) {
    __shared__ double sPf[kC][kPanelPitch];
    __shared__ int sIdx[kIdxCache];
    __shared__ double rsum[4][kBlockSize];
    __shared__ double sCoef[4][kBlockSize];
    __shared__ double sRed[4][kC][3];
    __shared__ double smax[4];
    const int wg = (blockIdx.x & 7) * groupsPerXcd + (blockIdx.x >> 3);
    if (wg >= nGroups) return;
    const int b = b0 + wg;
    const DevBlock blk = blocks[b];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, l16 = lane & 15, l4 = lane >> 4;
    const int nb = blk.nKept, np = blk.nPoints;
    if (nb <= 0) {
        if (t < np) eXC[blk.pStart + t] = 0.0;
        return;
    }
    const int *idxG = idxAll + blk.iOffset;
    for (int k = t; k < nb && k < kIdxCache; k += 256) sIdx[k] = idxG[k];
    __syncthreads();
    const int *idx = nb <= kIdxCache ? sIdx : idxG;
    const size_t R0 = static_cast<size_t>(blk.vOffset - rowBase) * kBlockSize;
    const double *F = B.p[0] + R0, *Fx = B.p[1] + R0, *Fy = B.p[2] + R0, *Fz = B.p[3] + R0;
    double *X = Xp + R0, *U = Up + R0;
    double *grad = gradCopies + static_cast<size_t>(wg % kGradCopies) * 3 * nAtoms;

    // Pass 1: X, rho and grad rho (DFTIntegrator_GridPointRho, DFTIntegrator.c:814-877).
    double pr[2] = {0.0, 0.0}, pgx[2] = {0.0, 0.0}, pgy[2] = {0.0, 0.0}, pgz[2] = {0.0, 0.0};
    reduced_product(P, n, idx, nb, F, sPf, [&](int I, v4d (&acc)[2][2]) {
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
            rsum[0][p] = pr[hn];
            rsum[1][p] = pgx[hn];
            rsum[2][p] = pgy[hn];
            rsum[3][p] = pgz[hn];
        }
    }
    __syncthreads();
    const bool valid = t < np;           // np <= 128: threads 128 .. 255 hold no point here
    const double rho = valid ? rsum[0][t] : 0.0, gx = valid ? 2.0 * rsum[1][t] : 0.0,
                 gy = valid ? 2.0 * rsum[2][t] : 0.0, gz = valid ? 2.0 * rsum[3][t] : 0.0;
    double mx = valid ? fabs(rho) : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) smax[wv] = mx;
    __syncthreads();
    if (!(fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3])) > rhoTol)) {   // the block is skipped (:332)
        if (valid) eXC[blk.pStart + t] = 0.0;
        return;
    }
    double u0 = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
    if (valid) {
        double eps, vr, vs;
        evaluate_functional(fs, rho, gx * gx + gy * gy + gz * gz, eps, vr, vs);
        const double ww = w[blk.pStart + t];
        vr *= ww;
        vs *= ww;
        eXC[blk.pStart + t] = eps * rho;
        u0 = 2.0 * vr;
        ux = 4.0 * vs * gx;
        uy = 4.0 * vs * gy;
        uz = 4.0 * vs * gz;
    }
    if (t < kBlockSize) {
        sCoef[0][t] = u0;
        sCoef[1][t] = ux;
        sCoef[2][t] = uy;
        sCoef[3][t] = uz;
    }
    __syncthreads();
    {
        // Thread t builds U at point t % 128 for rows t / 128, + 2, ...
        const int p = t & (kBlockSize - 1);
        const double c0 = sCoef[0][p], cx = sCoef[1][p], cy = sCoef[2][p], cz = sCoef[3][p];
        for (int i = t >> 7; i < nb; i += 2) {
            const size_t o = static_cast<size_t>(i) * kBlockSize + p;
            U[o] = c0 * F[o] + cx * Fx[o] + cy * Fy[o] + cz * Fz[o];
        }
    }
    __syncthreads();   // U and X are complete for the work-group

    // Pass 2: Z = P_red U and the per-row terms (DFTIntegrator_Gradients :643-747, _GradientsSigma :752-808).
    const double *Hxx = B.p[4] + R0, *Hxy = B.p[5] + R0, *Hxz = B.p[6] + R0, *Hyy = B.p[7] + R0, *Hyz = B.p[8] + R0,
                 *Hzz = B.p[9] + R0;
    double gs[3] = {0.0, 0.0, 0.0};
    reduced_product(P, n, idx, nb, U, sPf, [&](int I, v4d (&acc)[2][2]) {
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
                            const double ax = sCoef[1][p], ay = sCoef[2][p], az = sCoef[3][p];
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

int gradients(xq_grid *g, int nIDs, const int *ids, const double *Ppacked, double *grad) {
    FunctionalSet fs;
    XQ_TRY(functional_set(nIDs, ids, &fs));
    hipStream_t s = g->stream;
    XQ_TRY(upload_density(g, Ppacked, s));
    const int nA = g->nAtoms, nB = g->nBlocks;
    // Function -> atom (the reference's functionCenters).
    std::vector<int> atomOf(g->h_funcAtom);
    if (atomOf.empty()) atomOf.push_back(0);
    long long maxRows = 0;
    for (const DevBlock &b : g->blocks) maxRows = std::max<long long>(maxRows, b.nKept);
    const long long batchRows = std::max(kBatchRows, maxRows);
    std::vector<double> zeros(3 * static_cast<size_t>(nA), 0.0);
    std::vector<double> zeroCopies(zeros.size() * kGradCopies, 0.0);
    double *planes[12] = {};
    double *d_grad = nullptr, *d_eXC = nullptr, *d_copies = nullptr;
    int *d_atomOf = nullptr;
    int status = kOK;
    do {
        bool ok = true;
        for (double *&p : planes)
            if ((status = dev_alloc(g, &p, static_cast<size_t>(batchRows) * kBlockSize)) != kOK) { ok = false; break; }
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
        Planes10 B;
        for (int k = 0; k < 10; k++) B.p[k] = planes[k];
        for (int b0 = 0; b0 < nB && status == kOK;) {
            const long long rowBase = g->blocks[b0].vOffset;
            int b1 = b0;
            while (b1 < nB && g->blocks[b1].vOffset + g->blocks[b1].nKept - rowBase <= batchRows) b1++;
            if ((status = eval_basis2(g, b0, b1, rowBase, B, s)) != kOK) break;
            const int nGroups = b1 - b0, perXcd = (nGroups + 7) / 8;
            hipLaunchKernelGGL(grad_direct_kernel, dim3(8 * perXcd), dim3(256), 0, s, b0, nGroups, perXcd, rowBase,
                               g->d_blocks, g->d_idx, d_atomOf, g->nFunctions, g->d_P, g->d_w, fs, g->rhoTol, B, planes[10],
                               planes[11], d_eXC, d_copies, nA);
            status = check(hipGetLastError(), "grad_direct_kernel");
            b0 = b1;
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
    return status;
}

}  // namespace xq
