// XC nuclear gradients on the device, FP64, restricted closed shell with GGA functionals: DFTIntegrator_Integrate with a
// gradients argument and no Fock matrix (DFTIntegrator.c:159-509, as DFTGridIntegrator.Gradients calls it,
// DFTGridIntegrator.py:72-88).
//
// Per block, with the basis values, gradients and second derivatives of its kept functions (order + 1, :203-207):
//   rho, grad rho, the block skip and the functional as in the quadrature; vrho, vsigma *= w; eXC = eps rho (:332-350);
//   direct terms (DFTIntegrator_Gradients :643-747 and DFTIntegrator_GradientsSigma :752-808, factor 4 for the restricted
//   case): for each kept row i, g_i = sum_p [(2 vrho X + 4 vsigma Y) grad phi + 4 vsigma X (Hess phi . grad rho)]_ip with
//   X = P_red Phi and Y = P_red (grad Phi . grad rho); g_i is subtracted from the atom of function i and added to the
//   block's atom (UGradientContributions :909-941, the grid moving with its atom).
// The weight terms (DFTGridWeights_Derivatives) follow per point from eXC (weight_gradient_kernel, xq_grid.hip).
//
// Device layout: the basis data of a batch of consecutive blocks in ten planes of rows of 128 points (as the set-up's FP64
// values), plus X, grad Phi . grad rho and Y planes; batches keep the scratch bounded.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "xq_functional.h"
#include "xq_internal.h"

namespace xq {

namespace {

// Rows of basis data per batch (13 planes of 128 doubles per row: about 1.7 GB).
constexpr long long kBatchRows = 131072;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Direct terms of the blocks b0 + blockIdx.x (one work-group of 128 threads = points per block).  B: the ten basis planes;
// X, T, Y: scratch planes with the same rows; eXC[point] = eps rho (zero for skipped or empty blocks).
__global__ void __launch_bounds__(kBlockSize) grad_direct_kernel(
    int b0, long long rowBase, const DevBlock *__restrict__ blocks, const int *__restrict__ idx,
    const int *__restrict__ atomOf, int n, const double *__restrict__ P, const double *__restrict__ w, FunctionalSet fs,
    double rhoTol, Planes10 B, double *__restrict__ X, double *__restrict__ T, double *__restrict__ Y,
    double *__restrict__ eXC, double *__restrict__ grad) {
    __shared__ double red[2][3];
    __shared__ double smax[2];
    const int b = b0 + static_cast<int>(blockIdx.x);
    const DevBlock blk = blocks[b];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const bool valid = t < blk.nPoints;
    const int nb = blk.nKept;
    const int *kept = idx + blk.iOffset;
    const size_t R0 = static_cast<size_t>(blk.vOffset - rowBase);
    auto at = [&](const double *a, int i) { return a[(R0 + i) * kBlockSize + t]; };
    if (nb <= 0) {
        if (valid) eXC[blk.pStart + t] = 0.0;
        return;
    }
    // X = P_red Phi, rho and grad rho (DFTIntegrator_GridPointRho, :814-877).
    double rho = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
    for (int i = 0; i < nb; i++) {
        const double *Prow = P + static_cast<size_t>(kept[i]) * n;
        double x = 0.0;
        for (int j = 0; j < nb; j++) x += Prow[kept[j]] * at(B.p[0], j);
        X[(R0 + i) * kBlockSize + t] = x;
        rho += at(B.p[0], i) * x;
        ax += at(B.p[1], i) * x;
        ay += at(B.p[2], i) * x;
        az += at(B.p[3], i) * x;
    }
    const double gx = 2.0 * ax, gy = 2.0 * ay, gz = 2.0 * az;
    double m = valid ? fabs(rho) : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (lane == 0) smax[wv] = m;
    __syncthreads();
    if (!(fmax(smax[0], smax[1]) > rhoTol)) {      // the block is skipped (:332)
        if (valid) eXC[blk.pStart + t] = 0.0;
        return;
    }
    double vr = 0.0, vs = 0.0;
    if (valid) {
        double eps;
        evaluate_functional(fs, rho, gx * gx + gy * gy + gz * gz, eps, vr, vs);
        const double ww = w[blk.pStart + t];
        vr *= ww;
        vs *= ww;
        eXC[blk.pStart + t] = eps * rho;
    }
    // grad Phi . grad rho per row, then Y = P_red (grad Phi . grad rho); each thread uses its own column only.
    for (int i = 0; i < nb; i++)
        T[(R0 + i) * kBlockSize + t] = at(B.p[1], i) * gx + at(B.p[2], i) * gy + at(B.p[3], i) * gz;
    double gsum[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < nb; i++) {
        const double *Prow = P + static_cast<size_t>(kept[i]) * n;
        double y = 0.0;
        for (int j = 0; j < nb; j++) y += Prow[kept[j]] * T[(R0 + j) * kBlockSize + t];
        const double x = X[(R0 + i) * kBlockSize + t];
        const double w1 = 2.0 * vr * x + 4.0 * vs * y, w2 = 4.0 * vs * x;
        const double hx = at(B.p[4], i) * gx + at(B.p[5], i) * gy + at(B.p[6], i) * gz;
        const double hy = at(B.p[5], i) * gx + at(B.p[7], i) * gy + at(B.p[8], i) * gz;
        const double hz = at(B.p[6], i) * gx + at(B.p[8], i) * gy + at(B.p[9], i) * gz;
        double c[3] = {0.0, 0.0, 0.0};
        if (valid) {
            c[0] = w1 * at(B.p[1], i) + w2 * hx;
            c[1] = w1 * at(B.p[2], i) + w2 * hy;
            c[2] = w1 * at(B.p[3], i) + w2 * hz;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) c[d] = wave_sum(c[d]);
        if (lane == 0) { red[wv][0] = c[0]; red[wv][1] = c[1]; red[wv][2] = c[2]; }
        __syncthreads();
        if (t == 0) {
            const int atom = atomOf[kept[i]];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double gi = red[0][d] + red[1][d];
                atomicAdd(&grad[3 * atom + d], -gi);
                gsum[d] += gi;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int d = 0; d < 3; d++) atomicAdd(&grad[3 * blk.atom + d], gsum[d]);
    }
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
    double *planes[13] = {};
    double *d_grad = nullptr, *d_eXC = nullptr;
    int *d_atomOf = nullptr;
    int status = kOK;
    do {
        bool ok = true;
        for (double *&p : planes)
            if ((status = dev_alloc(g, &p, static_cast<size_t>(batchRows) * kBlockSize)) != kOK) { ok = false; break; }
        if (!ok) break;
        if ((status = dev_alloc(g, &d_grad, zeros.size())) != kOK) break;
        if ((status = dev_alloc(g, &d_eXC, static_cast<size_t>(g->nPoints > 0 ? g->nPoints : 1))) != kOK) break;
        if ((status = dev_alloc(g, &d_atomOf, atomOf.size())) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_grad, zeros.data(), zeros.size() * sizeof(double), hipMemcpyHostToDevice, s),
                            "zero gradients")) != kOK) break;
        if ((status = check(hipMemcpyAsync(d_atomOf, atomOf.data(), atomOf.size() * sizeof(int), hipMemcpyHostToDevice, s),
                            "copy function atoms")) != kOK) break;
        Planes10 B;
        for (int k = 0; k < 10; k++) B.p[k] = planes[k];
        for (int b0 = 0; b0 < nB && status == kOK;) {
            const long long rowBase = g->blocks[b0].vOffset;
            int b1 = b0;
            while (b1 < nB && g->blocks[b1].vOffset + g->blocks[b1].nKept - rowBase <= batchRows) b1++;
            if ((status = eval_basis2(g, b0, b1, rowBase, B, s)) != kOK) break;
            hipLaunchKernelGGL(grad_direct_kernel, dim3(b1 - b0), dim3(kBlockSize), 0, s, b0, rowBase, g->d_blocks, g->d_idx,
                               d_atomOf, g->nFunctions, g->d_P, g->d_w, fs, g->rhoTol, B, planes[10], planes[11],
                               planes[12], d_eXC, d_grad);
            status = check(hipGetLastError(), "grad_direct_kernel");
            b0 = b1;
        }
        if (status != kOK) break;
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
    dev_free(g, &d_eXC);
    dev_free(g, &d_atomOf);
    return status;
}

}  // namespace xq
