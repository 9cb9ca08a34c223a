// Unrestricted (UKS) quadrature: alpha and beta densities, alpha and beta V_xc (DFTIntegrator.c:281-402 with
// isSpinUnrestricted, as DFTIntegrator.pyx:82-107 calls it with alpha and beta matrices).
//
// The restricted path's kernels do the work, in either precision, with the two spins as two jobs of each launch: the
// density stage runs with each job's rhoOut set, which stores (rho_s, grad rho_s) per point and stops before the
// functional; spin_functional_kernel then evaluates the spin-polarised functional per block, skips a block whose |rho_a|
// and |rho_b| are all <= rhoTol, forms the block's E_xc and N_e sums and the W coefficients of both spins,
//   W_s = (w vrho_s / 2) Phi + (2 w vsigma_ss grad rho_s + w vsigma_ab grad rho_t) . grad Phi   (t the other spin),
// which is DFTIntegrator_Fock plus DFTIntegrator_FockSigma(:517-600) in the form V_s = Phi W_s^T + W_s Phi^T; the Fock
// stage then runs once for both spins.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "xq_functional.h"
#include "xq_internal.h"

namespace xq {

namespace {

// One work-group of 128 threads (one per point) per block.
__global__ void __launch_bounds__(kBlockSize) spin_functional_kernel(const DevBlock *__restrict__ blocks,
                                                                     const double *__restrict__ w, FunctionalSet fs,
                                                                     double rhoTol, const double *__restrict__ rhoA,
                                                                     const double *__restrict__ rhoB,
                                                                     double *__restrict__ wcA, double *__restrict__ wcB,
                                                                     double *__restrict__ blockEN, int *__restrict__ skip) {
    __shared__ double red[2][2];
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const DevBlock blk = blocks[b];
    const bool valid = t < blk.nPoints;
    const size_t p = static_cast<size_t>(blk.pStart) + t;
    double a[4] = {0.0, 0.0, 0.0, 0.0}, c[4] = {0.0, 0.0, 0.0, 0.0};
    if (valid && blk.nKept > 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) { a[k] = rhoA[4 * p + k]; c[k] = rhoB[4 * p + k]; }
    }
    // Skip test over both spins (RealArray2D_AbsoluteMaximum of the block's densities, DFTIntegrator.c:331).
    double mx = fmax(fabs(a[0]), fabs(c[0]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) red[wv][0] = mx;
    __syncthreads();
    const bool skipped = blk.nKept <= 0 || !(fmax(red[0][0], red[1][0]) > rhoTol);
    __syncthreads();
    double e = 0.0, ne = 0.0;
    double ca[4] = {0.0, 0.0, 0.0, 0.0}, cb[4] = {0.0, 0.0, 0.0, 0.0};
    if (valid && !skipped) {
        const double saa = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
        const double sab = a[1] * c[1] + a[2] * c[2] + a[3] * c[3];
        const double sbb = c[1] * c[1] + c[2] * c[2] + c[3] * c[3];
        PolOut f;
        evaluate_functional_pol(fs, a[0], c[0], saa, sab, sbb, f);
        const double ww = w[p], rt = a[0] + c[0];
        e = f.e * rt * ww;
        ne = rt * ww;
        const double va = f.va * ww, vb = f.vb * ww, vsaa = f.vsaa * ww, vsab = f.vsab * ww, vsbb = f.vsbb * ww;
        ca[0] = 0.5 * va;
        cb[0] = 0.5 * vb;
#pragma unroll
        for (int k = 1; k < 4; k++) {
            ca[k] = 2.0 * vsaa * a[k] + vsab * c[k];
            cb[k] = 2.0 * vsbb * c[k] + vsab * a[k];
        }
    }
    if (valid) {
#pragma unroll
        for (int k = 0; k < 4; k++) { wcA[4 * p + k] = ca[k]; wcB[4 * p + k] = cb[k]; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { e += __shfl_xor(e, o, 64); ne += __shfl_xor(ne, o, 64); }
    if (lane == 0) { red[wv][0] = e; red[wv][1] = ne; }
    __syncthreads();
    if (t == 0) {
        blockEN[2 * b] = red[0][0] + red[1][0];
        blockEN[2 * b + 1] = red[0][1] + red[1][1];
        skip[b] = skipped ? 1 : 0;
    }
}

}  // namespace

// The two spins run as two jobs of one launch per stage on the handle (the beta job with its own per-step arrays), in
// one queue of work with a single wait at the end: both densities in, both density stages in one launch, the
// polarised functional, both Fock stages in one launch, both V_xc out.
int integrate_uks(xq_grid *g, int nIDs, const int *ids, const double *Pa, const double *Pb, double *eXC, double *nE,
                  double *Va, double *Vb, double *stats) {
    FunctionalSet fs;
    XQ_TRY(functional_set(nIDs, ids, &fs));
    XQ_TRY(ensure_work(g));
    hipStream_t s = g->stream;
    const StageEvents ev{g->ev0, g->ev1, g->evRho0, g->evRho1, g->evFock1, g->evDens1};
    const int n = g->nFunctions;
    const size_t nn = static_cast<size_t>(n) * n, nPacked = static_cast<size_t>(n) * (n + 1) / 2;
    const size_t nP = static_cast<size_t>(g->nPoints > 0 ? g->nPoints : 1);
    const bool mixed = g->precision == XQ_PRECISION_MIXED;
    xq_grid *grids[2] = {g, g};
    const FunctionalSet fs2[2] = {fs, fs};
    StepArrays beta;                            // the beta job's P digits and W planes (handle's arrays: alpha)
    double *rhoA = nullptr, *rhoB = nullptr, *wcB = nullptr, *Pbd = nullptr, *Vbd = nullptr, *packedB = nullptr;
    double *stageB = nullptr;                   // pinned: beta's packed P in, packed V out
    MixedJob *h_tab = nullptr, *d_tab = nullptr;   // two 2-entry job tables: density stage, Fock stage
    int status = kOK;
    float msTotal = 0.0f;
    do {
        if ((status = mixed ? alloc_step_mixed(g, &beta) : alloc_step_f64(g, &beta)) != kOK) break;
        if ((status = dev_alloc(g, &rhoA, 4 * nP)) != kOK) break;
        if ((status = dev_alloc(g, &rhoB, 4 * nP)) != kOK) break;
        if ((status = dev_alloc(g, &wcB, 4 * nP)) != kOK) break;
        if ((status = dev_alloc(g, &Pbd, nn > 0 ? nn : 1)) != kOK) break;
        if ((status = dev_alloc(g, &Vbd, nn > 0 ? nn : 1)) != kOK) break;
        if ((status = dev_alloc(g, &packedB, nPacked > 0 ? nPacked : 1)) != kOK) break;
        if ((status = dev_alloc(g, &d_tab, 4)) != kOK) break;
        if ((status = check(cache_host_alloc(reinterpret_cast<void **>(&h_tab), 4 * sizeof(MixedJob)),
                            "hipHostMalloc")) != kOK) break;
        if ((status = check(cache_host_alloc(reinterpret_cast<void **>(&stageB), nPacked * sizeof(double) + 8),
                            "hipHostMalloc")) != kOK) break;
        beta.P = Pbd;
        beta.V = Vbd;
        // Both densities in, expanded to dense matrices.
        std::memcpy(g->h_stage, Pa, nPacked * sizeof(double));
        std::memcpy(stageB, Pb, nPacked * sizeof(double));
        if ((status = check(hipEventRecord(g->ev0, s), "event")) != kOK) break;
        if ((status = check(hipMemcpyAsync(g->d_packed, g->h_stage, nPacked * sizeof(double), hipMemcpyHostToDevice, s),
                            "copy P alpha")) != kOK) break;
        if ((status = check(hipMemcpyAsync(packedB, stageB, nPacked * sizeof(double), hipMemcpyHostToDevice, s),
                            "copy P beta")) != kOK) break;
        launch_expand(n, g->d_packed, g->d_P, s);
        launch_expand(n, packedB, Pbd, s);
        // Density stage of both spins in one launch: each job stores (rho_s, grad rho_s) per point.
        StepArrays dens[2];
        dens[0].rhoOut = rhoA;
        dens[1] = beta;
        dens[1].rhoOut = rhoB;
        status = mixed ? launch_mixed(s, ev, 2, grids, fs2, h_tab, d_tab, kStageDensity, dens)
                       : launch_f64(s, ev, 2, grids, fs2, h_tab, d_tab, kStageDensity, dens);
        if (status != kOK) break;
        if (g->nBlocks > 0)
            hipLaunchKernelGGL(spin_functional_kernel, dim3(g->nBlocks), dim3(kBlockSize), 0, s, g->d_blocks, g->d_w, fs,
                               g->rhoTol, rhoA, rhoB, g->d_wcoef, wcB, g->d_blockEN, g->d_skip);
        if ((status = check(hipGetLastError(), "spin_functional_kernel")) != kOK) break;
        // Fock stage of both spins in one launch (the second table: the first may still be in flight).
        launch_zero(n, g->d_V, s);
        launch_zero(n, Vbd, s);
        StepArrays fock[2];
        fock[1] = beta;
        fock[1].wcoef = wcB;
        status = mixed ? launch_mixed(s, ev, 2, grids, fs2, h_tab + 2, d_tab + 2, kStageFock, fock)
                       : launch_f64(s, ev, 2, grids, fs2, h_tab + 2, d_tab + 2, kStageFock, fock);
        if (status != kOK) break;
        launch_pack(n, g->d_V, g->d_packed, s);
        launch_pack(n, Vbd, packedB, s);
        if ((status = check(hipGetLastError(), "pack")) != kOK) break;
        if ((status = check(hipEventRecord(g->ev1, s), "event")) != kOK) break;
        if ((status = check(hipMemcpyAsync(g->h_stage, g->d_packed, nPacked * sizeof(double), hipMemcpyDeviceToHost, s),
                            "copy V alpha")) != kOK) break;
        if ((status = check(hipMemcpyAsync(stageB, packedB, nPacked * sizeof(double), hipMemcpyDeviceToHost, s),
                            "copy V beta")) != kOK) break;
        if (g->nBlocks > 0) {
            if ((status = check(hipMemcpyAsync(g->h_blockEN, g->d_blockEN, 2 * g->nBlocks * sizeof(double),
                                               hipMemcpyDeviceToHost, s), "copy block sums")) != kOK) break;
            if ((status = check(hipMemcpyAsync(g->h_skip, g->d_skip, g->nBlocks * sizeof(int), hipMemcpyDeviceToHost, s),
                                "copy skips")) != kOK) break;
        }
        if ((status = check(hipStreamSynchronize(s), "unrestricted quadrature")) != kOK) break;
        for (size_t k = 0; k < nPacked; k++) Va[k] += g->h_stage[k];
        for (size_t k = 0; k < nPacked; k++) Vb[k] += stageB[k];
        // FP64 reductions in block order (DFTIntegrator.c:350-352).
        double e = 0.0, ne = 0.0;
        long long skipped = 0;
        for (int b = 0; b < g->nBlocks; b++) {
            e += g->h_blockEN[2 * b];
            ne += g->h_blockEN[2 * b + 1];
            if (g->h_skip[b] && g->blocks[b].nKept > 0) skipped++;
        }
        *eXC = e;
        *nE = ne;
        (void)hipEventElapsedTime(&msTotal, g->ev0, g->ev1);
        if (stats != nullptr) {
            for (int k = 0; k < XQ_NSTATS; k++) stats[k] = 0.0;
            stats[0] = g->bfValues;
            stats[1] = static_cast<double>(skipped);
            stats[2] = static_cast<double>(msTotal);
            stats[5] = g->bfSquares;
            stats[8] = g->bfSquaresPadded;
        }
    } while (false);
    (void)hipStreamSynchronize(s);   // nothing below is released while a kernel may still use it
    if (mixed) free_step_mixed(g, &beta); else free_step_f64(g, &beta);
    dev_free(g, &rhoA);
    dev_free(g, &rhoB);
    dev_free(g, &wcB);
    dev_free(g, &Pbd);
    dev_free(g, &Vbd);
    dev_free(g, &packedB);
    dev_free(g, &d_tab);
    if (h_tab != nullptr) cache_host_free(h_tab);
    if (stageB != nullptr) cache_host_free(stageB);
    return status;
}

}  // namespace xq
