// One quadrature on the device: density and gradient, GGA functional, E_xc / N_e and V_xc, for every grid block.
//
// Follows DFTIntegrator_Integrate (pMolecule/QCModel/__extensions__/csource/DFTIntegrator.c:159-509), restricted closed
// shell with a Fock matrix and no gradients:
//   X = P_red Phi (:838), rho_p = sum_m Phi_mp X_mp (:839), grad rho_p = 2 sum_m grad Phi_mp X_mp, sigma = |grad rho|^2
//   (:843-850); a block is skipped when max_p |rho_p| <= rhoTol (:332); vrho and vsigma are scaled by the weights (:336-338);
//   E += sum_p w eps rho, N += sum_p w rho (:350-352); V_mn += sum_p Phi_mp W_np + W_mp Phi_np with
//   W = vrho Phi / 2 + 2 vsigma (grad rho . grad Phi), which is the sum of DFTIntegrator_Fock's three terms (:517-618).
//
// This file holds the host side of one quadrature (P upload and expansion, the precision's kernels, V packing and the
// FP64 reductions of E_xc and N_e in block order).  The kernels are in xq_fused.hip (FP64) and xq_mixed.hip (int8
// matrix cores).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "xq_functional.h"
#include "xq_internal.h"

namespace xq {

namespace {

__global__ void expand_packed_kernel(int n, const double *__restrict__ packed, double *__restrict__ dense) {
    const size_t total = static_cast<size_t>(n) * n;
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
        const int hi = i > j ? i : j, lo = i > j ? j : i;
        const double v = packed[static_cast<size_t>(hi) * (hi + 1) / 2 + lo];
        dense[e] = v;
    }
}

__global__ void pack_lower_kernel(int n, const double *__restrict__ dense, double *__restrict__ packed) {
    const size_t total = static_cast<size_t>(n) * (n + 1) / 2;
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        // Row i holds entries i(i+1)/2 .. i(i+1)/2 + i.
        int i = static_cast<int>((sqrt(8.0 * static_cast<double>(e) + 1.0) - 1.0) * 0.5);
        while (static_cast<size_t>(i) * (i + 1) / 2 > e) i--;
        while (static_cast<size_t>(i + 1) * (i + 2) / 2 <= e) i++;
        const int j = static_cast<int>(e - static_cast<size_t>(i) * (i + 1) / 2);
        packed[e] = dense[static_cast<size_t>(i) * n + j];
    }
}

__global__ void zero_kernel(size_t n, double *__restrict__ a) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x)
        a[e] = 0.0;
}

}  // namespace

void launch_fused_f64(xq_grid *g, const FunctionalSet &fs);
int launch_fused_mixed(xq_grid *g, const FunctionalSet &fs);

void free_work(xq_grid *g) {
    dev_free(&g->d_P);
    dev_free(&g->d_V);
    dev_free(&g->d_packed);
    dev_free(&g->d_blockEN);
    dev_free(&g->d_skip);
}

static int ensure_work(xq_grid *g) {
    if (g->d_P != nullptr) return kOK;
    const size_t n = static_cast<size_t>(g->nFunctions);
    XQ_TRY(dev_alloc(g, &g->d_P, n * n));
    XQ_TRY(dev_alloc(g, &g->d_V, n * n));
    XQ_TRY(dev_alloc(g, &g->d_packed, n * (n + 1) / 2));
    XQ_TRY(dev_alloc(g, &g->d_blockEN, 2 * static_cast<size_t>(g->nBlocks)));
    XQ_TRY(dev_alloc(g, &g->d_skip, static_cast<size_t>(g->nBlocks)));
    g->h_blockEN.assign(2 * static_cast<size_t>(g->nBlocks), 0.0);
    g->h_skip.assign(static_cast<size_t>(g->nBlocks), 0);
    return kOK;
}

int integrate(xq_grid *g, int nIDs, const int *ids, const double *Ppacked, double *eXC, double *nE, double *V,
              double *stats) {
    if (nIDs < 1 || nIDs > 4) { set_error("between one and four functionals are supported"); return kInvalidArgument; }
    FunctionalSet fs;
    fs.n = nIDs;
    for (int k = 0; k < nIDs; k++) {
        if (ids[k] != 101 && ids[k] != 106 && ids[k] != 130 && ids[k] != 131) {
            set_error("unsupported functional ID (supported: 101, 106, 130, 131)");
            return kInvalidArgument;   // DFTFunctionalModel.c:196-202
        }
        fs.id[k] = ids[k];
    }
    XQ_TRY(ensure_work(g));
    const size_t n = static_cast<size_t>(g->nFunctions), nPacked = n * (n + 1) / 2;
    hipStream_t s = g->stream;
    XQ_CHECK(hipMemcpyAsync(g->d_packed, Ppacked, nPacked * sizeof(double), hipMemcpyHostToDevice, s));
    XQ_CHECK(hipEventRecord(g->ev0, s));
    const int grid = static_cast<int>(std::min<size_t>((n * n + 255) / 256, 16384));
    hipLaunchKernelGGL(expand_packed_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, g->nFunctions, g->d_packed, g->d_P);
    hipLaunchKernelGGL(zero_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, n * n, g->d_V);
    XQ_CHECK(hipGetLastError());
    if (g->nBlocks > 0) {
        if (g->precision == XQ_PRECISION_FP64) {
            (void)hipEventRecord(g->evRho0, s);
            launch_fused_f64(g, fs);
            (void)hipEventRecord(g->evRho1, s);
            (void)hipEventRecord(g->evFock1, s);
        } else {
            XQ_TRY(launch_fused_mixed(g, fs));
        }
        XQ_CHECK(hipGetLastError());
    }
    const int pgrid = static_cast<int>(std::min<size_t>((nPacked + 255) / 256, 16384));
    hipLaunchKernelGGL(pack_lower_kernel, dim3(pgrid > 0 ? pgrid : 1), dim3(256), 0, s, g->nFunctions, g->d_V, g->d_packed);
    XQ_CHECK(hipGetLastError());
    XQ_CHECK(hipEventRecord(g->ev1, s));
    std::vector<double> vpacked(nPacked);
    XQ_CHECK(hipMemcpyAsync(vpacked.data(), g->d_packed, nPacked * sizeof(double), hipMemcpyDeviceToHost, s));
    if (g->nBlocks > 0) {
        XQ_CHECK(hipMemcpyAsync(g->h_blockEN.data(), g->d_blockEN, 2 * g->nBlocks * sizeof(double), hipMemcpyDeviceToHost, s));
        XQ_CHECK(hipMemcpyAsync(g->h_skip.data(), g->d_skip, g->nBlocks * sizeof(int), hipMemcpyDeviceToHost, s));
    }
    XQ_CHECK(hipStreamSynchronize(s));
    // FP64 reductions in block order (DFTIntegrator.c:350-352).
    double e = 0.0, ne = 0.0;
    long long skipped = 0;
    for (int b = 0; b < g->nBlocks; b++) {
        e += g->h_blockEN[2 * b];
        ne += g->h_blockEN[2 * b + 1];
        if (g->h_skip[b] && g->blocks[b].nKept > 0) skipped++;
    }
    for (size_t k = 0; k < nPacked; k++) V[k] += vpacked[k];
    *eXC = e;
    *nE = ne;
    if (stats != nullptr) {
        float ms = 0.0f, msRho = 0.0f, msFock = 0.0f;
        (void)hipEventElapsedTime(&ms, g->ev0, g->ev1);
        if (g->nBlocks > 0) {
            (void)hipEventElapsedTime(&msRho, g->evRho0, g->evRho1);
            (void)hipEventElapsedTime(&msFock, g->evRho1, g->evFock1);
        }
        stats[0] = g->bfValues;
        stats[1] = static_cast<double>(skipped);
        stats[2] = static_cast<double>(ms);
        stats[3] = static_cast<double>(msRho);
        stats[4] = static_cast<double>(msFock);
        stats[5] = g->bfSquares;
        // Algorithmic HBM bytes of the two stages: every stored operand read once, every result written once.
        const double nn = static_cast<double>(n) * static_cast<double>(n), G = static_cast<double>(g->nPoints),
                     B = static_cast<double>(g->nBlocks);
        if (g->precision == XQ_PRECISION_FP64) {
            // Fused kernel: values and gradients (4 x 8 bytes per stored value), P, weights, block sums, V.
            stats[6] = 32.0 * static_cast<double>(g->nValueRows) * kBlockSize + 8.0 * nn + 8.0 * G + 20.0 * B + 16.0 * nn;
            stats[7] = 0.0;
        } else {
            // Density: Phi digits (16 KB per chunk), fixed-point gradients (36 KB per chunk), P digits, point
            // exponents, weights in, W coefficients out, block sums.  Fock stage: W formation (Phi digits, gradients,
            // coefficients in, W digits out) and the Fock product (Phi and W digits in, V read and written).
            const double ch = static_cast<double>(g->nChunks);
            stats[6] = ch * (16384.0 + 36864.0) + 8.0 * nn + 512.0 * B + 40.0 * G + 20.0 * B;
            stats[7] = ch * (16384.0 + 36864.0 + 16384.0 + 32.0) + 32.0 * G + 384.0 * B + ch * 2.0 * 16384.0 + 16.0 * nn;
        }
    }
    return kOK;
}

}  // namespace xq
