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
// FP64 reductions of E_xc and N_e in block order).  The kernels are in xq_f64.hip (FP64 matrix cores) and xq_mixed.hip
// (int8 matrix cores).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

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

// Dense n x n P on the device from a packed lower triangle on the host (through the handle's pinned staging buffer).
// The handle's packed matrix buffer to its dense P, its dense V to zero, its dense V to the packed buffer.
void launch_expand_P(xq_grid *g, hipStream_t s) {
    const size_t n = static_cast<size_t>(g->nFunctions);
    const int grid = static_cast<int>(std::min<size_t>((n * n + 255) / 256, 16384));
    hipLaunchKernelGGL(expand_packed_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, g->nFunctions, g->d_packed,
                       g->d_P);
}

void launch_zero_V(xq_grid *g, hipStream_t s) {
    const size_t n = static_cast<size_t>(g->nFunctions);
    const int grid = static_cast<int>(std::min<size_t>((n * n + 255) / 256, 16384));
    hipLaunchKernelGGL(zero_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, n * n, g->d_V);
}

void launch_expand(int n, const double *packed, double *dense, hipStream_t s) {
    const size_t nn = static_cast<size_t>(n) * n;
    const int grid = static_cast<int>(std::min<size_t>((nn + 255) / 256, 16384));
    hipLaunchKernelGGL(expand_packed_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, n, packed, dense);
}

void launch_zero(int n, double *dense, hipStream_t s) {
    const size_t nn = static_cast<size_t>(n) * n;
    const int grid = static_cast<int>(std::min<size_t>((nn + 255) / 256, 16384));
    hipLaunchKernelGGL(zero_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, nn, dense);
}

void launch_pack(int n, const double *dense, double *packed, hipStream_t s) {
    const size_t nPacked = static_cast<size_t>(n) * (n + 1) / 2;
    const int grid = static_cast<int>(std::min<size_t>((nPacked + 255) / 256, 16384));
    hipLaunchKernelGGL(pack_lower_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, n, dense, packed);
}

void launch_pack_V(xq_grid *g, hipStream_t s) {
    const size_t n = static_cast<size_t>(g->nFunctions), nPacked = n * (n + 1) / 2;
    const int grid = static_cast<int>(std::min<size_t>((nPacked + 255) / 256, 16384));
    hipLaunchKernelGGL(pack_lower_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, g->nFunctions, g->d_V,
                       g->d_packed);
}

int upload_density(xq_grid *g, const double *Ppacked, hipStream_t s) {
    XQ_TRY(ensure_work(g));
    const size_t n = static_cast<size_t>(g->nFunctions), nPacked = n * (n + 1) / 2;
    std::memcpy(g->h_stage, Ppacked, nPacked * sizeof(double));
    XQ_CHECK(hipMemcpyAsync(g->d_packed, g->h_stage, nPacked * sizeof(double), hipMemcpyHostToDevice, s));
    const int grid = static_cast<int>(std::min<size_t>((n * n + 255) / 256, 16384));
    hipLaunchKernelGGL(expand_packed_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, g->nFunctions, g->d_packed, g->d_P);
    XQ_CHECK(hipGetLastError());
    return kOK;
}

// The per-call work arrays and the one-entry job table.  Each set is complete or absent: an allocation that fails part
// way releases what the set already holds, so the next call allocates it again from scratch.
static void free_work_arrays(xq_grid *g) {
    if (g->h_stage != nullptr) { cache_host_free(g->h_stage); g->h_stage = nullptr; }
    if (g->h_blockEN != nullptr) { cache_host_free(g->h_blockEN); g->h_blockEN = nullptr; }
    if (g->h_skip != nullptr) { cache_host_free(g->h_skip); g->h_skip = nullptr; }
    dev_free(g, &g->d_P);
    dev_free(g, &g->d_V);
    dev_free(g, &g->d_packed);
    dev_free(g, &g->d_blockEN);
    dev_free(g, &g->d_skip);
}

static void free_job_table(xq_grid *g) {
    if (g->h_job != nullptr) { cache_host_free(g->h_job); g->h_job = nullptr; }
    dev_free(g, &g->d_job);
}

void free_work(xq_grid *g) {
    free_job_table(g);
    free_work_arrays(g);
}

static int alloc_work_arrays(xq_grid *g) {
    const size_t n = static_cast<size_t>(g->nFunctions);
    XQ_TRY(dev_alloc(g, &g->d_P, n * n));
    XQ_TRY(dev_alloc(g, &g->d_V, n * n));
    XQ_TRY(dev_alloc(g, &g->d_packed, n * (n + 1) / 2));
    XQ_TRY(dev_alloc(g, &g->d_blockEN, 2 * static_cast<size_t>(g->nBlocks)));
    XQ_TRY(dev_alloc(g, &g->d_skip, static_cast<size_t>(g->nBlocks)));
    const size_t nb = static_cast<size_t>(g->nBlocks > 0 ? g->nBlocks : 1);
    XQ_CHECK(cache_host_alloc(reinterpret_cast<void **>(&g->h_stage), n * (n + 1) / 2 * sizeof(double) + 8));
    XQ_CHECK(cache_host_alloc(reinterpret_cast<void **>(&g->h_blockEN), 2 * nb * sizeof(double)));
    XQ_CHECK(cache_host_alloc(reinterpret_cast<void **>(&g->h_skip), nb * sizeof(int)));
    return kOK;
}

int ensure_work(xq_grid *g) {
    if (g->d_P != nullptr && g->d_V != nullptr && g->d_packed != nullptr && g->d_blockEN != nullptr &&
        g->d_skip != nullptr && g->h_stage != nullptr && g->h_blockEN != nullptr && g->h_skip != nullptr)
        return kOK;
    free_work_arrays(g);
    const int status = alloc_work_arrays(g);
    if (status != kOK) free_work_arrays(g);
    return status;
}

int ensure_job_table(xq_grid *g) {
    if (g->h_job != nullptr && g->d_job != nullptr) return kOK;
    free_job_table(g);
    int status = check(cache_host_alloc(reinterpret_cast<void **>(&g->h_job), sizeof(MixedJob)), "hipHostMalloc");
    if (status == kOK) status = dev_alloc(g, &g->d_job, 1);
    if (status != kOK) free_job_table(g);
    return status;
}

int functional_set(int nIDs, const int *ids, FunctionalSet *fs) {
    if (nIDs < 1 || nIDs > 4) { set_error("between one and four functionals are supported"); return kInvalidArgument; }
    fs->n = nIDs;
    int nLDA = 0;
    for (int k = 0; k < nIDs; k++) {
        const int id = ids[k];
        if (id != 1 && id != 7 && id != 8 && id != 101 && id != 106 && id != 130 && id != 131 && id != 402) {
            set_error("unsupported functional ID (supported: 1, 7, 8, 101, 106, 130, 131, 402)");
            return kInvalidArgument;   // DFTFunctionalModel.c:196-202
        }
        nLDA += (id == 1 || id == 7 || id == 8) ? 1 : 0;
        fs->id[k] = id;
    }
    if (nLDA > 0 && nLDA < nIDs) {
        // The named models are all LDA or all GGA (DFTDefinitions.py:212-222); a model mixing the two families is refused.
        set_error("LDA and GGA functionals cannot be combined in one model");
        return kInvalidArgument;
    }
    for (int k = nIDs; k < 4; k++) fs->id[k] = 0;
    return kOK;
}

// Algorithmic HBM bytes of the two stages of one handle's quadrature: every stored value of a kept function at a grid
// point (sum over blocks of kept functions x points, not the rows padded to chunks of 32) read once, every result
// written once.
void stage_bytes(const xq_grid *g, double *rho, double *fock, double *product) {
    const double n = static_cast<double>(g->nFunctions), nn = n * n, G = static_cast<double>(g->nPoints),
                 B = static_cast<double>(g->nBlocks), values = g->bfValues;
    if (g->precision == XQ_PRECISION_FP64) {
        // Density: Phi and grad Phi (32 bytes per value), P, weights in and W coefficients out (40 bytes per point),
        // block sums.  Fock stage: W formation (Phi and grad Phi in, W out: 40 bytes per value; the coefficients in) and
        // the Fock product (Phi and W in: 16 bytes per value; V read and written).
        *rho = 32.0 * values + 8.0 * nn + 40.0 * G + 20.0 * B;
        *fock = 56.0 * values + 32.0 * G + 16.0 * nn;
        *product = *rho;   // one density kernel, functional included
    } else {
        // Density: Phi as 5 digits and grad Phi as 3 x 32-bit fixed point (17 bytes per value), P digits (8 bytes per
        // entry), weights in, W coefficients out and the point exponents (44 bytes per point), block sums.  Fock stage,
        // over the superblocks' union rows: W formation (Phi digits and 24-bit grad Phi in, W digits out: 17 bytes per
        // union value; coefficients and gradient exponents in) and the Fock product (Phi and W digits in: 8 bytes per
        // union value; V read and written).
        const double union_values = g->bfUnion;
        *rho = 17.0 * values + 8.0 * nn + 44.0 * G + 20.0 * B;
        *fock = 25.0 * union_values + 35.0 * G + 16.0 * nn;
        // The density product kernel alone: Phi digits and grad Phi, P digits and row exponents, the point exponents
        // (4 bytes per point) in; each row chunk's share of rho and grad rho (4 x 8 bytes per point) out.
        *product = 17.0 * values + 8.0 * nn + 4.0 * n + 4.0 * G + 32.0 * kBlockSize * static_cast<double>(g->nChunks);
    }
}

static int integrate_jobs_queued(hipStream_t s, const StageEvents &ev, int nJobs, xq_grid *const *grids,
                                 const FunctionalSet *fs, const double *const *Ppacked, double *eXC, double *nE,
                                 double *const *V, double *stats, MixedJob *h_table, MixedJob *d_table) {
    const int precision = grids[0]->precision;
    for (int j = 0; j < nJobs; j++) {
        if (grids[j]->precision != precision) {
            set_error("the jobs of one batch must share their precision");
            return kInvalidArgument;
        }
        XQ_TRY(ensure_work(grids[j]));
    }
    // P crosses PCIe from each handle's pinned staging buffer (an asynchronous DMA; the buffer then receives V).
    for (int j = 0; j < nJobs; j++) {
        const xq_grid *g = grids[j];
        const size_t n = static_cast<size_t>(g->nFunctions), nPacked = n * (n + 1) / 2;
        std::memcpy(g->h_stage, Ppacked[j], nPacked * sizeof(double));
        XQ_CHECK(hipMemcpyAsync(g->d_packed, g->h_stage, nPacked * sizeof(double), hipMemcpyHostToDevice, s));
    }
    XQ_CHECK(hipEventRecord(ev.ev0, s));
    for (int j = 0; j < nJobs; j++) {
        const xq_grid *g = grids[j];
        const size_t n = static_cast<size_t>(g->nFunctions);
        const int grid = static_cast<int>(std::min<size_t>((n * n + 255) / 256, 16384));
        hipLaunchKernelGGL(expand_packed_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, g->nFunctions, g->d_packed,
                           g->d_P);
        hipLaunchKernelGGL(zero_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, n * n, g->d_V);
    }
    XQ_CHECK(hipGetLastError());
    if (precision == XQ_PRECISION_FP64) {
        XQ_TRY(launch_f64(s, ev, nJobs, grids, fs, h_table, d_table));
    } else {
        XQ_TRY(launch_mixed(s, ev, nJobs, grids, fs, h_table, d_table));
    }
    XQ_CHECK(hipGetLastError());
    for (int j = 0; j < nJobs; j++) {
        const xq_grid *g = grids[j];
        const size_t n = static_cast<size_t>(g->nFunctions), nPacked = n * (n + 1) / 2;
        const int pgrid = static_cast<int>(std::min<size_t>((nPacked + 255) / 256, 16384));
        hipLaunchKernelGGL(pack_lower_kernel, dim3(pgrid > 0 ? pgrid : 1), dim3(256), 0, s, g->nFunctions, g->d_V,
                           g->d_packed);
    }
    XQ_CHECK(hipGetLastError());
    XQ_CHECK(hipEventRecord(ev.ev1, s));
    // Results into the pinned buffers (asynchronous DMAs ordered after the kernels), then one wait.
    for (int j = 0; j < nJobs; j++) {
        xq_grid *g = grids[j];
        const size_t n = static_cast<size_t>(g->nFunctions), nPacked = n * (n + 1) / 2;
        XQ_CHECK(hipMemcpyAsync(g->h_stage, g->d_packed, nPacked * sizeof(double), hipMemcpyDeviceToHost, s));
        if (g->nBlocks > 0) {
            XQ_CHECK(hipMemcpyAsync(g->h_blockEN, g->d_blockEN, 2 * g->nBlocks * sizeof(double), hipMemcpyDeviceToHost,
                                    s));
            XQ_CHECK(hipMemcpyAsync(g->h_skip, g->d_skip, g->nBlocks * sizeof(int), hipMemcpyDeviceToHost, s));
        }
    }
    XQ_CHECK(hipStreamSynchronize(s));
    double bfv = 0.0, bfs = 0.0, bfp = 0.0, rhoBytes = 0.0, fockBytes = 0.0, productBytes = 0.0;
    long long skipped = 0;
    for (int j = 0; j < nJobs; j++) {
        const xq_grid *g = grids[j];
        // FP64 reductions in block order (DFTIntegrator.c:350-352).
        double e = 0.0, ne = 0.0;
        for (int b = 0; b < g->nBlocks; b++) {
            e += g->h_blockEN[2 * b];
            ne += g->h_blockEN[2 * b + 1];
            if (g->h_skip[b] && g->blocks[b].nKept > 0) skipped++;
        }
        const size_t n = static_cast<size_t>(g->nFunctions), nPacked = n * (n + 1) / 2;
        for (size_t k = 0; k < nPacked; k++) V[j][k] += g->h_stage[k];
        eXC[j] = e;
        nE[j] = ne;
        bfv += g->bfValues;
        bfs += g->bfSquares;
        bfp += g->bfSquaresPadded;
        double rb, fb, pb;
        stage_bytes(g, &rb, &fb, &pb);
        rhoBytes += rb;
        fockBytes += fb;
        productBytes += pb;
    }
    if (stats != nullptr) {
        float ms = 0.0f, msRho = 0.0f, msFock = 0.0f, msProduct = 0.0f;
        (void)hipEventElapsedTime(&ms, ev.ev0, ev.ev1);
        (void)hipEventElapsedTime(&msRho, ev.evRho0, ev.evRho1);
        (void)hipEventElapsedTime(&msFock, ev.evRho1, ev.evFock1);
        (void)hipEventElapsedTime(&msProduct, ev.evRho0, ev.evDens1);
        stats[0] = bfv;
        stats[1] = static_cast<double>(skipped);
        stats[2] = static_cast<double>(ms);
        stats[3] = static_cast<double>(msRho);
        stats[4] = static_cast<double>(msFock);
        stats[5] = bfs;
        stats[6] = rhoBytes;
        stats[7] = fockBytes;
        stats[8] = bfp;
        stats[9] = static_cast<double>(msProduct);
        stats[10] = productBytes;
    }
    return kOK;
}

// A failure part way leaves work queued on the stream that uses the handles' arrays; the stream is drained before the
// error returns, so that no array is released (or reused by another handle) while a kernel still reads it.
int integrate_jobs(hipStream_t s, const StageEvents &ev, int nJobs, xq_grid *const *grids, const FunctionalSet *fs,
                   const double *const *Ppacked, double *eXC, double *nE, double *const *V, double *stats,
                   MixedJob *h_table, MixedJob *d_table) {
    const int status = integrate_jobs_queued(s, ev, nJobs, grids, fs, Ppacked, eXC, nE, V, stats, h_table, d_table);
    if (status != kOK) (void)hipStreamSynchronize(s);
    return status;
}

int integrate(xq_grid *g, int nIDs, const int *ids, const double *Ppacked, double *eXC, double *nE, double *V,
              double *stats) {
    FunctionalSet fs;
    XQ_TRY(functional_set(nIDs, ids, &fs));
    XQ_TRY(ensure_job_table(g));
    const StageEvents ev{g->ev0, g->ev1, g->evRho0, g->evRho1, g->evFock1, g->evDens1};
    xq_grid *grids[1] = {g};
    const double *P[1] = {Ppacked};
    double *Vs[1] = {V};
    return integrate_jobs(g->stream, ev, 1, grids, &fs, P, eXC, nE, Vs, stats, g->h_job, g->d_job);
}

}  // namespace xq
