// The C-ABI of libxcquad (declarations and reference correspondences in include/xcquad.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "xq_internal.h"

namespace xq {

static thread_local std::string g_lastError;

void set_error(const std::string &message) { g_lastError = message; }

int check(hipError_t e, const char *what) {
    if (e == hipSuccess) return kOK;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? kOutOfMemory : kAlgorithmError;
}

long long estimated_points(int accuracy, int nAtoms, const int *Z);

}  // namespace xq

using namespace xq;

extern "C" {

int xq_grid_build(int accuracy, int nAtoms, const int *atomicNumbers, const double *xyzBohr, xq_grid **grid) {
    if (grid == nullptr) { set_error("grid is NULL"); return kInvalidArgument; }
    *grid = nullptr;
    if (accuracy < 0 || accuracy > 4 || nAtoms <= 0 || atomicNumbers == nullptr || xyzBohr == nullptr) {
        set_error("invalid grid arguments");
        return kInvalidArgument;
    }
    for (int a = 0; a < nAtoms; a++) {
        if (atomicNumbers[a] < 1 || atomicNumbers[a] > 118) { set_error("invalid atomic number"); return kInvalidArgument; }
        for (int c = 0; c < 3; c++)
            if (!std::isfinite(xyzBohr[3 * a + c])) { set_error("non-finite coordinate"); return kInvalidArgument; }
    }
    xq_grid *g = new (std::nothrow) xq_grid();
    if (g == nullptr) { set_error("host allocation failed"); return kOutOfMemory; }
    g->accuracy = accuracy;
    g->nAtoms = nAtoms;
    g->Z.assign(atomicNumbers, atomicNumbers + nAtoms);
    g->xyz.assign(xyzBohr, xyzBohr + 3 * nAtoms);
    int status = check(hipGetDevice(&g->device), "hipGetDevice");
    if (status == kOK) status = check(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (status == kOK) status = check(hipEventCreate(&g->ev0), "hipEventCreate");
    if (status == kOK) status = check(hipEventCreate(&g->ev1), "hipEventCreate");
    if (status == kOK) status = check(hipEventCreate(&g->evRho0), "hipEventCreate");
    if (status == kOK) status = check(hipEventCreate(&g->evRho1), "hipEventCreate");
    if (status == kOK) status = check(hipEventCreate(&g->evFock1), "hipEventCreate");
    if (status == kOK) status = check(hipEventCreate(&g->evDens1), "hipEventCreate");
    if (status == kOK) status = build_grid(g);
    if (status != kOK) { xq_free(g); return status; }
    *grid = g;
    return kOK;
}

int xq_grid_counts(const xq_grid *grid, int *nPoints, int *nBlocks) {
    if (grid == nullptr) { set_error("grid is NULL"); return kInvalidArgument; }
    if (nPoints) *nPoints = grid->nPoints;
    if (nBlocks) *nBlocks = grid->nBlocks;
    return kOK;
}

int xq_grid_points(const xq_grid *grid, double *xyzw, int *blockAtom, int *blockCount) {
    if (grid == nullptr) { set_error("grid is NULL"); return kInvalidArgument; }
    if (xyzw != nullptr && grid->nPoints > 0) {
        std::vector<double> c(grid->nPoints);
        const double *src[4] = {grid->d_x, grid->d_y, grid->d_z, grid->d_w};
        for (int k = 0; k < 4; k++) {
            XQ_CHECK(hipMemcpy(c.data(), src[k], grid->nPoints * sizeof(double), hipMemcpyDeviceToHost));
            for (int p = 0; p < grid->nPoints; p++) xyzw[4 * static_cast<size_t>(p) + k] = c[p];
        }
    }
    for (int b = 0; b < grid->nBlocks; b++) {
        if (blockAtom) blockAtom[b] = grid->blocks[b].atom;
        if (blockCount) blockCount[b] = grid->blocks[b].nPoints;
    }
    return kOK;
}

int xq_estimated_points(int accuracy, int nAtoms, const int *atomicNumbers, long long *nPoints) {
    if (accuracy < 0 || accuracy > 4 || nAtoms < 0 || nPoints == nullptr || (nAtoms > 0 && atomicNumbers == nullptr)) {
        set_error("invalid arguments");
        return kInvalidArgument;
    }
    *nPoints = estimated_points(accuracy, nAtoms, atomicNumbers);
    return kOK;
}

int xq_basis_set(xq_grid *grid, int nSpec, const double *spec, int spherical, int precision, int *nFunctions) {
    if (grid == nullptr || spec == nullptr || nSpec <= 0) { set_error("invalid basis arguments"); return kInvalidArgument; }
    int status = set_basis(grid, nSpec, spec, spherical, precision);
    if (status == kOK && nFunctions != nullptr) *nFunctions = grid->nFunctions;
    return status;
}

int xq_integrate_uks(xq_grid *grid, int nIDs, const int *ids, const double *densityA, const double *densityB,
                     double *eXC, double *nElectrons, double *fockA, double *fockB, double *stats) {
    if (grid == nullptr || ids == nullptr || densityA == nullptr || densityB == nullptr || eXC == nullptr ||
        nElectrons == nullptr || fockA == nullptr || fockB == nullptr) {
        set_error("invalid integrate arguments");
        return kInvalidArgument;
    }
    *eXC = 0.0;
    *nElectrons = 0.0;
    if (!grid->hasBasis) { set_error("no basis set on this grid"); return kInvalidArgument; }
    return integrate_uks(grid, nIDs, ids, densityA, densityB, eXC, nElectrons, fockA, fockB, stats);
}

int xq_integrate(xq_grid *grid, int nIDs, const int *ids, const double *densityP, double *eXC, double *nElectrons,
                 double *fockV, double *stats) {
    if (grid == nullptr || ids == nullptr || densityP == nullptr || eXC == nullptr || nElectrons == nullptr ||
        fockV == nullptr) {
        set_error("invalid integrate arguments");
        return kInvalidArgument;
    }
    *eXC = 0.0;
    *nElectrons = 0.0;
    if (!grid->hasBasis) { set_error("no basis set on this grid"); return kInvalidArgument; }
    return integrate(grid, nIDs, ids, densityP, eXC, nElectrons, fockV, stats);
}

int xq_integrate_noalloc(xq_grid *grid, int nIDs, const int *ids, const double *densityP, double *eXC,
                         double *nElectrons, double *fockV, double *stats) {
    if (grid == nullptr || ids == nullptr || densityP == nullptr || eXC == nullptr || nElectrons == nullptr ||
        fockV == nullptr) {
        set_error("invalid integrate arguments");
        return kInvalidArgument;
    }
    *eXC = 0.0;
    *nElectrons = 0.0;
    if (!grid->hasBasis) { set_error("no basis set on this grid"); return kInvalidArgument; }
    return integrate_noalloc(grid, nIDs, ids, densityP, eXC, nElectrons, fockV, stats);
}

int xq_noalloc_scratch(const xq_grid *grid, long long *neededBytes, long long *limitBytes) {
    if (limitBytes != nullptr) *limitBytes = static_cast<long long>(noalloc_scratch_bytes());
    if (neededBytes != nullptr) *neededBytes = grid == nullptr ? 0 : static_cast<long long>(noalloc_needed_bytes(grid));
    return kOK;
}

int xq_gradients(xq_grid *grid, int nIDs, const int *ids, const double *densityP, double *gradients3) {
    if (grid == nullptr || ids == nullptr || densityP == nullptr || gradients3 == nullptr) {
        set_error("invalid gradients arguments");
        return kInvalidArgument;
    }
    if (!grid->hasBasis) { set_error("no basis set on this grid"); return kInvalidArgument; }
    return gradients(grid, nIDs, ids, densityP, gradients3);
}

int xq_gradients_uks(xq_grid *grid, int nIDs, const int *ids, const double *densityA, const double *densityB,
                     double *gradients3) {
    if (grid == nullptr || ids == nullptr || densityA == nullptr || densityB == nullptr || gradients3 == nullptr) {
        set_error("invalid gradients arguments");
        return kInvalidArgument;
    }
    if (!grid->hasBasis) { set_error("no basis set on this grid"); return kInvalidArgument; }
    return gradients_uks(grid, nIDs, ids, densityA, densityB, gradients3);
}

int xq_estimate_bytes(long long nPoints, int nFunctions, double meanKeptFunctions, int precision, long long *bytes) {
    if (bytes == nullptr || nPoints < 0 || nFunctions < 0 || meanKeptFunctions < 0.0 ||
        (precision != XQ_PRECISION_FP64 && precision != XQ_PRECISION_MIXED)) {
        set_error("invalid arguments");
        return kInvalidArgument;
    }
    // Peak device bytes of one handle: during the basis set-up the FP64 values and gradients (32 bytes per stored
    // value) and the per-block kept-function scratch coexist with the mixed-precision planes, which then replace the
    // FP64 values.
    const double G = static_cast<double>(nPoints);
    const double blocks = static_cast<double>((nPoints + kBlockSize - 1) / kBlockSize);
    const double n = static_cast<double>(nFunctions);
    const double values = blocks * meanKeptFunctions * kBlockSize;
    const double grid = 32.0 * G + blocks * sizeof(DevBlock);
    const double kept = 4.0 * blocks * meanKeptFunctions + 4.0 * blocks * n;           // indices + set-up scratch
    const double work = 8.0 * (2.0 * n * n + n * (n + 1) / 2.0) + 20.0 * blocks;     // P, V, packed, block sums
    // Fock tile list and chunk tables (rows padded to chunks of 32: about 10% more stored values than kept).
    double b = grid + kept + work + 32.0 * values + 16.0 * blocks * 16.0 + 32.0 * G;
    if (precision == XQ_PRECISION_MIXED)
        // Phi digits in the density's order and 32-bit gradients (17 bytes per stored value), the density's row-chunk
        // shares of rho and grad rho (4 KB per chunk of 32 x 128 values: 1 byte per stored value), Phi and W digits and
        // 24-bit gradients over the superblocks' union rows (17 bytes per stored value, about 15% more rows), per-point
        // exponents, P digits.
        b += 1.1 * (18.0 + 1.15 * 17.0) * values + 512.0 * blocks + 8.0 * n * n + 4.0 * n;
    else
        // Phi in two operand orders, grad Phi, W (48 bytes per stored value).
        b += 1.1 * 48.0 * values;
    *bytes = static_cast<long long>(b);
    return kOK;
}

int xq_device_memory(long long *freeBytes, long long *totalBytes) {
    if (freeBytes == nullptr || totalBytes == nullptr) { set_error("invalid arguments"); return kInvalidArgument; }
    size_t f = 0, t = 0;
    const int status = check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    // Blocks the library holds idle for reuse are free for its handles.
    *freeBytes = static_cast<long long>(f + cache_idle_bytes());
    *totalBytes = static_cast<long long>(t);
    return status;
}

int xq_batch_create(int maxJobs, xq_batch **batch) {
    if (batch == nullptr || maxJobs < 1 || maxJobs > 4096) { set_error("invalid batch arguments"); return kInvalidArgument; }
    *batch = nullptr;
    xq_batch *b = new (std::nothrow) xq_batch();
    if (b == nullptr) { set_error("host allocation failed"); return kOutOfMemory; }
    b->maxJobs = maxJobs;
    int status = check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    hipEvent_t *evs[6] = {&b->ev0, &b->ev1, &b->evRho0, &b->evRho1, &b->evFock1, &b->evDens1};
    for (hipEvent_t *e : evs)
        if (status == kOK) status = check(hipEventCreate(e), "hipEventCreate");
    if (status == kOK)
        status = check(hipHostMalloc(reinterpret_cast<void **>(&b->h_jobs), maxJobs * sizeof(MixedJob)), "hipHostMalloc");
    if (status == kOK)
        status = check(hipMalloc(reinterpret_cast<void **>(&b->d_jobs), maxJobs * sizeof(MixedJob)), "hipMalloc");
    if (status != kOK) { xq_batch_free(b); return status; }
    *batch = b;
    return kOK;
}

int xq_integrate_batch(xq_batch *batch, int nJobs, xq_grid *const *grids, const int *nIDs, const int *ids,
                       const double *const *densityP, double *eXC, double *nElectrons, double *const *fockV,
                       double *stats) {
    if (batch == nullptr || grids == nullptr || nIDs == nullptr || ids == nullptr || densityP == nullptr ||
        eXC == nullptr || nElectrons == nullptr || fockV == nullptr || nJobs < 1 || nJobs > batch->maxJobs) {
        set_error("invalid batch integrate arguments");
        return kInvalidArgument;
    }
    std::vector<FunctionalSet> fs(nJobs);
    for (int j = 0; j < nJobs; j++) {
        eXC[j] = 0.0;
        nElectrons[j] = 0.0;
        if (grids[j] == nullptr || densityP[j] == nullptr || fockV[j] == nullptr) {
            set_error("invalid batch integrate arguments");
            return kInvalidArgument;
        }
        if (!grids[j]->hasBasis) { set_error("no basis set on a grid of the batch"); return kInvalidArgument; }
        for (int k = 0; k < j; k++)
            if (grids[k] == grids[j]) { set_error("a handle appears twice in one batch"); return kInvalidArgument; }
        const int status = functional_set(nIDs[j], ids + 4 * j, &fs[j]);
        if (status != kOK) return status;
    }
    const StageEvents ev{batch->ev0, batch->ev1, batch->evRho0, batch->evRho1, batch->evFock1, batch->evDens1};
    return integrate_jobs(batch->stream, ev, nJobs, grids, fs.data(), densityP, eXC, nElectrons, fockV, stats,
                          batch->h_jobs, batch->d_jobs);
}

void xq_batch_free(xq_batch *batch) {
    if (batch == nullptr) return;
    if (batch->stream) (void)hipStreamSynchronize(batch->stream);
    if (batch->h_jobs) (void)hipHostFree(batch->h_jobs);
    if (batch->d_jobs) (void)hipFree(batch->d_jobs);
    hipEvent_t evs[6] = {batch->ev0, batch->ev1, batch->evRho0, batch->evRho1, batch->evFock1, batch->evDens1};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    if (batch->stream) (void)hipStreamDestroy(batch->stream);
    delete batch;
}

int xq_device_bytes(const xq_grid *grid, long long *bytes) {
    if (grid == nullptr || bytes == nullptr) { set_error("invalid arguments"); return kInvalidArgument; }
    *bytes = grid->deviceBytes;
    return kOK;
}

void xq_free(xq_grid *grid) {
    if (grid == nullptr) return;
    if (grid->stream) (void)hipStreamSynchronize(grid->stream);
    free_work(grid);
    free_basis(grid);
    dev_free(grid, &grid->d_x);
    dev_free(grid, &grid->d_y);
    dev_free(grid, &grid->d_z);
    dev_free(grid, &grid->d_w);
    dev_free(grid, &grid->d_blocks);
    if (grid->ev0) (void)hipEventDestroy(grid->ev0);
    if (grid->ev1) (void)hipEventDestroy(grid->ev1);
    if (grid->evRho0) (void)hipEventDestroy(grid->evRho0);
    if (grid->evRho1) (void)hipEventDestroy(grid->evRho1);
    if (grid->evFock1) (void)hipEventDestroy(grid->evFock1);
    if (grid->evDens1) (void)hipEventDestroy(grid->evDens1);
    if (grid->stream) (void)hipStreamDestroy(grid->stream);
    delete grid;
}

const char *xq_last_error(void) { return g_lastError.c_str(); }

void xq_test_fail_allocations(long long count) { cache_fail_after(count); }

}  // extern "C"
