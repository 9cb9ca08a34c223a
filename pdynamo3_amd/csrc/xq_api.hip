// The C-ABI of libxcquad (declarations and reference correspondences in include/xcquad.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>

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
    int status = check(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (status == kOK) status = check(hipEventCreate(&g->ev0), "hipEventCreate");
    if (status == kOK) status = check(hipEventCreate(&g->ev1), "hipEventCreate");
    if (status == kOK) status = check(hipEventCreate(&g->evRho0), "hipEventCreate");
    if (status == kOK) status = check(hipEventCreate(&g->evRho1), "hipEventCreate");
    if (status == kOK) status = check(hipEventCreate(&g->evFock1), "hipEventCreate");
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

int xq_estimate_bytes(long long nPoints, int nFunctions, double meanKeptFunctions, int precision, long long *bytes) {
    if (bytes == nullptr || nPoints < 0 || nFunctions < 0 || meanKeptFunctions < 0.0) {
        set_error("invalid arguments");
        return kInvalidArgument;
    }
    const double blocks = static_cast<double>((nPoints + kBlockSize - 1) / kBlockSize);
    const double elt = precision == XQ_PRECISION_FP64 ? 8.0 : 4.0;
    const double n = static_cast<double>(nFunctions);
    double b = 32.0 * nPoints                                   // points and weights
             + blocks * sizeof(DevBlock)                       // block table
             + blocks * meanKeptFunctions * (4.0 + 4.0 * elt * kBlockSize)   // kept indices + values and gradients
             + 8.0 * (2.0 * n * n + n * (n + 1) / 2.0)         // dense P, V and the packed staging
             + (precision == XQ_PRECISION_FP64 ? 0.0 : 4.0 * n * n)
             + 48.0 * nPoints + 20.0 * blocks;                 // per-point and per-block work arrays
    *bytes = static_cast<long long>(b);
    return kOK;
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
    dev_free(&grid->d_x);
    dev_free(&grid->d_y);
    dev_free(&grid->d_z);
    dev_free(&grid->d_w);
    dev_free(&grid->d_blocks);
    if (grid->ev0) (void)hipEventDestroy(grid->ev0);
    if (grid->ev1) (void)hipEventDestroy(grid->ev1);
    if (grid->evRho0) (void)hipEventDestroy(grid->evRho0);
    if (grid->evRho1) (void)hipEventDestroy(grid->evRho1);
    if (grid->evFock1) (void)hipEventDestroy(grid->evFock1);
    if (grid->stream) (void)hipStreamDestroy(grid->stream);
    delete grid;
}

const char *xq_last_error(void) { return g_lastError.c_str(); }

}  // extern "C"
