// xq_integrate_noalloc: one restricted quadrature that allocates nothing.
//
// xq_integrate keeps its per-call work arrays on the handle (dense P and V, the packed staging, per-block E/N sums and
// skip flags, the job table) and allocates them on a handle's first call, with pinned host staging beside them.  This
// variant takes the same arrays from one file-scope device array, `scratch`, of kScratchBytes per GPU, and moves P in
// and V out through the caller's own (pageable) host arrays.  It issues no hipMalloc, hipHostMalloc or hipFree, so a
// caller that cannot allocate (code under graph capture, a serving loop that forbids allocation) can call it once the
// handle is set up (xq_grid_build and xq_basis_set allocate as before).
//
// Scratch layout (each region 256-byte aligned): the one-entry job table, the three reduction results, dense P (n^2),
// dense V (n^2), packed P (n(n+1)/2), packed V in and out (n(n+1)/2), block sums (2 B) and block flags (B).
//
// Results equal xq_integrate's on the same handle and density: the kernels are the same; V is incremented on the device
// by the same FP64 additions the host makes in xq_integrate; E_xc and N_e are summed in block order by one thread, the
// same sequence of FP64 additions as xq_integrate's host loop, so they are equal bit for bit.  V_xc itself comes from FP64
// atomic additions in the Fock stage, whose order varies from call to call in either entry point.
//
// Every call on one GPU uses that GPU's one scratch array: calls that overlap on one GPU (from two threads, or on two
// handles' streams at once) race and corrupt each other's results.  The array keeps the last call's contents; every
// region a call reads is written (or zeroed) by that call first.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>

#include "xq_internal.h"

namespace xq {

namespace {

constexpr size_t kScratchBytes = static_cast<size_t>(64) << 20;   // 64 MiB per GPU, held until the process ends
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratchLock;
void *g_scratchPtr[kMaxDevices] = {};   // device address of `scratch` on each GPU, from hipGetSymbolAddress

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Device address of this GPU's scratch array.  The array's name is a host-side stand-in: only the address that
// hipGetSymbolAddress returns for the current device may be handed to kernels and copies.
int scratch_address(int device, unsigned char **out) {
    std::lock_guard<std::mutex> lock(g_scratchLock);
    if (g_scratchPtr[device] == nullptr) {
        void *p = nullptr;
        XQ_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) { set_error("hipGetSymbolAddress returned no address for the scratch array"); return kAlgorithmError; }
        g_scratchPtr[device] = p;
    }
    *out = static_cast<unsigned char *>(g_scratchPtr[device]);
    return kOK;
}

struct Regions {
    MixedJob *job;
    double *sums;          // E_xc, N_e, skipped blocks
    double *P, *V;         // dense n x n
    double *Pin, *Vio;     // packed
    double *blockEN;
    int *skip;
    size_t bytes;
};

Regions carve(unsigned char *base, size_t n, size_t nBlocks) {
    Regions r{};
    const size_t nPacked = n * (n + 1) / 2, nb = nBlocks > 0 ? nBlocks : 1;
    size_t o = 0;
    auto take = [&](size_t bytes) { unsigned char *p = base == nullptr ? nullptr : base + o; o += align_up(bytes); return p; };
    r.job = reinterpret_cast<MixedJob *>(take(sizeof(MixedJob)));
    r.sums = reinterpret_cast<double *>(take(4 * sizeof(double)));
    r.P = reinterpret_cast<double *>(take(n * n * sizeof(double)));
    r.V = reinterpret_cast<double *>(take(n * n * sizeof(double)));
    r.Pin = reinterpret_cast<double *>(take(nPacked * sizeof(double)));
    r.Vio = reinterpret_cast<double *>(take(nPacked * sizeof(double)));
    r.blockEN = reinterpret_cast<double *>(take(2 * nb * sizeof(double)));
    r.skip = reinterpret_cast<int *>(take(nb * sizeof(int)));
    r.bytes = o;
    return r;
}

// io (packed lower triangle) += dense V: the addition xq_integrate makes on the host.
__global__ void pack_add_kernel(int n, const double *__restrict__ dense, double *__restrict__ io) {
    const size_t total = static_cast<size_t>(n) * (n + 1) / 2;
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        int i = static_cast<int>((sqrt(8.0 * static_cast<double>(e) + 1.0) - 1.0) * 0.5);
        while (static_cast<size_t>(i) * (i + 1) / 2 > e) i--;
        while (static_cast<size_t>(i + 1) * (i + 2) / 2 <= e) i++;
        const int j = static_cast<int>(e - static_cast<size_t>(i) * (i + 1) / 2);
        io[e] = io[e] + dense[static_cast<size_t>(i) * n + j];
    }
}

// E_xc and N_e as the FP64 sums of the block partial sums in block order (thread 0 and thread 1 of one work-group, from
// tiles staged in LDS), and the count of skipped blocks that keep functions (thread 2).
__global__ void block_sums_kernel(int nBlocks, const double *__restrict__ blockEN, const int *__restrict__ skip,
                                  const DevBlock *__restrict__ blocks, double *__restrict__ out) {
    __shared__ double sE[256], sN[256];
    __shared__ int sS[256];
    double acc = 0.0;
    long long skipped = 0;
    const int t = threadIdx.x;
    for (int b0 = 0; b0 < nBlocks; b0 += 256) {
        const int b = b0 + t;
        if (b < nBlocks) {
            sE[t] = blockEN[2 * static_cast<size_t>(b)];
            sN[t] = blockEN[2 * static_cast<size_t>(b) + 1];
            sS[t] = (skip[b] != 0 && blocks[b].nKept > 0) ? 1 : 0;
        }
        __syncthreads();
        const int m = nBlocks - b0 < 256 ? nBlocks - b0 : 256;
        if (t == 0) {
            for (int k = 0; k < m; k++) acc += sE[k];
        } else if (t == 1) {
            for (int k = 0; k < m; k++) acc += sN[k];
        } else if (t == 2) {
            for (int k = 0; k < m; k++) skipped += sS[k];
        }
        __syncthreads();
    }
    if (t < 2) out[t] = acc;
    if (t == 2) out[2] = static_cast<double>(skipped);
}

}  // namespace

size_t noalloc_scratch_bytes() { return kScratchBytes; }

size_t noalloc_needed_bytes(const xq_grid *g) {
    return carve(nullptr, static_cast<size_t>(g->nFunctions), static_cast<size_t>(g->nBlocks)).bytes;
}

// The queued part of one call; the caller drains the stream and restores the handle's work pointers afterwards.
static int integrate_noalloc_queued(xq_grid *g, const FunctionalSet &fs, const Regions &r, const double *Ppacked,
                                    double *V, double *host3) {
    const hipStream_t s = g->stream;
    const size_t n = static_cast<size_t>(g->nFunctions), nPacked = n * (n + 1) / 2;
    XQ_CHECK(hipMemcpyAsync(r.Pin, Ppacked, nPacked * sizeof(double), hipMemcpyHostToDevice, s));
    XQ_CHECK(hipMemcpyAsync(r.Vio, V, nPacked * sizeof(double), hipMemcpyHostToDevice, s));
    XQ_CHECK(hipEventRecord(g->ev0, s));
    launch_expand_P(g, s);   // the handle's work pointers are the scratch regions for this call
    launch_zero_V(g, s);   // V must start at zero; the density stage writes every block's sums and flag
    XQ_CHECK(hipGetLastError());
    MixedJob table[1];   // pageable: copied to the device before launch_* returns; the stream is drained before return
    std::memset(static_cast<void *>(table), 0, sizeof(table));
    xq_grid *grids[1] = {g};
    const StageEvents ev{g->ev0, g->ev1, g->evRho0, g->evRho1, g->evFock1, g->evDens1};
    if (g->precision == XQ_PRECISION_FP64) {
        XQ_TRY(launch_f64(s, ev, 1, grids, &fs, table, r.job));
    } else {
        XQ_TRY(launch_mixed(s, ev, 1, grids, &fs, table, r.job));
    }
    const int pgrid = static_cast<int>(std::min<size_t>((nPacked + 255) / 256, 16384));
    hipLaunchKernelGGL(pack_add_kernel, dim3(pgrid > 0 ? pgrid : 1), dim3(256), 0, s, g->nFunctions, r.V, r.Vio);
    XQ_CHECK(hipEventRecord(g->ev1, s));
    hipLaunchKernelGGL(block_sums_kernel, dim3(1), dim3(256), 0, s, g->nBlocks, r.blockEN, r.skip, g->d_blocks, r.sums);
    XQ_CHECK(hipGetLastError());
    XQ_CHECK(hipMemcpyAsync(V, r.Vio, nPacked * sizeof(double), hipMemcpyDeviceToHost, s));
    XQ_CHECK(hipMemcpyAsync(host3, r.sums, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    XQ_CHECK(hipStreamSynchronize(s));
    return kOK;
}

int integrate_noalloc(xq_grid *g, int nIDs, const int *ids, const double *Ppacked, double *eXC, double *nE, double *V,
                      double *stats) {
    FunctionalSet fs;
    XQ_TRY(functional_set(nIDs, ids, &fs));
    int device = -1;
    XQ_CHECK(hipGetDevice(&device));
    if (device != g->device) {
        set_error("xq_integrate_noalloc: the handle's device must be current");
        return kInvalidArgument;
    }
    if (device < 0 || device >= kMaxDevices) { set_error("xq_integrate_noalloc: device index out of range"); return kInvalidArgument; }
    const size_t need = noalloc_needed_bytes(g);
    if (need > kScratchBytes) {
        set_error("xq_integrate_noalloc: this molecule needs " + std::to_string(need) + " bytes of scratch (" +
                  std::to_string(g->nFunctions) + " basis functions, " + std::to_string(g->nBlocks) +
                  " blocks); the limit is " + std::to_string(kScratchBytes) + " bytes, about 1,600 functions");
        return kInvalidArgument;
    }
    unsigned char *base = nullptr;
    XQ_TRY(scratch_address(device, &base));
    const Regions r = carve(base, static_cast<size_t>(g->nFunctions), static_cast<size_t>(g->nBlocks));
    // The stages read the handle's work pointers: point them at the scratch regions for this call only.
    double *const oP = g->d_P, *const oV = g->d_V, *const oPacked = g->d_packed, *const oEN = g->d_blockEN;
    int *const oSkip = g->d_skip;
    g->d_P = r.P;
    g->d_V = r.V;
    g->d_packed = r.Pin;
    g->d_blockEN = r.blockEN;
    g->d_skip = r.skip;
    double host3[3] = {0.0, 0.0, 0.0};
    const int status = integrate_noalloc_queued(g, fs, r, Ppacked, V, host3);
    if (status != kOK) (void)hipStreamSynchronize(g->stream);
    g->d_P = oP;
    g->d_V = oV;
    g->d_packed = oPacked;
    g->d_blockEN = oEN;
    g->d_skip = oSkip;
    if (status != kOK) return status;
    *eXC = host3[0];
    *nE = host3[1];
    if (stats != nullptr) {
        float ms = 0.0f, msRho = 0.0f, msFock = 0.0f, msProduct = 0.0f;
        (void)hipEventElapsedTime(&ms, g->ev0, g->ev1);
        (void)hipEventElapsedTime(&msRho, g->evRho0, g->evRho1);
        (void)hipEventElapsedTime(&msFock, g->evRho1, g->evFock1);
        (void)hipEventElapsedTime(&msProduct, g->evRho0, g->evDens1);
        double rb, fb, pb;
        stage_bytes(g, &rb, &fb, &pb);
        stats[0] = g->bfValues;
        stats[1] = host3[2];
        stats[2] = static_cast<double>(ms);
        stats[3] = static_cast<double>(msRho);
        stats[4] = static_cast<double>(msFock);
        stats[5] = g->bfSquares;
        stats[6] = rb;
        stats[7] = fb;
        stats[8] = g->bfSquaresPadded;
        stats[9] = static_cast<double>(msProduct);
        stats[10] = pb;
    }
    return kOK;
}

}  // namespace xq
