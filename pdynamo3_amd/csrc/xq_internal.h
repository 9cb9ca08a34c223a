// Internal definitions of libxcquad: the handle, device data layout and error handling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/xcquad.h"
#include "xq_functional.h"

namespace xq {

// Status codes (pCore/__extensions__/cinclude/Status.h:24-33).
enum Status { kOK = 0, kAlgorithmError = 1, kInvalidArgument = 3, kNonConformable = 6, kOutOfMemory = 7 };

// Points per grid block (DFTGrid.c:82); block rows of basis values are stored with this stride.
constexpr int kBlockSize = 128;
// Largest shell supported by the evaluation kernels: up to l = 3 in one shell (SPDF) -> 20 Cartesian functions.
constexpr int kMaxCBF = 20;
constexpr int kMaxPrim = 16;
constexpr int kMaxL = 3;

void set_error(const std::string &message);

// Shell description on the device (one per shell of every atom).
struct DevShell {
    double x, y, z;            // centre (bohr)
    int atom;
    int lLow, lHigh;
    int nPrim, nCBF, nBasis;
    int fStart;                // index of the shell's first function
    int coef;                  // offset into the coefficient array: exponents[nPrim], then ccbf[nPrim][nCBF]
    int trans;                 // offset into the c2s array (nCBF x nBasis), -1 for Cartesian
    int pow;                   // offset into the Cartesian power array (3 ints per Cartesian function)
    double bound[kMaxPrim];    // screening prefactor per primitive: max_c |ccbf_qc| * max_s sum_c |c2s_cs|
};

// One grid block: at most kBlockSize consecutive points of one atom (DFTGrid.c:213-238).
struct DevBlock {
    int atom;
    int pStart;                // first point
    int nPoints;
    int nKept;                 // functions kept by the |phi| > bfTol filter
    long long vOffset;         // offset of the block's first value row (rows of kBlockSize entries)
    int iOffset;               // offset of the block's kept-function indices
    int pad;
};

// One job of a mixed-precision quadrature launch: a handle's device arrays, its functional, and the first work-group of
// the job in each stage's launch (0 for a single job; running offsets when several jobs share one launch).
struct MixedJob {
    const DevBlock *blocks;
    const int *chunkOffset, *chunkBlock, *idx;
    const int4 *tiles;
    const signed char *DA, *EA, *DB, *EB, *EG;
    const int *GF;
    const double *P, *w;
    int *eP;
    unsigned long long *Pq;
    double *wcoef, *blockEN, *V;
    double *rhoOut;                    // when set, the density kernels store (rho, grad rho) per point here and stop
    int *skip;
    signed char *DW, *EW;
    const double *FPB, *GPB, *FPA;     // FP64 path: Phi and grad Phi in the density's operand order, Phi in the Fock's
    double *WPA;                       // FP64 path: W in the Fock's operand order (per step)
    // Mixed path, Fock side: superblocks (up to four consecutive blocks of one atom) over the union of their kept
    // functions.  supers[s] = {first block, blocks, offset into uIdx, union size}; per block, its first union chunk;
    // per union chunk, its block; per union chunk, grad Phi as 24-bit fixed point in the W kernel's order (GHU, GLU).
    const int4 *supers, *uTiles;
    const int *uIdx, *uChunkOffset, *uChunkBlock;
    const short *GHU;
    const unsigned char *GLU;
    FunctionalSet fs;
    double rhoTol;
    int n, nBlocks, nChunks, nTiles;
    int rowBase, blockBase, chunkBase, groupBase;
    // Density launch: blocks order[xcdStart[x] .. xcdStart[x + 1]) run on XCD x, at that XCD's slots from xcdBase[x];
    // each entry {block, first chunk, first kept index, kept functions x 256 + points}.
    const int4 *order;
    int xcdStart[9], xcdBase[8];
    // Mixed density launch: one work-group per row chunk of a block; items[itemStart[x] .. itemStart[x + 1]) run on
    // XCD x from slot itemBase[x], each {chunk, kept functions x 1024 + row chunk, first kept index, block}; part
    // receives each chunk's share of rho and grad rho, [chunk][4][point].
    const int4 *items;
    int itemStart[9], itemBase[8];
    double *part;
};

}  // namespace xq

// The handle.  Device arrays are prefixed d_.
struct xq_grid {
    int accuracy = 2;
    int device = 0;                      // the GPU that was current when the handle was built
    int nAtoms = 0;
    std::vector<int> Z;
    std::vector<double> xyz;
    double bfTol = 1.0e-10, rhoTol = 1.0e-15;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evRho0 = nullptr, evRho1 = nullptr, evFock1 = nullptr, evDens1 = nullptr;

    // Grid.
    int nPoints = 0, nBlocks = 0;
    std::vector<xq::DevBlock> blocks;     // host copy
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr, *d_w = nullptr;
    xq::DevBlock *d_blocks = nullptr;
    // Density launch order: the blocks cut into 8 contiguous ranges of equal estimated cost (one per XCD, so that each
    // XCD's L2 holds one region's P rows), each range in order of decreasing cost (the longest blocks start first).
    std::vector<int4> h_order;
    int4 *d_order = nullptr;
    int xcdStart[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // The mixed density launch's items: each block of the order above cut into its row chunks, consecutive on the
    // block's XCD; per chunk, its share of rho and grad rho (d_part).
    std::vector<int4> h_items;
    int4 *d_items = nullptr;
    int itemStart[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double *d_part = nullptr;

    // Basis.
    int nFunctions = 0, nShells = 0, spherical = 1, precision = 0;
    bool hasBasis = false;
    xq::DevShell *d_shells = nullptr;
    double *d_coef = nullptr, *d_trans = nullptr;
    int *d_pow = nullptr;
    long long nValueRows = 0;             // sum over blocks of kept functions
    int *d_idx = nullptr;                 // kept-function indices, per block
    std::vector<int> h_shellOf;           // shell of each function (host), and its device copy for the XC gradients
    std::vector<int> h_funcAtom;          // atom of each function (host)
    int *d_shellOf = nullptr;
    void *d_f = nullptr, *d_fx = nullptr, *d_fy = nullptr, *d_fz = nullptr;   // values and gradients, rows of kBlockSize
    // Mixed precision (layouts in xq_mixed.hip): chunk tables, Fock tile list, digit planes and exponents of Phi in two
    // orders, fixed-point gradients, per-step W digits, P digits and W coefficients.
    int *d_chunkOffset = nullptr, *d_chunkBlock = nullptr;
    int4 *d_tiles = nullptr;
    signed char *d_DA = nullptr, *d_EA = nullptr, *d_DB = nullptr, *d_EB = nullptr;
    int *d_GF = nullptr;
    signed char *d_EG = nullptr;
    signed char *d_DW = nullptr, *d_EW = nullptr;
    unsigned long long *d_Pq = nullptr;
    int *d_eP = nullptr;
    double *d_wcoef = nullptr;
    double *d_rhoOut = nullptr;        // set only while an unrestricted quadrature runs its density stage
    long long nChunks = 0, nTiles = 0;
    // Mixed path, Fock side (xq_mixed.hip): superblock table, union indices, union chunk tables, tiles, grad Phi in
    // the union layout; the union row map is kept only during the set-up.
    int4 *d_supers = nullptr, *d_uTiles = nullptr;
    int *d_uIdx = nullptr, *d_uChunkOffset = nullptr, *d_uChunkBlock = nullptr, *d_uRow = nullptr;
    short *d_GHU = nullptr;
    unsigned char *d_GLU = nullptr;
    long long nSupers = 0, nChunksU = 0, nTilesU = 0, nUnionRows = 0;
    // FP64 path (layouts in xq_f64.hip): Phi and grad Phi planes in operand order, W per step.
    double *d_FPB = nullptr, *d_GPB = nullptr, *d_FPA = nullptr, *d_WPA = nullptr;
    double bfValues = 0.0;                // sum over blocks of kept functions x points
    double bfSquares = 0.0;               // sum over blocks of kept functions^2 x points
    double bfSquaresPadded = 0.0;         // sum over blocks of (kept functions rounded up to 32)^2 x 128
    double bfUnion = 0.0;                 // mixed path: sum over blocks of union functions of the superblock x points

    // Work arrays for a quadrature.
    double *d_P = nullptr, *d_V = nullptr, *d_packed = nullptr;   // dense n x n P and V, packed staging
    double *d_blockEN = nullptr;          // per block: E_xc and N_e partial sums
    int *d_skip = nullptr;                // per block: 1 if skipped by the density tolerance
    double *h_stage = nullptr;            // pinned: packed P in, packed V out
    double *h_blockEN = nullptr;          // pinned: per block E_xc and N_e partial sums
    int *h_skip = nullptr;                // pinned: per block skip flags
    xq::MixedJob *h_job = nullptr, *d_job = nullptr;   // one-entry job table of single calls (pinned host, device)
    long long deviceBytes = 0;
};

// A batch context: its own stream and events, and a job table of up to maxJobs entries (pinned host, device).
struct xq_batch {
    int maxJobs = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evRho0 = nullptr, evRho1 = nullptr, evFock1 = nullptr, evDens1 = nullptr;
    xq::MixedJob *h_jobs = nullptr, *d_jobs = nullptr;
};

namespace xq {

// Process-wide caching allocator (xq_alloc.hip): blocks freed by the library are kept for reuse instead of being
// returned with hipFree, which synchronises the device.  The caller must have completed all work on a block before it
// frees it; cache_free returns the bytes that were requested for the block (0 for a pointer it does not know).
// cache_idle_bytes: bytes held idle; cache_release: hipFree every idle block.
hipError_t cache_alloc(void **ptr, size_t bytes);
size_t cache_free(void *ptr);
hipError_t cache_host_alloc(void **ptr, size_t bytes);   // pinned host memory, cached the same way
void cache_host_free(void *ptr);
size_t cache_idle_bytes();
void cache_release();
void cache_fail_after(long long count);   // test hook: after `count` more device allocations, one fails (-1 disarms)
// Before a kernel whose private memory the runtime allocates at dispatch (the SCF integral kernels): at least `bytes` of
// free device memory, releasing the cache's idle blocks if needed; kOutOfMemory (with `what` in the message) if not.
int ensure_free_device_memory(size_t bytes, const char *what);

// Allocation helpers that account device bytes on the handle.  Every stage synchronises the handle's stream before it
// frees what it allocated.
template <typename T>
int dev_alloc(xq_grid *g, T **ptr, size_t count) {
    *ptr = nullptr;
    if (count == 0) count = 1;
    const size_t bytes = count * sizeof(T);
    hipError_t e = cache_alloc(reinterpret_cast<void **>(ptr), bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        *ptr = nullptr;
        return kOutOfMemory;
    }
    g->deviceBytes += static_cast<long long>(bytes);
    return kOK;
}

// Returns the array to the cache and takes its bytes off the handle's count.
template <typename T>
void dev_free(xq_grid *g, T **ptr) {
    if (*ptr != nullptr) {
        g->deviceBytes -= static_cast<long long>(cache_free(reinterpret_cast<void *>(*ptr)));
        *ptr = nullptr;
    }
}

int check(hipError_t e, const char *what);

// Stages.
int build_grid(xq_grid *g);
int set_basis(xq_grid *g, int nSpec, const double *spec, int spherical, int precision);
int integrate(xq_grid *g, int nIDs, const int *ids, const double *P, double *eXC, double *nE, double *V, double *stats);
// Quadratures of several handles on one stream; the mixed-precision stages run as one launch each over all the jobs.
struct StageEvents {
    hipEvent_t ev0, ev1, evRho0, evRho1, evFock1;
    hipEvent_t evDens1;   // after the density product kernel (before the mixed path's functional kernel)
};
int integrate_jobs(hipStream_t s, const StageEvents &ev, int nJobs, xq_grid *const *grids, const FunctionalSet *fs,
                   const double *const *P, double *eXC, double *nE, double *const *V, double *stats, MixedJob *h_table,
                   MixedJob *d_table);
int functional_set(int nIDs, const int *ids, FunctionalSet *fs);
// Algorithmic HBM bytes of the density and Fock stages of one quadrature on the handle (stats[6] and [7]).
void stage_bytes(const xq_grid *g, double *rho, double *fock, double *product);
// stages: kStageDensity (P digits and the density kernel), kStageFock (W and the Fock product), or both.
constexpr int kStageDensity = 1, kStageFock = 2;
// Per-step arrays of one job in place of its handle's (a null field keeps the handle's): the second spin of an
// unrestricted quadrature runs as a second job on the same handle with its own P, V, W and coefficients.
struct StepArrays {
    double *P = nullptr, *V = nullptr, *wcoef = nullptr, *rhoOut = nullptr, *WPA = nullptr, *part = nullptr;
    unsigned long long *Pq = nullptr;
    int *eP = nullptr;
    signed char *DW = nullptr, *EW = nullptr;
};
int launch_mixed(hipStream_t s, const StageEvents &ev, int nJobs, xq_grid *const *grids, const FunctionalSet *fs,
                 MixedJob *h_table, MixedJob *d_table, int stages = kStageDensity | kStageFock,
                 const StepArrays *over = nullptr);
int launch_f64(hipStream_t s, const StageEvents &ev, int nJobs, xq_grid *const *grids, const FunctionalSet *fs,
               MixedJob *h_table, MixedJob *d_table, int stages = kStageDensity | kStageFock,
               const StepArrays *over = nullptr);
int integrate_uks(xq_grid *g, int nIDs, const int *ids, const double *Pa, const double *Pb, double *eXC, double *nE,
                  double *Va, double *Vb, double *stats);
void launch_expand_P(xq_grid *g, hipStream_t s);
void launch_zero_V(xq_grid *g, hipStream_t s);
void launch_pack_V(xq_grid *g, hipStream_t s);
// The same on explicit arrays: packed lower triangle -> dense n x n, zero n x n, dense -> packed.
void launch_expand(int n, const double *packed, double *dense, hipStream_t s);
void launch_zero(int n, double *dense, hipStream_t s);
void launch_pack(int n, const double *dense, double *packed, hipStream_t s);
// Shared by both precisions: the job table entries of the given handles, the per-block chunk tables and Fock tile list.
// blockSlots (optional): the density launch's work-groups per XCD, the largest over the XCDs of the jobs' summed block
// ranges; itemSlots (optional): the same for the mixed density launch's row-chunk items.
void fill_jobs(int nJobs, xq_grid *const *grids, const FunctionalSet *fs, MixedJob *h_table, long long *rows,
               long long *blocks, long long *chunks, long long *groups, const StepArrays *over = nullptr,
               long long *blockSlots = nullptr, long long *itemSlots = nullptr);
// The density launch's block order (d_order, xcdStart) and the mixed launch's row-chunk items (d_items, itemStart),
// from the blocks' kept-function counts; at basis set-up.
int build_block_order(xq_grid *g, hipStream_t s);
int chunk_tables(xq_grid *g);
int prepare_f64(xq_grid *g, const double *f, const double *fx, const double *fy, const double *fz);
void free_f64(xq_grid *g);
void free_basis(xq_grid *g);
void free_work(xq_grid *g);
int prepare_mixed(xq_grid *g, const double *f, const double *fx, const double *fy, const double *fz);
// Ten FP64 planes of basis data (phi, its three first and six second derivatives), rows of kBlockSize points.
struct Planes10 {
    double *p[10];
};
int eval_basis2(xq_grid *g, int b0, int b1, long long rowBase, const Planes10 &out, hipStream_t s);
int gradients(xq_grid *g, int nIDs, const int *ids, const double *Ppacked, double *grad);
int gradients_uks(xq_grid *g, int nIDs, const int *ids, const double *Pa, const double *Pb, double *grad);
int weight_gradients(xq_grid *g, const double *d_eXC, double *d_grad, hipStream_t s);
// The per-call work arrays (dense P and V, packed staging, block sums and flags) and the one-entry job table of single
// calls: each is allocated complete or not at all.
int ensure_work(xq_grid *g);
int ensure_job_table(xq_grid *g);
int upload_density(xq_grid *g, const double *Ppacked, hipStream_t s);
void free_mixed(xq_grid *g);
int alloc_step_mixed(xq_grid *g, StepArrays *a);
void free_step_mixed(xq_grid *g, StepArrays *a);
int alloc_step_f64(xq_grid *g, StepArrays *a);
void free_step_f64(xq_grid *g, StepArrays *a);
// xq_integrate_noalloc (xq_noalloc.hip): the scratch array's size and what one call on the handle needs of it.
int integrate_noalloc(xq_grid *g, int nIDs, const int *ids, const double *P, double *eXC, double *nE, double *V,
                      double *stats);
size_t noalloc_scratch_bytes();
size_t noalloc_needed_bytes(const xq_grid *g);

}  // namespace xq

#define XQ_CHECK(call)                                   \
    do {                                                 \
        int _s = xq::check((call), #call);               \
        if (_s != xq::kOK) return _s;                    \
    } while (0)

#define XQ_TRY(call)                                     \
    do {                                                 \
        int _s = (call);                                 \
        if (_s != xq::kOK) return _s;                    \
    } while (0)
