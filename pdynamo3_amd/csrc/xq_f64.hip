// FP64 quadrature on FP64 matrix cores (v_mfma_f64_16x16x4f64), same mathematics as the reference
// (DFTIntegrator.c:159-509, 517-618, 814-901):
//   X = P_red Phi, rho = sum Phi X, grad rho = 2 sum grad Phi X, functional, W = (w vrho / 2) Phi + 2 w vsigma
//   (grad rho . grad Phi), V += Phi W^T + W Phi^T (lower triangle).
// The basis values are stored once per geometry in the lane order the products consume, so every operand load is one
// coalesced 512-byte row per wave.  The kernels follow the mixed path's structure (xq_mixed.hip): a per-block density
// kernel with P_red staged in LDS by row panels, a per-chunk W kernel, and a Fock kernel over a flat list of
// lower-triangle tiles with FP64 atomics; all read the same job table, so one launch serves one job or a batch.
//
// MFMA lane layout (16 x 16 x 4, FP64): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and receives
// D[4 v + l / 16][l % 16] for v = 0..3.
//
// Layouts (4096 doubles per chunk of 32 kept functions x 128 points; chunk C is chunk rc of block b):
//   FPB, GPB [C][g (4)][s (8)][h (2)][lane (64)]   Phi and grad Phi (GPB: 3 planes) in the density's B operand order:
//                                                  lane l holds row 4 s + l / 16 at point 32 g + 16 h + l % 16.  The
//                                                  density result register v of row half hm and point half hn is row
//                                                  16 hm + 4 v + l / 16, which is the operand entry (s = 4 hm + v, hn).
//   FPA, WPA [C][s (32)][h (2)][lane (64)]         Phi and W (per step) in the Fock product's A/B order: lane l holds
//                                                  row 16 h + l % 16 at point 4 s + l / 16.
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "xq_functional.h"
#include "xq_internal.h"

namespace xq {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int kC = 32;                   // chunk rows
constexpr int kChunk = 4096;             // doubles per chunk plane
constexpr int kPanelF = 4;               // chunks of P_red columns staged at once
constexpr int kPanelPitch = kPanelF * kC + 2;   // LDS row pitch in doubles (padded against bank conflicts)
constexpr int kIdxCache = 1024;          // kept-function indices of a block cached in LDS by the density kernel

__device__ __forceinline__ v4d mfma_f64(double a, double b, const v4d &c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename F>
__device__ __forceinline__ int job_of(const MixedJob *jobs, int nJobs, int g, F base) {
    int lo = 0, hi = nJobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base(jobs[mid]) <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Set-up, once per geometry: one chunk per work-group (256 threads) from the FP64 rows of f, fx, fy, fz.
__global__ void __launch_bounds__(256) f64_prepare_kernel(
    const DevBlock *__restrict__ blocks, const int *__restrict__ chunkOffset, const int *__restrict__ chunkBlock,
    const double *__restrict__ f, const double *__restrict__ fx, const double *__restrict__ fy,
    const double *__restrict__ fz, double *__restrict__ FPB, double *__restrict__ GPB, double *__restrict__ FPA) {
    const size_t C = blockIdx.x;
    const int b = chunkBlock[C];
    const DevBlock blk = blocks[b];
    const int nb = blk.nKept, rc = static_cast<int>(C) - chunkOffset[b];
    const size_t R = static_cast<size_t>(blk.vOffset);
    auto value = [&](const double *a, int r, int p) -> double {
        const int row = kC * rc + r;
        return row < nb ? a[(R + row) * kBlockSize + p] : 0.0;
    };
    for (int e = threadIdx.x; e < kChunk; e += 256) {
        const int lane = e & 63, h = (e >> 6) & 1, s = (e >> 7) & 7, g = e >> 10;
        const int r = 4 * s + (lane >> 4), p = 32 * g + 16 * h + (lane & 15);
        FPB[C * kChunk + e] = value(f, r, p);
        GPB[(3 * C) * kChunk + e] = value(fx, r, p);
        GPB[(3 * C + 1) * kChunk + e] = value(fy, r, p);
        GPB[(3 * C + 2) * kChunk + e] = value(fz, r, p);
        const int lane2 = e & 63, h2 = (e >> 6) & 1, s2 = e >> 7;
        FPA[C * kChunk + e] = value(f, 16 * h2 + (lane2 & 15), 4 * s2 + (lane2 >> 4));
    }
}

// Density kernel (per block, 256 threads = 4 waves; wave wv owns points 32 wv .. + 31): X = P_red Phi on FP64 MFMA in
// 32 x 32 tiles (2 x 2 MFMA tiles per wave), rho and grad rho, the functional, E_xc / N_e block sums and the W
// coefficients per point.  P_red is staged in LDS by row panels of kPanelF chunks; the Phi operands of the next chunk
// are loaded while the current one is multiplied.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
density_kernel_f64(const MixedJob *__restrict__ jobs, int nJobs) {
    __shared__ double sPf[kC][kPanelPitch];
    __shared__ int sIdx[kIdxCache];
    __shared__ double rsum[4][kBlockSize];
    __shared__ double wsum[4][2];
    // Work-group g runs on XCD x = g % 8 and takes slot g / 8 of that XCD's list: the jobs' block ranges x in job order
    // (each a contiguous region of its molecule, so that the XCD's L2 holds the P rows its blocks gather; the ranges
    // of one job have equal estimated cost), each range's blocks in order of decreasing cost.
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const MixedJob &J = jobs[job_of(jobs, nJobs, slot, [xcd](const MixedJob &j) { return j.xcdBase[xcd]; })];
    const int local = slot - J.xcdBase[xcd];
    if (local >= J.xcdStart[xcd + 1] - J.xcdStart[xcd]) return;
    const int b = J.order[J.xcdStart[xcd] + local].x;
    const DevBlock blk = J.blocks[b];
    const int nb = blk.nKept, np = blk.nPoints, n = J.n;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, l16 = lane & 15, l4 = lane >> 4;
    if (nb <= 0) {
        if (t == 0) { J.blockEN[2 * b] = 0.0; J.blockEN[2 * b + 1] = 0.0; J.skip[b] = 1; }
        if (J.rhoOut != nullptr && t < np)
            for (int c = 0; c < 4; c++) J.rhoOut[4 * (static_cast<size_t>(blk.pStart) + t) + c] = 0.0;
        return;
    }
    // The block's kept-function indices, cached in LDS for the P_red gathers (blocks with more kept functions than the
    // cache read them from memory).
    const int *idxG = J.idx + blk.iOffset;
    for (int k = t; k < nb && k < kIdxCache; k += 256) sIdx[k] = idxG[k];
    __syncthreads();
    const int *idx = nb <= kIdxCache ? sIdx : idxG;
    const size_t C = static_cast<size_t>(J.chunkOffset[b]);
    const int nChunks = (nb + kC - 1) / kC;
    const double *__restrict__ P = J.P;
    const double *__restrict__ FPB = J.FPB;
    const double *__restrict__ GPB = J.GPB;
    // Operand entry (s, h) of this lane within a chunk plane of FPB / GPB.
    auto op = [&](int s, int h) { return ((wv * 8 + s) * 2 + h) * 64 + lane; };
    auto fetchB = [&](int c, double (&bq)[8][2]) {
#pragma unroll
        for (int s = 0; s < 8; s++)
#pragma unroll
            for (int h = 0; h < 2; h++) bq[s][h] = FPB[(C + c) * kChunk + op(s, h)];
    };
    const int sr = t >> 3, s8 = t & 7;     // staging: row sr, columns 4 s8 .. + 3 of each chunk of the panel
    double bnext[8][2];
    fetchB(0, bnext);
    double pr[2] = {0.0, 0.0}, pgx[2] = {0.0, 0.0}, pgy[2] = {0.0, 0.0}, pgz[2] = {0.0, 0.0};
    for (int I = 0; I < nChunks; I++) {
        v4d acc[2][2];
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
            if (c0 == 0) {
#pragma unroll
                for (int hm = 0; hm < 2; hm++)
#pragma unroll
                    for (int hn = 0; hn < 2; hn++) acc[hm][hn] = v4d{0.0, 0.0, 0.0, 0.0};
            }
            __syncthreads();
            for (int cl = 0; cl < nc; cl++) {
                const int c = c0 + cl;
                double bq[8][2];
#pragma unroll
                for (int s = 0; s < 8; s++) { bq[s][0] = bnext[s][0]; bq[s][1] = bnext[s][1]; }
                if (c + 1 < nChunks) fetchB(c + 1, bnext);
                else if (I + 1 < nChunks) fetchB(0, bnext);
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
        // Epilogue: register v of (hm, hn) is row 16 hm + 4 v + l / 16 of chunk I at point 32 wv + 16 hn + l % 16.
#pragma unroll
        for (int hm = 0; hm < 2; hm++)
#pragma unroll
            for (int v = 0; v < 4; v++)
#pragma unroll
                for (int hn = 0; hn < 2; hn++) {
                    const int o = op(4 * hm + v, hn);
                    const double x = acc[hm][hn][v];
                    pr[hn] += FPB[(C + I) * kChunk + o] * x;
                    pgx[hn] += GPB[3 * (C + I) * kChunk + o] * x;
                    pgy[hn] += GPB[(3 * (C + I) + 1) * kChunk + o] * x;
                    pgz[hn] += GPB[(3 * (C + I) + 2) * kChunk + o] * x;
                }
    }
    // Sum over the four lanes that share a point (l / 16 = 0..3).
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
    double rho = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    if (t < np) { rho = rsum[0][t]; gx = 2.0 * rsum[1][t]; gy = 2.0 * rsum[2][t]; gz = 2.0 * rsum[3][t]; }
    if (J.rhoOut != nullptr) {   // one spin of an unrestricted quadrature: the functional follows in its own kernel
        if (t < np) {
            double *o = J.rhoOut + 4 * (static_cast<size_t>(blk.pStart) + t);
            o[0] = rho; o[1] = gx; o[2] = gy; o[3] = gz;
        }
        return;
    }
    double mx = t < np ? fabs(rho) : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) wsum[wv][0] = mx;
    __syncthreads();
    const bool skipped = !(fmax(fmax(wsum[0][0], wsum[1][0]), fmax(wsum[2][0], wsum[3][0])) > J.rhoTol);
    __syncthreads();
    double e = 0.0, ne = 0.0;
    if (t < np) {
        double vr = 0.0, vs = 0.0;
        if (!skipped) {
            double eps;
            evaluate_functional(J.fs, rho, gx * gx + gy * gy + gz * gz, eps, vr, vs);
            const double ww = J.w[blk.pStart + t];
            vr *= ww;
            vs *= ww;
            e = eps * rho * ww;
            ne = rho * ww;
        }
        double *cw = J.wcoef + 4 * (static_cast<size_t>(blk.pStart) + t);
        cw[0] = 0.5 * vr;
        cw[1] = 2.0 * vs * gx;
        cw[2] = 2.0 * vs * gy;
        cw[3] = 2.0 * vs * gz;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { e += __shfl_xor(e, o, 64); ne += __shfl_xor(ne, o, 64); }
    if (lane == 0) { wsum[wv][0] = e; wsum[wv][1] = ne; }
    __syncthreads();
    if (t == 0) {
        J.blockEN[2 * b] = ((wsum[0][0] + wsum[1][0]) + wsum[2][0]) + wsum[3][0];
        J.blockEN[2 * b + 1] = ((wsum[0][1] + wsum[1][1]) + wsum[2][1]) + wsum[3][1];
        J.skip[b] = skipped ? 1 : 0;
    }
}

// W (one chunk per work-group, 256 threads): W_rp = c0_p Phi_rp + c_p . grad Phi_rp from the density-order planes,
// transposed through LDS into the Fock order (WPA).
__global__ void __launch_bounds__(256) w_kernel_f64(const MixedJob *__restrict__ jobs, int nJobs) {
    __shared__ double sc[4][kBlockSize];
    __shared__ double sW[kC][kBlockSize + 1];
    const MixedJob &J = jobs[job_of(jobs, nJobs, blockIdx.x, [](const MixedJob &j) { return j.chunkBase; })];
    const size_t C = blockIdx.x - J.chunkBase;
    const int b = J.chunkBlock[C];
    if (J.skip[b]) return;
    const DevBlock blk = J.blocks[b];
    const int t = threadIdx.x;
    if (t < kBlockSize) {
#pragma unroll
        for (int k = 0; k < 4; k++) sc[k][t] = t < blk.nPoints ? J.wcoef[4 * (static_cast<size_t>(blk.pStart) + t) + k] : 0.0;
    }
    __syncthreads();
    for (int e = t; e < kChunk; e += 256) {
        const int lane = e & 63, h = (e >> 6) & 1, s = (e >> 7) & 7, g = e >> 10;
        const int r = 4 * s + (lane >> 4), p = 32 * g + 16 * h + (lane & 15);
        sW[r][p] = sc[0][p] * J.FPB[C * kChunk + e] + sc[1][p] * J.GPB[3 * C * kChunk + e] +
                   sc[2][p] * J.GPB[(3 * C + 1) * kChunk + e] + sc[3][p] * J.GPB[(3 * C + 2) * kChunk + e];
    }
    __syncthreads();
    for (int e = t; e < kChunk; e += 256) {
        const int lane = e & 63, h = (e >> 6) & 1, s = e >> 7;
        J.WPA[C * kChunk + e] = sW[16 * h + (lane & 15)][4 * s + (lane >> 4)];
    }
}

// Fock kernel: one wave per lower-triangle tile (I, J <= I), four tiles per work-group, no LDS and no barriers; K runs
// over the block's 128 points in steps of 4.  V += Phi_I W_J^T + W_I Phi_J^T with FP64 atomics.  Work-group g runs on XCD
// g % 8; tiles are dealt so that each XCD takes a contiguous range of the tile list.
__global__ void __launch_bounds__(256) fock_kernel_f64(const MixedJob *__restrict__ jobs, int nJobs, int groupsPerXcd,
                                                       int nGroups) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, l16 = lane & 15, l4 = lane >> 4;
    const int group = (blockIdx.x & 7) * groupsPerXcd + (blockIdx.x >> 3);
    if (group >= nGroups) return;
    const MixedJob &job = jobs[job_of(jobs, nJobs, group, [](const MixedJob &j) { return j.groupBase; })];
    const int tileIndex = 4 * (group - job.groupBase) + wv;
    if (tileIndex >= job.nTiles) return;
    const int4 tl = job.tiles[tileIndex];
    const int b = tl.x, I = tl.y, J = tl.z;
    if (job.skip[b]) return;
    const DevBlock blk = job.blocks[b];
    const int nb = blk.nKept;
    const int *idx = job.idx + blk.iOffset;
    const size_t CI = static_cast<size_t>(job.chunkOffset[b] + I), CJ = static_cast<size_t>(job.chunkOffset[b] + J);
    const double *fI = job.FPA + CI * kChunk + lane, *wI = job.WPA + CI * kChunk + lane;
    const double *fJ = job.FPA + CJ * kChunk + lane, *wJ = job.WPA + CJ * kChunk + lane;
    v4d acc[2][2];
#pragma unroll
    for (int hm = 0; hm < 2; hm++)
#pragma unroll
        for (int hn = 0; hn < 2; hn++) acc[hm][hn] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int s = 0; s < 32; s++) {
        double xf[2], xw[2], yf[2], yw[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int o = (s * 2 + h) * 64;
            xf[h] = fI[o];
            xw[h] = wI[o];
            yf[h] = fJ[o];
            yw[h] = wJ[o];
        }
#pragma unroll
        for (int hm = 0; hm < 2; hm++)
#pragma unroll
            for (int hn = 0; hn < 2; hn++) {
                acc[hm][hn] = mfma_f64(xf[hm], yw[hn], acc[hm][hn]);   // Phi_I W_J^T
                acc[hm][hn] = mfma_f64(xw[hm], yf[hn], acc[hm][hn]);   // W_I Phi_J^T
            }
    }
    const int n = job.n;
#pragma unroll
    for (int hn = 0; hn < 2; hn++) {
        const int gj = kC * J + 16 * hn + l16;
        if (gj >= nb) continue;
        const size_t col = static_cast<size_t>(idx[gj]);
#pragma unroll
        for (int hm = 0; hm < 2; hm++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int gi = kC * I + 16 * hm + 4 * v + l4;
                if (gi < nb && gj <= gi) atomicAdd(&job.V[static_cast<size_t>(idx[gi]) * n + col], acc[hm][hn][v]);
            }
    }
}

}  // namespace

int prepare_f64(xq_grid *g, const double *f, const double *fx, const double *fy, const double *fz) {
    XQ_TRY(chunk_tables(g));
    const size_t nc = static_cast<size_t>(g->nChunks > 0 ? g->nChunks : 1);
    XQ_TRY(dev_alloc(g, &g->d_FPB, nc * kChunk));
    XQ_TRY(dev_alloc(g, &g->d_GPB, 3 * nc * kChunk));
    XQ_TRY(dev_alloc(g, &g->d_FPA, nc * kChunk));
    XQ_TRY(dev_alloc(g, &g->d_WPA, nc * kChunk));
    XQ_TRY(dev_alloc(g, &g->d_wcoef, 4 * static_cast<size_t>(g->nPoints > 0 ? g->nPoints : 1)));
    hipStream_t s = g->stream;
    if (g->nChunks > 0)
        hipLaunchKernelGGL(f64_prepare_kernel, dim3(static_cast<unsigned>(g->nChunks)), dim3(256), 0, s, g->d_blocks,
                           g->d_chunkOffset, g->d_chunkBlock, f, fx, fy, fz, g->d_FPB, g->d_GPB, g->d_FPA);
    XQ_CHECK(hipGetLastError());
    XQ_CHECK(hipStreamSynchronize(s));
    return kOK;
}

// Per-step W planes of a second job on the handle (the beta spin of an unrestricted quadrature).
int alloc_step_f64(xq_grid *g, StepArrays *a) {
    return dev_alloc(g, &a->WPA, static_cast<size_t>(g->nChunks > 0 ? g->nChunks : 1) * kChunk);
}

void free_step_f64(xq_grid *g, StepArrays *a) { dev_free(g, &a->WPA); }

void free_f64(xq_grid *g) {
    dev_free(g, &g->d_FPB);
    dev_free(g, &g->d_GPB);
    dev_free(g, &g->d_FPA);
    dev_free(g, &g->d_WPA);
}

int launch_f64(hipStream_t s, const StageEvents &ev, int nJobs, xq_grid *const *grids, const FunctionalSet *fs,
               MixedJob *h_table, MixedJob *d_table, int stages, const StepArrays *over) {
    long long rows, blocks, chunks, groups, blockSlots;
    fill_jobs(nJobs, grids, fs, h_table, &rows, &blocks, &chunks, &groups, over, &blockSlots);
    if (blocks > INT32_MAX / 2 || chunks > INT32_MAX / 2 || groups > INT32_MAX / 16) {
        set_error("batch too large for one launch");
        return kInvalidArgument;
    }
    XQ_CHECK(hipMemcpyAsync(d_table, h_table, nJobs * sizeof(MixedJob), hipMemcpyHostToDevice, s));
    XQ_CHECK(hipEventRecord(ev.evRho0, s));
    if ((stages & kStageDensity) && blocks > 0) {
        hipLaunchKernelGGL(density_kernel_f64, dim3(static_cast<unsigned>(8 * blockSlots)), dim3(256), 0, s, d_table,
                           nJobs);
    }
    XQ_CHECK(hipEventRecord(ev.evDens1, s));
    XQ_CHECK(hipEventRecord(ev.evRho1, s));
    if ((stages & kStageFock) && chunks > 0) {
        hipLaunchKernelGGL(w_kernel_f64, dim3(static_cast<unsigned>(chunks)), dim3(256), 0, s, d_table, nJobs);
        const int perXcd = static_cast<int>((groups + 7) / 8);
        hipLaunchKernelGGL(fock_kernel_f64, dim3(static_cast<unsigned>(8 * perXcd)), dim3(256), 0, s, d_table, nJobs,
                           perXcd, static_cast<int>(groups));
    }
    XQ_CHECK(hipEventRecord(ev.evFock1, s));
    XQ_CHECK(hipGetLastError());
    return kOK;
}

}  // namespace xq
