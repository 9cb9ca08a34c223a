// One quadrature on the device: density and gradient, GGA functional, E_xc / N_e and V_xc, for every grid block.
//
// Follows DFTIntegrator_Integrate (pMolecule/QCModel/__extensions__/csource/DFTIntegrator.c:159-509), restricted closed
// shell with a Fock matrix and no gradients:
//   X = P_red Phi (:838), rho_p = sum_m Phi_mp X_mp (:839), grad rho_p = 2 sum_m grad Phi_mp X_mp, sigma = |grad rho|^2
//   (:843-850); a block is skipped when max_p |rho_p| <= rhoTol (:332); vrho and vsigma are scaled by the weights (:336-338);
//   E += sum_p w eps rho, N += sum_p w rho (:350-352); V_mn += sum_p Phi_mp W_np + W_mp Phi_np with
//   W = vrho Phi / 2 + 2 vsigma (grad rho . grad Phi), which is the sum of DFTIntegrator_Fock's three terms (:517-618).
//
// Kernels, each one work-group per grid block (blocks hold at most 128 points of one atom and their kept functions):
//   rho_functional_kernel : X in 64-row tiles, K chunks of 32 through LDS; the epilogue contracts X with Phi and grad Phi,
//                           then evaluates the functional in FP64 for the block's points and writes w vrho, w vsigma,
//                           grad rho, and the block's E_xc and N_e partial sums.
//   fock_kernel           : 64 x 64 tiles of the block's reduced V (lower triangle only), K = the block's points, added
//                           into the dense device V with FP64 atomics.
// Storage type T (double or float) is that of the basis values; A is the accumulation type of both contractions.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "xq_functional.h"
#include "xq_internal.h"

namespace xq {

namespace {

constexpr int kTile = 64;   // rows of X per pass, and the V tile edge
constexpr int kKC = 32;     // K chunk

__global__ void expand_packed_kernel(int n, const double *__restrict__ packed, double *__restrict__ dense,
                                     float *__restrict__ dense32) {
    const size_t total = static_cast<size_t>(n) * n;
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
        const int hi = i > j ? i : j, lo = i > j ? j : i;
        const double v = packed[static_cast<size_t>(hi) * (hi + 1) / 2 + lo];
        dense[e] = v;
        if (dense32) dense32[e] = static_cast<float>(v);
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

// 256 threads: ty = t / 16 owns X rows ty + 16 i (i < 4), tx = t % 16 owns points tx + 16 j (j < 8).
template <typename T, typename A>
__global__ void __launch_bounds__(256) rho_functional_kernel(
    const xq::DevBlock *__restrict__ blocks, const int *__restrict__ idxAll, const T *__restrict__ f,
    const T *__restrict__ fx, const T *__restrict__ fy, const T *__restrict__ fz, int n, const double *__restrict__ P,
    const float *__restrict__ P32, const double *__restrict__ w, FunctionalSet fs, double rhoTol,
    double *__restrict__ rhoOut, double *__restrict__ vrs, double *__restrict__ blockEN, int *__restrict__ skip) {
    __shared__ A Ps[kKC][kTile + 1];          // P_red tile, K-major
    __shared__ A Fs[kKC][kBlockSize];         // Phi rows of the K chunk
    __shared__ double red[16][kBlockSize];    // reduction over the 16 row groups
    __shared__ double res[4][kBlockSize];
    __shared__ double wred[4][2];
    const int b = blockIdx.x;
    const xq::DevBlock blk = blocks[b];
    const int nb = blk.nKept, np = blk.nPoints;
    const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
    if (nb <= 0) {
        if (t == 0) { blockEN[2 * b] = 0.0; blockEN[2 * b + 1] = 0.0; skip[b] = 1; }
        return;
    }
    const int *idx = idxAll + blk.iOffset;
    const size_t row0 = static_cast<size_t>(blk.vOffset);
    double pr[8], pgx[8], pgy[8], pgz[8];
#pragma unroll
    for (int j = 0; j < 8; j++) { pr[j] = 0.0; pgx[j] = 0.0; pgy[j] = 0.0; pgz[j] = 0.0; }
    for (int I = 0; I < nb; I += kTile) {
        A acc[4][8];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 8; j++) acc[i][j] = A(0);
        for (int K = 0; K < nb; K += kKC) {
            // P_red tile: rows I.., columns K..; 64 x 32 entries, 8 per thread.
            for (int e = t; e < kTile * kKC; e += 256) {
                const int r = e / kKC, c = e % kKC;
                A v = A(0);
                if (I + r < nb && K + c < nb)
                    v = static_cast<A>(P[static_cast<size_t>(idx[I + r]) * n + idx[K + c]]);
                Ps[c][r] = v;
            }
            // Phi rows K.. (32 x 128).
            for (int e = t; e < kKC * kBlockSize; e += 256) {
                const int r = e / kBlockSize, c = e % kBlockSize;
                Fs[r][c] = K + r < nb ? static_cast<A>(f[(row0 + K + r) * kBlockSize + c]) : A(0);
            }
            __syncthreads();
#pragma unroll 4
            for (int k = 0; k < kKC; k++) {
                A pv[4], fv[8];
#pragma unroll
                for (int i = 0; i < 4; i++) pv[i] = Ps[k][ty + 16 * i];
#pragma unroll
                for (int j = 0; j < 8; j++) fv[j] = Fs[k][tx + 16 * j];
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 8; j++) acc[i][j] += pv[i] * fv[j];
            }
            __syncthreads();
        }
        // Contract this X tile with Phi and grad Phi.
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = I + ty + 16 * i;
            if (r < nb) {
                const size_t o = (row0 + r) * kBlockSize;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int p = tx + 16 * j;
                    const double xv = static_cast<double>(acc[i][j]);
                    pr[j] += static_cast<double>(f[o + p]) * xv;
                    pgx[j] += static_cast<double>(fx[o + p]) * xv;
                    pgy[j] += static_cast<double>(fy[o + p]) * xv;
                    pgz[j] += static_cast<double>(fz[o + p]) * xv;
                }
            }
        }
    }
    // Sum the 16 row groups: rho and the three gradient components.
#pragma unroll
    for (int q = 0; q < 4; q++) {
#pragma unroll
        for (int j = 0; j < 8; j++) red[ty][tx + 16 * j] = q == 0 ? pr[j] : q == 1 ? pgx[j] : q == 2 ? pgy[j] : pgz[j];
        __syncthreads();
        if (t < kBlockSize) {
            double s = 0.0;
            for (int g = 0; g < 16; g++) s += red[g][t];
            res[q][t] = s;
        }
        __syncthreads();
    }
    // Functional, one thread per point.
    double e = 0.0, ne = 0.0, amax = 0.0;
    double rho = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, vr = 0.0, vs = 0.0;
    if (t < np) {
        rho = res[0][t];
        gx = 2.0 * res[1][t];
        gy = 2.0 * res[2][t];
        gz = 2.0 * res[3][t];
        amax = fabs(rho);
    }
    // Block maximum of |rho| (the skip test, DFTIntegrator.c:332).
    double m = amax;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((t & 63) == 0) wred[t >> 6][0] = m;
    __syncthreads();
    const double bmax = fmax(fmax(wred[0][0], wred[1][0]), fmax(wred[2][0], wred[3][0]));
    const bool skipped = !(bmax > rhoTol);
    __syncthreads();
    if (!skipped && t < np) {
        const double sigma = gx * gx + gy * gy + gz * gz;
        double eps;
        evaluate_functional(fs, rho, sigma, eps, vr, vs);
        const double ww = w[blk.pStart + t];
        vr *= ww;
        vs *= ww;
        e = eps * rho * ww;
        ne = rho * ww;
    }
    if (t < kBlockSize) {
        const size_t p = static_cast<size_t>(blk.pStart) + t;
        if (t < np) {
            rhoOut[4 * p] = rho;
            rhoOut[4 * p + 1] = gx;
            rhoOut[4 * p + 2] = gy;
            rhoOut[4 * p + 3] = gz;
            vrs[2 * p] = skipped ? 0.0 : vr;
            vrs[2 * p + 1] = skipped ? 0.0 : vs;
        }
    }
    // Block sums of E and N (fixed order: waves, then lanes by shuffles).
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { e += __shfl_xor(e, o, 64); ne += __shfl_xor(ne, o, 64); }
    if ((t & 63) == 0) { wred[t >> 6][0] = e; wred[t >> 6][1] = ne; }
    __syncthreads();
    if (t == 0) {
        blockEN[2 * b] = ((wred[0][0] + wred[1][0]) + wred[2][0]) + wred[3][0];
        blockEN[2 * b + 1] = ((wred[0][1] + wred[1][1]) + wred[2][1]) + wred[3][1];
        skip[b] = skipped ? 1 : 0;
    }
}

// W row r at point p: vrho/2 Phi + 2 vsigma (grad rho . grad Phi), with w already folded into vrho and vsigma.
template <typename T, typename A>
__device__ __forceinline__ A w_entry(const T *f, const T *fx, const T *fy, const T *fz, size_t o, A hv, A sv, A gx, A gy,
                                     A gz) {
    return hv * static_cast<A>(f[o]) +
           sv * (gx * static_cast<A>(fx[o]) + gy * static_cast<A>(fy[o]) + gz * static_cast<A>(fz[o]));
}

// 256 threads: ty = t / 16 owns V rows ty + 16 i, tx = t % 16 owns columns tx + 16 j (i, j < 4) of a 64 x 64 tile.
template <typename T, typename A>
__global__ void __launch_bounds__(256) fock_kernel(const xq::DevBlock *__restrict__ blocks,
                                                   const int *__restrict__ idxAll, const T *__restrict__ f,
                                                   const T *__restrict__ fx, const T *__restrict__ fy,
                                                   const T *__restrict__ fz, int n, const double *__restrict__ rhoIn,
                                                   const double *__restrict__ vrs, const int *__restrict__ skip,
                                                   double *__restrict__ V) {
    __shared__ A FI[kKC][kTile + 1], WI[kKC][kTile + 1], FJ[kKC][kTile + 1], WJ[kKC][kTile + 1];
    __shared__ A pv[4][kBlockSize];   // per point: vrho/2, 2 vsigma, grad rho
    const int b = blockIdx.x;
    if (skip[b]) return;
    const xq::DevBlock blk = blocks[b];
    const int nb = blk.nKept, np = blk.nPoints;
    const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
    const int *idx = idxAll + blk.iOffset;
    const size_t row0 = static_cast<size_t>(blk.vOffset);
    if (t < kBlockSize) {
        const size_t p = static_cast<size_t>(blk.pStart) + t;
        const bool valid = t < np;
        pv[0][t] = valid ? static_cast<A>(0.5 * vrs[2 * p]) : A(0);
        const double s2 = valid ? 2.0 * vrs[2 * p + 1] : 0.0;
        pv[1][t] = static_cast<A>(s2 * (valid ? rhoIn[4 * p + 1] : 0.0));
        pv[2][t] = static_cast<A>(s2 * (valid ? rhoIn[4 * p + 2] : 0.0));
        pv[3][t] = static_cast<A>(s2 * (valid ? rhoIn[4 * p + 3] : 0.0));
    }
    __syncthreads();
    for (int I = 0; I < nb; I += kTile) {
        for (int J = 0; J <= I; J += kTile) {
            A acc[4][4];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = A(0);
            for (int K = 0; K < np; K += kKC) {
                for (int e = t; e < kTile * kKC; e += 256) {
                    const int r = e / kKC, c = e % kKC, p = K + c;
                    A fiv = A(0), wiv = A(0), fjv = A(0), wjv = A(0);
                    if (p < np) {
                        if (I + r < nb) {
                            const size_t o = (row0 + I + r) * kBlockSize + p;
                            fiv = static_cast<A>(f[o]);
                            wiv = pv[0][p] * fiv + (pv[1][p] * static_cast<A>(fx[o]) + pv[2][p] * static_cast<A>(fy[o]) +
                                                    pv[3][p] * static_cast<A>(fz[o]));
                        }
                        if (J + r < nb) {
                            const size_t o = (row0 + J + r) * kBlockSize + p;
                            fjv = static_cast<A>(f[o]);
                            wjv = pv[0][p] * fjv + (pv[1][p] * static_cast<A>(fx[o]) + pv[2][p] * static_cast<A>(fy[o]) +
                                                    pv[3][p] * static_cast<A>(fz[o]));
                        }
                    }
                    FI[c][r] = fiv;
                    WI[c][r] = wiv;
                    FJ[c][r] = fjv;
                    WJ[c][r] = wjv;
                }
                __syncthreads();
#pragma unroll 4
                for (int k = 0; k < kKC; k++) {
                    A a1[4], a2[4], b1[4], b2[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) { a1[i] = FI[k][ty + 16 * i]; a2[i] = WI[k][ty + 16 * i]; }
#pragma unroll
                    for (int j = 0; j < 4; j++) { b1[j] = WJ[k][tx + 16 * j]; b2[j] = FJ[k][tx + 16 * j]; }
#pragma unroll
                    for (int i = 0; i < 4; i++)
#pragma unroll
                        for (int j = 0; j < 4; j++) acc[i][j] += a1[i] * b1[j] + a2[i] * b2[j];
                }
                __syncthreads();
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = I + ty + 16 * i;
                if (r >= nb) continue;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int c = J + tx + 16 * j;
                    if (c >= nb || c > r) continue;
                    atomicAdd(&V[static_cast<size_t>(idx[r]) * n + idx[c]], static_cast<double>(acc[i][j]));
                }
            }
        }
    }
}

template <typename T, typename A>
void launch_quadrature(xq_grid *g, const FunctionalSet &fs) {
    hipStream_t s = g->stream;
    const T *f = static_cast<const T *>(g->d_f), *fx = static_cast<const T *>(g->d_fx),
            *fy = static_cast<const T *>(g->d_fy), *fz = static_cast<const T *>(g->d_fz);
    (void)hipEventRecord(g->evRho0, s);
    hipLaunchKernelGGL((rho_functional_kernel<T, A>), dim3(g->nBlocks), dim3(256), 0, s, g->d_blocks, g->d_idx, f, fx, fy,
                       fz, g->nFunctions, g->d_P, g->d_P32, g->d_w, fs, g->rhoTol, g->d_rho, g->d_vrs, g->d_blockEN,
                       g->d_skip);
    (void)hipEventRecord(g->evRho1, s);
    hipLaunchKernelGGL((fock_kernel<T, A>), dim3(g->nBlocks), dim3(256), 0, s, g->d_blocks, g->d_idx, f, fx, fy, fz,
                       g->nFunctions, g->d_rho, g->d_vrs, g->d_skip, g->d_V);
    (void)hipEventRecord(g->evFock1, s);
}

}  // namespace

void launch_fused_f64(xq_grid *g, const FunctionalSet &fs);
int launch_fused_mixed(xq_grid *g, const FunctionalSet &fs);

void free_work(xq_grid *g) {
    dev_free(&g->d_P);
    dev_free(&g->d_P32);
    dev_free(&g->d_V);
    dev_free(&g->d_packed);
    dev_free(&g->d_rho);
    dev_free(&g->d_vrs);
    dev_free(&g->d_blockEN);
    dev_free(&g->d_skip);
}

static int ensure_work(xq_grid *g) {
    if (g->d_P != nullptr) return kOK;
    const size_t n = static_cast<size_t>(g->nFunctions);
    XQ_TRY(dev_alloc(g, &g->d_P, n * n));
    XQ_TRY(dev_alloc(g, &g->d_V, n * n));
    XQ_TRY(dev_alloc(g, &g->d_packed, n * (n + 1) / 2));
    XQ_TRY(dev_alloc(g, &g->d_rho, 4 * static_cast<size_t>(g->nPoints)));
    XQ_TRY(dev_alloc(g, &g->d_vrs, 2 * static_cast<size_t>(g->nPoints)));
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
    hipLaunchKernelGGL(expand_packed_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, g->nFunctions, g->d_packed, g->d_P,
                       g->d_P32);
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
    }
    return kOK;
}

}  // namespace xq
