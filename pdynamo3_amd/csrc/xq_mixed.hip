// Mixed-precision quadrature: both contractions on int8 matrix cores with exact integer accumulation (Ozaki-style
// slicing), FP64 everywhere else.
//
// Representation.  A value x of a group that shares the power of two 2^e (|x| < 2^e) is rounded once to the integer
// Q = round(x 2^(7D - e)), and |Q| is written in D base-128 digits that carry x's sign (d_k in [-127, 127]):
//     x = 2^e sum_k d_k 2^(-7(k+1)) + r,   |r| <= 2^(e - 7D - 1)   (one round-to-nearest, unbiased).
// A product of two sliced groups is sum_{k,l} 2^(-7(k+l+2)) sum d_k d_l.  The inner sums are int8 GEMMs with int32
// accumulation (v_mfma_i32_32x32x32_i8), exact for the sizes used here, and products of equal order k + l share one
// accumulator.  The accumulators of one result combine exactly into a 64-bit integer, sum_s acc_s 128^(S-1-s), which is
// converted to FP64 once and scaled by the two groups' powers of two.  Only r and the dropped products (order above the
// kept one) differ from FP64, both bounded relative to the groups' powers of two.
//
//   X = P_red Phi           P: 5 digits per element, one power of two per row of P (sliced once per step).
//                           Phi: 5 digits, one power of two per point of a block (over the block's kept functions).
//                           With 4 digits, the values more than 2^-28 below a point's largest round to zero; they lie
//                           in the functions' tails and all of the same sign, so rho came out low (E_xc about +1.2e-8
//                           Eh per water).  Orders k + l <= 4: a dropped product carries its pair's sign, so dropping
//                           products also shrinks the terms of rho toward zero, by about 2e-10 Eh per water in E_xc
//                           (order 5 would add three products per step for that).  rho = sum Phi X and
//                           grad rho = 2 sum grad Phi X in FP64, with Phi from its digits and grad Phi from its fixed
//                           point.
//   V += Phi W^T + W Phi^T  Phi and W: 4 digits, one power of two per function row of a block (over its 128 points).
//                           Orders k + l <= 3.  W = (w vrho / 2) Phi + 2 w vsigma (grad rho . grad Phi) is formed and
//                           sliced per step.
// The functional, the weights and every reduction are FP64.  grad Phi is stored as 32-bit fixed point relative to
// each point's largest component.  Rounding it shifts E_xc through sigma = |grad rho|^2 roughly with the square of the
// rounding unit: at 24 bits (or as FP32 values) by about +6e-9 Eh per water, at 32 bits by below 1e-11.
//
// Data layouts.  The kept functions of a block are cut into chunks of 32 rows; chunk C of the grid is chunk rc of block b
// (C = chunkOffset[b] + rc).  Rows past the block's kept count are zero.  The Fock side works on superblocks: up to
// kSuper consecutive blocks of one atom, whose rows are the union of their kept functions (uIdx).  Each block of a
// superblock holds its Phi and W digits over the union rows, zero for a row that the block does not keep, so that
// the superblock's V tiles sum all its blocks before one FP64 update of V; union chunk CU of the grid is union chunk rc
// of block b (CU = uChunkOffset[b] + rc); at set-up, uRow maps each union chunk row to its row in the block's kept list.
//   DA, DW [CU][q (4)][digit (4)][lane (64)][16 bytes]  Phi and W digits in the operand order of the Fock product: lane
//                                  (lr, lh) = (l % 32, l / 32) holds union row 32 rc + lr at points 32 q + 16 lh .. + 15,
//                                  which is its A operand (rows of tile I) and its B operand (rows of tile J).  Exponents
//                                  EA, EW [CU][32], one per row over the block's own 128 points.
//   GHU, GLU [CU][3][q (4)][lane (64)][16]        grad Phi of the union rows in DA's order as 24-bit fixed point:
//                                  value = (256 GHU + GLU) 2^(EG - 23), GHU int16, GLU uint8 (W enters V only, whose
//                                  error its 4-digit rows bound); the W kernel reads them with 16-byte loads.
//   DB     [C][digit (5)][point (128)][32 bytes]   Phi digits in the B operand order of the density product.  Position
//                                  16 h + v of a point holds row 32 rc + 8 (v % 4) + 4 h + v / 4 (chunk_fun): the row
//                                  whose result the MFMA lane (., h) receives in register v (P_red rows are staged in
//                                  the tile rows tile_row), so that the epilogue finds Phi for its own results in the
//                                  operand it loads.  P_red tiles use the same K order, in which each 4-byte group
//                                  holds rows q, q + 8, q + 16, q + 24.
//                                  Exponents EB [block][128].
//   GF     [C][3][point (128)][32 int32]           grad Phi in DB's order as 32-bit fixed point: value = GF 2^(EG - 31),
//                                  one power of two per point and component over the block's kept functions
//                                  (EG [block][3][128]).
//   Pq     [n][n] 64-bit words                     P digits 0..4 in bytes 0..4; row exponents eP [n].
//
// MFMA lane layout (32 x 32 x 32, int8): lane l supplies A[l % 32][16 (l / 32) .. + 15] and B[16 (l / 32) .. + 15][l % 32]
// (16 consecutive K bytes each) and receives D[(v % 4) + 4 (l / 32) + 8 (v / 4)][l % 32] for v = 0..15.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <vector>

#include "xq_functional.h"
#include "xq_internal.h"

namespace xq {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

namespace {

constexpr int kC = 32;                  // chunk length (functions or points) and MFMA tile edge
constexpr int kDPhi = 4;                // digits of Phi and W in the Fock product
constexpr int kDB = 5;                  // digits of Phi in the density product
constexpr int kDP = 5;                  // digits of P
constexpr int kXOrders = 5;             // digit orders 0..4 kept in X = P_red Phi
constexpr int kChunkBytes = 16384;      // one chunk of DA or DW: 4 digits x 32 rows x 128 points
constexpr int kChunkBytesB = kDB * kBlockSize * kC;   // one chunk of DB: 5 digits x 128 points x 32 rows
constexpr int kChunkGrad = 3 * kBlockSize * kC;     // one chunk of GF, GHU or GLU (elements)
constexpr int kGradBits = 31;           // GF: |value| < 2^EG is stored as a multiple of 2^(EG - 31)
constexpr int kGradBitsW = 23;          // GHU, GLU: |value| < 2^EG is stored as a multiple of 2^(EG - 23)
constexpr int kPRowDwords = 12;         // one staged P_red row in LDS: 8 words of K positions, padded against bank conflicts
constexpr int kIdxCache = 1024;         // kept-function indices of a block cached in LDS by the density kernel
constexpr int kSuper = 4;               // blocks per superblock of the Fock product

// Stored exponents are int8.  A group whose largest magnitude is below 2^-100 (contributions below 1e-30 of unit-sized
// terms) gets exponent 0, which rounds all its digits to zero.  Magnitudes here stay far below 2^127.
constexpr int kMinStoredExponent = -100;

__device__ __forceinline__ v16i mfma_i8(const v4i &a, const v4i &b, const v16i &c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// Power of two strictly above max|x| (frexp exponent), clamped as described above.
__device__ __forceinline__ int stored_exponent(double amax) {
    int e = 0;
    if (amax > 0.0) (void)frexp(amax, &e);
    return e < kMinStoredExponent ? 0 : e;
}

template <int D>
__device__ __forceinline__ void slice(double x, int e, signed char *d) {
    const long long Q = static_cast<long long>(rint(ldexp(x, 7 * D - e)));
    // |x| < 2^e, so |Q| <= 2^(7D); the one value that rounds up to 2^(7D) is clamped to the largest D-digit magnitude.
    unsigned long long a = static_cast<unsigned long long>(Q < 0 ? -Q : Q);
    a = a < (1ull << (7 * D)) ? a : (1ull << (7 * D)) - 1ull;
    const int sign = Q < 0 ? -1 : 1;
#pragma unroll
    for (int k = 0; k < D; k++) d[k] = static_cast<signed char>(sign * static_cast<int>((a >> (7 * (D - 1 - k))) & 127ull));
}

// Exact combination of digit-order accumulators: sum_s acc_s 128^(S-1-s) as a 64-bit integer.
template <int S>
__device__ __forceinline__ long long combine(const int *acc) {
    long long t = 0;
#pragma unroll
    for (int k = 0; k < S; k++) t = t * 128 + static_cast<long long>(acc[k]);
    return t;
}

// Signed byte v (0..15) of a 16-byte operand, and 16 bytes packed into one.
__device__ __forceinline__ int byte_of(const v4i &x, int v) {
    return static_cast<int>(static_cast<unsigned>(x[v >> 2]) << (24 - 8 * (v & 3))) >> 24;
}

__device__ __forceinline__ v4i pack16(const signed char (&d)[16]) {
    v4i r;
#pragma unroll
    for (int w = 0; w < 4; w++)
        r[w] = static_cast<int>((static_cast<unsigned>(d[4 * w]) & 255u) | ((static_cast<unsigned>(d[4 * w + 1]) & 255u) << 8) |
                                ((static_cast<unsigned>(d[4 * w + 2]) & 255u) << 16) |
                                ((static_cast<unsigned>(d[4 * w + 3]) & 255u) << 24));
    return r;
}

// 4-digit value of byte v of the digit operands d[0..3]: sum_j d_j 128^(3-j).
__device__ __forceinline__ int digits4(const v4i (&d)[kDPhi], int v) {
    return ((byte_of(d[0], v) * 128 + byte_of(d[1], v)) * 128 + byte_of(d[2], v)) * 128 + byte_of(d[3], v);
}

// 5-digit value of byte v of the digit operands d[0..4]: sum_j d_j 128^(4-j).
__device__ __forceinline__ long long digits5(const v4i (&d)[kDB], int v) {
    const int hi = ((byte_of(d[0], v) * 128 + byte_of(d[1], v)) * 128 + byte_of(d[2], v)) * 128 + byte_of(d[3], v);
    return static_cast<long long>(hi) * 128 + byte_of(d[4], v);
}

// Signed 16-bit element v (0..15) of 32 bytes, and unsigned byte v of 16 bytes.
__device__ __forceinline__ int short_of(const v4i (&h)[2], int v) {
    const unsigned w = static_cast<unsigned>(h[v >> 3][(v >> 1) & 3]);
    return static_cast<int>(w << (16 - 16 * (v & 1))) >> 16;
}
__device__ __forceinline__ int ubyte_of(const v4i &x, int v) {
    return static_cast<int>((static_cast<unsigned>(x[v >> 2]) >> (8 * (v & 3))) & 255u);
}

// Element v (0..15) of sixteen int32 held in four 16-byte operands.
__device__ __forceinline__ double int_of(const v4i (&g)[4], int v) { return static_cast<double>(g[v >> 2][v & 3]); }

// B-bit fixed point of g with |g| < 2^e: round(g 2^(B - e)) in [-(2^B - 1), 2^B - 1].
template <int B>
__device__ __forceinline__ int grad_encode(double g, int e) {
    long long Q = static_cast<long long>(rint(ldexp(g, B - e)));
    const long long lim = (1ll << B) - 1;
    Q = Q > lim ? lim : (Q < -lim ? -lim : Q);
    return static_cast<int>(Q);
}

// Row of a chunk held at density position 16 h + v (K position 4 q + j holds row 8 j + q).
__host__ __device__ constexpr int chunk_fun(int h, int v) { return 8 * (v & 3) + 4 * h + (v >> 2); }
// The MFMA row that holds row r of a P_red tile: the output row (v % 4) + 4 h + 8 (v / 4) of lane half h, register v,
// is row chunk_fun(h, v), the row whose Phi digits the lane holds at position 16 h + v.
__host__ __device__ constexpr int tile_row(int r) { return 8 * (r & 3) + 4 * ((r >> 2) & 1) + (r >> 3); }

// Job of work-group g in a launch over several jobs: the last job whose first work-group in this stage is <= g (jobs
// without work in the stage share the next job's first work-group and are never returned for it).
template <typename F>
__device__ __forceinline__ int job_of(const xq::MixedJob *jobs, int nJobs, int g, F base) {
    int lo = 0, hi = nJobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base(jobs[mid]) <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------------------------------------------------------------
// Set-up, once per geometry.
// ------------------------------------------------------------------------------------------------------------------------
// EB and EG: per point of a block, the powers of two above max |Phi| and above max |d Phi / dx_c| over the block's kept
// functions (128 threads = points).
__global__ void __launch_bounds__(kBlockSize) point_exponent_kernel(
    const xq::DevBlock *__restrict__ blocks, const double *__restrict__ f, const double *__restrict__ fx,
    const double *__restrict__ fy, const double *__restrict__ fz, signed char *__restrict__ EB,
    signed char *__restrict__ EG) {
    const int b = blockIdx.x, t = threadIdx.x;
    const xq::DevBlock blk = blocks[b];
    const size_t R = static_cast<size_t>(blk.vOffset);
    double m = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    for (int r = 0; r < blk.nKept; r++) {
        const size_t o = (R + r) * kBlockSize + t;
        m = fmax(m, fabs(f[o]));
        mx = fmax(mx, fabs(fx[o]));
        my = fmax(my, fabs(fy[o]));
        mz = fmax(mz, fabs(fz[o]));
    }
    EB[static_cast<size_t>(b) * kBlockSize + t] = static_cast<signed char>(stored_exponent(m));
    signed char *eg = EG + static_cast<size_t>(b) * 3 * kBlockSize + t;
    eg[0] = static_cast<signed char>(stored_exponent(mx));
    eg[kBlockSize] = static_cast<signed char>(stored_exponent(my));
    eg[2 * kBlockSize] = static_cast<signed char>(stored_exponent(mz));
}

// One union chunk per work-group (256 threads): EA, DA, GHU and GLU of the chunk's 32 union rows from the FP64 values of
// the block's kept rows (zero for the rows the block does not keep).
__global__ void __launch_bounds__(256) fock_operand_prepare_kernel(
    const xq::DevBlock *__restrict__ blocks, const int *__restrict__ uChunkBlock, const int *__restrict__ uRow,
    const double *__restrict__ f, const double *__restrict__ fx, const double *__restrict__ fy,
    const double *__restrict__ fz, const signed char *__restrict__ EG, signed char *__restrict__ DA,
    signed char *__restrict__ EA, short *__restrict__ GHU, unsigned char *__restrict__ GLU) {
    __shared__ int sE[kC];
    __shared__ int sRow[kC];
    const size_t C = blockIdx.x;
    const int b = uChunkBlock[C];
    const xq::DevBlock blk = blocks[b];
    const int t = threadIdx.x;
    const size_t R = static_cast<size_t>(blk.vOffset);
    if (t < kC) sRow[t] = uRow[C * kC + t];
    __syncthreads();
    auto value = [&](const double *a, int i, int p) -> double {
        const int k = sRow[i];
        return k >= 0 ? a[(R + k) * kBlockSize + p] : 0.0;
    };
    // EA: row i = t / 8 of the chunk, eight threads per row with 16 points each.
    {
        const int i = t >> 3, part = t & 7;
        double m = 0.0;
#pragma unroll 4
        for (int k = 0; k < 16; k++) m = fmax(m, fabs(value(f, i, 16 * part + k)));
#pragma unroll
        for (int o = 4; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 8));
        if (part == 0) {
            const int e = stored_exponent(m);
            sE[i] = e;
            EA[C * kC + i] = static_cast<signed char>(e);
        }
    }
    __syncthreads();
    // DA: thread = (q, lane).
    const int q = t >> 6, lane = t & 63, lr = lane & 31, lh = lane >> 5, e = sE[lr];
    signed char d[kDPhi][16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        signed char dd[kDPhi];
        slice<kDPhi>(value(f, lr, kC * q + 16 * lh + k), e, dd);
#pragma unroll
        for (int j = 0; j < kDPhi; j++) d[j][k] = dd[j];
    }
#pragma unroll
    for (int j = 0; j < kDPhi; j++)
        *reinterpret_cast<v4i *>(&DA[C * kChunkBytes + (q * kDPhi + j) * 1024 + lane * 16]) = pack16(d[j]);
    // GHU, GLU: the same thread's 16 values of each gradient component, scaled by each point's component exponent.
    const double *grads[3] = {fx, fy, fz};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        int Q[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int p = kC * q + 16 * lh + k;
            Q[k] = grad_encode<kGradBitsW>(value(grads[c], lr, p), EG[(static_cast<size_t>(b) * 3 + c) * kBlockSize + p]);
        }
        // Q = 256 hi + lo with lo in [0, 255]: hi = Q >> 8 (arithmetic), lo = Q & 255.
        v4i hq[2], lq;
#pragma unroll
        for (int w = 0; w < 8; w++)
            hq[w >> 2][w & 3] = static_cast<int>((static_cast<unsigned>(Q[2 * w] >> 8) & 0xffffu) |
                                                 (static_cast<unsigned>(Q[2 * w + 1] >> 8) << 16));
#pragma unroll
        for (int w = 0; w < 4; w++)
            lq[w] = static_cast<int>((static_cast<unsigned>(Q[4 * w]) & 255u) | ((static_cast<unsigned>(Q[4 * w + 1]) & 255u) << 8) |
                                     ((static_cast<unsigned>(Q[4 * w + 2]) & 255u) << 16) |
                                     ((static_cast<unsigned>(Q[4 * w + 3]) & 255u) << 24));
        const size_t o = C * kChunkGrad + ((c * 4 + q) * 64 + lane) * 16;
        *reinterpret_cast<v4i *>(GHU + o) = hq[0];
        *reinterpret_cast<v4i *>(GHU + o + 8) = hq[1];
        *reinterpret_cast<v4i *>(GLU + o) = lq;
    }
}

// One chunk per work-group (256 threads): DB and GF from the FP64 values of the chunk's rows.
__global__ void __launch_bounds__(256) density_operand_prepare_kernel(
    const xq::DevBlock *__restrict__ blocks, const int *__restrict__ chunkOffset, const int *__restrict__ chunkBlock,
    const double *__restrict__ f, const double *__restrict__ fx, const double *__restrict__ fy,
    const double *__restrict__ fz, const signed char *__restrict__ EB, const signed char *__restrict__ EG,
    signed char *__restrict__ DB, int *__restrict__ GF) {
    const size_t C = blockIdx.x;
    const int b = chunkBlock[C];
    const xq::DevBlock blk = blocks[b];
    const int nb = blk.nKept, rc = static_cast<int>(C) - chunkOffset[b], t = threadIdx.x;
    const size_t R = static_cast<size_t>(blk.vOffset);
    auto value = [&](const double *a, int row, int p) -> double {
        return row < nb ? a[(R + row) * kBlockSize + p] : 0.0;
    };
    // DB and GF: thread = (point p, half h) with the 16 rows chunk_fun(h, 0..15).
    {
        const int p = t & (kBlockSize - 1), h = t >> 7;
        const int e = EB[static_cast<size_t>(b) * kBlockSize + p];
        signed char d[kDB][16];
#pragma unroll
        for (int v = 0; v < 16; v++) {
            signed char dd[kDB];
            slice<kDB>(value(f, kC * rc + chunk_fun(h, v), p), e, dd);
#pragma unroll
            for (int j = 0; j < kDB; j++) d[j][v] = dd[j];
        }
#pragma unroll
        for (int j = 0; j < kDB; j++)
            *reinterpret_cast<v4i *>(&DB[C * kChunkBytesB + (j * kBlockSize + p) * kC + 16 * h]) = pack16(d[j]);
        const double *grads[3] = {fx, fy, fz};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int eg = EG[(static_cast<size_t>(b) * 3 + c) * kBlockSize + p];
            v4i gq[4];
#pragma unroll
            for (int v = 0; v < 16; v++)
                gq[v >> 2][v & 3] = grad_encode<kGradBits>(value(grads[c], kC * rc + chunk_fun(h, v), p), eg);
            const size_t o = C * kChunkGrad + (c * kBlockSize + p) * kC + 16 * h;
#pragma unroll
            for (int w = 0; w < 4; w++) *reinterpret_cast<v4i *>(GF + o + 4 * w) = gq[w];
        }
    }
}

// ------------------------------------------------------------------------------------------------------------------------
// Per step.
// ------------------------------------------------------------------------------------------------------------------------
// P row i (one work-group of 256 threads per row of every job): row exponent and the 5 digits of every element.
__global__ void __launch_bounds__(256) p_slice_kernel(const xq::MixedJob *__restrict__ jobs, int nJobs) {
    __shared__ double s[4];
    const xq::MixedJob &J = jobs[job_of(jobs, nJobs, blockIdx.x, [](const xq::MixedJob &j) { return j.rowBase; })];
    const int n = J.n;
    const double *__restrict__ P = J.P;
    int *__restrict__ eP = J.eP;
    unsigned long long *__restrict__ Pq = J.Pq;
    const size_t i = blockIdx.x - J.rowBase;
    const double *row = P + i * n;
    double m = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) m = fmax(m, fabs(row[k]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    const int e = stored_exponent(fmax(fmax(s[0], s[1]), fmax(s[2], s[3])));
    for (int k = threadIdx.x; k < n; k += 256) {
        signed char d[kDP];
        slice<kDP>(row[k], e, d);
        unsigned long long w = 0;
#pragma unroll
        for (int j = 0; j < kDP; j++) w |= (static_cast<unsigned long long>(d[j]) & 255ull) << (8 * j);
        Pq[i * n + k] = w;
    }
    if (threadIdx.x == 0) eP[i] = e;
}

// Density rows (one row chunk I of one block per work-group, 256 threads = 4 waves): X = P_red Phi for the chunk's 32
// rows on int8 MFMA, then the chunk's share of rho and grad rho at the block's points, stored in part[chunk][4][point].
// Wave wv owns points 32 wv .. + 31; lane (lr, lh) owns point 32 wv + lr and the rows chunk_fun(lh, v) of chunk I.
// Each step stages one column chunk of P_red's row panel (gathered) and that chunk's Phi digits (16-byte loads of the
// contiguous DB chunk) in LDS, both fetched by all 256 threads into registers one step ahead.  The row chunks of a
// block are consecutive items of one XCD, so that their work-groups run at about the same time and read the block's
// Phi digit chunks, in the same order, through that XCD's L2: Phi comes from HBM about once per block instead of once
// per row chunk.  Chunk I's digits, staged at step I, are kept in LDS for the epilogue.  52 KB of LDS and at most 168
// VGPRs let three work-groups share a CU (four, at 128 VGPRs, were slower: 3.56 against 2.75 ms on (H2O)64).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
density_rows_kernel(const xq::MixedJob *__restrict__ jobs, int nJobs, int interleaved) {
    __shared__ __attribute__((aligned(16))) int sP[kDP][kC][kPRowDwords];   // [digit][row][K positions / 4]
    __shared__ v4i sB[kChunkBytesB / 16];   // the step's Phi digits, in DB's order
    __shared__ v4i sBI[kChunkBytesB / 16];  // chunk I's Phi digits, kept from step I for the epilogue
    __shared__ int sPe[kC];
    __shared__ int sIdx[kIdxCache];
    // Work-group g runs on XCD x = g % 8 and takes slot g / 8 of that XCD's item list: the jobs' lists in job order, or,
    // for the two spins of one handle (interleaved), item slot / 2 of job slot % 2, so that both spins' work-groups of
    // a row chunk read its block's Phi chunks through L2 at about the same time.
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const xq::MixedJob &J = interleaved ? jobs[slot & 1]
                                        : jobs[job_of(jobs, nJobs, slot, [xcd](const xq::MixedJob &j) {
                                              return j.itemBase[xcd];
                                          })];
    const int local = interleaved ? slot >> 1 : slot - J.itemBase[xcd];
    if (local >= J.itemStart[xcd + 1] - J.itemStart[xcd]) return;
    const signed char *__restrict__ DB = J.DB;
    const int *__restrict__ GF = J.GF;
    const int n = J.n;
    const unsigned long long *__restrict__ Pq = J.Pq;
    const int *__restrict__ eP = J.eP;
    // The item's record ({chunk, kept functions x 1024 + row chunk, first kept index, block}): one load gives what the
    // first fetches need.
    const int4 ent = J.items[J.itemStart[xcd] + local];
    const size_t CI = static_cast<size_t>(ent.x);
    const int nb = ent.y >> 10, I = ent.y & 1023, b = ent.w;
    const size_t C = CI - I;
    const int nChunks = (nb + kC - 1) / kC;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int p = kC * wv + lr;
    const signed char *DBb = DB + C * kChunkBytesB;
    // P_red is staged one column chunk at a time: thread (row sr, slot s8) fetches columns 8 j + s8 (j = 0 .. 3) of
    // the chunk (eight lanes read eight consecutive kept functions) and writes them as one word per digit at K
    // positions 4 s8 .. + 3, in row tile_row(sr).
    const int sr = t >> 3, s8 = t & 7, srow = tile_row(sr);
    // The block's kept-function indices, cached in LDS for the P_red gathers (blocks with more kept functions than the
    // cache read them from memory).
    const int *idxG = J.idx + ent.z;
    for (int k = t; k < nb && k < kIdxCache; k += 256) sIdx[k] = idxG[k];
    __syncthreads();
    const int *idx = nb <= kIdxCache ? sIdx : idxG;
    if (t < kC) sPe[t] = kC * I + t < nb ? eP[idx[kC * I + t]] : 0;
    const int r = kC * I + sr;
    const size_t rowOff = r < nb ? static_cast<size_t>(idx[r]) * n : 0;
    v16i acc[kXOrders];
#pragma unroll
    for (int s = 0; s < kXOrders; s++) acc[s] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    // One chunk per step; the next step's P_red words and Phi digits are loaded into registers while the current
    // step's products run.
    constexpr int kPhiPerThread = kChunkBytesB / 16 / 256;
    const v4i *phiSrc = reinterpret_cast<const v4i *>(DBb);
    unsigned long long wd[4];
    v4i ph[kPhiPerThread];
    auto fetch = [&](int c) {
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const int k = kC * c + 8 * jj + s8;
            wd[jj] = (r < nb && k < nb) ? Pq[rowOff + idx[k]] : 0ull;
        }
#pragma unroll
        for (int i = 0; i < kPhiPerThread; i++)
            ph[i] = phiSrc[static_cast<size_t>(c) * (kChunkBytesB / 16) + t + 256 * i];
    };
    fetch(0);
    for (int c = 0; c < nChunks; c++) {
        __syncthreads();   // every wave is done with the previous chunk
#pragma unroll
        for (int i = 0; i < kPhiPerThread; i++) sB[t + 256 * i] = ph[i];
        if (c == I) {
#pragma unroll
            for (int i = 0; i < kPhiPerThread; i++) sBI[t + 256 * i] = ph[i];
        }
#pragma unroll
        for (int j = 0; j < kDP; j++) {
            unsigned x = 0;
#pragma unroll
            for (int jj = 0; jj < 4; jj++) x |= (static_cast<unsigned>(wd[jj] >> (8 * j)) & 255u) << (8 * jj);
            sP[j][srow][s8] = static_cast<int>(x);
        }
        __syncthreads();
        if (c + 1 < nChunks) fetch(c + 1);
        v4i a[kDP], bq[kDB];
#pragma unroll
        for (int j = 0; j < kDB; j++) bq[j] = sB[(j * kBlockSize + p) * 2 + lh];
#pragma unroll
        for (int j = 0; j < kDP; j++) a[j] = *reinterpret_cast<const v4i *>(&sP[j][lr][4 * lh]);
#pragma unroll
        for (int i = 0; i < kDP; i++)
#pragma unroll
            for (int j = 0; j < kDB; j++)
                if (i + j < kXOrders) acc[i + j] = mfma_i8(a[i], bq[j], acc[i + j]);
    }
    // Epilogue: X for rows chunk_fun(lh, v) of chunk I at point p; Phi from the chunk's own digits, grad Phi from its
    // fixed point (the sums over v are scaled once).  X in FP64 first, so that the accumulators are free before the
    // epilogue's operands arrive.
    const int eb = J.EB[static_cast<size_t>(b) * kBlockSize + p];
    double x[16];
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int s[kXOrders] = {acc[0][v], acc[1][v], acc[2][v], acc[3][v], acc[4][v]};
        x[v] = ldexp(static_cast<double>(combine<kXOrders>(s)), sPe[chunk_fun(lh, v)] + eb - 7 * (kXOrders + 1));
    }
    v4i dI[kDB];   // the lane's Phi digits of chunk I (written to sBI before the last step's barrier)
#pragma unroll
    for (int j = 0; j < kDB; j++) dI[j] = sBI[(j * kBlockSize + p) * 2 + lh];
    double sumR = 0.0, sumG[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 16; v++) sumR += static_cast<double>(digits5(dI, v)) * x[v];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const size_t o = CI * kChunkGrad + (c * kBlockSize + p) * kC + 16 * lh;
        v4i gq[4];
#pragma unroll
        for (int w = 0; w < 4; w++) gq[w] = *reinterpret_cast<const v4i *>(GF + o + 4 * w);
#pragma unroll
        for (int v = 0; v < 16; v++) sumG[c] += int_of(gq, v) * x[v];
    }
    double out[4] = {ldexp(sumR, eb - 7 * kDB), 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 3; c++)
        out[c + 1] = sumG[c] * ldexp(1.0, J.EG[(static_cast<size_t>(b) * 3 + c) * kBlockSize + p] - kGradBits);
#pragma unroll
    for (int c = 0; c < 4; c++) out[c] += __shfl_xor(out[c], 32, 64);
    if (lh == 0) {
        double *q = J.part + CI * 4 * kBlockSize;
#pragma unroll
        for (int c = 0; c < 4; c++) q[c * kBlockSize + p] = out[c];
    }
}

// Density functional (one block per work-group, 128 threads = one per point): rho and grad rho as the sums of the
// block's row-chunk shares in chunk order, the FP64 functional, E_xc / N_e block sums and the four W coefficients per
// point.  kRhoOut: the density stage of one spin of an unrestricted quadrature (the job's rhoOut receives rho and
// grad rho).
template <bool kRhoOut>
__global__ void __launch_bounds__(kBlockSize)
density_functional_kernel(const xq::MixedJob *__restrict__ jobs, int nJobs) {
    __shared__ double wsum[2][2];
    const xq::MixedJob &J = jobs[job_of(jobs, nJobs, blockIdx.x, [](const xq::MixedJob &j) { return j.blockBase; })];
    const int b = blockIdx.x - J.blockBase;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int pStart = J.blocks[b].pStart, np = J.blocks[b].nPoints;
    const int c0 = J.chunkOffset[b], T = J.chunkOffset[b + 1] - c0;
    double rho = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    if (t < np) {
        const double *q = J.part + static_cast<size_t>(c0) * 4 * kBlockSize + t;
        for (int I = 0; I < T; I++, q += 4 * kBlockSize) {
            rho += q[0];
            gx += q[kBlockSize];
            gy += q[2 * kBlockSize];
            gz += q[3 * kBlockSize];
        }
        gx *= 2.0;
        gy *= 2.0;
        gz *= 2.0;
    }
    if (kRhoOut) {   // one spin of an unrestricted quadrature: the functional follows in its own kernel
        if (t < np) {
            double *o = J.rhoOut + 4 * (static_cast<size_t>(pStart) + t);
            o[0] = rho; o[1] = gx; o[2] = gy; o[3] = gz;
        }
        return;
    }
    double mx = fabs(rho);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) wsum[wv][0] = mx;
    __syncthreads();
    const bool skipped = T == 0 || !(fmax(wsum[0][0], wsum[1][0]) > J.rhoTol);
    __syncthreads();
    double e = 0.0, ne = 0.0;
    if (t < np) {
        double vr = 0.0, vs = 0.0;
        if (!skipped) {
            double eps;
            evaluate_functional(J.fs, rho, gx * gx + gy * gy + gz * gz, eps, vr, vs);
            const double ww = J.w[pStart + t];
            vr *= ww;
            vs *= ww;
            e = eps * rho * ww;
            ne = rho * ww;
        }
        double *cw = J.wcoef + 4 * (static_cast<size_t>(pStart) + t);
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
        J.blockEN[2 * b] = wsum[0][0] + wsum[1][0];
        J.blockEN[2 * b + 1] = wsum[0][1] + wsum[1][1];
        J.skip[b] = skipped ? 1 : 0;
    }
}

// W digits (one union chunk per work-group, 256 threads = 4 waves = the 4 point chunks q):
// W_rp = c0_p Phi_rp + c_p . grad Phi_rp for the chunk's 32 union rows (zero for the rows the block does not keep), one
// power of two per row over the block's 128 points, written in DA's order (DW, EW).
__global__ void __launch_bounds__(256) w_digits_kernel(const xq::MixedJob *__restrict__ jobs, int nJobs) {
    __shared__ double sc[4][kBlockSize];
    __shared__ double sm[4][kC];
    const xq::MixedJob &J = jobs[job_of(jobs, nJobs, blockIdx.x, [](const xq::MixedJob &j) { return j.chunkBase; })];
    const xq::DevBlock *__restrict__ blocks = J.blocks;
    const int *__restrict__ uChunkBlock = J.uChunkBlock;
    const signed char *__restrict__ DA = J.DA;
    const signed char *__restrict__ EA = J.EA;
    const short *__restrict__ GHU = J.GHU;
    const unsigned char *__restrict__ GLU = J.GLU;
    const signed char *__restrict__ EG = J.EG;
    const double *__restrict__ wcoef = J.wcoef;
    const int *__restrict__ skip = J.skip;
    signed char *__restrict__ DW = J.DW;
    signed char *__restrict__ EW = J.EW;
    const size_t C = blockIdx.x - J.chunkBase;
    const int b = uChunkBlock[C];
    if (skip[b]) return;
    const xq::DevBlock blk = blocks[b];
    const int t = threadIdx.x, lane = t & 63, q = t >> 6, lr = lane & 31, lh = lane >> 5;
    // Per point: w vrho / 2, and 2 w vsigma grad rho scaled by the point's gradient powers of two.
    if (t < kBlockSize) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double c = t < blk.nPoints ? wcoef[4 * (static_cast<size_t>(blk.pStart) + t) + k] : 0.0;
            sc[k][t] = k == 0 ? c : ldexp(c, EG[(static_cast<size_t>(b) * 3 + (k - 1)) * kBlockSize + t] - kGradBitsW);
        }
    }
    __syncthreads();
    v4i d[kDPhi];
#pragma unroll
    for (int j = 0; j < kDPhi; j++)
        d[j] = *reinterpret_cast<const v4i *>(&DA[C * kChunkBytes + (q * kDPhi + j) * 1024 + lane * 16]);
    const int ea = EA[C * kC + lr];
    v4i gh[3][2], gl[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const size_t o = C * kChunkGrad + ((c * 4 + q) * 64 + lane) * 16;
        gh[c][0] = *reinterpret_cast<const v4i *>(GHU + o);
        gh[c][1] = *reinterpret_cast<const v4i *>(GHU + o + 8);
        gl[c] = *reinterpret_cast<const v4i *>(GLU + o);
    }
    auto grad = [&](int c, int v) { return static_cast<double>(short_of(gh[c], v) * 256 + ubyte_of(gl[c], v)); };
    double W[16];
    double m = 0.0;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int p = kC * q + 16 * lh + v;
        const double phi = ldexp(static_cast<double>(digits4(d, v)), ea - 28);
        W[v] = sc[0][p] * phi + sc[1][p] * grad(0, v) + sc[2][p] * grad(1, v) + sc[3][p] * grad(2, v);
        m = fmax(m, fabs(W[v]));
    }
    m = fmax(m, __shfl_xor(m, 32, 64));
    if (lh == 0) sm[q][lr] = m;
    __syncthreads();
    const int e = stored_exponent(fmax(fmax(sm[0][lr], sm[1][lr]), fmax(sm[2][lr], sm[3][lr])));
    signed char dw[kDPhi][16];
#pragma unroll
    for (int v = 0; v < 16; v++) {
        signed char dd[kDPhi];
        slice<kDPhi>(W[v], e, dd);
#pragma unroll
        for (int j = 0; j < kDPhi; j++) dw[j][v] = dd[j];
    }
#pragma unroll
    for (int j = 0; j < kDPhi; j++)
        *reinterpret_cast<v4i *>(&DW[C * kChunkBytes + (q * kDPhi + j) * 1024 + lane * 16]) = pack16(dw[j]);
    if (q == 0 && lh == 0) EW[C * kC + lr] = static_cast<signed char>(e);
}

// Fock kernel: one wave per lower-triangle tile (I, J <= I) of a superblock's reduced V, four tiles per work-group, no
// LDS and no barriers.  For each block of the superblock that is not skipped, K runs over its 128 points in four chunks
// of 32; each product accumulates in int32 over all of them (one power of two per row and block) and is promoted into
// the tile's FP64 sums; V += sum_b (Phi_I W_J^T + W_I Phi_J^T) is then added with one set of FP64 atomics.
// Work-group g runs on XCD g % 8; tiles are dealt so that each XCD takes a contiguous range of the tile list (tiles of
// one superblock share their chunks' digits in that XCD's L2).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
fock_kernel_mixed(const xq::MixedJob *__restrict__ jobs, int nJobs, int groupsPerXcd, int nGroups, int interleaved) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, lr = lane & 31, lh = lane >> 5;
    const int group = (blockIdx.x & 7) * groupsPerXcd + (blockIdx.x >> 3);
    if (group >= nGroups) return;
    // The two spins of one handle (interleaved): group g is group g / 2 of job g % 2, so that both spins' tiles of a
    // superblock share its Phi digits through L2.
    const xq::MixedJob &job = interleaved ? jobs[group & 1]
                                          : jobs[job_of(jobs, nJobs, group, [](const xq::MixedJob &j) {
                                                return j.groupBase;
                                            })];
    const int tileIndex = 4 * (interleaved ? group >> 1 : group - job.groupBase) + wv;
    if (tileIndex >= job.nTiles) return;
    const int *__restrict__ uChunkOffset = job.uChunkOffset;
    const signed char *__restrict__ DA = job.DA;
    const signed char *__restrict__ EA = job.EA;
    const signed char *__restrict__ DW = job.DW;
    const signed char *__restrict__ EW = job.EW;
    const int n = job.n;
    const int *__restrict__ skip = job.skip;
    double *__restrict__ V = job.V;
    const int4 tl = job.uTiles[tileIndex];
    const int I = tl.y, J = tl.z;
    const int4 sb = job.supers[tl.x];      // first block, blocks, union offset, union size
    const int nu = sb.w;
    const int *idx = job.uIdx + sb.z;
    const int gj = kC * J + lr;
    double x[16];
#pragma unroll
    for (int v = 0; v < 16; v++) x[v] = 0.0;
    bool any = false;
#pragma unroll 1
    for (int b = sb.x; b < sb.x + sb.y; b++) {
        if (skip[b]) continue;
        any = true;
        const size_t CI = static_cast<size_t>(uChunkOffset[b] + I), CJ = static_cast<size_t>(uChunkOffset[b] + J);
        // Two passes, Phi_I W_J^T then W_I Phi_J^T, each promoted into the FP64 sums before the next starts.
#pragma unroll 1
        for (int pass = 0; pass < 2; pass++) {
            const signed char *A = (pass == 0 ? DA : DW) + CI * kChunkBytes + lane * 16;
            const signed char *B = (pass == 0 ? DW : DA) + CJ * kChunkBytes + lane * 16;
            const signed char *eA = (pass == 0 ? EA : EW) + CI * kC, *eB = (pass == 0 ? EW : EA) + CJ * kC;
            v16i acc[4];
#pragma unroll
            for (int s = 0; s < 4; s++) acc[s] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            // All four point chunks' operands are loaded before their products (32 loads in flight per lane).
#pragma unroll
            for (int q = 0; q < 4; q++) {
                v4i xa[kDPhi], xb[kDPhi];
#pragma unroll
                for (int j = 0; j < kDPhi; j++) {
                    xa[j] = *reinterpret_cast<const v4i *>(A + (q * kDPhi + j) * 1024);
                    xb[j] = *reinterpret_cast<const v4i *>(B + (q * kDPhi + j) * 1024);
                }
#pragma unroll
                for (int i = 0; i < kDPhi; i++)
#pragma unroll
                    for (int j = 0; j < kDPhi; j++)
                        if (i + j <= 3) acc[i + j] = mfma_i8(xa[i], xb[j], acc[i + j]);
            }
            // Lane (lr, lh) holds rows 32 I + rl(v) and column 32 J + lr.
            const int eCol = eB[lr] - 35;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int rl = (v & 3) + 4 * lh + 8 * (v >> 2);
                const int tv[4] = {acc[0][v], acc[1][v], acc[2][v], acc[3][v]};
                x[v] += ldexp(static_cast<double>(combine<4>(tv)), eA[rl] + eCol);
            }
        }
    }
    if (!any) return;
    const size_t col = gj < nu ? static_cast<size_t>(idx[gj]) : 0;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int gi = kC * I + (v & 3) + 4 * lh + 8 * (v >> 2);
        if (gi < nu && gj < nu && gj <= gi) atomicAdd(&V[static_cast<size_t>(idx[gi]) * n + col], x[v]);
    }
}

}  // namespace

// Host side: digit storage at basis set-up and the per-step launches.
// Chunks of 32 kept functions per block (chunkOffset: first chunk of each block; chunkBlock: block of each chunk) and the
// lower-triangle Fock tiles (block, I, J <= I) of every block, for both precisions.
int chunk_tables(xq_grid *g) {
    if (g->d_chunkOffset != nullptr) return kOK;
    std::vector<int> chunkOffset(g->nBlocks + 1, 0), chunkBlock;
    std::vector<int4> tiles;
    long long chunks = 0;
    for (int b = 0; b < g->nBlocks; b++) {
        const int nb = g->blocks[b].nKept;
        // The density accumulators of the mixed path hold up to 5 x 127^2 x nb per order in int32 (nb <= 26630), and
        // their exact combination sum_s acc_s 128^(4-s) stays far inside int64.
        if (nb > 26000) { set_error("at most 26000 kept functions per block are supported"); return kInvalidArgument; }
        const int T = (nb + kC - 1) / kC;
        chunkOffset[b] = static_cast<int>(chunks);
        for (int I = 0; I < T; I++) {
            chunkBlock.push_back(b);
            for (int J = 0; J <= I; J++) tiles.push_back(make_int4(b, I, J, 0));
        }
        chunks += T;
    }
    chunkOffset[g->nBlocks] = static_cast<int>(chunks);
    if (chunks > INT32_MAX / 4 || tiles.size() > static_cast<size_t>(INT32_MAX / 4)) {
        set_error("too many chunks for one grid");
        return kInvalidArgument;
    }
    g->nChunks = chunks;
    g->nTiles = static_cast<long long>(tiles.size());
    const size_t nc = static_cast<size_t>(chunks > 0 ? chunks : 1);
    XQ_TRY(dev_alloc(g, &g->d_chunkOffset, chunkOffset.size()));
    XQ_TRY(dev_alloc(g, &g->d_chunkBlock, nc));
    XQ_TRY(dev_alloc(g, &g->d_tiles, tiles.empty() ? 1 : tiles.size()));
    hipStream_t s = g->stream;
    XQ_CHECK(hipMemcpyAsync(g->d_chunkOffset, chunkOffset.data(), chunkOffset.size() * sizeof(int), hipMemcpyHostToDevice, s));
    if (chunks > 0)
        XQ_CHECK(hipMemcpyAsync(g->d_chunkBlock, chunkBlock.data(), chunkBlock.size() * sizeof(int), hipMemcpyHostToDevice, s));
    if (!tiles.empty())
        XQ_CHECK(hipMemcpyAsync(g->d_tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice, s));
    // The host vectors must outlive the copies.
    XQ_CHECK(hipStreamSynchronize(s));
    return build_block_order(g, s);
}

// The density launch's block order.  A block of T row chunks costs about T^2 MFMA steps of the density kernel, plus
// one epilogue per row chunk and the functional (counted as one step each, and one for a block without kept
// functions).  The blocks are cut into 8 contiguous ranges of equal summed cost, one per XCD, each sorted by decreasing
// cost: on (H2O)64 the largest blocks take about 28 times the mean, and started last they set the length of the launch.
// The mixed density launch takes the same order with each block cut into its row chunks, consecutive on its XCD.
int build_block_order(xq_grid *g, hipStream_t s) {
    const int nB = g->nBlocks;
    std::vector<double> cost(static_cast<size_t>(nB > 0 ? nB : 1), 0.0);
    double total = 0.0;
    for (int b = 0; b < nB; b++) {
        const double T = static_cast<double>((g->blocks[b].nKept + kC - 1) / kC);
        cost[b] = T * T + T + 1.0;
        total += cost[b];
    }
    g->xcdStart[0] = 0;
    int b = 0;
    double acc = 0.0;
    for (int x = 1; x < 8; x++) {
        const double target = total * x / 8.0;
        while (b < nB && acc + 0.5 * cost[b] < target) acc += cost[b++];
        g->xcdStart[x] = b;
    }
    g->xcdStart[8] = nB;
    std::vector<int> order(static_cast<size_t>(nB > 0 ? nB : 1), 0), firstChunk(static_cast<size_t>(nB > 0 ? nB : 1), 0);
    for (int i = 0, c = 0; i < nB; i++) {
        order[i] = i;
        firstChunk[i] = c;
        c += (g->blocks[i].nKept + kC - 1) / kC;
    }
    for (int x = 0; x < 8; x++)
        std::stable_sort(order.begin() + g->xcdStart[x], order.begin() + g->xcdStart[x + 1],
                         [&](int u, int v) { return cost[u] > cost[v]; });
    g->h_order.assign(order.size(), make_int4(0, 0, 0, 0));
    g->h_items.clear();
    for (int x = 0; x < 8; x++) {
        g->itemStart[x] = static_cast<int>(g->h_items.size());
        for (int i = g->xcdStart[x]; i < g->xcdStart[x + 1]; i++) {
            const xq::DevBlock &blk = g->blocks[order[i]];
            g->h_order[i] = make_int4(order[i], firstChunk[order[i]], blk.iOffset, blk.nKept * 256 + blk.nPoints);
            // The block's row chunks, one item each (nKept <= 26000, so that row chunk I < 1024 fits beside it).
            for (int I = 0; I < (blk.nKept + kC - 1) / kC; I++)
                g->h_items.push_back(make_int4(firstChunk[order[i]] + I, blk.nKept * 1024 + I, blk.iOffset, order[i]));
        }
    }
    g->itemStart[8] = static_cast<int>(g->h_items.size());
    XQ_TRY(dev_alloc(g, &g->d_order, g->h_order.size()));
    XQ_TRY(dev_alloc(g, &g->d_items, g->h_items.empty() ? 1 : g->h_items.size()));
    XQ_CHECK(hipMemcpyAsync(g->d_order, g->h_order.data(), g->h_order.size() * sizeof(int4), hipMemcpyHostToDevice, s));
    if (!g->h_items.empty())
        XQ_CHECK(hipMemcpyAsync(g->d_items, g->h_items.data(), g->h_items.size() * sizeof(int4), hipMemcpyHostToDevice,
                                s));
    XQ_CHECK(hipStreamSynchronize(s));
    return kOK;
}

// Superblocks of the Fock product: runs of up to kSuper consecutive blocks of one atom, the union of their kept
// functions (sorted), each block's union chunks and the map from union rows to the block's kept rows, and the
// lower-triangle tiles (superblock, I, J <= I) over the union chunks.
static int union_tables(xq_grid *g) {
    const int nB = g->nBlocks;
    std::vector<int> idx(static_cast<size_t>(g->nValueRows > 0 ? g->nValueRows : 1));
    XQ_CHECK(hipStreamSynchronize(g->stream));
    if (g->nValueRows > 0)
        XQ_CHECK(hipMemcpyAsync(idx.data(), g->d_idx, g->nValueRows * sizeof(int), hipMemcpyDeviceToHost, g->stream));
    XQ_CHECK(hipStreamSynchronize(g->stream));
    std::vector<int4> supers, uTiles;
    std::vector<int> uIdx, uChunkOffset(static_cast<size_t>(nB) + 1, 0), uChunkBlock, uRow;
    std::vector<int> stamp(static_cast<size_t>(g->nFunctions > 0 ? g->nFunctions : 1), -1);
    long long chunks = 0;
    double unionValues = 0.0;
    for (int b0 = 0; b0 < nB;) {
        int b1 = b0 + 1;
        while (b1 < nB && b1 - b0 < kSuper && g->blocks[b1].atom == g->blocks[b0].atom) b1++;
        const int s = static_cast<int>(supers.size());
        const size_t u0 = uIdx.size();
        for (int b = b0; b < b1; b++)
            for (int k = 0; k < g->blocks[b].nKept; k++) {
                const int fn = idx[static_cast<size_t>(g->blocks[b].iOffset) + k];
                if (stamp[fn] != s) { stamp[fn] = s; uIdx.push_back(fn); }
            }
        std::sort(uIdx.begin() + u0, uIdx.end());
        const int nu = static_cast<int>(uIdx.size() - u0);
        if (nu > 32767) { set_error("at most 32767 kept functions per superblock are supported"); return kInvalidArgument; }
        supers.push_back(make_int4(b0, b1 - b0, static_cast<int>(u0), nu));
        const int T = (nu + kC - 1) / kC;
        for (int b = b0; b < b1; b++) {
            unionValues += static_cast<double>(nu) * g->blocks[b].nPoints;
            // Both lists are sorted: one merge gives each union row's place in the block's kept list.
            const int *kept = idx.data() + g->blocks[b].iOffset;
            const int nk = g->blocks[b].nKept;
            uChunkOffset[b] = static_cast<int>(chunks);
            int k = 0;
            for (int r = 0; r < T * kC; r++) {
                int m = -1;
                if (r < nu) {
                    const int fn = uIdx[u0 + r];
                    while (k < nk && kept[k] < fn) k++;
                    if (k < nk && kept[k] == fn) m = k;
                }
                uRow.push_back(m);
            }
            for (int I = 0; I < T; I++) uChunkBlock.push_back(b);
            chunks += T;
        }
        for (int I = 0; I < T; I++)
            for (int J = 0; J <= I; J++) uTiles.push_back(make_int4(s, I, J, 0));
        b0 = b1;
    }
    uChunkOffset[nB] = static_cast<int>(chunks);
    if (chunks > INT32_MAX / 4 || uTiles.size() > static_cast<size_t>(INT32_MAX / 4) ||
        uIdx.size() > static_cast<size_t>(INT32_MAX)) {
        set_error("too many union chunks for one grid");
        return kInvalidArgument;
    }
    g->nSupers = static_cast<long long>(supers.size());
    g->nChunksU = chunks;
    g->nTilesU = static_cast<long long>(uTiles.size());
    g->nUnionRows = static_cast<long long>(uIdx.size());
    g->bfUnion = unionValues;
    XQ_TRY(dev_alloc(g, &g->d_supers, supers.size()));
    XQ_TRY(dev_alloc(g, &g->d_uTiles, uTiles.size()));
    XQ_TRY(dev_alloc(g, &g->d_uIdx, uIdx.size()));
    XQ_TRY(dev_alloc(g, &g->d_uChunkOffset, uChunkOffset.size()));
    XQ_TRY(dev_alloc(g, &g->d_uChunkBlock, uChunkBlock.size()));
    XQ_TRY(dev_alloc(g, &g->d_uRow, uRow.size()));
    hipStream_t st = g->stream;
    auto up = [&](auto *dst, const auto &v) {
        return v.empty() ? hipSuccess : hipMemcpyAsync(dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice, st);
    };
    XQ_CHECK(up(g->d_supers, supers));
    XQ_CHECK(up(g->d_uTiles, uTiles));
    XQ_CHECK(up(g->d_uIdx, uIdx));
    XQ_CHECK(up(g->d_uChunkOffset, uChunkOffset));
    XQ_CHECK(up(g->d_uChunkBlock, uChunkBlock));
    XQ_CHECK(up(g->d_uRow, uRow));
    // The host vectors must outlive the copies.
    XQ_CHECK(hipStreamSynchronize(st));
    return kOK;
}

int prepare_mixed(xq_grid *g, const double *f, const double *fx, const double *fy, const double *fz) {
    XQ_TRY(chunk_tables(g));
    XQ_TRY(union_tables(g));
    const long long chunks = g->nChunks, chunksU = g->nChunksU;
    const size_t nc = static_cast<size_t>(chunks > 0 ? chunks : 1), ncU = static_cast<size_t>(chunksU > 0 ? chunksU : 1);
    XQ_TRY(dev_alloc(g, &g->d_DA, ncU * kChunkBytes));
    XQ_TRY(dev_alloc(g, &g->d_EA, ncU * kC));
    XQ_TRY(dev_alloc(g, &g->d_DB, nc * kChunkBytesB));
    XQ_TRY(dev_alloc(g, &g->d_EB, static_cast<size_t>(g->nBlocks > 0 ? g->nBlocks : 1) * kBlockSize));
    XQ_TRY(dev_alloc(g, &g->d_GF, nc * kChunkGrad));
    XQ_TRY(dev_alloc(g, &g->d_EG, static_cast<size_t>(g->nBlocks > 0 ? g->nBlocks : 1) * 3 * kBlockSize));
    XQ_TRY(dev_alloc(g, &g->d_DW, ncU * kChunkBytes));
    XQ_TRY(dev_alloc(g, &g->d_GHU, ncU * kChunkGrad));
    XQ_TRY(dev_alloc(g, &g->d_GLU, ncU * kChunkGrad));
    XQ_TRY(dev_alloc(g, &g->d_EW, ncU * kC));
    XQ_TRY(dev_alloc(g, &g->d_Pq, static_cast<size_t>(g->nFunctions) * g->nFunctions));
    XQ_TRY(dev_alloc(g, &g->d_eP, static_cast<size_t>(g->nFunctions)));
    XQ_TRY(dev_alloc(g, &g->d_wcoef, 4 * static_cast<size_t>(g->nPoints > 0 ? g->nPoints : 1)));
    XQ_TRY(dev_alloc(g, &g->d_part, nc * 4 * kBlockSize));
    hipStream_t s = g->stream;
    if (g->nBlocks > 0)
        hipLaunchKernelGGL(point_exponent_kernel, dim3(g->nBlocks), dim3(kBlockSize), 0, s, g->d_blocks, f, fx, fy, fz,
                           g->d_EB, g->d_EG);
    if (chunks > 0)
        hipLaunchKernelGGL(density_operand_prepare_kernel, dim3(static_cast<unsigned>(chunks)), dim3(256), 0, s,
                           g->d_blocks, g->d_chunkOffset, g->d_chunkBlock, f, fx, fy, fz, g->d_EB, g->d_EG, g->d_DB,
                           g->d_GF);
    if (chunksU > 0)
        hipLaunchKernelGGL(fock_operand_prepare_kernel, dim3(static_cast<unsigned>(chunksU)), dim3(256), 0, s,
                           g->d_blocks, g->d_uChunkBlock, g->d_uRow, f, fx, fy, fz, g->d_EG, g->d_DA, g->d_EA,
                           g->d_GHU, g->d_GLU);
    XQ_CHECK(hipGetLastError());
    XQ_CHECK(hipStreamSynchronize(s));
    dev_free(g, &g->d_uRow);   // used by the set-up only
    return kOK;
}

// Per-step arrays of a second job on the handle (the beta spin of an unrestricted quadrature): P digits and exponents,
// W digits and exponents, the density's row-chunk shares.  Complete or released.
int alloc_step_mixed(xq_grid *g, StepArrays *a) {
    const size_t ncU = static_cast<size_t>(g->nChunksU > 0 ? g->nChunksU : 1), n = static_cast<size_t>(g->nFunctions);
    const size_t nc = static_cast<size_t>(g->nChunks > 0 ? g->nChunks : 1);
    int status = dev_alloc(g, &a->DW, ncU * kChunkBytes);
    if (status == kOK) status = dev_alloc(g, &a->EW, ncU * kC);
    if (status == kOK) status = dev_alloc(g, &a->part, nc * 4 * kBlockSize);
    if (status == kOK) status = dev_alloc(g, &a->Pq, n * n);
    if (status == kOK) status = dev_alloc(g, &a->eP, n);
    if (status != kOK) free_step_mixed(g, a);
    return status;
}

void free_step_mixed(xq_grid *g, StepArrays *a) {
    dev_free(g, &a->DW);
    dev_free(g, &a->EW);
    dev_free(g, &a->Pq);
    dev_free(g, &a->eP);
    dev_free(g, &a->part);
}

void free_mixed(xq_grid *g) {
    dev_free(g, &g->d_order);
    dev_free(g, &g->d_items);
    dev_free(g, &g->d_part);
    g->h_order.clear();
    g->h_items.clear();
    for (int x = 0; x <= 8; x++) g->xcdStart[x] = g->itemStart[x] = 0;
    dev_free(g, &g->d_chunkOffset);
    dev_free(g, &g->d_chunkBlock);
    dev_free(g, &g->d_tiles);
    dev_free(g, &g->d_supers);
    dev_free(g, &g->d_uTiles);
    dev_free(g, &g->d_uIdx);
    dev_free(g, &g->d_uChunkOffset);
    dev_free(g, &g->d_uChunkBlock);
    dev_free(g, &g->d_uRow);
    dev_free(g, &g->d_GHU);
    dev_free(g, &g->d_GLU);
    dev_free(g, &g->d_DA);
    dev_free(g, &g->d_EA);
    dev_free(g, &g->d_DB);
    dev_free(g, &g->d_EB);
    dev_free(g, &g->d_GF);
    dev_free(g, &g->d_EG);
    dev_free(g, &g->d_DW);
    dev_free(g, &g->d_EW);
    dev_free(g, &g->d_Pq);
    dev_free(g, &g->d_eP);
    dev_free(g, &g->d_wcoef);
    g->nChunks = 0;
    g->nTiles = 0;
    g->nSupers = 0;
    g->nChunksU = 0;
    g->nTilesU = 0;
    g->nUnionRows = 0;
    g->bfUnion = 0.0;
}

// The job table entries of the given handles (each with its functional) and each stage's total work-groups.
void fill_jobs(int nJobs, xq_grid *const *grids, const FunctionalSet *fs, MixedJob *h_table, long long *rowsOut,
               long long *blocksOut, long long *chunksOut, long long *groupsOut, const StepArrays *over,
               long long *blockSlotsOut, long long *itemSlotsOut) {
    long long rows = 0, blocks = 0, chunks = 0, groups = 0, slots[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long itemSlots[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nJobs; j++) {
        const xq_grid *g = grids[j];
        MixedJob &J = h_table[j];
        J.blocks = g->d_blocks;
        J.order = g->d_order;
        for (int x = 0; x <= 8; x++) J.xcdStart[x] = g->xcdStart[x];
        for (int x = 0; x < 8; x++) {
            J.xcdBase[x] = static_cast<int>(slots[x]);
            slots[x] += g->xcdStart[x + 1] - g->xcdStart[x];
        }
        J.items = g->d_items;
        J.part = g->d_part;
        for (int x = 0; x <= 8; x++) J.itemStart[x] = g->itemStart[x];
        for (int x = 0; x < 8; x++) {
            J.itemBase[x] = static_cast<int>(itemSlots[x]);
            itemSlots[x] += g->itemStart[x + 1] - g->itemStart[x];
        }
        J.chunkOffset = g->d_chunkOffset;
        J.chunkBlock = g->d_chunkBlock;
        J.idx = g->d_idx;
        J.tiles = g->d_tiles;
        J.DA = g->d_DA;
        J.EA = g->d_EA;
        J.DB = g->d_DB;
        J.EB = g->d_EB;
        J.EG = g->d_EG;
        J.GF = g->d_GF;
        J.P = g->d_P;
        J.w = g->d_w;
        J.eP = g->d_eP;
        J.Pq = g->d_Pq;
        J.wcoef = g->d_wcoef;
        J.rhoOut = g->d_rhoOut;
        J.blockEN = g->d_blockEN;
        J.V = g->d_V;
        J.skip = g->d_skip;
        J.DW = g->d_DW;
        J.EW = g->d_EW;
        J.supers = g->d_supers;
        J.uTiles = g->d_uTiles;
        J.uIdx = g->d_uIdx;
        J.uChunkOffset = g->d_uChunkOffset;
        J.uChunkBlock = g->d_uChunkBlock;
        J.GHU = g->d_GHU;
        J.GLU = g->d_GLU;
        J.FPB = g->d_FPB;
        J.GPB = g->d_GPB;
        J.FPA = g->d_FPA;
        J.WPA = g->d_WPA;
        J.fs = fs[j];
        J.rhoTol = g->rhoTol;
        J.n = g->nFunctions;
        J.nBlocks = g->nBlocks;
        // The per-chunk and per-tile stages run over union chunks and superblock tiles on the mixed path.
        const bool mixed = g->precision == XQ_PRECISION_MIXED;
        J.nChunks = static_cast<int>(mixed ? g->nChunksU : g->nChunks);
        J.nTiles = static_cast<int>(mixed ? g->nTilesU : g->nTiles);
        J.rowBase = static_cast<int>(rows);
        J.blockBase = static_cast<int>(blocks);
        J.chunkBase = static_cast<int>(chunks);
        J.groupBase = static_cast<int>(groups);
        if (over != nullptr) {
            const StepArrays &o = over[j];
            if (o.P) J.P = o.P;
            if (o.V) J.V = o.V;
            if (o.wcoef) J.wcoef = o.wcoef;
            if (o.rhoOut) J.rhoOut = o.rhoOut;
            if (o.WPA) J.WPA = o.WPA;
            if (o.Pq) J.Pq = o.Pq;
            if (o.eP) J.eP = o.eP;
            if (o.DW) J.DW = o.DW;
            if (o.EW) J.EW = o.EW;
            if (o.part) J.part = o.part;
        }
        rows += J.n;
        blocks += J.nBlocks;
        chunks += J.nChunks;
        groups += (J.nTiles + 3) / 4;
    }
    *rowsOut = rows;
    *blocksOut = blocks;
    if (blockSlotsOut != nullptr) *blockSlotsOut = *std::max_element(slots, slots + 8);
    if (itemSlotsOut != nullptr) *itemSlotsOut = *std::max_element(itemSlots, itemSlots + 8);
    *chunksOut = chunks;
    *groupsOut = groups;
}

// The four stage launches of the mixed path, each one launch over all the jobs.  h_table (pinned) and d_table hold at
// least nJobs entries; the caller synchronises the stream before h_table is written again.
int launch_mixed(hipStream_t s, const StageEvents &ev, int nJobs, xq_grid *const *grids, const FunctionalSet *fs,
                 MixedJob *h_table, MixedJob *d_table, int stages, const StepArrays *over) {
    long long rows, blocks, chunks, groups, itemSlots;
    fill_jobs(nJobs, grids, fs, h_table, &rows, &blocks, &chunks, &groups, over, nullptr, &itemSlots);
    if (rows > INT32_MAX / 2 || blocks > INT32_MAX / 2 || chunks > INT32_MAX / 2 || groups > INT32_MAX / 16 ||
        itemSlots > INT32_MAX / 16) {
        set_error("batch too large for one launch");
        return kInvalidArgument;
    }
    XQ_CHECK(hipMemcpyAsync(d_table, h_table, nJobs * sizeof(MixedJob), hipMemcpyHostToDevice, s));
    // The two spins of one handle (an unrestricted quadrature) run interleaved, item by item and tile group by tile
    // group.
    const int interleaved = nJobs == 2 && grids[0] == grids[1] ? 1 : 0;
    if ((stages & kStageDensity) && rows > 0)
        hipLaunchKernelGGL(p_slice_kernel, dim3(static_cast<unsigned>(rows)), dim3(256), 0, s, d_table, nJobs);
    XQ_CHECK(hipEventRecord(ev.evRho0, s));
    if ((stages & kStageDensity) && blocks > 0) {
        // The two spins of one handle (an unrestricted quadrature) run interleaved item by item.
        if (itemSlots > 0)
            hipLaunchKernelGGL(density_rows_kernel, dim3(static_cast<unsigned>(8 * itemSlots)), dim3(256), 0, s, d_table,
                               nJobs, interleaved);
        XQ_CHECK(hipEventRecord(ev.evDens1, s));
        if (h_table[0].rhoOut != nullptr) {   // the density stage of one spin (every job of the call)
            hipLaunchKernelGGL(density_functional_kernel<true>, dim3(static_cast<unsigned>(blocks)), dim3(kBlockSize), 0,
                               s, d_table, nJobs);
        } else {
            hipLaunchKernelGGL(density_functional_kernel<false>, dim3(static_cast<unsigned>(blocks)), dim3(kBlockSize),
                               0, s, d_table, nJobs);
        }
    } else {
        XQ_CHECK(hipEventRecord(ev.evDens1, s));
    }
    XQ_CHECK(hipEventRecord(ev.evRho1, s));
    if ((stages & kStageFock) && chunks > 0) {
        hipLaunchKernelGGL(w_digits_kernel, dim3(static_cast<unsigned>(chunks)), dim3(256), 0, s, d_table, nJobs);
        const int perXcd = static_cast<int>((groups + 7) / 8);
        hipLaunchKernelGGL(fock_kernel_mixed, dim3(static_cast<unsigned>(8 * perXcd)), dim3(256), 0, s, d_table, nJobs,
                           perXcd, static_cast<int>(groups), interleaved);
    }
    XQ_CHECK(hipEventRecord(ev.evFock1, s));
    XQ_CHECK(hipGetLastError());
    return kOK;
}

}  // namespace xq
