// Integrals of the SCF wrapper on the GPU: the one-electron matrices (overlap, kinetic energy, electron-nuclear attraction)
// and the density-fitted Coulomb term, whose electron-fit integrals (mu nu|P) stay resident in device memory.
//
// Counterparts in the reference (QCModelDFT's energy closures, QCModelDFT.py:130-160, through
// GaussianBasisIntegralEvaluator.pyx):
//   qs_one_electron  <- GaussianBasisContainerIntegrals_f1KOf1i (GaussianBasisContainerIntegrals_f1Xg1.c:233-296) and
//                       GaussianBasisContainerIntegrals_f2Cm1V (GaussianBasisContainerIntegrals_f2Cp1.c:104-173), point nuclei
//   qs_fit_create    <- GaussianBasisContainerIntegrals_f1Xg2i (GaussianBasisContainerIntegrals_f1Xg2.c:43-128), Coulomb
//                       operator; values of magnitude <= 1e-12 are dropped, as its block storage drops them (:42, :67)
//   qs_fit_matrix    <- f1Xf1i_f1Oi with constraints (GaussianBasisIntegralEvaluator.pyx:260-302): (P|Q) from
//                       GaussianBasisContainerIntegrals_f1Cf1i, the last row the fit functions' integrals (f1Oi)
//   qs_fit_b, qs_fit_fock <- the two contractions of Fock_MakeFromFitIntegralsCoulomb (FockConstruction.c:31-108):
//                       b_P = sum_mn P_mn (mn|P) and F_mn += sum_P c_P (mn|P); the bordered solve between them is the caller's
// The integrals themselves are the McMurchie-Davidson forms of qs_md.h (the reference uses Rys quadrature; the values agree
// to rounding).  Functions are the reference's: contraction coefficients unnormalised as GaussianBasis_UnnormalizePrimitives,
// Cartesian factors sqrt((2l-1)!! / ((2x-1)!! (2y-1)!! (2z-1)!!)) (GaussianBasis.c:448-501, 728-751), spherical functions by
// the c2s transformation of xq_c2s.h.
//
// Layout of the electron-fit integrals: the significant orbital shell pairs (A >= B, pair prefactor above 1e-15) each own
// nA x nB consecutive rows of a row-major [rows][fit functions] FP64 array; a row's packed orbital index (mu >= nu) is in
// rowPacked, -1 for the upper-triangle duplicates of a diagonal shell pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qcscf.h"
#include "qs_shells.h"
#include "xq_c2s.h"
#include "xq_internal.h"

using xq::check;
using xq::kInvalidArgument;
using xq::kOK;
using xq::kOutOfMemory;
using xq::set_error;

namespace qs {

// The Boys function table, one copy per GPU (filled once per device by the first basis created on it).
__device__ double g_boys[kBoysPoints * kBoysTable];
// The c2s transformations of l = 0..2 ([l][6][5]), one copy per GPU.
__device__ double g_c2s[3 * 6 * 5];

namespace {

constexpr int kThreads = 128;

std::mutex g_boysMutex;
bool g_boysReady[64] = {};

// The Boys table and the c2s transformations on the current GPU (once per device) from ensure_boys.
int ensure_boys() {
    int device = 0;
    XQ_CHECK(hipGetDevice(&device));
    if (device < 0 || device >= 64) { set_error("device index out of range"); return kInvalidArgument; }
    std::lock_guard<std::mutex> lock(g_boysMutex);
    if (g_boysReady[device]) return kOK;
    std::vector<double> t;
    boys_table(t);
    XQ_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_boys), t.data(), t.size() * sizeof(double)));
    double c2s[90];
    c2s_table(c2s);
    XQ_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_c2s), c2s, sizeof(c2s)));
    g_boysReady[device] = true;
    return kOK;
}


// ------------------------------------------------------------------------------------------------------------------------
// One-electron matrices: one thread per shell pair (A >= B).
// ------------------------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) one_electron_kernel(int nPairs, const QPair *__restrict__ pairs,
                                                                const QShell *__restrict__ shells,
                                                                const double *__restrict__ exps,
                                                                const double *__restrict__ coefs, int nAtoms,
                                                                const double *__restrict__ atoms, int spherical,
                                                                double *__restrict__ S,
                                                                double *__restrict__ T, double *__restrict__ V) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nPairs) one_electron_pair(k, pairs, shells, exps, coefs, nAtoms, atoms, spherical, g_boys, g_c2s, S, T, V);
}

// ------------------------------------------------------------------------------------------------------------------------
// Fit matrix: (C|D) for fit shell pairs C >= D, one thread each; the fit functions' integrals in the last row.
// ------------------------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) fit_matrix_kernel(int nPairs, const QPair *__restrict__ pairs,
                                                              const QShell *__restrict__ shells,
                                                              const double *__restrict__ exps,
                                                              const double *__restrict__ coefs, int nFit, int spherical,
                                                              double *__restrict__ M) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nPairs) fit_matrix_pair(k, pairs, shells, exps, coefs, nFit, spherical, g_boys, g_c2s, M);
}

// ------------------------------------------------------------------------------------------------------------------------
// Electron-fit integrals: one thread per (significant orbital shell pair, fit shell).
// ------------------------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) fit_integrals_kernel(
    long long k0, long long nWork, int nFitShells, const PairEntry *__restrict__ pairs, const QShell *__restrict__ oShells,
    const double *__restrict__ oExps, const double *__restrict__ oCoefs, int oSpherical,
    const QShell *__restrict__ fShells, const double *__restrict__ fExps, const double *__restrict__ fCoefs, int fSpherical,
    int nFit, double *__restrict__ B) {
    const long long k = k0 + blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (k < nWork)
        fit_integrals_item(k, nFitShells, pairs, oShells, oExps, oCoefs, oSpherical, fShells, fExps, fCoefs, fSpherical, nFit,
                           g_boys, g_c2s, B);
}

// ------------------------------------------------------------------------------------------------------------------------
// Per SCF iteration: b = sum_rows w_row P_row B_row (partial sums over row chunks, then a fixed-order reduction) and
// F_row += B_row . c (one wave per row).  w_row = 2 off the diagonal, 1 on it, 0 for a duplicate row.
// ------------------------------------------------------------------------------------------------------------------------
constexpr int kRowChunk = 1024;

__global__ void __launch_bounds__(256) fit_b_partial_kernel(long long rows, int nFit, const double *__restrict__ B,
                                                           const long long *__restrict__ rowPacked,
                                                           const unsigned char *__restrict__ rowDiag,
                                                           const double *__restrict__ P, double *__restrict__ partial) {
    const int col = blockIdx.x * 256 + threadIdx.x;
    const long long r0 = static_cast<long long>(blockIdx.y) * kRowChunk;
    const long long r1 = r0 + kRowChunk < rows ? r0 + kRowChunk : rows;
    double acc = 0.0;
    if (col < nFit)
        for (long long r = r0; r < r1; r++) {
            const long long o = rowPacked[r];
            if (o < 0) continue;
            acc += (rowDiag[r] ? 1.0 : 2.0) * P[o] * B[r * nFit + col];
        }
    if (col < nFit) partial[static_cast<size_t>(blockIdx.y) * nFit + col] = acc;
}

__global__ void __launch_bounds__(256) fit_b_reduce_kernel(int nChunks, int nFit, const double *__restrict__ partial,
                                                          double *__restrict__ b) {
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= nFit) return;
    double acc = 0.0;
    for (int k = 0; k < nChunks; k++) acc += partial[static_cast<size_t>(k) * nFit + col];
    b[col] = acc;
}

__global__ void __launch_bounds__(256) fit_fock_kernel(long long rows, int nFit, const double *__restrict__ B,
                                                      const long long *__restrict__ rowPacked,
                                                      const double *__restrict__ c, double *__restrict__ F) {
    const long long r = (blockIdx.x * 256LL + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    const long long o = rowPacked[r];
    if (o < 0) return;                  // wave-uniform: the whole wave leaves together
    const double *row = B + r * nFit;
    double acc = 0.0;
    for (int k = lane; k < nFit; k += 64) acc += row[k] * c[k];
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) acc += __shfl_xor(acc, s, 64);
    if (lane == 0) F[o] += acc;
}

}  // namespace

}  // namespace qs

// ------------------------------------------------------------------------------------------------------------------------
// Handles.
// ------------------------------------------------------------------------------------------------------------------------
struct qs_basis : qs::QBasis {
    qs::QShell *d_shells = nullptr;
    double *d_exps = nullptr, *d_coefs = nullptr;
    hipStream_t stream = nullptr;      // the one-electron kernel's stream
};

struct qs_fit {
    int nFit = 0, nOrbital = 0;
    long long rows = 0, nPairs = 0;
    std::vector<qs::QPair> fitPairs;
    const qs_basis *orbital = nullptr, *fit = nullptr;
    double *d_B = nullptr;
    long long *d_rowPacked = nullptr;
    unsigned char *d_rowDiag = nullptr;
    double *d_P = nullptr, *d_F = nullptr, *d_b = nullptr, *d_c = nullptr, *d_partial = nullptr;
    hipStream_t stream = nullptr;
};

namespace qs {
namespace {

// Device arrays come from the quadrature library's caching allocator: a request loop sets up and releases molecules
// while other jobs compute, and hipFree would synchronise the whole device.
template <typename T>
int qs_alloc(T **p, size_t n) {
    *p = nullptr;
    if (xq::cache_alloc(reinterpret_cast<void **>(p), (n > 0 ? n : 1) * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        set_error("hipMalloc failed (" + std::to_string((n > 0 ? n : 1) * sizeof(T)) + " bytes)");
        return kOutOfMemory;
    }
    return kOK;
}

template <typename T>
void qs_release(T **p) {
    if (*p != nullptr) { (void)xq::cache_free(*p); *p = nullptr; }
}

// The integral kernels hold their McMurchie-Davidson tables in private memory (about 8 KB per lane for the fit
// integrals), and the runtime gives every hardware queue that runs one its own scratch area sized for the whole
// device.  Two such kernels in flight at once on two queues (set-up threads of a request loop) made the second queue's
// scratch allocation fail (HSA_STATUS_ERROR_OUT_OF_RESOURCES, queue abort), so at most one of them runs at a time in
// the process: each holds this lock from its launch until its stream has drained.  The same allocation fails when the
// device is nearly full (a request loop's caching allocator holds idle blocks), so each launch first makes sure that
// kScratchHeadroom bytes are free: the fit-integral kernel's 8,368 bytes per lane at 3 waves per SIMD on 256 CUs take
// about 1.6 GB.
std::mutex g_integralKernels;
constexpr size_t kScratchHeadroom = static_cast<size_t>(6) << 30;
// The runtime's scratch areas belong to hardware queues and come from a pool of fixed size, separate from the free
// device memory: several queues that each ran one of these kernels held areas sized for a whole-device dispatch and a
// further request failed with HSA_STATUS_ERROR_OUT_OF_RESOURCES with 217 GB free.  So every launch of them goes to one
// stream per device (one queue, one area), and no launch is larger than kSliceThreads threads, which bounds the area a
// dispatch asks for (about 1.8 GB at 13,760 bytes per lane).
constexpr int kMaxDevices = 64;
hipStream_t g_integralStream[kMaxDevices] = {};
constexpr long long kSliceThreads = 2048LL * 64;

// The integral stream of the current device (created on first use); the caller holds g_integralKernels.
int integral_stream(hipStream_t *out) {
    int dev = 0;
    XQ_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) { set_error("device index out of range"); return kInvalidArgument; }
    if (g_integralStream[dev] == nullptr)
        XQ_CHECK(hipStreamCreateWithFlags(&g_integralStream[dev], hipStreamNonBlocking));
    *out = g_integralStream[dev];
    return kOK;
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

int qs_basis_create(int nAtoms, const int *Z, const double *xyzBohr, int nSpec, const double *spec, int spherical,
                    qs_basis **basis) {
    if (basis == nullptr || nAtoms <= 0 || Z == nullptr || xyzBohr == nullptr || spec == nullptr || nSpec <= 0) {
        set_error("invalid basis arguments");
        return kInvalidArgument;
    }
    *basis = nullptr;
    XQ_TRY(ensure_boys());
    qs_basis *b = new qs_basis();
    b->nAtoms = nAtoms;
    b->spherical = spherical ? 1 : 0;
    b->Z.assign(Z, Z + nAtoms);
    b->xyz.assign(xyzBohr, xyzBohr + 3 * nAtoms);
    const std::string err = build_shells(b, nSpec, spec);
    int status = kOK;
    if (!err.empty()) { set_error(err); status = kInvalidArgument; }
    if (status == kOK) status = check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (status == kOK) status = qs_alloc(&b->d_shells, b->shells.size());
    if (status == kOK) status = qs_alloc(&b->d_exps, b->exps.size());
    if (status == kOK) status = qs_alloc(&b->d_coefs, b->coefs.size());
    if (status == kOK) status = check(hipMemcpyAsync(b->d_shells, b->shells.data(), b->shells.size() * sizeof(QShell),
                                                     hipMemcpyHostToDevice, b->stream), "copy shells");
    if (status == kOK) status = check(hipMemcpyAsync(b->d_exps, b->exps.data(), b->exps.size() * sizeof(double),
                                                     hipMemcpyHostToDevice, b->stream), "copy exponents");
    if (status == kOK) status = check(hipMemcpyAsync(b->d_coefs, b->coefs.data(), b->coefs.size() * sizeof(double),
                                                     hipMemcpyHostToDevice, b->stream), "copy coefficients");
    if (status == kOK) status = check(hipStreamSynchronize(b->stream), "copy basis");
    if (status != kOK) { qs_basis_free(b); return status; }
    *basis = b;
    return kOK;
}

int qs_basis_size(const qs_basis *basis, int *nFunctions) {
    if (basis == nullptr || nFunctions == nullptr) { set_error("invalid arguments"); return kInvalidArgument; }
    *nFunctions = basis->nFunctions;
    return kOK;
}

void qs_basis_free(qs_basis *basis) {
    if (basis == nullptr) return;
    qs_release(&basis->d_shells);
    qs_release(&basis->d_exps);
    qs_release(&basis->d_coefs);
    if (basis->stream) (void)hipStreamDestroy(basis->stream);
    delete basis;
}

int qs_one_electron(qs_basis *basis, double *S, double *T, double *V) {
    if (basis == nullptr || S == nullptr || T == nullptr || V == nullptr) {
        set_error("invalid one-electron arguments");
        return kInvalidArgument;
    }
    const size_t n = static_cast<size_t>(basis->nFunctions), size = n * (n + 1) / 2;
    std::vector<QPair> pairs = shell_pairs(basis, false);
    // Pairs of one class (angular momenta and primitive counts) side by side, so that a wave's pairs share their loop
    // bounds; each pair writes its own entries of S, T and V.
    auto cls = [&](const QPair &q) {
        const QShell &a = basis->shells[q.x], &b = basis->shells[q.y];
        return ((a.l * 8 + b.l) * 64 + a.nPrim) * 64 + b.nPrim;
    };
    std::stable_sort(pairs.begin(), pairs.end(), [&](const QPair &u, const QPair &v) { return cls(u) < cls(v); });
    std::vector<double> atoms(4 * static_cast<size_t>(basis->nAtoms));
    for (int a = 0; a < basis->nAtoms; a++) {
        for (int c = 0; c < 3; c++) atoms[4 * a + c] = basis->xyz[3 * a + c];
        atoms[4 * a + 3] = static_cast<double>(basis->Z[a]);
    }
    QPair *d_pairs = nullptr;
    double *d_atoms = nullptr, *d_S = nullptr, *d_T = nullptr, *d_V = nullptr;
    int status = qs_alloc(&d_pairs, pairs.size());
    if (status == kOK) status = qs_alloc(&d_atoms, atoms.size());
    if (status == kOK) status = qs_alloc(&d_S, size);
    if (status == kOK) status = qs_alloc(&d_T, size);
    if (status == kOK) status = qs_alloc(&d_V, size);
    {
        std::lock_guard<std::mutex> lock(g_integralKernels);
        hipStream_t s = nullptr;
        if (status == kOK) status = integral_stream(&s);
        if (status == kOK) status = check(hipMemcpyAsync(d_pairs, pairs.data(), pairs.size() * sizeof(QPair),
                                                         hipMemcpyHostToDevice, s), "copy pairs");
        if (status == kOK) status = check(hipMemcpyAsync(d_atoms, atoms.data(), atoms.size() * sizeof(double),
                                                         hipMemcpyHostToDevice, s), "copy atoms");
        if (status == kOK && !pairs.empty()) status = xq::ensure_free_device_memory(kScratchHeadroom, "one-electron integrals");
        for (long long k0 = 0; status == kOK && k0 < static_cast<long long>(pairs.size()); k0 += kSliceThreads) {
            const long long m = std::min<long long>(kSliceThreads, static_cast<long long>(pairs.size()) - k0);
            hipLaunchKernelGGL(one_electron_kernel, dim3(static_cast<unsigned>((m + kThreads - 1) / kThreads)),
                               dim3(kThreads), 0, s, static_cast<int>(m), d_pairs + k0, basis->d_shells, basis->d_exps,
                               basis->d_coefs, basis->nAtoms, d_atoms, basis->spherical, d_S, d_T, d_V);
            status = check(hipGetLastError(), "one_electron_kernel");
        }
        if (status == kOK) status = check(hipMemcpyAsync(S, d_S, size * sizeof(double), hipMemcpyDeviceToHost, s), "copy S");
        if (status == kOK) status = check(hipMemcpyAsync(T, d_T, size * sizeof(double), hipMemcpyDeviceToHost, s), "copy T");
        if (status == kOK) status = check(hipMemcpyAsync(V, d_V, size * sizeof(double), hipMemcpyDeviceToHost, s), "copy V");
        if (s != nullptr) {
            const int drained = check(hipStreamSynchronize(s), "one-electron integrals");
            if (status == kOK) status = drained;
        }
    }
    qs_release(&d_pairs);
    qs_release(&d_atoms);
    qs_release(&d_S);
    qs_release(&d_T);
    qs_release(&d_V);
    return status;
}

void qs_fit_free(qs_fit *fit) {
    if (fit == nullptr) return;
    if (fit->stream) (void)hipStreamSynchronize(fit->stream);
    qs_release(&fit->d_B);
    qs_release(&fit->d_rowPacked);
    qs_release(&fit->d_rowDiag);
    qs_release(&fit->d_P);
    qs_release(&fit->d_F);
    qs_release(&fit->d_b);
    qs_release(&fit->d_c);
    qs_release(&fit->d_partial);
    if (fit->stream) (void)hipStreamDestroy(fit->stream);
    delete fit;
}

int qs_fit_create(const qs_basis *orbital, const qs_basis *fitBasis, qs_fit **out) {
    if (orbital == nullptr || fitBasis == nullptr || out == nullptr) { set_error("invalid fit arguments"); return kInvalidArgument; }
    *out = nullptr;
    qs_fit *F = new qs_fit();
    F->orbital = orbital;
    F->fit = fitBasis;
    F->nFit = fitBasis->nFunctions;
    F->nOrbital = orbital->nFunctions;
    int status = check(hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking), "hipStreamCreate");
    // Significant shell pairs and their rows.
    std::vector<QPair> sp = status == kOK ? shell_pairs(orbital, true) : std::vector<QPair>();
    std::vector<PairEntry> pairs(sp.size());
    std::vector<long long> rowPacked;
    std::vector<unsigned char> rowDiag;
    long long rows = 0;
    for (size_t k = 0; k < sp.size(); k++) {
        const QShell &A = orbital->shells[sp[k].x], &B = orbital->shells[sp[k].y];
        pairs[k] = PairEntry{sp[k].x, sp[k].y, rows};
        for (int i = 0; i < A.nf; i++)
            for (int j = 0; j < B.nf; j++) {
                const int mu = A.f0 + i, nu = B.f0 + j;
                rowPacked.push_back(mu >= nu ? packed_index(mu, nu) : -1);
                rowDiag.push_back(mu == nu ? 1 : 0);
            }
        rows += static_cast<long long>(A.nf) * B.nf;
    }
    F->rows = rows;
    F->nPairs = static_cast<long long>(sp.size());
    const size_t nOrbPacked = static_cast<size_t>(F->nOrbital) * (F->nOrbital + 1) / 2;
    const long long nChunks = (rows + kRowChunk - 1) / kRowChunk;
    PairEntry *d_pairs = nullptr;
    // The fit shells in order of angular momentum for the integral launch: the consecutive work items of a wave take
    // one shell pair and consecutive fit shells, which then share their loop bounds (each shell writes its own
    // functions, f0 .. f0 + nf - 1, wherever it sits in this list).
    std::vector<QShell> fitSorted(fitBasis->shells);
    std::stable_sort(fitSorted.begin(), fitSorted.end(), [](const QShell &u, const QShell &v) { return u.l < v.l; });
    QShell *d_fitSorted = nullptr;
    if (status == kOK) status = qs_alloc(&F->d_B, static_cast<size_t>(rows) * F->nFit);
    if (status == kOK) status = qs_alloc(&F->d_rowPacked, rowPacked.size());
    if (status == kOK) status = qs_alloc(&F->d_rowDiag, rowDiag.size());
    if (status == kOK) status = qs_alloc(&F->d_P, nOrbPacked);
    if (status == kOK) status = qs_alloc(&F->d_F, nOrbPacked);
    if (status == kOK) status = qs_alloc(&F->d_b, F->nFit);
    if (status == kOK) status = qs_alloc(&F->d_c, F->nFit);
    if (status == kOK) status = qs_alloc(&F->d_partial, static_cast<size_t>(nChunks > 0 ? nChunks : 1) * F->nFit);
    if (status == kOK) status = qs_alloc(&d_pairs, pairs.size());
    if (status == kOK) status = qs_alloc(&d_fitSorted, fitSorted.empty() ? 1 : fitSorted.size());
    if (status == kOK && !fitSorted.empty())
        status = check(hipMemcpy(d_fitSorted, fitSorted.data(), fitSorted.size() * sizeof(QShell), hipMemcpyHostToDevice),
                       "copy fit shells");
    if (status == kOK && rows > 0) {
        status = check(hipMemcpy(F->d_rowPacked, rowPacked.data(), rowPacked.size() * sizeof(long long),
                                 hipMemcpyHostToDevice), "copy rows");
        if (status == kOK) status = check(hipMemcpy(F->d_rowDiag, rowDiag.data(), rowDiag.size(), hipMemcpyHostToDevice),
                                          "copy rows");
        if (status == kOK) status = check(hipMemcpy(d_pairs, pairs.data(), pairs.size() * sizeof(PairEntry),
                                                    hipMemcpyHostToDevice), "copy pairs");
        const long long nWork = F->nPairs * static_cast<long long>(fitBasis->shells.size());
        std::lock_guard<std::mutex> lock(g_integralKernels);
        hipStream_t is = nullptr;
        if (status == kOK) status = integral_stream(&is);
        if (status == kOK && nWork > 0) status = xq::ensure_free_device_memory(kScratchHeadroom, "fit integrals");
        for (long long k0 = 0; status == kOK && k0 < nWork; k0 += kSliceThreads) {
            const long long m = std::min<long long>(kSliceThreads, nWork - k0);
            hipLaunchKernelGGL(fit_integrals_kernel, dim3(static_cast<unsigned>((m + kThreads - 1) / kThreads)),
                               dim3(kThreads), 0, is, k0, nWork, static_cast<int>(fitBasis->shells.size()), d_pairs,
                               orbital->d_shells, orbital->d_exps, orbital->d_coefs, orbital->spherical,
                               d_fitSorted, fitBasis->d_exps, fitBasis->d_coefs, fitBasis->spherical, F->nFit,
                               F->d_B);
            status = check(hipGetLastError(), "fit_integrals_kernel");
        }
        if (is != nullptr) {
            const int drained = check(hipStreamSynchronize(is), "fit integrals");
            if (status == kOK) status = drained;
        }
    }
    qs_release(&d_pairs);
    qs_release(&d_fitSorted);
    if (status == kOK) F->fitPairs = shell_pairs(fitBasis, false);
    if (status != kOK) { qs_fit_free(F); return status; }
    *out = F;
    return kOK;
}

int qs_fit_matrix(qs_fit *fit, double *M) {
    if (fit == nullptr || M == nullptr) { set_error("invalid fit matrix arguments"); return kInvalidArgument; }
    const size_t n = static_cast<size_t>(fit->nFit) + 1, size = n * (n + 1) / 2;
    double *d_M = nullptr;
    QPair *d_pairs = nullptr;
    int status = qs_alloc(&d_M, size);
    if (status == kOK) status = qs_alloc(&d_pairs, fit->fitPairs.size());
    if (status == kOK) status = check(hipMemcpy(d_pairs, fit->fitPairs.data(), fit->fitPairs.size() * sizeof(QPair),
                                                hipMemcpyHostToDevice), "copy fit pairs");
    {
        std::lock_guard<std::mutex> lock(g_integralKernels);
        hipStream_t is = nullptr;
        if (status == kOK) status = integral_stream(&is);
        if (status == kOK) status = check(hipMemsetAsync(d_M, 0, size * sizeof(double), is), "clear M");
        if (status == kOK && !fit->fitPairs.empty()) status = xq::ensure_free_device_memory(kScratchHeadroom, "fit matrix");
        const long long nPairs = static_cast<long long>(fit->fitPairs.size());
        for (long long k0 = 0; status == kOK && k0 < nPairs; k0 += kSliceThreads) {
            const long long m = std::min<long long>(kSliceThreads, nPairs - k0);
            hipLaunchKernelGGL(fit_matrix_kernel, dim3(static_cast<unsigned>((m + kThreads - 1) / kThreads)),
                               dim3(kThreads), 0, is, static_cast<int>(m), d_pairs + k0, fit->fit->d_shells,
                               fit->fit->d_exps, fit->fit->d_coefs, fit->nFit, fit->fit->spherical, d_M);
            status = check(hipGetLastError(), "fit_matrix_kernel");
        }
        if (is != nullptr) {
            const int drained = check(hipStreamSynchronize(is), "fit matrix");
            if (status == kOK) status = drained;
        }
    }
    if (status == kOK) status = check(hipMemcpy(M, d_M, size * sizeof(double), hipMemcpyDeviceToHost), "copy M");
    qs_release(&d_M);
    qs_release(&d_pairs);
    return status;
}

int qs_fit_b(qs_fit *fit, const double *P, double *b) {
    if (fit == nullptr || P == nullptr || b == nullptr) { set_error("invalid fit arguments"); return kInvalidArgument; }
    const size_t nPacked = static_cast<size_t>(fit->nOrbital) * (fit->nOrbital + 1) / 2;
    const int nChunks = static_cast<int>((fit->rows + kRowChunk - 1) / kRowChunk);
    hipStream_t s = fit->stream;
    XQ_CHECK(hipMemcpyAsync(fit->d_P, P, nPacked * sizeof(double), hipMemcpyHostToDevice, s));
    if (nChunks > 0) {
        hipLaunchKernelGGL(fit_b_partial_kernel, dim3((fit->nFit + 255) / 256, nChunks), dim3(256), 0, s, fit->rows,
                           fit->nFit, fit->d_B, fit->d_rowPacked, fit->d_rowDiag, fit->d_P, fit->d_partial);
        hipLaunchKernelGGL(fit_b_reduce_kernel, dim3((fit->nFit + 255) / 256), dim3(256), 0, s, nChunks, fit->nFit,
                           fit->d_partial, fit->d_b);
        XQ_CHECK(hipGetLastError());
        XQ_CHECK(hipMemcpyAsync(b, fit->d_b, fit->nFit * sizeof(double), hipMemcpyDeviceToHost, s));
    } else {
        for (int k = 0; k < fit->nFit; k++) b[k] = 0.0;
    }
    XQ_CHECK(hipStreamSynchronize(s));
    return kOK;
}

int qs_fit_fock(qs_fit *fit, const double *c, double *Fock) {
    if (fit == nullptr || c == nullptr || Fock == nullptr) { set_error("invalid fit arguments"); return kInvalidArgument; }
    const size_t nPacked = static_cast<size_t>(fit->nOrbital) * (fit->nOrbital + 1) / 2;
    hipStream_t s = fit->stream;
    XQ_CHECK(hipMemcpyAsync(fit->d_c, c, fit->nFit * sizeof(double), hipMemcpyHostToDevice, s));
    XQ_CHECK(hipMemcpyAsync(fit->d_F, Fock, nPacked * sizeof(double), hipMemcpyHostToDevice, s));
    if (fit->rows > 0) {
        const long long waves = fit->rows;
        hipLaunchKernelGGL(fit_fock_kernel, dim3(static_cast<unsigned>((waves * 64 + 255) / 256)), dim3(256), 0, s,
                           fit->rows, fit->nFit, fit->d_B, fit->d_rowPacked, fit->d_c, fit->d_F);
        XQ_CHECK(hipGetLastError());
    }
    XQ_CHECK(hipMemcpyAsync(Fock, fit->d_F, nPacked * sizeof(double), hipMemcpyDeviceToHost, s));
    XQ_CHECK(hipStreamSynchronize(s));
    return kOK;
}

int qs_fit_estimate_bytes(int nAtoms, const int *Z, const double *xyzBohr, int nSpec, const double *spec, int spherical,
                          int nFitSpec, const double *fitSpec, int fitSpherical, long long *bytes) {
    if (bytes == nullptr || nAtoms <= 0 || Z == nullptr || xyzBohr == nullptr || spec == nullptr || nSpec <= 0 ||
        fitSpec == nullptr || nFitSpec <= 0) {
        set_error("invalid fit estimate arguments");
        return kInvalidArgument;
    }
    *bytes = 0;
    QBasis orb, fit;
    for (QBasis *b : {&orb, &fit}) {
        b->nAtoms = nAtoms;
        b->Z.assign(Z, Z + nAtoms);
        b->xyz.assign(xyzBohr, xyzBohr + 3 * nAtoms);
    }
    orb.spherical = spherical ? 1 : 0;
    fit.spherical = fitSpherical ? 1 : 0;
    std::string err = build_shells(&orb, nSpec, spec);
    if (err.empty()) err = build_shells(&fit, nFitSpec, fitSpec);
    if (!err.empty()) { set_error(err); return kInvalidArgument; }
    // The arrays qs_fit_create allocates: the integrals (rows x nFit), the row maps, P, F, b, c and the partial sums of b.
    long long rows = 0;
    for (const QPair &pq : shell_pairs(&orb, true))
        rows += static_cast<long long>(orb.shells[pq.x].nf) * orb.shells[pq.y].nf;
    const long long nFit = fit.nFunctions, n = orb.nFunctions, nPacked = n * (n + 1) / 2;
    const long long nChunks = (rows + kRowChunk - 1) / kRowChunk;
    *bytes = 8 * rows * nFit + 9 * rows + 16 * nPacked + 16 * nFit + 8 * (nChunks > 0 ? nChunks : 1) * nFit;
    return kOK;
}

int qs_fit_stats(const qs_fit *fit, long long *rows, long long *pairs, long long *deviceBytes) {
    if (fit == nullptr) { set_error("invalid arguments"); return kInvalidArgument; }
    if (rows) *rows = fit->rows;
    if (pairs) *pairs = fit->nPairs;
    if (deviceBytes) *deviceBytes = fit->rows * static_cast<long long>(fit->nFit) * 8;
    return kOK;
}

}  // extern "C"
