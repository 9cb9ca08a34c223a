// Dense steps of the SCF on the device: the orbitals and density of a Fock matrix, and the DIIS error vector.
//
// One handle per SCF holds the orthogonaliser X and the overlap S in device memory, a rocBLAS handle on a stream of its
// own, and the work matrices.  Each call moves a packed matrix in (expanded on the device), runs its GEMMs on rocBLAS and
// its eigenproblem on rocSOLVER (dsyevd), and moves the packed or dense result out; it returns when the result is on the
// host.  Calls on one handle must not overlap; handles on one device may be used from different threads at once.
//
// Storage.  rocBLAS works on column-major matrices; a row-major n x n array is the column-major array of the transpose,
// which is the same matrix for the symmetric F, P, S and X.  Products are formed in column-major terms and the one
// non-symmetric result, the error vector, is transposed back (below).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <string>

#include "../../include/qcscf.h"
#include "xq_internal.h"

struct qs_dense {
    int n = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    double *X = nullptr, *S = nullptr;        // orthogonaliser and overlap, n x n
    double *A = nullptr, *B = nullptr, *C = nullptr, *D = nullptr;   // work matrices, n x n
    double *packed = nullptr;                 // n (n + 1) / 2
    double *w = nullptr, *e = nullptr;        // eigenvalues, dsyevd's off-diagonal work
    rocblas_int *info = nullptr;
};

namespace {

int blas_check(rocblas_status st, const char *what) {
    if (st == rocblas_status_success) return xq::kOK;
    xq::set_error(std::string(what) + ": " + rocblas_status_to_string(st));
    return st == rocblas_status_memory_error ? xq::kOutOfMemory : xq::kAlgorithmError;
}

#define QS_BLAS(call, what)                                       \
    do {                                                          \
        const int st_ = blas_check((call), what);                 \
        if (st_ != xq::kOK) return st_;                           \
    } while (0)

void release(qs_dense *h) {
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->blas) rocblas_destroy_handle(h->blas);
    void *arrays[] = {h->X, h->S, h->A, h->B, h->C, h->D, h->packed, h->w, h->e, h->info};
    for (void *p : arrays) xq::cache_free(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

// kInvalidArgument unless the handle's device is current (its stream and rocBLAS handle belong to that device).
int check_device(const qs_dense *h) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != h->device) {
        xq::set_error("qs_dense: the device the handle was created on is not current");
        return xq::kInvalidArgument;
    }
    return xq::kOK;
}

// The packed lower triangle of a host matrix into the device dense matrix M (row-major, both triangles).
int upload_packed(qs_dense *h, const double *packedHost, double *M) {
    const size_t np = static_cast<size_t>(h->n) * (h->n + 1) / 2;
    XQ_CHECK(hipMemcpyAsync(h->packed, packedHost, np * sizeof(double), hipMemcpyHostToDevice, h->stream));
    xq::launch_expand(h->n, h->packed, M, h->stream);
    XQ_CHECK(hipGetLastError());
    return xq::kOK;
}

int aufbau(qs_dense *h, const double *fockF, int nOccupied, double *densityP, double *orbitalEnergies) {
    const int n = h->n;
    const double one = 1.0, zero = 0.0, two = 2.0;
    XQ_TRY(upload_packed(h, fockF, h->A));
    // B = F X, C = X B = X^T F X (X symmetric).
    QS_BLAS(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, h->A, n, h->X, n, &zero,
                          h->B, n), "dgemm F X");
    QS_BLAS(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, h->X, n, h->B, n, &zero,
                          h->C, n), "dgemm X F X");
    // Eigenvectors of C in its columns (lower triangle read), eigenvalues ascending.
    QS_BLAS(rocsolver_dsyevd(h->blas, rocblas_evect_original, rocblas_fill_lower, n, h->C, n, h->w, h->e, h->info),
            "rocsolver_dsyevd");
    rocblas_int info = 0;
    XQ_CHECK(hipMemcpyAsync(&info, h->info, sizeof(rocblas_int), hipMemcpyDeviceToHost, h->stream));
    XQ_CHECK(hipStreamSynchronize(h->stream));
    if (info != 0) {
        xq::set_error("rocsolver_dsyevd did not converge (info " + std::to_string(info) + ")");
        return xq::kAlgorithmError;
    }
    // D = X C_occ (n x nOccupied), then the upper triangle (column-major) of A = 2 D D^T, which is the lower triangle of
    // the row-major array that the packing kernel reads.
    if (nOccupied > 0) {
        QS_BLAS(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, n, nOccupied, n, &one, h->X, n,
                              h->C, n, &zero, h->D, n), "dgemm X C");
        QS_BLAS(rocblas_dsyrk(h->blas, rocblas_fill_upper, rocblas_operation_none, n, nOccupied, &two, h->D, n, &zero,
                              h->A, n), "dsyrk");
    } else {
        XQ_CHECK(hipMemsetAsync(h->A, 0, static_cast<size_t>(n) * n * sizeof(double), h->stream));
    }
    xq::launch_pack(n, h->A, h->packed, h->stream);
    XQ_CHECK(hipGetLastError());
    const size_t np = static_cast<size_t>(n) * (n + 1) / 2;
    XQ_CHECK(hipMemcpyAsync(densityP, h->packed, np * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (orbitalEnergies != nullptr)
        XQ_CHECK(hipMemcpyAsync(orbitalEnergies, h->w, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    XQ_CHECK(hipStreamSynchronize(h->stream));
    return xq::kOK;
}

// DIISSCFConverger.DIISCurrentFrame's error vector (DIISSCFConverger.py:291-303): e = X^T (F D S - S D F) X, returned as
// a dense row-major n x n array.  In column-major terms the device forms G = F D S, then X (G^T - G) X, whose row-major
// reading is X (G - G^T) X.
int commutator(qs_dense *h, const double *fockF, const double *densityP, double *error) {
    const int n = h->n;
    const double one = 1.0, zero = 0.0, minusOne = -1.0;
    XQ_TRY(upload_packed(h, fockF, h->A));
    XQ_TRY(upload_packed(h, densityP, h->B));
    // C = F D, D = C S = F D S.
    QS_BLAS(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, h->A, n, h->B, n, &zero,
                          h->C, n), "dgemm F D");
    QS_BLAS(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, h->C, n, h->S, n, &zero,
                          h->D, n), "dgemm F D S");
    // A = D^T - D, then C = A X and B = X C.
    QS_BLAS(rocblas_dgeam(h->blas, rocblas_operation_transpose, rocblas_operation_none, n, n, &one, h->D, n, &minusOne,
                          h->D, n, h->A, n), "dgeam");
    QS_BLAS(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, h->A, n, h->X, n, &zero,
                          h->C, n), "dgemm E X");
    QS_BLAS(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, h->X, n, h->C, n, &zero,
                          h->B, n), "dgemm X E X");
    XQ_CHECK(hipMemcpyAsync(error, h->B, static_cast<size_t>(n) * n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    XQ_CHECK(hipStreamSynchronize(h->stream));
    return xq::kOK;
}

}  // namespace

extern "C" {

int qs_dense_create(int n, const double *orthogonalizer, const double *overlapS, qs_dense **out) {
    if (out == nullptr || n <= 0 || orthogonalizer == nullptr || overlapS == nullptr) {
        xq::set_error("qs_dense_create: invalid argument");
        return xq::kInvalidArgument;
    }
    *out = nullptr;
    qs_dense *h = new qs_dense();
    h->n = n;
    int status = xq::kOK;
    do {
        if ((status = xq::check(hipGetDevice(&h->device), "hipGetDevice")) != xq::kOK) break;
        if ((status = xq::check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate")) != xq::kOK)
            break;
        if ((status = blas_check(rocblas_create_handle(&h->blas), "rocblas_create_handle")) != xq::kOK) break;
        if ((status = blas_check(rocblas_set_stream(h->blas, h->stream), "rocblas_set_stream")) != xq::kOK) break;
        const size_t nn = static_cast<size_t>(n) * n, np = static_cast<size_t>(n) * (n + 1) / 2;
        double **mats[] = {&h->X, &h->S, &h->A, &h->B, &h->C, &h->D};
        for (double **m : mats) {
            if (xq::cache_alloc(reinterpret_cast<void **>(m), nn * sizeof(double)) != hipSuccess) {
                xq::set_error("qs_dense_create: hipMalloc failed");
                status = xq::kOutOfMemory;
                break;
            }
        }
        if (status != xq::kOK) break;
        if (xq::cache_alloc(reinterpret_cast<void **>(&h->packed), np * sizeof(double)) != hipSuccess ||
            xq::cache_alloc(reinterpret_cast<void **>(&h->w), n * sizeof(double)) != hipSuccess ||
            xq::cache_alloc(reinterpret_cast<void **>(&h->e), n * sizeof(double)) != hipSuccess ||
            xq::cache_alloc(reinterpret_cast<void **>(&h->info), sizeof(rocblas_int)) != hipSuccess) {
            xq::set_error("qs_dense_create: hipMalloc failed");
            status = xq::kOutOfMemory;
            break;
        }
        if ((status = xq::check(hipMemcpyAsync(h->X, orthogonalizer, nn * sizeof(double), hipMemcpyHostToDevice, h->stream),
                                "copy X")) != xq::kOK) break;
        if ((status = upload_packed(h, overlapS, h->S)) != xq::kOK) break;
        status = xq::check(hipStreamSynchronize(h->stream), "qs_dense_create sync");
    } while (false);
    if (status != xq::kOK) {
        release(h);
        return status;
    }
    *out = h;
    return xq::kOK;
}

// QCOrbitals.MakeFromFock with the orthogonaliser (QCOrbitals.py:27-41: F' = X^T F X, EigenPairs, C = X C') and
// SymmetricMatrix.MakeFromEigenSystem for the occupied orbitals (DIISSCFConverger.py:434-446; occupancy 2 each).
int qs_dense_aufbau(qs_dense *h, const double *fockF, int nOccupied, double *densityP, double *orbitalEnergies) {
    if (h == nullptr || fockF == nullptr || densityP == nullptr || nOccupied < 0 || nOccupied > h->n) {
        xq::set_error("qs_dense_aufbau: invalid argument");
        return xq::kInvalidArgument;
    }
    XQ_TRY(check_device(h));
    // The stream is drained on every return, so that no queued copy touches the caller's arrays afterwards.
    const int status = aufbau(h, fockF, nOccupied, densityP, orbitalEnergies);
    (void)hipStreamSynchronize(h->stream);
    return status;
}

int qs_dense_commutator(qs_dense *h, const double *fockF, const double *densityP, double *error) {
    if (h == nullptr || fockF == nullptr || densityP == nullptr || error == nullptr) {
        xq::set_error("qs_dense_commutator: invalid argument");
        return xq::kInvalidArgument;
    }
    XQ_TRY(check_device(h));
    const int status = commutator(h, fockF, densityP, error);
    (void)hipStreamSynchronize(h->stream);
    return status;
}

void qs_dense_free(qs_dense *h) {
    if (h != nullptr) release(h);
}

}  // extern "C"
