/*==================================================================================================================================
! qcscf.h -- C-ABI of the SCF wrapper's integrals on the GPU (compiled into libxcquad.so beside the quadrature).
!
! The electronic structure around the XC quadrature of pDynamo3's QCModelDFT with a fit basis (QCModelDFT.py:130-160 and
! 197-250): one-electron matrices, and the Coulomb term by a charge-constrained density fit whose electron-fit integrals stay
! resident in device memory.  Entry points and the reference functions they replace:
!   qs_basis_create   <- GaussianBasisContainer.FromParameterDirectory + GaussianBasis_Finalize (GaussianBasisContainer.pyx:122-141;
!                        GaussianBasis.c:391-398): an orbital or fit basis on the atoms, in the element specification of
!                        xq_basis_set (include/xcquad.h); angular momenta 0 to 2
!   qs_one_electron   <- GaussianBasisContainerIntegrals_f1KOf1i (GaussianBasisContainerIntegrals_f1Xg1.c:233-296: kinetic and
!                        overlap) and GaussianBasisContainerIntegrals_f2Cm1V (GaussianBasisContainerIntegrals_f2Cp1.c:104-173:
!                        electron-nuclear attraction of Gaussian nuclei of charge Z, width 1e-4 bohr)
!   qs_fit_create     <- GaussianBasisContainerIntegrals_f1Xg2i (GaussianBasisContainerIntegrals_f1Xg2.c:43-128), Coulomb
!                        operator: (mu nu|P) computed once per geometry and kept in device memory
!   qs_fit_matrix     <- GaussianBasisIntegralEvaluator.f1Xf1i_f1Oi with constraints (GaussianBasisIntegralEvaluator.pyx:260-302)
!   qs_fit_b          <- the b vector of Fock_MakeCoefficientsFromFitIntegrals (FockConstruction.c:31-69), without the
!                        constraint entry (the caller appends the total charge)
!   qs_fit_fock       <- Fock_MakeFockFromFitIntegrals (FockConstruction.c:74-96)
!   qs_fit_estimate_bytes  (no counterpart: the device bytes of qs_fit_create, for admission control)
!   qs_dense_aufbau   <- QCOrbitals.MakeFromFock with the orthogonaliser (QCOrbitals.py:27-41) and
!                        SymmetricMatrix.MakeFromEigenSystem of the occupied orbitals (DIISSCFConverger.py:434-446)
!   qs_dense_commutator <- the DIIS error vector of DIISSCFConverger.DIISCurrentFrame (DIISSCFConverger.py:291-303)
! Conventions as include/xcquad.h: bohr, packed lower triangles (i(i+1)/2 + j), the reference's Status codes, host pointers.
!=================================================================================================================================*/
# ifndef QCSCF_H
# define QCSCF_H

# ifdef __cplusplus
extern "C" {
# endif

typedef struct qs_basis qs_basis ;
typedef struct qs_fit   qs_fit   ;
typedef struct qs_dense qs_dense ;

int  qs_basis_create ( int nAtoms, const int *atomicNumbers, const double *xyzBohr, int nSpec, const double *spec,
                       int spherical, qs_basis **basis ) ;
int  qs_basis_size   ( const qs_basis *basis, int *nFunctions ) ;
void qs_basis_free   ( qs_basis *basis ) ;

/* Overlap S, kinetic energy T and electron-nuclear attraction V (the basis's atoms as Gaussian nuclei of charge Z and
!  width 1e-4 bohr, the reference's default, GaussianNucleus.h), packed; all three are OVERWRITTEN. */
int  qs_one_electron ( qs_basis *basis, double *S, double *T, double *V ) ;

/* The electron-fit Coulomb integrals of an orbital and a fit basis on the same atoms (both handles must outlive the fit).
!  Shell pairs whose largest primitive prefactor is below 1e-15 are skipped; integrals of magnitude <= 1e-12 are dropped. */
int  qs_fit_create   ( const qs_basis *orbital, const qs_basis *fit, qs_fit **out ) ;
/* The constrained fit matrix, packed, extent nFit + 1: (P|Q), and in row nFit the integrals of the fit functions. */
int  qs_fit_matrix   ( qs_fit *fit, double *M ) ;
/* b_P = sum_mn P_mn (mn|P) for a packed density P (nFit values). */
int  qs_fit_b        ( qs_fit *fit, const double *densityP, double *b ) ;
/* F_mn += sum_P c_P (mn|P) (packed F, INCREMENTED; c holds nFit values). */
int  qs_fit_fock     ( qs_fit *fit, const double *c, double *fockF ) ;
/* Device bytes qs_fit_create would allocate for the orbital basis (spec, spherical) and the fit basis (fitSpec,
!  fitSpherical) on the given atoms, computed on the host with the same pair screen: admission control for a request loop.
!  No counterpart in the reference (its fit integrals are computed in host memory). */
int  qs_fit_estimate_bytes ( int nAtoms, const int *atomicNumbers, const double *xyzBohr, int nSpec, const double *spec,
                             int spherical, int nFitSpec, const double *fitSpec, int fitSpherical, long long *bytes ) ;
/* Rows of stored integrals (function pairs of the kept shell pairs), kept shell pairs, device bytes of the integrals. */
int  qs_fit_stats    ( const qs_fit *fit, long long *rows, long long *pairs, long long *deviceBytes ) ;
void qs_fit_free     ( qs_fit *fit ) ;

/* Dense SCF steps on the device (rocBLAS and rocSOLVER on a stream of the handle's own): one handle per SCF holds the
!  orthogonaliser X (dense n x n, row-major, symmetric: the reference's symmetric orthogonalisation, QCModelBase.py:313-333)
!  and the packed overlap S, on the device current at qs_dense_create, which must be current at every call (status 3
!  otherwise).  Every call is synchronous: it returns when its results are on the host.  Calls on one handle must not
!  overlap.
!  qs_dense_aufbau: P = 2 C_occ C_occ^T with C = X eigvec(X^T F X), the nOccupied lowest orbitals of the packed Fock
!  matrix F (packed P OVERWRITTEN; orbitalEnergies, n values ascending, may be NULL).
!  qs_dense_commutator: e = X^T (F P S - S P F) X for packed F and P (dense n x n, row-major, OVERWRITTEN). */
int  qs_dense_create     ( int n, const double *orthogonalizer, const double *overlapS, qs_dense **out ) ;
int  qs_dense_aufbau     ( qs_dense *dense, const double *fockF, int nOccupied, double *densityP,
                           double *orbitalEnergies ) ;
int  qs_dense_commutator ( qs_dense *dense, const double *fockF, const double *densityP, double *error ) ;
void qs_dense_free       ( qs_dense *dense ) ;

# ifdef __cplusplus
}
# endif
# endif
