/*==================================================================================================================================
! xcquad.h -- C-ABI of the MI355X exchange-correlation quadrature library (libxcquad.so).
!
! Drop-in for the C entry points behind pDynamo3's DFTGridIntegrator (pMolecule/QCModel/DFTGridIntegrator.py:25-93):
!   xq_grid_build        <- DFTGrid_Construct         (pMolecule/QCModel/__extensions__/cinclude/DFTGrid.h:59-62; DFTGrid.c:105)
!   xq_grid_counts       <- DFTGrid_NumberOfPoints / _NumberOfRecords (DFTGrid.h:74-75)
!   xq_grid_points       <- the grid point records (DFTGrid.h:28-34: atom, coordinates3, weights)
!   xq_estimated_points  <- DFTGrid_EstimatedPoints   (DFTGrid.h:67-69; DFTGrid.c:301-356)
!   xq_basis_set         <- the GaussianBasisContainer argument of DFTIntegrator_Integrate, built as
!                           GaussianBasis.pyx:114-137 + GaussianBasis_Finalize (GaussianBasis.c:391-398) build it
!   xq_integrate         <- DFTIntegrator_Integrate   (DFTIntegrator.h:55-68; DFTIntegrator.c:159-509), restricted
!   xq_integrate_uks     <- DFTIntegrator_Integrate   (the same, unrestricted: alpha and beta densities and Fock matrices)
!                           closed shell, Fock only (gradients3 = NULL)
!   xq_estimate_bytes    <- DFTGridIntegrator.EstimateMemory (DFTGridIntegrator.py:31-39), for this library's HBM layout
!   xq_free              <- DFTGrid_Deallocate        (DFTGrid.h:63-64)
!
! Conventions (as in the reference):
!   - Coordinates are bohr, n x 3 row-major FP64 (Coordinates3).  Matrices are packed lower triangles, entry (i >= j) at
!     i(i+1)/2 + j, FP64 (SymmetricMatrix.h:43-46).
!   - xq_integrate INCREMENTS V (the reference's Fock matrix is not initialised by the integrator, DFTIntegrator.pyx:9) and
!     OVERWRITES *eXC and *nElectrons (DFTIntegrator.c:174-175, 505-506).
!   - Return values are the reference's Status codes (pCore/__extensions__/cinclude/Status.h:24-33): 0 OK, 1 algorithm error
!     (used here for a HIP runtime failure), 3 invalid argument, 6 non-conformable arrays, 7 out of memory.  The caller turns
!     any non-zero code into an error, as DFTIntegrator.pyx:72 does.  xq_last_error() describes the last failure.
!   - All pointers are host pointers; the caller owns them.  The handle owns its device memory (grid, basis values,
!     work arrays), which stays resident across xq_integrate calls.  Calls on one handle must not overlap.
!=================================================================================================================================*/
# ifndef XCQUAD_H
# define XCQUAD_H

# include <stddef.h>

# ifdef __cplusplus
extern "C" {
# endif

typedef struct xq_grid xq_grid ;

/* Precision of the basis-value storage and of the two contractions (X = P Phi, V = Phi W^T + W Phi^T). */
enum {
    XQ_PRECISION_FP64  = 0 ,  /* FP64 storage, FP64 arithmetic                                                 */
    XQ_PRECISION_MIXED = 1    /* int8-digit basis values and P on int8 MFMA, exact int32 sums; FP64 functional   */
} ;

/* Grid accuracy: 0 VeryLow, 1 Low, 2 Medium, 3 High, 4 VeryHigh (DFTGrid.h:16-25). */
int         xq_grid_build       ( int accuracy, int nAtoms, const int *atomicNumbers, const double *xyzBohr, xq_grid **grid ) ;
int         xq_grid_counts      ( const xq_grid *grid, int *nPoints, int *nBlocks ) ;
int         xq_grid_points      ( const xq_grid *grid, double *xyzw, int *blockAtom, int *blockCount ) ;
int         xq_estimated_points ( int accuracy, int nAtoms, const int *atomicNumbers, long long *nPoints ) ;

/* Basis: element specification (flat FP64), per element Z, nShells, then per shell lLow, lHigh, nPrimitives, then per
!  primitive the exponent and one primitive-normalised coefficient per l in lLow..lHigh.  spherical != 0 selects real solid
!  harmonics (the reference default), 0 Cartesian functions.  Evaluates and stores the screened basis values and gradients of
!  every grid block on the device (the reference's inCore mode).  precision selects the storage (XQ_PRECISION_*). */
int         xq_basis_set        ( xq_grid *grid, int nSpec, const double *spec, int spherical, int precision, int *nFunctions ) ;

/* One quadrature.  ids: libxc functional IDs (1 Slater exchange, 7 VWN, 8 VWN-RPA, 101 PBE-x, 106 B88, 130 PBE-c, 131 LYP,
!  402 B3LYP's XC part; pMolecule/QCModel/DFTDefinitions.py), all LDA or all GGA.
!  densityP: packed total density (n(n+1)/2).  fockV: packed, incremented by V_xc.  Statistics (optional, may be NULL, else
!  XQ_NSTATS doubles): [0] sum over blocks of kept functions x points ("Quadrature BF Values"), [1] blocks skipped by the
!  density tolerance, [2] device ms from P expansion to packed V, [3] device ms of the density/functional kernel, [4] device
!  ms of the Fock stage, [5] sum over blocks of kept functions^2 x points (the contractions' algorithmic flops are
!  2 [5] each), [6] and [7] algorithmic HBM bytes of the density and Fock stages (each value of a kept function at a
!  point read once in the stage's storage format, each result written once), [8] sum over blocks of (kept functions
!  rounded up to 32)^2 x 128, the shape the matrix-core products run on (the mixed path's X = P_red Phi issues
!  2 x 15 x [8] int8 operations: 15 digit products), [9] device ms of the density product kernel alone (the mixed path's
!  X = P_red Phi with each row chunk's share of rho and grad rho; FP64: the density kernel, functional included, as [3]),
!  [10] its algorithmic HBM bytes (as [6]). */
# define XQ_NSTATS 11
int         xq_integrate        ( xq_grid *grid, int nIDs, const int *ids, const double *densityP, double *eXC,
                                  double *nElectrons, double *fockV, double *stats ) ;

/* xq_integrate without allocation.  No counterpart in the reference: it is for callers that cannot allocate device or
!  pinned host memory during a call (code under HIP graph capture, serving loops that forbid allocation).  Same arguments,
!  semantics and results as xq_integrate (E_xc and N_e bit for bit; V_xc up to the order of the Fock stage's FP64 atomic
!  additions), on a handle set up by xq_grid_build and xq_basis_set (which allocate).  The per-call work arrays live in one
!  static device array of 64 MiB per GPU, enough for about 1,600 basis functions; a larger molecule is refused with
!  status 3 before any work is queued.  P and V move through the caller's own (pageable) host arrays.  The handle's device
!  must be current.  Every call on one GPU shares that GPU's array: calls that overlap on one GPU race.
!  xq_noalloc_scratch reports the bytes a handle's call needs (neededBytes, 0 for a NULL handle) and the array's size. */
int         xq_integrate_noalloc ( xq_grid *grid, int nIDs, const int *ids, const double *densityP, double *eXC,
                                   double *nElectrons, double *fockV, double *stats ) ;
int         xq_noalloc_scratch   ( const xq_grid *grid, long long *neededBytes, long long *limitBytes ) ;

/* One unrestricted quadrature: DFTIntegrator_Integrate with densityQ, fockB and isSpinUnrestricted (DFTIntegrator.c:159-509,
!  the spin branches :301-330, :342-347, :366-402), as DFTIntegrator.pyx:82-107 calls it with alpha and beta matrices (the
!  caller converts total and spin matrices: alpha = (T + S)/2, beta = (T - S)/2; F_T += (F_a + F_b)/2, F_S += (F_a - F_b)/2).
!  densityA, densityB: packed alpha and beta densities; fockA, fockB: packed, INCREMENTED by V_xc of each spin; eXC and
!  nElectrons as xq_integrate; stats as xq_integrate ([3], [4], [6], [7] are 0).  The functionals are evaluated
!  spin-polarised (libxc with nspin = 2). */
int         xq_integrate_uks    ( xq_grid *grid, int nIDs, const int *ids, const double *densityA, const double *densityB,
                                  double *eXC, double *nElectrons, double *fockA, double *fockB, double *stats ) ;

/* XC nuclear gradients: DFTIntegrator_Integrate with a gradients3 argument and no Fock matrix (DFTIntegrator.c:159-509, the
!  doGradients branch :405-487 with DFTGridWeights_Derivatives, DFTGridWeights.c:98-215), as DFTGridIntegrator.Gradients
!  calls it (DFTGridIntegrator.py:72-88).  Restricted closed shell, FP64 on either precision's handle (the basis values and
!  second derivatives are evaluated afresh).  gradients3: nAtoms x 3 in hartree/bohr, INCREMENTED as the reference's
!  Coordinates3 is.  At most 1024 atoms. */
int         xq_gradients        ( xq_grid *grid, int nIDs, const int *ids, const double *densityP, double *gradients3 ) ;

/* Spin-unrestricted XC nuclear gradients: the same call with isSpinUnrestricted, alpha and beta densities and the
!  polarised functional (DFTIntegrator.c:405-487: DFTIntegrator_Gradients per spin and the two DFTIntegrator_GradientsSigma
!  cross calls with vSigmaPQ; the weight terms from eps (rho_a + rho_b)).  densityA, densityB: packed alpha and beta
!  densities (the caller converts total and spin matrices as for xq_integrate_uks); gradients3 as xq_gradients. */
int         xq_gradients_uks    ( xq_grid *grid, int nIDs, const int *ids, const double *densityA, const double *densityB,
                                  double *gradients3 ) ;

/* Peak device bytes of one handle for a molecule with the given grid size, function count and mean kept functions per
!  block, set-up included (admission control for the request loop); xq_device_bytes is what a handle holds now, and
!  xq_device_memory the device's free and total bytes (hipMemGetInfo; free includes the device blocks the library holds
!  idle for reuse by its handles). */
int         xq_estimate_bytes   ( long long nPoints, int nFunctions, double meanKeptFunctions, int precision, long long *bytes ) ;
int         xq_device_bytes     ( const xq_grid *grid, long long *bytes ) ;
int         xq_device_memory    ( long long *freeBytes, long long *totalBytes ) ;

void        xq_free             ( xq_grid *grid ) ;
const char *xq_last_error       ( void ) ;

/* Test hook, no counterpart in the reference: after `count` more device allocations by the library, the next one fails
!  as hipMalloc does when the device is full (once); a negative count disarms it.  Lets a test drive the out-of-memory
!  paths (status 7) and check that the handle is usable afterwards. */
void        xq_test_fail_allocations ( long long count ) ;

/* Batched quadratures for a request loop.  No counterpart in the reference (its integrator takes one molecule per call):
!  the quadratures of several handles (one per job, each with its own P, V and functional) run on the batch's stream, and
!  each mixed-precision stage is one launch over all the jobs' blocks.  Every job's results are those of xq_integrate
!  on its handle alone: E_xc and N_e bit for bit (per-block FP64 sums in block order), V_xc up to the order of the FP64
!  atomic additions.  ids holds 4 entries per job (the first nIDs[j] are used); stats (optional) is XQ_NSTATS doubles for
!  the whole batch.  All jobs must share one precision; a handle may appear once per batch, and must not be in use by
!  another call meanwhile.  Calls on one batch must not overlap. */
typedef struct xq_batch xq_batch ;
int         xq_batch_create     ( int maxJobs, xq_batch **batch ) ;
int         xq_integrate_batch  ( xq_batch *batch, int nJobs, xq_grid *const *grids, const int *nIDs, const int *ids,
                                  const double *const *densityP, double *eXC, double *nElectrons, double *const *fockV,
                                  double *stats ) ;
void        xq_batch_free       ( xq_batch *batch ) ;

# ifdef __cplusplus
}
# endif
# endif
