/*==================================================================================================================================
! Driver for the compiled pDynamo3 reference C path (test infrastructure only).
!
! This file is NOT part of the product.  It is linked, by oracle/Makefile, against the reference's own C sources compiled where
! they lie under the reference tree (DFTGrid.c, DFTGridWeights.c, DFTIntegrator.c, the vendored libxc "cxc", the Gaussian basis
! C files, pCore and pScientific arrays), producing oracle/_ref/libxcref.so.  It exists only to write golden vectors and to time
! the reference's own CPU path; no product code loads it.
!
! Each entry point reproduces what the reference's Python/Cython layer does before calling C:
!   - basis set construction follows GaussianBasis.pyx:__setstate__ (pyrex/pMolecule.QCModel.GaussianBases.GaussianBasis.pyx:114-137)
!     followed by Finalize (GaussianBasisContainer.pyx:136) and the container index arrays (GaussianBasisContainer.pyx:100-112);
!   - grid construction calls DFTGrid_Construct (DFTGrid.c:105) as DFTGrid.pyx:57-73 does;
!   - the quadrature calls DFTIntegrator_Integrate (DFTIntegrator.c:159) as DFTIntegrator.pyx:23-73 does (restricted case).
!=================================================================================================================================*/
# include <math.h>
# include <stdlib.h>
# include <string.h>
# include <time.h>

# include "BlockStorage.h"
# include "Coordinates3.h"
# include "DFTFunctionalModel.h"
# include "DFTGrid.h"
# include "DFTIntegrator.h"
# include "GaussianBasis.h"
# include "GaussianBasisContainer.h"
# include "GaussianBasisContainerIntegrals_f1Op1.h"
# include "GaussianBasisContainerIntegrals_f1X.h"
# include "GaussianBasisContainerIntegrals_f1Xg1.h"
# include "GaussianBasisContainerIntegrals_f1Xg2.h"
# include "GaussianBasisContainerIntegrals_f2Cp1.h"
# include "RealArray1D.h"
# include "GridFunctionDataBlock.h"
# include "IntegerArray1D.h"
# include "SymmetricMatrix.h"

typedef struct {
    Integer                 nAtoms      ;
    Integer                 nFunctions  ;
    Coordinates3           *coordinates ;
    IntegerArray1D         *atomicNumbers ;
    DFTGrid                *grid        ;
    GaussianBasisContainer *bases       ;
    GaussianBasis         **unique      ;
    Integer                 nUnique     ;
    DFTFunctionalModel     *functional  ;
    DFTFunctionalModel     *functionalU ;   /* the same IDs, spin-polarised (unrestricted models) */
    IntegerArray1D         *cfp, *fc    ;
} RefHandle ;

/* . Basis specification: a flat array of doubles, per element:
!    Z, nShells, then per shell: lLow, lHigh, nPrimitives, then per primitive: exponent, coefficient(lLow..lHigh) (normalized). */
static GaussianBasis *MakeBasis ( const double *spec, int *consumed, int spherical )
{
    int p = 0, s, q, c, z, nShells ;
    GaussianBasis *b ;
    z       = ( int ) spec[p++] ;
    nShells = ( int ) spec[p++] ;
    b = GaussianBasis_Allocate ( nShells ) ;
    b->atomicNumber = z ;
    b->isSpherical  = ( spherical ? True : False ) ;
    b->pNormalized  = True ;
    for ( s = 0 ; s < nShells ; s++ )
    {
        int lLow  = ( int ) spec[p++] ;
        int lHigh = ( int ) spec[p++] ;
        int nP    = ( int ) spec[p++] ;
        GaussianBasis_AllocateShell ( b, s, lHigh, lLow, nP ) ;
        for ( q = 0 ; q < nP ; q++ )
        {
            b->shells[s].primitives[q].exponent  = spec[p] ;
            b->shells[s].primitives[q].exponent0 = spec[p] ; p++ ;
            for ( c = 0 ; c < ( lHigh - lLow + 1 ) ; c++ )
            {
                b->shells[s].primitives[q].coefficients [c] = spec[p] ;
                b->shells[s].primitives[q].coefficients0[c] = spec[p] ; p++ ;
            }
        }
    }
    GaussianBasis_UnnormalizePrimitives ( b ) ;
    {
        Status st = Status_OK ;
        GaussianBasis_Finalize ( b, &st ) ;
    }
    *consumed = p ;
    return b ;
}

void *xcref_setup ( int accuracy, int nAtoms, const int *Z, const double *xyzBohr,
                    int nSpec, const double *spec, int spherical, int nIDs, const int *ids, int *status )
{
    RefHandle *h = ( RefHandle * ) calloc ( 1, sizeof ( RefHandle ) ) ;
    Status     st = Status_OK ;
    int        i, p, u ;
    h->nAtoms        = nAtoms ;
    h->coordinates   = Coordinates3_Allocate ( nAtoms, &st ) ;
    h->atomicNumbers = IntegerArray1D_AllocateWithExtent ( nAtoms, &st ) ;
    for ( i = 0 ; i < nAtoms ; i++ )
    {
        Coordinates3_SetRow ( h->coordinates, i, xyzBohr[3*i], xyzBohr[3*i+1], xyzBohr[3*i+2] ) ;
        Array1D_Item ( h->atomicNumbers, i ) = Z[i] ;
    }
    /* . Unique bases. */
    h->unique  = ( GaussianBasis ** ) calloc ( 128, sizeof ( GaussianBasis * ) ) ;
    h->nUnique = 0 ;
    for ( p = 0 ; p < nSpec ; )
    {
        int used = 0 ;
        GaussianBasis *b = MakeBasis ( &spec[p], &used, spherical ) ;
        h->unique[h->nUnique++] = b ;
        p += used ;
    }
    /* . Container: shares the unique entries (not the owner). */
    h->bases = GaussianBasisContainer_Allocate ( nAtoms, &st ) ;
    for ( i = 0 ; i < nAtoms ; i++ )
    {
        GaussianBasis *b = NULL ;
        for ( u = 0 ; u < h->nUnique ; u++ ) if ( h->unique[u]->atomicNumber == Z[i] ) b = h->unique[u] ;
        if ( b == NULL ) { *status = 3 ; return NULL ; }
        h->bases->entries[i] = b ;
    }
    h->nFunctions = GaussianBasisContainer_NumberOfFunctions ( h->bases ) ;
    h->cfp = IntegerArray1D_AllocateWithExtent ( nAtoms + 1   , &st ) ;
    h->fc  = IntegerArray1D_AllocateWithExtent ( h->nFunctions, &st ) ;
    GaussianBasisContainer_MakeIndexArrays ( h->bases, h->cfp, h->fc, &st ) ;
    /* . Functional. */
    {
        IntegerArray1D *idArray = IntegerArray1D_AllocateWithExtent ( nIDs, &st ) ;
        for ( i = 0 ; i < nIDs ; i++ ) Array1D_Item ( idArray, i ) = ids[i] ;
        h->functional  = DFTFunctionalModel_MakeFromIDs ( idArray, True , &st ) ;
        h->functionalU = DFTFunctionalModel_MakeFromIDs ( idArray, False, &st ) ;
        IntegerArray1D_Deallocate ( &idArray ) ;
    }
    /* . Grid. */
    h->grid = DFTGrid_Construct ( ( DFTGridAccuracy ) accuracy, h->atomicNumbers, h->coordinates, &st ) ;
    DFTGrid_MakeRecords ( h->grid, &st ) ;
    *status = ( int ) st ;
    return ( void * ) h ;
}

int xcref_nfunctions ( void *vh ) { return ( ( RefHandle * ) vh )->nFunctions ; }
int xcref_npoints    ( void *vh ) { return DFTGrid_NumberOfPoints  ( ( ( RefHandle * ) vh )->grid ) ; }
int xcref_nblocks    ( void *vh ) { return DFTGrid_NumberOfRecords ( ( ( RefHandle * ) vh )->grid ) ; }
int xcref_estimated_points ( void *vh )
{
    RefHandle *h = ( RefHandle * ) vh ;
    Status st = Status_OK ;
    return DFTGrid_EstimatedPoints ( h->grid->accuracy, h->atomicNumbers, &st ) ;
}

/* . Grid points (x,y,z,w per point, in record order) and per-block (atom, count). */
void xcref_grid ( void *vh, double *xyzw, int *blockAtom, int *blockCount )
{
    RefHandle *h = ( RefHandle * ) vh ;
    int r, i, n = 0 ;
    for ( r = 0 ; r < h->grid->numberOfRecords ; r++ )
    {
        DFTGridPointBlock *b = h->grid->records[r] ;
        blockAtom [r] = b->atom ;
        blockCount[r] = b->numberOfPoints ;
        for ( i = 0 ; i < b->numberOfPoints ; i++, n++ )
        {
            xyzw[4*n  ] = Coordinates3_Item ( b->coordinates3, i, 0 ) ;
            xyzw[4*n+1] = Coordinates3_Item ( b->coordinates3, i, 1 ) ;
            xyzw[4*n+2] = Coordinates3_Item ( b->coordinates3, i, 2 ) ;
            xyzw[4*n+3] = Array1D_Item ( b->weights, i ) ;
        }
    }
}

/* . One quadrature: returns E_xc and N_e; V (packed lower triangle) is incremented in place.  seconds = wall time of the call. */
int xcref_integrate ( void *vh, const double *P, int inCore, double *eXC, double *rho, double *V, double *seconds )
{
    RefHandle *h = ( RefHandle * ) vh ;
    Status st = Status_OK ;
    int n = h->nFunctions, k, size = ( n * ( n + 1 ) ) / 2 ;
    struct timespec t0, t1 ;
    SymmetricMatrix *dP = SymmetricMatrix_AllocateWithExtent ( n, &st ) ;
    SymmetricMatrix *fA = SymmetricMatrix_AllocateWithExtent ( n, &st ) ;
    for ( k = 0 ; k < size ; k++ ) { dP->data[k] = P[k] ; fA->data[k] = V[k] ; }
    clock_gettime ( CLOCK_MONOTONIC, &t0 ) ;
    DFTIntegrator_Integrate ( h->functional, h->grid, h->bases, h->coordinates, dP, NULL, ( inCore ? True : False ), False,
                              eXC, rho, fA, NULL, NULL, &st ) ;
    clock_gettime ( CLOCK_MONOTONIC, &t1 ) ;
    for ( k = 0 ; k < size ; k++ ) V[k] = fA->data[k] ;
    if ( seconds != NULL ) *seconds = ( t1.tv_sec - t0.tv_sec ) + 1.0e-9 * ( t1.tv_nsec - t0.tv_nsec ) ;
    SymmetricMatrix_Deallocate ( &dP ) ;
    SymmetricMatrix_Deallocate ( &fA ) ;
    return ( int ) st ;
}

/* . One unrestricted quadrature (DFTIntegrator.pyx:82-107: alpha and beta densities, alpha and beta Fock matrices): E_xc and
!    N_e; Va and Vb (packed) are incremented in place. */
int xcref_integrate_uks ( void *vh, const double *Pa, const double *Pb, int inCore, double *eXC, double *rho, double *Va,
                          double *Vb, double *seconds )
{
    RefHandle *h = ( RefHandle * ) vh ;
    Status st = Status_OK ;
    int n = h->nFunctions, k, size = ( n * ( n + 1 ) ) / 2 ;
    struct timespec t0, t1 ;
    SymmetricMatrix *dP = SymmetricMatrix_AllocateWithExtent ( n, &st ), *dQ = SymmetricMatrix_AllocateWithExtent ( n, &st ) ;
    SymmetricMatrix *fA = SymmetricMatrix_AllocateWithExtent ( n, &st ), *fB = SymmetricMatrix_AllocateWithExtent ( n, &st ) ;
    for ( k = 0 ; k < size ; k++ ) { dP->data[k] = Pa[k] ; dQ->data[k] = Pb[k] ; fA->data[k] = Va[k] ; fB->data[k] = Vb[k] ; }
    clock_gettime ( CLOCK_MONOTONIC, &t0 ) ;
    DFTIntegrator_Integrate ( h->functionalU, h->grid, h->bases, h->coordinates, dP, dQ, ( inCore ? True : False ), True,
                              eXC, rho, fA, fB, NULL, &st ) ;
    clock_gettime ( CLOCK_MONOTONIC, &t1 ) ;
    for ( k = 0 ; k < size ; k++ ) { Va[k] = fA->data[k] ; Vb[k] = fB->data[k] ; }
    if ( seconds != NULL ) *seconds = ( t1.tv_sec - t0.tv_sec ) + 1.0e-9 * ( t1.tv_nsec - t0.tv_nsec ) ;
    SymmetricMatrix_Deallocate ( &dP ) ; SymmetricMatrix_Deallocate ( &dQ ) ;
    SymmetricMatrix_Deallocate ( &fA ) ; SymmetricMatrix_Deallocate ( &fB ) ;
    return ( int ) st ;
}

/* . XC nuclear gradients: DFTIntegrator_Integrate with gradients3 and no Fock matrix, as DFTGridIntegrator.Gradients
!    (DFTGridIntegrator.py:72-88) calls it.  grad (nAtoms x 3) is incremented. */
int xcref_gradients ( void *vh, const double *P, int inCore, double *grad )
{
    RefHandle *h = ( RefHandle * ) vh ;
    Status st = Status_OK ;
    int n = h->nFunctions, k, i, size = ( n * ( n + 1 ) ) / 2 ;
    Real eXC = 0.0, rho = 0.0 ;
    SymmetricMatrix *dP = SymmetricMatrix_AllocateWithExtent ( n, &st ) ;
    Coordinates3 *g3 = Coordinates3_Allocate ( h->nAtoms, &st ) ;
    for ( k = 0 ; k < size ; k++ ) dP->data[k] = P[k] ;
    for ( i = 0 ; i < h->nAtoms ; i++ ) Coordinates3_SetRow ( g3, i, grad[3*i], grad[3*i+1], grad[3*i+2] ) ;
    DFTIntegrator_Integrate ( h->functional, h->grid, h->bases, h->coordinates, dP, NULL, ( inCore ? True : False ), False,
                              &eXC, &rho, NULL, NULL, g3, &st ) ;
    for ( i = 0 ; i < h->nAtoms ; i++ )
    {
        grad[3*i  ] = Coordinates3_Item ( g3, i, 0 ) ;
        grad[3*i+1] = Coordinates3_Item ( g3, i, 1 ) ;
        grad[3*i+2] = Coordinates3_Item ( g3, i, 2 ) ;
    }
    SymmetricMatrix_Deallocate ( &dP ) ;
    Coordinates3_Deallocate ( &g3 ) ;
    return ( int ) st ;
}

/* . Unrestricted XC nuclear gradients: the same with the alpha and beta densities and the polarised functional
!    (DFTIntegrator.c:405-487 with isSpinUnrestricted).  grad (nAtoms x 3) is incremented. */
int xcref_gradients_uks ( void *vh, const double *Pa, const double *Pb, int inCore, double *grad )
{
    RefHandle *h = ( RefHandle * ) vh ;
    Status st = Status_OK ;
    int n = h->nFunctions, k, i, size = ( n * ( n + 1 ) ) / 2 ;
    Real eXC = 0.0, rho = 0.0 ;
    SymmetricMatrix *dP = SymmetricMatrix_AllocateWithExtent ( n, &st ), *dQ = SymmetricMatrix_AllocateWithExtent ( n, &st ) ;
    Coordinates3 *g3 = Coordinates3_Allocate ( h->nAtoms, &st ) ;
    for ( k = 0 ; k < size ; k++ ) { dP->data[k] = Pa[k] ; dQ->data[k] = Pb[k] ; }
    for ( i = 0 ; i < h->nAtoms ; i++ ) Coordinates3_SetRow ( g3, i, grad[3*i], grad[3*i+1], grad[3*i+2] ) ;
    DFTIntegrator_Integrate ( h->functionalU, h->grid, h->bases, h->coordinates, dP, dQ, ( inCore ? True : False ), True,
                              &eXC, &rho, NULL, NULL, g3, &st ) ;
    for ( i = 0 ; i < h->nAtoms ; i++ )
    {
        grad[3*i  ] = Coordinates3_Item ( g3, i, 0 ) ;
        grad[3*i+1] = Coordinates3_Item ( g3, i, 1 ) ;
        grad[3*i+2] = Coordinates3_Item ( g3, i, 2 ) ;
    }
    SymmetricMatrix_Deallocate ( &dP ) ; SymmetricMatrix_Deallocate ( &dQ ) ;
    Coordinates3_Deallocate ( &g3 ) ;
    return ( int ) st ;
}

/* . Basis values and first derivatives of one block after the reference's |phi| > bfTol filter.
!    Outputs: nb, indices[nb], f/fx/fy/fz [nb x nPoints] row-major.  Returns nb. */
int xcref_block_values ( void *vh, int record, int *indices, double *f, double *fx, double *fy, double *fz )
{
    RefHandle *h = ( RefHandle * ) vh ;
    Status st = Status_OK ;
    DFTGridPointBlock *b = h->grid->records[record] ;
    GridFunctionDataBlock *d = GridFunctionDataBlock_Allocate ( h->nFunctions, b->numberOfPoints, 1, &st ) ;
    int i, p, nb ;
    GaussianBasisContainerIntegrals_f1Op1ir123 ( h->bases, h->coordinates, b->coordinates3, True, &(h->grid->bfTolerance), d, &st ) ;
    nb = d->numberOfFunctions ;
    for ( i = 0 ; i < nb ; i++ )
    {
        indices[i] = Array1D_Item ( d->indices, i ) ;
        for ( p = 0 ; p < b->numberOfPoints ; p++ )
        {
            f [i*b->numberOfPoints+p] = Array2D_Item ( d->f , i, p ) ;
            fx[i*b->numberOfPoints+p] = Array2D_Item ( d->fX, i, p ) ;
            fy[i*b->numberOfPoints+p] = Array2D_Item ( d->fY, i, p ) ;
            fz[i*b->numberOfPoints+p] = Array2D_Item ( d->fZ, i, p ) ;
        }
    }
    GridFunctionDataBlock_Deallocate ( &d ) ;
    return nb ;
}

/* . Functional values for an unpolarized density: eps (per particle), vrho, vsigma, summed over the functionals as
!    DFTFunctionalModel_Evaluate does (DFTFunctionalModel.c:108-154). */
int xcref_functional ( int nIDs, const int *ids, int n, const double *rho, const double *sigma,
                       double *eXC, double *vRho, double *vSigma )
{
    int f, i ;
    double *e = ( double * ) malloc ( n * sizeof ( double ) ), *vr = ( double * ) malloc ( n * sizeof ( double ) ),
           *vs = ( double * ) malloc ( n * sizeof ( double ) ) ;
    for ( i = 0 ; i < n ; i++ ) { eXC[i] = 0.0 ; vRho[i] = 0.0 ; vSigma[i] = 0.0 ; }
    for ( f = 0 ; f < nIDs ; f++ )
    {
        xc_func_type func ;
        if ( xc_func_init ( &func, ids[f], XC_UNPOLARIZED ) < 0 ) { free ( e ) ; free ( vr ) ; free ( vs ) ; return 3 ; }
        if ( func.info->family == XC_FAMILY_LDA )
        {
            xc_lda_exc_vxc ( &func, n, rho, e, vr ) ;
            for ( i = 0 ; i < n ; i++ ) vs[i] = 0.0 ;
        }
        else xc_gga_exc_vxc ( &func, n, rho, sigma, e, vr, vs ) ;
        for ( i = 0 ; i < n ; i++ ) { eXC[i] += e[i] ; vRho[i] += vr[i] ; vSigma[i] += vs[i] ; }
        xc_func_end ( &func ) ;
    }
    free ( e ) ; free ( vr ) ; free ( vs ) ;
    return 0 ;
}

/* . The same, spin-polarised: rho [n][2], sigma [n][3] -> eXC [n], vRho [n][2], vSigma [n][3]. */
int xcref_functional_pol ( int nIDs, const int *ids, int n, const double *rho, const double *sigma,
                           double *eXC, double *vRho, double *vSigma )
{
    int f, i ;
    double *e = ( double * ) malloc ( n * sizeof ( double ) ), *vr = ( double * ) malloc ( 2 * n * sizeof ( double ) ),
           *vs = ( double * ) malloc ( 3 * n * sizeof ( double ) ) ;
    for ( i = 0 ; i < n ; i++ ) eXC[i] = 0.0 ;
    for ( i = 0 ; i < 2 * n ; i++ ) vRho[i] = 0.0 ;
    for ( i = 0 ; i < 3 * n ; i++ ) vSigma[i] = 0.0 ;
    for ( f = 0 ; f < nIDs ; f++ )
    {
        xc_func_type func ;
        if ( xc_func_init ( &func, ids[f], XC_POLARIZED ) < 0 ) { free ( e ) ; free ( vr ) ; free ( vs ) ; return 3 ; }
        if ( func.info->family == XC_FAMILY_LDA )
        {
            xc_lda_exc_vxc ( &func, n, rho, e, vr ) ;
            for ( i = 0 ; i < 3 * n ; i++ ) vs[i] = 0.0 ;
        }
        else xc_gga_exc_vxc ( &func, n, rho, sigma, e, vr, vs ) ;
        for ( i = 0 ; i < n ; i++ ) eXC[i] += e[i] ;
        for ( i = 0 ; i < 2 * n ; i++ ) vRho[i] += vr[i] ;
        for ( i = 0 ; i < 3 * n ; i++ ) vSigma[i] += vs[i] ;
        xc_func_end ( &func ) ;
    }
    free ( e ) ; free ( vr ) ; free ( vs ) ;
    return 0 ;
}

/* . Cartesian to spherical transformation for lLow..lHigh (nc x ns, row-major). */
void xcref_c2s ( int lLow, int lHigh, double *out )
{
    Status st = Status_OK ;
    RealArray2D *t = GaussianBasis_TransformationCartesianToSpherical ( lLow, lHigh, &st ) ;
    int i, j ;
    for ( i = 0 ; i < View2D_Rows ( t ) ; i++ )
        for ( j = 0 ; j < View2D_Columns ( t ) ; j++ ) out[i*View2D_Columns ( t )+j] = Array2D_Item ( t, i, j ) ;
    RealArray2D_Deallocate ( &t ) ;
}

/*----------------------------------------------------------------------------------------------------------------------------------
! . Integrals of the SCF wrapper (electronic structure around the quadrature), as QCModelDFT's energy closures call them
!   (QCModelDFT.py:130-160 through GaussianBasisIntegralEvaluator.pyx).
!---------------------------------------------------------------------------------------------------------------------------------*/
/* . A basis-only handle: coordinates, container and index arrays (orbital or fit basis). */
void *xcref_basis_setup ( int nAtoms, const int *Z, const double *xyzBohr, int nSpec, const double *spec, int spherical,
                          int *status )
{
    RefHandle *h = ( RefHandle * ) calloc ( 1, sizeof ( RefHandle ) ) ;
    Status     st = Status_OK ;
    int        i, p, u ;
    h->nAtoms        = nAtoms ;
    h->coordinates   = Coordinates3_Allocate ( nAtoms, &st ) ;
    h->atomicNumbers = IntegerArray1D_AllocateWithExtent ( nAtoms, &st ) ;
    for ( i = 0 ; i < nAtoms ; i++ )
    {
        Coordinates3_SetRow ( h->coordinates, i, xyzBohr[3*i], xyzBohr[3*i+1], xyzBohr[3*i+2] ) ;
        Array1D_Item ( h->atomicNumbers, i ) = Z[i] ;
    }
    h->unique  = ( GaussianBasis ** ) calloc ( 128, sizeof ( GaussianBasis * ) ) ;
    h->nUnique = 0 ;
    for ( p = 0 ; p < nSpec ; )
    {
        int used = 0 ;
        h->unique[h->nUnique++] = MakeBasis ( &spec[p], &used, spherical ) ;
        p += used ;
    }
    h->bases = GaussianBasisContainer_Allocate ( nAtoms, &st ) ;
    for ( i = 0 ; i < nAtoms ; i++ )
    {
        GaussianBasis *b = NULL ;
        for ( u = 0 ; u < h->nUnique ; u++ ) if ( h->unique[u]->atomicNumber == Z[i] ) b = h->unique[u] ;
        if ( b == NULL ) { *status = 3 ; return NULL ; }
        h->bases->entries[i] = b ;
    }
    h->nFunctions = GaussianBasisContainer_NumberOfFunctions ( h->bases ) ;
    h->cfp = IntegerArray1D_AllocateWithExtent ( nAtoms + 1   , &st ) ;
    h->fc  = IntegerArray1D_AllocateWithExtent ( h->nFunctions, &st ) ;
    GaussianBasisContainer_MakeIndexArrays ( h->bases, h->cfp, h->fc, &st ) ;
    *status = ( int ) st ;
    return h ;
}

static SymmetricMatrix *NewSymmetric ( int n, Status *st )
{
    SymmetricMatrix *m = SymmetricMatrix_AllocateWithExtent ( n, st ) ;
    SymmetricMatrix_Set ( m, 0.0 ) ;
    return m ;
}

/* . Overlap, kinetic and electron-nuclear matrices (packed), as f1KOf1i and f2Cm1V fill them (point nuclei of charge Z). */
int xcref_one_electron ( void *vh, double *S, double *T, double *V )
{
    RefHandle       *h  = ( RefHandle * ) vh ;
    Status           st = Status_OK ;
    int              i, n = h->nFunctions, size = n * ( n + 1 ) / 2 ;
    SymmetricMatrix *s = NewSymmetric ( n, &st ), *t = NewSymmetric ( n, &st ), *v = NewSymmetric ( n, &st ) ;
    RealArray1D     *charges = RealArray1D_AllocateWithExtent ( h->nAtoms, &st ) ;
    for ( i = 0 ; i < h->nAtoms ; i++ ) Array1D_Item ( charges, i ) = ( double ) Array1D_Item ( h->atomicNumbers, i ) ;
    GaussianBasisContainerIntegrals_f1KOf1i ( h->bases, h->coordinates, t, s, &st ) ;
    GaussianBasisContainerIntegrals_f2Cm1V  ( h->bases, charges, NULL, NULL, h->coordinates, h->coordinates, NULL, v, &st ) ;
    for ( i = 0 ; i < size ; i++ ) { S[i] = s->data[i] ; T[i] = t->data[i] ; V[i] = v->data[i] ; }
    SymmetricMatrix_Deallocate ( &s ) ;
    SymmetricMatrix_Deallocate ( &t ) ;
    SymmetricMatrix_Deallocate ( &v ) ;
    RealArray1D_Deallocate ( &charges ) ;
    return ( int ) st ;
}

/* . Electron-fit Coulomb integrals (mu nu|P) as f1Xg2i stores them (the reference drops values below its underflow
!    threshold), scattered into a dense array [packed orbital pair][fit function]; the number of stored values. */
int xcref_fit_integrals ( void *vorb, void *vfit, double *dense, int *count )
{
    RefHandle    *o = ( RefHandle * ) vorb, *f = ( RefHandle * ) vfit ;
    Status        st = Status_OK ;
    BlockStorage *store = BlockStorage_Allocate ( &st ) ;
    Block        *block ;
    int           i, nFit = f->nFunctions, stored = 0 ;
    GaussianBasisContainerIntegrals_f1Xg2i ( o->bases, f->bases, o->coordinates, GaussianBasisOperator_Coulomb, store, &st ) ;
    List_Iterate_Initialize ( store->blocks ) ;
    while ( ( block = BlockStorage_Iterate ( store ) ) != NULL )
    {
        for ( i = 0 ; i < block->count ; i++ )
        {
            dense[( size_t ) block->indices32[i] * nFit + block->indices16[i]] = block->data[i] ;
            stored++ ;
        }
    }
    *count = stored ;
    BlockStorage_Deallocate ( &store ) ;
    return ( int ) st ;
}

/* . The fit matrix of f1Xf1i_f1Oi with constraints (GaussianBasisIntegralEvaluator.pyx:260-302): (P|Q) and, in the last row,
!    the fit functions' integrals; packed, extent nFit + 1. */
int xcref_fit_matrix ( void *vfit, double *M )
{
    RefHandle       *f  = ( RefHandle * ) vfit ;
    Status           st = Status_OK ;
    int              i, n = f->nFunctions, size = ( n + 1 ) * ( n + 2 ) / 2 ;
    SymmetricMatrix *eri = NewSymmetric ( n + 1, &st ) ;
    RealArray1D     *fso = RealArray1D_AllocateWithExtent ( n, &st ) ;
    GaussianBasisContainerIntegrals_f1Cf1i ( f->bases, f->coordinates, eri, &st ) ;
    GaussianBasisContainerIntegrals_f1Oi   ( f->bases, fso, &st ) ;
    for ( i = 0 ; i < n ; i++ ) SymmetricMatrix_Item ( eri, n, i ) = Array1D_Item ( fso, i ) ;
    for ( i = 0 ; i < size ; i++ ) M[i] = eri->data[i] ;
    SymmetricMatrix_Deallocate ( &eri ) ;
    RealArray1D_Deallocate ( &fso ) ;
    return ( int ) st ;
}

void xcref_free ( void *vh )
{
    RefHandle *h = ( RefHandle * ) vh ;
    int u ;
    Status st = Status_OK ;
    if ( h == NULL ) return ;
    DFTGrid_Deallocate ( &h->grid, &st ) ;
    DFTFunctionalModel_Deallocate ( &h->functional ) ;
    DFTFunctionalModel_Deallocate ( &h->functionalU ) ;
    for ( u = 0 ; u < h->nUnique ; u++ ) GaussianBasis_Deallocate ( &h->unique[u] ) ;
    free ( h->unique ) ;
    free ( h->bases->entries ) ;
    IntegerArray1D_Deallocate ( &h->cfp ) ;
    IntegerArray1D_Deallocate ( &h->fc ) ;
    free ( h->bases ) ;
    Coordinates3_Deallocate ( &h->coordinates ) ;
    IntegerArray1D_Deallocate ( &h->atomicNumbers ) ;
    free ( h ) ;
}
