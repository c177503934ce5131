/*
 * sla_blacs1x1.h -- the BLACS and TOOLS routines a ScaLAPACK LU program calls
 * around PDGESV, for a single-process 1 x 1 grid (libsla_blacs1x1.so).
 *
 * They exist so that a Fortran program such as the reference's
 * EXAMPLE/pdscaex.f:55-82 (BLACS_PINFO, BLACS_GET, BLACS_GRIDINIT,
 * DESCINIT, NUMROC, BLACS_GRIDINFO, BLACS_GRIDEXIT, BLACS_EXIT) links and runs
 * against libscalapack_amd.so without MPI.  A program that links a real
 * BLACS/ScaLAPACK ahead of this library gets the real ones instead.
 *
 * Semantics follow the reference on one process: one context is handed out
 * per BLACS_GRIDINIT of a 1 x 1 grid; any other shape yields a context whose
 * BLACS_GRIDINFO reports NPROW = -1 (the "illegal context" case the
 * computational routines turn into INFO = -(100*pos + CTXT_)).
 */
#ifndef SLA_BLACS1X1_H
#define SLA_BLACS1X1_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BLACS/SRC/blacs_pinfo_.c:14-27 */
void blacs_pinfo_(int *mypnum, int *nprocs);
/* BLACS/SRC/blacs_get_.c (WHAT = 0: default system context) */
void blacs_get_(const int *icontxt, const int *what, int *val);
/* BLACS/SRC/blacs_grid_.c (BLACS_GRIDINIT) */
void blacs_gridinit_(int *icontxt, const char *order, const int *nprow, const int *npcol,
                     size_t order_len);
/* BLACS/SRC/blacs_info_.c (BLACS_GRIDINFO) */
void blacs_gridinfo_(const int *icontxt, int *nprow, int *npcol, int *myrow, int *mycol);
void blacs_gridexit_(const int *icontxt);
void blacs_exit_(const int *notdone);
/* BLACS/SRC/blacs_abort_.c:15-19 */
void blacs_abort_(const int *icontxt, const int *errornum);
void blacs_barrier_(const int *icontxt, const char *scope, size_t scope_len);

/* TOOLS/numroc.f:53-76 */
int numroc_(const int *n, const int *nb, const int *iproc, const int *isrcproc,
            const int *nprocs);
/* TOOLS/descinit.f:152-186 */
void descinit_(int *desc, const int *m, const int *n, const int *mb, const int *nb,
               const int *irsrc, const int *icsrc, const int *ictxt, const int *lld, int *info);
/* TOOLS/indxg2p.f:48, TOOLS/indxg2l.f:48, TOOLS/indxl2g.f:49-50 */
int indxg2p_(const int *indxglob, const int *nb, const int *iproc, const int *isrcproc,
             const int *nprocs);
int indxg2l_(const int *indxglob, const int *nb, const int *iproc, const int *isrcproc,
             const int *nprocs);
int indxl2g_(const int *indxloc, const int *nb, const int *iproc, const int *isrcproc,
             const int *nprocs);
/* TOOLS/iceil.f:31 */
int iceil_(const int *inum, const int *idenom);

/* PDLAWRITE (TOOLS/pdlawrite.f:1-2): sub(A) to FILNAM as the reference
 * writes it -- a list-directed header " M N", then the values down the
 * columns one per record in D30.18 (pdlawrite.f:217).  IRWRIT / ICWRIT and
 * WORK are not needed on a 1 x 1 grid. */
void pdlawrite_(const char *filnam, const int *m, const int *n, const double *a, const int *ia,
                const int *ja, const int *desca, const int *irwrit, const int *icwrit, double *work,
                size_t filnam_len);
/* PDLAREAD (TOOLS/pdlaread.f:1): the matrix of FILNAM ("M N" header, values
 * down the columns, one per record, D or E exponents) into A(1:M, 1:N);
 * aborts when M > M_A or N > N_A as the reference does. */
void pdlaread_(const char *filnam, double *a, const int *desca, const int *irread, const int *icread,
               double *work, size_t filnam_len);

#ifdef __cplusplus
}
#endif

#endif /* SLA_BLACS1X1_H */
