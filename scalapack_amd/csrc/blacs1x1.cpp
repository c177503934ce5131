// blacs1x1.cpp -- single-process 1 x 1 BLACS grid and the TOOLS index helpers
// (libsla_blacs1x1.so).  Declarations and reference citations are in
// include/sla_blacs1x1.h.
#include "sla_blacs1x1.h"

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace {

std::mutex g_mu;
std::vector<int> g_ctx;  // per handed-out context: 1 = live 1 x 1 grid, 0 = exited/illegal

int mymax(int a, int b) { return a > b ? a : b; }
int mymin(int a, int b) { return a < b ? a : b; }

}  // namespace

extern "C" {

void blacs_pinfo_(int *mypnum, int *nprocs)
{
    *mypnum = 0;
    *nprocs = 1;
}

void blacs_get_(const int *, const int *, int *val) { *val = 0; }

void blacs_gridinit_(int *icontxt, const char *, const int *nprow, const int *npcol, size_t)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_ctx.push_back((*nprow == 1 && *npcol == 1) ? 1 : 0);
    if (!(*nprow == 1 && *npcol == 1))
        std::fprintf(stderr, "BLACS_GRIDINIT: a %d x %d grid needs %d processes; this "
                             "single-process BLACS provides 1 x 1 only\n",
                     *nprow, *npcol, *nprow * *npcol);
    *icontxt = (int)g_ctx.size() - 1;
}

void blacs_gridinfo_(const int *icontxt, int *nprow, int *npcol, int *myrow, int *mycol)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const int c = *icontxt;
    if (c >= 0 && c < (int)g_ctx.size() && g_ctx[c] == 1) {
        *nprow = *npcol = 1;
        *myrow = *mycol = 0;
    } else {
        *nprow = *npcol = *myrow = *mycol = -1;
    }
}

void blacs_gridexit_(const int *icontxt)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const int c = *icontxt;
    if (c >= 0 && c < (int)g_ctx.size()) g_ctx[c] = 0;
}

void blacs_exit_(const int *) {}

void blacs_abort_(const int *, const int *errornum)
{
    std::fprintf(stderr, "{0,0}, pnum=0, Contxt=0, killed other procs, exiting with error #%d.\n",
                 *errornum);
    std::fflush(stderr);
    std::exit(*errornum);
}

void blacs_barrier_(const int *, const char *, size_t) {}

int numroc_(const int *n, const int *nb, const int *iproc, const int *isrcproc, const int *nprocs)
{
    const int mydist = (*nprocs + *iproc - *isrcproc) % *nprocs;
    const int nblocks = *n / *nb;
    int r = (nblocks / *nprocs) * *nb;
    const int extrablks = nblocks % *nprocs;
    if (mydist < extrablks) r += *nb;
    else if (mydist == extrablks) r += *n % *nb;
    return r;
}

void descinit_(int *desc, const int *m, const int *n, const int *mb, const int *nb,
               const int *irsrc, const int *icsrc, const int *ictxt, const int *lld, int *info)
{
    int nprow, npcol, myrow, mycol;
    blacs_gridinfo_(ictxt, &nprow, &npcol, &myrow, &mycol);
    *info = 0;
    if (*m < 0) *info = -2;
    else if (*n < 0) *info = -3;
    else if (*mb < 1) *info = -4;
    else if (*nb < 1) *info = -5;
    else if (*irsrc < 0 || *irsrc >= nprow) *info = -6;
    else if (*icsrc < 0 || *icsrc >= npcol) *info = -7;
    else if (nprow == -1) *info = -8;
    else {
        int loc = numroc_(m, mb, &myrow, irsrc, &nprow);
        if (*lld < mymax(1, loc)) *info = -9;
    }
    if (*info != 0)
        std::printf("{%5d,%5d}:  On entry to DESCINIT parameter number %4d had an illegal value\n",
                    myrow, mycol, -*info);
    desc[0] = 1;
    desc[2] = mymax(0, *m);
    desc[3] = mymax(0, *n);
    desc[4] = mymax(1, *mb);
    desc[5] = mymax(1, *nb);
    desc[6] = mymax(0, mymin(*irsrc, nprow - 1));
    desc[7] = mymax(0, mymin(*icsrc, npcol - 1));
    desc[1] = *ictxt;
    int m0 = desc[2], mb0 = desc[4], rs = desc[6];
    desc[8] = mymax(*lld, mymax(1, nprow > 0 ? numroc_(&m0, &mb0, &myrow, &rs, &nprow) : 1));
}

int indxg2p_(const int *indxglob, const int *nb, const int *, const int *isrcproc, const int *nprocs)
{
    return (*isrcproc + (*indxglob - 1) / *nb) % *nprocs;
}

int indxg2l_(const int *indxglob, const int *nb, const int *, const int *, const int *nprocs)
{
    return *nb * ((*indxglob - 1) / (*nb * *nprocs)) + (*indxglob - 1) % *nb + 1;
}

int indxl2g_(const int *indxloc, const int *nb, const int *iproc, const int *isrcproc,
             const int *nprocs)
{
    return *nprocs * *nb * ((*indxloc - 1) / *nb) + (*indxloc - 1) % *nb +
           ((*nprocs + *iproc - *isrcproc) % *nprocs) * *nb + 1;
}

int iceil_(const int *inum, const int *idenom) { return (*inum + *idenom - 1) / *idenom; }

}  // extern "C"
