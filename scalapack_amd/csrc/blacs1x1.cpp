// blacs1x1.cpp -- single-process 1 x 1 BLACS grid and the TOOLS index helpers
// (libsla_blacs1x1.so).  Declarations and reference citations are in
// include/sla_blacs1x1.h.
#include "sla_blacs1x1.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

std::mutex g_mu;
std::vector<int> g_ctx;  // per handed-out context: 1 = live 1 x 1 grid, 0 = exited/illegal

int mymax(int a, int b) { return a > b ? a : b; }

// Fortran CHARACTER*(*) argument -> file name (trailing blanks dropped).
std::string fortran_string(const char *s, size_t len)
{
    std::string out(s, len);
    while (!out.empty() && (out.back() == ' ' || out.back() == '\0')) out.pop_back();
    return out;
}

// One value in the Fortran edit descriptor D30.18 as the reference's runtime
// writes it (TOOLS/pdlawrite.f:217): 0.d1..d18 x 10^e, exponent "D+ee" when
// |e| <= 99 and "+eee" without the letter otherwise, right-justified in 30.
std::string format_d30_18(double v)
{
    char body[64];
    if (std::isnan(v)) {
        std::snprintf(body, sizeof body, "NaN");
    } else if (std::isinf(v)) {
        std::snprintf(body, sizeof body, "%s", v < 0 ? "-Infinity" : "Infinity");
    } else {
        char e[64];
        std::snprintf(e, sizeof e, "%.17e", v);  // [-]d.ddddddddddddddddde[+-]XX, 18 digits
        const bool neg = e[0] == '-';
        const char *d = e + (neg ? 1 : 0);
        char digits[19];
        digits[0] = d[0];
        std::memcpy(digits + 1, d + 2, 17);
        digits[18] = '\0';
        int ex = std::atoi(std::strchr(d, 'e') + 1) + 1;
        if (v == 0.0) ex = 0;
        const int ax = ex < 0 ? -ex : ex;
        if (ax <= 99)
            std::snprintf(body, sizeof body, "%s0.%sD%c%02d", neg ? "-" : "", digits, ex < 0 ? '-' : '+', ax);
        else
            std::snprintf(body, sizeof body, "%s0.%s%c%03d", neg ? "-" : "", digits, ex < 0 ? '-' : '+', ax);
    }
    char out[64];
    std::snprintf(out, sizeof out, "%30s", body);
    return out;
}

// The first list-directed value of a record: separators blank or comma, a
// D exponent read as E.
bool parse_first_value(const std::string &line, double *v)
{
    size_t i = 0;
    while (i < line.size() && (std::isspace((unsigned char)line[i]) || line[i] == ',')) i++;
    if (i == line.size()) return false;
    size_t j = i;
    while (j < line.size() && !std::isspace((unsigned char)line[j]) && line[j] != ',') j++;
    std::string tok = line.substr(i, j - i);
    for (char &c : tok)
        if (c == 'D' || c == 'd' || c == 'q' || c == 'Q') c = 'E';
    // Fortran also takes an exponent without its letter ("0.123-299", as
    // D30.18 writes three-digit exponents): a sign after the first digit
    if (tok.find_first_of("Ee") == std::string::npos) {
        const size_t k = tok.find_last_of("+-");
        if (k != std::string::npos && k > 0 && std::isdigit((unsigned char)tok[k - 1]))
            tok.insert(k, 1, 'E');
    }
    char *end = nullptr;
    *v = std::strtod(tok.c_str(), &end);
    return end && *end == '\0';
}
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
    // an abort never reports success to the operating system
    std::exit(*errornum != 0 ? *errornum : 1);
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

void pdlawrite_(const char *filnam, const int *m, const int *n, const double *a, const int *ia,
                const int *ja, const int *desca, const int *irwrit, const int *icwrit, double *,
                size_t filnam_len)
{
    (void)irwrit;
    (void)icwrit;
    const std::string name = fortran_string(filnam, filnam_len);
    FILE *f = std::fopen(name.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "PDLAWRITE: cannot open %s\n", name.c_str());
        return;
    }
    std::fprintf(f, " %d %d\n", *m, *n);
    const int lld = desca[8];
    for (int j = 0; j < *n; j++)
        for (int i = 0; i < *m; i++) {
            const double v = a[(size_t)(*ia - 1 + i) + (size_t)(*ja - 1 + j) * lld];
            std::fprintf(f, "%s\n", format_d30_18(v).c_str());
        }
    std::fclose(f);
}

void pdlaread_(const char *filnam, double *a, const int *desca, const int *irread, const int *icread,
               double *, size_t filnam_len)
{
    (void)irread;
    (void)icread;
    const std::string name = fortran_string(filnam, filnam_len);
    FILE *f = std::fopen(name.c_str(), "r");
    if (!f) {
        std::fprintf(stderr, "PDLAREAD: cannot open %s\n", name.c_str());
        int ctx = desca[1], code = 0;
        blacs_abort_(&ctx, &code);
        return;
    }
    std::vector<std::string> lines;
    char buf[4096];
    while (std::fgets(buf, sizeof buf, f)) lines.emplace_back(buf);
    std::fclose(f);
    // header: the first two integers, over as many records as they take
    int hdr[2] = {0, 0}, got = 0;
    size_t li = 0;
    for (; li < lines.size() && got < 2; li++) {
        const char *p = lines[li].c_str();
        char *end = nullptr;
        while (got < 2) {
            while (*p && (std::isspace((unsigned char)*p) || *p == ',')) p++;
            if (!*p) break;
            const long v = std::strtol(p, &end, 10);
            if (end == p) break;
            hdr[got++] = (int)v;
            p = end;
        }
    }
    const int M = hdr[0], N = hdr[1];
    if (M <= 0 || N <= 0) return;  // pdlaread.f: quick return
    if (M > desca[2] || N > desca[3]) {  // pdlaread.f: too big, abort
        std::printf(" PDLAREAD: Matrix too big to fit in\n Abort ...\n");
        std::fflush(stdout);
        int ctx = desca[1], code = 0;
        blacs_abort_(&ctx, &code);
        return;
    }
    const int lld = desca[8];
    for (int j = 0; j < N; j++)
        for (int i = 0; i < M; i++) {
            double v = 0.0;
            while (li < lines.size() && !parse_first_value(lines[li], &v)) li++;
            if (li >= lines.size()) {
                std::fprintf(stderr, "PDLAREAD: %s ends before %d x %d values\n", name.c_str(), M, N);
                int ctx = desca[1], code = 0;
                blacs_abort_(&ctx, &code);
                return;
            }
            li++;
            a[(size_t)i + (size_t)j * lld] = v;
        }
}

}  // extern "C"
