/*
 * orc.c -- CPU oracle for Canard's compressible ILES time step.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h).  Restates the reference's FP64
 * arithmetic in the reference's own evaluation order (left-to-right sums,
 * products rounded separately; build with -ffp-contract=off), so the GPU
 * product can be compared against it to a tight tolerance.  Each function
 * cites the reference lines it follows (paths relative to the reference
 * repository root).
 */
#include "orc.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants: src/mo_parameters.f90:24-79 -------------------------- */
#define PI 3.141592653589793
static const double zero = 0.0, half = 0.5, one = 1.0, two = 2.0;
static const double sml = 1.0e-8;
enum { BC_NONREF = 10, BC_WALLI = 20, BC_WALLV = 25, BC_ICURV = 30,
       BC_ISTR = 35, BC_ISUB = 40, BC_PER = 45 };
#define GAM 1.4
static double gamm1, hamm1, hamhamm1, gamm1prndtli, twothirds;
static const double alpha = 4.0 / 9.0, beta = 1.0 / 36.0;
static const double aa = 20.0 / 27.0, ab = 25.0 / 216.0;
static const double alpha10 = 0.09166666665902128, alpha12 = 1.3500000000438646;
static const double beta13 = 0.2666666666794667, a10 = -0.3506944444280639;
static const double a12 = 0.8249999999974599, a13 = 0.7444444444773477;
static const double a14 = 0.014583333334139971;
static const double alpha01 = 8.000000000449464, beta02 = 6.000000000542;
static const double a01 = -6.666666667794732, a02 = 9.000000000052879;
static const double a03 = 1.333333333178555, a04 = -0.0833333333455545;

#define LMD 11 /* mo_numerics.f90:21 */
#define MBCI 4 /* mo_gcbc.f90:22 */
#define LJPL 100 /* mo_domdcomp.f90:15 */

static char g_err[256];
const char *orc_error(void) { return g_err; }
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

static void init_constants(void)
{
    gamm1 = GAM - one;
    hamm1 = one / gamm1;
    hamhamm1 = (one / GAM) * hamm1;
    gamm1prndtli = one / (gamm1 * 0.71);
    twothirds = two / 3.0;
}

/* ---- data ------------------------------------------------------------- */
typedef struct {
    int lxi, let, lze;
    int nbc[3][2], mcd[3][2];
    long npt;
    int ijk[3][3];   /* ijk[nn][r] = ijk(r+1,nn+1) of mo_domdcomp.f90:218-226 */
    int nbsize[3];
    int ndf[3][2][2]; /* [nn][side][nt] */
    double *xl[2][3], *xu[2][3];   /* [nt][nn]: (n,2) and (n,3) row-major */
    double *send[2][3];            /* [nt][nn]: ((side*NS+slot)*nb + jk)  */
    double *recv[2][3];            /* [nt][nn]: (((m*2+side)*NS+slot)*nb+jk) */
    double *drva[3], *drvb[3];     /* [nn]: ((side*5+m)*nb + jk)          */
    double *xyz, *xim, *etm, *zem, *yaco, *cm[3];
    double *qa, *qo, *de, *p, *ss, *rrv, *rrf, *tau, *avg;
    /* gcbc */
    long lq; double *sbcc; int *nrr;
    /* sponge */
    long lsz; long *lcsz; double *asz, *bsz;
    /* junctions (mo_domdcomp.f90:269-437) */
    int njp[8], jpcd[8][LJPL + 1];
    int njl[12], jlcd[12][LJPL + 1];
} blk_t;

struct orc_case {
    orc_params prm;
    int nblk;
    blk_t *b;
    double alphf, betf, fa, fb, fc;
    double albef[3][5][2];   /* albef(-2:2,0:2,0:1) -> [row][off+2][ns] */
    double fbc[3][5];        /* fbc(0:4,0:2) -> [row][k]                */
    double pbci[2][2][12], pbco[2][2][12], pbcot[2][2];  /* [nt][row][i] */
    double sbci[MBCI];
    double reoo, srefoo, srefp1dre, uoo[3], umf[3], dudtmf[3];
    int n, ndt; double dt, dts, dte, timo;
    /* junction tables shared by all ranks */
    int lpp, lqq;
    int *imjp, *jptag, *jjp;  /* jjp[(i)*(LJPL+1)+k] */
    int *imjl, *jltag, *jjl;
};

static inline long indx3(int i, int j, int k, int nn, int lxi, int let)
{   /* mo_utils.f90:54-69 (nn = 1..3) */
    switch (nn) {
    case 1: return ((long)k * (let + 1) + j) * (lxi + 1) + i;
    case 2: return ((long)j * (let + 1) + i) * (lxi + 1) + k;
    default: return ((long)i * (let + 1) + k) * (lxi + 1) + j;
    }
}
#define IDX(B, i, j, k, nn) indx3((i), (j), (k), (nn), (B)->lxi, (B)->let)

static void *xcalloc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "orc: out of memory\n"); abort(); }
    return p;
}

/* ---- mtrxi: Gauss-Jordan inverse with partial pivoting ----------------
 * mo_utils.f90:15-50.  Matrices are n x n, column-major a[i + n*j]. */
static void mtrxi(const double *ax, double *sx, int n)
{
    double *rx = xcalloc((size_t)n * n, sizeof(double));
    double *temp = xcalloc(n, sizeof(double));
    int *ipvt = xcalloc(n, sizeof(int));
    memcpy(rx, ax, sizeof(double) * n * n);
    for (int i = 0; i < n; i++) ipvt[i] = i;
    for (int ii = 0; ii < n; ii++) {
        int mk = ii; double best = fabs(rx[ii + n * ii]);
        for (int r = ii + 1; r < n; r++)      /* maxloc: first maximum */
            if (fabs(rx[r + n * ii]) > best) { best = fabs(rx[r + n * ii]); mk = r; }
        if (mk != ii) {
            int t = ipvt[mk]; ipvt[mk] = ipvt[ii]; ipvt[ii] = t;
            for (int c = 0; c < n; c++) {
                double x = rx[mk + n * c]; rx[mk + n * c] = rx[ii + n * c]; rx[ii + n * c] = x;
            }
        }
        double ra0k = one / rx[ii + n * ii];
        for (int r = 0; r < n; r++) temp[r] = rx[r + n * ii];
        for (int jj = 0; jj < n; jj++) {
            double ra1k = ra0k * rx[ii + n * jj];
            for (int r = 0; r < n; r++) rx[r + n * jj] = rx[r + n * jj] - ra1k * temp[r];
            rx[ii + n * jj] = ra1k;
        }
        for (int r = 0; r < n; r++) rx[r + n * ii] = -ra0k * temp[r];
        rx[ii + n * ii] = ra0k;
    }
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++) sx[r + n * ipvt[c]] = rx[r + n * c];
    free(rx); free(temp); free(ipvt);
}

/* C = A*B, n x n column-major, k-ordered sums */
static void matmul(const double *A, const double *B, double *C, int n)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s = s + A[i + n * k] * B[k + n * j];
            C[i + n * j] = s;
        }
}

/* ---- filter coefficients: mo_numerics.f90:306-326 (fcint) -------------- */
static void fcint(double fltk, double fltr, double *alphz, double *betz,
                  double *za, double *zb, double *zc)
{
    double c1 = cos(fltk), c2 = cos(2 * fltk), c3 = cos(3 * fltk);
    double fctrk = 1.0 / (((30.0 + (5.0 * (7.0 - 16.0 * fltr)) * c1)
                           + (2.0 * (1.0 + 8.0 * fltr)) * c2) - 3.0 * c3);
    *alphz = fctrk * (((20.0 * (2.0 * fltr - 1.0) - 30.0 * c1)
                       + (12.0 * (2.0 * fltr - 1.0)) * c2) - 2.0 * c3);
    *betz = (half * fctrk) * (((2.0 * (13.0 - 8.0 * fltr) + (33.0 - 48.0 * fltr) * c1)
                               + 6.0 * c2) - c3);
    double ch = cos(half * fltk), ch2 = ch * ch;
    *za = ((60.0 * (1.0 - fltr)) * (ch2 * ch2)) * fctrk;
    *zb = (-2.0 * *za) / 5.0;
    *zc = *za / 15.0;
}

/* ---- boundary filter coefficients: mo_numerics.f90:253-302 (fcbcm) ---- */
static void fcbcm(orc_case *C, double fltk, double fltrbc)
{
    double alphz, betz, za, zb, zc;
    double aok = log(fltrbc);
    fcint(fltk, half, &alphz, &betz, &za, &zb, &zc);
    double fctrk = one / ((one + alphz * fltrbc) + betz * (fltrbc * fltrbc));
    double r0[5] = { zero, zero, one, alphz * fctrk, betz * fctrk };
    for (int o = 0; o < 5; o++) C->albef[0][o][0] = r0[o];
    double resk = ((fltrbc - 1) * ((za + zc) + (fltrbc + 1) * (zb + fltrbc * zc))) / aok;
    C->fbc[0][0] = (za - (5 * resk) / 3) * fctrk;
    C->fbc[0][1] = (zb + (10 * resk) / 21) * fctrk;
    C->fbc[0][2] = (zc - (5 * resk) / 42) * fctrk;
    C->fbc[0][3] = ((5 * resk) / 252) * fctrk;
    C->fbc[0][4] = (-resk / 630) * fctrk;

    double r1[5] = { zero, alphz + betz * fltrbc, one, alphz, betz };
    for (int o = 0; o < 5; o++) C->albef[1][o][0] = r1[o];
    resk = ((fltrbc - 1) * (zb + zc * (fltrbc + 1))) / aok;
    C->fbc[1][0] = ((za + zb) + zc) + (1627 * resk) / 1260;
    C->fbc[1][1] = za + (10 * resk) / 21;
    C->fbc[1][2] = zb - (5 * resk) / 42;
    C->fbc[1][3] = zc + (5 * resk) / 252;
    C->fbc[1][4] = -resk / 630;

    double r2[5] = { betz, alphz, one, alphz, betz };
    for (int o = 0; o < 5; o++) C->albef[2][o][0] = r2[o];
    resk = (zc * (fltrbc - 1)) / aok;
    C->fbc[2][0] = (zb + zc) + (1627 * resk) / 1260;
    C->fbc[2][1] = za - (5 * resk) / 3;
    C->fbc[2][2] = za - (5 * resk) / 42;
    C->fbc[2][3] = zb + (5 * resk) / 252;
    C->fbc[2][4] = zc - resk / 630;
}

/* ---- quasi-disjoint interface tables: mo_numerics.f90:330-425 (sbcco) -- */
static void sbcco(orc_case *C)
{
    const int n = 2 * (LMD + 1); /* 24 */
    double *ax = xcalloc(n * n, sizeof(double)), *bx = xcalloc(n * n, sizeof(double));
    double *rx = xcalloc(n * n, sizeof(double)), *sx = xcalloc(n * n, sizeof(double));
    double *tx = xcalloc(n * n, sizeof(double));
#define AX(i, j) ax[((i)-1) + n * ((j)-1)]
#define BX(i, j) bx[((i)-1) + n * ((j)-1)]
#define RX(i, j) rx[((i)-1) + n * ((j)-1)]
    for (int nt = 0; nt < 2; nt++) {
        memset(ax, 0, sizeof(double) * n * n);
        memset(bx, 0, sizeof(double) * n * n);
        const int is = 1, ie = n;
        if (nt == 0) {
            AX(is, is) = one; AX(is, is + 1) = alpha01; AX(is, is + 2) = beta02;
            double b0[5] = { -(((a01 + a02) + a03) + a04), a01, a02, a03, a04 };
            for (int c = 0; c < 5; c++) BX(is, is + c) = b0[c];
            AX(is + 1, is) = alpha10; AX(is + 1, is + 1) = one;
            AX(is + 1, is + 2) = alpha12; AX(is + 1, is + 3) = beta13;
            double b1[5] = { a10, -(((a10 + a12) + a13) + a14), a12, a13, a14 };
            for (int c = 0; c < 5; c++) BX(is + 1, is + c) = b1[c];
            for (int i = is + 2; i <= ie - 2; i++) {
                double ar[5] = { beta, alpha, one, alpha, beta };
                double br[5] = { -ab, -aa, zero, aa, ab };
                for (int c = 0; c < 5; c++) { AX(i, i - 2 + c) = ar[c]; BX(i, i - 2 + c) = br[c]; }
            }
            for (int c = 0; c < 4; c++) AX(ie - 1, ie - c) = AX(is + 1, is + c);
            for (int c = 0; c < 5; c++) BX(ie - 1, ie - c) = -BX(is + 1, is + c);
            for (int c = 0; c < 3; c++) AX(ie, ie - c) = AX(is, is + c);
            for (int c = 0; c < 5; c++) BX(ie, ie - c) = -BX(is, is + c);
        } else {
            for (int c = 0; c < 3; c++) AX(is, is + c) = C->albef[0][2 + c][0];
            { int cols[5] = { 1, 2, 3, 4, 5 };
              for (int c = 0; c < 5; c++) BX(is, is + cols[c]) = C->fbc[0][c];
              double s = 0; for (int c = 0; c < 5; c++) s = s + C->fbc[0][c];
              BX(is, is) = -s; }
            for (int c = 0; c < 4; c++) AX(is + 1, is + c) = C->albef[1][1 + c][0];
            { int cols[5] = { 0, 2, 3, 4, 5 };
              for (int c = 0; c < 5; c++) BX(is + 1, is + cols[c]) = C->fbc[1][c];
              double s = 0; for (int c = 0; c < 5; c++) s = s + C->fbc[1][c];
              BX(is + 1, is + 1) = -s; }
            for (int c = 0; c < 5; c++) AX(is + 2, is + c) = C->albef[2][c][0];
            { int cols[5] = { 0, 1, 3, 4, 5 };
              for (int c = 0; c < 5; c++) BX(is + 2, is + cols[c]) = C->fbc[2][c];
              double s = 0; for (int c = 0; c < 5; c++) s = s + C->fbc[2][c];
              BX(is + 2, is + 2) = -s; }
            for (int i = is + 3; i <= ie - 3; i++) {
                double ar[5] = { C->betf, C->alphf, one, C->alphf, C->betf };
                double br[7] = { C->fc, C->fb, C->fa, -2 * ((C->fa + C->fb) + C->fc),
                                 C->fa, C->fb, C->fc };
                for (int c = 0; c < 5; c++) AX(i, i - 2 + c) = ar[c];
                for (int c = 0; c < 7; c++) BX(i, i - 3 + c) = br[c];
            }
            for (int c = 0; c < 5; c++) AX(ie - 2, ie - c) = AX(is + 2, is + c);
            for (int c = 0; c < 6; c++) BX(ie - 2, ie - c) = BX(is + 2, is + c);
            for (int c = 0; c < 4; c++) AX(ie - 1, ie - c) = AX(is + 1, is + c);
            for (int c = 0; c < 6; c++) BX(ie - 1, ie - c) = BX(is + 1, is + c);
            for (int c = 0; c < 3; c++) AX(ie, ie - c) = AX(is, is + c);
            for (int c = 0; c < 6; c++) BX(ie, ie - c) = BX(is, is + c);
        }
        mtrxi(ax, sx, n);
        memcpy(rx, ax, sizeof(double) * n * n);
        int ii = ie / 2 - 1;
        RX(ii, ii + 2) = 0; RX(ii + 1, ii + 2) = 0; RX(ii + 1, ii + 3) = 0;
        ii = ie / 2 + 2;
        RX(ii, ii - 2) = 0; RX(ii - 1, ii - 2) = 0; RX(ii - 1, ii - 3) = 0;
        matmul(sx, bx, tx, n);
        matmul(rx, tx, ax, n);
        for (int row = 0; row < 2; row++) {
            int r = ie / 2 + 1 + row;
            for (int i = 0; i <= LMD; i++) {
                C->pbco[nt][row][i] = AX(r, LMD + 1 - i);
                C->pbci[nt][row][i] = AX(r, LMD + 2 + i);
            }
        }
    }
#undef AX
#undef BX
#undef RX
    free(ax); free(bx); free(rx); free(sx); free(tx);
    for (int nt = 0; nt < 2; nt++)           /* mo_numerics.f90:127-132 */
        for (int r = 0; r < 2; r++) {
            double s = 0; for (int i = 0; i <= LMD; i++) s = s + C->pbco[nt][r][i];
            C->pbcot[nt][r] = s;
        }
}

/* ---- init_extracoeff_bounds: mo_numerics.f90:114-134 ------------------ */
static void init_extra(orc_case *C)
{
    double fltk = PI * C->prm.fltk;     /* mo_numerics.f90:107 */
    double za, zb, zc;
    fcbcm(C, fltk, C->prm.fltrbc);
    fcint(fltk, half, &C->alphf, &C->betf, &za, &zb, &zc);
    C->fa = za; C->fb = zb; C->fc = zc;
    double r0[5] = { zero, zero, one, C->alphf, C->betf };
    double r1[5] = { zero, C->alphf, one, C->alphf, C->betf };
    double r2[5] = { C->betf, C->alphf, one, C->alphf, C->betf };
    for (int o = 0; o < 5; o++) {
        C->albef[0][o][1] = r0[o]; C->albef[1][o][1] = r1[o]; C->albef[2][o][1] = r2[o];
    }
    sbcco(C);
}

/* ---- penta: LU factors of the pentadiagonal LHS, mo_numerics.f90:179-249
 * xl (n,2) and xu (n,3), row-major [i*2+c], [i*3+c]; ns/ne select the
 * boundary-row family (0: one-sided, 1: quasi-disjoint). */
static void penta(orc_case *C, double *xl, double *xu, int n, int ns, int ne, int nt)
{
    double albe[3][5][2], alpho, beto;
    if (nt == 0) {
        double r00[5] = { zero, zero, one, alpha01, beta02 };
        double r10[5] = { zero, alpha10, one, alpha12, beta13 };
        double r20[5] = { beta, alpha, one, alpha, beta };
        double r01[5] = { zero, zero, one, alpha, beta };
        double r11[5] = { zero, alpha, one, alpha, beta };
        for (int o = 0; o < 5; o++) {
            albe[0][o][0] = r00[o]; albe[1][o][0] = r10[o]; albe[2][o][0] = r20[o];
            albe[0][o][1] = r01[o]; albe[1][o][1] = r11[o]; albe[2][o][1] = r20[o];
        }
        alpho = alpha; beto = beta;
    } else {
        memcpy(albe, C->albef, sizeof albe);
        alpho = C->alphf; beto = C->betf;
    }
#define A(off, row, s) albe[row][(off) + 2][s]
#define XL(i, c) xl[(i)*2 + (c)-1]
#define XU(i, c) xu[(i)*3 + (c)-1]
    int is = 0, ie = n - 1, i;
    for (i = is; i <= ie; i++) { XL(i, 1) = XL(i, 2) = one; XU(i, 1) = XU(i, 2) = XU(i, 3) = one; }
    i = is;
    XU(i, 1) = one; XU(i, 2) = A(1, 0, ns); XU(i, 3) = A(2, 0, ns);
    i = is + 1;
    XL(i, 2) = A(-1, 1, ns) * XU(i - 1, 1);
    XU(i, 1) = one / (one - XU(i - 1, 2) * XL(i, 2));
    XU(i, 2) = A(1, 1, ns) - XU(i - 1, 3) * XL(i, 2);
    XU(i, 3) = A(2, 1, ns);
    i = is + 2;
    XL(i, 1) = A(-2, 2, ns) * XU(i - 2, 1);
    XL(i, 2) = (A(-1, 2, ns) - XU(i - 2, 2) * XL(i, 1)) * XU(i - 1, 1);
    XU(i, 1) = one / ((one - XU(i - 2, 3) * XL(i, 1)) - XU(i - 1, 2) * XL(i, 2));
    XU(i, 2) = A(1, 2, ns) - XU(i - 1, 3) * XL(i, 2);
    XU(i, 3) = A(2, 2, ns);
    for (i = is + 3; i <= ie - 3; i++) {
        XL(i, 1) = beto * XU(i - 2, 1);
        XL(i, 2) = (alpho - XU(i - 2, 2) * XL(i, 1)) * XU(i - 1, 1);
        XU(i, 1) = one / ((one - XU(i - 2, 3) * XL(i, 1)) - XU(i - 1, 2) * XL(i, 2));
        XU(i, 2) = alpho - XU(i - 1, 3) * XL(i, 2);
        XU(i, 3) = beto;
    }
    i = ie - 2;
    XL(i, 1) = A(2, 2, ne) * XU(i - 2, 1);
    XL(i, 2) = (A(1, 2, ne) - XU(i - 2, 2) * XL(i, 1)) * XU(i - 1, 1);
    XU(i, 1) = one / ((one - XU(i - 2, 3) * XL(i, 1)) - XU(i - 1, 2) * XL(i, 2));
    XU(i, 2) = A(-1, 2, ne) - XU(i - 1, 3) * XL(i, 2);
    XU(i, 3) = A(-2, 2, ne);
    i = ie - 1;
    XL(i, 1) = A(2, 1, ne) * XU(i - 2, 1);
    XL(i, 2) = (A(1, 1, ne) - XU(i - 2, 2) * XL(i, 1)) * XU(i - 1, 1);
    XU(i, 1) = one / ((one - XU(i - 2, 3) * XL(i, 1)) - XU(i - 1, 2) * XL(i, 2));
    XU(i, 2) = A(-1, 1, ne) - XU(i - 1, 3) * XL(i, 2);
    i = ie;
    XL(i, 1) = A(2, 0, ne) * XU(i - 2, 1);
    XL(i, 2) = (A(1, 0, ne) - XU(i - 2, 2) * XL(i, 1)) * XU(i - 1, 1);
    XU(i, 1) = one / ((one - XU(i - 2, 3) * XL(i, 1)) - XU(i - 1, 2) * XL(i, 2));
    for (i = is; i <= ie; i++) { XU(i, 2) = XU(i, 2) * XU(i, 1); XU(i, 3) = XU(i, 3) * XU(i, 1); }
#undef A
#undef XL
#undef XU
}

/* ---- partial sums for quasi-disjoint faces: mo_numerics.f90:429-705 ----
 * fld[b][nn] is the array whose values block b sends for direction nn.
 * Face exchange: the receiving face (nn,side) of block X gets what the
 * partner P = mcd(nn,side) computed on its face (nn,1-side); with every
 * rank posting in the same direction order this is what the reference's
 * tag matching delivers (mo_numerics.f90:512-522). */
static void mpigo(orc_case *C, int nt, const double *const *fld3 /* [b*3+nn] */,
                  int n45, int m)
{
    const int NS = 2 + nt;
    for (int bi = 0; bi < C->nblk; bi++) {
        blk_t *B = &C->b[bi];
        for (int nn = 0; nn < 3; nn++) {
            const double *f = fld3[bi * 3 + nn];
            int nb = B->nbsize[nn];
            for (int ip = 0; ip < 2; ip++) {
                if (B->ndf[nn][ip][nt] != 1) continue;
                int istart = ip * B->ijk[nn][0], iend = 1 - 2 * ip;
                double ra0k = 0; int iik = 1;
                switch (B->nbc[nn][ip]) {
                case BC_ISTR: ra0k = zero; iik = 1; break;
                case BC_ISUB: ra0k = zero; iik = 0; break;
                case BC_PER:  ra0k = n45;  iik = 1; break;
                }
                double *snd = B->send[nt][nn];
                for (int k = 0; k <= B->ijk[nn][2]; k++)
                    for (int j = 0; j <= B->ijk[nn][1]; j++) {
                        int jk = k * (B->ijk[nn][1] + 1) + j;
                        double resk = ra0k * f[IDX(B, istart, j, k, nn + 1)];
                        double sap[LMD + 1];
                        for (int i = 0; i <= LMD; i++)
                            sap[i] = f[IDX(B, istart + iend * (i + iik), j, k, nn + 1)];
                        double s0 = 0, s1 = 0;
                        for (int i = 0; i <= LMD; i++) s0 = s0 + C->pbco[nt][0][i] * sap[i];
                        for (int i = 0; i <= LMD; i++) s1 = s1 + C->pbco[nt][1][i] * sap[i];
                        snd[(ip * NS + 0) * nb + jk] = s0 - resk * C->pbcot[nt][0];
                        snd[(ip * NS + 1) * nb + jk] = s1 - resk * C->pbcot[nt][1];
                        if (nt == 1) {
                            double *s2 = &snd[(ip * NS + 2) * nb + jk];
                            if (C->prm.quirk_send2) *s2 = *s2 + ((sap[0] - resk) - *s2);
                            else *s2 = sap[0] - resk;
                        }
                    }
            }
        }
    }
    for (int bi = 0; bi < C->nblk; bi++) {
        blk_t *B = &C->b[bi];
        for (int nn = 0; nn < 3; nn++) {
            int nb = B->nbsize[nn];
            for (int ip = 0; ip < 2; ip++) {
                if (B->ndf[nn][ip][nt] != 1) continue;
                blk_t *P = &C->b[B->mcd[nn][ip]];
                const double *snd = P->send[nt][nn];
                double *rcv = B->recv[nt][nn];
                for (int s = 0; s < NS; s++)
                    memcpy(&rcv[((m * 2 + ip) * NS + s) * nb], &snd[((1 - ip) * NS + s) * nb],
                           sizeof(double) * nb);
            }
        }
    }
    if (n45 == 1) {
        for (int bi = 0; bi < C->nblk; bi++) {
            blk_t *B = &C->b[bi];
            for (int nn = 0; nn < 3; nn++) {
                const double *f = fld3[bi * 3 + nn];
                int nb = B->nbsize[nn];
                for (int ip = 0; ip < 2; ip++) {
                    if (B->nbc[nn][ip] != BC_PER) continue;
                    int istart = ip * B->ijk[nn][0];
                    double *rcv = B->recv[nt][nn];
                    for (int k = 0; k <= B->ijk[nn][2]; k++)
                        for (int j = 0; j <= B->ijk[nn][1]; j++) {
                            int jk = k * (B->ijk[nn][1] + 1) + j;
                            double fv = f[IDX(B, istart, j, k, nn + 1)];
                            rcv[((m * 2 + ip) * NS + 0) * nb + jk] += fv * C->pbcot[nt][0];
                            rcv[((m * 2 + ip) * NS + 1) * nb + jk] += fv * C->pbcot[nt][1];
                            rcv[((m * 2 + ip) * NS + nt + 1) * nb + jk] += nt * fv;
                        }
                }
            }
        }
    }
}

/* one-array helper: same field for every direction of every block */
static void mpigo_1(orc_case *C, int nt, double *const *fld, int n45, int m)
{
    const double **f3 = xcalloc(3 * C->nblk, sizeof(double *));
    for (int b = 0; b < C->nblk; b++) for (int nn = 0; nn < 3; nn++) f3[b * 3 + nn] = fld[b];
    mpigo(C, nt, f3, n45, m);
    free(f3);
}

/* ---- compact derivative along nn: mo_numerics.f90:709-886 -------------
 * in and out may alias (pencils are gathered before they are scattered). */
static void deriv(orc_case *C, blk_t *B, int nn /*0..2*/, int m,
                  const double *in, double *out)
{
    const int N = B->ijk[nn][0], nb = B->nbsize[nn];
    const int nstart = B->ndf[nn][0][0], nend = B->ndf[nn][1][0];
    const double *xl = B->xl[0][nn], *xu = B->xu[0][nn];
    const double *rcv = B->recv[0][nn];
    double *drva = B->drva[nn];
    const int nk = B->ijk[nn][2] + 1, nj = B->ijk[nn][1] + 1;
    static const double ab0[4] = { a01, a02, a03, a04 }, ab1[4] = { a10, a12, a13, a14 };
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int jk = 0; jk < nk * nj; jk++) {
        int k = jk / nj, j = jk % nj;
        double sa[N + 1], sb[N + 1];
        long li[N + 1];
        for (int i = 0; i <= N; i++) { li[i] = IDX(B, i, j, k, nn + 1); sa[i] = in[li[i]]; }
        if (nstart == 0) {
            static const int o0[4] = { 1, 2, 3, 4 }, o1[4] = { 0, 2, 3, 4 };
            double s = 0; for (int c = 0; c < 4; c++) s = s + ab0[c] * (sa[o0[c]] - sa[0]);
            sb[0] = s;
            s = 0; for (int c = 0; c < 4; c++) s = s + ab1[c] * (sa[o1[c]] - sa[1]);
            sb[1] = s;
        } else {
            double s = 0; for (int i = 0; i <= LMD; i++) s = s + C->pbci[0][0][i] * sa[i];
            sb[0] = s + rcv[((m * 2 + 0) * 2 + 0) * nb + jk];
            s = 0; for (int i = 0; i <= LMD; i++) s = s + C->pbci[0][1][i] * sa[i];
            sb[1] = s + rcv[((m * 2 + 0) * 2 + 1) * nb + jk];
        }
        for (int i = 2; i <= N - 2; i++)
            sb[i] = aa * (sa[i + 1] - sa[i - 1]) + ab * (sa[i + 2] - sa[i - 2]);
        if (nend == 0) {
            static const int o0[4] = { 1, 2, 3, 4 }, o1[4] = { 0, 2, 3, 4 };
            double s = 0; for (int c = 0; c < 4; c++) s = s + ab0[c] * (sa[N] - sa[N - o0[c]]);
            sb[N] = s;
            s = 0; for (int c = 0; c < 4; c++) s = s + ab1[c] * (sa[N - 1] - sa[N - o1[c]]);
            sb[N - 1] = s;
        } else {
            double s = 0; for (int i = 0; i <= LMD; i++) s = s + C->pbci[0][0][i] * sa[N - i];
            sb[N] = -s - rcv[((m * 2 + 1) * 2 + 0) * nb + jk];
            s = 0; for (int i = 0; i <= LMD; i++) s = s + C->pbci[0][1][i] * sa[N - i];
            sb[N - 1] = -s - rcv[((m * 2 + 1) * 2 + 1) * nb + jk];
        }
        sa[0] = sb[0];
        sa[1] = sb[1] - xl[1 * 2 + 1] * sa[0];
        for (int i = 2; i <= N; i++)
            sa[i] = (sb[i] - xl[i * 2 + 0] * sa[i - 2]) - xl[i * 2 + 1] * sa[i - 1];
        sb[N] = xu[N * 3 + 0] * sa[N];
        sb[N - 1] = xu[(N - 1) * 3 + 0] * sa[N - 1] - xu[(N - 1) * 3 + 1] * sb[N];
        for (int i = N - 2; i >= 0; i--)
            sb[i] = (xu[i * 3 + 0] * sa[i] - xu[i * 3 + 1] * sb[i + 1]) - xu[i * 3 + 2] * sb[i + 2];
        for (int i = 0; i <= N; i++) out[li[i]] = sb[i];
        drva[(0 * 5 + m) * nb + jk] = sb[0];
        drva[(1 * 5 + m) * nb + jk] = sb[N];
    }
}

/* ---- compact filter along nn: mo_numerics.f90:890-987 ----------------- */
static void filte(orc_case *C, blk_t *B, int nn /*0..2*/, int m, double *f)
{
    const int N = B->ijk[nn][0], nb = B->nbsize[nn];
    const int nstart = B->ndf[nn][0][1], nend = B->ndf[nn][1][1];
    const double *yl = B->xl[1][nn], *yu = B->xu[1][nn];
    const double *rcv = B->recv[1][nn];
    const double fa = C->fa, fb = C->fb, fc = C->fc;
    const int nk = B->ijk[nn][2] + 1, nj = B->ijk[nn][1] + 1;
#define RCV(side, slot) rcv[((m * 2 + (side)) * 3 + (slot)) * nb + jk]
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int jk = 0; jk < nk * nj; jk++) {
        int k = jk / nj, j = jk % nj;
        double sa[N + 1], sb[N + 1];
        long li[N + 1];
        for (int i = 0; i <= N; i++) { li[i] = IDX(B, i, j, k, nn + 1); sa[i] = f[li[i]]; }
        if (nstart == 0) {
            static const int o0[5] = { 1, 2, 3, 4, 5 }, o1[5] = { 0, 2, 3, 4, 5 }, o2[5] = { 0, 1, 3, 4, 5 };
            double s = 0; for (int c = 0; c < 5; c++) s = s + C->fbc[0][c] * (sa[o0[c]] - sa[0]);
            sb[0] = s;
            s = 0; for (int c = 0; c < 5; c++) s = s + C->fbc[1][c] * (sa[o1[c]] - sa[1]);
            sb[1] = s;
            s = 0; for (int c = 0; c < 5; c++) s = s + C->fbc[2][c] * (sa[o2[c]] - sa[2]);
            sb[2] = s;
        } else {
            double ra2k = sa[2] + sa[2];
            double s = 0; for (int i = 0; i <= LMD; i++) s = s + C->pbci[1][0][i] * sa[i];
            sb[0] = s + RCV(0, 0);
            s = 0; for (int i = 0; i <= LMD; i++) s = s + C->pbci[1][1][i] * sa[i];
            sb[1] = s + RCV(0, 1);
            sb[2] = (fa * ((sa[1] + sa[3]) - ra2k) + fb * ((sa[0] + sa[4]) - ra2k))
                    + fc * ((RCV(0, 2) + sa[5]) - ra2k);
        }
        for (int i = 3; i <= N - 3; i++) {
            double resk = sa[i] + sa[i];
            sb[i] = (fa * ((sa[i - 1] + sa[i + 1]) - resk) + fb * ((sa[i - 2] + sa[i + 2]) - resk))
                    + fc * ((sa[i - 3] + sa[i + 3]) - resk);
        }
        if (nend == 0) {
            static const int o0[5] = { 1, 2, 3, 4, 5 }, o1[5] = { 0, 2, 3, 4, 5 }, o2[5] = { 0, 1, 3, 4, 5 };
            double s = 0; for (int c = 0; c < 5; c++) s = s + C->fbc[0][c] * (sa[N - o0[c]] - sa[N]);
            sb[N] = s;
            s = 0; for (int c = 0; c < 5; c++) s = s + C->fbc[1][c] * (sa[N - o1[c]] - sa[N - 1]);
            sb[N - 1] = s;
            s = 0; for (int c = 0; c < 5; c++) s = s + C->fbc[2][c] * (sa[N - o2[c]] - sa[N - 2]);
            sb[N - 2] = s;
        } else {
            double ra2k = sa[N - 2] + sa[N - 2];
            double s = 0; for (int i = 0; i <= LMD; i++) s = s + C->pbci[1][0][i] * sa[N - i];
            sb[N] = s + RCV(1, 0);
            s = 0; for (int i = 0; i <= LMD; i++) s = s + C->pbci[1][1][i] * sa[N - i];
            sb[N - 1] = s + RCV(1, 1);
            sb[N - 2] = (fa * ((sa[N - 3] + sa[N - 1]) - ra2k) + fb * ((sa[N - 4] + sa[N]) - ra2k))
                        + fc * ((sa[N - 5] + RCV(1, 2)) - ra2k);
        }
        sa[0] = sb[0];
        sa[1] = sb[1] - yl[1 * 2 + 1] * sa[0];
        for (int i = 2; i <= N; i++)
            sa[i] = (sb[i] - yl[i * 2 + 0] * sa[i - 2]) - yl[i * 2 + 1] * sa[i - 1];
        sb[N] = yu[N * 3 + 0] * sa[N];
        sb[N - 1] = yu[(N - 1) * 3 + 0] * sa[N - 1] - yu[(N - 1) * 3 + 1] * sb[N];
        for (int i = N - 2; i >= 0; i--)
            sb[i] = (yu[i * 3 + 0] * sa[i] - yu[i * 3 + 1] * sb[i + 1]) - yu[i * 3 + 2] * sb[i + 2];
        for (int i = 0; i <= N; i++) f[li[i]] = f[li[i]] + sb[i];
    }
#undef RCV
}

/* ---- junction search: mo_domdcomp.f90:269-437, 563-585 --------------- */
static void trimm(int *itrim, int *lene, int lens)
{
    for (int k = 0; k <= lens - 1; k++)
        for (int kk = k + 1; kk <= lens; kk++)
            if (itrim[kk] - itrim[k] == 0) itrim[kk] = -1;
    *lene = lens;
    for (int k = 0; k <= lens; k++)
        if (itrim[k] == -1) {
            *lene = *lene - 1;
            itrim[k] = (k + 1 <= LJPL) ? itrim[k + 1] : -1;
            if (k + 1 <= LJPL) itrim[k + 1] = -1;
        }
}

static void junction_closure(int lp, int *imj, int *jj, int *tag)
{
    const int W = LJPL + 1;
    for (int i = 0; i <= lp; i++)
        for (int j = 0; j <= lp; j++) {
            int ll = 0;
            if ((i - j) * imj[i] * imj[j] != 0) {
                for (int k = 0; k <= imj[i]; k++)
                    for (int kk = 0; kk <= imj[j]; kk++)
                        if (jj[j * W + kk] - jj[i * W + k] == 0) ll++;
                if (ll != 0) {
                    int ks = imj[i] + 1, ke = ks + imj[j];
                    for (int c = 0; c <= ke - ks; c++) jj[i * W + ks + c] = jj[j * W + c];
                    trimm(&jj[i * W], &imj[i], ke);
                }
            }
        }
    for (int i = 0; i <= lp; i++) {
        int mn = jj[i * W];
        for (int k = 1; k <= imj[i]; k++) if (jj[i * W + k] < mn) mn = jj[i * W + k];
        tag[i] = mn;
    }
}

static void search_junctions(orc_case *C)
{
    const int W = LJPL + 1, mbk = C->nblk - 1;
    int ijp[4][3][2], ijl[4][3][2];
    C->lpp = 8 * mbk + 7; C->lqq = 12 * mbk + 11;
    C->imjp = xcalloc(C->lpp + 1, sizeof(int)); C->jptag = xcalloc(C->lpp + 1, sizeof(int));
    C->jjp = xcalloc((size_t)(C->lpp + 1) * W, sizeof(int));
    C->imjl = xcalloc(C->lqq + 1, sizeof(int)); C->jltag = xcalloc(C->lqq + 1, sizeof(int));
    C->jjl = xcalloc((size_t)(C->lqq + 1) * W, sizeof(int));
    for (int ip = 0; ip < 2; ip++) {
        int p1[4] = { 0, 2, 4, 6 }, p2[4] = { 0, 4, 1, 5 }, p3[4] = { 0, 1, 2, 3 };
        int l1[4] = { 0, 1, 8, 10 }, l2[4] = { 4, 5, 0, 2 }, l3[4] = { 8, 9, 4, 6 }, nj[4] = { 2, 2, 1, 1 };
        for (int i = 0; i < 4; i++) {
            ijp[i][0][ip] = p1[i] + ip; ijp[i][1][ip] = p2[i] + 2 * ip; ijp[i][2][ip] = p3[i] + 4 * ip;
            ijl[i][0][ip] = l1[i] + ip * nj[i]; ijl[i][1][ip] = l2[i] + ip * nj[i];
            ijl[i][2][ip] = l3[i] + ip * nj[i];
        }
    }
    for (int i = 0; i <= C->lpp; i++) { for (int k = 0; k < W; k++) C->jjp[i * W + k] = -1; C->jjp[i * W] = i; }
    for (int i = 0; i <= C->lqq; i++) { for (int k = 0; k < W; k++) C->jjl[i * W + k] = -1; C->jjl[i * W] = i; }
    for (int mm = 0; mm <= mbk; mm++)
        for (int nn = 0; nn < 3; nn++)
            for (int ip = 0; ip < 2; ip++) {
                int mp = C->b[mm].mcd[nn][ip];
                if (mp == -1) continue;
                for (int i = 0; i < 4; i++) {
                    int is = 8 * mm + ijp[i][nn][ip], ie = 8 * mp + ijp[i][nn][1 - ip];
                    C->imjp[is]++; C->jjp[is * W + C->imjp[is]] = C->jjp[ie * W];
                    is = 12 * mm + ijl[i][nn][ip]; ie = 12 * mp + ijl[i][nn][1 - ip];
                    C->imjl[is]++; C->jjl[is * W + C->imjl[is]] = C->jjl[ie * W];
                }
            }
    junction_closure(C->lpp, C->imjp, C->jjp, C->jptag);
    junction_closure(C->lqq, C->imjl, C->jjl, C->jltag);
    /* per-block tables; one rank per block, so idsd3(...) = block index */
    for (int mb = 0; mb <= mbk; mb++) {
        blk_t *B = &C->b[mb];
        for (int i = 0; i < 8; i++) {
            B->njp[i] = -1;
            for (int k = 0; k <= LJPL; k++) B->jpcd[i][k] = -1;
            int l = 8 * mb + i;
            if (C->imjp[l] >= 3) {
                B->njp[i] = l;
                for (int k = 1; k <= C->imjp[l]; k++) B->jpcd[i][k] = C->jjp[l * W + k] / 8;
            }
        }
        for (int i = 0; i < 12; i++) {
            B->njl[i] = -1;
            for (int k = 0; k <= LJPL; k++) B->jlcd[i][k] = -1;
            int l = 12 * mb + i;
            if (C->imjl[l] >= 2) {
                B->njl[i] = l;
                for (int k = 1; k <= C->imjl[l]; k++) B->jlcd[i][k] = C->jjl[l * W + k] / 12;
            }
        }
    }
}

void orc_junctions(orc_case *C, int b, int *np8, int *nl12)
{
    blk_t *B = &C->b[b];
    for (int i = 0; i < 8; i++) np8[i] = B->njp[i] == -1 ? -1 : C->imjp[B->njp[i]];
    for (int i = 0; i < 12; i++) nl12[i] = B->njl[i] == -1 ? -1 : C->imjl[B->njl[i]];
}

/* ---- message matching for junction averaging ---------------------------
 * Replays the per-rank posting order of isend/irecv; the k-th receive a
 * rank posts from source s with tag t matches the k-th send s posted to it
 * with tag t (MPI's non-overtaking rule). */
typedef struct { int dst, tag, src_item; } msg_t;

static void match_messages(int nranks, msg_t **sends, int *nsend, msg_t **recvs, int *nrecv,
                           int **match /* [rank][recv idx] -> global send id */)
{
    for (int r = 0; r < nranks; r++)
        for (int q = 0; q < nrecv[r]; q++) {
            int s = recvs[r][q].dst, t = recvs[r][q].tag;  /* dst field holds the source */
            int ord = 0;
            for (int q2 = 0; q2 < q; q2++)
                if (recvs[r][q2].dst == s && recvs[r][q2].tag == t) ord++;
            int found = -1, cnt = 0;
            for (int k = 0; k < nsend[s]; k++)
                if (sends[s][k].dst == r && sends[s][k].tag == t) {
                    if (cnt == ord) { found = sends[s][k].src_item; break; }
                    cnt++;
                }
            match[r][q] = found;
        }
}

/* average_point / average_line: mo_domdcomp.f90:439-559 */
static void average_junctions(orc_case *C, int lines)
{
    const int nr = C->nblk, W = LJPL + 1;
    (void)W;
    msg_t **sends = xcalloc(nr, sizeof(msg_t *)), **recvs = xcalloc(nr, sizeof(msg_t *));
    int *nsend = xcalloc(nr, sizeof(int)), *nrecv = xcalloc(nr, sizeof(int));
    int **match = xcalloc(nr, sizeof(int *));
    int nitem = lines ? 12 : 8;
    /* payloads: per rank, per item, pre-averaging values (edge: (len,5)) */
    double ***pay = xcalloc(nr, sizeof(double **));
    int **len = xcalloc(nr, sizeof(int *));
    for (int r = 0; r < nr; r++) {
        blk_t *B = &C->b[r];
        sends[r] = xcalloc(nitem * (LJPL + 1), sizeof(msg_t));
        recvs[r] = xcalloc(nitem * (LJPL + 1), sizeof(msg_t));
        match[r] = xcalloc(nitem * (LJPL + 1), sizeof(int));
        pay[r] = xcalloc(nitem, sizeof(double *));
        len[r] = xcalloc(nitem, sizeof(int));
        for (int i = 0; i < nitem; i++) {
            int ii = lines ? B->njl[i] : B->njp[i];
            if (ii == -1) continue;
            int itag = lines ? C->jltag[ii] : C->jptag[ii];
            int cnt = lines ? C->imjl[ii] : C->imjp[ii];
            int L;
            if (lines) {
                int nn = i / 4 + 1, ip = (i % 4) / 2, jp = (i % 4) % 2;
                L = B->ijk[nn - 1][2] + 1;
                pay[r][i] = xcalloc(5 * L, sizeof(double));
                for (int k = 0; k < L; k++) {
                    long l = IDX(B, ip * B->ijk[nn - 1][0], jp * B->ijk[nn - 1][1], k, nn);
                    for (int m = 0; m < 5; m++) pay[r][i][5 * k + m] = B->qa[m * B->npt + l];
                }
            } else {
                int ip = i % 2, jp = (i % 4) / 2, kp = i / 4;
                L = 1;
                pay[r][i] = xcalloc(5, sizeof(double));
                long l = IDX(B, ip * B->lxi, jp * B->let, kp * B->lze, 1);
                for (int m = 0; m < 5; m++) pay[r][i][m] = B->qa[m * B->npt + l];
            }
            len[r][i] = L;
            for (int j = 1; j <= cnt; j++) {
                int partner = lines ? B->jlcd[i][j] : B->jpcd[i][j];
                sends[r][nsend[r]++] = (msg_t){ partner, itag, r * 64 + i };
                recvs[r][nrecv[r]++] = (msg_t){ partner, itag, i };
            }
        }
    }
    match_messages(nr, sends, nsend, recvs, nrecv, match);
    for (int r = 0; r < nr; r++) {
        blk_t *B = &C->b[r];
        int q = 0;
        for (int i = 0; i < nitem; i++) {
            int ii = lines ? B->njl[i] : B->njp[i];
            if (ii == -1) continue;
            int cnt = lines ? C->imjl[ii] : C->imjp[ii];
            int L = len[r][i];
            double *acc = xcalloc(5 * L, sizeof(double));
            memcpy(acc, pay[r][i], sizeof(double) * 5 * L);
            for (int j = 1; j <= cnt; j++, q++) {
                int g = match[r][q];
                const double *src = NULL;
                if (g >= 0) {
                    int sr = g / 64, si = g % 64;
                    if (len[sr][si] == L) src = pay[sr][si];
                }
                if (!src) { fprintf(stderr, "orc: unmatched junction message\n"); abort(); }
                for (int c = 0; c < 5 * L; c++) acc[c] = acc[c] + src[c];
            }
            double fctr = one / (cnt + 1);
            if (lines) {
                int nn = i / 4 + 1, ip = (i % 4) / 2, jp = (i % 4) % 2;
                for (int k = 0; k < L; k++) {
                    long l = IDX(B, ip * B->ijk[nn - 1][0], jp * B->ijk[nn - 1][1], k, nn);
                    for (int m = 0; m < 5; m++) B->qa[m * B->npt + l] = fctr * acc[5 * k + m];
                }
            } else {
                int ip = i % 2, jp = (i % 4) / 2, kp = i / 4;
                long l = IDX(B, ip * B->lxi, jp * B->let, kp * B->lze, 1);
                for (int m = 0; m < 5; m++) B->qa[m * B->npt + l] = fctr * acc[m];
            }
            free(acc);
        }
    }
    for (int r = 0; r < nr; r++) {
        for (int i = 0; i < nitem; i++) free(pay[r][i]);
        free(pay[r]); free(len[r]); free(sends[r]); free(recvs[r]); free(match[r]);
    }
    free(pay); free(len); free(sends); free(recvs); free(match); free(nsend); free(nrecv);
}

/* ---- GCBC helpers: mo_gcbc.f90:504-611 -------------------------------- */
typedef struct { double ao, aoi, ve[3], hv2, vn, vs; } elem_t;

static void eleme(elem_t *e, const double cm[3], double qa1, const double qa24[3],
                  double pp, const double umf[3])
{
    double rhoi = one / qa1;
    e->ao = sqrt((GAM * rhoi) * pp);
    e->aoi = one / e->ao;
    for (int c = 0; c < 3; c++) e->ve[c] = rhoi * qa24[c];
    e->hv2 = half * ((e->ve[0] * e->ve[0] + e->ve[1] * e->ve[1]) + e->ve[2] * e->ve[2]);
    e->vn = (cm[0] * e->ve[0] + cm[1] * e->ve[1]) + cm[2] * e->ve[2];
    e->vs = (cm[0] * umf[0] + cm[1] * umf[1]) + cm[2] * umf[2];
}

static void xtq2r(const elem_t *e, const double cm[3], double xt[5][5])
{
    double ho = (gamm1 * e->aoi) * e->aoi, bo = 1 - ho * e->hv2, co = e->aoi * e->vn;
    double dm[3], rv[3];
    const double *ve = e->ve;
    for (int c = 0; c < 3; c++) { dm[c] = e->aoi * cm[c]; rv[c] = ho * ve[c]; }
    xt[0][0] = (bo * cm[0] + dm[1] * ve[2]) - dm[2] * ve[1];
    xt[0][1] = cm[0] * rv[0];
    xt[0][2] = cm[0] * rv[1] + dm[2];
    xt[0][3] = cm[0] * rv[2] - dm[1];
    xt[0][4] = -ho * cm[0];
    xt[1][0] = (bo * cm[1] + dm[2] * ve[0]) - dm[0] * ve[2];
    xt[1][1] = cm[1] * rv[0] - dm[2];
    xt[1][2] = cm[1] * rv[1];
    xt[1][3] = cm[1] * rv[2] + dm[0];
    xt[1][4] = -ho * cm[1];
    xt[2][0] = (bo * cm[2] + dm[0] * ve[1]) - dm[1] * ve[0];
    xt[2][1] = cm[2] * rv[0] + dm[1];
    xt[2][2] = cm[2] * rv[1] - dm[0];
    xt[2][3] = cm[2] * rv[2];
    xt[2][4] = -ho * cm[2];
    xt[3][0] = (one - bo) - co;
    xt[3][1] = dm[0] - rv[0];
    xt[3][2] = dm[1] - rv[1];
    xt[3][3] = dm[2] - rv[2];
    xt[3][4] = ho;
    xt[4][0] = (one - bo) + co;
    xt[4][1] = -dm[0] - rv[0];
    xt[4][2] = -dm[1] - rv[1];
    xt[4][3] = -dm[2] - rv[2];
    xt[4][4] = ho;
}

static void xtr2q(const elem_t *e, const double cm[3], double xt[5][5])
{
    double bo = e->hv2 + (hamm1 * e->ao) * e->ao, co = e->ao * e->vn, dm[3];
    const double *ve = e->ve;
    for (int c = 0; c < 3; c++) dm[c] = e->ao * cm[c];
    xt[0][0] = cm[0]; xt[0][1] = cm[1]; xt[0][2] = cm[2]; xt[0][3] = half; xt[0][4] = half;
    xt[1][0] = cm[0] * ve[0];
    xt[1][1] = cm[1] * ve[0] - dm[2];
    xt[1][2] = cm[2] * ve[0] + dm[1];
    xt[1][3] = half * (ve[0] + dm[0]);
    xt[1][4] = xt[1][3] - dm[0];
    xt[2][0] = cm[0] * ve[1] + dm[2];
    xt[2][1] = cm[1] * ve[1];
    xt[2][2] = cm[2] * ve[1] - dm[0];
    xt[2][3] = half * (ve[1] + dm[1]);
    xt[2][4] = xt[2][3] - dm[1];
    xt[3][0] = cm[0] * ve[2] - dm[1];
    xt[3][1] = cm[1] * ve[2] + dm[0];
    xt[3][2] = cm[2] * ve[2];
    xt[3][3] = half * (ve[2] + dm[2]);
    xt[3][4] = xt[3][3] - dm[2];
    xt[4][0] = (e->hv2 * cm[0] + dm[2] * ve[1]) - dm[1] * ve[2];
    xt[4][1] = (e->hv2 * cm[1] + dm[0] * ve[2]) - dm[2] * ve[0];
    xt[4][2] = (e->hv2 * cm[2] + dm[1] * ve[0]) - dm[0] * ve[1];
    xt[4][3] = half * (bo + co);
    xt[4][4] = xt[4][3] - co;
}

static void matvec5(double xt[5][5], const double v[5], double out[5])
{
    for (int i = 0; i < 5; i++) {
        double s = 0; for (int k = 0; k < 5; k++) s = s + xt[i][k] * v[k];
        out[i] = s;
    }
}

static int is_gcbc_face(int np)
{
    return np == BC_NONREF || np == BC_WALLI || np == BC_WALLV || np == BC_ICURV;
}
static int is_avg_face(int np) { return np == BC_ICURV || np == BC_ISTR || np == BC_PER; }

/* gcbc_init: mo_gcbc.f90:40-131 */
static void gcbc_init(orc_case *C)
{
    double cbca[MBCI * MBCI] = { 0 }, cbcs[MBCI * MBCI], rbci[MBCI] = { one, alpha10, beta, 0 };
#define CB(i, j) cbca[((i)-1) + MBCI * ((j)-1)]
    CB(1, 1) = alpha01; CB(1, 2) = beta02;
    CB(2, 1) = one; CB(2, 2) = alpha12; CB(2, 3) = beta13;
    CB(3, 1) = alpha; CB(3, 2) = one; CB(3, 3) = alpha; CB(3, 4) = beta;
    for (int i = 4; i <= MBCI; i++) {
        CB(i, i - 3) = beta; CB(i, i - 2) = alpha; CB(i, i - 1) = one; CB(i, i) = alpha;
        if (i < MBCI) CB(i, i + 1) = beta;
    }
#undef CB
    mtrxi(cbca, cbcs, MBCI);
    for (int i = 0; i < MBCI; i++) {
        double s = 0; for (int k = 0; k < MBCI; k++) s = s + cbcs[i + MBCI * k] * rbci[k];
        C->sbci[i] = -s;
    }
    double fctr = PI / (MBCI + 1), res = zero;
    for (int i = 0; i < MBCI; i++) {
        res = res + one;
        C->sbci[i] = (half * C->sbci[i]) * (one + cos(res * fctr));
    }
    for (int bi = 0; bi < C->nblk; bi++) {
        blk_t *B = &C->b[bi];
        double *cnt = xcalloc(B->npt, sizeof(double));
        long cap = 0;
        for (int nn = 0; nn < 3; nn++)
            for (int ip = 0; ip < 2; ip++)
                if (is_gcbc_face(B->nbc[nn][ip])) cap += (long)B->nbsize[nn] * (MBCI + 1);
        double *val = xcalloc(cap, sizeof(double));
        long *pt = xcalloc(cap, sizeof(long));
        B->nrr = xcalloc(B->npt, sizeof(int));
        long ll = -1;
        for (int nn = 0; nn < 3; nn++)
            for (int ip = 0; ip < 2; ip++) {
                int np = B->nbc[nn][ip], i = ip * B->ijk[nn][0], iq = 1 - 2 * ip;
                if (is_gcbc_face(np)) {
                    for (int k = 0; k <= B->ijk[nn][2]; k++)
                        for (int j = 0; j <= B->ijk[nn][1]; j++) {
                            long l = IDX(B, i, j, k, nn + 1);
                            double res2 = one / B->yaco[l];
                            ll++; cnt[l] += one; val[ll] = res2; pt[ll] = l;
                            for (int ii = 1; ii <= MBCI; ii++) {
                                l = IDX(B, i + iq * ii, j, k, nn + 1);
                                ll++; cnt[l] += one; val[ll] = res2 * C->sbci[ii - 1]; pt[ll] = l;
                            }
                        }
                }
                if (is_avg_face(np))
                    for (int k = 0; k <= B->ijk[nn][2]; k++)
                        for (int j = 0; j <= B->ijk[nn][1]; j++)
                            B->nrr[IDX(B, i, j, k, nn + 1)] = 1;
            }
        B->lq = ll;
        B->sbcc = xcalloc(ll + 1, sizeof(double));
        for (long q = 0; q <= ll; q++) B->sbcc[q] = val[q] / cnt[pt[q]];
        free(cnt); free(val); free(pt);
    }
}

/* gcbc_setup + gcbc_comm + gcbc_update: mo_gcbc.f90:133-340 */
static void gcbc_apply(orc_case *C)
{
    int any = 0;
    for (int bi = 0; bi < C->nblk; bi++)
        for (int nn = 0; nn < 3; nn++)
            for (int ip = 0; ip < 2; ip++) any |= is_gcbc_face(C->b[bi].nbc[nn][ip]);
    if (!any) return;
    for (int bi = 0; bi < C->nblk; bi++) {       /* gcbc_setup */
        blk_t *B = &C->b[bi];
        long np1 = B->npt;
        for (int nn = 0; nn < 3; nn++) {
            int nb = B->nbsize[nn];
            for (int ip = 0; ip < 2; ip++) {
                int np = B->nbc[nn][ip], i = ip * B->ijk[nn][0];
                if (!is_gcbc_face(np)) continue;
                double ra0 = (double)(((BC_WALLI - np) * (BC_WALLV - np) * (BC_ICURV - np)) / 3000);
                double ra1 = one - ra0;
                for (int k = 0; k <= B->ijk[nn][2]; k++)
                    for (int j = 0; j <= B->ijk[nn][1]; j++) {
                        int jk = k * (B->ijk[nn][1] + 1) + j;
                        long l = IDX(B, i, j, k, nn + 1);
                        double cm[3], q24[3], cha[5], yc[5], xt[5][5], out[5];
                        for (int c = 0; c < 3; c++) cm[c] = B->cm[nn][(ip * 3 + c) * nb + jk];
                        for (int c = 0; c < 3; c++) q24[c] = B->qa[(c + 1) * np1 + l];
                        elem_t e; eleme(&e, cm, B->qa[l], q24, B->p[l], C->umf);
                        xtq2r(&e, cm, xt);
                        for (int c = 0; c < 5; c++)
                            cha[c] = ra0 * B->drva[nn][(ip * 5 + c) * nb + jk] + ra1 * B->de[c * np1 + l];
                        for (int c = 0; c < 5; c++) yc[c] = B->yaco[l] * cha[c];
                        matvec5(xt, yc, out);
                        for (int c = 0; c < 5; c++) B->drva[nn][(ip * 5 + c) * nb + jk] = out[c];
                    }
            }
        }
    }
    for (int bi = 0; bi < C->nblk; bi++) {       /* gcbc_comm: BC 30 faces */
        blk_t *B = &C->b[bi];
        for (int nn = 0; nn < 3; nn++)
            for (int ip = 0; ip < 2; ip++)
                if (B->nbc[nn][ip] == BC_ICURV) {
                    blk_t *P = &C->b[B->mcd[nn][ip]];
                    int nb = B->nbsize[nn];
                    memcpy(&B->drvb[nn][ip * 5 * nb], &P->drva[nn][(1 - ip) * 5 * nb],
                           sizeof(double) * 5 * nb);
                }
    }
    for (int bi = 0; bi < C->nblk; bi++) {       /* gcbc_update */
        blk_t *B = &C->b[bi];
        long np1 = B->npt, ll = -1;
        for (int nn = 0; nn < 3; nn++) {
            int nb = B->nbsize[nn];
            for (int ip = 0; ip < 2; ip++) {
                int np = B->nbc[nn][ip], i = ip * B->ijk[nn][0], iq = 1 - 2 * ip;
                double ra0 = iq;
                if (!is_gcbc_face(np)) continue;
                double dtwi = one / (4 * C->dt + sml);
                for (int k = 0; k <= B->ijk[nn][2]; k++)
                    for (int j = 0; j <= B->ijk[nn][1]; j++) {
                        int jk = k * (B->ijk[nn][1] + 1) + j;
                        long l = IDX(B, i, j, k, nn + 1);
                        double cm[3], q24[3], cha[5], dha[5], dif[5], xt[5][5];
                        for (int c = 0; c < 3; c++) cm[c] = B->cm[nn][(ip * 3 + c) * nb + jk];
                        for (int c = 0; c < 3; c++) q24[c] = B->qa[(c + 1) * np1 + l];
                        elem_t e; eleme(&e, cm, B->qa[l], q24, B->p[l], C->umf);
                        for (int c = 0; c < 5; c++) {
                            cha[c] = B->drva[nn][(ip * 5 + c) * nb + jk];
                            dha[c] = B->drvb[nn][(ip * 5 + c) * nb + jk];
                        }
                        if (np == BC_NONREF) {
                            if (ra0 * ((e.vn + e.vs) + e.ao) > zero) cha[3] = zero;
                            if (ra0 * ((e.vn + e.vs) - e.ao) > zero) cha[4] = zero;
                        } else if (np == BC_WALLI || np == BC_WALLV) {
                            double sdu = 0;
                            for (int c = 0; c < 3; c++) sdu = sdu + cm[c] * C->dudtmf[c];
                            cha[3 + ip] = cha[4 - ip] + (((two * ra0) * e.aoi) * B->qa[l])
                                                        * (sdu + dtwi * (e.vn + e.vs));
                        } else { /* BC_ICURV */
                            if (ra0 * (e.vn + e.vs) > zero) { cha[0] = dha[0]; cha[1] = dha[1]; cha[2] = dha[2]; }
                            if (ra0 * ((e.vn + e.vs) + e.ao) > zero) cha[3] = dha[3];
                            if (ra0 * ((e.vn + e.vs) - e.ao) > zero) cha[4] = dha[4];
                        }
                        xtr2q(&e, cm, xt);
                        for (int c = 0; c < 5; c++) dif[c] = cha[c] - B->drva[nn][(ip * 5 + c) * nb + jk];
                        matvec5(xt, dif, dha);
                        for (int ii = 0; ii <= MBCI; ii++) {
                            long lp = IDX(B, i + iq * ii, j, k, nn + 1);
                            ll++;
                            for (int c = 0; c < 5; c++)
                                B->de[c * np1 + lp] = B->de[c * np1 + lp] + B->sbcc[ll] * dha[c];
                        }
                    }
            }
        }
    }
}

/* average_surface: mo_gcbc.f90:342-422 */
static void average_surface(orc_case *C)
{
    for (int bi = 0; bi < C->nblk; bi++) {
        blk_t *B = &C->b[bi];
        for (int nn = 0; nn < 3; nn++) {
            int nb = B->nbsize[nn];
            for (int ip = 0; ip < 2; ip++) {
                if (!is_avg_face(B->nbc[nn][ip])) continue;
                int i = ip * B->ijk[nn][0];
                for (int k = 0; k <= B->ijk[nn][2]; k++)
                    for (int j = 0; j <= B->ijk[nn][1]; j++) {
                        long l = IDX(B, i, j, k, nn + 1);
                        int jk = k * (B->ijk[nn][1] + 1) + j;
                        for (int m = 0; m < 5; m++) B->drva[nn][(ip * 5 + m) * nb + jk] = B->qa[m * B->npt + l];
                        B->avg[l] = one;
                    }
            }
        }
    }
    for (int bi = 0; bi < C->nblk; bi++) {
        blk_t *B = &C->b[bi];
        for (int nn = 0; nn < 3; nn++)
            for (int ip = 0; ip < 2; ip++)
                if (is_avg_face(B->nbc[nn][ip])) {
                    blk_t *P = &C->b[B->mcd[nn][ip]];
                    int nb = B->nbsize[nn];
                    memcpy(&B->drvb[nn][ip * 5 * nb], &P->drva[nn][(1 - ip) * 5 * nb],
                           sizeof(double) * 5 * nb);
                }
    }
    for (int bi = 0; bi < C->nblk; bi++) {
        blk_t *B = &C->b[bi];
        for (int nn = 0; nn < 3; nn++) {
            int nb = B->nbsize[nn];
            for (int ip = 0; ip < 2; ip++) {
                if (!is_avg_face(B->nbc[nn][ip])) continue;
                int i = ip * B->ijk[nn][0];
                for (int k = 0; k <= B->ijk[nn][2]; k++)
                    for (int j = 0; j <= B->ijk[nn][1]; j++) {
                        long l = IDX(B, i, j, k, nn + 1);
                        int jk = k * (B->ijk[nn][1] + 1) + j;
                        double nr = B->nrr[l];
                        B->avg[l] = B->avg[l] + nr;
                        double r2 = nr / B->avg[l];
                        for (int m = 0; m < 5; m++) {
                            double q = B->qa[m * B->npt + l];
                            B->qa[m * B->npt + l] = r2 * ((B->avg[l] - one) * q + B->drvb[nn][(ip * 5 + m) * nb + jk])
                                                    + (1 - nr) * q;
                        }
                    }
            }
        }
    }
}

/* wall_condition_update: mo_gcbc.f90:424-459 */
static void wall_update(orc_case *C)
{
    for (int bi = 0; bi < C->nblk; bi++) {
        blk_t *B = &C->b[bi];
        long N1 = B->npt;
        for (int nn = 0; nn < 3; nn++)
            for (int ip = 0; ip < 2; ip++) {
                int np = B->nbc[nn][ip], i = ip * B->ijk[nn][0];
                if (np != BC_WALLI && np != BC_WALLV) continue;
                for (int k = 0; k <= B->ijk[nn][2]; k++)
                    for (int j = 0; j <= B->ijk[nn][1]; j++) {
                        long l = IDX(B, i, j, k, nn + 1);
                        double *q = B->qa;
                        if (np == BC_WALLI) {
                            double s = 0; for (int c = 1; c <= 3; c++) s = s + q[c * N1 + l] * q[c * N1 + l];
                            q[4 * N1 + l] = 0 * q[4 * N1 + l]
                                            + 1 * (hamhamm1 * pow(q[l], GAM) + (half * s) / q[l]);
                        } else {
                            double fctr = 1 * q[l];
                            for (int c = 1; c <= 3; c++) q[c * N1 + l] = 0 * q[c * N1 + l] - fctr * C->umf[c - 1];
                            double s = 0; for (int c = 1; c <= 3; c++) s = s + q[c * N1 + l] * q[c * N1 + l];
                            q[4 * N1 + l] = 0 * q[4 * N1 + l] + 1 * (hamhamm1 * q[l] + (half * s) / q[l]);
                        }
                    }
            }
    }
}

/* ---- grid metrics: mo_grid.f90:52-189 --------------------------------- */
static void grid_metrics(orc_case *C)
{
    int nb = C->nblk;
    double **r3 = xcalloc(nb, sizeof(double *));
    double **d[3];
    for (int c = 0; c < 3; c++) d[c] = xcalloc(nb, sizeof(double *));
    for (int b = 0; b < nb; b++) {
        r3[b] = xcalloc(3 * C->b[b].npt, sizeof(double));
        for (int c = 0; c < 3; c++) d[c][b] = xcalloc(3 * C->b[b].npt, sizeof(double));
    }
    for (int m = 0; m < 3; m++) {
        for (int b = 0; b < nb; b++) memcpy(r3[b], C->b[b].xyz + m * C->b[b].npt, sizeof(double) * C->b[b].npt);
        mpigo_1(C, 0, r3, 1, m);
        for (int b = 0; b < nb; b++) {
            blk_t *B = &C->b[b];
            long N1 = B->npt;
            deriv(C, B, 2, m, r3[b], r3[b] + 2 * N1);
            deriv(C, B, 1, m, r3[b], r3[b] + 1 * N1);
            deriv(C, B, 0, m, r3[b], r3[b]);
            memcpy(d[m][b], r3[b], sizeof(double) * 3 * N1);
        }
    }
    for (int b = 0; b < nb; b++) {
        blk_t *B = &C->b[b];
        long N1 = B->npt;
        const double *qo = d[0][b], *qa = d[1][b], *de = d[2][b];
#define Q(a, c) a[((c)-1) * N1 + l]
        for (long l = 0; l < N1; l++) {
            B->xim[l] = Q(qa, 2) * Q(de, 3) - Q(de, 2) * Q(qa, 3);
            B->xim[N1 + l] = Q(de, 2) * Q(qo, 3) - Q(qo, 2) * Q(de, 3);
            B->xim[2 * N1 + l] = Q(qo, 2) * Q(qa, 3) - Q(qa, 2) * Q(qo, 3);
            B->etm[l] = Q(qa, 3) * Q(de, 1) - Q(de, 3) * Q(qa, 1);
            B->etm[N1 + l] = Q(de, 3) * Q(qo, 1) - Q(qo, 3) * Q(de, 1);
            B->etm[2 * N1 + l] = Q(qo, 3) * Q(qa, 1) - Q(qa, 3) * Q(qo, 1);
            B->zem[l] = Q(qa, 1) * Q(de, 2) - Q(de, 1) * Q(qa, 2);
            B->zem[N1 + l] = Q(de, 1) * Q(qo, 2) - Q(qo, 1) * Q(de, 2);
            B->zem[2 * N1 + l] = Q(qo, 1) * Q(qa, 2) - Q(qa, 1) * Q(qo, 2);
        }
    }
    double **fl = xcalloc(nb, sizeof(double *));
    for (int m = 0; m < 3; m++) {
        for (int w = 0; w < 3; w++) {
            for (int b = 0; b < nb; b++) {
                blk_t *B = &C->b[b];
                double *base = w == 0 ? B->xim : (w == 1 ? B->etm : B->zem);
                fl[b] = base + m * B->npt;
            }
            static const int nnf[3] = { 1, 2, 0 };   /* mo_numerics.f90:22 */
            for (int inn = 0; inn < 3; inn++) {
                mpigo_1(C, 1, fl, 0, m);
                for (int b = 0; b < nb; b++) filte(C, &C->b[b], nnf[inn], m, fl[b]);
            }
        }
    }
    free(fl);
    for (int b = 0; b < nb; b++) {
        blk_t *B = &C->b[b];
        long N1 = B->npt;
        const double *qo = d[0][b], *qa = d[1][b], *de = d[2][b];
        for (long l = 0; l < N1; l++) {
            double s = Q(qo, 1) * B->xim[l];
            s = s + Q(qo, 2) * B->etm[l];
            s = s + Q(qo, 3) * B->zem[l];
            s = s + Q(qa, 1) * B->xim[N1 + l];
            s = s + Q(qa, 2) * B->etm[N1 + l];
            s = s + Q(qa, 3) * B->zem[N1 + l];
            s = s + Q(de, 1) * B->xim[2 * N1 + l];
            s = s + Q(de, 2) * B->etm[2 * N1 + l];
            s = s + Q(de, 3) * B->zem[2 * N1 + l];
            B->yaco[l] = 3.0 / s;
        }
#undef Q
        for (int nn = 0; nn < 3; nn++) {
            const double *mt = nn == 0 ? B->xim : (nn == 1 ? B->etm : B->zem);
            int nbs = B->nbsize[nn];
            for (int ip = 0; ip < 2; ip++) {
                int i = ip * B->ijk[nn][0];
                for (int k = 0; k <= B->ijk[nn][2]; k++)
                    for (int j = 0; j <= B->ijk[nn][1]; j++) {
                        int jk = k * (B->ijk[nn][1] + 1) + j;
                        long l = IDX(B, i, j, k, nn + 1);
                        double rv[3];
                        for (int c = 0; c < 3; c++) rv[c] = B->yaco[l] * mt[c * N1 + l];
                        double fctr = one / sqrt((rv[0] * rv[0] + rv[1] * rv[1]) + rv[2] * rv[2]);
                        for (int c = 0; c < 3; c++) B->cm[nn][(ip * 3 + c) * nbs + jk] = fctr * rv[c];
                    }
            }
        }
    }
    for (int b = 0; b < nb; b++) { free(r3[b]); for (int c = 0; c < 3; c++) free(d[c][b]); }
    free(r3); for (int c = 0; c < 3; c++) free(d[c]);
}

/* ---- sponge: mo_sponge.f90:34-88 ---------------------------------------- */
static void sponge_init(orc_case *C)
{
    const orc_params *P = &C->prm;
    for (int bi = 0; bi < C->nblk; bi++) {
        blk_t *B = &C->b[bi];
        B->lsz = -1;
        if (!P->sponge) continue;
        double tmpa = PI / P->szth0, tmpb = PI / P->szth1;
        long cnt = 0;
        double *s1 = xcalloc(B->npt, sizeof(double)), *s2 = xcalloc(B->npt, sizeof(double));
        for (long l = 0; l < B->npt; l++) {
            double x = B->xyz[l], ra3 = zero * B->xyz[B->npt + l];
            double ra0 = tmpa * (x - ((ra3 - P->doml0) + P->szth0));
            double ra1 = tmpb * (((ra3 + P->doml1) - P->szth1) - x);
            double c0 = cos(fmax(fmin(ra0, PI), zero)), c1 = cos(fmax(fmin(ra1, PI), zero));
            s1[l] = (P->szco * half) * ((two + c0) + c1);
            s2[l] = (P->szco * half) * (one + c0);
            if (s1[l] > sml) cnt++;
        }
        B->lsz = cnt - 1;
        B->lcsz = xcalloc(cnt, sizeof(long));
        B->asz = xcalloc(cnt, sizeof(double)); B->bsz = xcalloc(cnt, sizeof(double));
        long q = 0;
        for (long l = 0; l < B->npt; l++)
            if (s1[l] > sml) {
                B->lcsz[q] = l; B->asz[q] = s1[l] / B->yaco[l]; B->bsz[q] = s2[l] / B->yaco[l]; q++;
            }
        free(s1); free(s2);
    }
}

static void sponge_go(orc_case *C)
{
    for (int bi = 0; bi < C->nblk; bi++) {
        blk_t *B = &C->b[bi];
        long N1 = B->npt;
        for (long q = 0; q <= B->lsz; q++) {
            long l = B->lcsz[q];
            B->de[l] = B->de[l] + B->asz[q] * (B->qa[l] - one);
            for (int c = 1; c <= 3; c++)
                B->de[c * N1 + l] = B->de[c * N1 + l] + B->bsz[q] * (B->qa[c * N1 + l] - zero);
            B->de[4 * N1 + l] = B->de[4 * N1 + l] + B->asz[q] * (B->qa[4 * N1 + l] - hamhamm1);
        }
    }
}

/* ---- moving frame: mo_physics.f90:144-169 ----------------------------- */
static void movef(orc_case *C, double dtko, double dtk, double timo)
{
    if (C->prm.nsmf == 0) {
        double ra0 = PI / C->prm.timf;
        double ra1 = ra0 * fmin(timo, C->prm.timf), ra2 = ra0 * fmin(timo + dtko, C->prm.timf);
        double fctr = one - cos(ra1), dfdt = ra0 * sin(ra2);
        double progmf = half * (fctr + dtk * dfdt);
        for (int c = 0; c < 3; c++) C->umf[c] = progmf * C->uoo[c];
        fctr = sin(ra1); dfdt = ra0 * cos(ra2);
        progmf = (half * ra0) * (fctr + dtk * dfdt);
        for (int c = 0; c < 3; c++) C->dudtmf[c] = progmf * C->uoo[c];
    } else {
        for (int c = 0; c < 3; c++) { C->umf[c] = C->uoo[c]; C->dudtmf[c] = zero; }
    }
}

/* ---- primitives: mo_canard_driver.f90:267-275 --------------------------- */
static void primitives(orc_case *C, blk_t *B)
{
    long N1 = B->npt;
    double *q = B->qa, *de = B->de;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long l = 0; l < N1; l++) {
        de[l] = one / q[l];
        de[N1 + l] = q[N1 + l] * de[l];
        de[2 * N1 + l] = q[2 * N1 + l] * de[l];
        de[3 * N1 + l] = q[3 * N1 + l] * de[l];
        B->p[l] = gamm1 * (q[4 * N1 + l] - half * ((q[N1 + l] * de[N1 + l] + q[2 * N1 + l] * de[2 * N1 + l])
                                                  + q[3 * N1 + l] * de[3 * N1 + l]));
        de[4 * N1 + l] = (GAM * B->p[l]) * de[l];
        B->ss[l] = (C->srefp1dre * pow(de[4 * N1 + l], 1.5)) / (de[4 * N1 + l] + C->srefoo);
    }
}

/* calc_time_step: mo_physics.f90:171-211 */
static double calc_time_step(orc_case *C)
{
    double res1 = 0, res2 = 0;
    int first = 1;
    for (int bi = 0; bi < C->nblk; bi++) {
        blk_t *B = &C->b[bi];
        long N1 = B->npt;
        for (long l = 0; l < N1; l++) {
            const double *xi = B->xim, *et = B->etm, *ze = B->zem, *de = B->de;
            double r1 = xi[l] * xi[l] + xi[N1 + l] * xi[N1 + l];
            r1 = r1 + xi[2 * N1 + l] * xi[2 * N1 + l];
            r1 = r1 + et[l] * et[l]; r1 = r1 + et[N1 + l] * et[N1 + l];
            r1 = r1 + et[2 * N1 + l] * et[2 * N1 + l];
            r1 = r1 + ze[l] * ze[l]; r1 = r1 + ze[N1 + l] * ze[N1 + l];
            r1 = r1 + ze[2 * N1 + l] * ze[2 * N1 + l];
            double u0 = de[N1 + l] + C->umf[0], u1 = de[2 * N1 + l] + C->umf[1], u2 = de[3 * N1 + l] + C->umf[2];
            double r2 = fabs((xi[l] * u0 + xi[N1 + l] * u1) + xi[2 * N1 + l] * u2)
                        + fabs((et[l] * u0 + et[N1 + l] * u1) + et[2 * N1 + l] * u2);
            r2 = r2 + fabs((ze[l] * u0 + ze[N1 + l] * u1) + ze[2 * N1 + l] * u2);
            double ssi = fabs(B->yaco[l]);
            double v1 = (sqrt(de[4 * N1 + l] * r1) + r2) * ssi;
            double v2 = (((de[l] * B->ss[l]) * r1) * ssi) * ssi;
            if (first || v1 > res1) res1 = v1;
            if (first || v2 > res2) res2 = v2;
            first = 0;
        }
    }
    double ra0 = C->prm.cfl / res1, ra1 = ra0;
    if (C->prm.viscous) ra1 = half / res2;
    return fmin(ra0, ra1);
}

/* calc_viscous_shear_stress: mo_physics.f90:216-301 */
static void viscous_stress(orc_case *C)
{
    int nb = C->nblk;
    /* rrv per block: (npt, 3, 4) for m = 2..5 */
    double **fld = xcalloc(nb, sizeof(double *));
    for (int b = 0; b < nb; b++) {
        blk_t *B = &C->b[b];
        long N1 = B->npt;
        memcpy(B->de, B->ss, sizeof(double) * N1);   /* de(:,1) = mu */
    }
    for (int m = 1; m <= 4; m++) {
        for (int b = 0; b < nb; b++) {
            blk_t *B = &C->b[b];
            long N1 = B->npt;
            double *r = B->rrv + (long)(m - 1) * 3 * N1;
            memcpy(r, B->de + m * N1, sizeof(double) * N1);
            fld[b] = r;
        }
        mpigo_1(C, 0, fld, 0, m);
    }
    free(fld);
    for (int b = 0; b < nb; b++) {
        blk_t *B = &C->b[b];
        long N1 = B->npt;
        double *g[4];
        for (int m = 1; m <= 4; m++) {
            double *r = B->rrv + (long)(m - 1) * 3 * N1;
            deriv(C, B, 2, m, r, r + 2 * N1);
            deriv(C, B, 1, m, r, r + N1);
            deriv(C, B, 0, m, r, r);
            g[m - 1] = r;
        }
        double *T = B->tau;  /* txx,tyy,tzz,txy,tyz,tzx,hxx,hyy,hzz */
        double *txx = T, *tyy = T + N1, *tzz = T + 2 * N1, *txy = T + 3 * N1, *tyz = T + 4 * N1,
               *tzx = T + 5 * N1, *hxx = T + 6 * N1, *hyy = T + 7 * N1, *hzz = T + 8 * N1;
        const double *xi = B->xim, *et = B->etm, *ze = B->zem;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (long l = 0; l < N1; l++) {
#define CT(c, G) ((xi[(c) * N1 + l] * G[l] + et[(c) * N1 + l] * G[N1 + l]) + ze[(c) * N1 + l] * G[2 * N1 + l])
            double u = B->de[N1 + l], v = B->de[2 * N1 + l], w = B->de[3 * N1 + l];
            double vtxx = CT(0, g[0]), vhzz = CT(1, g[0]), vtzx = CT(2, g[0]);
            double vtxy = CT(0, g[1]), vtyy = CT(1, g[1]), vhxx = CT(2, g[1]);
            double vhyy = CT(0, g[2]), vtyz = CT(1, g[2]), vtzz = CT(2, g[2]);
            double s1 = CT(0, g[3]), s2 = CT(1, g[3]), s3 = CT(2, g[3]);
#undef CT
            double r1 = B->de[l] * B->yaco[l];
            double r2 = gamm1prndtli * r1;
            double d5 = twothirds * ((vtxx + vtyy) + vtzz);
            B->de[4 * N1 + l] = d5;
            txx[l] = r1 * ((vtxx + vtxx) - d5);
            tyy[l] = r1 * ((vtyy + vtyy) - d5);
            tzz[l] = r1 * ((vtzz + vtzz) - d5);
            txy[l] = r1 * (vtxy + vhzz);
            tyz[l] = r1 * (vtyz + vhxx);
            tzx[l] = r1 * (vtzx + vhyy);
            hxx[l] = ((r2 * s1 + u * txx[l]) + v * txy[l]) + w * tzx[l];
            hyy[l] = ((r2 * s2 + u * txy[l]) + v * tyy[l]) + w * tyz[l];
            hzz[l] = ((r2 * s3 + u * tzx[l]) + v * tyz[l]) + w * tzz[l];
        }
    }
}

/* calc_fluxes: mo_physics.f90:306-402 */
static void fluxes(orc_case *C)
{
    int nb = C->nblk;
    const int vis = C->prm.viscous;
    const double *umf = C->umf;
    for (int b = 0; b < nb; b++) {
        blk_t *B = &C->b[b];
        long N1 = B->npt;
        const double *xi = B->xim, *et = B->etm, *ze = B->zem, *q = B->qa, *p = B->p;
        double *rr = B->rrf;  /* (npt,3,5) */
        const double *T = B->tau;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (long l = 0; l < N1; l++) {
            double *de = B->de;
            double r1 = de[N1 + l] + umf[0], r2 = de[2 * N1 + l] + umf[1], r3 = de[3 * N1 + l] + umf[2];
            double s[3], mt[3][3];
            for (int c = 0; c < 3; c++) { mt[0][c] = xi[c * N1 + l]; mt[1][c] = et[c * N1 + l]; mt[2][c] = ze[c * N1 + l]; }
            for (int d = 0; d < 3; d++) s[d] = (mt[d][0] * r1 + mt[d][1] * r2) + mt[d][2] * r3;
            double pl = p[l];
            double tx[3][3];  /* rows: momentum x,y,z -> (t_x., t_.y, t_.z) */
            if (vis) {
                tx[0][0] = T[l]; tx[0][1] = T[3 * N1 + l]; tx[0][2] = T[5 * N1 + l];
                tx[1][0] = T[3 * N1 + l]; tx[1][1] = T[N1 + l]; tx[1][2] = T[4 * N1 + l];
                tx[2][0] = T[5 * N1 + l]; tx[2][1] = T[4 * N1 + l]; tx[2][2] = T[2 * N1 + l];
            }
            for (int d = 0; d < 3; d++) {
                long o = (long)d * N1 + l;
                rr[0 * 3 * N1 + o] = q[l] * s[d];
                for (int c = 0; c < 3; c++) {
                    double v = q[(c + 1) * N1 + l] * s[d] + mt[d][c] * pl;
                    if (vis) {
                        v = v - mt[d][0] * tx[c][0];
                        v = v - mt[d][1] * tx[c][1];
                        v = v - mt[d][2] * tx[c][2];
                    }
                    rr[(c + 1) * 3 * N1 + o] = v;
                }
            }
            double e5 = q[4 * N1 + l] + pl;
            de[4 * N1 + l] = e5;
            for (int d = 0; d < 3; d++) {
                double v = e5 * s[d] - pl * ((umf[0] * mt[d][0] + umf[1] * mt[d][1]) + umf[2] * mt[d][2]);
                if (vis) {
                    v = v - mt[d][0] * T[6 * N1 + l];
                    v = v - mt[d][1] * T[7 * N1 + l];
                    v = v - mt[d][2] * T[8 * N1 + l];
                }
                rr[4 * 3 * N1 + (long)d * N1 + l] = v;
            }
        }
    }
    const double **f3 = xcalloc(3 * nb, sizeof(double *));
    for (int m = 0; m < 5; m++) {
        for (int b = 0; b < nb; b++) {
            long N1 = C->b[b].npt;
            for (int nn = 0; nn < 3; nn++) f3[b * 3 + nn] = C->b[b].rrf + (long)m * 3 * N1 + nn * N1;
        }
        mpigo(C, 0, f3, 0, m);
    }
    free(f3);
    for (int b = 0; b < nb; b++) {
        blk_t *B = &C->b[b];
        long N1 = B->npt;
        for (int m = 0; m < 5; m++) {
            double *r = B->rrf + (long)m * 3 * N1;
            deriv(C, B, 0, m, r, r);
            deriv(C, B, 1, m, r + N1, r + N1);
            deriv(C, B, 2, m, r + 2 * N1, r + 2 * N1);
            for (long l = 0; l < N1; l++) B->de[m * N1 + l] = (r[l] + r[N1 + l]) + r[2 * N1 + l];
        }
    }
}

/* ---- the time step: mo_canard_driver.f90:226-371 ---------------------- */
int orc_filter_phase(orc_case *C)
{
    static const int nnf[3] = { 1, 2, 0 };   /* eta, zeta, xi */
    double **fl = xcalloc(C->nblk, sizeof(double *));
    for (int m = 0; m < 5; m++) {
        for (int b = 0; b < C->nblk; b++) fl[b] = C->b[b].qa + m * C->b[b].npt;
        for (int inn = 0; inn < 3; inn++) {
            mpigo_1(C, 1, fl, 0, m);
            for (int b = 0; b < C->nblk; b++) filte(C, &C->b[b], nnf[inn], m, fl[b]);
        }
    }
    free(fl);
    for (int b = 0; b < C->nblk; b++) memcpy(C->b[b].qo, C->b[b].qa, sizeof(double) * 5 * C->b[b].npt);
    return 0;
}

int orc_stage(orc_case *C, int nk)
{
    const int nkrk = 4;
    double dtko = C->dt * (double)(nk - 2 > 0 ? (nk - 2 < 1 ? nk - 2 : 1) : 0) / (nkrk - nk + 3);
    double dtk = C->dt * (double)(nk - 1 < 1 ? nk - 1 : 1) / (nkrk - nk + 2);
    movef(C, dtko, dtk, C->timo);
    for (int b = 0; b < C->nblk; b++) primitives(C, &C->b[b]);
    if (nk == 1) {
        if (C->n % 10 == 1) {
            C->ndt = C->n;
            C->dts = C->dte;
            C->dte = calc_time_step(C);
        }
        double s = sin((0.05 * PI) * (double)(C->n - C->ndt));
        C->dt = C->dts + (C->dte - C->dts) * (s * s);
    }
    if (C->prm.viscous) viscous_stress(C);
    fluxes(C);
    gcbc_apply(C);
    if (C->prm.sponge) sponge_go(C);
    dtko = C->dt * (double)(nk - 1 < 1 ? nk - 1 : 1) / (nkrk - nk + 2);
    dtk = C->dt / (nkrk - nk + 1);
    movef(C, dtko, dtk, C->timo);
    for (int b = 0; b < C->nblk; b++) {
        blk_t *B = &C->b[b];
        long N1 = B->npt;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (long l = 0; l < N1; l++) {
            double r = dtk * B->yaco[l];
            for (int m = 0; m < 5; m++) B->qa[m * N1 + l] = B->qo[m * N1 + l] - r * B->de[m * N1 + l];
        }
    }
    wall_update(C);
    average_junctions(C, 0);
    average_junctions(C, 1);
    average_surface(C);
    return 0;
}

int orc_run(orc_case *C, int nsteps)
{
    for (int s = 0; s < nsteps; s++) {
        orc_filter_phase(C);
        for (int nk = 1; nk <= 4; nk++) orc_stage(C, nk);
        C->n = C->n + 1;
        C->timo = C->timo + C->dt;
    }
    return 0;
}

/* ---- setup ---------------------------------------------------------------- */
static int bad(const char *msg) { snprintf(g_err, sizeof g_err, "%s", msg); return -1; }

static int validate(const orc_blockdesc *d, int nblk)
{
    for (int b = 0; b < nblk; b++) {
        int n[3] = { d[b].lxi, d[b].let, d[b].lze };
        for (int nn = 0; nn < 3; nn++) {
            if (n[nn] < 11) return bad("every direction needs at least 12 points");
            for (int ip = 0; ip < 2; ip++) {
                int np = d[b].nbc[nn][ip], P = d[b].mcd[nn][ip];
                if (np != 10 && np != 20 && np != 25 && np != 30 && np != 35 && np != 40 && np != 45)
                    return bad("unknown boundary code");
                if (np == 30 || np == 35 || np == 40 || np == 45) {
                    if (P < 0 || P >= nblk) return bad("interface face without partner");
                    const orc_blockdesc *q = &d[P];
                    if (q->mcd[nn][1 - ip] != b) return bad("partner face does not point back");
                    if (q->nbc[nn][1 - ip] != np) return bad("partner face has another code");
                    int m[3] = { q->lxi, q->let, q->lze };
                    for (int t = 0; t < 3; t++) if (t != nn && m[t] != n[t]) return bad("partner face size differs");
                }
            }
        }
    }
    return 0;
}

orc_case *orc_create(const orc_params *p, const orc_blockdesc *d, int nblk)
{
    init_constants();
    if (validate(d, nblk)) return NULL;
    orc_case *C = xcalloc(1, sizeof(orc_case));
    C->prm = *p;
    C->nblk = nblk;
    C->b = xcalloc(nblk, sizeof(blk_t));
    init_extra(C);
    /* physics scalars: mo_physics.f90:101-111 */
    double amoo = sqrt((p->amach1 * p->amach1 + p->amach2 * p->amach2) + p->amach3 * p->amach3);
    C->reoo = p->reoo;
    if (amoo > sml) C->reoo = C->reoo / amoo;
    C->srefoo = 111 / p->tempoo;
    C->srefp1dre = (C->srefoo + one) / C->reoo;
    C->uoo[0] = p->amach1; C->uoo[1] = p->amach2; C->uoo[2] = p->amach3;
    for (int b = 0; b < nblk; b++) {
        blk_t *B = &C->b[b];
        B->lxi = d[b].lxi; B->let = d[b].let; B->lze = d[b].lze;
        memcpy(B->nbc, d[b].nbc, sizeof B->nbc); memcpy(B->mcd, d[b].mcd, sizeof B->mcd);
        B->npt = (long)(B->lxi + 1) * (B->let + 1) * (B->lze + 1);
        int t[3][3] = { { B->lxi, B->let, B->lze }, { B->let, B->lze, B->lxi }, { B->lze, B->lxi, B->let } };
        memcpy(B->ijk, t, sizeof t);
        for (int nn = 0; nn < 3; nn++) B->nbsize[nn] = (B->ijk[nn][1] + 1) * (B->ijk[nn][2] + 1);
        for (int nn = 0; nn < 3; nn++)
            for (int ip = 0; ip < 2; ip++) {
                int np = B->nbc[nn][ip];
                int v = (np == BC_ISTR || np == BC_ISUB || np == BC_PER) ? 1 : 0;
                B->ndf[nn][ip][0] = B->ndf[nn][ip][1] = v;
            }
        for (int nt = 0; nt < 2; nt++)
            for (int nn = 0; nn < 3; nn++) {
                int n = B->ijk[nn][0] + 1, nbs = B->nbsize[nn];
                B->xl[nt][nn] = xcalloc(2 * n, sizeof(double));
                B->xu[nt][nn] = xcalloc(3 * n, sizeof(double));
                penta(C, B->xl[nt][nn], B->xu[nt][nn], n, B->ndf[nn][0][nt], B->ndf[nn][1][nt], nt);
                B->send[nt][nn] = xcalloc((size_t)nbs * (2 + nt) * 2, sizeof(double));
                B->recv[nt][nn] = xcalloc((size_t)nbs * (2 + nt) * 2 * 5, sizeof(double));
            }
        for (int nn = 0; nn < 3; nn++) {
            B->drva[nn] = xcalloc((size_t)B->nbsize[nn] * 10, sizeof(double));
            B->drvb[nn] = xcalloc((size_t)B->nbsize[nn] * 10, sizeof(double));
            B->cm[nn] = xcalloc((size_t)B->nbsize[nn] * 6, sizeof(double));
        }
        long N1 = B->npt;
        B->xyz = xcalloc(3 * N1, sizeof(double));
        B->xim = xcalloc(3 * N1, sizeof(double)); B->etm = xcalloc(3 * N1, sizeof(double));
        B->zem = xcalloc(3 * N1, sizeof(double)); B->yaco = xcalloc(N1, sizeof(double));
        B->qa = xcalloc(5 * N1, sizeof(double)); B->qo = xcalloc(5 * N1, sizeof(double));
        B->de = xcalloc(5 * N1, sizeof(double)); B->p = xcalloc(N1, sizeof(double));
        B->ss = xcalloc(N1, sizeof(double)); B->avg = xcalloc(N1, sizeof(double));
        B->rrv = xcalloc(12 * N1, sizeof(double)); B->rrf = xcalloc(15 * N1, sizeof(double));
        B->tau = xcalloc(9 * N1, sizeof(double));
    }
    search_junctions(C);
    return C;
}

int orc_set_grid(orc_case *C, int b, const double *xyz)
{
    memcpy(C->b[b].xyz, xyz, sizeof(double) * 3 * C->b[b].npt);
    if (b == C->nblk - 1) {   /* all blocks set: metrics, gcbc, sponge */
        grid_metrics(C);
        gcbc_init(C);
        sponge_init(C);
    }
    return 0;
}

void orc_get_metrics(orc_case *C, int b, double *xim, double *etm, double *zem, double *yaco)
{
    blk_t *B = &C->b[b];
    if (xim) memcpy(xim, B->xim, sizeof(double) * 3 * B->npt);
    if (etm) memcpy(etm, B->etm, sizeof(double) * 3 * B->npt);
    if (zem) memcpy(zem, B->zem, sizeof(double) * 3 * B->npt);
    if (yaco) memcpy(yaco, B->yaco, sizeof(double) * B->npt);
}

void orc_set_state(orc_case *C, int b, const double *qa) { memcpy(C->b[b].qa, qa, sizeof(double) * 5 * C->b[b].npt); }
void orc_get_state(orc_case *C, int b, double *qa) { memcpy(qa, C->b[b].qa, sizeof(double) * 5 * C->b[b].npt); }

void orc_set_clock(orc_case *C, int n, int ndt, double dt, double dts, double dte, double timo)
{
    C->n = n; C->ndt = ndt; C->dt = dt; C->dts = dts; C->dte = dte; C->timo = timo;
}
void orc_get_clock(orc_case *C, int *n, int *ndt, double *dt, double *dts, double *dte, double *timo)
{
    *n = C->n; *ndt = C->ndt; *dt = C->dt; *dts = C->dts; *dte = C->dte; *timo = C->timo;
}

void orc_deriv_all(orc_case *C, int nn, const double *const *in, double *const *out)
{
    mpigo_1(C, 0, (double *const *)in, 0, 0);
    for (int b = 0; b < C->nblk; b++) deriv(C, &C->b[b], nn, 0, in[b], out[b]);
}

void orc_filter_all(orc_case *C, int nn, double *const *f)
{
    mpigo_1(C, 1, f, 0, 0);
    for (int b = 0; b < C->nblk; b++) filte(C, &C->b[b], nn, 0, f[b]);
}

int orc_coefs(orc_case *C, double *o, int nout)
{
    int k = 0;
#define PUT(x) do { if (k < nout) o[k] = (x); k++; } while (0)
    PUT(C->alphf); PUT(C->betf); PUT(C->fa); PUT(C->fb); PUT(C->fc);
    for (int r = 0; r < 3; r++) for (int c = 0; c < 5; c++) PUT(C->albef[r][c][0]);
    for (int r = 0; r < 3; r++) for (int c = 0; c < 5; c++) PUT(C->fbc[r][c]);
    for (int nt = 0; nt < 2; nt++) for (int r = 0; r < 2; r++) for (int i = 0; i < 12; i++) PUT(C->pbci[nt][r][i]);
    for (int nt = 0; nt < 2; nt++) for (int r = 0; r < 2; r++) for (int i = 0; i < 12; i++) PUT(C->pbco[nt][r][i]);
    for (int nt = 0; nt < 2; nt++) for (int r = 0; r < 2; r++) PUT(C->pbcot[nt][r]);
    for (int i = 0; i < MBCI; i++) PUT(C->sbci[i]);
#undef PUT
    return k;
}

void orc_penta(orc_case *C, int b, int nn, int nt, double *xl, double *xu)
{
    blk_t *B = &C->b[b];
    int n = B->ijk[nn][0] + 1;
    memcpy(xl, B->xl[nt][nn], sizeof(double) * 2 * n);
    memcpy(xu, B->xu[nt][nn], sizeof(double) * 3 * n);
}

/* ---- test hooks for pinning ------------------------------------------------
 * orc_gcbc_mats: the characteristic transforms xtq2r and xtr2q of
 * mo_gcbc.f90:504-611 at one state q (5 conserved variables), unit normal
 * cm and frame velocity umf; row-major 5x5 each.                          */
void orc_gcbc_mats(const double cm[3], const double q[5], const double umf[3], double *q2r, double *r2q)
{
    init_constants();
    double ri = one / q[0];
    double p = gamm1 * (q[4] - half * ((q[1] * (q[1] * ri) + q[2] * (q[2] * ri)) + q[3] * (q[3] * ri)));
    elem_t e;
    eleme(&e, cm, q[0], q + 1, p, umf);
    double xt[5][5];
    xtq2r(&e, cm, xt);
    memcpy(q2r, xt, sizeof xt);
    xtr2q(&e, cm, xt);
    memcpy(r2q, xt, sizeof xt);
}

/* orc_stage_tail: the end of an RK stage on the current state qa -- wall
 * condition update, point and line junction averaging, surface averaging
 * (mo_canard_driver.f90:340-355) -- with the frame velocity umf.           */
void orc_stage_tail(orc_case *C, const double umf[3])
{
    for (int c = 0; c < 3; c++) C->umf[c] = umf[c];
    wall_update(C);
    average_junctions(C, 0);
    average_junctions(C, 1);
    average_surface(C);
}

void orc_destroy(orc_case *C)
{
    if (!C) return;
    for (int b = 0; b < C->nblk; b++) {
        blk_t *B = &C->b[b];
        for (int nt = 0; nt < 2; nt++)
            for (int nn = 0; nn < 3; nn++) { free(B->xl[nt][nn]); free(B->xu[nt][nn]); free(B->send[nt][nn]); free(B->recv[nt][nn]); }
        for (int nn = 0; nn < 3; nn++) { free(B->drva[nn]); free(B->drvb[nn]); free(B->cm[nn]); }
        free(B->xyz); free(B->xim); free(B->etm); free(B->zem); free(B->yaco);
        free(B->qa); free(B->qo); free(B->de); free(B->p); free(B->ss); free(B->avg);
        free(B->rrv); free(B->rrf); free(B->tau);
        free(B->sbcc); free(B->nrr); free(B->lcsz); free(B->asz); free(B->bsz);
    }
    free(C->b);
    free(C->imjp); free(C->jptag); free(C->jjp); free(C->imjl); free(C->jltag); free(C->jjl);
    free(C);
}
