/*
 * CPU oracle for the FRiBoIS2D fractional-step fluid step (one MPI rank).
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference Fortran,
 * loop order and expression order kept, used as the parity checker by the
 * tests, by __graft_entry__.smoke() and as bench.py's cpu_baseline.  It is
 * never linked into or called by the product library.
 *
 * Every function cites the reference file:line it follows.  All REALs of the
 * reference are FP64 (src/makefile:22-23, -fdefault-real-8); this file is
 * compiled with -ffp-contract=off so no FMA contraction is introduced.
 *
 * LAPACK DGTTRF/DGTTRS and reference-BLAS DGEMM are restated from their
 * published netlib algorithms (LAPACK 3.8.0 / reference BLAS, the versions
 * pinned by moduleload.sh:7-8); OpenBLAS's kernels order the DGEMM sums
 * differently, so the GEMM part matches the reference only to rounding.
 *
 * One deliberate departure, shared with the GPU library: the Neumann null
 * mode of the z operator (eigenvalue wr ~ 0, phcalc_outflow.F90:37) makes its
 * j-system singular; the reference lets roundoff pick an additive constant.
 * Here that mode is gauge-fixed by setting its last unknown to zero.
 * Velocities and the mean-free pressure are independent of that constant.
 */
#include "fribo_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define DIMSD 9

struct orc_state {
    orc_params p;
    int n2, n3, n2m, n3m;
    double dx2, dx3, dx2q, dx3q, ren;
    double gam[4], rom[4], alm[4];
    double dt, dt_o, beta, al, ga, ro, cflm, q3avg;
    double qinf, qout;
    /* 1-D arrays, 1-based, with slack */
    double *rc, *rm, *zz, *zm, *g2rc, *g2rm, *g3rc, *g3rm;
    double *udx2c, *udx2m, *udx3c, *udx3m;
    int *jmv, *jpv, *jmc, *jpc, *kmv, *kpv, *kmc, *kpc;
    double *am2j, *ac2j, *ap2j, *am3j, *ac3j, *ap3j;
    double *am3ck, *ac3ck, *ap3ck, *am3sk, *ac3sk, *ap3sk;
    double *amphj, *acphj, *apphj, *amphk, *acphk, *apphk;
    double *zmm, *zmmt, *wr; /* zmm (n3m x n3m), wr 1-based */
    int null_mode;
    /* fields */
    double *q2, *q3, *pr, *dph, *q2f, *q3f; /* (1:n2,-2:n3m+3) */
    double *fey, *fez;                      /* (1:n2,-2:n3m+2) */
    double *ru2, *ru3, *rhs2, *rhs3, *rhs, *qcap; /* (1:n2,1:n3m) */
    double *qb2s, *qb3s, *qb2n, *qb3n, *dqb2s, *dqb3s, *dqb2n, *dqb3n;
    double *dq2x2o, *dq3x2o;
    /* scratch for the Poisson transform */
    double *hmn, *vmn, *hmnt;
    /* markers */
    int nlm;
    double *lmcc, *lmvel, *lmA;
    int *lmind, *eulins;
    double *phi, *phipred, *cl;
};

/* Fortran-layout index helpers. */
#define GX(s, j, k) ((size_t)((j) - 1) + (size_t)((k) + 2) * (size_t)(s)->n2)
#define RX(s, j, k) ((size_t)((j) - 1) + (size_t)((k) - 1) * (size_t)(s)->n2)
#define Q2(j, k) s->q2[GX(s, j, k)]
#define Q3(j, k) s->q3[GX(s, j, k)]
#define PR(j, k) s->pr[GX(s, j, k)]
#define DPH(j, k) s->dph[GX(s, j, k)]
#define Q2F(j, k) s->q2f[GX(s, j, k)]
#define Q3F(j, k) s->q3f[GX(s, j, k)]
#define FEY(j, k) s->fey[GX(s, j, k)]
#define FEZ(j, k) s->fez[GX(s, j, k)]
#define RU2(j, k) s->ru2[RX(s, j, k)]
#define RU3(j, k) s->ru3[RX(s, j, k)]
#define RHS2(j, k) s->rhs2[RX(s, j, k)]
#define RHS3(j, k) s->rhs3[RX(s, j, k)]
#define RHS(j, k) s->rhs[RX(s, j, k)]
#define QCAP(j, k) s->qcap[RX(s, j, k)]
/* markers: lmcc(nlm,2), lmind(nlm,2,2), eulins(nlm,2,2,9), phi(nlm,2,9), cl(nlm,2) */
#define LMCC(n, c) s->lmcc[(size_t)((n) - 1) + (size_t)((c) - 1) * s->nlm]
#define LMVEL(n, c) s->lmvel[(size_t)((n) - 1) + (size_t)((c) - 1) * s->nlm]
#define LMIND(n, d, c) s->lmind[(size_t)((n) - 1) + (size_t)((d) - 1) * s->nlm + (size_t)((c) - 1) * 2 * s->nlm]
#define EULINS(n, d, c, m)                                                            \
    s->eulins[(size_t)((n) - 1) + (size_t)((d) - 1) * s->nlm + (size_t)((c) - 1) * 2 * s->nlm + \
              (size_t)((m) - 1) * 4 * s->nlm]
#define PHI(n, d, m) s->phi[(size_t)((n) - 1) + (size_t)((d) - 1) * s->nlm + (size_t)((m) - 1) * 2 * s->nlm]
#define PHIPRED(n, d, m) \
    s->phipred[(size_t)((n) - 1) + (size_t)((d) - 1) * s->nlm + (size_t)((m) - 1) * 2 * s->nlm]
#define CL(n, d) s->cl[(size_t)((n) - 1) + (size_t)((d) - 1) * s->nlm]

static double *dalloc(size_t n) { return (double *)calloc(n + 4, sizeof(double)); }
static int *ialloc(size_t n) { return (int *)calloc(n + 4, sizeof(int)); }

/* ------------------------------------------------------------------------ */
/* LAPACK DGTTRF (netlib LAPACK 3.8.0 dgttrf.f), 1-based arrays.             */
/* dl(1:n-1), d(1:n), du(1:n-1), du2(1:n-2), ipiv(1:n).                      */
static void dgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv)
{
    int i;
    double fact, temp;
    for (i = 1; i <= n; ++i) ipiv[i] = i;
    for (i = 1; i <= n - 2; ++i) du2[i] = 0.0;
    for (i = 1; i <= n - 2; ++i) {
        if (fabs(d[i]) >= fabs(dl[i])) {
            /* no row interchange required, eliminate dl(i) */
            if (d[i] != 0.0) {
                fact = dl[i] / d[i];
                dl[i] = fact;
                d[i + 1] = d[i + 1] - fact * du[i];
            }
        } else {
            /* interchange rows i and i+1, eliminate dl(i) */
            fact = d[i] / dl[i];
            d[i] = dl[i];
            dl[i] = fact;
            temp = du[i];
            du[i] = d[i + 1];
            d[i + 1] = temp - fact * d[i + 1];
            du2[i] = du[i + 1];
            du[i + 1] = -fact * du[i + 1];
            ipiv[i] = i + 1;
        }
    }
    if (n > 1) {
        i = n - 1;
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] != 0.0) {
                fact = dl[i] / d[i];
                dl[i] = fact;
                d[i + 1] = d[i + 1] - fact * du[i];
            }
        } else {
            fact = d[i] / dl[i];
            d[i] = dl[i];
            dl[i] = fact;
            temp = du[i];
            du[i] = d[i + 1];
            d[i + 1] = temp - fact * d[i + 1];
            ipiv[i] = i + 1;
        }
    }
}

/* LAPACK DGTTRS('N', nrhs=1) -> DGTTS2 itrans=0 (netlib dgtts2.f). */
static void dgttrs(int n, const double *dl, const double *d, const double *du,
                   const double *du2, const int *ipiv, double *b)
{
    int i, ip;
    double temp;
    if (n == 0) return;
    for (i = 1; i <= n - 1; ++i) {
        ip = ipiv[i];
        temp = b[i + 1 - ip + i] - dl[i] * b[ip];
        b[i] = b[ip];
        b[i + 1] = temp;
    }
    b[n] = b[n] / d[n];
    if (n > 1) b[n - 1] = (b[n - 1] - du[n - 1] * b[n]) / d[n - 1];
    for (i = n - 2; i >= 1; --i)
        b[i] = (b[i] - du[i] * b[i + 1] - du2[i] * b[i + 2]) / d[i];
}

/* Gauge-fixed solve of the singular Poisson mode (see file header): the
 * same elimination as DGTTRF/DGTTRS without pivoting, last unknown set to 0,
 * rows 1..n-1 satisfied exactly. */
static void tri_null_solve(int n, const double *dl, const double *d0, const double *du,
                           double *b, double *work)
{
    int i;
    double *d = work;
    for (i = 1; i <= n; ++i) d[i] = d0[i];
    /* forward elimination over rows 2..n-1 only */
    for (i = 1; i <= n - 2; ++i) {
        double fact = dl[i] / d[i];
        d[i + 1] = d[i + 1] - fact * du[i];
        b[i + 1] = b[i + 1] - fact * b[i];
    }
    b[n] = 0.0;
    if (n > 1) b[n - 1] = (b[n - 1] - du[n - 1] * b[n]) / d[n - 1];
    for (i = n - 2; i >= 1; --i) b[i] = (b[i] - du[i] * b[i + 1]) / d[i];
}

/* ------------------------------------------------------------------------ */
/* indic.F90:4-44 */
static void indic(orc_state *s)
{
    int jc, kc, n2 = s->n2, n2m = s->n2m, n3m = s->n3m;
    for (kc = 1; kc <= n3m; ++kc) {
        s->kmv[kc] = kc - 1;
        s->kpv[kc] = kc + 1;
        if (kc == 1) s->kmv[kc] = kc;
        if (kc == n3m) s->kpv[kc] = kc;
    }
    for (kc = 1; kc <= n3m; ++kc) {
        s->kpc[kc] = s->kpv[kc] - kc;
        s->kmc[kc] = kc - s->kmv[kc];
    }
    for (jc = 1; jc <= n2; ++jc) {
        s->jmv[jc] = jc - 1;
        s->jpv[jc] = jc + 1;
        if (jc == 1) s->jmv[jc] = jc;
        if (jc == n2m) s->jpv[jc] = jc;
        if (jc == n2) s->jpv[jc] = jc;
    }
    for (jc = 1; jc <= n2m; ++jc) {
        s->jpc[jc] = s->jpv[jc] - jc;
        s->jmc[jc] = jc - s->jmv[jc];
    }
}

/* cordin.F90:105-123 and 196-215: metrics from the node coordinates */
static void cordin_metrics(orc_state *s)
{
    int j, k, n2 = s->n2, n2m = s->n2m, n3 = s->n3, n3m = s->n3m;
    double dx2 = s->dx2, dx3 = s->dx3;
    for (j = 1; j <= n2m; ++j) {
        s->rm[j] = (s->rc[j] + s->rc[j + 1]) * 0.5;
        s->g2rm[j] = (s->rc[j + 1] - s->rc[j]) * dx2;
    }
    for (j = 2; j <= n2m; ++j) s->g2rc[j] = (s->rc[j + 1] - s->rc[j - 1]) * dx2 * 0.5;
    s->g2rc[1] = (s->rc[2] - s->rc[1]) * dx2;
    s->g2rc[n2] = (s->rc[n2] - s->rc[n2m]) * dx2;
    s->rm[n2] = s->rc[n2];
    for (j = 1; j <= n2m; ++j) {
        s->udx2m[j] = dx2 / s->g2rm[j];
        s->udx2c[j] = dx2 / s->g2rc[j];
    }
    s->udx2c[n2] = dx2 / s->g2rc[n2];

    for (k = 1; k <= n3m; ++k) {
        s->zm[k] = (s->zz[k] + s->zz[k + 1]) * 0.5;
        s->g3rm[k] = (s->zz[k + 1] - s->zz[k]) * dx3;
    }
    for (k = 2; k <= n3m; ++k) s->g3rc[k] = (s->zz[k + 1] - s->zz[k - 1]) * dx3 * 0.5;
    s->g3rc[1] = (s->zz[2] - s->zz[1]) * dx3;
    s->g3rc[n3] = (s->zz[n3] - s->zz[n3m]) * dx3;
    s->zm[n3] = s->zz[n3];
    for (k = 1; k <= n3m; ++k) {
        s->udx3m[k] = dx3 / s->g3rm[k];
        s->udx3c[k] = dx3 / s->g3rc[k];
    }
    s->udx3c[n3] = dx3 / s->g3rc[n3];
    /* MLS_SD_ind reads one node past the end of its grid copy when the
     * support touches the last cell (MLSint.F90:89,93); the nearest defined
     * value is used here. */
    s->rm[n2 + 1] = s->rc[n2];
    s->zm[n3 + 1] = s->zz[n3];
}

/* coetar.F90:5-144 */
static void coetar(orc_state *s)
{
    int km, kc, kp, jm, jc, jp, n2 = s->n2, n2m = s->n2m, n3 = s->n3, n3m = s->n3m;
    int inwn = s->p.inslwn ? 1 : 0, inws = s->p.inslws ? 1 : 0;
    double a33, a33m, a33p, a22, a22m, a22p, dx2q = s->dx2q, dx3q = s->dx3q;

    s->am2j[1] = 0.0; s->ap2j[1] = 0.0; s->ac2j[1] = 1.0;
    s->am2j[n2] = 0.0; s->ap2j[n2] = 0.0; s->ac2j[n2] = 1.0;
    for (jc = 2; jc <= n2m; ++jc) {
        jm = jc - 1;
        a22 = dx2q / s->g2rc[jc];
        a22p = 1.0 / s->g2rm[jc];
        a22m = 1.0 / s->g2rm[jm];
        s->ap2j[jc] = a22 * a22p;
        s->am2j[jc] = a22 * a22m;
        s->ac2j[jc] = -(a22 * a22p + a22 * a22m);
    }
    for (jc = 2; jc <= n2m - 1; ++jc) {
        jp = s->jpv[jc];
        a22 = dx2q / s->g2rm[jc];
        a22p = a22 / s->g2rc[jp];
        a22m = a22 / s->g2rc[jc];
        s->ap3j[jc] = a22p;
        s->am3j[jc] = a22m;
        s->ac3j[jc] = -(a22p + a22m);
    }
    jc = n2m; jp = jc + 1;
    a22 = dx2q / s->g2rm[jc];
    a22p = +a22 / s->g2rc[jp];
    a22m = +a22 / s->g2rc[jc];
    s->am3j[jc] = a22m;
    s->ap3j[jc] = 2.0 * inwn * a22p;
    s->ac3j[jc] = -(a22m + inwn * a22p * 2.0);
    jc = 1; jp = jc + 1;
    a22 = dx2q / s->g2rm[jc];
    a22p = +a22 / s->g2rc[jp];
    a22m = +a22 / s->g2rc[jc];
    s->ap3j[jc] = a22p;
    s->am3j[jc] = 2.0 * a22m * inws;
    s->ac3j[jc] = -(a22p + 2.0 * inws * a22m);

    s->am3ck[1] = 0.0; s->ap3ck[1] = 0.0; s->ac3ck[1] = 1.0;
    s->am3ck[n3] = 0.0; s->ap3ck[n3] = 0.0; s->ac3ck[n3] = 1.0;
    for (kc = 2; kc <= n3m; ++kc) {
        km = kc - 1;
        a33 = dx3q / s->g3rc[kc];
        a33p = 1.0 / s->g3rm[kc];
        a33m = 1.0 / s->g3rm[km];
        s->ap3ck[kc] = a33 * a33p;
        s->am3ck[kc] = a33 * a33m;
        s->ac3ck[kc] = -(s->ap3ck[kc] + s->am3ck[kc]);
    }
    for (kc = 2; kc <= n3m - 1; ++kc) {
        kp = kc + 1;
        a33 = dx3q / s->g3rm[kc];
        a33p = +a33 / s->g3rc[kp];
        a33m = +a33 / s->g3rc[kc];
        s->ap3sk[kc] = a33p;
        s->am3sk[kc] = a33m;
        s->ac3sk[kc] = -(s->ap3sk[kc] + s->am3sk[kc]);
    }
    kc = 1; kp = kc + 1;
    a33 = dx3q / s->g3rm[kc];
    a33p = +a33 / s->g3rc[kp];
    a33m = +a33 / s->g3rc[kc];
    s->ap3sk[kc] = a33p;
    s->am3sk[kc] = 0.0;
    s->ac3sk[kc] = -(a33p + a33m * 2.0);
    kc = n3m; kp = kc + 1;
    a33 = dx3q / s->g3rm[kc];
    a33p = +a33 / s->g3rc[kp];
    a33m = +a33 / s->g3rc[kc];
    s->am3sk[kc] = a33m;
    s->ap3sk[kc] = 0.0;
    s->ac3sk[kc] = -(a33m + a33p * 2.0);
}

/* phini.F90:109-145 (tridiag_matrices) */
static void tridiag_matrices(orc_state *s)
{
    int jc, jp, kc, kp, n2m = s->n2m, n3m = s->n3m;
    double ugmmm, a33icc, a33icp, a22icc, a22icp;
    for (jc = 1; jc <= n2m; ++jc) {
        jp = s->jpv[jc];
        a22icc = s->jmc[jc] * s->dx2q / s->g2rc[jc];
        a22icp = s->jpc[jc] * s->dx2q / s->g2rc[jp];
        ugmmm = 1.0 / s->g2rm[jc];
        s->amphj[jc] = a22icc * ugmmm;
        s->apphj[jc] = a22icp * ugmmm;
        s->acphj[jc] = -(s->amphj[jc] + s->apphj[jc]);
    }
    for (kc = 1; kc <= n3m; ++kc) {
        kp = s->kpv[kc];
        a33icc = s->kmc[kc] * s->dx3q / s->g3rc[kc];
        a33icp = s->kpc[kc] * s->dx3q / s->g3rc[kp];
        ugmmm = 1.0 / s->g3rm[kc];
        s->amphk[kc] = a33icc * ugmmm;
        s->apphk[kc] = a33icp * ugmmm;
        s->acphk[kc] = -(s->amphk[kc] + s->apphk[kc]);
    }
}

/* ------------------------------------------------------------------------ */
orc_state *orc_create(const orc_params *p, const double *rc, const double *zz,
                      const double *zmm, const double *zmmt, const double *wr)
{
    int j, k, m;
    orc_state *s;
    if (!p || p->n2 < 5 || p->n3 < 5 || (p->nsst != 1 && p->nsst != 3)) return NULL;
    s = (orc_state *)calloc(1, sizeof(orc_state));
    s->p = *p;
    s->n2 = p->n2;
    s->n3 = p->n3;
    s->n2m = p->n2 - 1;
    s->n3m = p->n3 - 1;
    /* fluid_pre.F90:27-70 */
    s->dx2 = 1.0 / (p->rext2 / (double)s->n2m);
    s->dx3 = 1.0 / (p->alx3 / (double)s->n3m);
    s->dx2q = s->dx2 * s->dx2;
    s->dx3q = s->dx3 * s->dx3;
    s->ren = 1.0 / p->visc;
    if (p->nsst > 1) {
        s->gam[1] = 8.0 / 15.0; s->gam[2] = 5.0 / 12.0; s->gam[3] = 3.0 / 4.0;
        s->rom[1] = 0.0; s->rom[2] = -17.0 / 60.0; s->rom[3] = -5.0 / 12.0;
    } else {
        s->gam[1] = 1.5; s->gam[2] = 0.0; s->gam[3] = 0.0;
        s->rom[1] = -0.50; s->rom[2] = 0.0; s->rom[3] = 0.0;
    }
    for (m = 1; m <= p->nsst; ++m) s->alm[m] = s->gam[m] + s->rom[m];

    {
        size_t n2 = s->n2, n3 = s->n3;
        s->rc = dalloc(n2 + 2); s->rm = dalloc(n2 + 2); s->g2rc = dalloc(n2 + 2); s->g2rm = dalloc(n2 + 2);
        s->zz = dalloc(n3 + 2); s->zm = dalloc(n3 + 2); s->g3rc = dalloc(n3 + 2); s->g3rm = dalloc(n3 + 2);
        s->udx2c = dalloc(n2 + 2); s->udx2m = dalloc(n2 + 2); s->udx3c = dalloc(n3 + 2); s->udx3m = dalloc(n3 + 2);
        s->jmv = ialloc(n2 + 2); s->jpv = ialloc(n2 + 2); s->jmc = ialloc(n2 + 2); s->jpc = ialloc(n2 + 2);
        s->kmv = ialloc(n3 + 2); s->kpv = ialloc(n3 + 2); s->kmc = ialloc(n3 + 2); s->kpc = ialloc(n3 + 2);
        s->am2j = dalloc(n2 + 2); s->ac2j = dalloc(n2 + 2); s->ap2j = dalloc(n2 + 2);
        s->am3j = dalloc(n2 + 2); s->ac3j = dalloc(n2 + 2); s->ap3j = dalloc(n2 + 2);
        s->am3ck = dalloc(n3 + 2); s->ac3ck = dalloc(n3 + 2); s->ap3ck = dalloc(n3 + 2);
        s->am3sk = dalloc(n3 + 2); s->ac3sk = dalloc(n3 + 2); s->ap3sk = dalloc(n3 + 2);
        s->amphj = dalloc(n2 + 2); s->acphj = dalloc(n2 + 2); s->apphj = dalloc(n2 + 2);
        s->amphk = dalloc(n3 + 2); s->acphk = dalloc(n3 + 2); s->apphk = dalloc(n3 + 2);
        s->zmm = dalloc((size_t)s->n3m * s->n3m);
        s->zmmt = dalloc((size_t)s->n3m * s->n3m);
        s->wr = dalloc(n3 + 2);
        size_t gsz = n2 * (size_t)(s->n3m + 6), fsz = n2 * (size_t)(s->n3m + 5), rsz = n2 * (size_t)s->n3m;
        s->q2 = dalloc(gsz); s->q3 = dalloc(gsz); s->pr = dalloc(gsz); s->dph = dalloc(gsz);
        s->q2f = dalloc(gsz); s->q3f = dalloc(gsz);
        s->fey = dalloc(fsz); s->fez = dalloc(fsz);
        s->ru2 = dalloc(rsz); s->ru3 = dalloc(rsz); s->rhs2 = dalloc(rsz); s->rhs3 = dalloc(rsz);
        s->rhs = dalloc(rsz); s->qcap = dalloc(rsz);
        s->qb2s = dalloc(n2 + 2); s->qb3s = dalloc(n2 + 2); s->qb2n = dalloc(n2 + 2); s->qb3n = dalloc(n2 + 2);
        s->dqb2s = dalloc(n2 + 2); s->dqb3s = dalloc(n2 + 2); s->dqb2n = dalloc(n2 + 2); s->dqb3n = dalloc(n2 + 2);
        s->dq2x2o = dalloc(n2 + 2); s->dq3x2o = dalloc(n2 + 2);
        s->hmn = dalloc((size_t)s->n2m * s->n3m);
        s->vmn = dalloc((size_t)s->n2m * s->n3m);
        s->hmnt = dalloc((size_t)s->n2m * s->n3m);
    }
    for (j = 1; j <= s->n2; ++j) s->rc[j] = rc[j - 1];
    for (k = 1; k <= s->n3; ++k) s->zz[k] = zz[k - 1];
    for (k = 0; k < s->n3m * s->n3m; ++k) {
        s->zmm[k] = zmm[k];
        s->zmmt[k] = zmmt[k];
    }
    for (k = 1; k <= s->n3m; ++k) s->wr[k] = wr[k - 1];

    indic(s);
    cordin_metrics(s);
    tridiag_matrices(s);
    coetar(s);

    /* null mode: smallest |wr|, accepted if below 1e-9 of the largest */
    {
        double wmax = 0.0, wmin = 1e300;
        int imin = 0;
        for (k = 1; k <= s->n3m; ++k) {
            double a = fabs(s->wr[k]);
            if (a > wmax) wmax = a;
            if (a < wmin) { wmin = a; imin = k; }
        }
        s->null_mode = (wmin <= 1e-9 * wmax) ? imin : 0;
    }
    return s;
}

void orc_destroy(orc_state *s)
{
    if (!s) return;
    void *ptrs[] = {s->rc, s->rm, s->zz, s->zm, s->g2rc, s->g2rm, s->g3rc, s->g3rm, s->udx2c, s->udx2m,
                    s->udx3c, s->udx3m, s->jmv, s->jpv, s->jmc, s->jpc, s->kmv, s->kpv, s->kmc, s->kpc,
                    s->am2j, s->ac2j, s->ap2j, s->am3j, s->ac3j, s->ap3j, s->am3ck, s->ac3ck, s->ap3ck,
                    s->am3sk, s->ac3sk, s->ap3sk, s->amphj, s->acphj, s->apphj, s->amphk, s->acphk,
                    s->apphk, s->zmm, s->zmmt, s->wr, s->q2, s->q3, s->pr, s->dph, s->q2f, s->q3f,
                    s->fey, s->fez, s->ru2, s->ru3, s->rhs2, s->rhs3, s->rhs, s->qcap, s->qb2s, s->qb3s,
                    s->qb2n, s->qb3n, s->dqb2s, s->dqb3s, s->dqb2n, s->dqb3n, s->dq2x2o, s->dq3x2o,
                    s->hmn, s->vmn, s->hmnt, s->lmcc, s->lmvel, s->lmA, s->lmind, s->eulins, s->phi,
                    s->phipred, s->cl};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(s);
}

int orc_null_mode(orc_state *s) { return s ? s->null_mode : 0; }

int orc_set_markers(orc_state *s, int nlm, const double *lmcc, const double *lmvel,
                    const double *lmA)
{
    if (!s || nlm < 0) return ORC_ERR_ARG;
    if (nlm != s->nlm) {
        free(s->lmcc); free(s->lmvel); free(s->lmA); free(s->lmind); free(s->eulins);
        free(s->phi); free(s->phipred); free(s->cl);
        s->nlm = nlm;
        s->lmcc = dalloc(2 * (size_t)nlm);
        s->lmvel = dalloc(2 * (size_t)nlm);
        s->lmA = dalloc((size_t)nlm);
        s->lmind = ialloc(4 * (size_t)nlm);
        s->eulins = ialloc(4 * DIMSD * (size_t)nlm);
        s->phi = dalloc(2 * DIMSD * (size_t)nlm);
        s->phipred = dalloc(2 * DIMSD * (size_t)nlm);
        s->cl = dalloc(2 * (size_t)nlm);
    }
    if (lmcc) memcpy(s->lmcc, lmcc, 2 * sizeof(double) * nlm);
    if (lmvel) memcpy(s->lmvel, lmvel, 2 * sizeof(double) * nlm);
    if (lmA) memcpy(s->lmA, lmA, sizeof(double) * nlm);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* cfl.F90:4-51 */
static void cfl(orc_state *s)
{
    int j, k, n2m = s->n2m, n3m = s->n3m;
    double my_cflm = 1.0e-8, myq3avg = 0.0, qcf, udx2, udx3;
    for (k = 1; k <= n3m; ++k) {
        udx3 = s->udx3m[k];
        for (j = 1; j <= n2m; ++j) {
            udx2 = s->udx2m[j];
            qcf = fabs((Q2(j, k) + Q2(j + 1, k)) * 0.5 * udx2) +
                  fabs((Q3(j, k) + Q3(j, k + 1)) * 0.5 * udx3);
            my_cflm = fmax(my_cflm, qcf);
            myq3avg = myq3avg + Q3(j, k) * (s->zz[k + 1] - s->zz[k]) * (s->rc[j + 1] - s->rc[j]);
        }
    }
    s->cflm = my_cflm;
    s->q3avg = myq3avg / (s->p.rext2 * s->p.alx3);
}

/* iobcs.F90:5-77 (boucdq) */
static void boucdq(orc_state *s)
{
    int j, n2 = s->n2, n2m = s->n2m, n3 = s->n3, n3m = s->n3m;
    double dq2x2, dq3x2, area, darea, cor, dt = s->dt, ga = s->ga, ro = s->ro, cou = s->p.cou;
    for (j = 1; j <= n2; ++j) s->dqb2s[j] = 0.0;
    for (j = 1; j <= n2; ++j) s->dqb3s[j] = 0.0;
    for (j = 1; j <= n2m; ++j) {
        dq2x2 = (s->qb2n[j] - Q2(j, n3m)) * 2.0 * s->udx3c[n3];
        s->dqb2n[j] = -dt * (ga * dq2x2 + ro * s->dq2x2o[j]) * cou;
        dq3x2 = (s->qb3n[j] - Q3(j, n3m)) * s->udx3c[n3];
        s->dqb3n[j] = -dt * (ga * dq3x2 + ro * s->dq3x2o[j]) * cou;
        s->dq2x2o[j] = dq2x2;
        s->dq3x2o[j] = dq3x2;
    }
    area = 0.0; s->qout = 0.0; s->qinf = 0.0;
    for (j = 1; j <= n2m; ++j) {
        darea = s->g2rm[j] / s->dx2;
        area = area + darea;
        s->qout = s->qout + s->dqb3n[j] * darea;
        s->qinf = s->qinf + s->dqb3s[j] * darea;
    }
    cor = (s->qinf - s->qout) / area;
    for (j = 1; j <= n2m; ++j) s->dqb3n[j] = s->dqb3n[j] + cor;
}

/* iobcs.F90:81-108 (boucqt) */
static void boucqt(orc_state *s)
{
    int j;
    double darea;
    s->qout = 0.0; s->qinf = 0.0;
    for (j = 1; j <= s->n2m; ++j) {
        darea = s->g2rm[j] / s->dx2;
        s->qb2s[j] = s->qb2s[j] + s->dqb2s[j];
        s->qb3s[j] = s->qb3s[j] + s->dqb3s[j];
        s->qb2n[j] = s->qb2n[j] + s->dqb2n[j];
        s->qb3n[j] = s->qb3n[j] + s->dqb3n[j];
        s->qout = s->qout + s->qb3n[j] * darea;
        s->qinf = s->qinf + s->qb3s[j] * darea;
    }
}

/* HDNL.F90:5-110 (hdnl2), one rank: kstart=1, kend=n3m */
static void hdnl2(orc_state *s)
{
    int kc, kp, jp, jm, jc, kmm, kpp, n2m = s->n2m, n3m = s->n3m;
    double h22, h23;
    for (kc = 1; kc <= n3m; ++kc) {
        for (jc = 2; jc <= n2m; ++jc) {
            jm = jc - 1; jp = jc + 1;
            h22 = ((Q2(jp, kc) + Q2(jc, kc)) * (Q2(jp, kc) + Q2(jc, kc)) -
                   (Q2(jm, kc) + Q2(jc, kc)) * (Q2(jm, kc) + Q2(jc, kc))) * s->udx2c[jc] * 0.25;
            DPH(jc, kc) = -h22;
        }
    }
    for (kc = 2; kc <= n3m - 1; ++kc) {
        kmm = s->kmv[kc]; kpp = s->kpv[kc]; kp = kc + 1;
        for (jc = 2; jc <= n2m; ++jc) {
            jm = s->jmv[jc];
            h23 = ((Q3(jc, kp) + Q3(jm, kp)) * (Q2(jc, kpp) + Q2(jc, kc)) -
                   (Q3(jc, kc) + Q3(jm, kc)) * (Q2(jc, kc) + Q2(jc, kmm))) * s->udx3m[kc] * 0.25;
            DPH(jc, kc) = DPH(jc, kc) - h23;
        }
    }
    kc = n3m; kmm = kc - 1;
    for (jc = 2; jc <= n2m; ++jc) {
        jm = s->jmv[jc];
        h23 = (s->qb2n[jc] * (s->qb3n[jc] + s->qb3n[jm]) -
               (Q3(jc, kc) + Q3(jm, kc)) * (Q2(jc, kc) + Q2(jc, kmm)) * 0.5) * s->udx3m[kc] * 0.5;
        DPH(jc, kc) = DPH(jc, kc) - h23;
    }
    kc = 1; kp = kc + 1;
    for (jc = 2; jc <= n2m; ++jc) {
        jm = s->jmv[jc];
        h23 = ((Q3(jc, kp) + Q3(jm, kp)) * (Q2(jc, kc) + Q2(jc, kp)) * 0.5 -
               s->qb2s[jc] * (s->qb3s[jc] + s->qb3s[jm])) * s->udx3m[kc] * 0.5;
        DPH(jc, kc) = DPH(jc, kc) - h23;
    }
}

/* HDNL.F90:114-171 (hdnl3) */
static void hdnl3(orc_state *s)
{
    int jc, kc, kmm, km, kp, jp, jmm, jpp, n2m = s->n2m, n3m = s->n3m;
    double h32, h33;
    for (kc = 2; kc <= n3m - 1; ++kc) {
        kmm = kc - 1; km = s->kmv[kc]; kp = kc + 1;
        for (jc = 1; jc <= n2m; ++jc) {
            jmm = s->jmv[jc]; jpp = s->jpv[jc]; jp = jc + 1;
            h32 = (((Q2(jp, kc) + Q2(jp, km)) * (Q3(jpp, kc) + Q3(jc, kc))) -
                   ((Q2(jc, kc) + Q2(jc, km)) * (Q3(jc, kc) + Q3(jmm, kc)))) * s->udx2m[jc] * 0.25;
            h33 = ((Q3(jc, kp) + Q3(jc, kc)) * (Q3(jc, kp) + Q3(jc, kc)) -
                   (Q3(jc, kc) + Q3(jc, kmm)) * (Q3(jc, kc) + Q3(jc, kmm))) * s->udx3c[kc] * 0.25;
            QCAP(jc, kc) = -(h32 + h33);
        }
    }
}

/* INVTR.F90:4-152 (invtr2) */
static void invtr2(orc_state *s)
{
    int jc, kc, km, kp, jp, jm, n2m = s->n2m, n3m = s->n3m;
    double amm, app, acc, ammj, appj, accj, dcq2, dpx22, d22q2, d33q2, alre, udx2, codea, codeb;
    double ga = s->ga, ro = s->ro, dt = s->dt, al = s->al;
    alre = al / s->ren;
    for (kc = 2; kc <= n3m - 1; ++kc) {
        km = s->kmv[kc]; kp = s->kpv[kc];
        amm = s->am3sk[kc]; acc = s->ac3sk[kc]; app = s->ap3sk[kc];
        for (jc = 2; jc <= n2m; ++jc) {
            jm = jc - 1; jp = jc + 1;
            ammj = s->am2j[jc]; accj = s->ac2j[jc]; appj = s->ap2j[jc];
            udx2 = al * s->udx2c[jc];
            d22q2 = Q2(jp, kc) * appj + Q2(jc, kc) * accj + Q2(jm, kc) * ammj;
            d33q2 = Q2(jc, kp) * app + Q2(jc, kc) * acc + Q2(jc, km) * amm;
            dcq2 = d22q2 + d33q2;
            dpx22 = (PR(jc, kc) - PR(jm, kc)) * udx2;
            RHS2(jc, kc) = (ga * DPH(jc, kc) + ro * RU2(jc, kc) + alre * dcq2 - dpx22) * dt;
            RU2(jc, kc) = DPH(jc, kc);
        }
    }
    kc = 1; kp = kc + 1;
    codea = s->dx3q / (s->g3rc[kp] * s->g3rm[kc]);
    codeb = s->dx3q / (s->g3rc[kc] * s->g3rm[kc]);
    for (jc = 2; jc <= n2m; ++jc) {
        jp = jc + 1; jm = jc - 1;
        ammj = s->am2j[jc]; accj = s->ac2j[jc]; appj = s->ap2j[jc];
        d22q2 = Q2(jp, kc) * appj + Q2(jc, kc) * accj + Q2(jm, kc) * ammj;
        d33q2 = ((Q2(jc, kp) - Q2(jc, kc)) * codea - (Q2(jc, kc) - s->qb2s[jc]) * codeb * 2.0);
        dcq2 = d22q2 + d33q2;
        dpx22 = (PR(jc, kc) - PR(jm, kc)) * s->udx2c[jc] * al;
        RHS2(jc, kc) = +(ga * DPH(jc, kc) + ro * RU2(jc, kc) + alre * dcq2 - dpx22) * dt;
        RU2(jc, kc) = DPH(jc, kc);
    }
    kc = n3m; km = kc - 1; kp = kc + 1;
    codea = s->dx3q / (s->g3rc[kp] * s->g3rm[kc]);
    codeb = s->dx3q / (s->g3rc[kc] * s->g3rm[kc]);
    for (jc = 2; jc <= n2m; ++jc) {
        jm = jc - 1; jp = jc + 1;
        ammj = s->am2j[jc]; accj = s->ac2j[jc]; appj = s->ap2j[jc];
        d22q2 = Q2(jp, kc) * appj + Q2(jc, kc) * accj + Q2(jm, kc) * ammj;
        d33q2 = ((s->qb2n[jc] - Q2(jc, kc)) * codea * 2.0 - (Q2(jc, kc) - Q2(jc, km)) * codeb);
        dcq2 = d22q2 + d33q2;
        dpx22 = (PR(jc, kc) - PR(jm, kc)) * s->udx2c[jc] * al;
        RHS2(jc, kc) = +(ga * DPH(jc, kc) + ro * RU2(jc, kc) + alre * dcq2 - dpx22) * dt;
        RU2(jc, kc) = DPH(jc, kc);
    }
    for (kc = 1; kc <= n3m; ++kc)
        for (jc = 2; jc <= n2m; ++jc) Q2F(jc, kc) = Q2(jc, kc) + RHS2(jc, kc);
}

/* INVTR.F90:156-280 (invtr3) */
static void invtr3(orc_state *s)
{
    int jc, kc, km, kp, jp, jm, n2m = s->n2m, n3m = s->n3m;
    double udx3, dq32, dq33, dcq3, dpx33, app, acc, amm, appj, accj, ammj, alre;
    double ga = s->ga, ro = s->ro, dt = s->dt, al = s->al, q3lw = s->p.q3lw, q3uw = s->p.q3uw;
    alre = al / s->ren;
    for (kc = 2; kc <= n3m; ++kc) {
        km = kc - 1; kp = kc + 1;
        udx3 = al * s->udx3c[kc];
        amm = s->am3ck[kc]; acc = s->ac3ck[kc]; app = s->ap3ck[kc];
        jc = 1; jp = jc + 1;
        ammj = s->am3j[jc]; accj = s->ac3j[jc]; appj = s->ap3j[jc];
        dq32 = Q3(jp, kc) * appj + Q3(jc, kc) * accj + q3lw * ammj;
        dq33 = Q3(jc, kp) * app + Q3(jc, kc) * acc + Q3(jc, km) * amm;
        dcq3 = dq32 + dq33;
        dpx33 = (PR(jc, kc) - PR(jc, km)) * udx3;
        RHS3(jc, kc) = (ga * QCAP(jc, kc) + ro * RU3(jc, kc) + alre * dcq3 - dpx33) * dt;
        RU3(jc, kc) = QCAP(jc, kc);
        for (jc = 2; jc <= n2m - 1; ++jc) {
            jm = s->jmv[jc]; jp = s->jpv[jc];
            ammj = s->am3j[jc]; accj = s->ac3j[jc]; appj = s->ap3j[jc];
            dq32 = Q3(jp, kc) * appj + Q3(jc, kc) * accj + Q3(jm, kc) * ammj;
            dq33 = Q3(jc, kp) * app + Q3(jc, kc) * acc + Q3(jc, km) * amm;
            dcq3 = dq32 + dq33;
            dpx33 = (PR(jc, kc) - PR(jc, km)) * udx3;
            RHS3(jc, kc) = (ga * QCAP(jc, kc) + ro * RU3(jc, kc) + alre * dcq3 - dpx33) * dt;
            RU3(jc, kc) = QCAP(jc, kc);
        }
        jc = n2m; jm = jc - 1;
        ammj = s->am3j[jc]; accj = s->ac3j[jc]; appj = s->ap3j[jc];
        dq32 = q3uw * appj + Q3(jc, kc) * accj + Q3(jm, kc) * ammj;
        dq33 = Q3(jc, kp) * app + Q3(jc, kc) * acc + Q3(jc, km) * amm;
        dcq3 = dq32 + dq33;
        dpx33 = (PR(jc, kc) - PR(jc, km)) * udx3;
        RHS3(jc, kc) = (ga * QCAP(jc, kc) + ro * RU3(jc, kc) + alre * dcq3 - dpx33) * dt;
        RU3(jc, kc) = QCAP(jc, kc);
    }
    for (kc = 2; kc <= n3m; ++kc)
        for (jc = 1; jc <= n2m; ++jc) Q3F(jc, kc) = Q3(jc, kc) + RHS3(jc, kc);
}

/* ------------------------------------------------------------------------ */
/* Fortran MINLOC(|x(1:n) - v|): first index of the minimum, 1-based. */
static int minloc_abs(const double *x, int n, double v)
{
    int i, best = 1;
    double bv = fabs(x[1] - v);
    for (i = 2; i <= n; ++i) {
        double a = fabs(x[i] - v);
        if (a < bv) { bv = a; best = i; }
    }
    return best;
}

/* lgrdistribution.F90:4-82 */
static int lgrdistribution(orc_state *s)
{
    int i, ind, n2 = s->n2, n3 = s->n3, n2m = s->n2m, n3m = s->n3m, nlm = s->nlm;
    double ymax = -1e300, zmax = -1e300, ymin = 1e300, zmin = 1e300;
    if (nlm == 0) return ORC_OK;
    for (i = 1; i <= nlm; ++i) {
        if (LMCC(i, 1) > ymax) ymax = LMCC(i, 1);
        if (LMCC(i, 2) > zmax) zmax = LMCC(i, 2);
        if (LMCC(i, 1) < ymin) ymin = LMCC(i, 1);
        if (LMCC(i, 2) < zmin) zmin = LMCC(i, 2);
    }
    if (ymax > s->p.rext2 || ymin < 0.0 || zmax > s->p.alx3 || zmin < 0.0) return ORC_ERR_MARKER_OUTSIDE;
    for (i = 1; i <= nlm; ++i) {
        double y = LMCC(i, 1), z = LMCC(i, 2);
        LMIND(i, 1, 1) = minloc_abs(s->rc, n2, y);
        if (z < s->zm[1] && z > 0.0) ind = 1;
        else if (z > s->zm[n3m] && z < s->p.alx3) ind = n3m;
        else ind = minloc_abs(s->zm, n3, z);
        LMIND(i, 1, 2) = ind;
        if (y < s->rm[1] && y > 0.0) ind = 1;
        else if (y > s->rm[n2m] && y < s->p.rext2) ind = n2m;
        else ind = minloc_abs(s->rm, n2, y);
        LMIND(i, 2, 1) = ind;
        LMIND(i, 2, 2) = minloc_abs(s->zz, n3, z);
    }
    return ORC_OK;
}

/* MLSint.F90:4-104 (MLS_SD_ind) */
static void mls_sd_ind(orc_state *s, int idir, const int indse[2], double posk[2][DIMSD],
                       double dsk[2][DIMSD], int eul[2][DIMSD])
{
    int j, k, jj, kk, m, ind, n2m = s->n2m, n3m = s->n3m;
    const double *y, *z;
    int eulin[3][3][2];
    if (idir == 1) { y = s->rc; z = s->zm; }
    else { y = s->rm; z = s->zz; }
    for (j = 1; j <= 3; ++j)
        for (k = 1; k <= 3; ++k) {
            if (indse[0] == 1) jj = j + 1;
            else if (indse[0] == n2m) jj = j - 1;
            else jj = j;
            if (indse[1] == 1) kk = k + 1;
            else if (indse[1] == n3m) kk = k - 1;
            else kk = k;
            eulin[j - 1][k - 1][0] = indse[0] + jj - 2;
            eulin[j - 1][k - 1][1] = indse[1] + kk - 2;
        }
    m = 0;
    for (j = 1; j <= 3; ++j)
        for (k = 1; k <= 3; ++k) {
            ind = eulin[j - 1][k - 1][0];
            posk[0][m] = y[ind];
            dsk[0][m] = y[ind + 1] - y[ind];
            ind = eulin[j - 1][k - 1][1];
            posk[1][m] = z[ind];
            dsk[1][m] = z[ind + 1] - z[ind];
            eul[0][m] = eulin[j - 1][k - 1][0];
            eul[1][m] = eulin[j - 1][k - 1][1];
            ++m;
        }
}

/* One factor of weight_function (MLSint.F90:431-479) at the scaled distance
 * rcv = |dif| / (radiusSD * SDd): cubic spline (1), exponential with
 * wideness 0.3 (2), quartic spline (3), Roma's 3-point function (4). */
static double weight1d(double rcv, int wfun)
{
    const double wcon = 0.3;
    double rc2 = rcv * rcv, rc3 = rcv * rcv * rcv, rc4 = rcv * rcv * rcv * rcv;
    if (wfun == 1) {
        if (rcv > 1.0) return 0.0;
        if (rcv <= 0.5) return (2.0 / 3.0) - 4.0 * rc2 + 4.0 * rc3;
        return (4.0 / 3.0) - 4.0 * rcv + 4.0 * rc2 - (4.0 / 3.0) * rc3;
    } else if (wfun == 2) {
        double t = rcv / wcon;
        return (rcv <= 1.0) ? exp(-(t * t)) : 0.0;
    } else if (wfun == 3) {
        return (rcv <= 1.0) ? 1.0 - 6.0 * rc2 + 8.0 * rc3 - 3.0 * rc4 : 0.0;
    } else if (wfun == 4) {
        if (rcv > 1.0) return 0.0;
        if (rcv < 0.5) return 1.0 / 3.0 * (1.0 + sqrt(-3.0 * rc2 + 1.0));
        double om = 1.0 - rcv;
        return 1.0 / 6.0 * (5.0 - 3.0 * rcv - sqrt(-3.0 * (om * om) + 1.0));
    }
    return 0.0;
}

double orc_test_weight1d(double rcv, int wfun) { return weight1d(rcv, wfun); }

/* MLSint.F90:406-486 (weight_function) */
static void weight_function(const double dif[2][DIMSD], const double sdd[2][DIMSD], double *w,
                            int wfun, double radiusSD)
{
    for (int i = 0; i < DIMSD; ++i) {
        double wc[2];
        for (int k = 0; k < 2; ++k) {
            double sdds = radiusSD * sdd[k][i];
            wc[k] = weight1d(fabs(dif[k][i]) / sdds, wfun);
        }
        w[i] = wc[0] * wc[1];
    }
}

/* utils_math.f90:637-725 (GEsolve, scaled partial pivoting); returns info */
static int gesolve3(const double ai[3][3], const double bi[3], double x[3])
{
    const int n = 3;
    double a[3][3], b[3], sc[3], c, pivot, store;
    int i, j, k, l;
    for (i = 0; i < 3; ++i) {
        b[i] = bi[i];
        for (j = 0; j < 3; ++j) a[i][j] = ai[i][j];
    }
    for (k = 0; k < n - 1; ++k) {
        for (i = k; i < n; ++i) {
            sc[i] = 0.0;
            for (j = k; j < n; ++j) sc[i] = fmax(sc[i], fabs(a[i][j]));
        }
        pivot = fabs(a[k][k] / sc[k]);
        l = k;
        for (j = k + 1; j < n; ++j)
            if (fabs(a[j][k] / sc[j]) > pivot) {
                pivot = fabs(a[j][k] / sc[j]);
                l = j;
            }
        if (pivot == 0.0) return 0;
        if (l != k) {
            for (j = k; j < n; ++j) {
                store = a[k][j]; a[k][j] = a[l][j]; a[l][j] = store;
            }
            store = b[k]; b[k] = b[l]; b[l] = store;
        }
        for (i = k + 1; i < n; ++i) {
            c = a[i][k] / a[k][k];
            a[i][k] = 0.0;
            b[i] = b[i] - c * b[k];
            for (j = k + 1; j < n; ++j) a[i][j] = a[i][j] - c * a[k][j];
        }
    }
    x[n - 1] = b[n - 1] / a[n - 1][n - 1];
    for (i = n - 2; i >= 0; --i) {
        c = 0.0;
        for (j = i + 1; j < n; ++j) c = c + a[i][j] * x[j];
        x[i] = (b[i] - c) / a[i][i];
    }
    return 1;
}

/* MLSint.F90:115-166 (MLS_SF) with mls_basis_func 264-280 and compute_AB
 * 284-338; returns ORC_OK or ORC_ERR_PARTITION_UNITY. */
static int mls_sf(orc_state *s, const double pos[2], const double sdc[2][DIMSD],
                  const double sdd[2][DIMSD], double *philoc)
{
    double dif[2][DIMSD], w[DIMSD], pk[3][DIMSD], B[3][DIMSD], A[3][3], gps[3];
    /* GEsolve leaves x unset on a singular matrix; zeros are used here */
    double ptA[3] = {0.0, 0.0, 0.0};
    int i, j, k, info;
    for (i = 0; i < DIMSD; ++i) {
        pk[0][i] = 1.0;
        pk[1][i] = sdc[0][i];
        pk[2][i] = sdc[1][i];
        dif[0][i] = pos[0] - sdc[0][i];
        dif[1][i] = pos[1] - sdc[1][i];
    }
    weight_function(dif, sdd, w, s->p.wfun, s->p.radiusSD);
    for (i = 0; i < 3; ++i)
        for (j = 0; j < DIMSD; ++j) B[i][j] = pk[i][j] * w[j];
    for (j = 0; j < 3; ++j)
        for (k = 0; k < 3; ++k) A[j][k] = 0.0;
    for (i = 0; i < DIMSD; ++i)
        for (j = 0; j < 3; ++j)
            for (k = 0; k < 3; ++k) A[j][k] = A[j][k] + w[i] * (pk[j][i] * pk[k][i]);
    gps[0] = 1.0; gps[1] = pos[0]; gps[2] = pos[1];
    info = gesolve3((const double(*)[3])A, gps, ptA);
    (void)info; /* the reference only reports a singular matrix (MLSint.F90:153-155) */
    double sumphi = 0.0;
    for (j = 0; j < DIMSD; ++j) {
        philoc[j] = ptA[0] * B[0][j] + ptA[1] * B[1][j] + ptA[2] * B[2][j];
    }
    for (j = 0; j < DIMSD; ++j) sumphi = sumphi + philoc[j];
    if (fabs(1.0 - fabs(sumphi)) > 1.0e-10) return ORC_ERR_PARTITION_UNITY;
    return ORC_OK;
}

/* forcing_MLS.F90:135-182 (interp_spread) */
static void interp_spread(orc_state *s, int n)
{
    int m;
    double ucy = 0.0, ucz = 0.0, fcy, fcz;
    for (m = 1; m <= DIMSD; ++m) ucy = ucy + PHI(n, 1, m) * Q2F(EULINS(n, 1, 1, m), EULINS(n, 1, 2, m));
    fcy = (LMVEL(n, 1) - ucy) / s->dt;
    for (m = 1; m <= DIMSD; ++m) {
        int j = EULINS(n, 1, 1, m), k = EULINS(n, 1, 2, m);
        FEY(j, k) = FEY(j, k) + CL(n, 1) * PHI(n, 1, m) * fcy;
    }
    for (m = 1; m <= DIMSD; ++m) ucz = ucz + PHI(n, 2, m) * Q3F(EULINS(n, 2, 1, m), EULINS(n, 2, 2, m));
    fcz = (LMVEL(n, 2) - ucz) / s->dt;
    for (m = 1; m <= DIMSD; ++m) {
        int j = EULINS(n, 2, 1, m), k = EULINS(n, 2, 2, m);
        FEZ(j, k) = FEZ(j, k) + CL(n, 2) * PHI(n, 2, m) * fcz;
    }
}

/* forcing_MLS.F90:5-128, one rank (ghost adds of mpi_routines.F90:295-387) */
static int forcing_mls(orc_state *s, int substep)
{
    int fs, n, idir, i, m, rc = ORC_OK;
    size_t fsz = (size_t)s->n2 * (s->n3m + 5);
    for (fs = 1; fs <= s->p.forst; ++fs) {
        memset(s->fey, 0, fsz * sizeof(double));
        memset(s->fez, 0, fsz * sizeof(double));
        if (substep == 1 && fs == 1) {
            memset(s->phipred, 0, sizeof(double) * 2 * DIMSD * s->nlm);
            memset(s->cl, 0, sizeof(double) * 2 * s->nlm);
            memset(s->phi, 0, sizeof(double) * 2 * DIMSD * s->nlm);
        }
        for (n = 1; n <= s->nlm; ++n) {
            double pos[2] = {LMCC(n, 1), LMCC(n, 2)};
            double areal = s->lmA[n - 1];
            if (!(s->zz[1] < pos[1] && pos[1] < s->zz[s->n3])) continue;
            if (substep == 1 && fs == 1) {
                int eulin[2][2][DIMSD];
                for (idir = 1; idir <= 2; ++idir) {
                    int indse[2] = {LMIND(n, idir, 1), LMIND(n, idir, 2)};
                    double sdc[2][DIMSD], sdd[2][DIMSD], philoc[DIMSD];
                    mls_sd_ind(s, idir, indse, sdc, sdd, eulin[idir - 1]);
                    int e = mls_sf(s, pos, (const double(*)[DIMSD])sdc, (const double(*)[DIMSD])sdd, philoc);
                    if (e != ORC_OK && rc == ORC_OK) rc = e;
                    for (m = 0; m < DIMSD; ++m) PHI(n, idir, m + 1) = philoc[m];
                    double dvel = 0.0, lgr_h = 0.0;
                    for (i = 0; i < DIMSD; ++i) {
                        double vol = sdd[0][i] * sdd[1][i];
                        dvel = dvel + philoc[i] * vol;
                        lgr_h = lgr_h + 1.0 / 2.0 * philoc[i] * (sdd[0][i] + sdd[1][i]);
                    }
                    CL(n, idir) = areal * lgr_h / dvel;
                }
                for (idir = 1; idir <= 2; ++idir)
                    for (i = 1; i <= 2; ++i)
                        for (m = 1; m <= DIMSD; ++m) EULINS(n, idir, i, m) = eulin[idir - 1][i - 1][m - 1];
            }
            interp_spread(s, n);
        }
        if (substep == 1 && fs == 1)
            memcpy(s->phipred, s->phi, sizeof(double) * 2 * DIMSD * s->nlm);
        /* update_add_upper_ghost on the top (only) rank: plane kend is
         * overwritten with the never-received buffer, i.e. zeroed. */
        for (int j = 1; j <= s->n2; ++j) {
            FEY(j, s->n3m) = 0.0;
            FEZ(j, s->n3m) = 0.0;
        }
    }
    return rc;
}

/* ------------------------------------------------------------------------ */
/* solve_fluid.F90:53-103 (solq2j) */
static void solq2j(orc_state *s, double *dl, double *d, double *du, double *du2, int *ipiv, double *f,
                   double *am, double *ap)
{
    int jc, kc, n2 = s->n2, n2m = s->n2m, n3m = s->n3m;
    double betadx = s->beta * s->al, b;
    for (kc = 1; kc <= n3m; ++kc) {
        ap[1] = 0.0; am[1] = 0.0; ap[n2] = 0.0; am[n2] = 0.0;
        f[1] = 0.0;
        for (jc = 2; jc <= n2m; ++jc) {
            b = 1.0 / (1.0 - betadx * s->ac2j[jc]);
            ap[jc] = -betadx * s->ap2j[jc] * b;
            am[jc] = -betadx * s->am2j[jc] * b;
            f[jc] = RHS(jc, kc) * b;
        }
        f[n2] = 0.0;
        for (jc = 1; jc <= n2 - 1; ++jc) { dl[jc] = am[jc + 1]; du[jc] = ap[jc]; }
        for (jc = 1; jc <= n2; ++jc) d[jc] = 1.0;
        dgttrf(n2, dl, d, du, du2, ipiv);
        dgttrs(n2, dl, d, du, du2, ipiv, f);
        for (jc = 1; jc <= n2; ++jc) RHS(jc, kc) = f[jc];
    }
}

/* solve_fluid.F90:4-47 (solq3j) */
static void solq3j(orc_state *s, double *dl, double *d, double *du, double *du2, int *ipiv, double *f,
                   double *am, double *ap)
{
    int jc, kc, n2m = s->n2m, n3m = s->n3m;
    double betadx = s->beta * s->al, b;
    for (kc = 1; kc <= n3m; ++kc) {
        for (jc = 1; jc <= n2m; ++jc) {
            b = 1.0 / (1.0 - betadx * s->ac3j[jc]);
            ap[jc] = -betadx * s->ap3j[jc] * b;
            am[jc] = -betadx * s->am3j[jc] * b;
            f[jc] = RHS(jc, kc) * b;
        }
        for (jc = 1; jc <= n2m - 1; ++jc) { dl[jc] = am[jc + 1]; du[jc] = ap[jc]; }
        for (jc = 1; jc <= n2m; ++jc) d[jc] = 1.0;
        dgttrf(n2m, dl, d, du, du2, ipiv);
        dgttrs(n2m, dl, d, du, du2, ipiv, f);
        for (jc = 1; jc <= n2m; ++jc) RHS(jc, kc) = f[jc];
    }
}

/* solve_fluid.F90:107-172 (solq2k); the transposes of matrix_trans.F90 are
 * plain index swaps on one rank */
static void solq2k(orc_state *s, double *dl, double *d, double *du, double *du2, int *ipiv, double *f,
                   double *am, double *ap)
{
    int jc, kc, n2m = s->n2m, n3m = s->n3m;
    double betadx = s->beta * s->al, b, ugkks, ugkkn;
    ugkks = 1.0 / (s->g3rm[1] * s->g3rc[1]) * s->dx3q * 2.0 * betadx;
    ugkkn = 1.0 / (s->g3rm[n3m] * s->g3rc[s->n3]) * s->dx3q * 2.0 * betadx;
    for (jc = 1; jc <= n2m; ++jc) {
        for (kc = 1; kc <= n3m; ++kc) {
            b = 1.0 / (1.0 - s->ac3sk[kc] * betadx);
            am[kc] = -s->am3sk[kc] * betadx * b;
            ap[kc] = -s->ap3sk[kc] * betadx * b;
            f[kc] = RHS(jc, kc) * b;
        }
        for (kc = 1; kc <= n3m - 1; ++kc) { dl[kc] = am[kc + 1]; du[kc] = ap[kc]; }
        for (kc = 1; kc <= n3m; ++kc) d[kc] = 1.0;
        dgttrf(n3m, dl, d, du, du2, ipiv);
        f[1] = f[1] + s->dqb2s[jc] * ugkks;
        f[n3m] = f[n3m] + s->dqb2n[jc] * ugkkn;
        dgttrs(n3m, dl, d, du, du2, ipiv, f);
        for (kc = 1; kc <= n3m; ++kc) RHS(jc, kc) = f[kc];
    }
    for (kc = 1; kc <= n3m; ++kc)
        for (jc = 1; jc <= n2m; ++jc) Q2(jc, kc) = Q2(jc, kc) + RHS(jc, kc);
}

/* solve_fluid.F90:176-255 (solq3k) */
static void solq3k(orc_state *s, double *dl, double *d, double *du, double *du2, int *ipiv, double *f,
                   double *am, double *ap)
{
    int jc, kc, n2m = s->n2m, n3 = s->n3, n3m = s->n3m;
    double betadx = s->beta * s->al, b;
    for (jc = 1; jc <= n2m; ++jc) {
        am[1] = 0.0; ap[1] = 0.0; am[n3] = 0.0; ap[n3] = 0.0;
        f[1] = s->dqb3s[jc];
        for (kc = 2; kc <= n3m; ++kc) {
            b = 1.0 / (1.0 - s->ac3sk[kc] * betadx);
            am[kc] = -s->am3sk[kc] * betadx * b;
            ap[kc] = -s->ap3sk[kc] * betadx * b;
            f[kc] = RHS(jc, kc) * b;
        }
        f[n3] = s->dqb3n[jc];
        for (kc = 1; kc <= n3 - 1; ++kc) { dl[kc] = am[kc + 1]; du[kc] = ap[kc]; }
        for (kc = 1; kc <= n3; ++kc) d[kc] = 1.0;
        dgttrf(n3, dl, d, du, du2, ipiv);
        dgttrs(n3, dl, d, du, du2, ipiv, f);
        for (kc = 1; kc <= n3m; ++kc) RHS(jc, kc) = f[kc];
    }
    for (kc = 1; kc <= n3m; ++kc)
        for (jc = 1; jc <= n2m; ++jc) Q3(jc, kc) = Q3(jc, kc) + RHS(jc, kc);
    for (jc = 1; jc <= n2m; ++jc) Q3(jc, n3) = Q3(jc, n3) + s->dqb3n[jc];
}

/* helsolve.F90:4-56 */
static void helsolve(orc_state *s)
{
    int jc, kc, n2 = s->n2, n2m = s->n2m, n3m = s->n3m, nmax = (s->n2 > s->n3 ? s->n2 : s->n3) + 4;
    double *dl = dalloc(nmax), *d = dalloc(nmax), *du = dalloc(nmax), *du2 = dalloc(nmax);
    double *f = dalloc(nmax), *am = dalloc(nmax), *ap = dalloc(nmax);
    int *ipiv = ialloc(nmax);
    for (kc = 1; kc <= n3m; ++kc)
        for (jc = 2; jc <= n2m; ++jc) RHS(jc, kc) = RHS2(jc, kc) + FEY(jc, kc) * s->dt;
    solq2j(s, dl, d, du, du2, ipiv, f, am, ap);
    solq2k(s, dl, d, du, du2, ipiv, f, am, ap);
    for (kc = -2; kc <= n3m + 3; ++kc) {
        Q2(n2, kc) = 0.0;
        Q2(1, kc) = 0.0;
    }
    for (kc = 2; kc <= n3m; ++kc)
        for (jc = 1; jc <= n2m; ++jc) RHS(jc, kc) = RHS3(jc, kc) + FEZ(jc, kc) * s->dt;
    solq3j(s, dl, d, du, du2, ipiv, f, am, ap);
    solq3k(s, dl, d, du, du2, ipiv, f, am, ap);
    free(dl); free(d); free(du); free(du2); free(f); free(am); free(ap); free(ipiv);
}

/* divergence.F90:15-42 (divg) */
static void divg(orc_state *s)
{
    int jc, kc, n2m = s->n2m, n3m = s->n3m;
    double usdtal = 1.0 / (s->dt * s->al), dqcap;
    for (kc = 1; kc <= n3m; ++kc)
        for (jc = 1; jc <= n2m; ++jc) {
            dqcap = (Q2(jc + 1, kc) - Q2(jc, kc)) * s->udx2m[jc] + (Q3(jc, kc + 1) - Q3(jc, kc)) * s->udx3m[kc];
            DPH(jc, kc) = dqcap * usdtal;
        }
}

/* phcalc_outflow.F90:4-71 with reference-BLAS DGEMM loop order */
static void phcalc(orc_state *s)
{
    int i, j, k, m, n2m = s->n2m, n3m = s->n3m;
    double *hmn = s->hmn, *vmn = s->vmn, *hmnt = s->hmnt; /* hmn(n2m,n3m), vmn/hmnt(n3m,n2m) */
    double *dl = dalloc(n2m + 2), *d = dalloc(n2m + 2), *du = dalloc(n2m + 2), *du2 = dalloc(n2m + 2);
    double *f = dalloc(n2m + 2), *ap = dalloc(n2m + 2), *am = dalloc(n2m + 2), *work = dalloc(n2m + 2);
    int *ipiv = ialloc(n2m + 2);
    /* DGEMM('n','t'): hmn = dphR * zmmt^T */
    for (m = 1; m <= n3m; ++m) {
        double *c = hmn + (size_t)(m - 1) * n2m;
        for (i = 0; i < n2m; ++i) c[i] = 0.0;
        for (k = 1; k <= n3m; ++k) {
            double temp = s->zmmt[(size_t)(m - 1) + (size_t)(k - 1) * n3m];
            const double *a = &DPH(1, k);
            for (i = 0; i < n2m; ++i) c[i] = c[i] + temp * a[i];
        }
    }
    /* per-mode tridiagonal along j */
    for (k = 1; k <= n3m; ++k) {
        for (j = 1; j <= n2m; ++j) {
            double b = 1.0 / (s->acphj[j] + s->wr[k]);
            d[j] = 1.0;
            ap[j] = s->apphj[j] * b;
            am[j] = s->amphj[j] * b;
            f[j] = hmn[(size_t)(j - 1) + (size_t)(k - 1) * n2m] * b;
        }
        for (j = 1; j <= n2m - 1; ++j) { dl[j] = am[j + 1]; du[j] = ap[j]; }
        if (k == s->null_mode) {
            tri_null_solve(n2m, dl, d, du, f, work);
        } else {
            dgttrf(n2m, dl, d, du, du2, ipiv);
            dgttrs(n2m, dl, d, du, du2, ipiv, f);
        }
        for (j = 1; j <= n2m; ++j) vmn[(size_t)(k - 1) + (size_t)(j - 1) * n3m] = f[j];
    }
    /* DGEMM('n','n'): hmnt = zmm * vmnR */
    for (j = 1; j <= n2m; ++j) {
        double *c = hmnt + (size_t)(j - 1) * n3m;
        for (i = 0; i < n3m; ++i) c[i] = 0.0;
        for (m = 1; m <= n3m; ++m) {
            double temp = vmn[(size_t)(m - 1) + (size_t)(j - 1) * n3m];
            const double *a = s->zmm + (size_t)(m - 1) * n3m;
            for (i = 0; i < n3m; ++i) c[i] = c[i] + temp * a[i];
        }
    }
    for (k = 1; k <= n3m; ++k)
        for (j = 1; j <= n2m; ++j) DPH(j, k) = hmnt[(size_t)(k - 1) + (size_t)(j - 1) * n3m];
    free(dl); free(d); free(du); free(du2); free(f); free(ap); free(am); free(work); free(ipiv);
}

/* updvp.F90:5-59 (updvp) */
static void updvp(orc_state *s)
{
    int jc, kc, jm, km, n2m = s->n2m, n3m = s->n3m;
    double usukm, usujm, locdph, al = s->al, dt = s->dt;
    for (kc = 1; kc <= n3m; ++kc)
        for (jc = 1; jc <= n2m; ++jc) {
            jm = s->jmv[jc];
            usujm = al * dt * s->udx2c[jc];
            locdph = DPH(jc, kc);
            Q2(jc, kc) = Q2(jc, kc) - (locdph - DPH(jm, kc)) * usujm;
        }
    for (kc = 2; kc <= n3m; ++kc) {
        km = kc - 1;
        usukm = al * dt * s->udx3c[kc];
        for (jc = 1; jc <= n2m; ++jc) Q3(jc, kc) = Q3(jc, kc) - (DPH(jc, kc) - DPH(jc, km)) * usukm;
    }
}

/* updvp.F90:68-105 (prcalc) */
static void prcalc(orc_state *s)
{
    int kp, km, jm, jp, jc, kc, n2m = s->n2m, n3m = s->n3m;
    double be = s->al * s->beta, ammk, acck, appk, ammj, accj, appj;
    for (kc = 1; kc <= n3m; ++kc) {
        kp = s->kpv[kc]; km = s->kmv[kc];
        ammk = s->amphk[kc]; acck = s->acphk[kc]; appk = s->apphk[kc];
        for (jc = 1; jc <= n2m; ++jc) {
            jm = s->jmv[jc]; jp = s->jpv[jc];
            ammj = s->amphj[jc]; accj = s->acphj[jc]; appj = s->apphj[jc];
            PR(jc, kc) = PR(jc, kc) + DPH(jc, kc) -
                         be * ((DPH(jp, kc) * appj + DPH(jc, kc) * accj + DPH(jm, kc) * ammj) +
                               (DPH(jc, kp) * appk + DPH(jc, kc) * acck + DPH(jc, km) * ammk));
        }
    }
}

/* updvp.F90:112-144 (pr_submean) */
static void pr_submean(orc_state *s)
{
    int j, k, n2m = s->n2m, n3m = s->n3m;
    double pravg = 0.0;
    for (k = 1; k <= n3m; ++k)
        for (j = 1; j <= n2m; ++j)
            pravg = pravg + PR(j, k) * (s->zz[k + 1] - s->zz[k]) * (s->rc[j + 1] - s->rc[j]);
    pravg = pravg / (s->p.rext2 * s->p.alx3);
    size_t gsz = (size_t)s->n2 * (s->n3m + 6);
    for (size_t i = 0; i < gsz; ++i) s->pr[i] = s->pr[i] - pravg;
}

/* ------------------------------------------------------------------------ */
/* fluid_tsch.F90:25-57 */
int orc_cfl_dt(orc_state *s, int ns, double *dt, double dt_o, double *cflm)
{
    if (!s || !dt) return ORC_ERR_ARG;
    s->dt = *dt;
    s->dt_o = dt_o;
    cfl(s);
    if (!s->p.idtv) {
        if (ns == 0) {
            s->cflm = s->cflm * s->dt;
        } else if (ns > 0) {
            s->dt = s->p.cflmax / s->cflm;
            if (s->dt > s->p.dtmax) s->dt = s->p.dtmax;
        }
        if (s->dt < 1.0e-8) {
            *dt = s->dt;
            if (cflm) *cflm = s->cflm;
            return ORC_ERR_DT_TOO_SMALL;
        }
    } else {
        s->cflm = s->cflm * s->dt;
        if (s->cflm > s->p.cfllim) {
            if (cflm) *cflm = s->cflm;
            return ORC_ERR_CFL_TOO_LARGE;
        }
    }
    s->beta = s->dt / s->ren * 0.5;
    *dt = s->dt;
    if (cflm) *cflm = s->cflm;
    return ORC_OK;
}

/* fluid_tsch.F90:62-166, substep n (1-based) */
int orc_substep(orc_state *s, int n, int ns)
{
    int rc;
    if (!s || n < 1 || n > s->p.nsst) return ORC_ERR_ARG;
    if (s->p.nsst == 3) {
        s->al = s->alm[n];
        s->ga = s->gam[n];
        s->ro = s->rom[n];
    } else {
        if (ns == 0) {
            s->ga = 1.0;
            s->ro = 0.0;
        } else {
            s->ga = 1.0 + 0.5 * s->dt / s->dt_o;
            s->ro = -0.5 * s->dt / s->dt_o;
        }
        s->al = s->ga + s->ro;
    }
    boucdq(s);
    hdnl2(s);
    hdnl3(s);
    invtr2(s);
    invtr3(s);
    if (s->p.ibact) {
        if (n == 1) {
            rc = lgrdistribution(s);
            if (rc != ORC_OK) return rc;
        }
        rc = forcing_mls(s, n);
        if (rc != ORC_OK) return rc;
    }
    helsolve(s);
    divg(s);
    phcalc(s);
    updvp(s);
    prcalc(s);
    pr_submean(s);
    boucqt(s);
    return ORC_OK;
}

int orc_fluid_tsch(orc_state *s, int ns, double *dt, double dt_o, double *cflm)
{
    int rc = orc_cfl_dt(s, ns, dt, dt_o, cflm);
    if (rc != ORC_OK) return rc;
    for (int n = 1; n <= s->p.nsst; ++n) {
        rc = orc_substep(s, n, ns);
        if (rc != ORC_OK) return rc;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Named array access (reference layout). */
typedef struct {
    const char *name;
    int is_int;
    size_t off; /* offset of the pointer member */
    int kind;   /* size class */
    int base;   /* 1 for 1-based 1-D arrays (copy from index 1) */
} arr_desc;

enum { SZ_G, SZ_F, SZ_R, SZ_N2, SZ_N3, SZ_N3M, SZ_Z, SZ_M2, SZ_M1, SZ_MIND, SZ_MEUL, SZ_MPHI, SZ_MCL };

#define AD(nm, member, isint, kind, base) {nm, isint, offsetof(orc_state, member), kind, base}
#include <stddef.h>
static const arr_desc g_arrays[] = {
    AD("q2", q2, 0, SZ_G, 0), AD("q3", q3, 0, SZ_G, 0), AD("pr", pr, 0, SZ_G, 0),
    AD("dph", dph, 0, SZ_G, 0), AD("q2f", q2f, 0, SZ_G, 0), AD("q3f", q3f, 0, SZ_G, 0),
    AD("fey", fey, 0, SZ_F, 0), AD("fez", fez, 0, SZ_F, 0),
    AD("ru2", ru2, 0, SZ_R, 0), AD("ru3", ru3, 0, SZ_R, 0), AD("rhs2", rhs2, 0, SZ_R, 0),
    AD("rhs3", rhs3, 0, SZ_R, 0), AD("rhs", rhs, 0, SZ_R, 0), AD("qcap", qcap, 0, SZ_R, 0),
    AD("qb2s", qb2s, 0, SZ_N2, 1), AD("qb3s", qb3s, 0, SZ_N2, 1), AD("qb2n", qb2n, 0, SZ_N2, 1),
    AD("qb3n", qb3n, 0, SZ_N2, 1), AD("dqb2s", dqb2s, 0, SZ_N2, 1), AD("dqb3s", dqb3s, 0, SZ_N2, 1),
    AD("dqb2n", dqb2n, 0, SZ_N2, 1), AD("dqb3n", dqb3n, 0, SZ_N2, 1),
    AD("dq2x2o", dq2x2o, 0, SZ_N2, 1), AD("dq3x2o", dq3x2o, 0, SZ_N2, 1),
    AD("rc", rc, 0, SZ_N2, 1), AD("rm", rm, 0, SZ_N2, 1), AD("g2rc", g2rc, 0, SZ_N2, 1),
    AD("g2rm", g2rm, 0, SZ_N2, 1), AD("udx2c", udx2c, 0, SZ_N2, 1), AD("udx2m", udx2m, 0, SZ_N2, 1),
    AD("zz", zz, 0, SZ_N3, 1), AD("zm", zm, 0, SZ_N3, 1), AD("g3rc", g3rc, 0, SZ_N3, 1),
    AD("g3rm", g3rm, 0, SZ_N3, 1), AD("udx3c", udx3c, 0, SZ_N3, 1), AD("udx3m", udx3m, 0, SZ_N3, 1),
    AD("am2j", am2j, 0, SZ_N2, 1), AD("ac2j", ac2j, 0, SZ_N2, 1), AD("ap2j", ap2j, 0, SZ_N2, 1),
    AD("am3j", am3j, 0, SZ_N2, 1), AD("ac3j", ac3j, 0, SZ_N2, 1), AD("ap3j", ap3j, 0, SZ_N2, 1),
    AD("am3ck", am3ck, 0, SZ_N3, 1), AD("ac3ck", ac3ck, 0, SZ_N3, 1), AD("ap3ck", ap3ck, 0, SZ_N3, 1),
    AD("am3sk", am3sk, 0, SZ_N3, 1), AD("ac3sk", ac3sk, 0, SZ_N3, 1), AD("ap3sk", ap3sk, 0, SZ_N3, 1),
    AD("amphj", amphj, 0, SZ_N2, 1), AD("acphj", acphj, 0, SZ_N2, 1), AD("apphj", apphj, 0, SZ_N2, 1),
    AD("amphk", amphk, 0, SZ_N3, 1), AD("acphk", acphk, 0, SZ_N3, 1), AD("apphk", apphk, 0, SZ_N3, 1),
    AD("jmv", jmv, 1, SZ_N2, 1), AD("jpv", jpv, 1, SZ_N2, 1), AD("kmv", kmv, 1, SZ_N3, 1),
    AD("kpv", kpv, 1, SZ_N3, 1),
    AD("lmcc", lmcc, 0, SZ_M2, 0), AD("lmvel", lmvel, 0, SZ_M2, 0), AD("lmA", lmA, 0, SZ_M1, 0),
    AD("lmind", lmind, 1, SZ_MIND, 0), AD("eulins", eulins, 1, SZ_MEUL, 0),
    AD("phi", phi, 0, SZ_MPHI, 0), AD("phipred", phipred, 0, SZ_MPHI, 0), AD("cl", cl, 0, SZ_MCL, 0),
};

static const arr_desc *find_arr(const char *name)
{
    for (size_t i = 0; i < sizeof(g_arrays) / sizeof(g_arrays[0]); ++i)
        if (strcmp(g_arrays[i].name, name) == 0) return &g_arrays[i];
    return NULL;
}

static long arr_len(orc_state *s, const arr_desc *a)
{
    switch (a->kind) {
    case SZ_G: return (long)s->n2 * (s->n3m + 6);
    case SZ_F: return (long)s->n2 * (s->n3m + 5);
    case SZ_R: return (long)s->n2 * s->n3m;
    case SZ_N2: return s->n2;
    case SZ_N3: return s->n3;
    case SZ_N3M: return s->n3m;
    case SZ_M2: return 2L * s->nlm;
    case SZ_M1: return s->nlm;
    case SZ_MIND: return 4L * s->nlm;
    case SZ_MEUL: return 4L * DIMSD * s->nlm;
    case SZ_MPHI: return 2L * DIMSD * s->nlm;
    case SZ_MCL: return 2L * s->nlm;
    }
    return 0;
}

long orc_size(orc_state *s, const char *name)
{
    const arr_desc *a;
    if (!s || !name) return 0;
    if (!strcmp(name, "zmm") || !strcmp(name, "zmmt")) return (long)s->n3m * s->n3m;
    if (!strcmp(name, "wr")) return s->n3m;
    a = find_arr(name);
    return a ? arr_len(s, a) : 0;
}

static void *arr_ptr(orc_state *s, const arr_desc *a) { return *(void **)((char *)s + a->off); }

int orc_get(orc_state *s, const char *name, void *out)
{
    const arr_desc *a;
    if (!s || !name || !out) return ORC_ERR_ARG;
    if (!strcmp(name, "scalars")) {
        double *o = (double *)out;
        o[0] = s->dt; o[1] = s->beta; o[2] = s->al; o[3] = s->ga; o[4] = s->ro; o[5] = s->cflm;
        o[6] = s->q3avg; o[7] = s->qinf; o[8] = s->qout; o[9] = s->dx2; o[10] = s->dx3; o[11] = s->ren;
        return ORC_OK;
    }
    a = find_arr(name);
    if (!a) return ORC_ERR_ARG;
    long n = arr_len(s, a);
    size_t esz = a->is_int ? sizeof(int) : sizeof(double);
    char *src = (char *)arr_ptr(s, a) + (a->base ? esz : 0);
    memcpy(out, src, (size_t)n * esz);
    return ORC_OK;
}

int orc_set(orc_state *s, const char *name, const void *in)
{
    const arr_desc *a;
    if (!s || !name || !in) return ORC_ERR_ARG;
    a = find_arr(name);
    if (!a) return ORC_ERR_ARG;
    long n = arr_len(s, a);
    size_t esz = a->is_int ? sizeof(int) : sizeof(double);
    char *dst = (char *)arr_ptr(s, a) + (a->base ? esz : 0);
    memcpy(dst, in, (size_t)n * esz);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Loads on the body and the forcing check (called after the step by
 * fluid_tsch.F90:167-177 when ibact/cbody are set).                         */

/* MLSint.F90:492-585 (weight_function_d): W and its derivatives w.r.t. the
 * probe position, w[i][0] = W, w[i][1] = dW/dy, w[i][2] = dW/dz. */
static void weight_function_d(const double dif[2][DIMSD], const double sdd[2][DIMSD], double w[DIMSD][3],
                              int wfun, double radiusSD)
{
    const double wcon = 0.3, eps = 1.0e-20;
    for (int i = 0; i < DIMSD; ++i) {
        double wc[2] = {0.0, 0.0}, dwcdc[2] = {0.0, 0.0};
        for (int k = 0; k < 2; ++k) {
            const double difc = dif[k][i], sdds = radiusSD * sdd[k][i];
            const double drdc = (fabs(difc) <= eps) ? 0.0 : (difc / fabs(difc)) / sdds;
            const double rcv = fabs(difc) / sdds;
            const double rc2 = rcv * rcv, rc3 = rcv * rcv * rcv, rc4 = rcv * rcv * rcv * rcv;
            if (wfun == 1) {
                if (rcv > 1.0) { wc[k] = 0.0; dwcdc[k] = 0.0; }
                else if (rcv <= 0.5) {
                    wc[k] = (2.0 / 3.0) - 4.0 * rc2 + 4.0 * rc3;
                    dwcdc[k] = (-8.0 * rcv + 12.0 * rc2) * drdc;
                } else {
                    wc[k] = (4.0 / 3.0) - 4.0 * rcv + 4.0 * rc2 - (4.0 / 3.0) * rc3;
                    dwcdc[k] = (-4.0 + 8.0 * rcv - 4.0 * rc2) * drdc;
                }
            } else if (wfun == 2) {
                if (rcv <= 1.0) {
                    const double t = rcv / wcon;
                    wc[k] = exp(-(t * t));
                    dwcdc[k] = (-1.0 / (wcon * wcon) * 2.0 * rcv * exp(-(t * t))) * drdc;
                }
            } else if (wfun == 3) {
                if (rcv <= 1.0) {
                    wc[k] = 1.0 - 6.0 * rc2 + 8.0 * rc3 - 3.0 * rc4;
                    dwcdc[k] = -12.0 * rcv + 24.0 * rc2 - 12.0 * rc3;   /* no drdc, as the reference */
                }
            } else if (wfun == 4) {
                if (rcv > 1.0) { wc[k] = 0.0; dwcdc[k] = 0.0; }
                else if (rcv < 0.5) {
                    wc[k] = 1.0 / 3.0 * (1.0 + sqrt(-3.0 * rc2 + 1.0));
                    dwcdc[k] = -rcv * pow(-3.0 * rc2 - 1.0, -0.5);    /* NaN, as the reference */
                } else {
                    const double om = 1.0 - rcv;
                    wc[k] = 1.0 / 6.0 * (5.0 - 3.0 * rcv - sqrt(-3.0 * (om * om) + 1.0));
                    dwcdc[k] = -(1.0 - rcv) / (2.0 * sqrt(1.0 - 3.0 * (om * om))) - 0.5;
                }
            }
        }
        w[i][0] = wc[0] * wc[1];
        w[i][1] = wc[1] * dwcdc[0];
        w[i][2] = wc[0] * dwcdc[1];
    }
}

/* MLSint.F90:173-258 (MLS_SF_d) with mls_basis_func (264-280) and
 * compute_AB_d (342-402): phid[0][i] = phi_i, phid[1][i] = dphi_i/dy,
 * phid[2][i] = dphi_i/dz at the probe.  Returns ORC_ERR_PARTITION_UNITY when
 * |1-|sum phi|| > 1e-10 (MLSint.F90:249-255). */
static int mls_sf_d(orc_state *s, const double pos[2], const double sdc[2][DIMSD], const double sdd[2][DIMSD],
                    double phid[3][DIMSD])
{
    double dif[2][DIMSD], w[DIMSD][3], pk[3][DIMSD], B[3][DIMSD][3], A[3][3][3], Aa[3][3];
    double gp[3][3], gam[3][3] = {{0.0}}, C[3], x[3];
    int i, j, k, l;
    for (i = 0; i < DIMSD; ++i) {
        pk[0][i] = 1.0;
        pk[1][i] = sdc[0][i];
        pk[2][i] = sdc[1][i];
        dif[0][i] = pos[0] - sdc[0][i];
        dif[1][i] = pos[1] - sdc[1][i];
    }
    weight_function_d(dif, sdd, w, s->p.wfun, s->p.radiusSD);
    for (i = 0; i < 3; ++i)
        for (j = 0; j < DIMSD; ++j)
            for (k = 0; k < 3; ++k) B[i][j][k] = pk[i][j] * w[j][k];
    for (j = 0; j < 3; ++j)
        for (k = 0; k < 3; ++k)
            for (l = 0; l < 3; ++l) A[j][k][l] = 0.0;
    for (i = 0; i < DIMSD; ++i)
        for (j = 0; j < 3; ++j)
            for (k = 0; k < 3; ++k) {
                const double pp = pk[j][i] * pk[k][i];
                for (l = 0; l < 3; ++l) A[j][k][l] = A[j][k][l] + w[i][l] * pp;
            }
    /* gp: identity with gp(1,2:3) = pos */
    for (i = 0; i < 3; ++i)
        for (j = 0; j < 3; ++j) gp[i][j] = (i == j) ? 1.0 : 0.0;
    gp[0][1] = pos[0];
    gp[0][2] = pos[1];
    for (j = 0; j < 3; ++j)
        for (k = 0; k < 3; ++k) Aa[j][k] = A[j][k][0];
    x[0] = x[1] = x[2] = 0.0;
    gesolve3((const double(*)[3])Aa, gp[0], x);
    for (j = 0; j < 3; ++j) gam[j][0] = x[j];
    for (l = 1; l <= 2; ++l) {
        /* C = gp(l+1,:) - matmul(A(:,:,l+1), gam(:,1)) */
        for (i = 0; i < 3; ++i) {
            double m = 0.0;
            for (k = 0; k < 3; ++k) m = m + A[i][k][l] * gam[k][0];
            C[i] = gp[l][i] - m;
        }
        x[0] = x[1] = x[2] = 0.0;
        gesolve3((const double(*)[3])Aa, C, x);
        for (j = 0; j < 3; ++j) gam[j][l] = x[j];
    }
    double sumphi = 0.0;
    for (i = 0; i < DIMSD; ++i) {
        phid[0][i] = 0.0;
        phid[1][i] = 0.0;
        phid[2][i] = 0.0;
        for (j = 0; j < 3; ++j) {
            phid[0][i] = phid[0][i] + gam[j][0] * B[j][i][0];
            phid[1][i] = phid[1][i] + gam[j][1] * B[j][i][0] + gam[j][0] * B[j][i][1];
            phid[2][i] = phid[2][i] + gam[j][2] * B[j][i][0] + gam[j][0] * B[j][i][2];
        }
    }
    for (i = 0; i < DIMSD; ++i) sumphi = sumphi + phid[0][i];
    if (fabs(1.0 - fabs(sumphi)) > 1.0e-10) return ORC_ERR_PARTITION_UNITY;
    return ORC_OK;
}

/* loads_transfer.F90:168-206 (probeind): first minimum of |x - p| over the
 * 11 nodes around the marker's own node; nodes outside the grid array are
 * never selected (the reference reads past its array there). */
static int window_minloc(const double *x, int n, int c, double v)
{
    int best = 0;
    double bv = 0.0;
    for (int i = c - 5; i <= c + 5; ++i) {
        if (i < 1 || i > n) continue;
        const double a = fabs(x[i] - v);
        if (best == 0 || a < bv) { bv = a; best = i; }
    }
    return best;
}

static void probeind(orc_state *s, const int markind[2], const double probe[2], int ind[2], int idir)
{
    const double *y = (idir == 1) ? s->rc : s->rm;
    const double *z = (idir == 2) ? s->zz : s->zm;
    ind[0] = window_minloc(y, s->n2, markind[0], probe[0]);
    ind[1] = window_minloc(z, s->n3, markind[1], probe[1]);
}

/* MLS_SD_ind (MLSint.F90:4-104) for all three grids: idir 3 is the pressure
 * grid (rm, zm). */
static void mls_sd_ind3(orc_state *s, int idir, const int indse[2], double posk[2][DIMSD], double dsk[2][DIMSD],
                        int eul[2][DIMSD])
{
    if (idir != 3) {
        mls_sd_ind(s, idir, indse, posk, dsk, eul);
        return;
    }
    int j, k, jj, kk, m = 0;
    for (j = 1; j <= 3; ++j)
        for (k = 1; k <= 3; ++k) {
            jj = (indse[0] == 1) ? j + 1 : (indse[0] == s->n2m ? j - 1 : j);
            kk = (indse[1] == 1) ? k + 1 : (indse[1] == s->n3m ? k - 1 : k);
            const int ej = indse[0] + jj - 2, ek = indse[1] + kk - 2;
            posk[0][m] = s->rm[ej];
            dsk[0][m] = s->rm[ej + 1] - s->rm[ej];
            posk[1][m] = s->zm[ek];
            dsk[1][m] = s->zm[ek + 1] - s->zm[ek];
            eul[0][m] = ej;
            eul[1][m] = ek;
            ++m;
        }
}

/* loads_transfer.F90:210-267 (getpbup): value and first/second derivative
 * combinations of q2, q3, pr at the probes.  The backward differences at
 * j = 1 read f(0,k), which in the reference's column-major storage is
 * f(n2,k-1); GX reproduces that address. */
static void getpbup(orc_state *s, const int eul[3][2][DIMSD], const double phi[3][3][DIMSD], double ucy[7],
                    double ucz[7], double pc[7])
{
    const double *fld[3] = {s->q2, s->q3, s->pr};
    double *out[3] = {ucy, ucz, pc};
    for (int d = 0; d < 3; ++d) {
        double *u = out[d];
        for (int i = 0; i < 7; ++i) u[i] = 0.0;
        for (int n = 0; n < DIMSD; ++n) {
            const int cy = eul[d][0][n], cz = eul[d][1][n];
            const double *f = fld[d];
            double uk[3];
            uk[0] = f[GX(s, cy, cz)];
            uk[1] = (f[GX(s, cy, cz)] - f[GX(s, cy - 1, cz)]) * s->dx2 / s->g2rc[cy];
            uk[2] = (f[GX(s, cy, cz)] - f[GX(s, cy, cz - 1)]) * s->dx3 / s->g3rc[cz];
            for (int i = 0; i < 3; ++i) u[i] = u[i] + phi[d][i][n] * uk[0];
            for (int i = 0; i < 2; ++i) u[i + 3] = u[i + 3] + phi[d][1][n] * uk[1 + i];
            for (int i = 0; i < 2; ++i) u[i + 5] = u[i + 5] + phi[d][2][n] * uk[1 + i];
        }
    }
}

/* loads_transfer.F90:4-162 (loads_transfer_c) on one rank.  lmnor, lmacc are
 * (nlm,2) as rigid() leaves them; press (nlm) and tau (nlm,2) out.  hpr is
 * ddmin*kprobe with ddmin already scaled by kprobe in cordin.F90:217-233.
 * lmtan keeps its value from the previous marker when the probe velocity
 * has no tangential part (loads_transfer.F90:94-96). */
int orc_loads_transfer(orc_state *s, double kprobe, const double *lmnor, const double *lmacc, double *press,
                       double *tau)
{
    int n, nlm = s->nlm, rc = ORC_OK;
    double dymin = 1.0e+6, dzmin = 1.0e+6, lmtan[2] = {0.0, 0.0};
    for (int i = 2; i <= s->n2; ++i) dymin = fmin(dymin, s->rc[i] - s->rc[i - 1]);
    for (int i = 2; i <= s->n3; ++i) dzmin = fmin(dzmin, s->zz[i] - s->zz[i - 1]);
    const double ddmin = fmin(dzmin, dymin) * kprobe;
    const double hpr = ddmin * kprobe, visc = s->p.visc;
#define NOR(n, c) lmnor[(size_t)((n) - 1) + (size_t)((c) - 1) * nlm]
#define ACC(n, c) lmacc[(size_t)((n) - 1) + (size_t)((c) - 1) * nlm]
    for (n = 1; n <= nlm; ++n) {
        press[n - 1] = 0.0;
        tau[n - 1] = 0.0;
        tau[n - 1 + nlm] = 0.0;
    }
    for (n = 1; n <= nlm; ++n) {
        const double pos[2] = {LMCC(n, 1), LMCC(n, 2)};
        const double probe[2] = {pos[0] + NOR(n, 1) * hpr, pos[1] + NOR(n, 2) * hpr};
        if (!(s->zz[1] < pos[1] && pos[1] < s->zz[s->n3m + 1])) continue;
        int eul[3][2][DIMSD];
        double phi[3][3][DIMSD];
        for (int idir = 1; idir <= 3; ++idir) {
            int tmp[2], indse[2];
            double sdc[2][DIMSD], sdd[2][DIMSD];
            if (idir == 3) { tmp[0] = LMIND(n, 2, 1); tmp[1] = LMIND(n, 1, 2); }
            else { tmp[0] = LMIND(n, idir, 1); tmp[1] = LMIND(n, idir, 2); }
            probeind(s, tmp, probe, indse, idir);
            mls_sd_ind3(s, idir, indse, sdc, sdd, eul[idir - 1]);
            const int e = mls_sf_d(s, probe, (const double(*)[DIMSD])sdc, (const double(*)[DIMSD])sdd,
                                   phi[idir - 1]);
            if (e && rc == ORC_OK) rc = e;
        }
        double ucy[7], ucz[7], pc[7];
        getpbup(s, (const int(*)[2][DIMSD])eul, (const double(*)[3][DIMSD])phi, ucy, ucz, pc);
        const double vel[2] = {ucy[0], ucz[0]};
        const double veln_md = vel[0] * NOR(n, 1) + vel[1] * NOR(n, 2);
        const double velt[2] = {vel[0] - veln_md * NOR(n, 1), vel[1] - veln_md * NOR(n, 2)};
        const double velt_md = sqrt(velt[0] * velt[0] + velt[1] * velt[1]);
        if (velt_md != 0.0) {
            const double nrm = sqrt(velt[0] * velt[0] + velt[1] * velt[1]);
            lmtan[0] = velt[0] / nrm;
            lmtan[1] = velt[1] / nrm;
        }
        const double nory = NOR(n, 1), norz = NOR(n, 2), tany = lmtan[0], tanz = lmtan[1];
        double dpdnw = ACC(n, 1) * NOR(n, 1) + ACC(n, 2) * NOR(n, 2);
        dpdnw = -1.0 * dpdnw;
        const double dpdne = pc[1] * NOR(n, 1) + pc[2] * NOR(n, 2);
        press[n - 1] = pc[0] - 0.5 * (dpdnw + dpdne) * hpr;
        const double ducdn = (ucy[1] * tany * nory + ucy[2] * tany * norz + ucz[1] * tanz * nory + ucz[2] * tanz * norz);
        const double d3ucdn3 = (pc[3] * nory * nory + pc[4] * nory * norz + pc[5] * norz * nory + pc[6] * norz * norz) *
                               1.0 / visc;
        const double d2ucdn2 = (ucy[3] * tany * nory * nory + ucy[4] * tany * nory * norz +
                                ucy[5] * tany * norz * nory + ucy[6] * tany * norz * norz +
                                ucz[3] * tanz * nory * nory + ucz[4] * tanz * nory * norz +
                                ucz[5] * tanz * norz * nory + ucz[6] * tanz * norz * norz);
        const double ducdnm = ducdn - d2ucdn2 * hpr - 0.5 * d3ucdn3 * (hpr * hpr);
        tau[n - 1] = visc * ducdnm * lmtan[0];
        tau[n - 1 + nlm] = visc * ducdnm * lmtan[1];
    }
#undef NOR
#undef ACC
    return rc;
}

/* ckbcs.F90:5-86 (checkbcs): velocity error at the markers after the step,
 * interpolated with phipred over the forcing supports; nbcv/nbcw are the
 * Euclidean norms divided by nlm. */
int orc_checkbcs(orc_state *s, double *bcerrv, double *bcerrw, double *nbcv, double *nbcw)
{
    int np, n, nlm = s->nlm;
    double sv = 0.0, sw = 0.0;
    for (np = 1; np <= nlm; ++np) {
        bcerrv[np - 1] = 0.0;
        bcerrw[np - 1] = 0.0;
        const double z = LMCC(np, 2);
        if (!(s->zz[1] < z && z < s->zz[s->n3m + 1])) continue;
        double ucy = 0.0, ucz = 0.0;
        for (n = 1; n <= DIMSD; ++n) ucy = ucy + PHIPRED(np, 1, n) * Q2(EULINS(np, 1, 1, n), EULINS(np, 1, 2, n));
        bcerrv[np - 1] = LMVEL(np, 1) - ucy;
        for (n = 1; n <= DIMSD; ++n) ucz = ucz + PHIPRED(np, 2, n) * Q3(EULINS(np, 2, 1, n), EULINS(np, 2, 2, n));
        bcerrw[np - 1] = LMVEL(np, 2) - ucz;
    }
    for (np = 0; np < nlm; ++np) {
        sv = sv + bcerrv[np] * bcerrv[np];
        sw = sw + bcerrw[np] * bcerrw[np];
    }
    *nbcv = sqrt(sv) / (double)nlm;
    *nbcw = sqrt(sw) / (double)nlm;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Diagnostics of ttime (ttime.F90:29-69,136) on the current state:
 * out[0..1] vmaxv (vcheck.F90:3-36), out[2..3] vminv (vcheck.F90:42-73),
 * out[4] divgck (divergence.F90:49-78), out[5] kinen (utils_proc.F90:474-515),
 * out[6] getintvort (utils_proc.F90:582-629), out[7] cflm and out[8] q3avg
 * of cfl (cfl.F90:4-51).                                                    */
int orc_diagnostics(orc_state *s, double *out)
{
    int j, k, n2m = s->n2m, n3m = s->n3m;
    double vmax2 = -100.0, vmax3 = -100.0, vmin2 = 100.0, vmin3 = 100.0, qmax = -100.0, ek = 0.0;
    for (k = 1; k <= n3m; ++k)
        for (j = 1; j <= n2m; ++j) {
            vmax2 = fmax(vmax2, fabs(Q2(j, k)));
            vmax3 = fmax(vmax3, fabs(Q3(j, k)));
            vmin2 = fmin(vmin2, Q2(j, k));
            vmin3 = fmin(vmin3, Q3(j, k));
        }
    for (k = 1; k <= n3m; ++k)
        for (j = 1; j <= n2m; ++j) {
            const double dq = (Q2(j + 1, k) - Q2(j, k)) * s->udx2m[j] + (Q3(j, k + 1) - Q3(j, k)) * s->udx3m[k];
            qmax = fmax(fabs(dq), qmax);
        }
    for (k = 1; k <= n3m; ++k)
        for (j = 1; j <= n2m; ++j) {
            const double vol = (s->zz[k + 1] - s->zz[k]) * (s->rc[j + 1] - s->rc[j]);
            const double vccv = (Q2(j, k) + Q2(j + 1, k)) * 0.5;
            const double vccw = (Q3(j, k) + Q3(j, k + 1)) * 0.5;
            const double vel = vccv * vccv + vccw * vccw;
            ek = ek + vel * vol;
        }
    ek = 0.5 * ek / (s->p.rext2 * s->p.alx3);
    /* vorticity integral along the z = 18.5 line (first node nearest to it) */
    double vi = 0.0;
    {
        const double locz = 18.5;
        int indz = 1;
        double best = fabs(s->zz[1] - locz);
        for (k = 2; k <= s->n3; ++k)
            if (fabs(s->zz[k] - locz) < best) { best = fabs(s->zz[k] - locz); indz = k; }
        if (indz <= n3m) {
            const double dz = s->zz[indz + 1] - s->zz[indz];
            for (j = 1; j <= n2m; ++j) {
                const double dwdy = (Q3(j + 1, indz) - Q3(j, indz)) / (s->rc[j + 1] - s->rc[j]);
                const double dvdz = (Q2(j, indz + 1) - Q2(j, indz)) / dz;
                vi = vi + (dwdy - dvdz) * (s->rc[j + 1] - s->rc[j]);
            }
        }
    }
    const double cflm0 = s->cflm, q3avg0 = s->q3avg;
    cfl(s);
    out[0] = vmax2; out[1] = vmax3; out[2] = vmin2; out[3] = vmin3;
    out[4] = qmax; out[5] = ek; out[6] = vi; out[7] = s->cflm; out[8] = s->q3avg;
    s->cflm = cflm0;
    s->q3avg = q3avg0;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Test hooks: the restated LAPACK / GEsolve / Poisson pieces in isolation.  */
int orc_test_dgtsv(int n, const double *dl0, const double *d0, const double *du0, double *b0)
{
    double *dl = dalloc(n + 2), *d = dalloc(n + 2), *du = dalloc(n + 2), *du2 = dalloc(n + 2), *b = dalloc(n + 2);
    int *ipiv = ialloc(n + 2);
    for (int i = 1; i <= n; ++i) { d[i] = d0[i - 1]; b[i] = b0[i - 1]; }
    for (int i = 1; i <= n - 1; ++i) { dl[i] = dl0[i - 1]; du[i] = du0[i - 1]; }
    dgttrf(n, dl, d, du, du2, ipiv);
    dgttrs(n, dl, d, du, du2, ipiv, b);
    for (int i = 1; i <= n; ++i) b0[i - 1] = b[i];
    free(dl); free(d); free(du); free(du2); free(b); free(ipiv);
    return ORC_OK;
}

int orc_test_gesolve(const double *a_colmajor, const double *b, double *x)
{
    double a[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = a_colmajor[i + 3 * j];
    return gesolve3((const double(*)[3])a, b, x);
}

/* phcalc_outflow on a given right-hand side dph(1:n2m,1:n3m) (column-major) */
int orc_test_phcalc(orc_state *s, double *field)
{
    for (int k = 1; k <= s->n3m; ++k)
        for (int j = 1; j <= s->n2m; ++j) DPH(j, k) = field[(size_t)(j - 1) + (size_t)(k - 1) * s->n2m];
    phcalc(s);
    for (int k = 1; k <= s->n3m; ++k)
        for (int j = 1; j <= s->n2m; ++j) field[(size_t)(j - 1) + (size_t)(k - 1) * s->n2m] = DPH(j, k);
    return ORC_OK;
}

/* The step scalars {dt, beta, al, ga, ro} a stage reads (fluid_tsch.F90:57,
 * 62-77), set directly for the stage-level pins. */
int orc_set_scalars(orc_state *s, const double *v)
{
    if (!s || !v) return ORC_ERR_ARG;
    s->dt = v[0]; s->beta = v[1]; s->al = v[2]; s->ga = v[3]; s->ro = v[4];
    return ORC_OK;
}

/* One stage of the substep on the current state, by name: "cfl", "boucdq",
 * "boucqt", "explicit" (hdnl2, hdnl3, invtr2, invtr3), "lgrdistribution",
 * "forcing1" (lgrdistribution + forcing_MLS of substep 1), "helsolve", "divg", "updvp",
 * "prcalc", "pr_submean". */
int orc_test_stage(orc_state *s, const char *name)
{
    if (!s || !name) return ORC_ERR_ARG;
    if (!strcmp(name, "cfl")) cfl(s);
    else if (!strcmp(name, "boucdq")) boucdq(s);
    else if (!strcmp(name, "boucqt")) boucqt(s);
    else if (!strcmp(name, "explicit")) { hdnl2(s); hdnl3(s); invtr2(s); invtr3(s); }
    else if (!strcmp(name, "lgrdistribution")) return lgrdistribution(s);
    else if (!strcmp(name, "forcing1")) {
        int rc = lgrdistribution(s);
        if (rc != ORC_OK) return rc;
        return forcing_mls(s, 1);
    }
    else if (!strcmp(name, "helsolve")) helsolve(s);
    else if (!strcmp(name, "divg")) divg(s);
    else if (!strcmp(name, "updvp")) updvp(s);
    else if (!strcmp(name, "prcalc")) prcalc(s);
    else if (!strcmp(name, "pr_submean")) pr_submean(s);
    else return ORC_ERR_ARG;
    return ORC_OK;
}
