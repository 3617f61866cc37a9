/* CPU restatement of the DMCBS walker-propagation path (test infrastructure only).
 *
 * This file is the oracle.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, and only as the checker or the CPU baseline.  It is pinned by
 * the golden fixtures that tests/golden/make_golden.py writes from the reference itself
 * (tests/test_oracle_golden.py), and by closed forms and invariants where the reference has
 * no counterpart (the single-particle-move sweep: re-inversion, full-recompute ratios and
 * finite differences, tests/test_oracle_invariants.py).
 *
 * Every function cites the reference routine (file:line) it restates.  Arithmetic is FP64;
 * the LU factorisation is a plain partial-pivoting Doolittle factorisation (LAPACK dgetf2
 * algorithm) followed by forward/back substitution for the inverse, so it agrees with the
 * reference's dgetrf/dgetri to rounding.
 */
#include "dmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXK 512

/* Planted mistakes for the tests that show the single-particle checks catch them (built
 * only by those tests, into a separate library): 1 = sign of the down-spin determinant
 * gradient, 2 = factor of the Green's-function term, 3 = a dropped factor 2 on the Jastrow
 * part of the acceptance.  0 (the default) is the oracle. */
#ifndef DMCO_PLANT
#define DMCO_PLANT 0
#endif

struct dmco_ctx {
    dmco_params p;
    int npart, n;               /* n = particles per spin */
    double el, eli, eunit, etrial, nfac, hbar_t;  /* hbar_t = hbar^2/m for the tables */
    double sigma, sig2, range, scale;
    int ntab, ntab0;
    double *vt, *ut, *dut, *d2ut;   /* itab = 1 (opposite spin) tables [0..ntab] */
    double *vt2, *ut2, *dut2, *d2ut2; /* itab = 2 (equal spin) tables */
    int ieq;
    /* beta */
    double b, d, c, c0, c1, c2, c3, c4, small, rangeb, scaleb;
    double *bt, *dbt, *d2bt;
    int nobeta;
    /* k vectors */
    int nk;
    int32_t numsh[DMCO_MAXORB];
    double ak[MAXK][3], ak2[MAXK];
    double vorb[DMCO_MAXORB];
    double pi;
};

/* ---------------------------------------------------------------- random numbers */
/* ran.f90:8-27 ran1, 48-bit LCG; the 24-bit split of the reference equals the product
 * modulo 2^48, which unsigned 64-bit arithmetic gives directly. */
#define MASK48 ((1ULL << 48) - 1ULL)
double dmco_ran1(int64_t *s) {
    uint64_t x = (uint64_t)*s;
    x = (x * 44485709377909ULL + 96309754297ULL) & MASK48;
    *s = (int64_t)x;
    return (double)(x | 1ULL) * 0x1p-48;
}
/* ran.f90:29-48 ran2 */
double dmco_ran2(int64_t *s) {
    uint64_t x = (uint64_t)*s;
    x = (x * 34522712143931ULL + 55789347517ULL) & MASK48;
    *s = (int64_t)x;
    return (double)(x | 1ULL) * 0x1p-48;
}
/* ran.f90:74-104 gaussian(n): Box-Muller on pairs; an odd last value uses cos */
void dmco_gaussian(int64_t *s, int32_t n, double *g) {
    const double pi = 4.0 * atan(1.0);
    int m = 2 * ((n + 1) / 2);
    double rn[m];
    for (int i = 0; i < m; i++) rn[i] = dmco_ran1(s);
    for (int i = 0; i < n / 2; i++) {
        double arg = 2.0 * pi * rn[2 * i];
        double x3 = sqrt(-2.0 * log(rn[2 * i + 1]));
        g[2 * i] = sin(arg) * x3;
        g[2 * i + 1] = cos(arg) * x3;
    }
    if (n % 2) {
        double arg = 2.0 * pi * rn[n - 1];
        g[n - 1] = cos(arg) * sqrt(-2.0 * log(rn[n]));
    }
}

/* Fortran nint (round half away from zero) wrapping: x - L*nint(x/L) */
static inline double wrap(double x, double el, double eli) { return x - el * round(x * eli); }
static inline double wrapd(double x, double el) { return x - el * round(x / el); }

/* ---------------------------------------------------------------- k shells */
/* kshell.f90:20-156 kstates3d: the first nsh0 shells of half-space k-vectors, ak = 2 pi m / L;
 * numsh0[s] vectors in shell s.  Returns the number of vectors, or -1 ("Too many states")
 * or -2 (more than maxk vectors). */
static int kshell_core(double el, int nsh0, int32_t *numsh0, double (*akv)[3], double *ak2v,
                       int maxk) {
    int i2max = 2 * nsh0;
    int imax = (int)lroundf(sqrtf((float)i2max));
    int cap = (imax + 1) * (imax + 1) * (imax + 1) * 4 + 8;
    int (*ik)[3] = malloc(sizeof(int[3]) * cap);
    int nk = 0;
    for (int ix = 0; ix <= imax; ix++)
        for (int iy = 0; iy <= imax; iy++)
            for (int iz = 0; iz <= imax; iz++) {
                int i2 = ix * ix + iy * iy + iz * iz;
                if (i2 > i2max) continue;
                ik[nk][0] = ix; ik[nk][1] = iy; ik[nk][2] = iz; nk++;
                if (ix == 0 && iy != 0 && iz != 0) { ik[nk][0] = ix; ik[nk][1] = iy; ik[nk][2] = -iz; nk++; }
                if (iy == 0 && ix != 0 && iz != 0) { ik[nk][0] = ix; ik[nk][1] = iy; ik[nk][2] = -iz; nk++; }
                if (iz == 0 && ix != 0 && iy != 0) { ik[nk][0] = ix; ik[nk][1] = -iy; ik[nk][2] = iz; nk++; }
                if (ix != 0 && iy != 0 && iz != 0) {
                    ik[nk][0] = ix; ik[nk][1] = iy; ik[nk][2] = -iz; nk++;
                    ik[nk][0] = ix; ik[nk][1] = -iy; ik[nk][2] = iz; nk++;
                    ik[nk][0] = ix; ik[nk][1] = -iy; ik[nk][2] = -iz; nk++;
                }
            }
    for (int i = 0; i < nk - 1; i++)
        for (int j = i + 1; j < nk; j++) {
            int a = ik[i][0] * ik[i][0] + ik[i][1] * ik[i][1] + ik[i][2] * ik[i][2];
            int b = ik[j][0] * ik[j][0] + ik[j][1] * ik[j][1] + ik[j][2] * ik[j][2];
            if (a >= b) {
                for (int d = 0; d < 3; d++) { int t = ik[i][d]; ik[i][d] = ik[j][d]; ik[j][d] = t; }
            }
        }
    int *numsh = calloc(nk + 2, sizeof(int));
    int i2m = 0, nsh = 1;
    for (int i = 0; i < nk; i++) {
        int i2 = ik[i][0] * ik[i][0] + ik[i][1] * ik[i][1] + ik[i][2] * ik[i][2];
        if (i2 > i2m) { nsh++; i2m = i2; }
        numsh[nsh]++;
    }
    for (int i = 1; i <= nsh; i++) numsh[i] *= 2;
    if (nsh0 >= nsh) { free(ik); free(numsh); return -1; } /* "Too many states" */
    double piel = 2.0 * (4.0 * atan(1.0)) / el;
    int k = 0;
    for (int i = 1; i <= nsh0; i++) {
        numsh0[i - 1] = numsh[i] / 2;
        int m = numsh[i] / 2 > 1 ? numsh[i] / 2 : 1;
        for (int j = 0; j < m; j++) {
            if (k >= maxk) { free(ik); free(numsh); return -2; }
            for (int d = 0; d < 3; d++) akv[k][d] = piel * ik[k][d];
            ak2v[k] = akv[k][0] * akv[k][0] + akv[k][1] * akv[k][1] + akv[k][2] * akv[k][2];
            k++;
        }
    }
    free(ik);
    free(numsh);
    return k;
}

static int kstates3d(dmco_ctx *c, int nsh0) {
    int k = kshell_core(c->el, nsh0, c->numsh, c->ak, c->ak2, MAXK);
    if (k < 0) return -1;
    c->nk = k;
    return 0;
}

int dmco_kshells(double el, int32_t nsh0, int32_t *numsh, double *ak, int32_t maxk) {
    double (*akv)[3] = malloc(sizeof(double[3]) * (maxk > 0 ? maxk : 1));
    double *ak2v = malloc(sizeof(double) * (maxk > 0 ? maxk : 1));
    int k = kshell_core(el, nsh0, numsh, akv, ak2v, maxk);
    for (int i = 0; i < k; i++)
        for (int d = 0; d < 3; d++) ak[3 * i + d] = akv[i][d];
    free(akv);
    free(ak2v);
    return k;
}

/* ---------------------------------------------------------------- beta(r) */
/* betafun.f90:127-149 beta0 */
static void beta0(const dmco_ctx *c, double r, double *b1, double *b1p, double *b1pp) {
    double x = c->b * r, d = c->d, cc = c->c;
    if (r < c->small) {
        *b1 = c->c0 + x * (c->c1 + x * (c->c2 + x * (c->c3 + x * c->c4)));
        *b1p = c->c1 + x * (2.0 * c->c2 + x * (3.0 * c->c3 + 4.0 * c->c4 * x));
        *b1pp = 2.0 * c->c2 + x * (6.0 * c->c3 + 12.0 * c->c4 * x);
    } else {
        double ex = exp(-x), ex10 = exp(-d * x);
        *b1 = (1.0 + cc * x) * (1.0 - ex10) * ex / x;
        *b1p = ex * (-1.0 - x * (1.0 + cc * x) + ex10 * (1.0 + x * (1.0 + d) * (1.0 + cc * x))) / (x * x);
        *b1pp = ex * (2.0 + x * (2.0 + x * (1.0 + cc * x)) -
                      ex10 * (2.0 + x * (1.0 + d) * (2.0 + x * (1.0 + d) * (1.0 + x * cc)))) / (x * x * x);
    }
    *b1p = c->b * *b1p;
    *b1pp = c->b * c->b * *b1pp;
    *b1 /= d;
    *b1p /= d;
    *b1pp /= d;
}
/* betafun.f90:151-171 beta */
void dmco_beta(const dmco_ctx *c, double r, double out[3]) {
    if (r > c->el * 0.5) { out[0] = out[1] = out[2] = 0.0; return; }
    double bet0, t1, t2, bet1, bet1p, bet1pp, bet2, bet2p, bet2pp;
    beta0(c, c->el * 0.5, &bet0, &t1, &t2);
    beta0(c, r, &bet1, &bet1p, &bet1pp);
    beta0(c, c->el - r, &bet2, &bet2p, &bet2pp);
    out[0] = bet1 + bet2 - 2.0 * bet0;
    out[1] = (r < 1.0e-14) ? 0.0 : bet1p - bet2p;
    out[2] = bet1pp + bet2pp;
}
/* betafun.f90:74-125 initbeta */
static void initbeta(dmco_ctx *c, double bb, double db) {
    double b = bb, d = db, el = c->el;
    c->b = b;
    c->d = d;
    c->c = 0.5 * (2.0 + 2.0 * d * b * el + (d * b * el) * (d * b * el) * exp(b * el * (1.0 + d)) +
                  2.0 * d * b * b * el * el * exp(b * el * (1.0 + d)) + 2.0 * b * el -
                  2.0 * exp(d * b * el) * b * el - 2.0 * exp(d * b * el)) /
           (el * el * (exp(d * b * el) - 1.0 - d + d * exp(b * el * (1.0 + d))) * b * b);
    double cc = c->c;
    c->c0 = d;
    c->small = 1.0e-3 / b;
    c->c1 = -d + cc * d - d * d / 2.0;
    c->c2 = d / 2.0 + d * d / 2.0 + d * d * d / 6.0 + cc * (-d - d * d / 2.0);
    c->c3 = -d / 6.0 - d * d / 4.0 - d * d * d / 6.0 - d * d * d * d / 24.0 +
            cc * (d / 2.0 + d * d / 2.0 + d * d * d / 6.0);
    c->c4 = d / 24.0 + d * d / 12.0 + d * d * d / 12.0 + d * d * d * d / 24.0 +
            d * d * d * d * d / 120.0 - cc * (d / 6.0 + d * d / 4.0 + d * d * d / 6.0 + d * d * d * d / 24.0);
    int nt = c->ntab;
    double dr = 0.5 * el / nt;
    int nj = 0;
    double betan = 10.0, o[3];
    while (fabs(betan) > 1.0e-10) {
        nj++;
        dmco_beta(c, el * nj, o);
        betan = o[0];
    }
    for (int i = 0; i <= nt; i++) {
        double r = i * dr, b1[3], b2[3], b3[3];
        dmco_beta(c, r, b1);
        c->bt[i] = b1[0];
        c->dbt[i] = b1[1];
        c->d2bt[i] = b1[2];
        for (int j = 1; j <= nj; j++) {
            dmco_beta(c, j * el - r, b2);
            dmco_beta(c, j * el + r, b3);
            c->bt[i] += b2[0] + b3[0];
            c->dbt[i] += -b2[1] + b3[1];
            c->d2bt[i] += b2[2] + b3[2];
        }
    }
    double last = c->bt[nt];
    for (int i = 0; i <= nt; i++) {
        double r = i * dr;
        c->bt[i] -= last;
        if (i != 0) c->d2bt[i] += 2.0 * c->dbt[i] / r;
        else c->d2bt[i] = 3.0 * c->d2bt[i];
    }
    /* setbeta bfact scaling (betafun.f90:62-64) */
    double f = fabs(c->p.bfact);
    for (int i = 0; i <= nt; i++) { c->bt[i] *= f; c->dbt[i] *= f; c->d2bt[i] *= f; }
}

/* 4-point Lagrange interpolation weights (coshpot.f90:172-179, betafun.f90:183-190) */
static inline int lagrange(double s, int ntab, double w[4]) {
    int idx = (int)s;
    if (idx > ntab - 2) idx = ntab - 2;
    if (idx < 1) idx = 1;
    double dr = s - idx;
    w[0] = -dr * (dr - 1.0) * (dr - 2.0) / 6.0;
    w[1] = (dr + 1.0) * (dr - 1.0) * (dr - 2.0) / 2.0;
    w[2] = -(dr + 1.0) * dr * (dr - 2.0) / 2.0;
    w[3] = (dr + 1.0) * dr * (dr - 1.0) / 6.0;
    return idx;
}
static inline double interp(const double *t, int idx, const double w[4]) {
    return w[0] * t[idx - 1] + w[1] * t[idx] + w[2] * t[idx + 1] + w[3] * t[idx + 2];
}

/* betafun.f90:173-194 radialfun */
void dmco_radial(const dmco_ctx *c, double r, double out[3]) {
    if (c->nobeta || r > c->rangeb) { out[0] = out[1] = out[2] = 0.0; return; }
    double w[4];
    int idx = lagrange(c->scaleb * r, c->ntab, w);
    out[0] = interp(c->bt, idx, w);
    out[1] = interp(c->dbt, idx, w);
    out[2] = interp(c->d2bt, idx, w);
}

/* betafun.f90:196-327 derbeta: d beta / d(b, d) of the closed form of beta (betafun.f90:
 * 151-171 composed of beta0, 127-149, large-x branch), including the dependence of the
 * constant c on (b, d) (betafun.f90:80-83).  The reference's expression is machine
 * generated; here the same derivative is taken exactly by forward-mode dual numbers. */
typedef struct { double v, b, d; } dual;
static inline dual dc(double v) { return (dual){v, 0.0, 0.0}; }
static inline dual dadd(dual x, dual y) { return (dual){x.v + y.v, x.b + y.b, x.d + y.d}; }
static inline dual dsub(dual x, dual y) { return (dual){x.v - y.v, x.b - y.b, x.d - y.d}; }
static inline dual dmul(dual x, dual y) {
    return (dual){x.v * y.v, x.b * y.v + x.v * y.b, x.d * y.v + x.v * y.d};
}
static inline dual ddiv(dual x, dual y) {
    double q = x.v / y.v;
    return (dual){q, (x.b - q * y.b) / y.v, (x.d - q * y.d) / y.v};
}
static inline dual dexp(dual x) { double e = exp(x.v); return (dual){e, e * x.b, e * x.d}; }
static dual cfun_d(dual b, dual d, double el) {
    dual E = dc(el), one = dc(1.0), two = dc(2.0);
    dual dbe = dmul(dmul(d, b), E);                     /* d b L */
    dual e1 = dexp(dmul(dmul(b, E), dadd(one, d)));     /* exp(b L (1+d)) */
    dual e2 = dexp(dbe);                                /* exp(d b L) */
    dual num = dadd(dadd(dadd(two, dmul(two, dbe)), dmul(dmul(dbe, dbe), e1)),
                    dmul(dmul(dmul(two, d), dmul(dmul(b, b), dc(el * el))), e1));
    num = dsub(dsub(dadd(num, dmul(two, dmul(b, E))), dmul(dmul(two, e2), dmul(b, E))),
               dmul(two, e2));
    dual den = dmul(dmul(dc(el * el), dadd(dsub(dsub(e2, one), d), dmul(d, e1))), dmul(b, b));
    return dmul(dc(0.5), ddiv(num, den));
}
static dual bfun0_d(double r, dual b, dual d, dual c) {
    dual x = dmul(b, dc(r)), one = dc(1.0);
    dual t = dmul(dmul(dadd(one, dmul(c, x)), dsub(one, dexp(dmul(dc(-1.0), dmul(d, x))))),
                  dexp(dmul(dc(-1.0), x)));
    return ddiv(t, dmul(x, d));
}
void dmco_derbeta(const dmco_ctx *c, double r, double out[2]) {
    out[0] = out[1] = 0.0;
    if (c->nobeta || r > c->rangeb) return;
    double el = c->el;
    dual b = {c->b, 1.0, 0.0}, d = {c->d, 0.0, 1.0};
    dual cc = cfun_d(b, d, el);
    dual f = dadd(bfun0_d(r, b, d, cc), bfun0_d(el - r, b, d, cc));
    f = dsub(f, dmul(dc(2.0), bfun0_d(0.5 * el, b, d, cc)));
    out[0] = f.b;
    out[1] = f.d;
}

/* ---------------------------------------------------------------- Jastrow tables */
/* coshpot.f90:145-152 vpot (hbar here is hbar^2/(2 m_r) = the table's hbar) */
static double vpot_v0(const dmco_ctx *c, double v0, double r) {
    double ch = cosh(c->p.amu * r);
    return -2.0 * v0 * c->p.amu * c->p.amu * c->hbar_t / (ch * ch);
}
double dmco_vpot(const dmco_ctx *c, double r) { return vpot_v0(c, c->p.v0, r); }

/* rungekutta.f90:127-150 rk3d */
static void rk3d(const dmco_ctx *c, double v0, double *psi, double *dpsi, double ene,
                 double hb, double dr, int nt, int *ncross) {
    psi[0] = 0.0;
    dpsi[0] = 1.0;
    *ncross = 0;
    for (int i = 1; i <= nt; i++) {
        double r = (i - 1) * dr;
        double vr = vpot_v0(c, v0, r), vh = vpot_v0(c, v0, r + 0.5 * dr), vd = vpot_v0(c, v0, r + dr);
        double k1a = -(ene - vr) * psi[i - 1] / hb;
        double k1b = dpsi[i - 1];
        double k2a = -(ene - vh) * (psi[i - 1] + 0.5 * dr * k1b) / hb;
        double k2b = dpsi[i - 1] + 0.5 * dr * k1a;
        double k3a = -(ene - vh) * (psi[i - 1] + 0.5 * dr * k2b) / hb;
        double k3b = dpsi[i - 1] + 0.5 * dr * k2a;
        double k4a = -(ene - vd) * (psi[i - 1] + dr * k3b) / hb;
        double k4b = dpsi[i - 1] + dr * k3a;
        dpsi[i] = dpsi[i - 1] + (k1a + 2.0 * k2a + 2.0 * k3a + k4a) * dr / 6.0;
        psi[i] = psi[i - 1] + (k1b + 2.0 * k2b + 2.0 * k3b + k4b) * dr / 6.0;
        if (psi[i] != copysign(psi[i], psi[i - 1])) (*ncross)++;
    }
}
/* rungekutta.f90:8-100 orbital (3D branch, nodes = 0) */
static int orbital(const dmco_ctx *c, double v0, double *psi, double *dpsi, double *d2psi,
                   double hb, int nt, double dr) {
    int ncross;
    double emax = vpot_v0(c, v0, 0.0), emin = emax;
    for (int i = 1; i <= nt; i++) {
        double v = vpot_v0(c, v0, i * dr);
        if (v > emax) emax = v;
        if (v < emin) emin = v;
    }
    emax = 0.001;
    rk3d(c, v0, psi, dpsi, emax, hb, dr, nt, &ncross);
    while (ncross <= 0) {
        emax = 2.0 * emax;
        rk3d(c, v0, psi, dpsi, emax, hb, dr, nt, &ncross);
    }
    double ene = 0.5 * (emin + emax);
    int iter = 0;
    for (;;) {
        rk3d(c, v0, psi, dpsi, ene, hb, dr, nt, &ncross);
        double r = dr * nt;
        double der = (dpsi[nt] - psi[nt] / r) / r;
        der = der / psi[nt];
        if (der < 0.0) emax = ene; else emin = ene;
        if (fabs(der) < 1.0e-8) break;
        ene = 0.5 * (emin + emax);
        iter++;
        if (iter > 100) return -1; /* "orbital: no convergence" */
    }
    psi[0] = dpsi[0];
    d2psi[0] = (vpot_v0(c, v0, 0.0) - ene) * dpsi[0] / (3.0 * hb);
    dpsi[0] = 0.0;
    for (int i = 1; i <= nt; i++) {
        double r = i * dr;
        dpsi[i] = (dpsi[i] - psi[i] / r) / r;
        psi[i] = psi[i] / r;
        d2psi[i] = (vpot_v0(c, v0, r) - ene) * psi[i] / hb - 2.0 * dpsi[i] / r;
    }
    return 0;
}
/* coshpot.f90:14-143 setphi (3D, periodic box) */
static int setphi(dmco_ctx *c) {
    int nt = c->ntab;
    c->range = 0.5 * c->el;
    double dr = c->range / nt;
    c->scale = 1.0 / dr;
    double v0op = c->p.v0op != 0.0 ? c->p.v0op : c->p.v0;
    double v0eq = c->p.v0eq != 0.0 ? c->p.v0eq : c->p.v0;
    for (int i = 0; i <= nt; i++) {
        c->vt[i] = vpot_v0(c, c->p.v0, i * dr);
        c->vt2[i] = 0.0;
    }
    int ntab0 = (int)(nt * c->p.djas / (0.5 * c->el));
    c->ntab0 = ntab0;
    double *f = calloc(ntab0 + 1, sizeof(double)), *df = calloc(ntab0 + 1, sizeof(double)),
           *d2f = calloc(ntab0 + 1, sizeof(double));
    double cj[2] = {c->p.cjas1, c->p.cjas2};
    double v0s[2] = {v0op, v0eq};
    double *U[2] = {c->ut, c->ut2}, *DU[2] = {c->dut, c->dut2}, *D2U[2] = {c->d2ut, c->d2ut2};
    for (int ic = 0; ic < 2; ic++) {
        for (int i = 0; i <= nt; i++) U[ic][i] = DU[ic][i] = D2U[ic][i] = 0.0;
        if (cj[ic] != 0.0) {
            if (orbital(c, v0s[ic], f, df, d2f, c->hbar_t, ntab0, dr)) { free(f); free(df); free(d2f); return -1; }
            for (int i = 0; i <= ntab0 && i <= nt; i++) {
                U[ic][i] = f[i] / f[ntab0];
                DU[ic][i] = df[i] / f[ntab0];
                D2U[ic][i] = d2f[i] / f[ntab0];
            }
        }
        for (int i = ntab0 + 1; i <= nt; i++) { U[ic][i] = 1.0; DU[ic][i] = 0.0; D2U[ic][i] = 0.0; }
        for (int i = 0; i <= nt; i++) {
            if (cj[ic] != 0.0) {
                DU[ic][i] = DU[ic][i] / U[ic][i];
                D2U[ic][i] = D2U[ic][i] / U[ic][i] - DU[ic][i] * DU[ic][i];
                U[ic][i] = log(U[ic][i]);
            }
            U[ic][i] *= cj[ic];
            DU[ic][i] *= cj[ic];
            D2U[ic][i] *= cj[ic];
        }
    }
    c->ieq = c->p.cjas2 != 0.0;
    free(f); free(df); free(d2f);
    return 0;
}

/* ---------------------------------------------------------------- set-up */
dmco_ctx *dmco_create(const dmco_params *p) {
    dmco_ctx *c = calloc(1, sizeof(dmco_ctx));
    c->p = *p;
    c->pi = 4.0 * atan(1.0);
    c->npart = p->npart;
    c->n = p->npart / 2;
    c->ntab = p->ntab;
    /* psislaterbcs.f90:70 and wavefunction.f90:91-92: L from the density */
    c->el = pow(p->npart / p->rho, 1.0 / 3.0);
    c->eli = 1.0 / c->el;
    double mu = p->mass * p->mass / (p->mass + p->mass);
    /* dmcbcs.f90:143 setphi(..., hbar/(2 mu)); coshpot.f90:33 doubles it */
    c->hbar_t = 2.0 * (p->hbar / (2.0 * mu));
    /* dmcbcs.f90:146-166 units */
    if (p->units) {
        double rho = p->npart / (c->el * c->el * c->el);
        double kunit = pow(3.0 * c->pi * c->pi * rho, 1.0 / 3.0);
        c->eunit = 3.0 / 5.0 * p->hbar / (2.0 * mu) * kunit * kunit;
    } else {
        c->eunit = 1.0;
    }
    c->etrial = p->etrial * c->eunit * p->npart;
    c->nfac = 1.0 / c->eunit / p->npart; /* step.f90:24-25 */
    /* step.f90:29-30 */
    c->sigma = sqrt(2.0 * p->hbar * p->dt) / sqrt(p->mass);
    c->sig2 = c->sigma * c->sigma;
    int nt = p->ntab;
    double *blk = calloc(11 * (nt + 1), sizeof(double));
    c->vt = blk; c->ut = blk + (nt + 1); c->dut = blk + 2 * (nt + 1); c->d2ut = blk + 3 * (nt + 1);
    c->vt2 = blk + 4 * (nt + 1); c->ut2 = blk + 5 * (nt + 1); c->dut2 = blk + 6 * (nt + 1);
    c->d2ut2 = blk + 7 * (nt + 1);
    c->bt = blk + 8 * (nt + 1); c->dbt = blk + 9 * (nt + 1); c->d2bt = blk + 10 * (nt + 1);
    for (int i = 0; i < p->norb; i++) c->vorb[i] = p->vorb[i];
    if (kstates3d(c, p->norb)) { dmco_destroy(c); return NULL; }
    /* betafun.f90:12-72 setbeta (bfact >= 0 branch) */
    c->nobeta = p->bfact == 0.0;
    c->rangeb = 0.5 * c->el;
    c->scaleb = nt / c->rangeb;
    if (!c->nobeta) initbeta(c, p->bbeta, p->dbeta);
    if (setphi(c)) { dmco_destroy(c); return NULL; }
    return c;
}

void dmco_destroy(dmco_ctx *c) {
    if (!c) return;
    free(c->vt);
    free(c);
}

void dmco_info(const dmco_ctx *c, double out[8]) {
    out[0] = c->el; out[1] = c->eunit; out[2] = c->etrial; out[3] = c->nfac;
    out[4] = c->nk; out[5] = c->ntab0; out[6] = c->sigma; out[7] = c->scale;
}

int dmco_kvec(const dmco_ctx *c, double *ak, double *ak2, int32_t *numsh) {
    for (int k = 0; k < c->nk; k++) {
        for (int d = 0; d < 3; d++) ak[3 * k + d] = c->ak[k][d];
        ak2[k] = c->ak2[k];
    }
    for (int i = 0; i < c->p.norb; i++) numsh[i] = c->numsh[i];
    return c->nk;
}

void dmco_tables(const dmco_ctx *c, double *vt, double *ut, double *dut, double *d2ut,
                 double *bt, double *dbt, double *d2bt) {
    size_t s = sizeof(double) * (c->ntab + 1);
    memcpy(vt, c->vt, s); memcpy(ut, c->ut, s); memcpy(dut, c->dut, s); memcpy(d2ut, c->d2ut, s);
    memcpy(bt, c->bt, s); memcpy(dbt, c->dbt, s); memcpy(d2bt, c->d2bt, s);
}

void dmco_set_params(dmco_ctx *c, const double *vorb, double bbeta, double dbeta) {
    for (int i = 0; i < c->p.norb; i++) c->vorb[i] = vorb[i];
    if (!c->nobeta && (bbeta != c->b || dbeta != c->d)) initbeta(c, bbeta, dbeta);
}

/* ---------------------------------------------------------------- Jastrow + V */
/* one pair's table values: returns 0 if r >= range (coshpot.f90:167-193) */
static inline int pairtab(const dmco_ctx *c, double r, int itab, double *v, double *u,
                          double *du, double *d2u) {
    if (!(r < c->range)) return 0;
    double w[4];
    int idx = lagrange(c->scale * r, c->ntab, w);
    if (itab == 1) {
        *v = interp(c->vt, idx, w); *u = interp(c->ut, idx, w);
        *du = interp(c->dut, idx, w); *d2u = interp(c->d2ut, idx, w);
    } else {
        *v = interp(c->vt2, idx, w); *u = interp(c->ut2, idx, w);
        *du = interp(c->dut2, idx, w); *d2u = interp(c->d2ut2, idx, w);
    }
    return 1;
}
static inline int active_pair(const dmco_ctx *c, int i, int j) {
    int si = i < c->n, sj = j < c->n;
    return (si != sj) || c->ieq; /* iop is always set (coshpot.f90:88) */
}
/* coshpot.f90:154-205 getphi */
void dmco_getphi(const dmco_ctx *c, const double *x, double *u, double *du, double *d2u,
                 double *v) {
    int np = c->npart;
    double m = c->p.mass, sm = sqrt(m), mu = m * m / (m + m);
    *u = *d2u = *v = 0.0;
    for (int i = 0; i < 3 * np; i++) du[i] = 0.0;
    for (int i = 0; i < np - 1; i++)
        for (int j = i + 1; j < np; j++) {
            if (!active_pair(c, i, j)) continue;
            double dx[3], r2 = 0.0;
            for (int d = 0; d < 3; d++) {
                dx[d] = wrap(x[3 * i + d] - x[3 * j + d], c->el, c->eli);
                r2 += dx[d] * dx[d];
            }
            double r = fmax(1.0e-14, sqrt(r2));
            double vp, uj, duj, d2uj;
            int itab = ((i < c->n) != (j < c->n)) ? 1 : 2;
            if (!pairtab(c, r, itab, &vp, &uj, &duj, &d2uj)) continue;
            *v += vp;
            *u += uj;
            for (int d = 0; d < 3; d++) {
                du[3 * i + d] += dx[d] * duj / (r * sm);
                du[3 * j + d] -= dx[d] * duj / (r * sm);
            }
            *d2u += (d2uj + 2.0 * duj / r) / mu;
        }
    double s = 0.0;
    for (int i = 0; i < 3 * np; i++) s += du[i] * du[i];
    *d2u += s;
}

/* ---------------------------------------------------------------- pairing */
/* phi(r_i - r_j) and its gradient/laplacian w.r.t. r_i: the matrix element of allpair
 * (psislaterbcs.f90:248-293) written per pair: alpha_1 + sum_s 2 alpha_s sum_{k in s}
 * cos(k.(r_i - r_j)) + beta(|r_i - r_j|_mi). */
static void pairphi(const dmco_ctx *c, const double *xi, const double *xj, double *f,
                    double g[3], double *l) {
    *f = c->vorb[0];
    g[0] = g[1] = g[2] = 0.0;
    *l = 0.0;
    int k = 1;
    for (int ksh = 1; ksh < c->p.norb; ksh++) {
        double v = 2.0 * c->vorb[ksh];
        for (int j = 0; j < c->numsh[ksh]; j++, k++) {
            /* allpair evaluates cos/sin of k.r for each particle and recombines */
            double ai = c->ak[k][0] * xi[0] + c->ak[k][1] * xi[1] + c->ak[k][2] * xi[2];
            double aj = c->ak[k][0] * xj[0] + c->ak[k][1] * xj[1] + c->ak[k][2] * xj[2];
            double ci = cos(ai), si = sin(ai), cj = cos(aj), sj = sin(aj);
            double cc = ci * cj + si * sj, ss = si * cj - ci * sj;
            *f += v * cc;
            for (int d = 0; d < 3; d++) g[d] += -c->ak[k][d] * v * ss;
            *l += -c->ak2[k] * v * cc;
        }
    }
    double dx[3], r2 = 0.0;
    for (int d = 0; d < 3; d++) {
        dx[d] = wrapd(xi[d] - xj[d], c->el);
        r2 += dx[d] * dx[d];
    }
    double r = sqrt(r2), rf[3];
    dmco_radial(c, r, rf);
    *f += rf[0];
    for (int d = 0; d < 3; d++) g[d] += rf[1] * dx[d] / r;
    *l += rf[2];
}

/* LU with partial pivoting (dgetf2), log|det|, sign and inverse (matrixlapack.f90:13-41).
 * a is n x n column-major a[i + n*j]; on success a holds the inverse.  Returns 0, or
 * -1 for an exactly singular matrix (then detl = -1e30 and is is left unchanged). */
static int matinv(double *a, int n, double *detl, int32_t *is) {
    int piv[128];
    double lu[128 * 128];
    memcpy(lu, a, sizeof(double) * n * n);
    int sing = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = fabs(lu[k + n * k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(lu[i + n * k]) > mx) { mx = fabs(lu[i + n * k]); p = i; }
        piv[k] = p;
        if (lu[p + n * k] == 0.0) { sing = 1; continue; }
        if (p != k)
            for (int j = 0; j < n; j++) { double t = lu[k + n * j]; lu[k + n * j] = lu[p + n * j]; lu[p + n * j] = t; }
        double rp = 1.0 / lu[k + n * k];
        for (int i = k + 1; i < n; i++) lu[i + n * k] *= rp;
        for (int j = k + 1; j < n; j++) {
            double t = lu[k + n * j];
            for (int i = k + 1; i < n; i++) lu[i + n * j] -= lu[i + n * k] * t;
        }
    }
    if (sing) { *detl = -1.0e30; return -1; }
    double dl = 0.0;
    int s = 1;
    for (int i = 0; i < n; i++) {
        dl += log(fabs(lu[i + n * i]));
        if (lu[i + n * i] < 0.0) s = -s;
        if (piv[i] != i) s = -s;
    }
    *detl = dl;
    *is = s;
    /* inverse: solve A X = I column by column */
    for (int col = 0; col < n; col++) {
        double b[128];
        for (int i = 0; i < n; i++) b[i] = (i == col) ? 1.0 : 0.0;
        for (int k = 0; k < n; k++)
            if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
        for (int i = 0; i < n; i++)
            for (int k = 0; k < i; k++) b[i] -= lu[i + n * k] * b[k];
        for (int i = n - 1; i >= 0; i--) {
            for (int k = i + 1; k < n; k++) b[i] -= lu[i + n * k] * b[k];
            b[i] /= lu[i + n * i];
        }
        for (int i = 0; i < n; i++) a[i + n * col] = b[i];
    }
    return 0;
}

/* pair matrix S(i,j) = phi(r_up_i - r_dn_j) with gradient (w.r.t. r_up_i) and laplacian;
 * column-major s[i + n*j], ds[3*(i + n*j) + d] (psislaterbcs.f90:227-294). */
static void allpair(const dmco_ctx *c, const double *x, double *s, double *ds, double *d2s) {
    int n = c->n;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double f, g[3], l;
            pairphi(c, x + 3 * i, x + 3 * (n + j), &f, g, &l);
            s[i + n * j] = f;
            if (ds) for (int d = 0; d < 3; d++) ds[3 * (i + n * j) + d] = g[d];
            if (d2s) d2s[i + n * j] = l;
        }
}

/* psislaterbcs.f90:155-225 bcs (no unpaired orbitals) */
void dmco_bcs(const dmco_ctx *c, const double *x, double *phil, int32_t *is, double *dphi,
              double *d2phi, double *sinv) {
    int n = c->n;
    double *s = malloc(sizeof(double) * n * n * 5), *ds = s + n * n, *d2s = s + 4 * n * n;
    allpair(c, x, s, ds, d2s);
    *is = 1;
    double sm = sqrt(c->p.mass), m = c->p.mass;
    for (int i = 0; i < 3 * 2 * n; i++) dphi[i] = 0.0;
    *d2phi = 0.0;
    if (matinv(s, n, phil, is)) { free(s); return; }
    for (int i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) {
            double acc = 0.0;
            for (int j = 0; j < n; j++) acc += s[j + n * i] * ds[3 * (i + n * j) + d];
            dphi[3 * i + d] = acc / sm;
        }
        double acc = 0.0;
        for (int j = 0; j < n; j++) acc += s[j + n * i] * d2s[i + n * j];
        *d2phi += acc / m;
    }
    for (int i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) {
            double acc = 0.0;
            for (int j = 0; j < n; j++) acc += s[i + n * j] * ds[3 * (j + n * i) + d];
            dphi[3 * (n + i) + d] = -acc / sm;
        }
        double acc = 0.0;
        for (int j = 0; j < n; j++) acc += s[i + n * j] * d2s[j + n * i];
        *d2phi += acc / m;
    }
    if (sinv) memcpy(sinv, s, sizeof(double) * n * n);
    free(s);
}

/* wavefunction.f90:161-201 hpsi with a2 = 0, no one-body or long-range terms */
void dmco_hpsi(const dmco_ctx *c, const double *x, double *out, double *dpsi) {
    int np = c->npart;
    double u, d2u, v, phil, d2phi;
    int32_t is;
    double *du = malloc(sizeof(double) * 6 * np), *dphi = du + 3 * np;
    dmco_getphi(c, x, &u, du, &d2u, &v);
    dmco_bcs(c, x, &phil, &is, dphi, &d2phi, NULL);
    double dot = 0.0;
    for (int i = 0; i < 3 * np; i++) {
        dpsi[i] = dphi[i] + du[i];
        dot += du[i] * dphi[i];
    }
    out[0] = phil + u;
    out[1] = d2phi + d2u + 2.0 * dot;
    out[2] = v;
    out[3] = is;
    free(du);
}

/* coshpot.f90:207-264 getphione: exp(sum_j u(r_i' - r_j) - u(r_i - r_j)) */
static double jastrow_ratio(const dmco_ctx *c, const double *x, int i, const double *xnew) {
    double u = 0.0;
    for (int j = 0; j < c->npart; j++) {
        if (j == i || !active_pair(c, i, j)) continue;
        int itab = ((i < c->n) != (j < c->n)) ? 1 : 2;
        for (int pass = 0; pass < 2; pass++) {
            const double *xi = pass ? xnew : x + 3 * i;
            double r2 = 0.0;
            for (int d = 0; d < 3; d++) {
                double dx = wrap(xi[d] - x[3 * j + d], c->el, c->eli);
                r2 += dx * dx;
            }
            double r = fmax(1.0e-14, sqrt(r2)), vp, uj, duj, d2uj;
            if (pairtab(c, r, itab, &vp, &uj, &duj, &d2uj)) u += pass ? uj : -uj;
        }
    }
    return exp(u);
}
/* psislaterbcs.f90:296-376 bcsone: det ratio from the stored inverse */
static double det_ratio(const dmco_ctx *c, const double *x, const double *sinv, int i,
                        const double *xnew) {
    int n = c->n;
    double rat = 0.0, f, g[3], l;
    if (i < n) {
        for (int j = 0; j < n; j++) {
            pairphi(c, xnew, x + 3 * (n + j), &f, g, &l);
            rat += sinv[j + n * i] * f;
        }
    } else {
        int jd = i - n;
        for (int j = 0; j < n; j++) {
            pairphi(c, x + 3 * j, xnew, &f, g, &l);
            rat += sinv[jd + n * j] * f;
        }
    }
    return rat;
}
/* wavefunction.f90:393-404 move1 */
double dmco_ratio1(const dmco_ctx *c, const double *x, int32_t ip, const double *xnew) {
    int n = c->n, np = c->npart;
    double *sinv = malloc(sizeof(double) * (n * n + 3 * np)), *dphi = sinv + n * n;
    double phil, d2phi;
    int32_t is;
    dmco_bcs(c, x, &phil, &is, dphi, &d2phi, sinv);
    double r = jastrow_ratio(c, x, ip, xnew) * det_ratio(c, x, sinv, ip, xnew);
    free(sinv);
    return r;
}

/* psislaterbcs.f90:449-524 getderbcs with optsw and, when optbeta is set, optbeta:
 * d log psi / d (alpha_1 .. alpha_norb [, b, d]) = sum_ij Sinv(j, i) dS(i, j), with
 * dS = 1 for alpha_1, 2 sum_{k in shell} cos(k.r_ij) for the other shells and the closed-form
 * derbeta (betafun.f90:196-327) for b and d (psislaterbcs.f90:510-517). */
void dmco_getder(const dmco_ctx *c, const double *x, int32_t optbeta, double *dlogpsi) {
    int n = c->n, norb = c->p.norb, np = norb + (optbeta ? 2 : 0);
    double *s = malloc(sizeof(double) * n * n), phil;
    int32_t is = 1;
    allpair(c, x, s, NULL, NULL);
    for (int p = 0; p < np; p++) dlogpsi[p] = 0.0;
    if (matinv(s, n, &phil, &is)) { free(s); return; }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double dx[3];
            for (int d = 0; d < 3; d++) dx[d] = wrapd(x[3 * i + d] - x[3 * (n + j) + d], c->el);
            double sij = s[j + n * i];
            dlogpsi[0] += sij;
            int k = 1;
            for (int ksh = 1; ksh < norb; ksh++) {
                double acc = 0.0;
                for (int kk = 0; kk < c->numsh[ksh]; kk++, k++)
                    acc += 2.0 * cos(c->ak[k][0] * dx[0] + c->ak[k][1] * dx[1] + c->ak[k][2] * dx[2]);
                dlogpsi[ksh] += sij * acc;
            }
            if (optbeta) {
                double db[2];
                dmco_derbeta(c, sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]), db);
                dlogpsi[norb] += sij * db[0];
                dlogpsi[norb + 1] += sij * db[1];
            }
        }
    free(s);
}

/* ---------------------------------------------------------------- walkers */
/* dmcbcs.f90:237-301 with isite = .true. */
void dmco_init_walkers(const dmco_ctx *c, int64_t seed, int32_t nwalk, double *x,
                       int64_t *irn) {
    int np = c->npart;
    int64_t mod = seed & (int64_t)MASK48; /* setrn masks to 48 bits */
    int64_t loc = seed;
    double *cfg = malloc(sizeof(double) * 3 * np);
    for (int i = 0; i < 3 * np; i++) cfg[i] = c->el * (dmco_ran1(&mod) - 0.5);
    for (int k = 0; k < nwalk; k++) {
        for (int i = 0; i < np; i++) {
            double u = dmco_ran1(&mod);
            int j = (int)(np * u + 1.0) - 1;
            for (int d = 0; d < 3; d++) {
                double t = cfg[3 * i + d];
                cfg[3 * i + d] = cfg[3 * j + d];
                cfg[3 * j + d] = t;
            }
        }
        memcpy(x + (size_t)3 * np * k, cfg, sizeof(double) * 3 * np);
        dmco_ran2(&loc);
        irn[k] = loc;
    }
    free(cfg);
}

void dmco_init_hpsi(const dmco_ctx *c, int32_t nw, const double *x, double *dpsi, double *sc) {
    int np = c->npart;
#pragma omp parallel for schedule(dynamic, 4)
    for (int w = 0; w < nw; w++) {
        double o[4];
        dmco_hpsi(c, x + (size_t)3 * np * w, o, dpsi + (size_t)3 * np * w);
        double *s = sc + 6 * w;
        s[0] = 1.0; s[1] = o[0]; s[2] = o[1]; s[3] = o[2]; s[4] = o[3]; s[5] = o[3];
    }
}

static inline void addval(double *est, int i, double val, double wt) {
    est[3 * (i - 1)] += val * wt;
    est[3 * (i - 1) + 1] += val * val * wt;
    est[3 * (i - 1) + 2] += wt;
}

/* step.f90:56-76,127-160 (idmc = 1) for one walker; per-walker estimator contributions
 * go to est (14 x 3) */
static void step1_one(const dmco_ctx *c, const double *x, const double *dpsi, const double *sc,
                      int64_t irn, int opcomp, double *est, double *xo, double *dpsio,
                      double *sco, int64_t *irno, double *wto) {
    int np = c->npart;
    double hbar = c->p.hbar, el = c->el;
    double *g = malloc(sizeof(double) * 3 * np), *x1 = malloc(sizeof(double) * 3 * np);
    for (int i = 0; i < 3 * np; i++) x1[i] = wrapd(x[i], el);
    int64_t s = irn;
    dmco_gaussian(&s, 3 * np, g);
    double sig = c->sigma, sm = sqrt(c->p.mass);
    for (int i = 0; i < 3 * np; i++) {
        double gs = sig * g[i];
        double drift = dpsi[i] * (sig * sig) * sm; /* dcut < 0: no cut */
        if (c->p.dcut >= 0.0) {
            double dc = sig * c->p.dcut;
            drift = fmin(dc, fmax(-dc, drift));
        }
        double dx = drift + gs;
        xo[i] = wrapd(x1[i] + dx, el);
    }
    double o[4];
    dmco_hpsi(c, xo, o, dpsio);
    double psil2 = o[0], d2psi2 = o[1], v2 = o[2];
    int is2 = (int)o[3];
    double en = -hbar * d2psi2 + v2;
    double ego = -hbar * sc[2] + sc[3];
    double egn = en;
    double wt = exp(-(0.5 * (ego + egn) - c->etrial) * c->p.dt);
    double weight = fmin(c->p.wtcut, wt);
    if (is2 != (int)sc[4]) {
        if (c->p.fn) { weight = 0.0; wt = 0.0; }
        addval(est, 5, 1.0, 1.0);
    } else {
        addval(est, 5, 0.0, 1.0);
    }
    double wt0 = sc[5];
    double wtg = wt * exp(psil2 - psil2) * is2 * wt0;
    if (opcomp) {
        double nfac = c->nfac;
        addval(est, 1, en * nfac, wtg);
        addval(est, 2, (en - v2) * nfac, wtg);
        addval(est, 3, v2 * nfac, wtg);
        addval(est, 4, 0.0, wtg);
        addval(est, 6, 0.5 * (ego + egn) * nfac, 1.0);
        addval(est, 7, exp(psil2 - psil2), 1.0);
        addval(est, 8, is2 * exp(psil2), 1.0);
        addval(est, 9, exp(psil2), 1.0);
        addval(est, 10, 1.0 * is2, 1.0);
        addval(est, 11, wt, 1.0);
        addval(est, 12, wtg, 1.0);
        addval(est, 13, copysign(1.0, wtg), 1.0);
        addval(est, 14, copysign(1.0, wtg), wtg);
    }
    sco[0] = weight; sco[1] = psil2; sco[2] = d2psi2; sco[3] = v2; sco[4] = is2; sco[5] = wt0;
    *irno = s;
    if (wto) *wto = wt;
    free(g);
    free(x1);
}

void dmco_step1(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                const double *sc, const int64_t *irn, int32_t opcomp, double *est,
                double *x_out, double *dpsi_out, double *sc_out, int64_t *irn_out,
                double *wt_out, int32_t nthreads) {
    int np = c->npart;
    double *pe = calloc((size_t)nw * 42, sizeof(double));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int w = 0; w < nw; w++) {
        size_t o = (size_t)3 * np * w;
        step1_one(c, x + o, dpsi + o, sc + 6 * w, irn[w], opcomp, pe + 42 * w, x_out + o,
                  dpsi_out + o, sc_out + 6 * w, irn_out + w, wt_out ? wt_out + w : NULL);
    }
    /* fixed-order reduction of the per-walker estimator contributions */
    for (int w = 0; w < nw; w++)
        for (int k = 0; k < 42; k++) est[k] += pe[42 * w + k];
    free(pe);
}

/* step.f90:56-76,78-122 step1 with idmc = 0 (variational): the same proposal as step1,
 * then the Metropolis test with the drift-diffusion Green's functions and one more ran1
 * draw; the walker kept is the new one (wt0 = is) or the old one, with the RNG state after
 * that draw.  Estimators 1-6 of dmcbcs.f90:180-186 every step. */
static void vmc_one(const dmco_ctx *c, const double *x, const double *dpsi, const double *sc,
                    int64_t irn, double *est, double *xo, double *dpsio, double *sco,
                    int64_t *irno) {
    int np = c->npart;
    double hbar = c->p.hbar, el = c->el;
    double *g = malloc(sizeof(double) * 3 * np), *x1 = malloc(sizeof(double) * 3 * np);
    double *dro = malloc(sizeof(double) * 3 * np), *dx = malloc(sizeof(double) * 3 * np);
    double *xn = malloc(sizeof(double) * 3 * np), *dn = malloc(sizeof(double) * 3 * np);
    for (int i = 0; i < 3 * np; i++) x1[i] = wrapd(x[i], el);
    int64_t s = irn;
    dmco_gaussian(&s, 3 * np, g);
    double sig = c->sigma, sm = sqrt(c->p.mass), dc = sig * c->p.dcut;
    for (int i = 0; i < 3 * np; i++) {
        double drift = dpsi[i] * (sig * sig) * sm;
        if (c->p.dcut >= 0.0) drift = fmin(dc, fmax(-dc, drift));
        dro[i] = drift;
        dx[i] = drift + sig * g[i];
        xn[i] = wrapd(x1[i] + dx[i], el);
    }
    double o[4];
    dmco_hpsi(c, xn, o, dn);
    double prob = 2.0 * (o[0] - sc[1]);
    for (int i = 0; i < np; i++) {
        double dot = 0.0;
        for (int d = 0; d < 3; d++) {
            int t = 3 * i + d;
            double drn = dn[t] * (sig * sig) * sm;
            if (c->p.dcut >= 0.0) drn = fmin(dc, fmax(-dc, drn));
            dot += (dro[t] + drn) * (dro[t] - drn - 2.0 * dx[t]);
        }
        prob += dot / (2.0 * sig * sig);
    }
    prob = exp(fmin(0.0, prob));
    double rn = dmco_ran1(&s);
    int acc = rn < prob;
    const double *xk = acc ? xn : x1, *dk = acc ? dn : dpsi;
    double d2 = acc ? o[1] : sc[2], v = acc ? o[2] : sc[3], dot = 0.0;
    for (int t = 0; t < 3 * np; t++) dot += dk[t] * dk[t] / c->p.mass;
    double en = -hbar * d2 + v, ejf = -0.5 * hbar * (d2 - dot) + v, nfac = c->nfac;
    double wgt = sc[0];
    addval(est, 1, en * nfac, wgt);
    addval(est, 2, (en - v) * nfac, wgt);
    addval(est, 3, v * nfac, wgt);
    addval(est, 4, 0.0, wgt);
    addval(est, 5, ejf * nfac, wgt);
    addval(est, 6, acc ? 1.0 : 0.0, 1.0);
    memcpy(xo, xk, sizeof(double) * 3 * np);
    memcpy(dpsio, dk, sizeof(double) * 3 * np);
    sco[0] = wgt;
    sco[1] = acc ? o[0] : sc[1];
    sco[2] = d2;
    sco[3] = v;
    sco[4] = acc ? o[3] : sc[4];
    sco[5] = acc ? o[3] : sc[5];
    *irno = s;
    free(g); free(x1); free(dro); free(dx); free(xn); free(dn);
}

void dmco_vmc_step1(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                    const double *sc, const int64_t *irn, double *est, double *x_out,
                    double *dpsi_out, double *sc_out, int64_t *irn_out, int32_t nthreads) {
    int np = c->npart;
    double *pe = calloc((size_t)nw * 42, sizeof(double));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int w = 0; w < nw; w++) {
        size_t o = (size_t)3 * np * w;
        vmc_one(c, x + o, dpsi + o, sc + 6 * w, irn[w], pe + 42 * w, x_out + o, dpsi_out + o,
                sc_out + 6 * w, irn_out + w);
    }
    for (int w = 0; w < nw; w++)
        for (int k = 0; k < 42; k++) est[k] += pe[42 * w + k];
    free(pe);
}

/* step.f90:56-76,161-226 step1 with idmc = 2 (Metropolis DMC) for one walker: the proposal
 * and the Metropolis test of vmc_one; an accepted walker is weighted exp(-((E_old + E_new)/2 -
 * E_T) dt), capped, with the node test, a rejected one exp(-(E_old - E_T) dt) times its weight,
 * capped; estimators 1-13 of dmcbcs.f90:204-218 (node crossing 4 on accepted moves only,
 * acceptance 5 always, the rest on opcomp). */
static void mdmc_one(const dmco_ctx *c, const double *x, const double *dpsi, const double *sc,
                     int64_t irn, int opcomp, double *est, double *xo, double *dpsio,
                     double *sco, int64_t *irno, double *wto) {
    int np = c->npart;
    double hbar = c->p.hbar, el = c->el, nfac = c->nfac;
    double *g = malloc(sizeof(double) * 3 * np), *x1 = malloc(sizeof(double) * 3 * np);
    double *dro = malloc(sizeof(double) * 3 * np), *dx = malloc(sizeof(double) * 3 * np);
    double *xn = malloc(sizeof(double) * 3 * np), *dn = malloc(sizeof(double) * 3 * np);
    for (int i = 0; i < 3 * np; i++) x1[i] = wrapd(x[i], el);
    int64_t s = irn;
    dmco_gaussian(&s, 3 * np, g);
    double sig = c->sigma, sm = sqrt(c->p.mass), dc = sig * c->p.dcut;
    for (int i = 0; i < 3 * np; i++) {
        double drift = dpsi[i] * (sig * sig) * sm;
        if (c->p.dcut >= 0.0) drift = fmin(dc, fmax(-dc, drift));
        dro[i] = drift;
        dx[i] = drift + sig * g[i];
        xn[i] = wrapd(x1[i] + dx[i], el);
    }
    double o[4];
    dmco_hpsi(c, xn, o, dn);
    double prob = 2.0 * (o[0] - sc[1]);
    for (int i = 0; i < np; i++) {
        double dot = 0.0;
        for (int d = 0; d < 3; d++) {
            int t = 3 * i + d;
            double drn = dn[t] * (sig * sig) * sm;
            if (c->p.dcut >= 0.0) drn = fmin(dc, fmax(-dc, drn));
            dot += (dro[t] + drn) * (dro[t] - drn - 2.0 * dx[t]);
        }
        prob += dot / (2.0 * sig * sig);
    }
    prob = exp(fmin(0.0, prob));
    double rn = dmco_ran1(&s);
    int acc = rn < prob;
    double eo = -hbar * sc[2] + sc[3], wt0 = sc[5], wt, weight, e, v, psil;
    int is;
    if (acc) {
        e = -hbar * o[1] + o[2];
        v = o[2];
        psil = o[0];
        is = (int)o[3];
        wt = exp(-(0.5 * (eo + e) - c->etrial) * c->p.dt);
        weight = fmin(c->p.wtcut, wt);
        if (is != (int)sc[4]) {
            if (c->p.fn) { weight = 0.0; wt = 0.0; }
            addval(est, 4, 1.0, 1.0);
        } else {
            addval(est, 4, 0.0, 1.0);
        }
    } else {
        e = eo;
        v = sc[3];
        psil = sc[1];
        is = (int)sc[4];
        wt = exp(-(eo - c->etrial) * c->p.dt);
        weight = fmin(c->p.wtcut, wt * sc[0]);
    }
    double wtg = wt * is * wt0;
    if (opcomp) {
        addval(est, 1, e * nfac, wtg);
        addval(est, 2, (e - v) * nfac, wtg);
        addval(est, 3, v * nfac, wtg);
        addval(est, 6, 1.0, 1.0);
        addval(est, 7, is * exp(psil), 1.0);
        addval(est, 8, exp(psil), 1.0);
        addval(est, 9, 1.0 * is, 1.0);
        addval(est, 10, wt, 1.0);
        addval(est, 11, wtg, 1.0);
        addval(est, 12, copysign(1.0, wtg), 1.0);
        addval(est, 13, copysign(1.0, wtg), wtg);
    }
    addval(est, 5, acc ? 1.0 : 0.0, 1.0);
    memcpy(xo, acc ? xn : x1, sizeof(double) * 3 * np);
    memcpy(dpsio, acc ? dn : dpsi, sizeof(double) * 3 * np);
    sco[0] = weight;
    sco[1] = psil;
    sco[2] = acc ? o[1] : sc[2];
    sco[3] = v;
    sco[4] = is;
    sco[5] = wt0;
    *irno = s;
    if (wto) *wto = wt;
    free(g); free(x1); free(dro); free(dx); free(xn); free(dn);
}

void dmco_mdmc_step1(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                     const double *sc, const int64_t *irn, int32_t opcomp, double *est,
                     double *x_out, double *dpsi_out, double *sc_out, int64_t *irn_out,
                     double *wt_out, int32_t nthreads) {
    int np = c->npart;
    double *pe = calloc((size_t)nw * 42, sizeof(double));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int w = 0; w < nw; w++) {
        size_t o = (size_t)3 * np * w;
        mdmc_one(c, x + o, dpsi + o, sc + 6 * w, irn[w], opcomp, pe + 42 * w, x_out + o,
                 dpsi_out + o, sc_out + 6 * w, irn_out + w, wt_out ? wt_out + w : NULL);
    }
    for (int w = 0; w < nw; w++)
        for (int k = 0; k < 42; k++) est[k] += pe[42 * w + k];
    free(pe);
}

/* step.f90:287-315 branch */
int dmco_branch(int32_t nw, int32_t npart, const double *x, const double *dpsi,
                const double *sc, const int64_t *irn, double factor, int32_t cap,
                double *x_out, double *dpsi_out, double *sc_out, int64_t *irn_out) {
    int nt = 0;
    size_t m = (size_t)3 * npart;
    for (int w = 0; w < nw; w++) {
        double wt = sc[6 * w] * factor;
        int64_t s = irn[w];
        double u = dmco_ran1(&s);
        int iwt = (int)(wt + u);
        for (int k = 0; k < iwt; k++) {
            dmco_ran2(&s);
            if (nt >= cap) return -1; /* stack overflow (stack.f90:88-91) */
            memcpy(x_out + m * nt, x + m * w, sizeof(double) * m);
            memcpy(dpsi_out + m * nt, dpsi + m * w, sizeof(double) * m);
            memcpy(sc_out + 6 * nt, sc + 6 * w, sizeof(double) * 6);
            sc_out[6 * nt] = copysign(1.0, sc[6 * w]);
            irn_out[nt] = s;
            nt++;
        }
    }
    return nt;
}

typedef void (*step1_fn)(const dmco_ctx *, int32_t, const double *, const double *,
                         const double *, const int64_t *, int32_t, double *, double *, double *,
                         double *, int64_t *, double *, int32_t);

/* dmcbcs.f90:437-445 with the step1 of idmc = 1 (dmco_dmc_step) or idmc = 2 (dmco_mdmc_step) */
static int dmc_step_with(step1_fn step1, const dmco_ctx *c, int32_t nw, const double *x,
                         const double *dpsi, const double *sc, const int64_t *irn,
                         int32_t opcomp, double *est, int32_t cap, double *x_out,
                         double *dpsi_out, double *sc_out, int64_t *irn_out, int32_t *nt,
                         int32_t nthreads);

int dmco_dmc_step(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                  const double *sc, const int64_t *irn, int32_t opcomp, double *est,
                  int32_t cap, double *x_out, double *dpsi_out, double *sc_out,
                  int64_t *irn_out, int32_t *nt, int32_t nthreads) {
    return dmc_step_with(dmco_step1, c, nw, x, dpsi, sc, irn, opcomp, est, cap, x_out,
                         dpsi_out, sc_out, irn_out, nt, nthreads);
}

int dmco_mdmc_step(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                   const double *sc, const int64_t *irn, int32_t opcomp, double *est,
                   int32_t cap, double *x_out, double *dpsi_out, double *sc_out,
                   int64_t *irn_out, int32_t *nt, int32_t nthreads) {
    return dmc_step_with(dmco_mdmc_step1, c, nw, x, dpsi, sc, irn, opcomp, est, cap, x_out,
                         dpsi_out, sc_out, irn_out, nt, nthreads);
}

static int dmc_step_with(step1_fn step1, const dmco_ctx *c, int32_t nw, const double *x,
                         const double *dpsi, const double *sc, const int64_t *irn,
                         int32_t opcomp, double *est, int32_t cap, double *x_out,
                         double *dpsi_out, double *sc_out, int64_t *irn_out, int32_t *nt,
                         int32_t nthreads) {
    int np = c->npart;
    size_t m = (size_t)3 * np;
    double *x1 = malloc(sizeof(double) * m * nw), *d1 = malloc(sizeof(double) * m * nw);
    double *s1 = malloc(sizeof(double) * 6 * nw);
    int64_t *r1 = malloc(sizeof(int64_t) * nw);
    step1(c, nw, x, dpsi, sc, irn, opcomp, est, x1, d1, s1, r1, NULL, nthreads);
    int n1 = dmco_branch(nw, np, x1, d1, s1, r1, 1.0, cap, x_out, dpsi_out, sc_out, irn_out);
    free(x1); free(d1); free(s1); free(r1);
    if (n1 < 0) return -1;
    nt[0] = n1;
    nt[1] = n1;
    int nwalk = c->p.nwalk;
    if (n1 > c->p.popup * nwalk || n1 < c->p.popdn * nwalk) {
        double *x2 = malloc(sizeof(double) * m * n1), *d2 = malloc(sizeof(double) * m * n1);
        double *s2 = malloc(sizeof(double) * 6 * n1);
        int64_t *r2 = malloc(sizeof(int64_t) * n1);
        memcpy(x2, x_out, sizeof(double) * m * n1);
        memcpy(d2, dpsi_out, sizeof(double) * m * n1);
        memcpy(s2, sc_out, sizeof(double) * 6 * n1);
        memcpy(r2, irn_out, sizeof(int64_t) * n1);
        int n2 = dmco_branch(n1, np, x2, d2, s2, r2, 1.0 * nwalk / n1, cap, x_out, dpsi_out,
                             sc_out, irn_out);
        free(x2); free(d2); free(s2); free(r2);
        if (n2 < 0) return -1;
        nt[1] = n2;
    }
    return 0;
}

/* ---------------------------------------------------------------- single-particle moves */
/* gradient of ln det w.r.t. particle ip at position xp, through inverse column/row:
 * up: sum_j sinv(j,ip) grad phi(xp - r_dn_j); down: -sum_i sinv(jd,i) grad phi(r_up_i - xp) */
static void det_grad(const dmco_ctx *c, const double *x, const double *sinv, int ip,
                     const double *xp, double scale, double g[3]) {
    int n = c->n;
    double f, gg[3], l;
    g[0] = g[1] = g[2] = 0.0;
    if (ip < n) {
        for (int j = 0; j < n; j++) {
            pairphi(c, xp, x + 3 * (n + j), &f, gg, &l);
            for (int d = 0; d < 3; d++) g[d] += sinv[j + n * ip] * gg[d];
        }
    } else {
        int jd = ip - n;
        for (int i = 0; i < n; i++) {
            pairphi(c, x + 3 * i, xp, &f, gg, &l);
#if DMCO_PLANT == 1
            for (int d = 0; d < 3; d++) g[d] += sinv[jd + n * i] * gg[d];
#else
            for (int d = 0; d < 3; d++) g[d] -= sinv[jd + n * i] * gg[d];
#endif
        }
    }
    for (int d = 0; d < 3; d++) g[d] *= scale;
}
/* Jastrow gradient of particle ip at position xp (coshpot.f90:195-196 per pair) */
static void jas_grad(const dmco_ctx *c, const double *x, int ip, const double *xp, double g[3]) {
    g[0] = g[1] = g[2] = 0.0;
    for (int j = 0; j < c->npart; j++) {
        if (j == ip || !active_pair(c, ip, j)) continue;
        int itab = ((ip < c->n) != (j < c->n)) ? 1 : 2;
        double dx[3], r2 = 0.0;
        for (int d = 0; d < 3; d++) {
            dx[d] = wrap(xp[d] - x[3 * j + d], c->el, c->eli);
            r2 += dx[d] * dx[d];
        }
        double r = fmax(1.0e-14, sqrt(r2)), vp, uj, duj, d2uj;
        if (!pairtab(c, r, itab, &vp, &uj, &duj, &d2uj)) continue;
        for (int d = 0; d < 3; d++) g[d] += dx[d] * duj / r;
    }
}
/* Sherman-Morrison update of sinv after particle ip moved to xp with det ratio rho */
static void sm_update(const dmco_ctx *c, const double *x, double *sinv, int ip,
                      const double *xp, double rho) {
    int n = c->n;
    double f, gg[3], l, snew[128], z[128], col[128];
    if (ip < n) {
        for (int j = 0; j < n; j++) { pairphi(c, xp, x + 3 * (n + j), &f, gg, &l); snew[j] = f; }
        for (int k = 0; k < n; k++) {
            double acc = 0.0;
            for (int j = 0; j < n; j++) acc += snew[j] * sinv[j + n * k];
            z[k] = acc;
        }
        for (int j = 0; j < n; j++) col[j] = sinv[j + n * ip];
        for (int k = 0; k < n; k++) {
            double fk = (z[k] - (k == ip ? 1.0 : 0.0)) / rho;
            for (int j = 0; j < n; j++) sinv[j + n * k] -= col[j] * fk;
        }
    } else {
        int jd = ip - n;
        for (int i = 0; i < n; i++) { pairphi(c, x + 3 * i, xp, &f, gg, &l); snew[i] = f; }
        for (int m = 0; m < n; m++) {
            double acc = 0.0;
            for (int i = 0; i < n; i++) acc += sinv[m + n * i] * snew[i];
            z[m] = acc;
        }
        for (int i = 0; i < n; i++) col[i] = sinv[jd + n * i];
        for (int m = 0; m < n; m++) {
            double fm = (z[m] - (m == jd ? 1.0 : 0.0)) / rho;
            for (int i = 0; i < n; i++) sinv[m + n * i] -= fm * col[i];
        }
    }
}

static void sm_one(const dmco_ctx *c, const double *x, const double *dpsi, const double *sc,
                   int64_t irn, int kref, double *xo, double *dpsio, double *sco,
                   int64_t *irno, double *wto, int32_t *nacc, double *tr) {
    int np = c->npart, n = c->n;
    double el = c->el, hbar = c->p.hbar, sig = c->sigma, sig2 = sig * sig, sm = sqrt(c->p.mass);
    double *sinv = malloc(sizeof(double) * (n * n + 3 * np)), *dphi = sinv + n * n;
    double phil, d2phi;
    int32_t is;
    for (int i = 0; i < 3 * np; i++) xo[i] = wrapd(x[i], el);
    dmco_bcs(c, xo, &phil, &is, dphi, &d2phi, sinv);
    int64_t s = irn;
    int acc = 0;
    for (int ip = 0; ip < np; ip++) {
        if (ip > 0 && ip % kref == 0) dmco_bcs(c, xo, &phil, &is, dphi, &d2phi, sinv);
        double *xi = xo + 3 * ip;
        double gd[3], gj[3], go[3], gn[3], g3[3], xn[3], dx[3], dro[3], drn[3];
        double *t = tr ? tr + (size_t)DMCO_SM_TRACE * ip : NULL;
        det_grad(c, xo, sinv, ip, xi, 1.0, gd);
        jas_grad(c, xo, ip, xi, gj);
        for (int d = 0; d < 3; d++) go[d] = gd[d] + gj[d];
        dmco_gaussian(&s, 3, g3);
        for (int d = 0; d < 3; d++) {
            /* drift = dpsi * sigma^2 * sqrt(m) with dpsi = grad ln psi / sqrt(m), cut to
             * +-sigma*dcut when dcut >= 0 (step.f90:66-67) */
            dro[d] = go[d] * sig2;
            if (c->p.dcut >= 0.0) dro[d] = fmin(sig * c->p.dcut, fmax(-sig * c->p.dcut, dro[d]));
            dx[d] = dro[d] + sig * g3[d];
            xn[d] = wrapd(xi[d] + dx[d], el);
        }
        double rho_d = det_ratio(c, xo, sinv, ip, xn);
        double rho_j = jastrow_ratio(c, xo, ip, xn);
        double u = dmco_ran1(&s);
        if (t) {
            t[0] = ip;
            for (int d = 0; d < 3; d++) {
                t[1 + d] = xi[d];
                t[4 + d] = go[d];
                t[7 + d] = dro[d];
                t[10 + d] = dx[d];
                t[13 + d] = xn[d];
                t[21 + d] = NAN;
            }
            t[16] = rho_d;
            t[17] = rho_j;
            t[18] = u;
            t[19] = NAN;
            t[20] = 0.0;
        }
        if (c->p.fn && !(rho_d > 0.0)) continue; /* node crossing: reject */
        det_grad(c, xo, sinv, ip, xn, 1.0 / rho_d, gd);
        jas_grad(c, xo, ip, xn, gj);
#if DMCO_PLANT == 3
        double logp = 2.0 * log(fabs(rho_d)) + log(rho_j);
#else
        double logp = 2.0 * (log(fabs(rho_d)) + log(rho_j));
#endif
        for (int d = 0; d < 3; d++) {
            gn[d] = gd[d] + gj[d];
            drn[d] = gn[d] * sig2;
            if (c->p.dcut >= 0.0) drn[d] = fmin(sig * c->p.dcut, fmax(-sig * c->p.dcut, drn[d]));
#if DMCO_PLANT == 2
            logp += (dro[d] + drn[d]) * (dro[d] - drn[d] - 2.0 * dx[d]) / sig2;
#else
            logp += (dro[d] + drn[d]) * (dro[d] - drn[d] - 2.0 * dx[d]) / (2.0 * sig2);
#endif
        }
        double prob = exp(fmin(0.0, logp));
        const int ok = u < prob;
        if (t) {
            t[19] = logp;
            t[20] = ok;
            for (int d = 0; d < 3; d++) t[21 + d] = gn[d];
        }
        if (ok) {
            sm_update(c, xo, sinv, ip, xn, rho_d);
            for (int d = 0; d < 3; d++) xi[d] = xn[d];
            acc++;
        }
    }
    (void)sm;
    double o[4];
    dmco_hpsi(c, xo, o, dpsio);
    double en = -hbar * o[1] + o[2], eo = -hbar * sc[2] + sc[3];
    double wt = exp(-(0.5 * (eo + en) - c->etrial) * c->p.dt);
    double weight = fmin(c->p.wtcut, wt);
    if ((int)o[3] != (int)sc[4] && c->p.fn) { weight = 0.0; wt = 0.0; }
    sco[0] = weight; sco[1] = o[0]; sco[2] = o[1]; sco[3] = o[2]; sco[4] = o[3]; sco[5] = sc[5];
    *irno = s;
    *wto = wt;
    *nacc = acc;
    free(sinv);
}

void dmco_sm_step1(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                   const double *sc, const int64_t *irn, int32_t krefresh, double *x_out,
                   double *dpsi_out, double *sc_out, int64_t *irn_out, double *wt_out,
                   int32_t *nacc_out, int32_t nthreads) {
    dmco_sm_step1_trace(c, nw, x, dpsi, sc, irn, krefresh, x_out, dpsi_out, sc_out, irn_out,
                        wt_out, nacc_out, NULL, nthreads);
}

void dmco_sm_step1_trace(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                         const double *sc, const int64_t *irn, int32_t krefresh, double *x_out,
                         double *dpsi_out, double *sc_out, int64_t *irn_out, double *wt_out,
                         int32_t *nacc_out, double *trace, int32_t nthreads) {
    int np = c->npart;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 2)
    for (int w = 0; w < nw; w++) {
        size_t o = (size_t)3 * np * w;
        sm_one(c, x + o, dpsi + o, sc + 6 * w, irn[w], krefresh, x_out + o, dpsi_out + o,
               sc_out + 6 * w, irn_out + w, wt_out + w, nacc_out + w,
               trace ? trace + (size_t)DMCO_SM_TRACE * np * w : NULL);
    }
}

/* ---------------------------------------------------------------- observables */
/* gofr.f90:52-98 addgofr with ndim = 3: pair-distance histogram by spin channel and the
 * one-body density profile along each axis, of one configuration x[npart][3] (up spins
 * 0..nup-1).  Layouts follow the reference's column-major arrays: gnow/gerr [ntab+1][3]
 * (channel 0 up-down, 1 up-up, 2 down-down), dnow/derr [2 ntab+1][6] (bin b is index
 * b - ntab; channels x, y, z of spin up, then of spin down).  dr = L/(2 ntab)
 * (gofr.f90:35-36). */
void dmco_addgofr(int32_t npart, int32_t nup, double el, int32_t ntab, const double *x,
                  double weight, double *gnow, double *gerr, double *dnow, double *derr,
                  double *wsblk) {
    double dr = 0.5 * el / ntab, scale = 1.0 / dr;
    for (int i = 0; i < npart - 1; i++)
        for (int j = i + 1; j < npart; j++) {
            int itab = 1;
            if (i < nup && j < nup) itab = 2;
            if (i >= nup && j >= nup) itab = 3;
            double dx[3], r2 = 0.0;
            for (int d = 0; d < 3; d++) {
                dx[d] = x[3 * i + d] - x[3 * j + d];
                dx[d] = dx[d] - el * round(dx[d] / el);
            }
            for (int d = 0; d < 3; d++) r2 += dx[d] * dx[d];
            double r = sqrt(r2);
            int index = r < 0.5 * el ? (int)(scale * r) : ntab;
            gnow[3 * index + itab - 1] += 2.0 * weight;
            gerr[3 * index + itab - 1] += 4.0 * weight;
        }
    *wsblk += weight;
    for (int i = 0; i < npart; i++) {
        for (int is = 0; is < 3; is++) {
            double dx = x[3 * i + is] - 0.0;
            dx = dx - el * round(dx / el);
            int itab = i < nup ? is : is + 3;
            int index = (int)(scale * dx);
            int e = 6 * (index + ntab) + itab;
            if (index == 0) {
                dnow[e] += 0.5 * weight;
                derr[e] += 0.25 * weight;
            } else {
                dnow[e] += weight;
                derr[e] += weight;
            }
        }
    }
}

/* sofq.f90:40-65 addsofq: rho_up(q), rho_dn(q) = sum exp(-i q.x) over each spin, and
 * snow/serr [nq][3] += (|rho_up|^2, |rho_dn|^2, Re rho_up conj(rho_dn)) * weight and their
 * squares (real parts) * weight; ak[nq][3] from setupk. */
void dmco_addsofq(int32_t npart, int32_t nup, int32_t nq, const double *ak, const double *x,
                  double weight, double *snow, double *serr, double *wsblk) {
    for (int iq = 0; iq < nq; iq++) {
        double ur = 0.0, ui = 0.0, dr_ = 0.0, di = 0.0;
        for (int i = 0; i < npart; i++) {
            double arg = ak[3 * iq] * x[3 * i] + ak[3 * iq + 1] * x[3 * i + 1] +
                         ak[3 * iq + 2] * x[3 * i + 2];
            double c = cos(arg), s = -sin(arg);
            if (i < nup) { ur += c; ui += s; } else { dr_ += c; di += s; }
        }
        /* z1 = rho_up conj(rho_up), z2 = rho_dn conj(rho_dn), z3 = rho_up conj(rho_dn) */
        double z1r = ur * ur + ui * ui, z1i = ui * ur - ur * ui;
        double z2r = dr_ * dr_ + di * di, z2i = di * dr_ - dr_ * di;
        double z3r = ur * dr_ + ui * di, z3i = ui * dr_ - ur * di;
        snow[3 * iq] += z1r * weight;
        serr[3 * iq] += (z1r * z1r - z1i * z1i) * weight;
        snow[3 * iq + 1] += z2r * weight;
        serr[3 * iq + 1] += (z2r * z2r - z2i * z2i) * weight;
        snow[3 * iq + 2] += z3r * weight;
        serr[3 * iq + 2] += (z3r * z3r - z3i * z3i) * weight;
    }
    *wsblk += weight;
}

/* momentum3d.f90:13-79 setupenk with ndim = 3: the octant vectors (ix, iy, iz >= 0,
 * |m|^2 <= ntab) in the reference's enumeration order, put in |m|^2 order by its exchange
 * pass, and grouped in shells of equal |m|^2: ik[ntot][3], indk[nsh] (first vector of each
 * shell), nk[nsh].  Returns nsh, or -1 when more than maxk vectors. */
int dmco_enk_shells(int32_t ntab, int32_t maxk, int32_t *ik, int32_t *indk, int32_t *nk,
                    int32_t *ntot_out) {
    int nnk = (int)sqrt((double)ntab) + 1, ntot = 0;
    for (int ix = 0; ix <= nnk; ix++)
        for (int iy = 0; iy <= nnk; iy++)
            for (int iz = 0; iz <= nnk; iz++) {
                if (ix * ix + iy * iy + iz * iz > ntab) continue;
                if (ntot >= maxk) return -1;
                ik[3 * ntot] = ix; ik[3 * ntot + 1] = iy; ik[3 * ntot + 2] = iz;
                ntot++;
            }
    int *k2 = malloc(sizeof(int) * (ntot > 0 ? ntot : 1));
    for (int i = 0; i < ntot; i++)
        k2[i] = ik[3 * i] * ik[3 * i] + ik[3 * i + 1] * ik[3 * i + 1] + ik[3 * i + 2] * ik[3 * i + 2];
    for (int i = 1; i < ntot; i++)
        for (int j = 0; j < i; j++)
            if (k2[i] < k2[j]) {
                int t = k2[i]; k2[i] = k2[j]; k2[j] = t;
                for (int d = 0; d < 3; d++) { t = ik[3 * i + d]; ik[3 * i + d] = ik[3 * j + d]; ik[3 * j + d] = t; }
            }
    int itab = 0;
    indk[0] = 0;
    nk[0] = 1;
    for (int i = 1; i < ntot; i++) {
        if (k2[i] != k2[i - 1]) {
            itab++;
            indk[itab] = i;
            nk[itab] = 0;
        }
        nk[itab]++;
    }
    free(k2);
    *ntot_out = ntot;
    return itab + 1;
}

/* momentum3d.f90:89-120 addenk: for every particle i and nsamp = 10 samples, a displacement
 * dx = (u - 1/2) L from three ran1 draws of the walker's stream (set from irn and not saved
 * back), the ratio psirat of move1 (the inverse of hpsi, wavefunction.f90:393-404), and
 * enktmp(shell) += psirat * prod_d cos(dk m_d dx_d) over the shell's vectors; then
 * enknow += enktmp/nk * wt, enkerr += (enktmp/nk)^2 * wt, wsblk += wt * nsamp. */
void dmco_addenk(const dmco_ctx *c, int32_t nsh, const int32_t *ik, const int32_t *indk,
                 const int32_t *nk, const double *x, int64_t irn, double wt, double *enknow,
                 double *enkerr, double *wsblk) {
    const int nsamp = 10, n = c->n, np = c->npart;
    double dk = 2.0 * (4.0 * atan(1.0)) / c->el;
    double *sinv = malloc(sizeof(double) * (n * n + 3 * np)), *dphi = sinv + n * n;
    double *tmp = calloc(nsh, sizeof(double));
    double phil, d2phi;
    int32_t is;
    dmco_bcs(c, x, &phil, &is, dphi, &d2phi, sinv);
    int64_t s = irn & (int64_t)MASK48;
    for (int i = 0; i < np; i++)
        for (int ii = 0; ii < nsamp; ii++) {
            double dx[3], xnew[3];
            for (int d = 0; d < 3; d++) {
                dx[d] = (dmco_ran1(&s) - 0.5) * c->el;
                xnew[d] = x[3 * i + d] + dx[d];
            }
            double psirat = jastrow_ratio(c, x, i, xnew) * det_ratio(c, x, sinv, i, xnew);
            for (int k = 0; k < nsh; k++)
                for (int j = 0; j < nk[k]; j++) {
                    const int32_t *m = ik + 3 * (indk[k] + j);
                    tmp[k] += psirat * cos(dk * m[0] * dx[0]) * cos(dk * m[1] * dx[1]) *
                              cos(dk * m[2] * dx[2]);
                }
        }
    for (int k = 0; k < nsh; k++) {
        enknow[k] += tmp[k] / nk[k] * wt;
        enkerr[k] += (tmp[k] / nk[k]) * (tmp[k] / nk[k]) * wt;
    }
    *wsblk += wt * nsamp;
    free(tmp);
    free(sinv);
}
