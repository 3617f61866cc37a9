/* CPU restatement of the DMCBS walker-propagation path (test infrastructure).
 *
 * This library is the oracle: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, always as the checker.  Every function follows the reference
 * routine cited beside it; the golden fixtures written from the reference itself
 * (tests/golden/*.npz) pin it.
 *
 * Conventions: positions x are [npart][3] (particle-major, x,y,z), up spins first, the
 * same memory order as the reference's x(3,npart).  FP64 throughout.
 */
#ifndef DMC_ORACLE_H
#define DMC_ORACLE_H
#include <stdint.h>

#define DMCO_MAXORB 16

typedef struct dmco_params {
    int32_t npart;      /* total particles; nup = ndn = npart/2 */
    int32_t norb;       /* number of s-wave k-shells (alpha_1..alpha_norb) */
    int32_t ntab;       /* table points */
    int32_t fn;         /* fixed node */
    int32_t units;      /* energies in E_F units */
    int32_t nwalk;      /* target population */
    double rho, hbar, mass, dt, etrial, dcut, wtcut, popup, popdn;
    double bbeta, dbeta, bfact, v0, amu, cjas1, cjas2, djas, v0op, v0eq;
    double vorb[DMCO_MAXORB];
} dmco_params;

typedef struct dmco_ctx dmco_ctx;

/* set-up: dmcbcs.f90:142-177 (setpsi, setphi, setstep) */
dmco_ctx *dmco_create(const dmco_params *p);
void dmco_destroy(dmco_ctx *c);
/* el, eunit, etrial(internal), nfac, nk, ntab0, sigma, scale */
void dmco_info(const dmco_ctx *c, double out[8]);
/* k-vectors: ak[nk][3], ak2[nk], numsh[norb] (kshell.f90:20-156) */
int dmco_kvec(const dmco_ctx *c, double *ak, double *ak2, int32_t *numsh);
/* tables [0..ntab]: v, u, du, d2u (coshpot.f90:77-133) and b, db, d2b (betafun.f90:74-125) */
void dmco_tables(const dmco_ctx *c, double *vt, double *ut, double *dut, double *d2ut,
                 double *bt, double *dbt, double *d2bt);
void dmco_radial(const dmco_ctx *c, double r, double out[3]);   /* betafun.f90:173-194 */
void dmco_beta(const dmco_ctx *c, double r, double out[3]);     /* betafun.f90:151-171 */
void dmco_derbeta(const dmco_ctx *c, double r, double out[2]);  /* betafun.f90:196-327 */
double dmco_vpot(const dmco_ctx *c, double r);                  /* coshpot.f90:145-152 */

/* wavefunction pieces on one configuration */
void dmco_getphi(const dmco_ctx *c, const double *x, double *u, double *du, double *d2u,
                 double *v);                                     /* coshpot.f90:154-205 */
void dmco_bcs(const dmco_ctx *c, const double *x, double *phil, int32_t *is, double *dphi,
              double *d2phi, double *sinv);                      /* psislaterbcs.f90:155-225 */
/* hpsi: out = {psil, d2psi, v, is}; dpsi[npart][3]  (wavefunction.f90:161-201, a2 = 0) */
void dmco_hpsi(const dmco_ctx *c, const double *x, double *out, double *dpsi);
/* psi(x with particle ip moved to xnew)/psi(x), ip 0-based (wavefunction.f90:393-404) */
double dmco_ratio1(const dmco_ctx *c, const double *x, int32_t ip, const double *xnew);
/* d log psi / d alpha_s, s = 1..norb (psislaterbcs.f90:449-524, optsw, no optbeta) */
void dmco_getder(const dmco_ctx *c, const double *x, int32_t optbeta, double *dlogpsi);
/* set alpha_1..alpha_norb and (b, d) of beta; rebuilds the beta table (setbcspar) */
void dmco_set_params(dmco_ctx *c, const double *vorb, double bbeta, double dbeta);

/* random numbers (ran.f90) */
double dmco_ran1(int64_t *s);
double dmco_ran2(int64_t *s);
void dmco_gaussian(int64_t *s, int32_t n, double *g);

/* initial walkers from one uniform configuration (dmcbcs.f90:237-301) */
void dmco_init_walkers(const dmco_ctx *c, int64_t seed, int32_t nwalk, double *x,
                       int64_t *irn);

/* Walker population.  sc[w][6] = weight, psil, d2psi, v, is, wt0 (dump order of the
 * reference driver).  est[14][3] accumulates (sum val*wt, sum val^2*wt, sum wt). */
/* initial hpsi pass (dmcbcs.f90:314-345): fills dpsi, sc */
void dmco_init_hpsi(const dmco_ctx *c, int32_t nw, const double *x, double *dpsi,
                    double *sc);
/* one DMC step: step1 (idmc=1) + branch(1) + population control (dmcbcs.f90:437-445).
 * Writes at most cap walkers to the *_out arrays; nt[0] = after branch, nt[1] = after
 * population control.  Returns 0, or -1 on capacity overflow. */
int dmco_dmc_step(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                  const double *sc, const int64_t *irn, int32_t opcomp, double *est,
                  int32_t cap, double *x_out, double *dpsi_out, double *sc_out,
                  int64_t *irn_out, int32_t *nt, int32_t nthreads);
/* the same with step1 of idmc = 2, Metropolis DMC (step.f90:161-226); est holds
 * estimators 1-13 of dmcbcs.f90:204-218 */
int dmco_mdmc_step(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                   const double *sc, const int64_t *irn, int32_t opcomp, double *est,
                   int32_t cap, double *x_out, double *dpsi_out, double *sc_out,
                   int64_t *irn_out, int32_t *nt, int32_t nthreads);
void dmco_mdmc_step1(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                     const double *sc, const int64_t *irn, int32_t opcomp, double *est,
                     double *x_out, double *dpsi_out, double *sc_out, int64_t *irn_out,
                     double *wt_out, int32_t nthreads);
/* the step1 part alone for every walker (no branching): new state + per-walker
 * uncapped weight wt (0 after a node crossing with fn) */
void dmco_step1(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                const double *sc, const int64_t *irn, int32_t opcomp, double *est,
                double *x_out, double *dpsi_out, double *sc_out, int64_t *irn_out,
                double *wt_out, int32_t nthreads);
/* branch with factor: returns new count (or -1 if > cap) (step.f90:287-315) */
void dmco_vmc_step1(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                    const double *sc, const int64_t *irn, double *est, double *x_out,
                    double *dpsi_out, double *sc_out, int64_t *irn_out, int32_t nthreads);
int dmco_branch(int32_t nw, int32_t npart, const double *x, const double *dpsi,
                const double *sc, const int64_t *irn, double factor, int32_t cap,
                double *x_out, double *dpsi_out, double *sc_out, int64_t *irn_out);

/* Single-particle-move DMC sweep with Sherman-Morrison inverse updates (the algorithm of
 * the device's single-particle path; no reference counterpart).  One step of one walker: particles moved in order
 * 0..npart-1; the inverse is rebuilt from scratch every krefresh accepted-or-rejected moves
 * and at the end of the sweep.  st[w] state as for the population arrays.  Outputs the
 * new state, the uncapped weight and the number of accepted moves. */
void dmco_sm_step1(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                   const double *sc, const int64_t *irn, int32_t krefresh, double *x_out,
                   double *dpsi_out, double *sc_out, int64_t *irn_out, double *wt_out,
                   int32_t *nacc_out, int32_t nthreads);
/* The same sweep with a record of every move: trace[w][ip][DMCO_SM_TRACE] =
 *   0 ip, 1-3 the mover's position before the move, 4-6 grad_ip ln psi there (determinant
 *   through the carried inverse + Jastrow), 7-9 the drift after the cut, 10-12 the
 *   displacement (drift + sigma * gaussian), 13-15 the proposed position (wrapped),
 *   16 the determinant ratio, 17 the Jastrow ratio, 18 the acceptance draw, 19 the log of
 *   the Metropolis ratio (NaN for a move refused by the node test), 20 accepted (0/1),
 *   21-23 grad_ip ln psi at the proposed position (NaN for a refused move). */
#define DMCO_SM_TRACE 24
void dmco_sm_step1_trace(const dmco_ctx *c, int32_t nw, const double *x, const double *dpsi,
                         const double *sc, const int64_t *irn, int32_t krefresh, double *x_out,
                         double *dpsi_out, double *sc_out, int64_t *irn_out, double *wt_out,
                         int32_t *nacc_out, double *trace, int32_t nthreads);

/* Observables (the operators of operators.f90:92-108) on one configuration x[npart][3]. */
/* setupk with ndim = 3 (kshell.f90:6-156) for nsh0 shells of box side el: numsh[nsh0],
 * ak[k][3]; returns the number of vectors, -1 when nsh0 exceeds the enumerated shells,
 * -2 when more than maxk vectors */
int dmco_kshells(double el, int32_t nsh0, int32_t *numsh, double *ak, int32_t maxk);
/* addgofr (gofr.f90:52-98): gnow/gerr [ntab+1][3], dnow/derr [2 ntab+1][6] */
void dmco_addgofr(int32_t npart, int32_t nup, double el, int32_t ntab, const double *x,
                  double weight, double *gnow, double *gerr, double *dnow, double *derr,
                  double *wsblk);
/* addsofq (sofq.f90:40-65): snow/serr [nq][3] */
void dmco_addsofq(int32_t npart, int32_t nup, int32_t nq, const double *ak, const double *x,
                  double weight, double *snow, double *serr, double *wsblk);
/* setupenk (momentum3d.f90:13-79): octant vectors ik[ntot][3] in shells indk[nsh], nk[nsh];
 * returns nsh (-1: more than maxk vectors) */
int dmco_enk_shells(int32_t ntab, int32_t maxk, int32_t *ik, int32_t *indk, int32_t *nk,
                    int32_t *ntot);
/* addenk (momentum3d.f90:89-120): enknow/enkerr [nsh] */
void dmco_addenk(const dmco_ctx *c, int32_t nsh, const int32_t *ik, const int32_t *indk,
                 const int32_t *nk, const double *x, int64_t irn, double wt, double *enknow,
                 double *enkerr, double *wsblk);
#endif
