// Kernel entry points and shared device constants.
#pragma once
#include "dmc_device.h"

namespace dmc {

constexpr uint64_t kMask48 = (1ULL << 48) - 1ULL;
constexpr uint64_t kM1 = 44485709377909ULL, kA1 = 96309754297ULL;  // ran1 (ran.f90:15,18)
constexpr uint64_t kM2 = 34522712143931ULL, kA2 = 55789347517ULL;  // ran2 (ran.f90:36,39)
constexpr double kPi = 3.14159265358979323846;

// Host launchers.  Each returns a hipError_t value.
int launch_eval(int nup, const DevSys &S, const double *x, int nw, double *psil, int *is,
                double *dpsi, double *d2psi, double *v, char *sinv, hipStream_t st);
int launch_ratio(int nup, const DevSys &S, const double *x, const int *ip, const double *xnew,
                 int nw, double *ratio, hipStream_t st);
int launch_step(int nup, const DevSys &S, const DevState &A, const DevState &B,
                const DevStep &T, int opcomp, int grid, hipStream_t st);
// variational (Metropolis) step of state A into state B, no branching
int launch_vmc(int nup, const DevSys &S, const DevState &A, const DevState &B, const DevStep &T,
               int grid, hipStream_t st);
// Metropolis DMC step (step1 with idmc = 2) of state A into state B with the branch draw
int launch_mdmc(int nup, const DevSys &S, const DevState &A, const DevState &B,
                const DevStep &T, int opcomp, int grid, hipStream_t st);
// single-particle moves m0..m1-1 of every walker, from positions/RNG xin/irn_in and the
// stored inverse sinv, to xout/irn_out (move kernel)
int launch_sm_moves(int nup, const DevSys &S, const double *xin, double *xout,
                    const uint64_t *irn_in, uint64_t *irn_out, char *sinv, int m0, int m1,
                    const DevStep &T, int grid, hipStream_t st);
// FP64 rebuild of the inverse of the positions x into sinv (refresh kernel)
int launch_sm_refresh(int nup, const DevSys &S, const double *x, char *sinv, const DevStep &T,
                      int grid, hipStream_t st);
// end of the sweep: FP64 hpsi of B.x, inverse into B.sinv, weights against A, branch draw
int launch_sm_end(int nup, const DevSys &S, const DevState &A, const DevState &B,
                  const DevStep &T, int opcomp, int grid, hipStream_t st);
// 1 if a workgroup of the end kernel that stages the radial tables in LDS fits the device
int sm_end_tab_fits(int nup, const DevSys &S);
// 1 if a workgroup of the two-walker move kernel that stages its pair records in LDS fits
int sm_moves_tab_fits(int nup, const DevSys &S);
// parameter adjoint of nw configurations: d log psi and d E_L per parameter, [nw][norb + 2]
int launch_adjoint(int nup, const DevSys &S, const double *x, int nw, int norb, double *dlogpsi,
                   double *delocal, hipStream_t st);
// branching: multiplicities of state A with factor (population control), prefix sums,
// the destination map and the gather into state B
int launch_branch_count(const DevSys &S, const DevState &A, const DevStep &T, double factor,
                        int grid, hipStream_t st);
int launch_scan(const DevStep &T, hipStream_t st);
int launch_map_gather(const DevSys &S, const DevState &src, const DevState &dst,
                      const DevStep &T, int nsrc, int ndst, int capacity, hipStream_t st);
// grid < 0: the count in nt[0]
int launch_est_reduce(const DevStep &T, int grid, double *acc, hipStream_t st);
// branch and population control with every size on the device (graph-capturable): the
// population enters in B with its multiplicities and leaves in A; nt[2] = population
int launch_branch_device(const DevSys &S, const DevState &A, const DevState &B, const DevStep &T,
                         int capacity, hipStream_t st);
// optimiser sums (calder) of the walkers of A with their parameter derivatives O [nw][P]
int launch_opt_accumulate(const DevSys &S, const DevState &A, const double *O, int nw, int P,
                          double *acc, hipStream_t st);
// nt[0] = nt[2] at the start of a device-driven step
int launch_pop_begin(const DevStep &T, hipStream_t st);

bool nup_supported(int nup);

// Observables of a walker population (addgofr, addsofq): per-block partial sums
// part[nblk][nsum], entry 0 the summed weight, then (ntab > 0) the pair counts times weight
// [ntab+1][3] and the density counts times weight [2 ntab+1][6], then (nq > 0)
// [nq][6] = |rho_up|^2, |rho_dn|^2, Re rho_up conj(rho_dn) and their squares (real parts),
// each times weight.
struct ObsArgs {
    const double *x, *weight, *wt0;
    const int *is;
    int nw, npart, nup;
    double el, eli, half, scale;  // box side, 1/L, L/2, 1/dr with dr = L/(2 ntab)
    int ntab;                // 0: no g(r) and density
    const int4 *qm;          // S(q) vectors in units of 2 pi / L
    int nq, mmax;            // 0: no S(q); largest |component| of the vectors
    double twopi_l;
    int nsum, nblk;
    int part_global;  // partial sums in part (each block its row) instead of LDS
    double *part;
};
constexpr int kObsBlocks = 1024;
size_t obs_lds_bytes(int npart, int ntab, int nq, int mmax, bool part_global);
int obs_blocks(int nw);
// k_obs into a.part, then a fixed-order reduction over the blocks: tmp holds
// ceil(nblk/32) rows of nsum, out one row
int launch_obs(const ObsArgs &a, double *tmp, double *out, hipStream_t st);
// out[e] = sum over the rows r of in[r][e] in row order, in two passes: segments of seg
// rows into tmp [ceil(nrow/seg)][ncol], then the segments
int launch_rows_reduce(const double *in, int nrow, int ncol, int seg, double *tmp, double *out,
                       hipStream_t st);

// n(k) of walkers w0 .. w0 + nw - 1 of a population (addenk): one row per walker
// rows[w][1 + 2 nsh] = (wt nsamp, enktmp/nk wt [nsh], (enktmp/nk)^2 wt [nsh]).  Three
// launches: per walker the FP64 inverse, the projections of the partner phases, the partner
// coefficients and the old Jastrow sums into per-walker buffers; then one wave per 64
// displacements of a walker (the shell sums of each wave into part); then per walker the
// sums over its waves and its row.
struct NofkArgs {
    const double *x, *weight, *wt0;
    const int *is;
    const uint64_t *irn;
    int w0, nw, nsamp, nsh, nchunk;
    const int4 *ik;          // octant vectors in shell order
    const int2 *shell;       // (first vector, count) of each shell
    const uint64_t *ja, *jc; // ran1 state after 3 m draws = ja[m] s + jc[m], m < npart nsamp
    double dk;               // 2 pi / L
    double2 *proj;           // [nw][npart][nq] sum_j Sinv exp(i q.r_j) over the partners of i
    double *coef;            // [nw][npart][npart/2] Sinv of i with each partner
    double *uold;            // [nw][npart] Jastrow sum of i at its position
    int *sing;               // [nw] singular pairing matrix
    double *part;            // [nw][nchunk][nsh] shell sums of each wave
    double *rows;            // [w0 + nw][1 + 2 nsh]
};
size_t nofk_prep_lds_bytes(int nup, int nq);
int launch_nofk(int nup, const DevSys &S, const NofkArgs &a, hipStream_t st);

}  // namespace dmc

namespace dmc {
size_t allpart_lds_bytes(int nup, int nq);
}
