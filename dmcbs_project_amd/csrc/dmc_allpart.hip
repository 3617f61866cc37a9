// All-particle (reference-replay) kernels: hpsi of a walker in one 64-lane workgroup, the
// fused DMC step (Gaussian drift-diffusion + hpsi + weight + node test + estimators +
// branch multiplicity), single-particle ratios, and the parameter adjoint.
//
// Work split inside a walker: the 64 lanes form an 8 x 8 grid; lane (r, c) owns up
// particles i = r + 8a and down particles j = c + 8b (a, b < TR = ceil(n/8)), i.e. a
// TR x TR tile of the n x n pairing matrix S(i, j) = phi(r_up_i - r_dn_j).  The tile stays
// in registers through the Gauss-Jordan inversion; only the pivot row and the pivot column
// travel through LDS.  Plane-wave phases exp(i q.r) of every particle are built from three
// sincos per particle and staged in LDS.
#include "dmc_wave.h"

namespace dmc {

// ------------------------------------------------------------------------------------
// kernels

template <int NUP>
__global__ __launch_bounds__(64, 2) void k_eval(DevSys S, const double *__restrict__ x, int nw,
                                             double *psil, int *is, double *dpsi,
                                             double *d2psi, double *v, char *sinv) {
    extern __shared__ double sm[];
    const Layout<NUP> L(S.nq, true);
    const int w = blockIdx.x, tid = threadIdx.x;
    if (w >= nw) return;
    constexpr int M = 6 * NUP;
    load_q<NUP>(S, sm, L, tid);
    for (int t = tid; t < M; t += 64) sm[L.xs + t] = x[size_t(w) * M + t];
    wave_sync();
    const Tile<NUP> T(tid);
    PairTile<NUP> rad;
    const JSums J = jastrow_terms<NUP>(S, sm, L, T, tid, &rad);
    const DetInfo D = build_inverse<NUP>(S, sm, L, T, tid, &rad);
    if (sinv) store_sinv<NUP>(S, sm, L, D.singular, sinv + size_t(w) * sinv_stride(S), tid);
    grad_lap<NUP>(S, sm, L, T, D, J, tid);
    for (int t = tid; t < M; t += 64) dpsi[size_t(w) * M + t] = sm[L.dp + t];
    if (tid == 0) {
        psil[w] = sm[L.scal];
        d2psi[w] = sm[L.scal + 1];
        v[w] = sm[L.scal + 2];
        is[w] = int(sm[L.scal + 3]);
    }
}

// The fused DMC step of the reference (step.f90:56-76,127-160 with idmc = 1) for walker w
// of state A, written to the same slot of state B, followed by the branch draw
// (step.f90:302-305): B.irn holds the state after that draw and T.count[w] = iwt.
template <int NUP>
__global__ __launch_bounds__(64, 2) void k_step(DevSys S, DevState A, DevState B, DevStep T,
                                             int opcomp) {
    extern __shared__ double sm[];
    const Layout<NUP> L(S.nq, true);
    const int w = blockIdx.x, tid = threadIdx.x;
    if (w >= T.nt[0]) return;
    constexpr int M = 6 * NUP;                // 3 * npart coordinates
    constexpr int ND = 2 * ((M + 1) / 2);     // uniforms drawn by gaussian(M)
    load_q<NUP>(S, sm, L, tid);
    const uint64_t s0 = A.irn[w];
    const double sig = S.sigma;
    const double dfac = S.sig2 * S.sqrtm;
    const double *xo = A.x + size_t(w) * M, *go = A.dpsi + size_t(w) * M;
    // Box-Muller pairs (ran.f90:85-102); draw t uses the jump-ahead state after t+1 draws
    for (int m = tid; 2 * m < M; m += 64) {
        const int t0 = 2 * m, t1 = 2 * m + 1;
        const uint64_t a0 = (S.ja1[t0] * s0 + S.jc1[t0]) & kMask48;
        const uint64_t a1 = (S.ja1[t1] * s0 + S.jc1[t1]) & kMask48;
        const double u1 = double(a0 | 1ull) * 0x1p-48, u2 = double(a1 | 1ull) * 0x1p-48;
        double sn, cs;
        sincos(2.0 * kPi * u1, &sn, &cs);
        const double rad = sqrt(-2.0 * log(u2));
        // odd count: the last value uses cos (ran.f90:95-102)
        const double g0 = (t1 == M ? cs : sn) * rad, g1 = cs * rad;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int t = t0 + e;
            if (t < M) {
                double drift = go[t] * dfac;
                if (S.cutdrift) drift = fmin(S.dcut, fmax(-S.dcut, drift));
                const double x1 = wrapd(xo[t], S);
                sm[L.xs + t] = wrapd(x1 + (drift + sig * (e ? g1 : g0)), S);
            }
        }
    }
    wave_sync();
    hpsi_block<NUP>(S, sm, L, tid);
    // new RNG state after gaussian(M)
    const uint64_t s1 = (S.ja1[ND - 1] * s0 + S.jc1[ND - 1]) & kMask48;
    finish_walker<NUP>(S, sm, L, A, B, T, w, tid, s1, opcomp);
}

// Gaussian pair m of gaussian(M) for the walker whose RNG state is s0 (ran.f90:85-102): the
// values of coordinates 2m and 2m+1 (the last value of an odd count uses cos)
__device__ __forceinline__ void gauss_pair(const DevSys &S, uint64_t s0, int m, int M, double &g0,
                                           double &g1) {
    const int t0 = 2 * m, t1 = 2 * m + 1;
    const uint64_t a0 = (S.ja1[t0] * s0 + S.jc1[t0]) & kMask48;
    const uint64_t a1 = (S.ja1[t1] * s0 + S.jc1[t1]) & kMask48;
    const double u1 = double(a0 | 1ull) * 0x1p-48, u2 = double(a1 | 1ull) * 0x1p-48;
    double sn, cs;
    sincos(2.0 * kPi * u1, &sn, &cs);
    const double rad = sqrt(-2.0 * log(u2));
    g0 = (t1 == M ? cs : sn) * rad;
    g1 = cs * rad;
}

// The variational step of the reference (step1 with idmc = 0, step.f90:56-76,78-122): the
// same all-particle drift-diffusion proposal as k_step, then the Metropolis test with the
// ratio of the drift-diffusion Green's functions, one more ran1 draw, and the six VMC
// estimators (total, kinetic, potential, external potential, Jackson-Feenberg energy,
// acceptance) of the state kept.  Walker w of A goes to slot w of B, accepted or not; there
// is no branching.
template <int NUP>
__global__ __launch_bounds__(64, 2) void k_vmc(DevSys S, DevState A, DevState B, DevStep T) {
    extern __shared__ double sm[];
    const Layout<NUP> L(S.nq, true);
    const int w = blockIdx.x, tid = threadIdx.x;
    if (w >= T.nt[0]) return;
    constexpr int M = 6 * NUP;
    constexpr int ND = 2 * ((M + 1) / 2);
    load_q<NUP>(S, sm, L, tid);
    const uint64_t s0 = A.irn[w];
    const double sig = S.sigma, dfac = S.sig2 * S.sqrtm;
    const double *xo = A.x + size_t(w) * M, *go = A.dpsi + size_t(w) * M;
    for (int m = tid; 2 * m < M; m += 64) {
        double g[2];
        gauss_pair(S, s0, m, M, g[0], g[1]);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int t = 2 * m + e;
            if (t < M) {
                double drift = go[t] * dfac;
                if (S.cutdrift) drift = fmin(S.dcut, fmax(-S.dcut, drift));
                sm[L.xs + t] = wrapd(wrapd(xo[t], S) + (drift + sig * g[e]), S);
            }
        }
    }
    wave_sync();
    hpsi_block<NUP>(S, sm, L, tid);
    // Green's-function term sum_i (do_i + dn_i).(do_i - dn_i - 2 dx_i) / (2 sigma^2), and the
    // squared gradients of both states for the Jackson-Feenberg energy
    double gterm = 0.0, dot_o = 0.0, dot_n = 0.0;
    for (int m = tid; 2 * m < M; m += 64) {
        double g[2];
        gauss_pair(S, s0, m, M, g[0], g[1]);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int t = 2 * m + e;
            if (t < M) {
                double dro = go[t] * dfac, drn = sm[L.dp + t] * dfac;
                if (S.cutdrift) {
                    dro = fmin(S.dcut, fmax(-S.dcut, dro));
                    drn = fmin(S.dcut, fmax(-S.dcut, drn));
                }
                const double dx = dro + sig * g[e];
                gterm += (dro + drn) * (dro - drn - 2.0 * dx);
                dot_o += go[t] * go[t];
                dot_n += sm[L.dp + t] * sm[L.dp + t];
            }
        }
    }
    gterm = wave_sum(gterm);
    dot_o = wave_sum(dot_o) * S.invm;
    dot_n = wave_sum(dot_n) * S.invm;
    const double psil2 = sm[L.scal], d2psi2 = sm[L.scal + 1], v2 = sm[L.scal + 2];
    const int is2 = int(sm[L.scal + 3]);
    const double prob = exp(fmin(0.0, 2.0 * (psil2 - A.psil[w]) + gterm / (2.0 * S.sig2)));
    uint64_t s1 = (S.ja1[ND - 1] * s0 + S.jc1[ND - 1]) & kMask48;  // after gaussian(M)
    s1 = (s1 * kM1 + kA1) & kMask48;                                 // randn(1)
    const bool acc = double(s1 | 1ull) * 0x1p-48 < prob;
    double *xn = B.x + size_t(w) * M, *gn = B.dpsi + size_t(w) * M;
    for (int t = tid; t < M; t += 64) {
        xn[t] = acc ? sm[L.xs + t] : wrapd(xo[t], S);
        gn[t] = acc ? sm[L.dp + t] : go[t];
    }
    // the six estimators of the kept state (addval: val*wt, val^2*wt, wt), in LDS first
    double *est = sm + L.est;
    if (tid < 42) est[tid] = 0.0;
    wave_sync();
    if (tid == 0) {
        const double d2 = acc ? d2psi2 : A.d2psi[w], v = acc ? v2 : A.v[w];
        const double dot = acc ? dot_n : dot_o, wgt = A.weight[w], nf = S.nfac;
        const double en = -S.hbar * d2 + v;
        const double ejf = -0.5 * S.hbar * (d2 - dot) + v;
        const double val[6] = {en * nf, (en - v) * nf, v * nf, 0.0, ejf * nf, acc ? 1.0 : 0.0};
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            const double wt = e == 5 ? 1.0 : wgt;
            est[3 * e] = val[e] * wt;
            est[3 * e + 1] = val[e] * val[e] * wt;
            est[3 * e + 2] = wt;
        }
        B.irn[w] = s1;
        B.weight[w] = wgt;
        B.psil[w] = acc ? psil2 : A.psil[w];
        B.d2psi[w] = d2;
        B.v[w] = v;
        B.is[w] = acc ? is2 : A.is[w];
        B.wt0[w] = acc ? double(is2) : A.wt0[w];
        T.nacc[w] = acc ? 1 : 0;
        if (acc) atomicAdd(reinterpret_cast<unsigned long long *>(&T.moves[0]), 1ull);
        atomicAdd(reinterpret_cast<unsigned long long *>(&T.moves[1]), 1ull);
    }
    wave_sync();
    if (tid < 42) T.est[size_t(w) * 42 + tid] = est[tid];
}

// The Metropolis DMC step of the reference (step1 with idmc = 2, step.f90:56-76,161-226): the
// all-particle drift-diffusion proposal and Metropolis test of k_vmc, then the weight of the
// state kept: exp(-((E_old + E_new)/2 - E_T) dt), capped, for an accepted move (zero after a
// node crossing under the fixed-node rule), exp(-(E_old - E_T) dt) times the walker's weight,
// capped, for a rejected one; the 13 estimators of idmc = 2 (dmcbcs.f90:204-218: energies,
// node crossing, acceptance, |psit|/psig, psit, psig, sign, wt, wtg, sign(wtg)) and the
// branch draw (step.f90:302-305) as in k_step.
template <int NUP>
__global__ __launch_bounds__(64, 2) void k_mdmc(DevSys S, DevState A, DevState B, DevStep T,
                                             int opcomp) {
    extern __shared__ double sm[];
    const Layout<NUP> L(S.nq, true);
    const int w = blockIdx.x, tid = threadIdx.x;
    if (w >= T.nt[0]) return;
    constexpr int M = 6 * NUP;
    constexpr int ND = 2 * ((M + 1) / 2);
    load_q<NUP>(S, sm, L, tid);
    const uint64_t s0 = A.irn[w];
    const double sig = S.sigma, dfac = S.sig2 * S.sqrtm;
    const double *xo = A.x + size_t(w) * M, *go = A.dpsi + size_t(w) * M;
    for (int m = tid; 2 * m < M; m += 64) {
        double g[2];
        gauss_pair(S, s0, m, M, g[0], g[1]);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int t = 2 * m + e;
            if (t < M) {
                double drift = go[t] * dfac;
                if (S.cutdrift) drift = fmin(S.dcut, fmax(-S.dcut, drift));
                sm[L.xs + t] = wrapd(wrapd(xo[t], S) + (drift + sig * g[e]), S);
            }
        }
    }
    wave_sync();
    hpsi_block<NUP>(S, sm, L, tid);
    double gterm = 0.0;
    for (int m = tid; 2 * m < M; m += 64) {
        double g[2];
        gauss_pair(S, s0, m, M, g[0], g[1]);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int t = 2 * m + e;
            if (t < M) {
                double dro = go[t] * dfac, drn = sm[L.dp + t] * dfac;
                if (S.cutdrift) {
                    dro = fmin(S.dcut, fmax(-S.dcut, dro));
                    drn = fmin(S.dcut, fmax(-S.dcut, drn));
                }
                const double dx = dro + sig * g[e];
                gterm += (dro + drn) * (dro - drn - 2.0 * dx);
            }
        }
    }
    gterm = wave_sum(gterm);
    const double psil2 = sm[L.scal], d2psi2 = sm[L.scal + 1], v2 = sm[L.scal + 2];
    const int is2 = int(sm[L.scal + 3]);
    const double prob = exp(fmin(0.0, 2.0 * (psil2 - A.psil[w]) + gterm / (2.0 * S.sig2)));
    uint64_t s1 = (S.ja1[ND - 1] * s0 + S.jc1[ND - 1]) & kMask48;  // after gaussian(M)
    s1 = (s1 * kM1 + kA1) & kMask48;                                 // randn(1)
    const bool acc = double(s1 | 1ull) * 0x1p-48 < prob;
    double *xn = B.x + size_t(w) * M, *gn = B.dpsi + size_t(w) * M;
    for (int t = tid; t < M; t += 64) {
        xn[t] = acc ? sm[L.xs + t] : wrapd(xo[t], S);
        gn[t] = acc ? sm[L.dp + t] : go[t];
    }
    double *est = sm + L.est;
    if (tid < 42) est[tid] = 0.0;
    wave_sync();
    if (tid == 0) {
        auto put = [&](int e, double val, double wgt) {
            est[3 * e] = val * wgt;
            est[3 * e + 1] = val * val * wgt;
            est[3 * e + 2] = wgt;
        };
        const double eo = -S.hbar * A.d2psi[w] + A.v[w], nf = S.nfac, wt0 = A.wt0[w];
        double wt, weight, e, v, psil;
        int is;
        if (acc) {
            e = -S.hbar * d2psi2 + v2;
            v = v2;
            psil = psil2;
            is = is2;
            wt = exp(-(0.5 * (eo + e) - S.etrial) * S.dt);
            weight = fmin(S.wtcut, wt);
            const bool cross = is2 != A.is[w];
            if (cross && S.fn) {
                weight = 0.0;
                wt = 0.0;
            }
            put(3, cross ? 1.0 : 0.0, 1.0);  // node crossing, accepted moves only
        } else {
            e = eo;
            v = A.v[w];
            psil = A.psil[w];
            is = A.is[w];
            wt = exp(-(eo - S.etrial) * S.dt);
            weight = fmin(S.wtcut, wt * A.weight[w]);
        }
        put(4, acc ? 1.0 : 0.0, 1.0);  // acceptance
        const double wtg = wt * double(is) * wt0;
        if (opcomp) {
            const double sg = copysign(1.0, wtg);
            put(0, e * nf, wtg);
            put(1, (e - v) * nf, wtg);
            put(2, v * nf, wtg);
            put(5, 1.0, 1.0);
            put(6, is * exp(psil), 1.0);
            put(7, exp(psil), 1.0);
            put(8, double(is), 1.0);
            put(9, wt, 1.0);
            put(10, wtg, 1.0);
            put(11, sg, 1.0);
            put(12, sg, wtg);
        }
        s1 = (s1 * kM1 + kA1) & kMask48;  // the branch draw
        const double ub = double(s1 | 1ull) * 0x1p-48;
        const int iwt = int(weight + ub);
        T.count[w] = iwt;
        if (iwt >= 0 && iwt < 500)
            atomicAdd(reinterpret_cast<unsigned long long *>(&T.hist[iwt]), 1ull);
        B.irn[w] = s1;
        B.psil[w] = psil;
        B.d2psi[w] = acc ? d2psi2 : A.d2psi[w];
        B.v[w] = v;
        B.is[w] = is;
        B.wt0[w] = wt0;
        B.weight[w] = weight;
        T.nacc[w] = acc ? 1 : 0;
        if (acc) atomicAdd(reinterpret_cast<unsigned long long *>(&T.moves[0]), 1ull);
        atomicAdd(reinterpret_cast<unsigned long long *>(&T.moves[1]), 1ull);
    }
    wave_sync();
    if (tid < 42) T.est[size_t(w) * 42 + tid] = est[tid];
}

// single-particle ratio psi(x')/psi(x) (move1, wavefunction.f90:393-404): determinant
// ratio from the inverse (bcsone, psislaterbcs.f90:296-376) times the Jastrow ratio
template <int NUP>
__global__ __launch_bounds__(64) void k_ratio(DevSys S, const double *__restrict__ x,
                                              const int *ip, const double *xnew, int nw,
                                              double *ratio) {
    extern __shared__ double sm[];
    const Layout<NUP> L(S.nq);
    constexpr int SP = Layout<NUP>::SP, NP = Layout<NUP>::NP;
    const int w = blockIdx.x, tid = threadIdx.x;
    if (w >= nw) return;
    constexpr int M = 6 * NUP;
    load_q<NUP>(S, sm, L, tid);
    for (int t = tid; t < M; t += 64) sm[L.xs + t] = x[size_t(w) * M + t];
    double *xn = sm + L.xn;
    if (tid < 3) xn[tid] = xnew[3 * w + tid];
    wave_sync();
    const Tile<NUP> T(tid);
    const DetInfo D = build_inverse<NUP>(S, sm, L, T, tid);
    const int i = ip[w];
    const double2 *ph = reinterpret_cast<const double2 *>(sm + L.work);
    double2 *phn = reinterpret_cast<double2 *>(sm + L.xn + 4);
    const int4 *qm = reinterpret_cast<const int4 *>(sm + L.qm);
    if (tid == 0) phases_of(S, qm, xn, phn, 1);
    wave_sync();
    const double *qc = sm + L.qc, *sinv = sm + L.sinv, *xs = sm + L.xs;
    double part = 0.0;
    if (tid < NUP) {
        const int j = tid;
        if (i < NUP) {
            const double f = phi_pair(S, qc, phn, 1, ph + NUP + j, NP, xn, xs + 3 * (NUP + j));
            part = sinv[j * SP + i] * f;
        } else {
            const double f = phi_pair(S, qc, ph + j, NP, phn, 1, xs + 3 * j, xn);
            part = sinv[(i - NUP) * SP + j] * f;
        }
    }
    const double rdet = wave_sum(part);
    const double lj = jastrow_logratio<NUP>(S, xs, i, xn, tid);
    if (tid == 0) ratio[w] = D.singular ? 0.0 : exp(lj) * rdet;
}

// ------------------------------------------------------------------------------------
// launchers

bool nup_supported(int nup) { return nup == 1 || nup == 7 || nup == 19 || nup == 33; }

#define DMC_DISPATCH(NUPV, CALL)          \
    switch (NUPV) {                       \
        case 1: { constexpr int N_ = 1; CALL; } break;   \
        case 7: { constexpr int N_ = 7; CALL; } break;   \
        case 19: { constexpr int N_ = 19; CALL; } break; \
        case 33: { constexpr int N_ = 33; CALL; } break; \
        default: return hipErrorInvalidValue;            \
    }

int launch_eval(int nup, const DevSys &S, const double *x, int nw, double *psil, int *is,
                double *dpsi, double *d2psi, double *v, char *sinv, hipStream_t st) {
    if (nw <= 0) return hipSuccess;
    DMC_DISPATCH(nup, hipLaunchKernelGGL(k_eval<N_>, dim3(nw), dim3(64), lds_bytes<N_>(S.nq, true), st,
                                         S, x, nw, psil, is, dpsi, d2psi, v, sinv));
    return hipGetLastError();
}

int launch_ratio(int nup, const DevSys &S, const double *x, const int *ip, const double *xnew,
                 int nw, double *ratio, hipStream_t st) {
    if (nw <= 0) return hipSuccess;
    DMC_DISPATCH(nup, hipLaunchKernelGGL(k_ratio<N_>, dim3(nw), dim3(64), lds_bytes<N_>(S.nq), st,
                                         S, x, ip, xnew, nw, ratio));
    return hipGetLastError();
}

int launch_step(int nup, const DevSys &S, const DevState &A, const DevState &B,
                const DevStep &T, int opcomp, int grid, hipStream_t st) {
    if (grid <= 0) return hipSuccess;
    DMC_DISPATCH(nup, hipLaunchKernelGGL(k_step<N_>, dim3(grid), dim3(64), lds_bytes<N_>(S.nq, true), st,
                                         S, A, B, T, opcomp));
    return hipGetLastError();
}

int launch_vmc(int nup, const DevSys &S, const DevState &A, const DevState &B, const DevStep &T,
               int grid, hipStream_t st) {
    if (grid <= 0) return hipSuccess;
    DMC_DISPATCH(nup, hipLaunchKernelGGL(k_vmc<N_>, dim3(grid), dim3(64), lds_bytes<N_>(S.nq, true), st,
                                         S, A, B, T));
    return hipGetLastError();
}

int launch_mdmc(int nup, const DevSys &S, const DevState &A, const DevState &B,
                const DevStep &T, int opcomp, int grid, hipStream_t st) {
    if (grid <= 0) return hipSuccess;
    DMC_DISPATCH(nup, hipLaunchKernelGGL(k_mdmc<N_>, dim3(grid), dim3(64), lds_bytes<N_>(S.nq, true), st,
                                         S, A, B, T, opcomp));
    return hipGetLastError();
}

size_t allpart_lds_bytes(int nup, int nq) {
    switch (nup) {
        case 1: return lds_bytes<1>(nq, true);
        case 7: return lds_bytes<7>(nq, true);
        case 19: return lds_bytes<19>(nq, true);
        case 33: return lds_bytes<33>(nq, true);
    }
    return 0;
}

}  // namespace dmc
