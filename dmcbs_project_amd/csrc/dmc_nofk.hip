// Momentum distribution n(k) of the resident population (addenk, momentum3d.f90:89-120):
// for every particle i and nsamp samples, a uniform displacement dx in the box drawn from the
// walker's own ran1 stream (set from irn and not saved back), the single-particle ratio
// psi(x with r_i + dx)/psi(x) of move1 (wavefunction.f90:393-404) with the inverse of hpsi,
// and per shell of octant vectors the sum of psirat * prod_d cos(dk m_d dx_d).
//
// One 64-lane wave per walker:
//   * FP64 inverse of the pairing matrix and the phases of every particle (build_inverse);
//   * the projections P_i(q) = sum_j Sinv(j, i) exp(i q.r_j) over the partners of i, so
//     that the plane-wave part of a ratio is sum_q c_q Re(exp(i q.x') conj P_i(q)) -- nq
//     products per move instead of n nq; the down-spin P in their own area, the up-spin P
//     over the up-spin phases, which are no longer needed once the down-spin P are formed;
//   * the Jastrow sum of each particle at its old position;
//   * lanes over the N nsamp moves: three ran1 draws from the jump-ahead of 3 m draws, the
//     radial part sum_j Sinv (alpha_0 + beta(r')) and the new Jastrow sum over the n
//     partners, the plane-wave part, then cos(dk m dx_d) for m = 0..7 by the Chebyshev
//     recurrence and the shell sums into a per-lane column of LDS;
//   * lanes over shells sum the 64 columns in order and write the walker's row
//     (wt nsamp, enktmp/nk wt, (enktmp/nk)^2 wt), which a fixed-order reduction over the
//     walkers turns into the block sums.
#include "dmc_wave.h"

namespace dmc {
namespace {

constexpr int kAccStride = 65;  // shell accumulators [nsh][65]: lanes in distinct banks

// cos(m theta) from the table, m wave-uniform (the same vector for every lane)
__device__ __forceinline__ double sel8(const double (&c)[8], int m) {
    switch (__builtin_amdgcn_readfirstlane(m)) {
        case 0: return c[0];
        case 1: return c[1];
        case 2: return c[2];
        case 3: return c[3];
        case 4: return c[4];
        case 5: return c[5];
        case 6: return c[6];
        default: return c[7];
    }
}

// Jastrow sum over the opposite-spin partners of particle i at position pos (getphione,
// coshpot.f90:207-264), one lane
template <int NUP>
__device__ __forceinline__ double jastrow_sum(const DevSys &S, const double *xs, int i,
                                              const double *pos) {
    const int p0 = i < NUP ? NUP : 0;
    double u = 0.0;
    for (int j = 0; j < NUP; ++j) {
        const double *xj = xs + 3 * (p0 + j);
        double r2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double dx = wrapd(pos[d] - xj[d], S);
            r2 += dx * dx;
        }
        const double rj = fmax(1.0e-14, sqrt(r2));
        if (rj < S.range) {
            double w0, w1, w2, w3;
            const int idx = lagrange(S.scale * rj, S.ntab, w0, w1, w2, w3);
            u += lookup4(S.jt, idx, w0, w1, w2, w3).y;
        }
    }
    return u;
}

template <int NUP>
__global__ __launch_bounds__(64) void k_nofk(DevSys S, NofkArgs a) {
    extern __shared__ double sm[];
    const Layout<NUP> L(S.nq);
    constexpr int SP = Layout<NUP>::SP, NP = Layout<NUP>::NP, M = 6 * NUP;
    const int w = blockIdx.x, tid = threadIdx.x;
    if (w >= a.nw) return;
    const int nq = S.nq, nsh = a.nsh, nsamp = a.nsamp;
    double2 *pdn = reinterpret_cast<double2 *>(sm + ((L.total + 1) & ~1));  // [NUP][nq]
    double *uold = reinterpret_cast<double *>(pdn + NUP * nq);               // [NP]
    double *acc = uold + NP;                                                  // [nsh][65]

    load_q<NUP>(S, sm, L, tid);
    for (int t = tid; t < M; t += 64) sm[L.xs + t] = a.x[size_t(w) * M + t];
    for (int e = tid; e < nsh * kAccStride; e += 64) acc[e] = 0.0;
    wave_sync();
    const Tile<NUP> T(tid);
    const DetInfo D = build_inverse<NUP>(S, sm, L, T, tid);
    double2 *ph = reinterpret_cast<double2 *>(sm + L.work);  // exp(i q.r_p) at [q][NP]
    const double *sinv = sm + L.sinv, *xs = sm + L.xs, *qc = sm + L.qc;
    const int4 *qm = reinterpret_cast<const int4 *>(sm + L.qm);

    // down-spin particle i: rows of the ratio are Sinv(i - n, j) phi(r_j - x')
    for (int e = tid; e < NUP * nq; e += 64) {
        const int i = e / nq, q = e - i * nq;
        double re = 0.0, im = 0.0;
        for (int j = 0; j < NUP; ++j) {
            const double s = sinv[i * SP + j];
            const double2 z = ph[q * NP + j];
            re = fma(s, z.x, re);
            im = fma(s, z.y, im);
        }
        pdn[i * nq + q] = make_double2(re, im);
    }
    wave_sync();
    // up-spin particle i: Sinv(j, i) phi(x' - r_{n+j}); written over the up-spin phases
    for (int e = tid; e < NUP * nq; e += 64) {
        const int i = e / nq, q = e - i * nq;
        double re = 0.0, im = 0.0;
        for (int j = 0; j < NUP; ++j) {
            const double s = sinv[j * SP + i];
            const double2 z = ph[q * NP + NUP + j];
            re = fma(s, z.x, re);
            im = fma(s, z.y, im);
        }
        ph[q * NP + i] = make_double2(re, im);
    }
    for (int i = tid; i < NP; i += 64) uold[i] = jastrow_sum<NUP>(S, xs, i, xs + 3 * i);
    wave_sync();

    const uint64_t s0 = a.irn[w] & kMask48;
    const int nmove = NP * nsamp;
    for (int m = tid; m < nmove; m += 64) {
        const int i = m / nsamp;
        uint64_t s = (a.ja[m] * s0 + a.jc[m]) & kMask48;  // state after 3 m draws
        double dx[3], xn[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s = (s * kM1 + kA1) & kMask48;
            dx[d] = (double(s | 1ULL) * 0x1p-48 - 0.5) * S.el;
            xn[d] = xs[3 * i + d] + dx[d];
        }
        const bool up = i < NUP;
        const int p0 = up ? NUP : 0;
        double unew = 0.0, rad = 0.0;
        for (int j = 0; j < NUP; ++j) {
            const double *xj = xs + 3 * (p0 + j);
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double dd = wrapd(xn[d] - xj[d], S);
                r2 += dd * dd;
            }
            const double r = sqrt(r2);
            const double rj = fmax(1.0e-14, r);
            if (rj < S.range) {
                double w0, w1, w2, w3;
                const int idx = lagrange(S.scale * rj, S.ntab, w0, w1, w2, w3);
                unew += lookup4(S.jt, idx, w0, w1, w2, w3).y;
            }
            const double sij = up ? sinv[j * SP + i] : sinv[(i - NUP) * SP + j];
            rad = fma(sij, S.alpha0 + beta_val(S, r), rad);
        }
        double pc[3][5], ps[3][5];
        axis_powers(S, xn, pc, ps);
        const double2 *P = up ? ph + i : pdn + (i - NUP) * nq;
        const int pst = up ? NP : 1;
        double pl = 0.0;
        for (int q = 0; q < nq; ++q) {
            const int4 mq = qm[q];
            double cx, sx, cy, sy, cz, sz;
            cpow(pc[0], ps[0], mq.x, cx, sx);
            cpow(pc[1], ps[1], mq.y, cy, sy);
            cpow(pc[2], ps[2], mq.z, cz, sz);
            const double cxy = cx * cy - sx * sy, sxy = sx * cy + cx * sy;
            const double zr = cxy * cz - sxy * sz, zi = sxy * cz + cxy * sz;
            const double2 pv = P[q * pst];
            pl = fma(qc[5 * q], zr * pv.x + zi * pv.y, pl);
        }
        const double psirat = D.singular ? 0.0 : exp(unew - uold[i]) * (rad + pl);
        // cos(dk m dx_d), m = 0..7
        double cc[3][8];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double c1 = cos(a.dk * dx[d]);
            cc[d][0] = 1.0;
            cc[d][1] = c1;
#pragma unroll
            for (int k = 2; k < 8; ++k) cc[d][k] = fma(2.0 * c1, cc[d][k - 1], -cc[d][k - 2]);
        }
        for (int k = 0; k < nsh; ++k) {
            const int2 sh = a.shell[k];
            double f = 0.0;
            for (int v = sh.x; v < sh.x + sh.y; ++v) {
                const int4 mv = a.ik[v];
                f = fma(sel8(cc[0], mv.x) * sel8(cc[1], mv.y), sel8(cc[2], mv.z), f);
            }
            acc[k * kAccStride + tid] = fma(psirat, f, acc[k * kAccStride + tid]);
        }
    }
    wave_sync();
    const double wt = a.weight[w] * double(a.is[w]) * a.wt0[w];
    double *row = a.rows + size_t(w) * (1 + 2 * nsh);
    for (int k = tid; k < nsh; k += 64) {
        double e = 0.0;
        for (int l = 0; l < 64; ++l) e += acc[k * kAccStride + l];
        const double en = e / double(a.shell[k].y);
        row[1 + k] = en * wt;
        row[1 + nsh + k] = en * en * wt;
    }
    if (tid == 0) row[0] = wt * double(nsamp);
}

}  // namespace

size_t nofk_lds_bytes(int nup, int nq, int nsh) {
    size_t base;
    switch (nup) {
        case 1: base = lds_bytes<1>(nq); break;
        case 7: base = lds_bytes<7>(nq); break;
        case 19: base = lds_bytes<19>(nq); break;
        case 33: base = lds_bytes<33>(nq); break;
        default: return 0;
    }
    base = (base + 15) & ~size_t(15);
    return base + sizeof(double) * (2 * size_t(nup) * nq + 2 * size_t(nup) +
                                    size_t(nsh) * kAccStride);
}

int launch_nofk(int nup, const DevSys &S, const NofkArgs &a, hipStream_t st) {
    if (a.nw <= 0) return hipSuccess;
    const size_t lds = nofk_lds_bytes(nup, S.nq, a.nsh);
    switch (nup) {
        case 1: hipLaunchKernelGGL(k_nofk<1>, dim3(a.nw), dim3(64), lds, st, S, a); break;
        case 7: hipLaunchKernelGGL(k_nofk<7>, dim3(a.nw), dim3(64), lds, st, S, a); break;
        case 19: hipLaunchKernelGGL(k_nofk<19>, dim3(a.nw), dim3(64), lds, st, S, a); break;
        case 33: hipLaunchKernelGGL(k_nofk<33>, dim3(a.nw), dim3(64), lds, st, S, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dmc
