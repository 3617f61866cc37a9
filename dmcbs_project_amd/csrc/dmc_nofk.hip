// Momentum distribution n(k) of the resident population (addenk, momentum3d.f90:89-120):
// for every particle i and nsamp samples, a uniform displacement dx in the box drawn from the
// walker's own ran1 stream (set from irn and not saved back), the single-particle ratio
// psi(x with r_i + dx)/psi(x) of move1 (wavefunction.f90:393-404) with the inverse of hpsi,
// and per shell of octant vectors the sum of psirat * prod_d cos(dk m_d dx_d).
//
// Three launches, so that the displacements run at full occupancy:
//   * k_nofk_prep, one 64-lane wave per walker: FP64 inverse of the pairing matrix and the
//     phases of every particle (build_inverse, 35 KB of LDS at N = 66), then into per-walker
//     buffers in HBM: the projections P_i(q) = sum_j Sinv exp(i q.r_j) over the partners of
//     i (the plane-wave part of a ratio is then sum_q c_q Re(exp(i q.x') conj P_i(q)), nq
//     products per displacement instead of n nq), the inverse entries of i with each
//     partner, and the Jastrow sum of i at its old position;
//   * k_nofk_moves, one wave per 64 displacements of a walker (1.6 KB of LDS: positions):
//     three ran1 draws from the jump-ahead of 3 m draws, the radial part
//     sum_j Sinv (alpha_0 + beta(r')) and the new Jastrow sum over the n partners, the
//     plane-wave part, cos(dk m dx_d) for m = 0..7 by the Chebyshev recurrence, and per
//     shell the wave sum of psirat times the shell's cosine products;
//   * k_nofk_rows, one thread per walker: the sum over its waves in order and its row
//     (wt nsamp, enktmp/nk wt, (enktmp/nk)^2 wt), reduced over the walkers in a fixed order
//     afterwards.
#include "dmc_wave.h"

namespace dmc {
namespace {

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

// u of the Jastrow/potential table at distance r (getphione, coshpot.f90:207-264)
__device__ __forceinline__ double jastrow_u(const DevSys &S, double r2) {
    const double rj = fmax(1.0e-14, sqrt(r2));
    if (rj >= S.range) return 0.0;
    double w0, w1, w2, w3;
    const int idx = lagrange(S.scale * rj, S.ntab, w0, w1, w2, w3);
    if (idx >= S.jfar) return 0.0;  // u is zero on stencils past the healing distance
    return w0 * S.jt[idx - 1].y + w1 * S.jt[idx].y + w2 * S.jt[idx + 1].y +
           w3 * S.jt[idx + 2].y;
}

template <int NUP>
__global__ __launch_bounds__(64) void k_nofk_prep(DevSys S, NofkArgs a) {
    extern __shared__ double sm[];
    const Layout<NUP> L(S.nq);
    constexpr int SP = Layout<NUP>::SP, NP = Layout<NUP>::NP, M = 6 * NUP;
    const int b = blockIdx.x, tid = threadIdx.x;
    if (b >= a.nw) return;
    const int w = a.w0 + b, nq = S.nq;
    load_q<NUP>(S, sm, L, tid);
    for (int t = tid; t < M; t += 64) sm[L.xs + t] = a.x[size_t(w) * M + t];
    wave_sync();
    const Tile<NUP> T(tid);
    const DetInfo D = build_inverse<NUP>(S, sm, L, T, tid);
    const double2 *ph = reinterpret_cast<const double2 *>(sm + L.work);  // [q][NP]
    const double *sinv = sm + L.sinv, *xs = sm + L.xs;
    double2 *P = a.proj + size_t(b) * NP * nq;
    // up-spin i: Sinv(j, i) phi(x' - r_{n+j}); down-spin i: Sinv(i - n, j) phi(r_j - x')
    for (int e = tid; e < NP * nq; e += 64) {
        const int i = e / nq, q = e - i * nq;
        double re = 0.0, im = 0.0;
        if (i < NUP) {
            for (int j = 0; j < NUP; ++j) {
                const double s = sinv[j * SP + i];
                const double2 z = ph[q * NP + NUP + j];
                re = fma(s, z.x, re);
                im = fma(s, z.y, im);
            }
        } else {
            for (int j = 0; j < NUP; ++j) {
                const double s = sinv[(i - NUP) * SP + j];
                const double2 z = ph[q * NP + j];
                re = fma(s, z.x, re);
                im = fma(s, z.y, im);
            }
        }
        P[e] = make_double2(re, im);
    }
    double *C = a.coef + size_t(b) * NP * NUP;
    for (int e = tid; e < NP * NUP; e += 64) {
        const int i = e / NUP, j = e - i * NUP;
        C[e] = i < NUP ? sinv[j * SP + i] : sinv[(i - NUP) * SP + j];
    }
    for (int i = tid; i < NP; i += 64) {
        const int p0 = i < NUP ? NUP : 0;
        double u = 0.0;
        for (int j = 0; j < NUP; ++j) {
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double dx = wrapd(xs[3 * i + d] - xs[3 * (p0 + j) + d], S);
                r2 += dx * dx;
            }
            u += jastrow_u(S, r2);
        }
        a.uold[size_t(b) * NP + i] = u;
    }
    if (tid == 0) a.sing[b] = D.singular ? 1 : 0;
}

template <int NUP>
__global__ __launch_bounds__(64) void k_nofk_moves(DevSys S, NofkArgs a) {
    constexpr int NP = 2 * NUP, M = 6 * NUP;
    __shared__ double xs[M];
    const int tid = threadIdx.x, nchunk = a.nchunk;
    const int b = blockIdx.x / nchunk, c = blockIdx.x - b * nchunk;
    const int w = a.w0 + b, nq = S.nq, nsamp = a.nsamp;
    for (int t = tid; t < M; t += 64) xs[t] = a.x[size_t(w) * M + t];
    wave_sync();
    const int m = c * 64 + tid;
    const bool live = m < NP * nsamp;
    const int i = live ? m / nsamp : 0;
    double dx[3] = {0.0, 0.0, 0.0};
    double psirat = 0.0;
    if (live) {
        uint64_t s = (a.ja[m] * (a.irn[w] & kMask48) + a.jc[m]) & kMask48;  // after 3 m draws
        double xn[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s = (s * kM1 + kA1) & kMask48;
            dx[d] = (double(s | 1ULL) * 0x1p-48 - 0.5) * S.el;
            xn[d] = xs[3 * i + d] + dx[d];
        }
        const int p0 = i < NUP ? NUP : 0;
        const double *C = a.coef + (size_t(b) * NP + i) * NUP;
        double unew = 0.0, rad = 0.0;
        for (int j = 0; j < NUP; ++j) {
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double dd = wrapd(xn[d] - xs[3 * (p0 + j) + d], S);
                r2 += dd * dd;
            }
            unew += jastrow_u(S, r2);
            rad = fma(C[j], S.alpha0 + beta_val(S, sqrt(r2)), rad);
        }
        double pc[3][5], ps[3][5];
        axis_powers(S, xn, pc, ps);
        const double2 *P = a.proj + (size_t(b) * NP + i) * nq;
        double pl = 0.0;
        for (int q = 0; q < nq; ++q) {
            const int4 mq = S.qm[q];
            double cx, sx, cy, sy, cz, sz;
            cpow(pc[0], ps[0], mq.x, cx, sx);
            cpow(pc[1], ps[1], mq.y, cy, sy);
            cpow(pc[2], ps[2], mq.z, cz, sz);
            const double cxy = cx * cy - sx * sy, sxy = sx * cy + cx * sy;
            const double zr = cxy * cz - sxy * sz, zi = sxy * cz + cxy * sz;
            const double2 pv = P[q];
            pl = fma(S.wq[q], zr * pv.x + zi * pv.y, pl);
        }
        psirat = a.sing[b] ? 0.0 : exp(unew - a.uold[size_t(b) * NP + i]) * (rad + pl);
    }
    // cos(dk m dx_d), m = 0..7 (idle lanes: dx = 0, psirat = 0)
    double cc[3][8];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double c1 = cos(a.dk * dx[d]);
        cc[d][0] = 1.0;
        cc[d][1] = c1;
#pragma unroll
        for (int k = 2; k < 8; ++k) cc[d][k] = fma(2.0 * c1, cc[d][k - 1], -cc[d][k - 2]);
    }
    double *out = a.part + (size_t(b) * nchunk + c) * a.nsh;
    for (int k = 0; k < a.nsh; ++k) {
        const int2 sh = a.shell[k];
        double f = 0.0;
        for (int v = sh.x; v < sh.x + sh.y; ++v) {
            const int4 mv = a.ik[v];
            f = fma(sel8(cc[0], mv.x) * sel8(cc[1], mv.y), sel8(cc[2], mv.z), f);
        }
        const double tot = wave_sum(psirat * f);
        if (tid == 0) out[k] = tot;
    }
}

__global__ __launch_bounds__(256) void k_nofk_rows(NofkArgs a) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= a.nw) return;
    const int w = a.w0 + b, nsh = a.nsh, nchunk = a.nchunk;
    const double wt = a.weight[w] * double(a.is[w]) * a.wt0[w];
    double *row = a.rows + size_t(w) * (1 + 2 * nsh);
    const double *pw = a.part + size_t(b) * nchunk * nsh;
    row[0] = wt * double(a.nsamp);
    for (int k = 0; k < nsh; ++k) {
        double e = 0.0;
        for (int c = 0; c < nchunk; ++c) e += pw[c * nsh + k];
        const double en = e / double(a.shell[k].y);
        row[1 + k] = en * wt;
        row[1 + nsh + k] = en * en * wt;
    }
}

}  // namespace

size_t nofk_prep_lds_bytes(int nup, int nq) {
    switch (nup) {
        case 1: return lds_bytes<1>(nq);
        case 7: return lds_bytes<7>(nq);
        case 19: return lds_bytes<19>(nq);
        case 33: return lds_bytes<33>(nq);
        default: return 0;
    }
}

int launch_nofk(int nup, const DevSys &S, const NofkArgs &a, hipStream_t st) {
    if (a.nw <= 0) return hipSuccess;
    const size_t lds = nofk_prep_lds_bytes(nup, S.nq);
    const dim3 gm(unsigned(a.nw) * unsigned(a.nchunk));
    switch (nup) {
        case 1:
            hipLaunchKernelGGL(k_nofk_prep<1>, dim3(a.nw), dim3(64), lds, st, S, a);
            hipLaunchKernelGGL(k_nofk_moves<1>, gm, dim3(64), 0, st, S, a);
            break;
        case 7:
            hipLaunchKernelGGL(k_nofk_prep<7>, dim3(a.nw), dim3(64), lds, st, S, a);
            hipLaunchKernelGGL(k_nofk_moves<7>, gm, dim3(64), 0, st, S, a);
            break;
        case 19:
            hipLaunchKernelGGL(k_nofk_prep<19>, dim3(a.nw), dim3(64), lds, st, S, a);
            hipLaunchKernelGGL(k_nofk_moves<19>, gm, dim3(64), 0, st, S, a);
            break;
        case 33:
            hipLaunchKernelGGL(k_nofk_prep<33>, dim3(a.nw), dim3(64), lds, st, S, a);
            hipLaunchKernelGGL(k_nofk_moves<33>, gm, dim3(64), 0, st, S, a);
            break;
        default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_nofk_rows, dim3((a.nw + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace dmc
