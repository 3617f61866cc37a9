// Device building blocks of the single-particle-move sweep: FP32/FP64 pair evaluation,
// move-kernel LDS layout and the per-walker bodies of the move, refresh and end kernels
// (one walker per 64-lane wave, LDS base and walker index given by the caller).
#pragma once
#include "dmc_wave.h"

namespace dmc {

template <typename R>
struct Vec2;
template <>
struct Vec2<double> {
    using T = double2;
};
template <>
struct Vec2<float> {
    using T = float2;
};

// minimum image with round-to-nearest-even (a tie has measure zero for these positions)
__device__ __forceinline__ double wrapr(double x, const DevSys &S) {
    return x - S.el * rint(x * S.eli);
}

// square root and reciprocal: correctly rounded in FP64, the hardware approximations
// (about 1 ulp) in FP32
template <typename R>
__device__ __forceinline__ R sqrt_r(R x) {
    if constexpr (sizeof(R) == 8) return sqrt(x);
    else return __builtin_amdgcn_sqrtf(x);
}
template <typename R>
__device__ __forceinline__ R rcp_r(R x) {
    if constexpr (sizeof(R) == 8) return 1.0 / x;
    else return __builtin_amdgcn_rcpf(x);
}

// distance of the pair (a, b), the minimum-image difference formed in FP64 and the rest in R
template <typename R>
__device__ __forceinline__ R pair_dist(const DevSys &S, const double *xa, const double *xb,
                                       R (&dx)[3]) {
    R r2 = R(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        dx[d] = R(wrapr(xa[d] - xb[d], S));
        r2 += dx[d] * dx[d];
    }
    return sqrt_r<R>(r2);
}

template <typename R, typename V>
__device__ __forceinline__ R field(const V &v, int k) {
    return k == 0 ? R(v.x) : k == 1 ? R(v.y) : k == 2 ? R(v.z) : R(v.w);
}

// R-typed interpolation of two fields of a table record at s (4-point Lagrange,
// coshpot.f90:172-192); fields (0,1) = (b, db) of the beta table, (1,2) = (u, du) of the
// Jastrow table
template <typename R, int F0, int F1>
__device__ __forceinline__ void tab2(const double4 *t64, const float4 *t32, R s, int ntab, R &x,
                                     R &y) {
    int idx = int(s);
    idx = idx > ntab - 2 ? ntab - 2 : idx;
    idx = idx < 1 ? 1 : idx;
    const R t = s - R(idx);
    const R w0 = -t * (t - R(1)) * (t - R(2)) * R(1.0 / 6.0);
    const R w1 = (t + R(1)) * (t - R(1)) * (t - R(2)) * R(0.5);
    const R w2 = -(t + R(1)) * t * (t - R(2)) * R(0.5);
    const R w3 = (t + R(1)) * t * (t - R(1)) * R(1.0 / 6.0);
    if constexpr (sizeof(R) == 8) {
        const double4 a = t64[idx - 1], b = t64[idx], c = t64[idx + 1], d = t64[idx + 2];
        x = w0 * field<R>(a, F0) + w1 * field<R>(b, F0) + w2 * field<R>(c, F0) + w3 * field<R>(d, F0);
        y = w0 * field<R>(a, F1) + w1 * field<R>(b, F1) + w2 * field<R>(c, F1) + w3 * field<R>(d, F1);
    } else {
        const float4 a = t32[idx - 1], b = t32[idx], c = t32[idx + 1], d = t32[idx + 2];
        x = w0 * field<R>(a, F0) + w1 * field<R>(b, F0) + w2 * field<R>(c, F0) + w3 * field<R>(d, F0);
        y = w0 * field<R>(a, F1) + w1 * field<R>(b, F1) + w2 * field<R>(c, F1) + w3 * field<R>(d, F1);
    }
}

// LDS of the move kernel (offsets in doubles)
template <int NUP, typename R>
struct MvLayout {
    // row stride of the inverse: >= NUP and = 1 (mod 16), so that lanes walking a column of
    // the inverse hit distinct banks
    static constexpr int SP = ((NUP + 14) / 16) * 16 + 1, NP = 2 * NUP;
    static constexpr int RD = sizeof(R) == 8 ? 1 : 2;  // R values per double
    int xs, qc, qm, sv, ph, phn, pw, buf, total;
    __host__ __device__ explicit MvLayout(int nq) {
        xs = 0;
        qc = xs + 3 * NP;
        qm = qc + (5 * nq + RD - 1) / RD;
        sv = qm + 2 * nq;
        ph = sv + (NUP * SP + RD - 1) / RD;
        phn = ph + (2 * NP * nq + RD - 1) / RD;
        pw = phn + (2 * nq + RD - 1) / RD;      // exp(i m theta_d), d < 3, m = 0..4
        buf = pw + (2 * 15 + RD - 1) / RD;
        total = buf + (3 * SP + RD - 1) / RD;
    }
};

template <int NUP, typename R>
size_t mv_lds_bytes(int nq) { return sizeof(double) * MvLayout<NUP, R>(nq).total; }

// exp(i m theta_d) for m = 0..4 of one coordinate, R-typed
template <typename R>
__device__ __forceinline__ void powers(const DevSys &S, double x, R (&c)[5], R (&s)[5]) {
    R s1, c1;
    if constexpr (sizeof(R) == 8) {
        sincos(x * S.twopi_l, &s1, &c1);
    } else {
        sincosf(float(x * S.twopi_l), &s1, &c1);
    }
    c[0] = R(1);
    s[0] = R(0);
    c[1] = c1;
    s[1] = s1;
#pragma unroll
    for (int k = 2; k < 5; ++k) {
        c[k] = c[k - 1] * c1 - s[k - 1] * s1;
        s[k] = s[k - 1] * c1 + c[k - 1] * s1;
    }
}

// product of the three factors exp(i m_d theta_d), read from a power table pw[d][m]
template <typename R>
__device__ __forceinline__ typename Vec2<R>::T qphase(const typename Vec2<R>::T *pw,
                                                      const int4 m) {
    const int mm[3] = {m.x, m.y, m.z};
    R c[3], s[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int am = mm[d] < 0 ? -mm[d] : mm[d];
        const typename Vec2<R>::T v = pw[5 * d + am];
        c[d] = v.x;
        s[d] = mm[d] < 0 ? -v.y : v.y;
    }
    const R cxy = c[0] * c[1] - s[0] * s[1], sxy = s[0] * c[1] + c[0] * s[1];
    typename Vec2<R>::T z;
    z.x = cxy * c[2] - sxy * s[2];
    z.y = sxy * c[2] + cxy * s[2];
    return z;
}

// moves m0..m1-1 of walker w: positions and RNG state from xin/irn_in, the stored inverse
// sinv_g; results to xout/irn_out (LDS at sm, MvLayout)
template <int NUP, typename R>
__device__ __forceinline__ void sm_moves_walker(const DevSys &S, double *sm, int w, int tid,
                                                const double *__restrict__ xin, double *xout,
                                                const uint64_t *irn_in, uint64_t *irn_out,
                                                char *sinv_g, int m0, int m1, const DevStep &T) {
    using R2 = typename Vec2<R>::T;
    const MvLayout<NUP, R> X(S.nq);
    constexpr int M = 6 * NUP, NP = 2 * NUP, SP = MvLayout<NUP, R>::SP;
    const int nq = S.nq;
    double *xs = sm + X.xs;
    R *qc = reinterpret_cast<R *>(sm + X.qc);
    int4 *qm = reinterpret_cast<int4 *>(sm + X.qm);
    R *sv = reinterpret_cast<R *>(sm + X.sv);
    R2 *ph = reinterpret_cast<R2 *>(sm + X.ph);
    R2 *phn = reinterpret_cast<R2 *>(sm + X.phn);
    R2 *pw = reinterpret_cast<R2 *>(sm + X.pw);
    R *phinew = reinterpret_cast<R *>(sm + X.buf);
    R *colbuf = phinew + SP;
    R *fbuf = colbuf + SP;

    for (int q = tid; q < nq; q += 64) {
        const int4 m = S.qm[q];
        const double wq = S.wq[q], tp = S.twopi_l;
        qm[q] = m;
        qc[5 * q] = R(wq);
        qc[5 * q + 1] = R(-wq * tp * m.x);
        qc[5 * q + 2] = R(-wq * tp * m.y);
        qc[5 * q + 3] = R(-wq * tp * m.z);
    }
    for (int t = tid; t < M; t += 64) xs[t] = wrapd(xin[size_t(w) * M + t], S);
    {
        const R *g = reinterpret_cast<const R *>(sinv_g + size_t(w) * sinv_bytes(S));
        for (int e = tid; e < NUP * NUP; e += 64) {
            const int j = e / NUP, i = e - j * NUP;
            sv[j * SP + i] = g[e];
        }
    }
    wave_sync();
    // q-major phases ph[q * NP + p] of every particle (lanes over particles)
    for (int p = tid; p < NP; p += 64) {
        R c[3][5], sn[3][5];
#pragma unroll
        for (int d = 0; d < 3; ++d) powers<R>(S, xs[3 * p + d], c[d], sn[d]);
        for (int q = 0; q < nq; ++q) {
            const int4 m = qm[q];
            const int mm[3] = {m.x, m.y, m.z};
            R cq[3], sq[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                // select kept apart from an indexed private load (no scratch)
                const int am = mm[d] < 0 ? -mm[d] : mm[d];
                R cc = c[d][0], ss = sn[d][0];
#pragma unroll
                for (int k = 1; k < 5; ++k) {
                    const bool t = am == k;
                    cc = t ? c[d][k] : cc;
                    ss = t ? sn[d][k] : ss;
                    __builtin_amdgcn_sched_barrier(0);
                }
                cq[d] = cc;
                sq[d] = mm[d] < 0 ? -ss : ss;
            }
            const R cxy = cq[0] * cq[1] - sq[0] * sq[1], sxy = sq[0] * cq[1] + cq[0] * sq[1];
            R2 z;
            z.x = cxy * cq[2] - sxy * sq[2];
            z.y = sxy * cq[2] + cxy * sq[2];
            ph[q * NP + p] = z;
        }
    }
    wave_sync();

    const double sig = S.sigma, sig2 = S.sig2, half_isig2 = 0.5 / S.sig2;
    const R alpha0 = R(S.alpha0);
    uint64_t s = irn_in[w];
    int nacc = 0;
    const int r0 = tid / NUP, c0 = tid - (tid / NUP) * NUP;
    for (int ip = m0; ip < m1; ++ip) {
        const bool up = ip < NUP;
        const int a = up ? ip : ip - NUP;
        const double xa[3] = {xs[3 * ip], xs[3 * ip + 1], xs[3 * ip + 2]};
        // old gradient: lanes b < NUP over the opposite-spin partners
        R go0 = 0, go1 = 0, go2 = 0, uo = 0, wgt = 0;
        int pb = 0;
        if (tid < NUP) {
            const int b = tid;
            pb = up ? NUP + b : b;
            wgt = up ? sv[b * SP + a] : sv[a * SP + b];
            R dx[3];
            const R ro = pair_dist<R>(S, xa, xs + 3 * pb, dx);
            // table values first, so that their loads overlap the plane-wave sum
            R bv = 0, db = 0, uu = 0, du = 0;
            const R rj = ro > R(1.0e-14) ? ro : R(1.0e-14);
            if (!S.nobeta && ro <= R(S.rangeb)) tab2<R, 0, 1>(S.bt, S.bt32, R(S.scaleb) * ro, S.ntab, bv, db);
            if (rj < R(S.range)) tab2<R, 1, 2>(S.jt, S.jt32, R(S.scale) * rj, S.ntab, uu, du);
#pragma unroll 4
            for (int q = 0; q < nq; ++q) {
                const R2 u = ph[q * NP + ip], v = ph[q * NP + pb];
                const R sab = u.y * v.x - u.x * v.y;
                go0 += qc[5 * q + 1] * sab;
                go1 += qc[5 * q + 2] * sab;
                go2 += qc[5 * q + 3] * sab;
            }
            const R t = db * rcp_r<R>(ro), tj = du * rcp_r<R>(rj);
            uo = uu;
            // determinant part weighted by the inverse, plus the Jastrow part
            go0 = wgt * (go0 + t * dx[0]) + dx[0] * tj;
            go1 = wgt * (go1 + t * dx[1]) + dx[1] * tj;
            go2 = wgt * (go2 + t * dx[2]) + dx[2] * tj;
        }
        const double go[3] = {double(wave_sum(go0)), double(wave_sum(go1)), double(wave_sum(go2))};
        // gaussian(3) then the acceptance draw: five ran1 draws per move
        double u[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            s = (s * kM1 + kA1) & kMask48;
            u[k] = double(s | 1ull) * 0x1p-48;
        }
        double g3[3];
        {
            R s1, c1, s3, c3;
            if constexpr (sizeof(R) == 8) {
                sincos(2.0 * kPi * u[0], &s1, &c1);
                sincos(2.0 * kPi * u[2], &s3, &c3);
            } else {
                sincosf(float(2.0 * kPi * u[0]), &s1, &c1);
                sincosf(float(2.0 * kPi * u[2]), &s3, &c3);
            }
            const R r12 = sqrt(R(-2) * log(R(u[1]))), r34 = sqrt(R(-2) * log(R(u[3])));
            g3[0] = double(s1 * r12);
            g3[1] = double(c1 * r12);
            g3[2] = double(c3 * r34);
        }
        double dro[3], dxm[3], xn[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            dro[d] = go[d] * sig2;
            if (S.cutdrift) dro[d] = fmin(S.dcut, fmax(-S.dcut, dro[d]));
            dxm[d] = dro[d] + sig * g3[d];
            xn[d] = wrapr(xa[d] + dxm[d], S);
        }
        // phases of the proposed position: powers per axis (lanes 0..2), then lanes over q
        if (tid < 3) {
            R c[5], sn[5];
            powers<R>(S, xn[tid], c, sn);
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                pw[5 * tid + k].x = c[k];
                pw[5 * tid + k].y = sn[k];
            }
        }
        wave_sync();
        if (tid < nq) phn[tid] = qphase<R>(pw, qm[tid]);
        wave_sync();
        // pair values and gradients at the proposed position
        double rho_part = 0.0;
        R gn0 = 0, gn1 = 0, gn2 = 0, jn0 = 0, jn1 = 0, jn2 = 0, dU = 0;
        if (tid < NUP) {
            R dx[3];
            const R rn = pair_dist<R>(S, xn, xs + 3 * pb, dx);
            R bv = 0, db = 0, uu = 0, du = 0;
            const R rj = rn > R(1.0e-14) ? rn : R(1.0e-14);
            if (!S.nobeta && rn <= R(S.rangeb)) tab2<R, 0, 1>(S.bt, S.bt32, R(S.scaleb) * rn, S.ntab, bv, db);
            if (rj < R(S.range)) tab2<R, 1, 2>(S.jt, S.jt32, R(S.scale) * rj, S.ntab, uu, du);
            R f = 0, g0 = 0, g1 = 0, g2 = 0;
#pragma unroll 4
            for (int q = 0; q < nq; ++q) {
                const R2 un = phn[q], v = ph[q * NP + pb];
                const R cab = un.x * v.x + un.y * v.y;
                const R sab = un.y * v.x - un.x * v.y;
                f += qc[5 * q] * cab;
                g0 += qc[5 * q + 1] * sab;
                g1 += qc[5 * q + 2] * sab;
                g2 += qc[5 * q + 3] * sab;
            }
            f += alpha0 + bv;
            const R t = db * rcp_r<R>(rn);
            g0 += t * dx[0];
            g1 += t * dx[1];
            g2 += t * dx[2];
            phinew[tid] = f;
            rho_part = double(wgt) * double(f);
            gn0 = wgt * g0;
            gn1 = wgt * g1;
            gn2 = wgt * g2;
            const R tj = du * rcp_r<R>(rj);
            jn0 = dx[0] * tj;
            jn1 = dx[1] * tj;
            jn2 = dx[2] * tj;
            dU = uu - uo;
        }
        const double rho = wave_sum(rho_part);  // FP64 sum: the node test uses its sign
        const double gn[3] = {double(wave_sum(gn0)), double(wave_sum(gn1)), double(wave_sum(gn2))};
        const double jn[3] = {double(wave_sum(jn0)), double(wave_sum(jn1)), double(wave_sum(jn2))};
        const double dlogj = double(wave_sum(dU));
        bool accept = !(S.fn && !(rho > 0.0)) && rho != 0.0;
        if (accept) {
            double logp = 2.0 * (log(fabs(rho)) + dlogj);
            const double irho = 1.0 / rho;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double drn = (gn[d] * irho + jn[d]) * sig2;
                if (S.cutdrift) drn = fmin(S.dcut, fmax(-S.dcut, drn));
                logp += (dro[d] + drn) * (dro[d] - drn - 2.0 * dxm[d]) * half_isig2;
            }
            accept = u[4] < exp(fmin(0.0, logp));
        }
        if (accept) {
            ++nacc;
            const R irr = R(1.0 / rho);
            // capture column a (up mover) or row a (down mover) of the inverse
            if (tid < NUP) colbuf[tid] = up ? sv[tid * SP + a] : sv[a * SP + tid];
            wave_sync();
            if (tid < NUP) {
                const int k = tid;
                R z = 0;
                if (up) {
#pragma unroll
                    for (int b = 0; b < NUP; ++b) z += phinew[b] * sv[b * SP + k];
                } else {
#pragma unroll
                    for (int b = 0; b < NUP; ++b) z += sv[k * SP + b] * phinew[b];
                }
                fbuf[k] = (z - (k == a ? R(1) : R(0))) * irr;
            }
            wave_sync();
            // rank-1 update over the n x n elements, 64 at a time; (r, c) of element
            // e0 + tid advanced by 64 = kq n + kr without a division
            {
                constexpr int kq = 64 / NUP, kr = 64 % NUP;
                int r = r0, c = c0;
#pragma unroll
                for (int e0 = 0; e0 < NUP * NUP; e0 += 64) {
                    if (e0 + tid < NUP * NUP)
                        sv[r * SP + c] -= up ? colbuf[r] * fbuf[c] : fbuf[r] * colbuf[c];
                    r += kq;
                    c += kr;
                    if (c >= NUP) {
                        c -= NUP;
                        ++r;
                    }
                }
            }
            if (tid < 3) xs[3 * ip + tid] = xn[tid];
            if (tid < nq) ph[tid * NP + ip] = phn[tid];
        }
        wave_sync();
    }
    for (int t = tid; t < M; t += 64) xout[size_t(w) * M + t] = xs[t];
    if (tid == 0) {
        irn_out[w] = s;
        T.nacc[w] = (m0 == 0 ? 0 : T.nacc[w]) + nacc;
        atomicAdd(reinterpret_cast<unsigned long long *>(&T.moves[0]), (unsigned long long)nacc);
        atomicAdd(reinterpret_cast<unsigned long long *>(&T.moves[1]),
                  (unsigned long long)(m1 - m0));
    }
}

// FP64 rebuild of the inverse of positions x (every krefresh moves)
template <int NUP>
__device__ __forceinline__ void sm_refresh_walker(const DevSys &S, double *sm, int w, int tid,
                                                  const double *__restrict__ x, char *sinv_g) {
    const Layout<NUP> L(S.nq);
    constexpr int M = 6 * NUP;
    load_q<NUP>(S, sm, L, tid);
    for (int t = tid; t < M; t += 64) sm[L.xs + t] = x[size_t(w) * M + t];
    wave_sync();
    const Tile<NUP> Tl(tid);
    const DetInfo D = build_inverse<NUP>(S, sm, L, Tl, tid);
    store_sinv<NUP>(S, sm, L, D.singular, sinv_g + size_t(w) * sinv_bytes(S), tid);
}

// end of the sweep: FP64 hpsi of the final positions B.x, the fresh inverse into B.sinv,
// weight against the energy of A, estimators and branch draw (state after the moves in
// B.irn)
template <int NUP>
__device__ __forceinline__ void sm_end_walker(const DevSys &S, double *sm, int w, int tid,
                                              const DevState &A, const DevState &B,
                                              const DevStep &T, int opcomp) {
    const Layout<NUP> L(S.nq);
    constexpr int M = 6 * NUP;
    load_q<NUP>(S, sm, L, tid);
    for (int t = tid; t < M; t += 64) sm[L.xs + t] = B.x[size_t(w) * M + t];
    wave_sync();
    const Tile<NUP> Tl(tid);
    const DetInfo D = build_inverse<NUP>(S, sm, L, Tl, tid);
    store_sinv<NUP>(S, sm, L, D.singular, B.sinv + size_t(w) * sinv_bytes(S), tid);
    grad_lap<NUP>(S, sm, L, Tl, D, tid);
    finish_walker<NUP>(S, sm, L, A, B, T, w, tid, B.irn[w], opcomp);
}

}  // namespace dmc
