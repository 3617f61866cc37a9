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

// natural log of a positive normal FP64 number: x = m 2^e with m in [1/sqrt 2, sqrt 2)
// (frexp), ln m = 2 atanh(z) = 2 (z + z^3/3 + ... + z^19/19) with z = (m - 1)/(m + 1),
// |z| <= 0.1716 (truncation below 1e-18), and e ln 2 with ln 2 split in a high part exact in
// 32 bits and a low part
__device__ __forceinline__ double log_pos(double x) {
    int e;
    double m = frexp(x, &e);  // [0.5, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        --e;
    }
    const double d = m + 1.0;
    double r = __builtin_amdgcn_rcp(d);  // about 1 ulp, then two Newton steps
    r = fma(fma(-d, r, 1.0), r, r);
    r = fma(fma(-d, r, 1.0), r, r);
    const double z = (m - 1.0) * r, z2 = z * z;
    double p = fma(z2, 1.0 / 19.0, 1.0 / 17.0);
    p = fma(z2, p, 1.0 / 15.0);
    p = fma(z2, p, 1.0 / 13.0);
    p = fma(z2, p, 1.0 / 11.0);
    p = fma(z2, p, 1.0 / 9.0);
    p = fma(z2, p, 1.0 / 7.0);
    p = fma(z2, p, 1.0 / 5.0);
    p = fma(z2, p, 1.0 / 3.0);
    const double lm = 2.0 * fma(z * z2, p, z);
    const double fe = double(e);
    return fma(fe, 0.693147180369123816490, fma(fe, 1.90821492927058770002e-10, lm));
}

// exp(y) for y <= 0 in FP64: y = k ln 2 + r with k = rint(y / ln 2), r = y - k ln 2 from the
// split ln 2 (Cody-Waite), |r| <= ln 2 / 2; exp(r) by its Taylor polynomial to r^13
// (truncation below 5e-18) and the factor 2^k by ldexp; below -1100 the result is 0
__device__ __forceinline__ double exp_nonpos(double y) {
    y = fmax(y, -1100.0);
    const double k = rint(y * 1.44269504088896340736);
    const double r = fma(-k, 1.90821492927058770002e-10, fma(-k, 0.693147180369123816490, y));
    double p = fma(r, 1.0 / 6227020800.0, 1.0 / 479001600.0);  // 1/13!, 1/12!
    p = fma(r, p, 1.0 / 39916800.0);
    p = fma(r, p, 1.0 / 3628800.0);
    p = fma(r, p, 1.0 / 362880.0);
    p = fma(r, p, 1.0 / 40320.0);
    p = fma(r, p, 1.0 / 5040.0);
    p = fma(r, p, 1.0 / 720.0);
    p = fma(r, p, 1.0 / 120.0);
    p = fma(r, p, 1.0 / 24.0);
    p = fma(r, p, 1.0 / 6.0);
    p = fma(r, p, 0.5);
    p = fma(r, p, 1.0);
    p = fma(r, p, 1.0);
    return ldexp(p, int(k));
}

// natural log (positive argument) and exp (non-positive argument, as the moves use them):
// the polynomials above in FP64, the hardware log2/exp2 (about 1 ulp) scaled by ln 2 /
// log2 e in FP32
template <typename R>
__device__ __forceinline__ R log_r(R x) {
    if constexpr (sizeof(R) == 8) return log_pos(x);
    else return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
}
template <typename R>
__device__ __forceinline__ R exp_r(R x) {
    if constexpr (sizeof(R) == 8) return exp_nonpos(x);
    else return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
}

// sin and cos of 2 pi t for a turn count t given in FP64: FP64 reduction to the nearest
// turn, then the FP64 or FP32 polynomials above
template <typename R>
__device__ __forceinline__ void sincos_2pi(double t, R &s, R &c) {
    t -= rint(t);
    if constexpr (sizeof(R) == 8) sincos_turns(t, s, c);
    else sincos_turns(float(t), s, c);
}

// the value of v in lane l, wave-uniform
__device__ __forceinline__ float lane_value(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_value(double v, int l) { return readlane_d(v, l); }
__device__ __forceinline__ uint64_t lane_value(uint64_t v, int l) {
    const unsigned lo = __builtin_amdgcn_readlane(unsigned(v & 0xffffffffull), l);
    const unsigned hi = __builtin_amdgcn_readlane(unsigned(v >> 32), l);
    return (uint64_t(hi) << 32) | lo;
}

// distance of the pair (a, b): the difference of the FP64 positions, then the minimum image
// in FP64 (R = double) or, rounded to FP32, in FP32 (the wrapped difference then carries an
// absolute error of about ulp(L))
template <typename R>
__device__ __forceinline__ R pair_dist(const DevSys &S, const double *xa, const double *xb,
                                       R (&dx)[3]) {
    R r2 = R(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if constexpr (sizeof(R) == 8) {
            dx[d] = wrapr(xa[d] - xb[d], S);
        } else {
            const float t = float(xa[d] - xb[d]);
            dx[d] = fmaf(-float(S.el), rintf(t * float(S.eli)), t);
        }
        r2 += dx[d] * dx[d];
    }
    return sqrt_r<R>(r2);
}

template <typename R>
struct Vec4;
template <>
struct Vec4<double> {
    using T = double4;
};
template <>
struct Vec4<float> {
    using T = float4;
};

// The pair table of the moves: record i = {beta, beta', u, u'} at r = i / scale (the beta
// and Jastrow tables share the grid [0, L/2] of ntab intervals), interpolated with the
// 4-point Lagrange weights of coshpot.f90:172-192.  Beyond L/2 every field is zero.
template <typename R>
__device__ __forceinline__ typename Vec4<R>::T pair_tab(const DevSys &S, R r) {
    using R4 = typename Vec4<R>::T;
    R4 o;
    o.x = o.y = o.z = o.w = R(0);
    if (!(r <= R(S.rangeb))) return o;
    const R s = R(S.scaleb) * r;
    const int ntab = S.ntab;
    int idx = int(s);
    idx = idx > ntab - 2 ? ntab - 2 : idx;
    idx = idx < 1 ? 1 : idx;
    R w0, w1, w2, w3;
    lagrange_weights<R>(s - R(idx), w0, w1, w2, w3);
    const R4 *tab;
    if constexpr (sizeof(R) == 8) tab = S.mt64;
    else tab = S.mt32;
    const R4 a = tab[idx - 1], b = tab[idx], c = tab[idx + 1], d = tab[idx + 2];
    o.x = w0 * a.x + w1 * b.x + w2 * c.x + w3 * d.x;
    o.y = w0 * a.y + w1 * b.y + w2 * c.y + w3 * d.y;
    o.z = w0 * a.z + w1 * b.z + w2 * c.z + w3 * d.z;
    o.w = w0 * a.w + w1 * b.w + w2 * c.w + w3 * d.w;
    return o;
}

// Pair terms of the mover at xa with the partner at xp: the minimum-image difference dx,
// beta'(r)/r, u'(r)/r and u(r) (r floored at 1e-14 for the Jastrow part, coshpot.f90:167)
// and, if wanted, beta(r)
template <typename R>
__device__ __forceinline__ void pair_terms(const DevSys &S, const double *xa, const double *xp,
                                           R (&dx)[3], R &tb, R &tj, R &u, R &bv) {
    const R r = pair_dist<R>(S, xa, xp, dx);
    const typename Vec4<R>::T v = pair_tab<R>(S, r);
    const R rj = r > R(1.0e-14) ? r : R(1.0e-14);
    bv = v.x;
    tb = v.y * rcp_r<R>(r);
    tj = v.w * rcp_r<R>(rj);
    u = v.z;
}

// v(lane) + v(lane ^ 32) in every lane (the two half-waves exchanged by a permlane swap)
__device__ __forceinline__ float half_sum(float v) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v),
                                                    false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ double half_sum(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = unsigned(b & 0xffffffffll), hi = unsigned(b >> 32);
    const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return dbl_of(int(l[0]), int(h[0])) + dbl_of(int(l[1]), int(h[1]));
}

// LDS of the move kernel (offsets in doubles, every area 16-byte aligned)
template <int NUP, typename R>
struct MvLayout {
    static constexpr int NP = 2 * NUP;
    static constexpr int RD = sizeof(R) == 8 ? 1 : 2;  // R values per double
    // the column broadcast is split in two halves (partners [0, HB) and [HB, NUP)), each
    // starting on a 16-byte boundary
    static constexpr int HB = (NUP + 1) / 2, HBP = (HB + 3) / 4 * 4;
    // row stride (complex values) of the phase planes P[q][p] and of the transposition
    // area: odd, so that lanes reading one partner at different q, or one row of the
    // transposition each, hit distinct banks; in FP32 also at least 2 HB, so that the two
    // half-waves of the projection read whole halves (the columns past NUP are zero)
    static constexpr int ST = sizeof(R) == 8 ? (NUP | 1) : ((2 * HB) | 1);
    static constexpr int NB = (NUP + 3) / 4 * 4;
    static __host__ __device__ constexpr int dbl(int nr) { return ((nr + RD - 1) / RD + 1) / 2 * 2; }
    int xs, pu, pd, qk, qm, wb, fb, ub, jmp, total;
    __host__ __device__ explicit MvLayout(int nq) {
        xs = 0;                                  // positions, [NP][3] FP64
        const int plane = dbl(2 * nq * ST);      // phases of one spin, [nq][ST] complex
        const int trans = dbl(NUP * ST);         // transposition of the inverse (overlay)
        pu = xs + 3 * NP;
        pd = pu + plane;
        const int pend = pu + (2 * plane > trans ? 2 * plane : trans);
        qk = pend;                               // per q: w_q, -w_q q (R)
        qm = qk + dbl(4 * nq);                   // integer q vectors (int4)
        wb = qm + 2 * nq;                        // inverse column, two halves (R)
        fb = wb + dbl(2 * HBP);                  // new pair values (R)
        ub = fb + dbl(NB);                       // w_q exp(i q.x') (R)
        jmp = ub + dbl(2 * nq + 4);              // jump-ahead a, c of draws 1..5 (uint64)
        total = jmp + 10;
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

// phases exp(i q.r) of every particle into the planes pu (up) / pd (down), [q][ST]
template <int NUP, typename R>
__device__ __forceinline__ void move_phases(const DevSys &S, const double *xs, const int4 *qm,
                                            typename Vec2<R>::T *pu, typename Vec2<R>::T *pd,
                                            int tid) {
    using R2 = typename Vec2<R>::T;
    constexpr int NP = 2 * NUP, ST = MvLayout<NUP, R>::ST, PADC = ST - NUP;
    const int nq = S.nq;
    if constexpr (PADC > 0) {
        for (int t = tid; t < 2 * nq * PADC; t += 64) {
            const int row = t / PADC, col = NUP + (t - row * PADC);
            R2 z;
            z.x = z.y = R(0);
            (row < nq ? pu : pd)[(row < nq ? row : row - nq) * ST + col] = z;
        }
    }
    for (int p = tid; p < NP; p += 64) {
        R2 *pl = p < NUP ? pu + p : pd + (p - NUP);
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
            pl[q * ST] = z;
        }
    }
}

// Moves m0..m1-1 of walker w: positions and RNG state from xin/irn_in, the stored inverse
// sinv_g (dense [j][i] = Sinv(j, i)); results to xout/irn_out (LDS at sm, MvLayout).
//
// Mover a with the n opposite-spin partners b.  Lane k < n holds column k of the inverse
// in registers while up particles move (Sinv(b, k), b = 0..n-1) and row k while down
// particles move (Sinv(k, b)): in both halves the mover's determinant weights are
// w_b = "column a" of what the lanes hold, and the Sherman-Morrison step is the same
// rank-1 update.  The plane-wave parts of the drift at the old and the proposed position
// and of the ratio go through the projections X_q = sum_b w_b exp(i q.r_b) (64 lanes: one
// real component of one q per lane, the partners split over the two half-waves):
//   sum_b w_b grad_a cos(q.(r_a - r_b)) = -q Im(exp(i q.r_a) conj(X_q)).
// Lanes over the partners add the radial (beta) and Jastrow terms from one table lookup
// per pair, and the new pair values phi(r'_a - r_b) that the update needs.  The pair terms
// of the next mover at its current position are formed while the current move is
// evaluated: moves of one spin never change them.
template <int NUP, typename R, int NQC>
__device__ __forceinline__ void sm_moves_walker(const DevSys &S, double *sm, int w, int tid,
                                                const double *__restrict__ xin, double *xout,
                                                const uint64_t *irn_in, uint64_t *irn_out,
                                                char *sinv_g, int m0, int m1, const DevStep &T) {
    using R2 = typename Vec2<R>::T;
    using L = MvLayout<NUP, R>;
    const L X(S.nq);
    constexpr int M = 6 * NUP, ST = L::ST, HB = L::HB, HBP = L::HBP;
    const int nq = S.nq, no = 2 * nq;
    double *xs = sm + X.xs;
    R2 *pu = reinterpret_cast<R2 *>(sm + X.pu);
    R2 *pd = reinterpret_cast<R2 *>(sm + X.pd);
    R *tr = reinterpret_cast<R *>(sm + X.pu);
    R *qk = reinterpret_cast<R *>(sm + X.qk);
    int4 *qm = reinterpret_cast<int4 *>(sm + X.qm);
    R *wb = reinterpret_cast<R *>(sm + X.wb);
    R *fb = reinterpret_cast<R *>(sm + X.fb);
    R *ub = reinterpret_cast<R *>(sm + X.ub);

    for (int q = tid; q < nq; q += 64) {
        const int4 m = S.qm[q];
        const double wq = S.wq[q], tp = S.twopi_l;
        qm[q] = m;
        qk[4 * q] = R(wq);
        qk[4 * q + 1] = R(-wq * tp * m.x);
        qk[4 * q + 2] = R(-wq * tp * m.y);
        qk[4 * q + 3] = R(-wq * tp * m.z);
    }
    for (int t = tid; t < M; t += 64) xs[t] = wrapd(xin[size_t(w) * M + t], S);
    wave_sync();
    move_phases<NUP, R>(S, xs, qm, pu, pd, tid);

    // projection lanes: output o = (lane & 31) + 32 c is component (o & 1) of q = o >> 1
    const int h = tid >> 5, ol = tid & 31;
    R cw[NQC], cg[NQC][3];
    int4 cm[NQC];
#pragma unroll
    for (int c = 0; c < NQC; ++c) {
        const int o = ol + 32 * c, q = o >> 1;
        const bool ok = o < no;
        cw[c] = ok ? qk[4 * q] : R(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) cg[c][d] = ok ? qk[4 * q + 1 + d] : R(0);
        cm[c] = ok ? qm[q] : make_int4(0, 0, 0, 0);
    }

    // the inverse: column (up movers) or row (down movers) k = lane
    bool colmode = m0 < NUP;
    R Ac[NUP];
    {
        const R *g = reinterpret_cast<const R *>(sinv_g + size_t(w) * sinv_stride(S));
        const int k = tid < NUP ? tid : 0;
#pragma unroll
        for (int b = 0; b < NUP; ++b) Ac[b] = g[colmode ? b * NUP + k : k * NUP + b];
    }
    // partner position of lane b (the partners do not move while one spin moves)
    double xp[3];
    {
        const int pb = (colmode ? NUP : 0) + (tid < NUP ? tid : 0);
#pragma unroll
        for (int d = 0; d < 3; ++d) xp[d] = xs[3 * pb + d];
    }
    wave_sync();

    for (int t = tid; t < 2 * HBP; t += 64) wb[t] = R(0);  // padding of the second half stays 0
    const double sig = S.sigma, sig2 = S.sig2, half_isig2 = 0.5 / S.sig2;
    // drift cut (an infinite cut when the drift is not cut)
    const double dcut = S.cutdrift ? S.dcut : __builtin_huge_val();
    const R alpha0 = R(S.alpha0);
    uint64_t s = irn_in[w];
    int nacc = 0;
    // jump-ahead coefficients of this lane's draws (state after t + 1 draws = ja1[t] s +
    // jc1[t]): lane l < 2 draws 2l, 2l + 1; lane 2 draw 4; other lanes none (zero, the draw
    // then stays finite and unused).  Staged in LDS, read where used.
    uint64_t *jmp = reinterpret_cast<uint64_t *>(sm + X.jmp);
    if (tid < 5) {
        jmp[tid] = S.ja1[tid];
        jmp[5 + tid] = S.jc1[tid];
    }
    wave_sync();
    const int jl0 = tid < 3 ? 2 * tid : -1, jl1 = tid < 2 ? 2 * tid + 1 : -1;
    // pair terms of the mover at its current position (lanes < NUP)
    R ob = 0, oj = 0, ou = 0, odx[3] = {0, 0, 0};
    bool have_old = false;
    for (int ip = m0; ip < m1; ++ip) {
        if (ip == NUP && colmode) {
            // down movers from here: rows of the inverse into the lanes through LDS (over
            // the phase planes, rebuilt afterwards), the up particles become the partners
            if (tid < NUP) {
#pragma unroll
                for (int b = 0; b < NUP; ++b) tr[b * ST + tid] = Ac[b];
            }
            wave_sync();
            if (tid < NUP) {
#pragma unroll
                for (int b = 0; b < NUP; ++b) Ac[b] = tr[tid * ST + b];
            }
            wave_sync();
            move_phases<NUP, R>(S, xs, qm, pu, pd, tid);
            const int pb = tid < NUP ? tid : 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) xp[d] = xs[3 * pb + d];
            colmode = false;
            have_old = false;
            wave_sync();
        }
        const bool up = ip < NUP;
        const int a = up ? ip : ip - NUP;
        R2 *Pm = up ? pu : pd;        // phases of the mover's spin
        const R2 *Pp = up ? pd : pu;  // phases of the partners
        const double xa[3] = {xs[3 * ip], xs[3 * ip + 1], xs[3 * ip + 2]};
        // the mover's determinant weights w_b to every lane
        if (tid == a) {
#pragma unroll
            for (int b = 0; b < NUP; ++b) wb[b < HB ? b : HBP + b - HB] = Ac[b];
        }
        wave_sync();
        R wgt = R(0);
        if (tid < NUP) wgt = wb[tid < HB ? tid : HBP + tid - HB];
        if (!have_old && tid < NUP) {
            R bv;
            pair_terms<R>(S, xa, xp, odx, ob, oj, ou, bv);
        }
        // projections X_o, partners [h HB, h HB + HB) in half-wave h
        R Xo[NQC];
#pragma unroll
        for (int c = 0; c < NQC; ++c) {
            const int o = ol + 32 * c;
            R acc = R(0);
            if (o < no) {
                const R *pl = reinterpret_cast<const R *>(Pp + (o >> 1) * ST + h * HB) + (o & 1);
                const R *wh = wb + h * HBP;
                if constexpr (ST >= 2 * HB) {
                    // padded rows (FP32): both half-waves run all HB terms
#pragma unroll
                    for (int bb = 0; bb < HB; ++bb) acc += wh[bb] * pl[2 * bb];
                } else {
                    // unpadded rows (FP64, where the padding would cost occupancy)
#pragma unroll
                    for (int bb = 0; bb < HB; ++bb)
                        if (h * HB + bb < NUP) acc += wh[bb] * pl[2 * bb];
                }
            }
            Xo[c] = half_sum(acc);
        }
        // drift at the current position
        R g0 = R(0), g1 = R(0), g2 = R(0);
        if (tid < NUP) {
            const R f = wgt * ob + oj;
            g0 = f * odx[0];
            g1 = f * odx[1];
            g2 = f * odx[2];
        }
        if (h == 0) {
#pragma unroll
            for (int c = 0; c < NQC; ++c) {
                const int o = ol + 32 * c;
                if (o < no) {
                    const R2 u = Pm[(o >> 1) * ST + a];
                    const R t = (o & 1) ? -u.x * Xo[c] : u.y * Xo[c];
                    g0 += cg[c][0] * t;
                    g1 += cg[c][1] * t;
                    g2 += cg[c][2] * t;
                }
            }
        }
        if constexpr (sizeof(R) == 4) wave_sum2(g0, g1);
        else {
            g0 = wave_sum(g0);
            g1 = wave_sum(g1);
        }
        const R go[3] = {g0, g1, wave_sum(g2)};
        // gaussian(3) then the acceptance draw: five ran1 draws per move, formed by jump-ahead
        // from the walker state in separate lanes (lane l < 2: the Box-Muller pair of draws
        // 2l+1, 2l+2; lane 2: the fifth draw, the walker's next state) and broadcast
        R g3[3];
        double u4;
        {
            const uint64_t ja0 = jl0 >= 0 ? jmp[jl0] : 0, jc0 = jl0 >= 0 ? jmp[5 + jl0] : 0;
            const uint64_t ja1 = jl1 >= 0 ? jmp[jl1] : 0, jc1 = jl1 >= 0 ? jmp[5 + jl1] : 0;
            const uint64_t sa = (ja0 * s + jc0) & kMask48, sb = (ja1 * s + jc1) & kMask48;
            const double ua = double(sa | 1ull) * 0x1p-48, ub = double(sb | 1ull) * 0x1p-48;
            R sn, cs;
            if constexpr (sizeof(R) == 8) sincos_turns(ua - rint(ua), sn, cs);
            else sincos_turns(float(ua - rint(ua)), sn, cs);
            const R rad = sqrt_r<R>(R(-2) * log_r<R>(R(ub)));
            const R gx = sn * rad, gy = cs * rad;
            g3[0] = lane_value(gx, 0);
            g3[1] = lane_value(gy, 0);
            g3[2] = lane_value(gy, 1);
            s = lane_value(sa, 2);
            u4 = double(s | 1ull) * 0x1p-48;
        }
        // drift and displacement in R (FP32 for the mixed moves, whose gradient and draws are
        // FP32), added to the FP64 position
        R dro[3], dxm[3];
        double xn[3];
        {
            const R rs2 = R(sig2), rsig = R(sig), rcut = R(dcut);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                dro[d] = fmin(rcut, fmax(-rcut, go[d] * rs2));
                dxm[d] = fma(rsig, g3[d], dro[d]);
                xn[d] = wrapr(xa[d] + double(dxm[d]), S);
            }
        }
        // phases of the proposed position, one q per lane pair: the argument reduced to
        // a fraction of a turn in FP64
        R unc[NQC], uns[NQC];
#pragma unroll
        for (int c = 0; c < NQC; ++c) {
            const int o = ol + 32 * c;
            const int4 m = cm[c];
            R sn, cs;
            sincos_2pi<R>((m.x * xn[0] + m.y * xn[1] + m.z * xn[2]) * S.eli, sn, cs);
            unc[c] = cs;
            uns[c] = sn;
            if (h == 0 && o < no) ub[o] = cw[c] * ((o & 1) ? sn : cs);
        }
        wave_sync();
        // pair values at the proposed position (lanes < NUP)
        double rho_part = 0.0;
        R gn0 = 0, gn1 = 0, gn2 = 0, jn0 = 0, jn1 = 0, jn2 = 0, dU = 0, fnew = 0;
        if (tid < NUP) {
            R dx[3], tb, tj, uu, bv;
            pair_terms<R>(S, xn, xp, dx, tb, tj, uu, bv);
            R f = alpha0 + bv;
            const R2 *pp = Pp + tid;
#pragma unroll 4
            for (int q = 0; q < nq; ++q) {
                const R2 v = pp[q * ST];
                f = fma(ub[2 * q], v.x, f);
                f = fma(ub[2 * q + 1], v.y, f);
            }
            fnew = f;
            rho_part = double(wgt) * double(f);
            const R wt = wgt * tb;
            gn0 = wt * dx[0];
            gn1 = wt * dx[1];
            gn2 = wt * dx[2];
            jn0 = tj * dx[0];
            jn1 = tj * dx[1];
            jn2 = tj * dx[2];
            dU = uu - ou;
        }
        if (h == 0) {
#pragma unroll
            for (int c = 0; c < NQC; ++c) {
                const int o = ol + 32 * c;
                if (o < no) {
                    const R t = (o & 1) ? -unc[c] * Xo[c] : uns[c] * Xo[c];
                    gn0 += cg[c][0] * t;
                    gn1 += cg[c][1] * t;
                    gn2 += cg[c][2] * t;
                }
            }
        }
        // pair terms of the next mover of this spin at its current position
        have_old = false;
        if (ip + 1 < m1 && ip + 1 != NUP) {
            if (tid < NUP) {
                R bv;
                pair_terms<R>(S, xs + 3 * (ip + 1), xp, odx, ob, oj, ou, bv);
            }
            have_old = true;
        }
        const double rho = wave_sum(rho_part);  // FP64 sum: the node test uses its sign
        // grad ln psi at the proposed position, sum_b (g_b / rho + j_b): every lane divides
        // its determinant part by the ratio before the sums (an infinite or NaN sum when
        // rho = 0 is never used: such a move is rejected)
        const R irho = rcp_r<R>(R(rho));
        R dn0 = fma(gn0, irho, jn0), dn1 = fma(gn1, irho, jn1), dn2 = fma(gn2, irho, jn2);
        if constexpr (sizeof(R) == 4) {
            wave_sum2(dn0, dn1);
            wave_sum2(dn2, dU);
        } else {
            dn0 = wave_sum(dn0);
            dn1 = wave_sum(dn1);
            dn2 = wave_sum(dn2);
            dU = wave_sum(dU);
        }
        const R dn[3] = {dn0, dn1, dn2};
        bool accept = !(S.fn && !(rho > 0.0)) && rho != 0.0;
        double logp_rec = __builtin_nan("");
        if (accept) {
            // Metropolis exponent of the drift-diffusion Green's functions: FP64 moves in FP64,
            // mixed moves in FP32 (its inputs carry FP32 errors already)
            if constexpr (sizeof(R) == 8) {
                double lq = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double drn = fmin(dcut, fmax(-dcut, dn[d] * sig2));
                    lq += (dro[d] + drn) * (dro[d] - drn - 2.0 * dxm[d]);
                }
                const double logp = 2.0 * (log_r<double>(fabs(rho)) + dU) + lq * half_isig2;
                logp_rec = logp;
                accept = u4 < exp_r<double>(fmin(0.0, logp));
            } else {
                const float fs2 = float(sig2), fcut = float(dcut);
                float lq = 0.0f;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const float drn = fminf(fcut, fmaxf(-fcut, dn[d] * fs2));
                    lq += (dro[d] + drn) * (dro[d] - drn - 2.0f * dxm[d]);
                }
                const float logp = 2.0f * (log_r<float>(fabsf(float(rho))) + dU) +
                                   lq * float(half_isig2);
                logp_rec = logp;
                accept = u4 < double(exp_r<float>(fminf(0.0f, logp)));
            }
        }
        if (T.trace && tid == 0) {
            double *rec = T.trace + (size_t(w) * S.npart + ip) * kTraceRec;
            rec[0] = rho;
            rec[1] = double(dU);
            rec[2] = accept ? 1.0 : 0.0;
            rec[3] = u4;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                rec[4 + d] = xn[d];
                rec[7 + d] = double(dro[d]);
            }
            rec[10] = logp_rec;
            rec[11] = 0.0;
        }
        if (accept) {
            ++nacc;
            const R irr = irho;
            if (tid < NUP) fb[tid] = fnew;
            wave_sync();
            if (tid < NUP) {
                R z = R(0);
#pragma unroll
                for (int b = 0; b < NUP; ++b) z += fb[b] * Ac[b];
                const R fz = (z - (tid == a ? R(1) : R(0))) * irr;
#pragma unroll
                for (int b = 0; b < NUP; ++b) Ac[b] -= wb[b < HB ? b : HBP + b - HB] * fz;
            }
            if (tid < 3) xs[3 * ip + tid] = xn[tid];
            if (h == 0) {
#pragma unroll
                for (int c = 0; c < NQC; ++c) {
                    const int o = ol + 32 * c;
                    if (o < no && !(o & 1)) {
                        R2 z;
                        z.x = unc[c];
                        z.y = uns[c];
                        Pm[(o >> 1) * ST + a] = z;
                    }
                }
            }
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

// the move kernel's chunk count of projection outputs (32 per chunk) for nq plane waves
__host__ __device__ inline int mv_chunks(int nq) { return nq <= 16 ? 1 : 4; }

template <int NUP, typename R>
__device__ __forceinline__ void sm_moves_walker_any(const DevSys &S, double *sm, int w, int tid,
                                                    const double *__restrict__ xin, double *xout,
                                                    const uint64_t *irn_in, uint64_t *irn_out,
                                                    char *sinv_g, int m0, int m1,
                                                    const DevStep &T) {
    if (S.nq <= 16)
        sm_moves_walker<NUP, R, 1>(S, sm, w, tid, xin, xout, irn_in, irn_out, sinv_g, m0, m1, T);
    else
        sm_moves_walker<NUP, R, 4>(S, sm, w, tid, xin, xout, irn_in, irn_out, sinv_g, m0, m1, T);
}


// FP64 rebuild of the inverse of positions x (every krefresh moves)
template <int NUP>
__device__ __forceinline__ void sm_refresh_walker(const DevSys &S, double *sm, int w, int tid,
                                                  const double *__restrict__ x, char *sinv_g) {
    const Layout<NUP> L(S.nq, true);
    constexpr int M = 6 * NUP;
    load_q<NUP>(S, sm, L, tid);
    for (int t = tid; t < M; t += 64) sm[L.xs + t] = x[size_t(w) * M + t];
    wave_sync();
    const Tile<NUP> Tl(tid);
    const DetInfo D = build_inverse<NUP>(S, sm, L, Tl, tid);
    store_sinv<NUP>(S, sm, L, D.singular, sinv_g + size_t(w) * sinv_stride(S), tid);
}

// end of the sweep: FP64 hpsi of the final positions B.x, the fresh inverse into B.sinv,
// weight against the energy of A, estimators and branch draw (state after the moves in
// B.irn).  TB: the view of the radial tables (TabL: staged in the workgroup's LDS, with the
// walker's area in the compact layout).
template <int NUP, class TB>
__device__ __forceinline__ void sm_end_walker(const DevSys &S, const TB &tb, double *sm, int w,
                                              int tid, const DevState &A, const DevState &B,
                                              const DevStep &T, int opcomp) {
    const Layout<NUP> L(S.nq, true, TB::kCompact);
    constexpr int M = 6 * NUP;
    load_q<NUP>(S, sm, L, tid);
    for (int t = tid; t < M; t += 64) sm[L.xs + t] = B.x[size_t(w) * M + t];
    wave_sync();
    const Tile<NUP> Tl(tid);
    PairTile<NUP> rad;
    const JSums J = jastrow_terms<NUP>(S, tb, sm, L, Tl, tid, &rad);
    const DetInfo D = build_inverse<NUP>(S, tb, sm, L, Tl, tid, &rad);
    store_sinv<NUP>(S, sm, L, D.singular, B.sinv + size_t(w) * sinv_stride(S), tid);
    grad_lap<NUP>(S, tb, sm, L, Tl, D, J, tid);
    finish_walker<NUP>(S, sm, L, A, B, T, w, tid, B.irn[w], opcomp);
}

template <int NUP>
__device__ __forceinline__ void sm_end_walker(const DevSys &S, double *sm, int w, int tid,
                                              const DevState &A, const DevState &B,
                                              const DevStep &T, int opcomp) {
    sm_end_walker<NUP>(S, TabG(S), sm, w, tid, A, B, T, opcomp);
}

}  // namespace dmc
