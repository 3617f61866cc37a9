// Device building blocks of one walker per 64-lane workgroup: LDS layout, the 8 x 8 lane
// tile, pairing-matrix build and Gauss-Jordan inversion, gradient/laplacian contraction,
// Jastrow and potential lookups.  Included by the kernel translation units.
#pragma once


#include <type_traits>

#include "dmc_kernels.h"

namespace dmc {

// ------------------------------------------------------------------------------------
// small helpers

// Ordering of the LDS traffic of one walker.  Every kernel gives a walker one 64-lane wave
// (and a workgroup may hold several walkers, one per wave), so the lanes that exchange data
// through LDS are always those of one wave: a wavefront-scope fence, which also keeps the
// compiler from moving LDS accesses across it, is the synchronisation they need.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// f(std::integral_constant<int, I>) for I = B .. E-1 (a loop over a compile-time index)
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Keeps v computed at this point: an empty asm statement that reads and writes it.  Without
// it the compiler may sink a sum to a distant use and hold all of its addends live in between.
__device__ __forceinline__ void pin(double &v) { asm volatile("" : "+v"(v)); }

// 1/sqrt(x) for x > 0 in the normal range: the hardware estimate refined by two Newton
// steps y <- y + y (1/2 - x y^2 / 2), each roughly squaring the relative error (a few ulp at
// the end; 8 operations instead of a correctly rounded square root and a division)
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double e = fma(-h * y, y, 0.5);
    y = fma(y, e, y);
    e = fma(-h * y, y, 0.5);
    return fma(y, e, y);
}

// 1/x: the hardware reciprocal estimate refined by two Newton steps y <- y + y (1 - x y)
// (about 1 ulp; 5 dependent operations instead of the division sequence)
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// sin and cos of 2 pi t for |t| <= 1/2 (t in turns), FP32: reduction to the nearest quarter
// turn (|x| <= pi/4), the Taylor polynomials of sin to x^9 and of cos to x^10 (truncation
// errors 1.8e-9 and 1.1e-10 at pi/4), then the rotation by that quarter turn
__device__ __forceinline__ void sincos_turns(float t, float &s, float &c) {
    const float q = rintf(4.0f * t);
    const float x = fmaf(-0.25f, q, t) * 6.28318530717958648f;
    const float x2 = x * x;
    float sp = fmaf(x2, 2.75573192e-6f, -1.98412698e-4f);
    sp = fmaf(x2, sp, 8.33333333e-3f);
    sp = fmaf(x2, sp, -1.66666667e-1f);
    const float sn = fmaf(x * x2, sp, x);
    float cp = fmaf(x2, -2.75573192e-7f, 2.48015873e-5f);
    cp = fmaf(x2, cp, -1.38888889e-3f);
    cp = fmaf(x2, cp, 4.16666667e-2f);
    cp = fmaf(x2, cp, -0.5f);
    const float cs = fmaf(x2, cp, 1.0f);
    // quarter turn iq (mod 4): odd iq swaps sin and cos; sin changes sign for iq = 2, 3 and
    // cos for iq = 1, 2 (selects and sign flips, no branches)
    const unsigned iq = unsigned(int(q));
    const bool sw = iq & 1u;
    const float s0 = sw ? cs : sn, c0 = sw ? sn : cs;
    s = __uint_as_float(__float_as_uint(s0) ^ ((iq & 2u) << 30));
    c = __uint_as_float(__float_as_uint(c0) ^ (((iq ^ (iq >> 1)) & 1u) << 31));
}

// The same in FP64: t - q/4 is exact (Sterbenz), x = 2 pi (t - q/4) in [-pi/4, pi/4] carries
// one rounding; the Taylor polynomials of sin to x^17 and of cos to x^18 (coefficients
// +-1/n!, truncation errors below 1e-19) in Horner form in x^2; the quarter turn as above
__device__ __forceinline__ void sincos_turns(double t, double &s, double &c) {
    const double q = rint(4.0 * t);
    const double x = fma(-0.25, q, t) * 6.283185307179586477;
    const double x2 = x * x;
    double sp = fma(x2, 1.0 / 355687428096000.0, -1.0 / 1307674368000.0);  // 1/17!, -1/15!
    sp = fma(x2, sp, 1.0 / 6227020800.0);
    sp = fma(x2, sp, -1.0 / 39916800.0);
    sp = fma(x2, sp, 1.0 / 362880.0);
    sp = fma(x2, sp, -1.0 / 5040.0);
    sp = fma(x2, sp, 1.0 / 120.0);
    sp = fma(x2, sp, -1.0 / 6.0);
    const double sn = fma(x * x2, sp, x);
    double cp = fma(x2, -1.0 / 6402373705728000.0, 1.0 / 20922789888000.0);  // -1/18!, 1/16!
    cp = fma(x2, cp, -1.0 / 87178291200.0);
    cp = fma(x2, cp, 1.0 / 479001600.0);
    cp = fma(x2, cp, -1.0 / 3628800.0);
    cp = fma(x2, cp, 1.0 / 40320.0);
    cp = fma(x2, cp, -1.0 / 720.0);
    cp = fma(x2, cp, 1.0 / 24.0);
    cp = fma(x2, cp, -0.5);
    const double cs = fma(x2, cp, 1.0);
    const unsigned iq = unsigned(int(q));
    const bool sw = iq & 1u;
    const double s0 = sw ? cs : sn, c0 = sw ? sn : cs;
    s = __longlong_as_double(__double_as_longlong(s0) ^ ((long long)((iq >> 1) & 1u) << 63));
    c = __longlong_as_double(__double_as_longlong(c0) ^
                             ((long long)((iq ^ (iq >> 1)) & 1u) << 63));
}

__device__ __forceinline__ double wrapd(double x, const DevSys &S) {
    return x - S.el * round(x * S.eli);  // Fortran nint: half away from zero
}

// Weights of the 4-point Lagrange interpolation on the nodes -1, 0, 1, 2 at t
// (coshpot.f90:172-179 / betafun.f90:183-190): w0 = -t(t-1)(t-2)/6 and w3 = (t+1)t(t-1)/6
// from p = t(t-1); w2 and w1 from the identities sum w = 1 and sum w x = t that the cubic
// weights satisfy (10 operations instead of 16)
template <typename R>
__device__ __forceinline__ void lagrange_weights(R t, R &w0, R &w1, R &w2, R &w3) {
    const R p = fma(t, t, -t);
    w0 = p * fma(t, R(-1.0 / 6.0), R(1.0 / 3.0));
    w3 = p * fma(t, R(1.0 / 6.0), R(1.0 / 6.0));
    w2 = fma(R(-2), w3, t + w0);
    w1 = R(1) - (w0 + w3) - w2;
}

// interval of s on the table grid, clamped to [1, ntab-2], and its 4-point weights
__device__ __forceinline__ int lagrange(double s, int ntab, double &w0, double &w1, double &w2,
                                        double &w3) {
    int idx = int(s);
    idx = idx > ntab - 2 ? ntab - 2 : idx;
    idx = idx < 1 ? 1 : idx;
    lagrange_weights<double>(s - idx, w0, w1, w2, w3);
    return idx;
}

__device__ __forceinline__ double4 lookup4(const double4 *tab, int idx, double w0, double w1,
                                           double w2, double w3) {
    const double4 a = tab[idx - 1], b = tab[idx], c = tab[idx + 1], d = tab[idx + 2];
    return make_double4(w0 * a.x + w1 * b.x + w2 * c.x + w3 * d.x,
                        w0 * a.y + w1 * b.y + w2 * c.y + w3 * d.y,
                        w0 * a.z + w1 * b.z + w2 * c.z + w3 * d.z,
                        w0 * a.w + w1 * b.w + w2 * c.w + w3 * d.w);
}

// beta(r) and (beta', lap beta) from the beta table (radialfun, betafun.f90:173-194)
__device__ __forceinline__ double beta_val(const DevSys &S, double r) {
    if (S.nobeta || r > S.rangeb) return 0.0;
    double w0, w1, w2, w3;
    const int idx = lagrange(S.scaleb * r, S.ntab, w0, w1, w2, w3);
    return w0 * S.bt1[idx - 1] + w1 * S.bt1[idx] + w2 * S.bt1[idx + 1] + w3 * S.bt1[idx + 2];
}

// ------------------------------------------------------------------------------------
// Views of the radial tables that the hpsi pair passes read: beta b [ntab+1], {beta',
// lap beta} [ntab+1], the potential [ntab+1] and the Jastrow/potential records {v, u, u',
// u''} of the intervals before jfar.
//
// TabG reads the arrays of the system in global memory (through the vector L1), one walker
// per workgroup.  TabL reads a copy that the workgroup stages in LDS once for all of its
// walkers (one per wave), with the walkers' own areas in the compact layout beside it:
// a 64-lane gather of random table entries is one LDS access per lane instead of one L1
// cache-line access per lane, at LDS latency.
struct TabG {
    static constexpr bool kCompact = false;
    const double *b1, *v1;
    const double2 *b2;
    const double4 *j4;
    __device__ explicit TabG(const DevSys &S) : b1(S.bt1), v1(S.vt1), b2(S.bt2), j4(S.jt) {}
    __device__ double bt1(int i) const { return b1[i]; }
    __device__ double vt1(int i) const { return v1[i]; }
    __device__ double2 bt2(int i) const { return b2[i]; }
    __device__ double4 jt(int i) const { return j4[i]; }
};

typedef __attribute__((address_space(3))) const double lds_cdouble;

// offsets (in doubles) of the staged tables at the start of a workgroup's LDS: bt1, vt1,
// bt2 (pairs), then the records 0 .. jfar + 1 that a stencil before jfar reads (all of them
// when the Jastrow reaches the end of the table); the walkers' areas start at `total`,
// a multiple of 16 doubles
struct TabLayout {
    int n1, nrec, b1, v1, b2, j4, total;
    __host__ __device__ explicit TabLayout(int ntab, int jfar) {
        n1 = ntab + 1;
        nrec = jfar + 2 < n1 ? jfar + 2 : n1;
        b1 = 0;
        v1 = b1 + n1;
        b2 = (v1 + n1 + 1) / 2 * 2;  // 16-byte aligned pairs
        j4 = b2 + 2 * n1;
        total = (j4 + 4 * nrec + 15) / 16 * 16;
    }
};

struct TabL {
    static constexpr bool kCompact = true;
    lds_cdouble *b1, *v1, *b2, *j4;
    __device__ TabL(const DevSys &S, const double *base) {
        const TabLayout TL(S.ntab, S.jfar);
        lds_cdouble *p = (lds_cdouble *)(base);
        b1 = p + TL.b1;
        v1 = p + TL.v1;
        b2 = p + TL.b2;
        j4 = p + TL.j4;
    }
    __device__ double bt1(int i) const { return b1[i]; }
    __device__ double vt1(int i) const { return v1[i]; }
    __device__ double2 bt2(int i) const { return make_double2(b2[2 * i], b2[2 * i + 1]); }
    __device__ double4 jt(int i) const {
        return make_double4(j4[4 * i], j4[4 * i + 1], j4[4 * i + 2], j4[4 * i + 3]);
    }
};

// copy of the tables into the workgroup's LDS at base, by all of its threads; the caller
// synchronises the workgroup before the first read
__device__ __forceinline__ void stage_tables(const DevSys &S, double *base, int tid, int nthr) {
    const TabLayout TL(S.ntab, S.jfar);
    for (int i = tid; i < TL.n1; i += nthr) {
        base[TL.b1 + i] = S.bt1[i];
        base[TL.v1 + i] = S.vt1[i];
        const double2 v = S.bt2[i];
        base[TL.b2 + 2 * i] = v.x;
        base[TL.b2 + 2 * i + 1] = v.y;
    }
    for (int i = tid; i < TL.nrec; i += nthr) {
        const double4 v = S.jt[i];
        base[TL.j4 + 4 * i] = v.x;
        base[TL.j4 + 4 * i + 1] = v.y;
        base[TL.j4 + 4 * i + 2] = v.z;
        base[TL.j4 + 4 * i + 3] = v.w;
    }
}

// walkers (one per wave) per workgroup of the kernels that stage the tables in LDS
constexpr int kTabWaves = 8;

template <class TB>
__device__ __forceinline__ double4 lookup4j(const TB &tb, int idx, double w0, double w1,
                                            double w2, double w3) {
    const double4 a = tb.jt(idx - 1), b = tb.jt(idx), c = tb.jt(idx + 1), d = tb.jt(idx + 2);
    return make_double4(w0 * a.x + w1 * b.x + w2 * c.x + w3 * d.x,
                        w0 * a.y + w1 * b.y + w2 * c.y + w3 * d.y,
                        w0 * a.z + w1 * b.z + w2 * c.z + w3 * d.z,
                        w0 * a.w + w1 * b.w + w2 * c.w + w3 * d.w);
}

template <class TB>
__device__ __forceinline__ double beta_val(const TB &tb, const DevSys &S, double r) {
    if (S.nobeta || r > S.rangeb) return 0.0;
    double w0, w1, w2, w3;
    const int idx = lagrange(S.scaleb * r, S.ntab, w0, w1, w2, w3);
    return w0 * tb.bt1(idx - 1) + w1 * tb.bt1(idx) + w2 * tb.bt1(idx + 1) + w3 * tb.bt1(idx + 2);
}

// Sum over the 64 lanes of a wave, returned as a wave-uniform value.  Within each 16-lane
// row by DPP lane exchanges (xor 1, xor 2, mirror within 8, mirror within 16), then rows
// combined by row_bcast:15 (rows 1, 3 add the last lane of the row before) and row_bcast:31
// (rows 2, 3 add lane 31), so that lane 63 holds the total, read back as a scalar.  All 64
// lanes must be active at the call.
__device__ __forceinline__ int dpp_row(int v, int k) {
    switch (k) {
        case 0: return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
        case 1: return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
        case 2: return __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false);  // row_half_mirror
        default: return __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false); // row_mirror
    }
}
__device__ __forceinline__ int dpp_bcast(int v, int k) {
    return k == 0 ? __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false)   // row_bcast:15
                  : __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31
}
// the same row exchanges with bound_ctrl set (every source lane of these patterns is in the
// row), so that the compiler folds each exchange into the add (v_add_f32_dpp).  The adds are
// kept apart from the product that may form v (no contraction into an FMA): a fused first
// step would round the lane's own term and its partner's differently, and the lanes of a row
// would end with totals that differ in the last bits.
__device__ __forceinline__ float row_sum16(float v) {
#pragma clang fp contract(off)
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, true));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, true));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, true));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, true));
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
    v = row_sum16(v);
#pragma unroll
    for (int k = 0; k < 2; ++k) v += __int_as_float(dpp_bcast(__float_as_int(v), k));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// Sums over the 64 lanes of two values at once: a permlane swap of the half-waves puts the
// half-wave pair sums of a in lanes 0-31 and of b in lanes 32-63, rows of 16 are summed,
// and row_bcast:15 joins rows 0+1 (a, lane 31) and rows 2+3 (b, lane 63).
__device__ __forceinline__ void wave_sum2(float &a, float &b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                    false);
    float v = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    v = row_sum16(v);
    v += __int_as_float(dpp_bcast(__float_as_int(v), 0));
    a = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
    b = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double dbl_of(int lo, int hi) {
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// the value of v in lane l, wave-uniform
__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    return dbl_of(__builtin_amdgcn_readlane(int(b & 0xffffffffll), l),
                  __builtin_amdgcn_readlane(int(b >> 32), l));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long long b = __double_as_longlong(v);
        v += dbl_of(dpp_row(int(b & 0xffffffffll), k), dpp_row(int(b >> 32), k));
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const long long b = __double_as_longlong(v);
        v += dbl_of(dpp_bcast(int(b & 0xffffffffll), k), dpp_bcast(int(b >> 32), k));
    }
    const long long b = __double_as_longlong(v);
    return dbl_of(__builtin_amdgcn_readlane(int(b & 0xffffffffll), 63),
                  __builtin_amdgcn_readlane(int(b >> 32), 63));
}

// exp(i m theta) from the precomputed powers, |m| <= 4.  m is wave-uniform (the same q
// for every lane): the power is picked by a register index, no per-lane selects.
__device__ __forceinline__ void cpow(const double (&pc)[5], const double (&ps)[5], int m,
                                     double &c, double &s) {
    m = __builtin_amdgcn_readfirstlane(m);
    int am = m < 0 ? -m : m;
    am = am > 4 ? 4 : am;
    c = pc[am];  // a uniform index into registers (relative moves, no branches)
    s = ps[am];
    s = m < 0 ? -s : s;
}

// exp(i m 2 pi x_d / L), m = 0..4, of the three coordinates of one position
__device__ __forceinline__ void axis_powers(const DevSys &S, const double *pos, double (&pc)[3][5],
                                            double (&ps)[3][5]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double s1, c1;
        const double t = pos[d] * S.eli;  // turns; the positions are wrapped into the box
        sincos_turns(t - rint(t), s1, c1);
        pc[d][0] = 1.0; ps[d][0] = 0.0;
        pc[d][1] = c1; ps[d][1] = s1;
#pragma unroll
        for (int m = 2; m < 5; ++m) {
            pc[d][m] = pc[d][m - 1] * c1 - ps[d][m - 1] * s1;
            ps[d][m] = ps[d][m - 1] * c1 + pc[d][m - 1] * s1;
        }
    }
}

// exp(i q.r) for all q of one position (q-vectors wave-uniform)
__device__ __forceinline__ void phases_of(const DevSys &S, const int4 *qm, const double *pos,
                                          double2 *out, int stride) {
    double pc[3][5], ps[3][5];
    axis_powers(S, pos, pc, ps);
    // the q vectors through scalar loads of the system's read-only table (wave-uniform, so
    // the powers are picked by a register index without a move from vector registers)
    typedef int v4i __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(4))) const v4i cv4i;
    const cv4i *qk = (const cv4i *)(S.qm);
    (void)qm;
    for (int q = 0; q < S.nq; ++q) {
        const v4i mv = qk[q];
        const int4 m = make_int4(mv.x, mv.y, mv.z, mv.w);
        double cx, sx, cy, sy, cz, sz;
        cpow(pc[0], ps[0], m.x, cx, sx);
        cpow(pc[1], ps[1], m.y, cy, sy);
        cpow(pc[2], ps[2], m.z, cz, sz);
        const double cxy = cx * cy - sx * sy, sxy = sx * cy + cx * sy;
        out[q * stride] = make_double2(cxy * cz - sxy * sz, sxy * cz + cxy * sz);
    }
}

// exp(i q.r) of every particle into ph[q * NP + p]: lanes over the first 64 particles
// through their per-axis powers; particles beyond the 64th (N > 64) one (particle, q) pair
// per lane with a direct sincos of q.r, so that the wave makes one pass of the slow path
template <int NUP>
__device__ __forceinline__ void phases_all(const DevSys &S, const double *xs, const int4 *qm,
                                           double2 *ph, int tid) {
    constexpr int NP = 2 * NUP;
    if (tid < NP) phases_of(S, qm, xs + 3 * tid, ph + tid, NP);
    if constexpr (NP > 64) {
        const int nq = S.nq;
        for (int t = tid; t < (NP - 64) * nq; t += 64) {
            const int e = t / nq, q = t - e * nq, p = 64 + e;
            const int4 m = qm[q];
            const double *pos = xs + 3 * p;
            const double tq = S.eli * (m.x * pos[0] + m.y * pos[1] + m.z * pos[2]);
            double sn, cs;
            sincos_turns(tq - rint(tq), sn, cs);
            ph[q * NP + p] = make_double2(cs, sn);
        }
    }
}

// exp(i q.r) of every particle for the q vectors q0 <= q < q1 into ph[(q - q0) * NP + p]:
// the same arithmetic per element as phases_all (one chunk of the q vectors at a time, for
// the compact layout)
template <int NUP>
__device__ __forceinline__ void phases_range(const DevSys &S, const double *xs, const int4 *qm,
                                             double2 *ph, int q0, int q1, int tid) {
    constexpr int NP = 2 * NUP;
    if (tid < NP) {
        double pc[3][5], ps[3][5];
        axis_powers(S, xs + 3 * tid, pc, ps);
        for (int q = q0; q < q1; ++q) {
            const int4 m = qm[q];
            double cx, sx, cy, sy, cz, sz;
            cpow(pc[0], ps[0], m.x, cx, sx);
            cpow(pc[1], ps[1], m.y, cy, sy);
            cpow(pc[2], ps[2], m.z, cz, sz);
            const double cxy = cx * cy - sx * sy, sxy = sx * cy + cx * sy;
            ph[(q - q0) * NP + tid] = make_double2(cxy * cz - sxy * sz, sxy * cz + cxy * sz);
        }
    }
    if constexpr (NP > 64) {
        const int nqc = q1 - q0;
        for (int t = tid; t < (NP - 64) * nqc; t += 64) {
            const int e = t / nqc, qq = t - e * nqc, p = 64 + e;
            const int4 m = qm[q0 + qq];
            const double *pos = xs + 3 * p;
            const double tq = S.eli * (m.x * pos[0] + m.y * pos[1] + m.z * pos[2]);
            double sn, cs;
            sincos_turns(tq - rint(tq), sn, cs);
            ph[qq * NP + p] = make_double2(cs, sn);
        }
    }
}

// ------------------------------------------------------------------------------------
// LDS layout of one walker (offsets in doubles)
//
// Compact layout (hpsi kernels that share LDS-staged tables between the walkers of a
// workgroup): the lean layout with the phases of qch q vectors at a time and the partial
// sums of one spin at a time, so that eight walkers and the tables fit in the LDS of a CU.
//
// Full layout (single-particle ratio, parameter adjoint): the phases of every particle and
// the inverse stay in LDS after build_inverse.  Lean layout (hpsi, step and refresh
// kernels): phases, elimination scratch, inverse and partial sums share one area in turn,
// and the gradient pass rebuilds the phases once the inverse is in registers, so that two
// walkers fit where one did.

template <int NUP>
struct Layout {
    static constexpr int TR = (NUP + 7) / 8, R8 = 8 * TR, NP = 2 * NUP;
    // row stride of the inverse and of the partial sums: = 8 (mod 16), so that the 8 x 8
    // lane grid reading a tile (rows c + 8b, columns r + 8a) hits distinct banks
    static constexpr int SP = (NUP + 7) / 16 * 16 + 8;
    // q vectors per chunk of phases in the compact layout
    static constexpr int kQChunk = 8;
    int xs, qc, qm, dp, xn, mcol, ints, work, sinv, est, scal, total;
    int qch;  // q vectors whose phases the work area holds at once
    bool lean;
    __host__ __device__ explicit Layout(int nq, bool lean_ = false, bool compact = false)
        : lean(lean_ || compact) {
        qch = compact && nq > kQChunk ? kQChunk : nq;
        const int phs = 2 * NP * qch;         // phases, [qch][NP] complex
        const int inv = NUP * SP;             // Sinv(j, i) at [j * SP + i]
        // partial sums, [spin][3][8][SP]: both spins, or one at a time (compact)
        const int part = (compact ? 3 : 6) * 8 * SP;
        xs = 0;
        qc = xs + 3 * NP;                     // per q: w, -w q (3), -w q^2
        qm = qc + 5 * nq;                     // integer q vectors (int4)
        if (!lean) {
            dp = qm + 2 * nq;
            xn = dp + 3 * NP;                 // one extra position (4) and its phases (2 nq)
            mcol = xn + 4 + 2 * nq;
            ints = mcol + R8;                 // perm[64], pinv[64]
            work = ints + 64;
            sinv = work + phs;
            est = work;
            scal = work + (phs + inv > part ? phs + inv : part);
        } else {
            work = qm + 2 * nq;
            sinv = work;
            ints = sinv + inv;                // live until the inverse is scattered
            mcol = ints + 64;
            dp = work;                        // written once the partial sums are read
            est = dp + 3 * NP;
            xn = -1;                          // no extra position
            int w = phs;
            if (mcol + R8 - work > w) w = mcol + R8 - work;
            if (part > w) w = part;
            if (3 * NP + 42 > w) w = 3 * NP + 42;
            scal = work + w;
        }
        total = scal + 16;
    }
};

template <int NUP>
size_t lds_bytes(int nq, bool lean = false) { return sizeof(double) * Layout<NUP>(nq, lean).total; }

// per-walker stride (doubles) of the compact layout: whole 128-byte lines
template <int NUP>
__host__ __device__ inline int compact_stride(int nq) {
    return (Layout<NUP>(nq, true, true).total + 15) / 16 * 16;
}

// LDS of a workgroup of the table-staging kernels: the tables, then one walker area in the
// compact layout per wave
template <int NUP>
size_t tab_lds_bytes(const DevSys &S) {
    return sizeof(double) * (size_t(TabLayout(S.ntab, S.jfar).total) +
                             size_t(kTabWaves) * compact_stride<NUP>(S.nq));
}

// Load q coefficients into LDS (all lanes)
template <int NUP>
__device__ __forceinline__ void load_q(const DevSys &S, double *sm, const Layout<NUP> &L,
                                       int tid) {
    int4 *qm = reinterpret_cast<int4 *>(sm + L.qm);
    const double tp = S.twopi_l;
    for (int q = tid; q < S.nq; q += 64) {
        const int4 m = S.qm[q];
        const double w = S.wq[q];
        const double qx = tp * m.x, qy = tp * m.y, qz = tp * m.z;
        double *c = sm + L.qc + 5 * q;
        c[0] = w;
        c[1] = -w * qx;
        c[2] = -w * qy;
        c[3] = -w * qz;
        c[4] = -w * (qx * qx + qy * qy + qz * qz);
        qm[q] = m;
    }
}

// ------------------------------------------------------------------------------------
// Lane tile of the 8 x 8 grid
template <int NUP>
struct Tile {
    static constexpr int TR = (NUP + 7) / 8;
    int r, c, ri[TR], cj[TR];
    bool rv[TR], cv[TR];
    __device__ explicit Tile(int tid) {
        r = tid >> 3;
        c = tid & 7;
#pragma unroll
        for (int a = 0; a < TR; ++a) {
            const int i = r + 8 * a;
            rv[a] = i < NUP;
            ri[a] = rv[a] ? i : NUP - 1;
            const int j = c + 8 * a;
            cv[a] = j < NUP;
            cj[a] = cv[a] ? j : NUP - 1;
        }
    }
};

// Largest 32-bit key over the lanes that share lane & 7 (the 8 lanes of a tile column group):
// integer maxima with the lanes ^ 8 (DPP row_ror:8), ^ 16 and ^ 32 (permlane swaps of a
// register with itself, which give both lanes of a pair the same two values); every lane of
// the group ends with the group's maximum.
__device__ __forceinline__ int group_max_key(int key) {
    key = max(key, __builtin_amdgcn_update_dpp(0, key, 0x128, 0xF, 0xF, false));
    const auto s16 = __builtin_amdgcn_permlane16_swap(unsigned(key), unsigned(key), false, false);
    key = max(int(s16[0]), int(s16[1]));
    const auto s32 = __builtin_amdgcn_permlane32_swap(unsigned(key), unsigned(key), false, false);
    return max(int(s32[0]), int(s32[1]));
}

// per-lane partial sums of the particles of one spin into the work area, [3][8][SP]: the up
// particles ri[a] of the lane's tile rows by lane column (spin 0), or the down particles
// cj[a] of its tile columns by lane row (spin 1); a lane over particles then sums the 8
// entries of its particle
template <int NUP>
__device__ __forceinline__ void put_spin_part(double *part, const Tile<NUP> &T, int spin,
                                              const double (*p)[3]) {
    constexpr int TR = Layout<NUP>::TR, SP = Layout<NUP>::SP;
#pragma unroll
    for (int a = 0; a < TR; ++a) {
        const bool v = spin ? T.cv[a] : T.rv[a];
        const int k = spin ? T.cj[a] : T.ri[a], l = spin ? T.r : T.c;
        if (v) {
#pragma unroll
            for (int d = 0; d < 3; ++d) part[(d * 8 + l) * SP + k] = p[a][d];
        }
    }
}

// the radial part alpha_1 + beta(r) of a lane's tile of the pairing matrix
template <int NUP>
struct PairTile {
    double v[Layout<NUP>::TR][Layout<NUP>::TR];
};

struct DetInfo {
    double logdet;
    int sgn;
    bool singular;
};

// Phases A-D: plane-wave phases of the positions in LDS (xs), the pairing matrix, its
// inverse, log|det| and sign.  Leaves the phases (work area) and the inverse (sinv,
// Sinv(j, i) at [j * SP + i]) in LDS.
// TB: the view of the beta table (when rad is not given); with a compact-layout view the
// phases are built and contracted qch q vectors at a time.
template <int NUP, class TB>
__device__ __forceinline__ DetInfo build_inverse(const DevSys &S, const TB &tb, double *sm,
                                                 const Layout<NUP> &L, const Tile<NUP> &T,
                                                 int tid, const PairTile<NUP> *rad = nullptr) {
    constexpr int TR = Layout<NUP>::TR, SP = Layout<NUP>::SP, NP = Layout<NUP>::NP;
    const int nq = S.nq;
    const int r = T.r, c = T.c;
    const double *xs = sm + L.xs;
    double2 *ph = reinterpret_cast<double2 *>(sm + L.work);
    const int4 *qm = reinterpret_cast<const int4 *>(sm + L.qm);
    int *perm = reinterpret_cast<int *>(sm + L.ints);
    int *pinv = perm + 64;
    const double *qc = sm + L.qc;
    const int *ri = T.ri, *cj = T.cj;
    const bool *rv = T.rv, *cv = T.cv;

    // Phase A: plane-wave phases of every particle
    // q-major: ph[q * NP + p], so that lanes over particles read consecutive words
    if constexpr (!TB::kCompact) {
        phases_all<NUP>(S, xs, qm, ph, tid);
        wave_sync();
    }

    // Phase B: pairing-matrix tile S(i, j) = alpha_1 + beta(|r_ij|) + sum_q w_q cos(q.r_ij);
    // the radial part alpha_1 + beta comes from the Jastrow pass when it ran first (rad)
    double Sm[TR][TR];
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TR; ++b) {
            if (rad) {
                Sm[a][b] = rad->v[a][b];
                continue;
            }
            const double *xi = xs + 3 * ri[a], *xj = xs + 3 * (NUP + cj[b]);
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double dx = wrapd(xi[d] - xj[d], S);
                r2 += dx * dx;
            }
            Sm[a][b] = S.alpha0 + beta_val(tb, S, sqrt(r2));
        }
    // the plane-wave terms of q vectors q0 <= q < q1, whose phases are at ph[(q - q0) * NP]
    auto add_waves = [&](int q0, int q1) __attribute__((always_inline)) {
#pragma unroll 2
        for (int q = q0; q < q1; ++q) {
            const double w = qc[5 * q];
            const double2 *pq = ph + (q - q0) * NP;
            double2 zi[TR], zj[TR];
#pragma unroll
            for (int a = 0; a < TR; ++a) {
                zi[a] = pq[ri[a]];
                zi[a].x *= w;
                zi[a].y *= w;
                zj[a] = pq[NUP + cj[a]];
            }
#pragma unroll
            for (int a = 0; a < TR; ++a)
#pragma unroll
                for (int b = 0; b < TR; ++b) {
                    Sm[a][b] = fma(zi[a].x, zj[b].x, Sm[a][b]);
                    Sm[a][b] = fma(zi[a].y, zj[b].y, Sm[a][b]);
                }
        }
    };
    if constexpr (TB::kCompact) {
        for (int q0 = 0; q0 < nq; q0 += L.qch) {
            const int q1 = q0 + L.qch < nq ? q0 + L.qch : nq;
            phases_range<NUP>(S, xs, qm, ph, q0, q1, tid);
            wave_sync();
            add_waves(q0, q1);
            wave_sync();  // the next chunk's phases, or the elimination scratch, go there
        }
    } else {
        add_waves(0, nq);
        wave_sync();  // the lean layout puts the elimination scratch over the phases
    }

    // Phase C: Gauss-Jordan inversion with partial pivoting, rows never moved
    // The pivot row goes through the area of the inverse, which is written only once the
    // elimination ends.
    double *prow = sm + L.sinv, *mcol = sm + L.mcol;
    // low bits of the pivot keys: 63 - i for a row i not yet a pivot row; the sign bit is set
    // once it is one (and for the padding rows), which makes every key of that row negative
    int klo[TR];
#pragma unroll
    for (int a = 0; a < TR; ++a) klo[a] = rv[a] ? 63 - (r + 8 * a) : int(0x80000000u);
    double prod = 1.0;
    int expo = 0, sgn = 1;
    bool singular = false;
    int permk = 0;  // lane k: the pivot row of step k
    // the parity of the permutation's inversions, counted as the pivots are chosen: step k's
    // pivot row p precedes, in the pivot order, exactly the rows still unused below it
    // (scalar registers: p is wave-uniform)
    static_assert(NUP <= 64, "rows in a 64-bit mask");
    uint64_t unused = NUP == 64 ? ~0ull : (1ull << NUP) - 1;
    int inv_odd = 0;
    // The pivots in blocks of eight, one per tile column BK (a compile-time index, so the
    // pivot column is read and overwritten in place), one pivot per iteration of a rolled
    // loop within a block (a compact body whose registers are those of one step).  Tile row
    // ap of the pivot row is wave-uniform, picked by scalar branches.
    auto pivot_block = [&](auto bkc) __attribute__((always_inline)) {
        constexpr int BK = decltype(bkc)::value;
        constexpr int KN = NUP - 8 * BK < 8 ? NUP - 8 * BK : 8;
#pragma unroll 1
        for (int ck = 0; ck < KN; ++ck) {
            const int k = 8 * BK + ck;
            // Partial pivoting with 32-bit keys: the high word of |S(i, k)| (monotonic in the
            // magnitude) with its 6 lowest bits replaced by 63 - i.  The integer maximum is a
            // largest candidate to 2^-14 relative, ties going to the lower row (any such
            // pivot keeps the multipliers at most 1 + 2^-13 in magnitude).
            // The pivot column goes to LDS unscaled and first, as it depends on no pivot choice;
            // its readers scale it by the pivot's reciprocal (the same products).
            int key = -1;
            if (c == ck) {
#pragma unroll
                for (int a = 0; a < TR; ++a) mcol[r + 8 * a] = Sm[a][BK];
#pragma unroll
                for (int a = 0; a < TR; ++a) {
                    const unsigned hi = unsigned(__double_as_longlong(Sm[a][BK]) >> 32);
                    key = max(key, int((hi & 0x7fffffc0u) | unsigned(klo[a])));
                }
            }
            key = group_max_key(key);
            const int p = 63 - (__builtin_amdgcn_readlane(key, ck) & 63);
            const int ap = p >> 3;
            // the pivot from the lane that holds it (tile row ap of a lane of column group
            // ck) and its reciprocal, formed while the pivot row goes through LDS
            double pe = Sm[0][BK];
#pragma unroll
            for (int a = 1; a < TR; ++a)
                if (a == ap) pe = Sm[a][BK];
            const double piv = readlane_d(pe, 8 * (p & 7) + ck);
            double rp = rcp_nr(piv);
            pin(rp);  // issued ahead of the pivot-row stores, which cover its latency
            // the pivot row from tile row ap, in a scalar branch of its own (the empty asm
            // statement, different in each branch, keeps them from merging into selects)
            auto put_row = [&](auto ac) __attribute__((always_inline)) {
                constexpr int A = decltype(ac)::value;
                if (A == ap) {
                    if (r == (p & 7)) {
#pragma unroll
                        for (int b = 0; b < TR; ++b) prow[c + 8 * b] = Sm[A][b];
                    }
                    asm volatile("; pivot row %0" ::"n"(A) : "memory");
                }
            };
            static_for<0, TR>(put_row);
            wave_sync();
            {
                // sign and zero test on the pivot's bits (wave-uniform, scalar operations)
                const long long pb = __double_as_longlong(piv);
                singular |= (pb & 0x7fffffffffffffffll) == 0;
                sgn = pb < 0 ? -sgn : sgn;
            }
            prod *= fabs(piv);
            double f[TR], pv[TR];
#pragma unroll
            for (int a = 0; a < TR; ++a) f[a] = mcol[r + 8 * a] * rp;  // multipliers
#pragma unroll
            for (int b = 0; b < TR; ++b) pv[b] = prow[c + 8 * b];
            permk = tid == k ? p : permk;
            inv_odd ^= __popcll(unused & ((1ull << p) - 1)) & 1;
            unused &= ~(1ull << p);
            // elimination of every element, then the pivot column (lanes c == ck, tile
            // column BK) and the pivot row (lanes r == p & 7, tile row ap) overwritten
#pragma unroll
            for (int a = 0; a < TR; ++a)
#pragma unroll
                for (int b = 0; b < TR; ++b) Sm[a][b] -= f[a] * pv[b];
            const bool colk_lane = c == ck, prow_lane = r == (p & 7);
#pragma unroll
            for (int a = 0; a < TR; ++a) Sm[a][BK] = colk_lane ? -f[a] : Sm[a][BK];
#pragma unroll
            for (int a = 0; a < TR; ++a)
                if (a == ap) {
#pragma unroll
                    for (int b = 0; b < TR; ++b) {
                        const double prw = (b == BK && colk_lane) ? rp : pv[b] * rp;
                        Sm[a][b] = prow_lane ? prw : Sm[a][b];
                    }
                    if (prow_lane) klo[a] |= int(0x80000000u);
                }
            wave_sync();
        }
        // the running product of |pivots| renormalised after every block (exact)
        int e;
        prod = frexp(prod, &e);
        expo += e;
    };
    static_for<0, TR>(pivot_block);
    if (tid < NUP) {
        perm[tid] = permk;
        pinv[permk] = tid;
    }
    wave_sync();

    if (inv_odd) sgn = -sgn;  // the permutation's parity
    const double logdet = singular ? -1.0e30 : log(prod) + expo * 0.69314718055994530942;
    if (singular) sgn = 1;

    // Phase D: scatter the inverse: Sinv(pinv[alpha], perm[beta]) = B(alpha, beta)
    double *sinv = sm + L.sinv;
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TR; ++b)
            if (rv[a] && cv[b]) sinv[pinv[ri[a]] * SP + perm[cj[b]]] = Sm[a][b];
    wave_sync();
    return DetInfo{logdet, sgn, singular};
}

template <int NUP>
__device__ __forceinline__ DetInfo build_inverse(const DevSys &S, double *sm,
                                                 const Layout<NUP> &L, const Tile<NUP> &T,
                                                 int tid, const PairTile<NUP> *rad = nullptr) {
    return build_inverse<NUP>(S, TabG(S), sm, L, T, tid, rad);
}

// Jastrow and potential terms of the pair loop of getphi (coshpot.f90:165-202), opposite
// spins: per particle, lane p < NUP holds the Jastrow gradient of up particle p (j[0]) and
// down particle p (j[1]); the sums U, V and d2U are wave-uniform.  Runs before the pairing
// matrix is built (the lean layout's work area is free then), so that its partial sums are
// never live together with the inverse; with rad it also gives the radial part of the
// lane's tile of the pairing matrix from the same pair distances.
struct JSums {
    double j[2][3];
    double U, V, D2U;
};

template <int NUP, class TB>
__device__ __forceinline__ JSums jastrow_terms(const DevSys &S, const TB &tb, double *sm,
                                               const Layout<NUP> &L, const Tile<NUP> &T,
                                               int tid, PairTile<NUP> *rad = nullptr) {
    constexpr int TR = Layout<NUP>::TR, SP = Layout<NUP>::SP;
    constexpr int PS = 3 * 8 * SP;
    const double *xs = sm + L.xs;
    double *part = sm + L.work;
    const int r = T.r, c = T.c;
    double U = 0.0, V = 0.0, D2U = 0.0;
    double pu[TR][3], pdn[TR][3];
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) pu[a][d] = pdn[a][d] = 0.0;
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TR; ++b) {
            if (!(T.rv[a] && T.cv[b])) continue;
            const double *xi = xs + 3 * T.ri[a], *xj = xs + 3 * (NUP + T.cj[b]);
            // minimum image by round-to-nearest-even: at an exact tie both images have the
            // same distance and every pair term has zero slope there (r = L/2)
            double dx[3], r2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                dx[d] = xi[d] - xj[d];
                dx[d] -= S.el * rint(dx[d] * S.eli);
                r2 += dx[d] * dx[d];
            }
            // r and 1/r from one reciprocal square root, r floored at 1e-14 (coshpot.f90:167;
            // the floor also keeps coincident particles finite in the beta lookup)
            const double irj = rsqrt_nr(fmax(r2, 1.0e-28)), rj = r2 * irj, rr = rj;
            double w0, w1, w2, w3, v0, v1, v2, v3;
            int idx;
            // the beta sum is pinned before alpha_1 is added, so that its roundings do not
            // depend on how the code around it is arranged (both table views below give the
            // same bits)
            if constexpr (!TB::kCompact) {
                // tables in global memory: the beta and potential nodes of the pair are loaded
                // together, before either sum (intervals clamped to the table; out-of-range
                // terms are dropped below), so that one L1 round trip covers both
                double wb0, wb1, wb2, wb3;
                const int ib = lagrange(S.scaleb * rr, S.ntab, wb0, wb1, wb2, wb3);
                idx = lagrange(S.scale * rj, S.ntab, w0, w1, w2, w3);
                const double b0 = tb.bt1(ib - 1), b1 = tb.bt1(ib), b2 = tb.bt1(ib + 1),
                             b3 = tb.bt1(ib + 2);
                v0 = tb.vt1(idx - 1);
                v1 = tb.vt1(idx);
                v2 = tb.vt1(idx + 1);
                v3 = tb.vt1(idx + 2);
                if (rad) {
                    double bs = (S.nobeta || rr > S.rangeb)
                                    ? 0.0
                                    : wb0 * b0 + wb1 * b1 + wb2 * b2 + wb3 * b3;
                    pin(bs);
                    rad->v[a][b] = S.alpha0 + bs;
                }
            } else {
                // tables in LDS (short latency): each lookup where it is used, which keeps
                // fewer values live across the pair
                if (rad) {
                    double bs = beta_val(tb, S, rr);
                    pin(bs);
                    rad->v[a][b] = S.alpha0 + bs;
                }
                idx = lagrange(S.scale * rj, S.ntab, w0, w1, w2, w3);
                v0 = v1 = v2 = v3 = 0.0;
                if (rj < S.range && idx >= S.jfar) {
                    v0 = tb.vt1(idx - 1);
                    v1 = tb.vt1(idx);
                    v2 = tb.vt1(idx + 1);
                    v3 = tb.vt1(idx + 2);
                }
            }
            if (rj < S.range) {
                if (idx >= S.jfar) {
                    // past the healing distance u, u' and u'' are zero on the whole stencil:
                    // the potential alone, the same arithmetic as the record's first field
                    V += w0 * v0 + w1 * v1 + w2 * v2 + w3 * v3;
                } else {
                    const double4 jv = lookup4j(tb, idx, w0, w1, w2, w3);
                    V += jv.x;
                    U += jv.y;
                    const double t = jv.z * irj * S.invsqrtm;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        pu[a][d] += dx[d] * t;
                        pdn[b][d] -= dx[d] * t;
                    }
                    D2U += (jv.w + 2.0 * jv.z * irj) * S.invmu;
                }
            }
            // a scheduling fence per pair: the table loads of one pair in flight at a time
            // (unfenced, the straight-line loop issues every pair's loads at once)
            __builtin_amdgcn_sched_barrier(0);
        }
    JSums J;
    const int pl = tid < NUP ? tid : 0;
    if constexpr (TB::kCompact) {
        // one spin at a time through the work area: the same sums in the same order
#pragma unroll
        for (int spin = 0; spin < 2; ++spin) {
            put_spin_part<NUP>(part, T, spin, spin ? pdn : pu);
            wave_sync();
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double v = 0.0;
#pragma unroll
                for (int l = 0; l < 8; ++l) v += part[(d * 8 + l) * SP + pl];
                pin(v);
                J.j[spin][d] = v;
            }
            wave_sync();
        }
    } else {
#pragma unroll
        for (int a = 0; a < TR; ++a) {
            if (T.rv[a]) {
#pragma unroll
                for (int d = 0; d < 3; ++d) part[(d * 8 + c) * SP + T.ri[a]] = pu[a][d];
            }
            if (T.cv[a]) {
#pragma unroll
                for (int d = 0; d < 3; ++d) part[PS + (d * 8 + r) * SP + T.cj[a]] = pdn[a][d];
            }
        }
        wave_sync();
#pragma unroll
        for (int spin = 0; spin < 2; ++spin)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double v = 0.0;
#pragma unroll
                for (int l = 0; l < 8; ++l) v += part[spin * PS + (d * 8 + l) * SP + pl];
                pin(v);
                J.j[spin][d] = v;
            }
    }
    J.U = wave_sum(U);
    J.V = wave_sum(V);
    J.D2U = wave_sum(D2U);
    wave_sync();
    return J;
}

template <int NUP>
__device__ __forceinline__ JSums jastrow_terms(const DevSys &S, double *sm,
                                               const Layout<NUP> &L, const Tile<NUP> &T,
                                               int tid, PairTile<NUP> *rad = nullptr) {
    return jastrow_terms<NUP>(S, TabG(S), sm, L, T, tid, rad);
}

// Phases E-G: gradients and laplacian of log det through the inverse, Jastrow and
// potential, and the hpsi combination (wavefunction.f90:171-201), on the lean layout.
// Results in LDS: scal[0..4] = psil, d2psi, v, is, log|det|; dp = dpsi.  The inverse goes
// into the lane tile first; the Jastrow gradients are summed per particle through the work
// area, then the phases are rebuilt there for the plane-wave contraction, so that the lanes
// hold one set of partial sums at a time.  Partial sums travel as [quantity][lane row or
// column][particle] (row stride SP): the stores of the lane grid and the loads of lanes over
// particles both hit distinct banks.
template <int NUP, class TB>
__device__ void grad_lap(const DevSys &S, const TB &tb, double *sm, const Layout<NUP> &L,
                         const Tile<NUP> &T,
                         const DetInfo &D, const JSums &J, int tid) {
    constexpr int TR = Layout<NUP>::TR, SP = Layout<NUP>::SP, NP = Layout<NUP>::NP;
    constexpr int PS = 3 * 8 * SP;  // partial sums of one spin, [3][8][SP]
    const int nq = S.nq;
    const int r = T.r, c = T.c;
    const double *xs = sm + L.xs;
    const double2 *ph = reinterpret_cast<const double2 *>(sm + L.work);
    const double *qc = sm + L.qc;
    const double *sinv = sm + L.sinv;
    double *part = sm + L.work;
    const int *ri = T.ri, *cj = T.cj;
    const bool *rv = T.rv, *cv = T.cv;
    const bool singular = D.singular;
    const double logdet = D.logdet;
    const int sgn = D.sgn;
    double sv[TR][TR];
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TR; ++b)
            sv[a][b] = (rv[a] && cv[b] && !singular) ? sinv[cj[b] * SP + ri[a]] : 0.0;
    wave_sync();  // the inverse is in registers: the work area is free

    // radial part of the pairing functions, summed per particle through the work area
    const int pl = tid < NUP ? tid : 0;
    double lap = 0.0;
    double gs[2][3];
    {
        double pu[TR][3], pdn[TR][3];
#pragma unroll
        for (int a = 0; a < TR; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) pu[a][d] = pdn[a][d] = 0.0;
#pragma unroll
        for (int a = 0; a < TR; ++a)
#pragma unroll
            for (int b = 0; b < TR; ++b) {
                if (!(rv[a] && cv[b])) continue;
                const double *xi = xs + 3 * ri[a], *xj = xs + 3 * (NUP + cj[b]);
                double dx[3], r2 = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    dx[d] = xi[d] - xj[d];
                    dx[d] -= S.el * rint(dx[d] * S.eli);
                    r2 += dx[d] * dx[d];
                }
                const double ir = rsqrt_nr(fmax(r2, 1.0e-28)), rr = r2 * ir;
                if (!S.nobeta && rr <= S.rangeb) {
                    double w0, w1, w2, w3;
                    const int idx = lagrange(S.scaleb * rr, S.ntab, w0, w1, w2, w3);
                    const double2 A = tb.bt2(idx - 1), B = tb.bt2(idx), C = tb.bt2(idx + 1),
                                  D = tb.bt2(idx + 2);
                    const double2 bv = make_double2(w0 * A.x + w1 * B.x + w2 * C.x + w3 * D.x,
                                                    w0 * A.y + w1 * B.y + w2 * C.y + w3 * D.y);
                    const double t = sv[a][b] * bv.x * ir;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        pu[a][d] += t * dx[d];
                        pdn[b][d] -= t * dx[d];
                    }
                    lap += sv[a][b] * bv.y;
                }
                __builtin_amdgcn_sched_barrier(0);  // one pair's table loads in flight
            }
        if constexpr (TB::kCompact) {
#pragma unroll
            for (int spin = 0; spin < 2; ++spin) {
                put_spin_part<NUP>(part, T, spin, spin ? pdn : pu);
                wave_sync();
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    double v = 0.0;
#pragma unroll
                    for (int l = 0; l < 8; ++l) v += part[(d * 8 + l) * SP + pl];
                    gs[spin][d] = v;
                }
                wave_sync();
            }
        } else {
#pragma unroll
            for (int a = 0; a < TR; ++a) {
                if (rv[a]) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) part[(d * 8 + c) * SP + ri[a]] = pu[a][d];
                }
                if (cv[a]) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) part[PS + (d * 8 + r) * SP + cj[a]] = pdn[a][d];
                }
            }
        }
    }
    if constexpr (!TB::kCompact) {
        wave_sync();
#pragma unroll
        for (int spin = 0; spin < 2; ++spin)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double v = 0.0;
#pragma unroll
                for (int l = 0; l < 8; ++l) v += part[spin * PS + (d * 8 + l) * SP + pl];
                gs[spin][d] = v;
            }
    }
    wave_sync();
    const int4 *qmv = reinterpret_cast<const int4 *>(sm + L.qm);
    double2 *phw = reinterpret_cast<double2 *>(sm + L.work);
    if constexpr (!TB::kCompact) {
        phases_all<NUP>(S, xs, qmv, phw, tid);
        wave_sync();
    }

    double gu[TR][3], gd[TR][3];
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) gu[a][d] = gd[a][d] = 0.0;
    // plane-wave part: Sum_j Sinv(j, i) sin(q.r_ij) = s_i X_i - c_i Y_i with X = C^T Sinv
    // for the up particles, and the same with Z, W for the down particles, one pass over q
    // (the compact layout: the phases of chunks of qch q vectors, built in turn)
    const int qstep = TB::kCompact ? L.qch : nq;
    for (int q0 = 0; q0 < nq; q0 += qstep) {
    const int q1 = q0 + qstep < nq ? q0 + qstep : nq;
    if constexpr (TB::kCompact) {
        if (q0 > 0) wave_sync();  // done with the previous chunk
        phases_range<NUP>(S, xs, qmv, phw, q0, q1, tid);
        wave_sync();
    }
#pragma unroll 1
    for (int q = q0; q < q1; ++q) {
        const double *cq = qc + 5 * q;
        double2 zi[TR], zj[TR];
#pragma unroll
        for (int a = 0; a < TR; ++a) {
            zi[a] = ph[(q - q0) * NP + ri[a]];
            zj[a] = ph[(q - q0) * NP + NUP + cj[a]];
        }
        double ls = 0.0;
#pragma unroll
        for (int a = 0; a < TR; ++a) {
            double X = 0.0, Y = 0.0;
#pragma unroll
            for (int b = 0; b < TR; ++b) {
                X = fma(sv[a][b], zj[b].x, X);
                Y = fma(sv[a][b], zj[b].y, Y);
            }
            const double t = zi[a].y * X - zi[a].x * Y;
            gu[a][0] += cq[1] * t;
            gu[a][1] += cq[2] * t;
            gu[a][2] += cq[3] * t;
            ls = fma(zi[a].x, X, ls);
            ls = fma(zi[a].y, Y, ls);
        }
        lap += cq[4] * ls;
#pragma unroll
        for (int b = 0; b < TR; ++b) {
            double Z = 0.0, W = 0.0;
#pragma unroll
            for (int a = 0; a < TR; ++a) {
                Z += sv[a][b] * zi[a].y;
                W += sv[a][b] * zi[a].x;
            }
            const double u = zj[b].x * Z - zj[b].y * W;
            gd[b][0] -= cq[1] * u;
            gd[b][1] -= cq[2] * u;
            gd[b][2] -= cq[3] * u;
        }
    }
    }
    wave_sync();  // done with the phases

    // determinant gradients per particle
    if constexpr (TB::kCompact) {
#pragma unroll
        for (int spin = 0; spin < 2; ++spin) {
            put_spin_part<NUP>(part, T, spin, spin ? gd : gu);
            wave_sync();
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double g = 0.0;
#pragma unroll
                for (int l = 0; l < 8; ++l) g += part[(d * 8 + l) * SP + pl];
                gs[spin][d] += g;
                __builtin_amdgcn_sched_barrier(0);
            }
            wave_sync();
        }
    } else {
#pragma unroll
        for (int a = 0; a < TR; ++a) {
            if (rv[a]) {
#pragma unroll
                for (int d = 0; d < 3; ++d) part[(d * 8 + c) * SP + ri[a]] = gu[a][d];
            }
            if (cv[a]) {
#pragma unroll
                for (int d = 0; d < 3; ++d) part[PS + (d * 8 + r) * SP + cj[a]] = gd[a][d];
            }
        }
        wave_sync();
#pragma unroll
        for (int spin = 0; spin < 2; ++spin)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double g = 0.0;
#pragma unroll
                for (int l = 0; l < 8; ++l) g += part[spin * PS + (d * 8 + l) * SP + pl];
                gs[spin][d] += g;
                __builtin_amdgcn_sched_barrier(0);
            }
        wave_sync();  // dp shares the work area
    }
    double du2 = 0.0, dudp = 0.0;
    if (tid < NUP) {
#pragma unroll
        for (int spin = 0; spin < 2; ++spin) {
            const int p = spin * NUP + tid;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double g = singular ? 0.0 : gs[spin][d] * S.invsqrtm, j = J.j[spin][d];
                sm[L.dp + 3 * p + d] = g + j;
                du2 += j * j;
                dudp += j * g;
            }
        }
    }
    lap = wave_sum(lap);
    du2 = wave_sum(du2);
    dudp = wave_sum(dudp);
    if (tid == 0) {
        const double d2phi = singular ? 0.0 : 2.0 * lap * S.invm;
        sm[L.scal + 0] = logdet + J.U;
        sm[L.scal + 1] = d2phi + J.D2U + du2 + 2.0 * dudp;
        sm[L.scal + 2] = J.V;
        sm[L.scal + 3] = double(sgn);
        sm[L.scal + 4] = logdet;
    }
    wave_sync();
}

template <int NUP>
__device__ __forceinline__ void grad_lap(const DevSys &S, double *sm, const Layout<NUP> &L,
                                         const Tile<NUP> &T, const DetInfo &D, const JSums &J,
                                         int tid) {
    grad_lap<NUP>(S, TabG(S), sm, L, T, D, J, tid);
}

// copy the inverse left in LDS by build_inverse to global memory, dense [j][i], as FP64
// or FP32 (zeros for a singular matrix), in 16-byte stores
template <int NUP>
__device__ void store_sinv(const DevSys &S, const double *sm, const Layout<NUP> &L, bool singular,
                           char *dst, int tid) {
    constexpr int SP = Layout<NUP>::SP;
    const double *sv = sm + L.sinv;
    constexpr int NE = NUP * NUP;
    // element e of the dense slot, zero beyond the matrix (the slot is padded to whole
    // 128-byte lines, so the 16-byte groups below stay inside it)
    auto elem = [&](int e) __attribute__((always_inline)) {
        const int j = e / NUP, i = e - j * NUP;
        return (e < NE && !singular) ? sv[j * SP + i] : 0.0;
    };
    if (S.sinv_f32) {
        float4 *d = reinterpret_cast<float4 *>(dst);
        for (int g = tid; g < (NE + 3) / 4; g += 64)
            d[g] = make_float4(float(elem(4 * g)), float(elem(4 * g + 1)), float(elem(4 * g + 2)),
                               float(elem(4 * g + 3)));
    } else {
        double2 *d = reinterpret_cast<double2 *>(dst);
        for (int g = tid; g < (NE + 1) / 2; g += 64) d[g] = make_double2(elem(2 * g), elem(2 * g + 1));
    }
}

__host__ __device__ inline size_t sinv_bytes(const DevSys &S) {
    return size_t(S.n) * S.n * (S.sinv_f32 ? 4 : 8);
}

// per-walker slot of the stored inverse: the dense matrix padded to whole 128-byte lines (the
// branch copies slots with 16-byte vector loads)
__host__ __device__ inline size_t sinv_stride(const DevSys &S) {
    return (sinv_bytes(S) + 127) / 128 * 128;
}

template <int NUP, class TB = TabG>
__device__ __forceinline__ void hpsi_block(const DevSys &S, double *sm, const Layout<NUP> &L,
                                           int tid, const TB &tb) {
    const Tile<NUP> T(tid);
    PairTile<NUP> rad;
    const JSums J = jastrow_terms<NUP>(S, tb, sm, L, T, tid, &rad);
    const DetInfo D = build_inverse<NUP>(S, tb, sm, L, T, tid, &rad);
    grad_lap<NUP>(S, tb, sm, L, T, D, J, tid);
}

template <int NUP>
__device__ __forceinline__ void hpsi_block(const DevSys &S, double *sm, const Layout<NUP> &L,
                                           int tid) {
    hpsi_block<NUP>(S, sm, L, tid, TabG(S));
}

// Jastrow ratio exp(sum_j u(|x' - r_j|) - u(|r_i - r_j|)) over the opposite-spin partners
// of particle i (getphione, coshpot.f90:207-264); lanes over partners, wave-reduced.
template <int NUP>
__device__ double jastrow_logratio(const DevSys &S, const double *xs, int i, const double *xn,
                                   int tid) {
    constexpr int NP = 2 * NUP;
    double du = 0.0;
    for (int j = tid; j < NP; j += 64) {
        if ((j < NUP) == (i < NUP)) continue;
        for (int pass = 0; pass < 2; ++pass) {
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double xi = pass ? xn[d] : xs[3 * i + d];
                const double dx = wrapd(xi - xs[3 * j + d], S);
                r2 += dx * dx;
            }
            const double rj = fmax(1.0e-14, sqrt(r2));
            if (rj < S.range) {
                double w0, w1, w2, w3;
                const int idx = lagrange(S.scale * rj, S.ntab, w0, w1, w2, w3);
                if (idx < S.jfar) {  // u is zero on stencils past the healing distance
                    const double u = lookup4(S.jt, idx, w0, w1, w2, w3).y;
                    du += pass ? u : -u;
                }
            }
        }
    }
    return wave_sum(du);
}

// pair function phi(xa - xb) from the phases of both positions (allpair element)
__device__ __forceinline__ double phi_pair(const DevSys &S, const double *qc, const double2 *za,
                                           int sa, const double2 *zb, int sb, const double *xa,
                                           const double *xb) {
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double dx = wrapd(xa[d] - xb[d], S);
        r2 += dx * dx;
    }
    double f = S.alpha0 + beta_val(S, sqrt(r2));
    for (int q = 0; q < S.nq; ++q)
        f += qc[5 * q] * (za[q * sa].x * zb[q * sb].x + za[q * sa].y * zb[q * sb].y);
    return f;
}


// End of a DMC step for walker w, after hpsi of its new positions (in LDS): write the new
// state to B, the weight (step.f90:128-141), the estimator contributions (step.f90:
// 137-156) and the branch draw (step.f90:302-305).  s1 is the RNG state after the moves.
template <int NUP>
__device__ void finish_walker(const DevSys &S, double *sm, const Layout<NUP> &L,
                              const DevState &A, const DevState &B, const DevStep &T, int w,
                              int tid, uint64_t s1, int opcomp) {
    constexpr int M = 6 * NUP;
    double *xn = B.x + size_t(w) * M, *gn = B.dpsi + size_t(w) * M;
    for (int t = tid; t < M; t += 64) {
        xn[t] = sm[L.xs + t];
        gn[t] = sm[L.dp + t];
    }
    // the 42 estimator values are assembled in LDS (the work area is free by now) and
    // stored once, so that no two lanes store to one global address
    double *est = sm + L.est;
    if (tid < 42) est[tid] = 0.0;
    wave_sync();
    if (tid == 0) {
        const double psil2 = sm[L.scal], d2psi2 = sm[L.scal + 1], v2 = sm[L.scal + 2];
        const int is2 = int(sm[L.scal + 3]);
        const double en = -S.hbar * d2psi2 + v2;
        const double ego = -S.hbar * A.d2psi[w] + A.v[w];
        double wt = exp(-(0.5 * (ego + en) - S.etrial) * S.dt);
        double weight = fmin(S.wtcut, wt);
        const bool cross = is2 != A.is[w];
        if (cross && S.fn) {
            weight = 0.0;
            wt = 0.0;
        }
        est[12] = cross ? 1.0 : 0.0;  // node crossing (estimator 5): val, val^2, wt
        est[13] = cross ? 1.0 : 0.0;
        est[14] = 1.0;
        const double wt0 = A.wt0[w];
        const double wtg = wt * double(is2) * wt0;
        if (opcomp) {
            // estimators 1-4 and 6-14 of step.f90:142-156 (addval: val*wt, val^2*wt, wt)
            const double nf = S.nfac, sg = copysign(1.0, wtg);
            auto put = [&](int e, double val, double wgt) {
                est[3 * e] = val * wgt;
                est[3 * e + 1] = val * val * wgt;
                est[3 * e + 2] = wgt;
            };
            put(0, en * nf, wtg);
            put(1, (en - v2) * nf, wtg);
            put(2, v2 * nf, wtg);
            put(3, 0.0, wtg);
            put(5, 0.5 * (ego + en) * nf, 1.0);
            put(6, 1.0, 1.0);
            put(7, is2 * exp(psil2), 1.0);
            put(8, exp(psil2), 1.0);
            put(9, double(is2), 1.0);
            put(10, wt, 1.0);
            put(11, wtg, 1.0);
            put(12, sg, 1.0);
            put(13, sg, wtg);
        }
        s1 = (s1 * kM1 + kA1) & kMask48;  // the branch draw
        const double ub = double(s1 | 1ull) * 0x1p-48;
        const int iwt = int(weight + ub);
        T.count[w] = iwt;
        if (iwt >= 0 && iwt < 500)
            atomicAdd(reinterpret_cast<unsigned long long *>(&T.hist[iwt]), 1ull);
        B.irn[w] = s1;
        B.psil[w] = psil2;
        B.d2psi[w] = d2psi2;
        B.v[w] = v2;
        B.is[w] = is2;
        B.wt0[w] = wt0;
        B.weight[w] = weight;
    }
    wave_sync();
    if (tid < 42) T.est[size_t(w) * 42 + tid] = est[tid];
}

}  // namespace dmc
