// Mixed-precision single-particle moves with two walkers per 64-lane wave.
//
// The sweep is the one of sm_moves_walker (dmc_sm.h): the same proposal, acceptance, node
// rule and Sherman-Morrison update, in FP32 with FP64 positions, the FP64 ratio sum and the
// jump-ahead draws.  What changes is the lane layout.  Half-wave h (lanes 32h .. 32h + 31)
// carries walker 2p + h, and lane k of a half:
//  * holds column k of the FP32 inverse while up particles move (row k while down
//    particles move), 33 rows; with n = 33 the 33rd column (row) is spread over the lanes,
//    lane k holding its element k, and the corner element sits in every lane;
//  * evaluates the pair of the mover with partner k (the proposed position in one pass, the
//    next mover's current position in another);
//  * forms the projection output k (one real component of one q) over all partners, and
//    the phases exp(i q.x) of its q: even lanes at the proposed position, odd lanes at the
//    next mover's position, exchanged between the two lanes of a pair.
// With n = 33 the pairs of every mover with partners 31 and 32 at its current position are
// formed once per half-sweep (lane k for mover k; mover 32 just before it moves), so that
// the pass over the next mover's partners takes lanes 0-30 and lane 31 evaluates the
// proposed position with partner 32: two passes of pair evaluations per move.  Sums run
// over half-waves (DPP within rows of 16, then a permlane16 swap), so the two walkers never
// mix; every value a walker's decision depends on is bitwise identical in the lanes of its
// half.  The phases of the mover's own spin are never stored: the next mover's phases come
// from the odd lanes of the previous move.
#pragma once
#include "dmc_sm.h"

namespace dmc {

// Sum over the 32 lanes of this lane's half-wave, in every lane of the half (rows of 16 by
// DPP, then the two rows of the half joined by a permlane16 swap).  Every lane must end with
// the same bits: the lanes of a walker use their own copy of the total (the Sherman-Morrison
// step of the spread column, the border pair value), so the adds are never contracted with
// the product that forms v.
__device__ __forceinline__ float half_total(float v) {
#pragma clang fp contract(off)
    v = row_sum16(v);
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false,
                                                    false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// two half-wave sums at once: a permlane16 swap packs the pair sums of a into the even rows
// and of b into the odd rows, each row is summed, and a second swap hands both totals to
// every lane of the half
__device__ __forceinline__ void half_total2(float &a, float &b) {
#pragma clang fp contract(off)
    auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false,
                                              false);
    float v = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    v = row_sum16(v);
    r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ __forceinline__ double half_total(double v) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long long b = __double_as_longlong(v);
        v += dbl_of(dpp_row(int(b & 0xffffffffll), k), dpp_row(int(b >> 32), k));
    }
    const long long b = __double_as_longlong(v);
    const unsigned lo = unsigned(b & 0xffffffffll), hi = unsigned(b >> 32);
    const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return dbl_of(int(l[0]), int(h[0])) + dbl_of(int(l[1]), int(h[1]));
}

// the value of lane j (< 32) of this lane's half-wave
__device__ __forceinline__ float half_lane(float v, int j, bool upper) {
    const float a = lane_value(v, j), b = lane_value(v, j + 32);
    return upper ? b : a;
}
__device__ __forceinline__ uint64_t half_lane(uint64_t v, int j, bool upper) {
    const uint64_t a = lane_value(v, j), b = lane_value(v, j + 32);
    return upper ? b : a;
}
__device__ __forceinline__ double half_lane(double v, int j, bool upper) {
    const double a = lane_value(v, j), b = lane_value(v, j + 32);
    return upper ? b : a;
}

// v of the other lane of an even/odd lane pair (DPP quad_perm 1,0,3,2)
__device__ __forceinline__ float pair_swap(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}

// LDS of one wave (offsets in doubles): two walker areas, then what the walkers share
template <int NUP>
struct PairLayout {
    static constexpr int NP = 2 * NUP;
    static constexpr int ST = NUP | 1;  // row stride (complex) of the partner phases, odd
    static constexpr int TS = NUP | 1;  // row stride (floats) of the transposition, odd
    static constexpr int NQMAX = 16;    // plane waves: 2 nq projection outputs on 32 lanes
    // per walker: positions [NP][3] FP64; the partners' phases [q][ST] complex FP32, which
    // the transposition of the inverse [NUP][TS] FP32 overlays at the switch of spins
    // (reaching into the weights, which are dead there); the mover's weights w_b and the
    // new pair values f_b (33 FP32 each, in 144 B); w_q exp(i q.x') (32 FP32); the draws of
    // two batches of 10 moves (the current one and the next), 32 B per move (the three
    // gaussians as FP32, the acceptance draw as FP64)
    static constexpr int XS = 0;
    static constexpr int PL = XS + 3 * NP;
    static constexpr int WB = PL + NQMAX * ST;
    static constexpr int FB = WB + 18;
    static constexpr int UB = FB + 18;
    static constexpr int RB = UB + 16;
    static constexpr int WALKER = RB + 2 * 4 * 10;  // two batch buffers
    static_assert(PL + (NUP * TS + 1) / 2 <= WALKER, "transposition area");
    static_assert(PL % 2 == 0 && WB % 2 == 0 && WALKER % 2 == 0, "16-byte alignment");
    // shared: per q the weights w_q, -w_q q (FP32), the integer q vectors, the jump-ahead
    // coefficients of draws 1..NJ (the draws of a batch of kRngBatch moves)
    static constexpr int NJ = 5 * 10;
    static constexpr int QK = 2 * WALKER;
    static constexpr int QM = QK + 2 * NQMAX;
    static constexpr int JMP = QM + 2 * NQMAX;
    static constexpr int TOTAL = JMP + 2 * NJ;
};

template <int NUP>
size_t mv_pair_lds_bytes() { return sizeof(double) * PairLayout<NUP>::TOTAL; }

// whether the two-walker moves take a system: FP32 inverse, 17..33 particles of each spin,
// at most 16 plane waves
__host__ __device__ inline bool mv_pair_fits(int nup, int nq) {
    return nup >= 17 && nup <= 33 && nq <= PairLayout<33>::NQMAX;
}

// phases exp(i q.r_b) of the partners b of this half's walker into its plane [q][ST]
// (lane k: partners k, k + 32)
template <int NUP>
__device__ __forceinline__ void partner_phases(const DevSys &S, const double *xs,
                                               const int4 *qm, float2 *pl, int k,
                                               bool colmode) {
    constexpr int ST = PairLayout<NUP>::ST;
    const int nq = S.nq;
    for (int b = k; b < NUP; b += 32) {
        const int p = (colmode ? NUP : 0) + b;
        float c[3][5], sn[3][5];
#pragma unroll
        for (int d = 0; d < 3; ++d) powers<float>(S, xs[3 * p + d], c[d], sn[d]);
        for (int q = 0; q < nq; ++q) {
            const int4 m = qm[q];
            const int mm[3] = {m.x, m.y, m.z};
            float cq[3], sq[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int am = mm[d] < 0 ? -mm[d] : mm[d];
                float cc = c[d][0], ss = sn[d][0];
#pragma unroll
                for (int j = 1; j < 5; ++j) {
                    const bool t = am == j;
                    cc = t ? c[d][j] : cc;
                    ss = t ? sn[d][j] : ss;
                    __builtin_amdgcn_sched_barrier(0);
                }
                cq[d] = cc;
                sq[d] = mm[d] < 0 ? -ss : ss;
            }
            const float cxy = cq[0] * cq[1] - sq[0] * sq[1], sxy = sq[0] * cq[1] + cq[0] * sq[1];
            float2 z;
            z.x = cxy * cq[2] - sxy * sq[2];
            z.y = sxy * cq[2] + cxy * sq[2];
            pl[q * ST + b] = z;
        }
    }
}

// Views of the moves' FP32 pair records {b, db, u, du} [ntab+1]: MtG reads them from global
// memory (the system's copy, through the vector L1), MtL from a copy staged in the LDS of a
// workgroup of several waves.
struct MtG {
    const float4 *t;
    __device__ float4 at(int i) const { return t[i]; }
};
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const v4f lds_cv4f;
struct MtL {
    lds_cv4f *t;
    __device__ float4 at(int i) const {
        const v4f v = t[i];
        return make_float4(v.x, v.y, v.z, v.w);
    }
};

// Pair terms of the mover at pa with a partner at pb (both FP64 positions in registers), as
// pair_terms of dmc_sm.h but without a branch: a pair beyond the table range (about half of
// all pairs of a uniform walker, scattered over the lanes) reads the last interval and takes
// zero weights, so that the lanes never split and the pair passes stay in one block.
template <class MT>
__device__ __forceinline__ void pair_terms_v(const DevSys &S, const MT &mt, const double (&pa)[3],
                                             const double (&pb)[3], float (&dx)[3], float &tb,
                                             float &tj, float &u, float &bv) {
    const float r = pair_dist<float>(S, pa, pb, dx);
    const bool in = r <= float(S.rangeb);
    const float sc = in ? float(S.scaleb) * r : 0.0f;
    const int ntab = S.ntab;
    int idx = int(sc);
    idx = idx > ntab - 2 ? ntab - 2 : idx;
    idx = idx < 1 ? 1 : idx;
    float w0, w1, w2, w3;
    lagrange_weights<float>(sc - float(idx), w0, w1, w2, w3);
    const float z = in ? 1.0f : 0.0f;
    w0 *= z;
    w1 *= z;
    w2 *= z;
    w3 *= z;
    const float4 a = mt.at(idx - 1), b = mt.at(idx), c = mt.at(idx + 1), d = mt.at(idx + 2);
    bv = w0 * a.x + w1 * b.x + w2 * c.x + w3 * d.x;
    const float db = w0 * a.y + w1 * b.y + w2 * c.y + w3 * d.y;
    u = w0 * a.z + w1 * b.z + w2 * c.z + w3 * d.z;
    const float du = w0 * a.w + w1 * b.w + w2 * c.w + w3 * d.w;
    const float rj = r > 1.0e-14f ? r : 1.0e-14f;
    tb = db * rcp_r<float>(r);
    tj = du * rcp_r<float>(rj);
}

// Moves m0..m1-1 of walkers 2 wp and 2 wp + 1 (nw walkers in all; a half whose walker does
// not exist repeats the last walker and writes nothing)
template <int NUP, class MT>
__device__ __forceinline__ void sm_moves_pair(const DevSys &S, const MT &mt, double *sm, int wp,
                                              int nw, int tid, const double *__restrict__ xin,
                                              double *xout, const uint64_t *irn_in,
                                              uint64_t *irn_out, char *sinv_g, int m0, int m1,
                                              const DevStep &T) {
    static_assert(NUP >= 2 && NUP <= 33, "two-walker moves: at most 33 partners");
    using L = PairLayout<NUP>;
    constexpr bool BD = NUP == 33;           // partners 31 and 32 apart
    constexpr int NA = NUP < 32 ? NUP : 32;  // partners at the proposed position, lanes
    constexpr int NB = BD ? 31 : NUP;        // partners of the next mover, lanes
    constexpr int M = 6 * NUP, ST = L::ST, TS = L::TS;
    const int nq = S.nq, no = 2 * nq;
    const bool upper = tid >= 32;
    const int k = tid & 31;
    const int wr = 2 * wp + (upper ? 1 : 0);
    const bool active = wr < nw;
    const int w = active ? wr : nw - 1;
    double *wa = sm + (upper ? L::WALKER : 0);
    double *xs = wa + L::XS;
    float2 *pl = reinterpret_cast<float2 *>(wa + L::PL);
    const float *plf = reinterpret_cast<const float *>(wa + L::PL);
    float *tr = reinterpret_cast<float *>(wa + L::PL);
    float *wb = reinterpret_cast<float *>(wa + L::WB);
    float *fb = reinterpret_cast<float *>(wa + L::FB);
    float *ub = reinterpret_cast<float *>(wa + L::UB);
    double *rb = wa + L::RB;
    float *qk = reinterpret_cast<float *>(sm + L::QK);
    int4 *qm = reinterpret_cast<int4 *>(sm + L::QM);
    uint64_t *jmp = reinterpret_cast<uint64_t *>(sm + L::JMP);

    for (int q = tid; q < nq; q += 64) {
        const int4 m = S.qm[q];
        const double wq = S.wq[q], tp = S.twopi_l;
        qm[q] = m;
        qk[4 * q] = float(wq);
        qk[4 * q + 1] = float(-wq * tp * m.x);
        qk[4 * q + 2] = float(-wq * tp * m.y);
        qk[4 * q + 3] = float(-wq * tp * m.z);
    }
    // (the system's jump-ahead table holds the first 1024 draws)
    if (tid < L::NJ) {
        jmp[tid] = S.ja1[tid];
        jmp[L::NJ + tid] = S.jc1[tid];
    }
    for (int t = k; t < M; t += 32) xs[t] = wrapd(xin[size_t(w) * M + t], S);
    wave_sync();

    // lane k is projection output k: component k & 1 of q = k >> 1
    const int cp = k & 1, qo = k >> 1;
    const bool ok = k < no;
    const float cw = ok ? qk[4 * qo] : 0.0f;
    float cg[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) cg[d] = ok ? qk[4 * qo + 1 + d] : 0.0f;
    const int4 cm = ok ? qm[qo] : make_int4(0, 0, 0, 0);
    const int pbase = ok ? 2 * ST * qo + cp : cp;  // component of P_b(q) at plf[pbase + 2 b]

    // the inverse: column (up movers) or row (down movers) k of this half's walker
    bool colmode = m0 < NUP;
    float Ac[NUP];
    float Ab = 0.0f, Acc = 0.0f;  // n = 33: element k of column (row) 32, and the corner
    {
        const float *g = reinterpret_cast<const float *>(sinv_g + size_t(w) * sinv_stride(S));
        const int kk = k < NUP ? k : 0;
#pragma unroll
        for (int b = 0; b < NUP; ++b) Ac[b] = g[colmode ? b * NUP + kk : kk * NUP + b];
        if constexpr (BD) {
            Ab = g[colmode ? kk * NUP + 32 : 32 * NUP + kk];
            Acc = g[32 * NUP + 32];
        }
    }

    double xp[3];  // position of partner k
    // n = 33: lane k's mover k with partner 31 (e1) and partner 32 (e2) at its current
    // position; mover 32's are put in lanes 0 (e1) and 1 (e2) just before it moves
    float e1b = 0, e1j = 0, e1u = 0, e1d[3] = {0, 0, 0};
    float e2b = 0, e2j = 0, e2u = 0, e2d[3] = {0, 0, 0};
    auto setup_half = [&]() {
        partner_phases<NUP>(S, xs, qm, pl, k, colmode);
        const int p0 = colmode ? NUP : 0, pb = p0 + (k < NUP ? k : 0);
#pragma unroll
        for (int d = 0; d < 3; ++d) xp[d] = xs[3 * pb + d];
        if constexpr (BD) {
            const int mv = (colmode ? 0 : NUP) + k;
            double pa[3], q31[3], q32[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                pa[d] = xs[3 * mv + d];
                q31[d] = xs[3 * (p0 + 31) + d];
                q32[d] = xs[3 * (p0 + 32) + d];
            }
            float bv;
            pair_terms_v(S, mt, pa, q31, e1d, e1b, e1j, e1u, bv);
            pair_terms_v(S, mt, pa, q32, e2d, e2b, e2j, e2u, bv);
        }
    };
    setup_half();

    const double sig = S.sigma, sig2 = S.sig2, half_isig2 = 0.5 / S.sig2;
    const double dcut = S.cutdrift ? S.dcut : __builtin_huge_val();
    const float alpha0 = float(S.alpha0);
    uint64_t s = irn_in[w];
    int nacc = 0;
    // the draws of a batch of up to 10 moves, formed every 10th move: lane 3m + i of the
    // half (m < 10) holds those of the batch's move m, lane 3m draws 5m + 1, 5m + 2 (a
    // gaussian pair), lane 3m + 1 draws 5m + 3, 5m + 4, lane 3m + 2 draw 5m + 5 (the
    // acceptance draw; the walker's state after the move)
    constexpr int kRngBatch = 10;
    static_assert(3 * kRngBatch <= 32 && 5 * kRngBatch == L::NJ, "draw batch");
    const int bm = k / 3, bi = k - 3 * bm;
    // the draws of cnt (1 .. kRngBatch) moves from the walker state s into batch buffer buf
    // (move m's 32 B at rb[4 (kRngBatch buf + m)]); s advances to the state after them
    auto fill_batch = [&](int buf, int cnt) __attribute__((always_inline)) {
        const bool on = bm < cnt;
        const bool on1 = on && bi < 2;
        const int t0 = on ? 5 * bm + 2 * bi : 0, t1 = on1 ? t0 + 1 : 0;  // < NJ
        const uint64_t ja0 = on ? jmp[t0] : 0, jc0 = on ? jmp[L::NJ + t0] : 0;
        const uint64_t ja1 = on1 ? jmp[t1] : 0, jc1 = on1 ? jmp[L::NJ + t1] : 0;
        const uint64_t sa = (ja0 * s + jc0) & kMask48, sb = (ja1 * s + jc1) & kMask48;
        const double ua = double(sa | 1ull) * 0x1p-48, ubd = double(sb | 1ull) * 0x1p-48;
        float sn, cs;
        sincos_turns(float(ua - rint(ua)), sn, cs);
        const float rad = sqrt_r<float>(-2.0f * log_r<float>(float(ubd)));
        const float gx = sn * rad, gy = cs * rad;
        double *rm = rb + 4 * (kRngBatch * buf + bm);
        float *rf = reinterpret_cast<float *>(rm);
        if (on) {
            if (bi == 0) {
                rf[0] = gx;
                rf[1] = gy;
            } else if (bi == 1) {
                rf[2] = gy;
            } else {
                rm[2] = ua;  // the acceptance draw
            }
        }
        s = half_lane(sa, 3 * cnt - 1, upper);  // the state after the batch
        wave_sync();
    };
    float ob = 0, oj = 0, ou = 0, odx[3] = {0, 0, 0};  // mover with partner k (k < NB)
    float pc = 0, ps = 0;  // exp(i q.x_a) of the mover's current position (this lane's q)
    bool have_old = false;
    wave_sync();
    if (m1 > m0) fill_batch(0, m1 - m0 < kRngBatch ? m1 - m0 : kRngBatch);  // the first batch
    for (int ip = m0; ip < m1; ++ip) {
        if (ip == NUP && colmode) {
            // down movers from here: rows of the inverse into the lanes through LDS (over the
            // phase plane, rebuilt afterwards), the up particles become the partners
            if (k < NUP) {
#pragma unroll
                for (int b = 0; b < NUP; ++b) tr[b * TS + k] = Ac[b];
                if constexpr (BD) tr[k * TS + 32] = Ab;
            }
            wave_sync();
            if (k < NUP) {
#pragma unroll
                for (int b = 0; b < NUP; ++b) Ac[b] = tr[k * TS + b];
                if constexpr (BD) Ab = tr[32 * TS + k];
            }
            wave_sync();
            colmode = false;
            setup_half();
            have_old = false;
            wave_sync();
        }
        const bool up = ip < NUP;
        const int a = up ? ip : ip - NUP;
        const int p0 = up ? NUP : 0;  // first partner (global index)
        const double xa[3] = {xs[3 * ip], xs[3 * ip + 1], xs[3 * ip + 2]};
        // the mover's determinant weights w_b to every lane of the half
        if (a < 32) {
            if (k == a) {
#pragma unroll
                for (int b = 0; b < NUP; ++b) wb[b] = Ac[b];
            }
        } else {
            wb[k] = Ab;
            if (k == 0) wb[32] = Acc;
        }
        if constexpr (BD) {
            if (a == 32) {
                // mover 32 with partners 31 (lane 0) and 32 (lane 1)
                if (k < 2) {
                    double q3[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) q3[d] = xs[3 * (p0 + 31 + k) + d];
                    float d3[3], b3, j3, u3, bv;
                    pair_terms_v(S, mt, xa, q3, d3, b3, j3, u3, bv);
                    if (k == 0) {
                        e1b = b3; e1j = j3; e1u = u3;
#pragma unroll
                        for (int d = 0; d < 3; ++d) e1d[d] = d3[d];
                    } else {
                        e2b = b3; e2j = j3; e2u = u3;
#pragma unroll
                        for (int d = 0; d < 3; ++d) e2d[d] = d3[d];
                    }
                }
            }
        }
        if (!have_old) {
            if (k < NB) {
                float bv;
                pair_terms_v(S, mt, xa, xp, odx, ob, oj, ou, bv);
            }
            float sn, cs;
            sincos_2pi<float>((cm.x * xa[0] + cm.y * xa[1] + cm.z * xa[2]) * S.eli, sn, cs);
            pc = cs;
            ps = sn;
        }
        wave_sync();
        const float wgt = k < NUP ? wb[k] : 0.0f;
        float w31 = 0.0f, w32 = 0.0f;
        if constexpr (BD) {
            w31 = wb[31];
            w32 = wb[32];
        }
        // the lanes that hold the mover's pairs with partners 31 and 32
        const bool own1 = BD && k == (a < 32 ? a : 0);
        const bool own2 = BD && k == (a < 32 ? a : 1);
        // projection X_k = sum_b w_b P_b(q) (component k & 1 of q = k >> 1)
        float X = 0.0f;
#pragma unroll
        for (int b = 0; b < NUP; ++b) X += wb[b] * plf[pbase + 2 * b];
        // drift at the current position
        float g0, g1, g2;
        {
            const float f = k < NB ? wgt * ob + oj : 0.0f;
            g0 = f * odx[0];
            g1 = f * odx[1];
            g2 = f * odx[2];
        }
        if constexpr (BD) {
            const float f1 = own1 ? w31 * e1b + e1j : 0.0f;
            const float f2 = own2 ? w32 * e2b + e2j : 0.0f;
            g0 += f1 * e1d[0] + f2 * e2d[0];
            g1 += f1 * e1d[1] + f2 * e2d[1];
            g2 += f1 * e1d[2] + f2 * e2d[2];
        }
        {
            const float t = cp ? -pc * X : ps * X;
            g0 += cg[0] * t;
            g1 += cg[1] * t;
            g2 += cg[2] * t;
        }
        half_total2(g0, g1);
        g2 = half_total(g2);
        const float go[3] = {g0, g1, g2};
        // gaussian(3) then the acceptance draw: five ran1 draws per move by jump-ahead from
        // the walker state, formed for a batch of moves at once and broadcast over the half
        float g3[3];
        double u4;
        {
            const int ml = (ip - m0) % kRngBatch;  // wave-uniform
            const int buf = ((ip - m0) / kRngBatch) & 1;
            // at a batch's first move, the next batch's draws into the other buffer, a batch
            // ahead of their use (off the chain of the moves that read them)
            if (ml == 0 && ip + kRngBatch < m1)
                fill_batch(buf ^ 1, m1 - ip - kRngBatch < kRngBatch ? m1 - ip - kRngBatch
                                                                     : kRngBatch);
            const double *rbb = rb + 4 * kRngBatch * buf;
            const float4 gq = reinterpret_cast<const float4 *>(rbb)[2 * ml];
            g3[0] = gq.x;
            g3[1] = gq.y;
            g3[2] = gq.z;
            u4 = rbb[4 * ml + 2];
        }
        float dro[3], dxm[3];
        double xn[3];
        {
            const float rs2 = float(sig2), rsig = float(sig), rcut = float(dcut);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                dro[d] = fminf(rcut, fmaxf(-rcut, go[d] * rs2));
                dxm[d] = fmaf(rsig, g3[d], dro[d]);
                xn[d] = wrapr(xa[d] + double(dxm[d]), S);
            }
        }
        const bool nxt = ip + 1 < m1 && ip + 1 != NUP;
        // Both pair passes right after the proposal, before the phases and the LDS round trip
        // of w_q exp(i q.x'): their table loads are then in flight while the phases are formed.
        // Pass A: the proposed position with partner k (lanes < NA).
        float adx[3] = {0, 0, 0}, atb = 0, atj = 0, auu = 0, abv = 0;
        if (k < NA) pair_terms_v(S, mt, xn, xp, adx, atb, atj, auu, abv);
        const float ou_cur = ou;  // the mover's old u with partner k (pass B overwrites it)
        // Pass B: the next mover of this spin at its current position (lanes < NB), and with
        // n = 33 the proposed position with partner 32 (lane 31)
        const bool lb = BD && k == 31;
        float xdx[3] = {0, 0, 0}, xtb = 0, xtj = 0, xu = 0, xbv = 0;
        if (nxt || BD) {  // wave-uniform; the lanes select their pair
            double pa[3], pb[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                pa[d] = lb ? xn[d] : xs[3 * (ip + 1) + d];
                pb[d] = lb ? xs[3 * (p0 + 32 < 2 * NUP ? p0 + 32 : 0) + d] : xp[d];
            }
            float d3[3], b3, j3, u3, bv;
            pair_terms_v(S, mt, pa, pb, d3, b3, j3, u3, bv);
            const bool keep = nxt && k < NB;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                xdx[d] = lb ? d3[d] : 0.0f;
                odx[d] = keep ? d3[d] : odx[d];
            }
            xtb = lb ? b3 : 0.0f;
            xtj = lb ? j3 : 0.0f;
            xu = lb ? u3 : 0.0f;
            xbv = lb ? bv : 0.0f;
            ob = keep ? b3 : ob;
            oj = keep ? j3 : oj;
            ou = keep ? u3 : ou;
        }
        // phases: even lanes at the proposed position, odd lanes at the next mover's
        float unc, uns, nxc, nxs;
        {
            double xq[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) xq[d] = (cp && nxt) ? xs[3 * (ip + 1) + d] : xn[d];
            float sn, cs;
            sincos_2pi<float>((cm.x * xq[0] + cm.y * xq[1] + cm.z * xq[2]) * S.eli, sn, cs);
            const float so = pair_swap(sn), co = pair_swap(cs);
            unc = cp ? co : cs;
            uns = cp ? so : sn;
            nxc = cp ? cs : co;
            nxs = cp ? sn : so;
            ub[k] = cw * (cp ? uns : unc);  // zero past the 2 nq outputs
        }
        wave_sync();
        // pass A's new pair values f_k and gradient terms
        double rho_part = 0.0;
        float gn0 = 0, gn1 = 0, gn2 = 0, jn0 = 0, jn1 = 0, jn2 = 0, dU = 0, fnew = 0;
        if (k < NA) {
            float f = alpha0 + abv;
            const float2 *pp = pl + k;
#pragma unroll 4
            for (int q = 0; q < nq; ++q) {
                const float2 v = pp[q * ST];
                f = fmaf(ub[2 * q], v.x, f);
                f = fmaf(ub[2 * q + 1], v.y, f);
            }
            fnew = f;
            rho_part = double(wgt) * double(f);
            const float wt = wgt * atb;
            gn0 = wt * adx[0];
            gn1 = wt * adx[1];
            gn2 = wt * adx[2];
            jn0 = atj * adx[0];
            jn1 = atj * adx[1];
            jn2 = atj * adx[2];
            dU = auu;
        }
        dU -= k < NB ? ou_cur : 0.0f;
        if constexpr (BD) dU -= (own1 ? e1u : 0.0f) + (own2 ? e2u : 0.0f);
        {
            const float t = cp ? -unc * X : uns * X;
            gn0 += cg[0] * t;
            gn1 += cg[1] * t;
            gn2 += cg[2] * t;
        }
        have_old = nxt;
        float f32 = 0.0f;
        if constexpr (BD) {
            // the new pair value of partner 32: its plane-wave part over the lanes of the half
            float t32 = ok ? ub[k] * plf[pbase + 64] : 0.0f;
            t32 = half_total(t32);
            f32 = alpha0 + half_lane(xbv, 31, upper) + t32;
            // lane 31 adds partner 32 (its x* terms are zero in the other lanes)
            rho_part += lb ? double(w32) * double(f32) : 0.0;
            const float wt = w32 * xtb;
            gn0 += wt * xdx[0];
            gn1 += wt * xdx[1];
            gn2 += wt * xdx[2];
            jn0 += xtj * xdx[0];
            jn1 += xtj * xdx[1];
            jn2 += xtj * xdx[2];
            dU += xu;
        }
        const double rho = half_total(rho_part);  // FP64 sum: the node test uses its sign
        const float irho = rcp_r<float>(float(rho));
        float dn0 = fmaf(gn0, irho, jn0), dn1 = fmaf(gn1, irho, jn1), dn2 = fmaf(gn2, irho, jn2);
        half_total2(dn0, dn1);
        half_total2(dn2, dU);
        const float dn[3] = {dn0, dn1, dn2};
        bool accept = !(S.fn && !(rho > 0.0)) && rho != 0.0;
        double logp_rec = __builtin_nan("");
        if (accept) {
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
        if (T.trace && k == 0 && active) {
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
        // Sherman-Morrison step (a rejected move updates by zero)
        fb[k] = fnew;  // past NUP: never read
        if constexpr (BD) fb[32] = f32;  // the same value from every lane of the half
        wave_sync();
        {
            float z = 0.0f;
#pragma unroll
            for (int b = 0; b < NUP; ++b) z += fb[b] * Ac[b];
            float fz = (z - (k == a ? 1.0f : 0.0f)) * irho;
            fz = accept ? fz : 0.0f;
#pragma unroll
            for (int b = 0; b < NUP; ++b) Ac[b] -= wb[b] * fz;
            if constexpr (BD) {
                const float z32 = half_total(fnew * Ab) + f32 * Acc;
                float fz32 = (z32 - (a == 32 ? 1.0f : 0.0f)) * irho;
                fz32 = accept ? fz32 : 0.0f;
                Ab -= wgt * fz32;
                Acc -= w32 * fz32;
            }
        }
        if (accept) {
            ++nacc;
            if (k < 3) xs[3 * ip + k] = k == 0 ? xn[0] : (k == 1 ? xn[1] : xn[2]);
        }
        pc = nxc;
        ps = nxs;
        wave_sync();
    }
    if (active) {
        for (int t = k; t < M; t += 32) xout[size_t(w) * M + t] = xs[t];
        if (k == 0) {
            irn_out[w] = s;
            T.nacc[w] = (m0 == 0 ? 0 : T.nacc[w]) + nacc;
        }
    }
    const bool both = 2 * wp + 1 < nw;
    const int n0 = __builtin_amdgcn_readlane(nacc, 0), n1 = __builtin_amdgcn_readlane(nacc, 32);
    if (tid == 0) {
        atomicAdd(reinterpret_cast<unsigned long long *>(&T.moves[0]),
                  (unsigned long long)(n0 + (both ? n1 : 0)));
        atomicAdd(reinterpret_cast<unsigned long long *>(&T.moves[1]),
                  (unsigned long long)((m1 - m0) * (both ? 2 : 1)));
    }
}

template <int NUP>
__device__ __forceinline__ void sm_moves_pair(const DevSys &S, double *sm, int wp, int nw,
                                              int tid, const double *__restrict__ xin,
                                              double *xout, const uint64_t *irn_in,
                                              uint64_t *irn_out, char *sinv_g, int m0, int m1,
                                              const DevStep &T) {
    sm_moves_pair<NUP>(S, MtG{S.mt32}, sm, wp, nw, tid, xin, xout, irn_in, irn_out, sinv_g, m0,
                       m1, T);
}

// waves (walker pairs) per workgroup of the move kernel that stages the pair records in LDS
constexpr int kMvWaves = 4;

// its LDS: the pair records [ntab+1] (16-byte aligned), then one two-walker area per wave
template <int NUP>
size_t mv_pair_tab_lds_bytes(int ntab) {
    return size_t(ntab + 1) * sizeof(float4) + kMvWaves * mv_pair_lds_bytes<NUP>();
}

}  // namespace dmc
