// Single-particle-move DMC sweep with Sherman-Morrison inverse updates, one walker per
// 64-lane workgroup.
//
// A step moves the particles one at a time in order (0..npart-1).  For particle a the
// drift comes from grad ln psi at the current configuration (determinant part through
// column/row a of the inverse, Jastrow part over the partners), the proposal is
// r' = r + drift + sigma g with g from gaussian(3) of the walker's own ran1 stream, and the
// move is accepted with the Metropolis probability of the drift-diffusion Green's function
// (the form of step.f90:80-93), the fixed-node rule rejecting any move that changes the sign
// of the determinant.  An accepted move updates the inverse by a rank-1 Sherman-Morrison
// step: an up particle changes row a of S(i, j) = phi(r_up_i - r_dn_j), a down particle
// column a.  Every krefresh moves, and at the end of the sweep, the inverse and log|det|
// are rebuilt from scratch in FP64; the end-of-sweep rebuild also gives the FP64 local
// energy, weight, estimators and branch draw exactly as the all-particle step.
//
// R = double: every move in FP64.  R = float (mixed precision): the pair functions, Jastrow
// lookups, gradients and the Sherman-Morrison updates run in FP32 on FP32 copies of the
// inverse and phases; the refresh, the determinant-ratio sum used for the node test,
// positions, weights and accumulators stay FP64.
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

template <typename R>
__device__ __forceinline__ R wsum(R v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// R-typed table record (v, u, du, d2u) or (b, db, lap b, 0) interpolated at s
template <typename R>
__device__ __forceinline__ void tab_at(const double4 *t64, const float4 *t32, R s, int ntab,
                                       R &x, R &y, R &z, R &w) {
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
        x = w0 * a.x + w1 * b.x + w2 * c.x + w3 * d.x;
        y = w0 * a.y + w1 * b.y + w2 * c.y + w3 * d.y;
        z = w0 * a.z + w1 * b.z + w2 * c.z + w3 * d.z;
        w = w0 * a.w + w1 * b.w + w2 * c.w + w3 * d.w;
    } else {
        const float4 a = t32[idx - 1], b = t32[idx], c = t32[idx + 1], d = t32[idx + 2];
        x = w0 * a.x + w1 * b.x + w2 * c.x + w3 * d.x;
        y = w0 * a.y + w1 * b.y + w2 * c.y + w3 * d.y;
        z = w0 * a.z + w1 * b.z + w2 * c.z + w3 * d.z;
        w = w0 * a.w + w1 * b.w + w2 * c.w + w3 * d.w;
    }
}

// LDS of the sweep: the all-particle layout plus, for R = float, FP32 copies of the
// inverse, the phases and the q coefficients.
template <int NUP, typename R>
struct SmLayout {
    static constexpr int SP = Layout<NUP>::SP, R8 = Layout<NUP>::R8, NP = Layout<NUP>::NP;
    Layout<NUP> L;
    int sv, ph, qc, total;  // offsets in doubles
    __host__ __device__ explicit SmLayout(int nq) : L(nq) {
        if (sizeof(R) == 8) {
            sv = L.sinv;
            ph = L.work;
            qc = L.qc;
            total = L.total;
        } else {
            sv = L.total;
            ph = sv + (R8 * SP + 1) / 2;
            qc = ph + NP * nq;
            total = qc + (5 * nq + 1) / 2;
        }
    }
};

template <int NUP, typename R>
size_t sm_lds_bytes(int nq) { return sizeof(double) * SmLayout<NUP, R>(nq).total; }

// FP64 rebuild of the inverse (and phases); FP32 copies for the mixed variant
template <int NUP, typename R>
__device__ bool sm_refresh(const DevSys &S, double *sm, const SmLayout<NUP, R> &X,
                           const Tile<NUP> &T, int tid) {
    const DetInfo D = build_inverse<NUP>(S, sm, X.L, T, tid);
    if constexpr (sizeof(R) == 4) {
        constexpr int SP = Layout<NUP>::SP, NP = Layout<NUP>::NP;
        float *sv = reinterpret_cast<float *>(sm + X.sv);
        const double *s64 = sm + X.L.sinv;
        for (int e = tid; e < NUP * NUP; e += 64) {
            const int j = e / NUP, i = e - j * NUP;
            sv[j * SP + i] = float(s64[j * SP + i]);
        }
        float2 *ph = reinterpret_cast<float2 *>(sm + X.ph);
        const double2 *p64 = reinterpret_cast<const double2 *>(sm + X.L.work);
        for (int e = tid; e < NP * S.nq; e += 64) ph[e] = make_float2(float(p64[e].x), float(p64[e].y));
        __syncthreads();
    }
    return D.singular;
}

template <int NUP, typename R>
__global__ __launch_bounds__(64) void k_sm_step(DevSys S, DevState A, DevState B, DevStep T,
                                                int kref, int opcomp) {
    using R2 = typename Vec2<R>::T;
    extern __shared__ double sm[];
    const SmLayout<NUP, R> X(S.nq);
    const Layout<NUP> &L = X.L;
    constexpr int M = 6 * NUP, NP = 2 * NUP, SP = Layout<NUP>::SP;
    const int w = blockIdx.x, tid = threadIdx.x;
    if (w >= T.nt[0]) return;
    const int nq = S.nq;
    load_q<NUP>(S, sm, L, tid);
    if constexpr (sizeof(R) == 4) {
        float *q32 = reinterpret_cast<float *>(sm + X.qc);
        __syncthreads();
        for (int e = tid; e < 5 * nq; e += 64) q32[e] = float(sm[L.qc + e]);
    }
    double *xs = sm + L.xs;
    for (int t = tid; t < M; t += 64) xs[t] = wrapd(A.x[size_t(w) * M + t], S);
    __syncthreads();
    const Tile<NUP> Tl(tid);
    bool singular = sm_refresh<NUP, R>(S, sm, X, Tl, tid);

    R *sv = reinterpret_cast<R *>(sm + X.sv);
    R2 *ph = reinterpret_cast<R2 *>(sm + X.ph);
    const R *qc = reinterpret_cast<const R *>(sm + X.qc);
    R2 *phn = reinterpret_cast<R2 *>(sm + L.xn + 4);
    R *phinew = reinterpret_cast<R *>(sm + L.prow);
    R *colbuf = reinterpret_cast<R *>(sm + L.mcol);
    R *fbuf = reinterpret_cast<R *>(sm + L.dp);
    const R alpha0 = R(S.alpha0);
    const double sig = S.sigma, sig2 = S.sig2;
    uint64_t s = A.irn[w];
    int nacc = 0;

    for (int ip = 0; ip < NP; ++ip) {
        if (ip > 0 && ip % kref == 0) singular = sm_refresh<NUP, R>(S, sm, X, Tl, tid);
        const bool up = ip < NUP;
        const int a = up ? ip : ip - NUP;
        const double xa[3] = {xs[3 * ip], xs[3 * ip + 1], xs[3 * ip + 2]};
        // old gradient of particle ip: lanes b < NUP over the opposite-spin partners
        R go0 = 0, go1 = 0, go2 = 0, jo0 = 0, jo1 = 0, jo2 = 0, uo = 0, wgt = 0;
        int pb = 0;
        if (tid < NUP) {
            const int b = tid;
            pb = up ? NUP + b : b;
            wgt = up ? sv[b * SP + a] : sv[a * SP + b];
            double dx[3], r2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                dx[d] = wrapd(xa[d] - xs[3 * pb + d], S);
                r2 += dx[d] * dx[d];
            }
            const R ro = R(sqrt(r2));
            const R2 *za = ph + ip * nq, *zb = ph + pb * nq;
            for (int q = 0; q < nq; ++q) {
                const R2 u = za[q], v = zb[q];
                const R sab = u.y * v.x - u.x * v.y;
                go0 += qc[5 * q + 1] * sab;
                go1 += qc[5 * q + 2] * sab;
                go2 += qc[5 * q + 3] * sab;
            }
            if (!S.nobeta && ro <= R(S.rangeb)) {
                R bv, db, lb, pad;
                tab_at<R>(S.bt, S.bt32, R(S.scaleb) * ro, S.ntab, bv, db, lb, pad);
                const R t = db / ro;
                go0 += t * R(dx[0]);
                go1 += t * R(dx[1]);
                go2 += t * R(dx[2]);
            }
            go0 *= wgt;
            go1 *= wgt;
            go2 *= wgt;
            const R rj = ro > R(1.0e-14) ? ro : R(1.0e-14);
            if (rj < R(S.range)) {
                R vv, uu, du, d2u;
                tab_at<R>(S.jt, S.jt32, R(S.scale) * rj, S.ntab, vv, uu, du, d2u);
                uo = uu;
                const R t = du / rj;
                jo0 = R(dx[0]) * t;
                jo1 = R(dx[1]) * t;
                jo2 = R(dx[2]) * t;
            }
        }
        const double go[3] = {double(wsum(go0)), double(wsum(go1)), double(wsum(go2))};
        const double jo[3] = {double(wsum(jo0)), double(wsum(jo1)), double(wsum(jo2))};
        // proposal: gaussian(3) then the acceptance draw (5 ran1 draws per move)
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
            dro[d] = (go[d] + jo[d]) * sig2;
            if (S.cutdrift) dro[d] = fmin(S.dcut, fmax(-S.dcut, dro[d]));
            dxm[d] = dro[d] + sig * g3[d];
            xn[d] = wrapd(xa[d] + dxm[d], S);
        }
        // phases of the proposed position: lanes over q
        if (tid < nq) {
            double pc[3][5], ps[3][5];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if constexpr (sizeof(R) == 8) {
                    sincos(xn[d] * S.twopi_l, &ps[d][1], &pc[d][1]);
                } else {
                    float sf, cf;
                    sincosf(float(xn[d] * S.twopi_l), &sf, &cf);
                    ps[d][1] = sf;
                    pc[d][1] = cf;
                }
                pc[d][0] = 1.0;
                ps[d][0] = 0.0;
#pragma unroll
                for (int m = 2; m < 5; ++m) {
                    pc[d][m] = pc[d][m - 1] * pc[d][1] - ps[d][m - 1] * ps[d][1];
                    ps[d][m] = ps[d][m - 1] * pc[d][1] + pc[d][m - 1] * ps[d][1];
                }
            }
            const int4 m = reinterpret_cast<const int4 *>(sm + L.ints + 64)[tid];
            double cx, sx, cy, sy, cz, sz;
            cpow(pc[0], ps[0], m.x, cx, sx);
            cpow(pc[1], ps[1], m.y, cy, sy);
            cpow(pc[2], ps[2], m.z, cz, sz);
            const double cxy = cx * cy - sx * sy, sxy = sx * cy + cx * sy;
            phn[tid].x = R(cxy * cz - sxy * sz);
            phn[tid].y = R(sxy * cz + cxy * sz);
        }
        __syncthreads();
        // pair values and gradients at the proposed position
        double rho_part = 0.0;
        R gn0 = 0, gn1 = 0, gn2 = 0, jn0 = 0, jn1 = 0, jn2 = 0, dU = 0;
        if (tid < NUP) {
            double dx[3], r2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                dx[d] = wrapd(xn[d] - xs[3 * pb + d], S);
                r2 += dx[d] * dx[d];
            }
            const R rn = R(sqrt(r2));
            const R2 *zb = ph + pb * nq;
            R f = 0, g0 = 0, g1 = 0, g2 = 0;
            for (int q = 0; q < nq; ++q) {
                const R2 un = phn[q], v = zb[q];
                const R cab = un.x * v.x + un.y * v.y;
                const R sab = un.y * v.x - un.x * v.y;
                f += qc[5 * q] * cab;
                g0 += qc[5 * q + 1] * sab;
                g1 += qc[5 * q + 2] * sab;
                g2 += qc[5 * q + 3] * sab;
            }
            f += alpha0;
            if (!S.nobeta && rn <= R(S.rangeb)) {
                R bv, db, lb, pad;
                tab_at<R>(S.bt, S.bt32, R(S.scaleb) * rn, S.ntab, bv, db, lb, pad);
                f += bv;
                const R t = db / rn;
                g0 += t * R(dx[0]);
                g1 += t * R(dx[1]);
                g2 += t * R(dx[2]);
            }
            phinew[tid] = f;
            rho_part = double(wgt) * double(f);
            gn0 = wgt * g0;
            gn1 = wgt * g1;
            gn2 = wgt * g2;
            const R rj = rn > R(1.0e-14) ? rn : R(1.0e-14);
            R un = 0;
            if (rj < R(S.range)) {
                R vv, uu, du, d2u;
                tab_at<R>(S.jt, S.jt32, R(S.scale) * rj, S.ntab, vv, uu, du, d2u);
                un = uu;
                const R t = du / rj;
                jn0 = R(dx[0]) * t;
                jn1 = R(dx[1]) * t;
                jn2 = R(dx[2]) * t;
            }
            dU = un - uo;
        }
        const double rho = wave_sum(rho_part);  // FP64 sum: the node test uses its sign
        const double gn[3] = {double(wsum(gn0)), double(wsum(gn1)), double(wsum(gn2))};
        const double jn[3] = {double(wsum(jn0)), double(wsum(jn1)), double(wsum(jn2))};
        const double dlogj = double(wsum(dU));
        bool accept = !singular && !(S.fn && !(rho > 0.0)) && rho != 0.0;
        if (accept) {
            double logp = 2.0 * (log(fabs(rho)) + dlogj);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double drn = (gn[d] / rho + jn[d]) * sig2;
                if (S.cutdrift) drn = fmin(S.dcut, fmax(-S.dcut, drn));
                logp += (dro[d] + drn) * (dro[d] - drn - 2.0 * dxm[d]) / (2.0 * sig2);
            }
            accept = u[4] < exp(fmin(0.0, logp));
        }
        if (accept) {
            ++nacc;
            const R rr = R(rho);
            // capture column a (up mover) or row a (down mover) of the inverse
            if (tid < NUP) colbuf[tid] = up ? sv[tid * SP + a] : sv[a * SP + tid];
            __syncthreads();
            if (tid < NUP) {
                const int k = tid;
                R z = 0;
                if (up) {
                    for (int b = 0; b < NUP; ++b) z += phinew[b] * sv[b * SP + k];
                } else {
                    for (int b = 0; b < NUP; ++b) z += sv[k * SP + b] * phinew[b];
                }
                fbuf[k] = (z - (k == a ? R(1) : R(0))) / rr;
            }
            __syncthreads();
            for (int e = tid; e < NUP * NUP; e += 64) {
                const int r = e / NUP, c = e - r * NUP;
                sv[r * SP + c] -= up ? colbuf[r] * fbuf[c] : fbuf[r] * colbuf[c];
            }
            if (tid < 3) xs[3 * ip + tid] = xn[tid];
            if (tid < nq) ph[ip * nq + tid] = phn[tid];
        }
        __syncthreads();
    }
    // end of sweep: FP64 hpsi at the final positions, weights, estimators, branch draw
    hpsi_block<NUP>(S, sm, L, tid);
    if (tid == 0) {
        T.nacc[w] = nacc;
        atomicAdd(reinterpret_cast<unsigned long long *>(&T.moves[0]), (unsigned long long)nacc);
        atomicAdd(reinterpret_cast<unsigned long long *>(&T.moves[1]), (unsigned long long)NP);
    }
    finish_walker<NUP>(S, sm, L, A, B, T, w, tid, s, opcomp);
}

template <int NUP>
int sm_launch(int mixed, const DevSys &S, const DevState &A, const DevState &B,
              const DevStep &T, int krefresh, int opcomp, int grid, hipStream_t st) {
    if (mixed) {
        const size_t lds = sm_lds_bytes<NUP, float>(S.nq);
        hipLaunchKernelGGL((k_sm_step<NUP, float>), dim3(grid), dim3(64), lds, st, S, A, B, T,
                           krefresh, opcomp);
    } else {
        const size_t lds = sm_lds_bytes<NUP, double>(S.nq);
        hipLaunchKernelGGL((k_sm_step<NUP, double>), dim3(grid), dim3(64), lds, st, S, A, B, T,
                           krefresh, opcomp);
    }
    return hipGetLastError();
}

int launch_sm_step(int nup, int mixed, const DevSys &S, const DevState &A, const DevState &B,
                   const DevStep &T, int krefresh, int opcomp, int grid, hipStream_t st) {
    if (grid <= 0) return hipSuccess;
    switch (nup) {
        case 1: return sm_launch<1>(mixed, S, A, B, T, krefresh, opcomp, grid, st);
        case 7: return sm_launch<7>(mixed, S, A, B, T, krefresh, opcomp, grid, st);
        case 19: return sm_launch<19>(mixed, S, A, B, T, krefresh, opcomp, grid, st);
        case 33: return sm_launch<33>(mixed, S, A, B, T, krefresh, opcomp, grid, st);
    }
    return hipErrorInvalidValue;
}

int launch_adjoint(int, const DevSys &, const double *, int, const double *, double *, double *,
                   hipStream_t) {
    return hipErrorNotSupported;
}

}  // namespace dmc
