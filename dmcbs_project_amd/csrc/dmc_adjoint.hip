// Parameter adjoint: d log psi / d theta and d E_L / d theta for
// theta = (alpha_1 .. alpha_norb, b, d) on a batch of configurations, one walker per 64-lane
// workgroup, FP64.  The inverse is rebuilt here (build_inverse), never stored.
//
// With S(i, j) = phi(r_up_i - r_dn_j), G = grad_i S, L = lap S:
//   d log psi / d theta = sum_ij Sinv(j, i) dS(i, j)            (getderbcs,
//                                                              psislaterbcs.f90:449-524)
// and, since only the determinant part of E_L = -hbar (d2phi + d2u + |du|^2 + 2 du.dphi) + v
// depends on theta,
//   d2phi + 2 du.dphi = sum_ij Sinv(j, i) H(i, j),
//   H(i, j) = (2/m) L(i, j) + (2/sqrt m) (du_i - du_j) . G(i, j)
// (du the Jastrow gradient, already divided by sqrt m), so that with K = Sinv H Sinv
//   d E_L / d theta = -hbar [ sum_ij Sinv(j, i) dH(i, j) - sum_ij dS(i, j) K(j, i) ].
// dS, dG, dL are 1, 0, 0 for alpha_1; 2 sum_{q in shell} (cos, -q sin, -q^2 cos)(q.r_ij)
// for alpha_s; and the b- and d-derivatives of the beta table (an exact forward-mode
// derivative of the table build) for b and d.
#include "dmc_wave.h"

namespace dmc {

template <int NUP>
struct AdjLayout {
    static constexpr int R8 = Layout<NUP>::R8, SP = Layout<NUP>::SP, NP = Layout<NUP>::NP;
    Layout<NUP> L;
    int part, du, hm, tm, res, total;
    __host__ __device__ AdjLayout(int nq, int npar) : L(nq) {
        part = L.total;               // Jastrow partial sums [6][NUP][8]
        du = part + 48 * NUP;         // Jastrow gradient per particle [NP][3]
        hm = du + 3 * NP;             // H(i, j) at [i * SP + j]
        tm = hm + R8 * SP;            // T = H Sinv at [i * SP + k]
        res = tm + R8 * SP;           // [3][npar]: dlog, A, B
        total = res + 3 * npar;
    }
};

template <int NUP>
size_t adj_lds_bytes(int nq, int npar) { return sizeof(double) * AdjLayout<NUP>(nq, npar).total; }

template <int NUP>
__global__ __launch_bounds__(64) void k_adjoint(DevSys S, const double *__restrict__ x, int nw,
                                                int norb, double *dlog, double *delo) {
    extern __shared__ double sm[];
    constexpr int TR = Layout<NUP>::TR, SP = Layout<NUP>::SP, NP = Layout<NUP>::NP;
    const int npar = norb + 2;
    const AdjLayout<NUP> X(S.nq, npar);
    const Layout<NUP> &L = X.L;
    const int w = blockIdx.x, tid = threadIdx.x;
    if (w >= nw) return;
    constexpr int M = 6 * NUP;
    const int nq = S.nq;
    load_q<NUP>(S, sm, L, tid);
    for (int t = tid; t < M; t += 64) sm[L.xs + t] = x[size_t(w) * M + t];
    double *res = sm + X.res;
    for (int t = tid; t < 3 * npar; t += 64) res[t] = 0.0;
    wave_sync();
    const Tile<NUP> T(tid);
    const DetInfo D = build_inverse<NUP>(S, sm, L, T, tid);
    const double *xs = sm + L.xs, *sinv = sm + L.sinv, *qc = sm + L.qc;
    const double2 *ph = reinterpret_cast<const double2 *>(sm + L.work);
    const int4 *qm = reinterpret_cast<const int4 *>(sm + L.qm);
    const int r = T.r, c = T.c;

    // Jastrow gradient of every particle (the pair loop of getphi, coshpot.f90:165-202)
    double *pj = sm + X.part, *du = sm + X.du;
    {
        double ju[TR][3], jd[TR][3];
#pragma unroll
        for (int a = 0; a < TR; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) ju[a][d] = jd[a][d] = 0.0;
#pragma unroll
        for (int a = 0; a < TR; ++a)
#pragma unroll
            for (int b = 0; b < TR; ++b) {
                if (!(T.rv[a] && T.cv[b])) continue;
                double dx[3], r2 = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    dx[d] = wrapd(xs[3 * T.ri[a] + d] - xs[3 * (NUP + T.cj[b]) + d], S);
                    r2 += dx[d] * dx[d];
                }
                const double rj = fmax(1.0e-14, sqrt(r2));
                if (rj < S.range) {
                    double w0, w1, w2, w3;
                    const int idx = lagrange(S.scale * rj, S.ntab, w0, w1, w2, w3);
                    if (idx < S.jfar) {  // u' is zero on stencils past the healing distance
                        const double t = lookup4(S.jt, idx, w0, w1, w2, w3).z / (rj * S.sqrtm);
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            ju[a][d] += dx[d] * t;
                            jd[b][d] -= dx[d] * t;
                        }
                    }
                }
            }
#pragma unroll
        for (int a = 0; a < TR; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (T.rv[a]) pj[(d * NUP + T.ri[a]) * 8 + c] = ju[a][d];
                if (T.cv[a]) pj[((3 + d) * NUP + T.cj[a]) * 8 + r] = jd[a][d];
            }
        wave_sync();
        for (int p = tid; p < NP; p += 64) {
            const int off = p < NUP ? 0 : 3, pp = p < NUP ? p : p - NUP;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double s = 0.0;
#pragma unroll
                for (int l = 0; l < 8; ++l) s += pj[((off + d) * NUP + pp) * 8 + l];
                du[3 * p + d] = s;
            }
        }
        wave_sync();
    }

    const double cL = 2.0 * S.invm, cG = 2.0 * S.invsqrtm;
    // du_p . v for the Jastrow gradient of particle p (read from LDS where needed, so that
    // the register tiles below stay within the register file)
    auto dudot = [&](int p, double vx, double vy, double vz) {
        return du[3 * p] * vx + du[3 * p + 1] * vy + du[3 * p + 2] * vz;
    };

    // H(i, j) of the lane tile, written to LDS:
    //   H = cL lap S + cG (du_i - du_j) . grad S, accumulated per q as
    //   cL lq cos(q.r) + cG sin(q.r) (du_i - du_j) . gq  (gq = -w q, lq = -w |q|^2)
    double *hm = sm + X.hm, *tm = sm + X.tm;
    {
        double H[TR][TR];
#pragma unroll
        for (int a = 0; a < TR; ++a)
#pragma unroll
            for (int b = 0; b < TR; ++b) H[a][b] = 0.0;
        for (int q = 0; q < nq; ++q) {
            const double *cq = qc + 5 * q;
            double2 zi[TR], zj[TR];
            double gi[TR], gj[TR];
#pragma unroll
            for (int a = 0; a < TR; ++a) {
                zi[a] = ph[q * NP + T.ri[a]];
                zj[a] = ph[q * NP + NUP + T.cj[a]];
                gi[a] = cG * dudot(T.ri[a], cq[1], cq[2], cq[3]);
                gj[a] = cG * dudot(NUP + T.cj[a], cq[1], cq[2], cq[3]);
            }
            const double lq = cL * cq[4];
#pragma unroll
            for (int a = 0; a < TR; ++a)
#pragma unroll
                for (int b = 0; b < TR; ++b) {
                    const double cab = zi[a].x * zj[b].x + zi[a].y * zj[b].y;
                    const double sab = zi[a].y * zj[b].x - zi[a].x * zj[b].y;
                    H[a][b] += lq * cab + (gi[a] - gj[b]) * sab;
                }
        }
#pragma unroll
        for (int a = 0; a < TR; ++a)
#pragma unroll
            for (int b = 0; b < TR; ++b) {
                double dx[3], r2 = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    dx[d] = wrapd(xs[3 * T.ri[a] + d] - xs[3 * (NUP + T.cj[b]) + d], S);
                    r2 += dx[d] * dx[d];
                }
                const double rr = sqrt(r2);
                double h = H[a][b];
                if (!S.nobeta && rr <= S.rangeb) {
                    double w0, w1, w2, w3;
                    const int idx = lagrange(S.scaleb * rr, S.ntab, w0, w1, w2, w3);
                    const double4 bv = lookup4(S.bt, idx, w0, w1, w2, w3);
                    const double g = dudot(T.ri[a], dx[0], dx[1], dx[2]) -
                                     dudot(NUP + T.cj[b], dx[0], dx[1], dx[2]);
                    h += cL * bv.z + cG * bv.y * g / rr;
                }
                if (T.rv[a] && T.cv[b]) hm[T.ri[a] * SP + T.cj[b]] = h;
            }
    }
    wave_sync();
    // T = H Sinv (up x up), then KT(i, j) = K(j, i) = sum_m Sinv(j, m) T(m, i)
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TR; ++b) {
            double s = 0.0;
            for (int j = 0; j < NUP; ++j) s += hm[T.ri[a] * SP + j] * sinv[j * SP + T.cj[b]];
            if (T.rv[a] && T.cv[b]) tm[T.ri[a] * SP + T.cj[b]] = s;  // T(i = row a, k = col b)
        }
    wave_sync();
    // K(j, i) of the tile into the H area (H is consumed), the inverse tile into registers
    double *km = hm;
    double sv[TR][TR];
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TR; ++b) {
            double s = 0.0;
            const int i = T.ri[a], j = T.cj[b];
            for (int m = 0; m < NUP; ++m) s += sinv[j * SP + m] * tm[m * SP + i];
            const bool ok = T.rv[a] && T.cv[b] && !D.singular;
            sv[a][b] = ok ? sinv[j * SP + i] : 0.0;
            if (T.rv[a] && T.cv[b]) km[i * SP + j] = ok ? s : 0.0;
        }
    wave_sync();
    auto kt = [&](int a, int b) {
        return (T.rv[a] && T.cv[b]) ? km[T.ri[a] * SP + T.cj[b]] : 0.0;
    };

    // alpha_1 (dS = 1) and the beta parameters b, d (derivative tables)
    {
        double dl0 = 0.0, a0 = 0.0, dlb = 0.0, ab = 0.0, bb = 0.0, dld = 0.0, ad = 0.0, bd = 0.0;
#pragma unroll
        for (int a = 0; a < TR; ++a)
#pragma unroll
            for (int b = 0; b < TR; ++b) {
                dl0 += sv[a][b];
                a0 += kt(a, b);
                if (!(T.rv[a] && T.cv[b])) continue;
                double dx[3], r2 = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    dx[d] = wrapd(xs[3 * T.ri[a] + d] - xs[3 * (NUP + T.cj[b]) + d], S);
                    r2 += dx[d] * dx[d];
                }
                const double rr = sqrt(r2);
                if (S.nobeta || rr > S.rangeb) continue;
                double w0, w1, w2, w3;
                const int idx = lagrange(S.scaleb * rr, S.ntab, w0, w1, w2, w3);
                const double4 vb = lookup4(S.btb, idx, w0, w1, w2, w3);
                const double4 vd = lookup4(S.btd, idx, w0, w1, w2, w3);
                const double gdu = (dudot(T.ri[a], dx[0], dx[1], dx[2]) -
                                    dudot(NUP + T.cj[b], dx[0], dx[1], dx[2])) / rr;
                dlb += sv[a][b] * vb.x;
                ab += kt(a, b) * vb.x;
                bb += sv[a][b] * (cL * vb.z + cG * gdu * vb.y);
                dld += sv[a][b] * vd.x;
                ad += kt(a, b) * vd.x;
                bd += sv[a][b] * (cL * vd.z + cG * gdu * vd.y);
            }
        dl0 = wave_sum(dl0);
        a0 = wave_sum(a0);
        dlb = wave_sum(dlb);
        ab = wave_sum(ab);
        bb = wave_sum(bb);
        dld = wave_sum(dld);
        ad = wave_sum(ad);
        bd = wave_sum(bd);
        if (tid == 0) {
            res[0] += dl0;
            res[npar] += a0;
            res[norb] += dlb;
            res[npar + norb] += ab;
            res[2 * npar + norb] += bb;
            res[norb + 1] += dld;
            res[npar + norb + 1] += ad;
            res[2 * npar + norb + 1] += bd;
        }
    }
    // alpha_s, s >= 2: plane waves, accumulated per q and added to their shell
    for (int q = 0; q < nq; ++q) {
        const int4 m = qm[q];
        const int sh = m.w;
        // -q from the integer vector (independent of alpha_s, which may be zero)
        const double nx = -S.twopi_l * m.x, ny = -S.twopi_l * m.y, nz = -S.twopi_l * m.z;
        double2 zi[TR], zj[TR];
        double qi[TR], qj[TR];
#pragma unroll
        for (int a = 0; a < TR; ++a) {
            zi[a] = ph[q * NP + T.ri[a]];
            zj[a] = ph[q * NP + NUP + T.cj[a]];
            qi[a] = dudot(T.ri[a], nx, ny, nz);
            qj[a] = dudot(NUP + T.cj[a], nx, ny, nz);
        }
        const double lq = -(nx * nx + ny * ny + nz * nz);
        double dls = 0.0, as = 0.0, bs = 0.0;
#pragma unroll
        for (int a = 0; a < TR; ++a)
#pragma unroll
            for (int b = 0; b < TR; ++b) {
                const double cab = zi[a].x * zj[b].x + zi[a].y * zj[b].y;
                const double sab = zi[a].y * zj[b].x - zi[a].x * zj[b].y;
                dls += sv[a][b] * cab;
                as += kt(a, b) * cab;
                bs += sv[a][b] * (cL * lq * cab + cG * (qi[a] - qj[b]) * sab);
            }
        dls = wave_sum(dls);
        as = wave_sum(as);
        bs = wave_sum(bs);
        if (tid == 0) {
            res[sh] += 2.0 * dls;
            res[npar + sh] += 2.0 * as;
            res[2 * npar + sh] += 2.0 * bs;
        }
    }
    wave_sync();
    for (int p = tid; p < npar; p += 64) {
        const bool ok = !D.singular;
        dlog[size_t(w) * npar + p] = ok ? res[p] : 0.0;
        if (delo)
            delo[size_t(w) * npar + p] = ok ? -S.hbar * (res[2 * npar + p] - res[npar + p]) : 0.0;
    }
}

int launch_adjoint(int nup, const DevSys &S, const double *x, int nw, int norb, double *dlogpsi,
                   double *delocal, hipStream_t st) {
    if (nw <= 0) return hipSuccess;
    const int npar = norb + 2;
    switch (nup) {
        case 1: hipLaunchKernelGGL(k_adjoint<1>, dim3(nw), dim3(64), adj_lds_bytes<1>(S.nq, npar), st, S, x, nw, norb, dlogpsi, delocal); break;
        case 7: hipLaunchKernelGGL(k_adjoint<7>, dim3(nw), dim3(64), adj_lds_bytes<7>(S.nq, npar), st, S, x, nw, norb, dlogpsi, delocal); break;
        case 19: hipLaunchKernelGGL(k_adjoint<19>, dim3(nw), dim3(64), adj_lds_bytes<19>(S.nq, npar), st, S, x, nw, norb, dlogpsi, delocal); break;
        case 33: hipLaunchKernelGGL(k_adjoint<33>, dim3(nw), dim3(64), adj_lds_bytes<33>(S.nq, npar), st, S, x, nw, norb, dlogpsi, delocal); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dmc
