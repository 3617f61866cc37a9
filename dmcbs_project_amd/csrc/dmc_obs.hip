// Observables of the resident population: the pair-distance histogram and the one-body
// density profile of addgofr (gofr.f90:52-98) and the spin-resolved structure factor of
// addsofq (sofq.f90:40-65), each walker weighted as in addoperators (operators.f90:102:
// weight * is * wt0; psil = psigl on this path).
//
// One 256-thread workgroup takes a contiguous range of walkers, one walker at a time:
//   * pairs and density: integer bin counts with LDS atomics, then each thread adds
//     weight * count to the FP64 partial sums of the bins it owns;
//   * S(q): exp(-i m 2 pi x_d / L) for m = 0..mmax of every coordinate in LDS (one sincos
//     per coordinate, powers by recurrence), then one thread per (q, spin) multiplies its
//     three axis factors per particle and sums rho(q) of that spin; the two threads of one
//     q are neighbouring lanes, and the up-spin lane takes rho_dn by a lane shuffle and adds
//     the six products to the partial sums of q, which it owns.
// Every partial sum has one owner thread and a fixed walker order, and the block partials
// are reduced in a fixed order, so results are reproducible run to run.
#include <hip/hip_runtime.h>

#include "dmc_kernels.h"

namespace dmc {
namespace {

constexpr int kObsThreads = 256;

// PG: partial sums in the block's row of the output instead of LDS (a template parameter,
// so that each variant addresses its partial sums in one known memory space)
template <bool PG>
__global__ __launch_bounds__(kObsThreads) void k_obs(ObsArgs a) {
    extern __shared__ double sm[];
    const int t = threadIdx.x;
    const int np = a.npart, nup = a.nup, ntab = a.ntab, nq = a.nq, mp = a.mmax + 1;
    const int ng = ntab > 0 ? 3 * (ntab + 1) + 6 * (2 * ntab + 1) : 0;
    // partial sums [nsum] in LDS, or (tables too large for it) in the block's own row of the
    // output, read and written by the same owner threads
    double *part = PG ? a.part + size_t(blockIdx.x) * a.nsum : sm;
    double *xs = PG ? sm : part + ((a.nsum + 1) & ~1);  // [np][3]
    double2 *pw = reinterpret_cast<double2 *>(xs + ((3 * np + 1) & ~1));  // [np][3][mp]
    int *cnt = reinterpret_cast<int *>(pw + (nq > 0 ? np * 3 * mp : 0));  // [ng]

    const int b = blockIdx.x;
    const int w0 = int(int64_t(a.nw) * b / a.nblk), w1 = int(int64_t(a.nw) * (b + 1) / a.nblk);
    for (int e = t; e < a.nsum; e += kObsThreads) part[e] = 0.0;
    for (int e = t; e < ng; e += kObsThreads) cnt[e] = 0;

    // pair p = (i, i + k mod np) with i = p mod np, k = p / np + 1, p < np (np - 1) / 2: k runs
    // to (np - 1) / 2, and for even np the last k = np / 2 stops at i = np / 2 - 1, so each
    // unordered pair comes once.  A thread's pairs step by kObsThreads: (i, k) advance by
    // (kObsThreads mod np, kObsThreads / np) with a carry, without a division per pair.
    const int npair = np * (np - 1) / 2, di = kObsThreads % np, dk = kObsThreads / np;
    const int i_start = t % np, k_start = t / np + 1;
    // a walker's coordinates (3 np <= 2 kObsThreads) are loaded one walker ahead
    double xr[2];
    auto load_walker = [&](int w) {
        const double *xw = a.x + size_t(w) * 3 * np;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int e = t + r * kObsThreads;
            xr[r] = e < 3 * np ? xw[e] : 0.0;
        }
    };
    if (w0 < w1) load_walker(w0);
    for (int w = w0; w < w1; ++w) {
        const double wt = a.weight[w] * double(a.is[w]) * a.wt0[w];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int e = t + r * kObsThreads;
            if (e < 3 * np) xs[e] = xr[r];
        }
        __syncthreads();
        if (w + 1 < w1) load_walker(w + 1);
        if (ntab > 0) {
            int i = i_start, k = k_start;
            for (int p = t; p < npair; p += kObsThreads) {
                int j = i + k;
                if (j >= np) j -= np;
                double r2 = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    // the minimum-image distance does not depend on how a tie (|dx| = L/2)
                    // is rounded, so rint serves for the reference's nint here
                    double dx = xs[3 * i + d] - xs[3 * j + d];
                    dx = dx - a.el * rint(dx * a.eli);
                    r2 += dx * dx;
                }
                const double r = sqrt(r2);
                int idx = r < a.half ? int(a.scale * r) : ntab;
                idx = idx < 0 ? 0 : (idx > ntab ? ntab : idx);
                const bool ui = i < nup, uj = j < nup;
                const int ch = ui && uj ? 1 : (!ui && !uj ? 2 : 0);
                atomicAdd(&cnt[3 * idx + ch], 1);
                i += di;
                k += dk;
                if (i >= np) {
                    i -= np;
                    ++k;
                }
            }
            // density profile along each axis (truncation toward zero, as the reference's
            // integer conversion)
            for (int e = t; e < 3 * np; e += kObsThreads) {
                const int i = e / 3, d = e - 3 * i;
                double dx = xs[e];
                dx = dx - a.el * round(dx * a.eli);
                int idx = int(a.scale * dx);
                idx = idx < -ntab ? -ntab : (idx > ntab ? ntab : idx);
                const int ch = i < nup ? d : d + 3;
                atomicAdd(&cnt[3 * (ntab + 1) + 6 * (idx + ntab) + ch], 1);
            }
        }
        if (nq > 0) {
            const int o = 1 + (ntab > 0 ? ng : 0);
            for (int e = t; e < 3 * np; e += kObsThreads) {
                double sn, cs;
                sincos(a.twopi_l * xs[e], &sn, &cs);
                double2 *pe = pw + e * mp;
                const double2 p1 = make_double2(cs, -sn);  // exp(-i theta)
                double2 pm = make_double2(1.0, 0.0);
                pe[0] = pm;
                for (int m = 1; m < mp; ++m) {
                    pm = make_double2(pm.x * p1.x - pm.y * p1.y, pm.x * p1.y + pm.y * p1.x);
                    pe[m] = pm;
                }
            }
            __syncthreads();
            // task k = 2 q + spin; 2 nq is even, so both lanes of a q run the same iterations
            for (int k = t; k < 2 * nq; k += kObsThreads) {
                const int q = k >> 1, s = k & 1;
                const int i0 = s ? nup : 0, i1 = s ? np : nup;
                const int4 m = a.qm[q];
                const int ax = m.x < 0 ? -m.x : m.x, ay = m.y < 0 ? -m.y : m.y,
                          az = m.z < 0 ? -m.z : m.z;
                const double sx = m.x < 0 ? -1.0 : 1.0, sy = m.y < 0 ? -1.0 : 1.0,
                             sz = m.z < 0 ? -1.0 : 1.0;
                double re = 0.0, im = 0.0;
                for (int i = i0; i < i1; ++i) {
                    const double2 *pi = pw + 3 * i * mp;
                    const double2 fx = pi[ax], fy = pi[mp + ay], fz = pi[2 * mp + az];
                    // a negative m takes the conjugate factor
                    const double fxi = sx * fx.y, fyi = sy * fy.y, fzi = sz * fz.y;
                    const double xyr = fx.x * fy.x - fxi * fyi, xyi = fx.x * fyi + fxi * fy.x;
                    re += xyr * fz.x - xyi * fzi;
                    im += xyr * fzi + xyi * fz.x;
                }
                const double dr = __shfl_xor(re, 1), di = __shfl_xor(im, 1);
                if (s == 0) {
                    const double z1 = re * re + im * im, z2 = dr * dr + di * di;
                    const double z3r = re * dr + im * di, z3i = im * dr - re * di;
                    double *ps = part + o + 6 * q;
                    ps[0] += z1 * wt;
                    ps[1] += z2 * wt;
                    ps[2] += z3r * wt;
                    ps[3] += z1 * z1 * wt;
                    ps[4] += z2 * z2 * wt;
                    ps[5] += (z3r * z3r - z3i * z3i) * wt;
                }
            }
        }
        __syncthreads();
        for (int e = t; e < ng; e += kObsThreads) {
            const int c = cnt[e];
            if (c) {
                part[1 + e] += wt * double(c);
                cnt[e] = 0;
            }
        }
        if (t == 0) part[0] += wt;
        __syncthreads();
    }
    if constexpr (!PG) {
        double *out = a.part + size_t(b) * a.nsum;
        for (int e = t; e < a.nsum; e += kObsThreads) out[e] = part[e];
    }
}

// out[s][e] = sum over rows r in segment s of in[r][e], rows in order
__global__ __launch_bounds__(kObsThreads) void k_obs_reduce(const double *in, int nrow,
                                                           int ncol, int seg, double *out) {
    const int e = blockIdx.x * kObsThreads + threadIdx.x;
    const int s = blockIdx.y;
    if (e >= ncol) return;
    const int r0 = s * seg, r1 = r0 + seg < nrow ? r0 + seg : nrow;
    double acc = 0.0;
#pragma unroll 8
    for (int r = r0; r < r1; ++r) acc += in[size_t(r) * ncol + e];
    out[size_t(s) * ncol + e] = acc;
}

}  // namespace

size_t obs_lds_bytes(int npart, int ntab, int nq, int mmax, bool part_global) {
    const int ng = ntab > 0 ? 3 * (ntab + 1) + 6 * (2 * ntab + 1) : 0;
    const int nsum = 1 + ng + 6 * nq;
    size_t d = (part_global ? 0 : size_t((nsum + 1) & ~1)) + size_t((3 * npart + 1) & ~1);
    d += 2 * size_t(nq > 0 ? npart * 3 * (mmax + 1) : 0);
    return d * sizeof(double) + sizeof(int) * size_t(ng);
}

int obs_blocks(int nw) { return nw < kObsBlocks ? nw : kObsBlocks; }

int launch_obs(const ObsArgs &a, double *tmp, double *out, hipStream_t st) {
    const size_t lds = obs_lds_bytes(a.npart, a.ntab, a.nq, a.mmax, a.part_global != 0);
    if (a.part_global)
        hipLaunchKernelGGL(k_obs<true>, dim3(a.nblk), dim3(kObsThreads), lds, st, a);
    else
        hipLaunchKernelGGL(k_obs<false>, dim3(a.nblk), dim3(kObsThreads), lds, st, a);
    // block partials [nblk][nsum] -> segments of 32 rows -> one row
    return launch_rows_reduce(a.part, a.nblk, a.nsum, 32, tmp, out, st);
}

int launch_rows_reduce(const double *in, int nrow, int ncol, int seg, double *tmp, double *out,
                       hipStream_t st) {
    const int nseg = (nrow + seg - 1) / seg;
    const int gx = (ncol + kObsThreads - 1) / kObsThreads;
    hipLaunchKernelGGL(k_obs_reduce, dim3(gx, nseg), dim3(kObsThreads), 0, st, in, nrow, ncol,
                       seg, tmp);
    hipLaunchKernelGGL(k_obs_reduce, dim3(gx, 1), dim3(kObsThreads), 0, st, tmp, nseg, ncol,
                       nseg, out);
    return int(hipGetLastError());
}

}  // namespace dmc
