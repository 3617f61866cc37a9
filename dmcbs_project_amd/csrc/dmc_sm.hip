// Single-particle-move DMC sweep with Sherman-Morrison inverse updates.
//
// A step moves the particles one at a time in order (0..npart-1).  For particle a the
// drift comes from grad ln psi at the current configuration (determinant part through
// column/row a of the inverse, Jastrow part over the opposite-spin partners), cut to
// +-sigma*dcut per component when dcut >= 0 (step.f90:32-37,66-67); the proposal is
// r' = r + drift + sigma g with g from gaussian(3) of the walker's own ran1 stream, and the
// move is accepted with the Metropolis probability of the drift-diffusion Green's function
// (the form of step.f90:80-93); the fixed-node rule rejects any move that changes the sign
// of the determinant.  An accepted move updates the inverse by a rank-1 Sherman-Morrison
// step: an up particle changes row a of S(i, j) = phi(r_up_i - r_dn_j), a down particle
// column a.
//
// Three kernels, so that each gets the registers it needs:
//  * k_sm_moves32/64: moves m0..m1-1 of one walker per 64-lane workgroup, from the inverse
//    the walker carries (FP32 or FP64);
//  * k_sm_refresh: FP64 rebuild of the inverse (every krefresh moves inside a sweep);
//  * k_sm_end: at the end of the sweep the FP64 hpsi of the final positions gives the
//    local energy, weight, estimators and branch draw exactly as the all-particle step, and
//    the fresh FP64 inverse that the walker carries into the next step.
//
// R = double: every move in FP64.  R = float (mixed precision): pair functions, Jastrow
// lookups, gradients and the Sherman-Morrison updates run in FP32 on an FP32 inverse and
// FP32 phases; the refresh, the determinant-ratio sum behind the node test, positions,
// weights and accumulators stay FP64.
#include "dmc_sm2.h"

namespace dmc {

// Kernels of one system: walker w = workgroup index, one 64-lane wave per walker.
// FP32 moves: the compiler's own register budget (139 VGPRs at N = 66, three waves per SIMD)
// times best; FP64 moves: two waves per SIMD asked for (254 VGPRs, no spills), 8.2 ms
// against 12.6 ms at the one wave per SIMD the compiler picks by itself.
template <int NUP, int NQC>
__global__ __launch_bounds__(64) void k_sm_moves32(DevSys S, const double *__restrict__ xin,
                                                   double *xout, const uint64_t *irn_in,
                                                   uint64_t *irn_out, char *sinv_g, int m0,
                                                   int m1, DevStep T) {
    extern __shared__ double sm[];
    const int w = blockIdx.x;
    if (w >= T.nt[0]) return;
    sm_moves_walker<NUP, float, NQC>(S, sm, w, threadIdx.x, xin, xout, irn_in, irn_out, sinv_g,
                                     m0, m1, T);
}
template <int NUP, int NQC>
__global__ __launch_bounds__(64, 2) void k_sm_moves64(DevSys S, const double *__restrict__ xin,
                                                      double *xout, const uint64_t *irn_in,
                                                      uint64_t *irn_out, char *sinv_g, int m0,
                                                      int m1, DevStep T) {
    extern __shared__ double sm[];
    const int w = blockIdx.x;
    if (w >= T.nt[0]) return;
    sm_moves_walker<NUP, double, NQC>(S, sm, w, threadIdx.x, xin, xout, irn_in, irn_out, sinv_g,
                                      m0, m1, T);
}

// FP32 moves of two walkers per wave (workgroup p: walkers 2p and 2p + 1)
template <int NUP>
__global__ __launch_bounds__(64) void k_sm_moves32x2(DevSys S, const double *__restrict__ xin,
                                                     double *xout, const uint64_t *irn_in,
                                                     uint64_t *irn_out, char *sinv_g, int m0,
                                                     int m1, DevStep T) {
    extern __shared__ double sm[];
    const int nw = T.nt[0];
    const int wp = blockIdx.x;
    if (2 * wp >= nw) return;
    sm_moves_pair<NUP>(S, sm, wp, nw, threadIdx.x, xin, xout, irn_in, irn_out, sinv_g, m0, m1, T);
}

// The same moves with the FP32 pair records staged in LDS: workgroups of kMvWaves waves (two
// walkers each) copy the 16 KB table once, then each wave moves its pair of walkers.
template <int NUP>
__global__ __launch_bounds__(64 * kMvWaves) void k_sm_moves32x2_tab(
    DevSys S, const double *__restrict__ xin, double *xout, const uint64_t *irn_in,
    uint64_t *irn_out, char *sinv_g, int m0, int m1, DevStep T) {
    extern __shared__ double smem[];
    const int n1 = S.ntab + 1;
    float4 *tab = reinterpret_cast<float4 *>(smem);
    for (int i = threadIdx.x; i < n1; i += 64 * kMvWaves) tab[i] = S.mt32[i];
    __syncthreads();
    const int nw = T.nt[0];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wp = blockIdx.x * kMvWaves + wave;
    if (2 * wp >= nw) return;
    double *sm = smem + 2 * n1 + wave * PairLayout<NUP>::TOTAL;
    const MtL mt{(lds_cv4f *)(smem)};
    sm_moves_pair<NUP>(S, mt, sm, wp, nw, threadIdx.x & 63, xin, xout, irn_in, irn_out, sinv_g,
                       m0, m1, T);
}

// FP64 rebuild of the inverse of positions x (every krefresh moves)
template <int NUP>
__global__ __launch_bounds__(64) void k_sm_refresh(DevSys S, const double *__restrict__ x,
                                                   char *sinv_g, DevStep T) {
    extern __shared__ double sm[];
    const int w = blockIdx.x;
    if (w >= T.nt[0]) return;
    sm_refresh_walker<NUP>(S, sm, w, threadIdx.x, x, sinv_g);
}

// end of the sweep: FP64 hpsi of the final positions B.x, the fresh inverse into B.sinv,
// weight against the energy of A, estimators and branch draw (state after the moves in
// B.irn).  Two waves per SIMD asked for: the compiler spills some of the gradient pass to
// scratch, which costs less than the latency a single wave per SIMD leaves exposed (at N=66
// the kernel took 3.0 ms at one wave per SIMD, 2.0 ms at two).
template <int NUP>
__global__ __launch_bounds__(64, 2) void k_sm_end(DevSys S, DevState A, DevState B, DevStep T,
                                               int opcomp) {
    extern __shared__ double sm[];
    const int w = blockIdx.x;
    if (w >= T.nt[0]) return;
    sm_end_walker<NUP>(S, sm, w, threadIdx.x, A, B, T, opcomp);
}

// end of the sweep with the radial tables staged in LDS: workgroups of kTabWaves walkers (one
// per wave) that copy the tables once for all of them.  The walkers of a workgroup share one
// copy; the walker areas are in the compact layout.  (A persistent form, each wave looping
// over walkers, lets the compiler hoist the walker's invariant values out of the loop, where
// they stay live through the whole walker and spill: 2.1 ms against 1.76 ms for this form at
// N = 66 with 32k walkers.)
template <int NUP>
__global__ __launch_bounds__(64 * kTabWaves) void k_sm_end_tab(DevSys S, DevState A, DevState B,
                                                               DevStep T, int opcomp) {
    extern __shared__ double sm[];
    stage_tables(S, sm, threadIdx.x, 64 * kTabWaves);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), tid = threadIdx.x & 63;
    const int w = blockIdx.x * kTabWaves + wave;
    if (w >= T.nt[0]) return;
    const TabL tb(S, sm);
    double *smw = sm + TabLayout(S.ntab, S.jfar).total + wave * compact_stride<NUP>(S.nq);
    sm_end_walker<NUP>(S, tb, smw, w, tid, A, B, T, opcomp);
}

template <int NUP, typename R, int NQC>
int sm_moves_tq(const DevSys &S, const double *xin, double *xout, const uint64_t *irn_in,
                uint64_t *irn_out, char *sinv, int m0, int m1, const DevStep &T, int grid,
                hipStream_t st) {
    const size_t lds = mv_lds_bytes<NUP, R>(S.nq);
    if constexpr (sizeof(R) == 8)
        hipLaunchKernelGGL((k_sm_moves64<NUP, NQC>), dim3(grid), dim3(64), lds, st, S, xin, xout,
                           irn_in, irn_out, sinv, m0, m1, T);
    else
        hipLaunchKernelGGL((k_sm_moves32<NUP, NQC>), dim3(grid), dim3(64), lds, st, S, xin, xout,
                           irn_in, irn_out, sinv, m0, m1, T);
    return hipGetLastError();
}

template <int NUP>
int sm_moves_t(const DevSys &S, const double *xin, double *xout, const uint64_t *irn_in,
               uint64_t *irn_out, char *sinv, int m0, int m1, const DevStep &T, int grid,
               hipStream_t st) {
    if (S.nq > 64) return hipErrorInvalidValue;  // four projection chunks of 32
    if constexpr (NUP >= 17) {
        if (S.sinv_f32 && mv_pair_fits(NUP, S.nq)) {
            const int pairs = (grid + 1) / 2;
            if (S.mv_lds)
                hipLaunchKernelGGL(k_sm_moves32x2_tab<NUP>,
                                   dim3((pairs + kMvWaves - 1) / kMvWaves), dim3(64 * kMvWaves),
                                   mv_pair_tab_lds_bytes<NUP>(S.ntab), st, S, xin, xout, irn_in,
                                   irn_out, sinv, m0, m1, T);
            else
                hipLaunchKernelGGL(k_sm_moves32x2<NUP>, dim3(pairs), dim3(64),
                                   mv_pair_lds_bytes<NUP>(), st, S, xin, xout, irn_in, irn_out,
                                   sinv, m0, m1, T);
            return hipGetLastError();
        }
    }
    const bool one = mv_chunks(S.nq) == 1;
    if (S.sinv_f32)
        return one ? sm_moves_tq<NUP, float, 1>(S, xin, xout, irn_in, irn_out, sinv, m0, m1, T, grid, st)
                   : sm_moves_tq<NUP, float, 4>(S, xin, xout, irn_in, irn_out, sinv, m0, m1, T, grid, st);
    return one ? sm_moves_tq<NUP, double, 1>(S, xin, xout, irn_in, irn_out, sinv, m0, m1, T, grid, st)
               : sm_moves_tq<NUP, double, 4>(S, xin, xout, irn_in, irn_out, sinv, m0, m1, T, grid, st);
}

template <int NUP>
int sm_refresh_t(const DevSys &S, const double *x, char *sinv, const DevStep &T, int grid,
                 hipStream_t st) {
    hipLaunchKernelGGL(k_sm_refresh<NUP>, dim3(grid), dim3(64), lds_bytes<NUP>(S.nq, true), st, S, x,
                       sinv, T);
    return hipGetLastError();
}

template <int NUP>
int sm_end_t(const DevSys &S, const DevState &A, const DevState &B, const DevStep &T,
             int opcomp, int grid, hipStream_t st) {
    if (S.tab_lds) {
        const int blocks = (grid + kTabWaves - 1) / kTabWaves;
        hipLaunchKernelGGL(k_sm_end_tab<NUP>, dim3(blocks), dim3(64 * kTabWaves),
                           tab_lds_bytes<NUP>(S), st, S, A, B, T, opcomp);
    } else {
        hipLaunchKernelGGL(k_sm_end<NUP>, dim3(grid), dim3(64), lds_bytes<NUP>(S.nq, true), st,
                           S, A, B, T, opcomp);
    }
    return hipGetLastError();
}

// does a workgroup of the table-staging end kernel fit the device (its LDS)?
template <int NUP>
int sm_end_tab_fits_t(const DevSys &S) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(k_sm_end_tab<NUP>), 64 * kTabWaves,
            tab_lds_bytes<NUP>(S)) != hipSuccess)
        return 0;
    return nb > 0;
}

int launch_sm_moves(int nup, const DevSys &S, const double *xin, double *xout,
                    const uint64_t *irn_in, uint64_t *irn_out, char *sinv, int m0, int m1,
                    const DevStep &T, int grid, hipStream_t st) {
    if (grid <= 0) return hipSuccess;
    switch (nup) {
        case 1: return sm_moves_t<1>(S, xin, xout, irn_in, irn_out, sinv, m0, m1, T, grid, st);
        case 7: return sm_moves_t<7>(S, xin, xout, irn_in, irn_out, sinv, m0, m1, T, grid, st);
        case 19: return sm_moves_t<19>(S, xin, xout, irn_in, irn_out, sinv, m0, m1, T, grid, st);
        case 33: return sm_moves_t<33>(S, xin, xout, irn_in, irn_out, sinv, m0, m1, T, grid, st);
    }
    return hipErrorInvalidValue;
}

template <int NUP>
int sm_moves_tab_fits_t(const DevSys &S) {
    if constexpr (NUP >= 17) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, reinterpret_cast<const void *>(k_sm_moves32x2_tab<NUP>), 64 * kMvWaves,
                mv_pair_tab_lds_bytes<NUP>(S.ntab)) != hipSuccess)
            return 0;
        return nb > 0;
    }
    return 0;
}

int sm_moves_tab_fits(int nup, const DevSys &S) {
    switch (nup) {
        case 19: return sm_moves_tab_fits_t<19>(S);
        case 33: return sm_moves_tab_fits_t<33>(S);
    }
    return 0;
}

int sm_end_tab_fits(int nup, const DevSys &S) {
    switch (nup) {
        case 1: return sm_end_tab_fits_t<1>(S);
        case 7: return sm_end_tab_fits_t<7>(S);
        case 19: return sm_end_tab_fits_t<19>(S);
        case 33: return sm_end_tab_fits_t<33>(S);
    }
    return 0;
}

int launch_sm_refresh(int nup, const DevSys &S, const double *x, char *sinv, const DevStep &T,
                      int grid, hipStream_t st) {
    if (grid <= 0) return hipSuccess;
    switch (nup) {
        case 1: return sm_refresh_t<1>(S, x, sinv, T, grid, st);
        case 7: return sm_refresh_t<7>(S, x, sinv, T, grid, st);
        case 19: return sm_refresh_t<19>(S, x, sinv, T, grid, st);
        case 33: return sm_refresh_t<33>(S, x, sinv, T, grid, st);
    }
    return hipErrorInvalidValue;
}

int launch_sm_end(int nup, const DevSys &S, const DevState &A, const DevState &B,
                  const DevStep &T, int opcomp, int grid, hipStream_t st) {
    if (grid <= 0) return hipSuccess;
    switch (nup) {
        case 1: return sm_end_t<1>(S, A, B, T, opcomp, grid, st);
        case 7: return sm_end_t<7>(S, A, B, T, opcomp, grid, st);
        case 19: return sm_end_t<19>(S, A, B, T, opcomp, grid, st);
        case 33: return sm_end_t<33>(S, A, B, T, opcomp, grid, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace dmc
