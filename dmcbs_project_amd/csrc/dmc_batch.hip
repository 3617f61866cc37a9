// Several resident systems (different particle numbers, each with its own tables and
// population) stepped together: the move, refresh and end work of all systems runs in one
// persistent launch per phase.  A workgroup holds wpg walkers, one per 64-lane wave; every
// wave takes the next walker from a device counter until all are done, walkers of the
// largest system first (largest-first order keeps the tail short).  Sizes come from the
// device population words, so that the whole step can be captured in a graph.
#include "dmc_batch.h"
#include "dmc_sm2.h"

namespace dmc {

__device__ __forceinline__ int next_item(int *counter, int tid) {
    int item = 0;
    if (tid == 0) item = atomicAdd(counter, 1);
    return __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
}

// The persistent loop of one phase: the systems in order (largest first), and within a
// system the walkers taken one per wave from the system's own counter.  A wave moves on to
// the next system when the current one has no walker left, so that all systems share one
// launch.  Each particle number has its own loop, so that the compiler allocates the
// registers of one particle number at a time.
#define DMC_BATCH_LOOP(CALL)                                                                \
    for (int s = 0; s < a.nsys; ++s) {                                                      \
        const BatchSys &b = a.sys[s];                                                       \
        if (chunk >= 0 && chunk * b.k >= b.npart) continue;                                 \
        const int n = b.T.nt[0];                                                            \
        int *ctr = counters + s;                                                            \
        switch (b.nup) {                                                                    \
            case 1: for (int w; (w = next_item(ctr, tid)) < n;) { CALL(1); wave_sync(); } break;   \
            case 7: for (int w; (w = next_item(ctr, tid)) < n;) { CALL(7); wave_sync(); } break;   \
            case 19: for (int w; (w = next_item(ctr, tid)) < n;) { CALL(19); wave_sync(); } break; \
            case 33: for (int w; (w = next_item(ctr, tid)) < n;) { CALL(33); wave_sync(); } break; \
        }                                                                                   \
    }

// moves of chunk c (moves c*k .. min((c+1)*k, npart) - 1 of every walker).  FP32 systems
// that the two-walker moves take (sm_moves_pair) are dequeued by pairs of walkers, so that
// every walker runs the code it runs in a per-system launch.
template <typename R, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_batch_moves(BatchArgs a, int chunk, int *counters,
                                                          int stride) {
    extern __shared__ double smem[];
    const int tid = threadIdx.x & 63;
    double *sm = smem + size_t(threadIdx.x >> 6) * stride;
    if constexpr (sizeof(R) == 4) {
        for (int s = 0; s < a.nsys; ++s) {
            const BatchSys &b = a.sys[s];
            if (chunk * b.k >= b.npart || !mv_pair_fits(b.nup, b.S.nq)) continue;
            const int n = b.T.nt[0];
            int *ctr = counters + s;
            const int m0 = chunk * b.k, m1 = min(m0 + b.k, b.npart);
#define DMC_PAIR(NUP)                                                                        \
    for (int p; 2 * (p = next_item(ctr, tid)) < n;) {                                        \
        sm_moves_pair<NUP>(b.S, sm, p, n, tid, m0 ? b.B.x : b.A.x, b.B.x,                    \
                           m0 ? b.B.irn : b.A.irn, b.B.irn, b.A.sinv, m0, m1, b.T);          \
        wave_sync();                                                                         \
    }
            switch (b.nup) {
                case 19: DMC_PAIR(19) break;
                case 33: DMC_PAIR(33) break;
            }
#undef DMC_PAIR
        }
    }
#define DMC_MOVES(NUP)                                                                      \
    {                                                                                       \
        const int m0 = chunk * b.k, m1 = min(m0 + b.k, b.npart);                            \
        if (sizeof(R) == 4 && mv_pair_fits(b.nup, b.S.nq)) break;                           \
        sm_moves_walker_any<NUP, R>(b.S, sm, w, tid, m0 ? b.B.x : b.A.x, b.B.x,             \
                                m0 ? b.B.irn : b.A.irn, b.B.irn, b.A.sinv, m0, m1, b.T);    \
    }
    DMC_BATCH_LOOP(DMC_MOVES)
#undef DMC_MOVES
}

// FP64 refresh of the working inverse from the positions after chunk c - 1
template <int WPG>
__global__ __launch_bounds__(64 * WPG) void k_batch_refresh(BatchArgs a, int chunk,
                                                            int *counters, int stride) {
    extern __shared__ double smem[];
    const int tid = threadIdx.x & 63;
    double *sm = smem + size_t(threadIdx.x >> 6) * stride;
#define DMC_REFRESH(NUP) sm_refresh_walker<NUP>(b.S, sm, w, tid, b.B.x, b.A.sinv)
    DMC_BATCH_LOOP(DMC_REFRESH)
#undef DMC_REFRESH
}

// end of the sweep of every walker of every system
template <int WPG>
__global__ __launch_bounds__(64 * WPG) void k_batch_end(BatchArgs a, int opcomp, int *counters,
                                                        int stride) {
    extern __shared__ double smem[];
    const int tid = threadIdx.x & 63;
    double *sm = smem + size_t(threadIdx.x >> 6) * stride;
    const int chunk = -1;
#define DMC_END(NUP) sm_end_walker<NUP>(b.S, sm, w, tid, b.A, b.B, b.T, opcomp)
    DMC_BATCH_LOOP(DMC_END)
#undef DMC_END
}
#undef DMC_BATCH_LOOP

// start of a step: the work counters of every phase are cleared and every system processes
// its resident population
__global__ void k_batch_begin(BatchArgs a, int *counters, int ncounters) {
    const int t = threadIdx.x;
    for (int i = t; i < ncounters; i += blockDim.x) counters[i] = 0;
    if (t < a.nsys) a.sys[t].T.nt[0] = a.sys[t].T.nt[2];
}

// LDS of one wave (doubles) for a phase: the largest over the systems
size_t batch_wave_lds(const BatchHost &h, BatchPhase ph) {
    size_t m = 0;
    for (const BatchSys &b : h.sys) {
        size_t v = 0;
        const int nq = b.S.nq;
        if (ph == BatchPhase::Moves) {
            if (h.fp32) {
                switch (b.nup) {
                    case 1: v = mv_lds_bytes<1, float>(nq); break;
                    case 7: v = mv_lds_bytes<7, float>(nq); break;
                    case 19: v = mv_pair_fits(19, nq) ? mv_pair_lds_bytes<19>() : mv_lds_bytes<19, float>(nq); break;
                    case 33: v = mv_pair_fits(33, nq) ? mv_pair_lds_bytes<33>() : mv_lds_bytes<33, float>(nq); break;
                }
            } else {
                switch (b.nup) {
                    case 1: v = mv_lds_bytes<1, double>(nq); break;
                    case 7: v = mv_lds_bytes<7, double>(nq); break;
                    case 19: v = mv_lds_bytes<19, double>(nq); break;
                    case 33: v = mv_lds_bytes<33, double>(nq); break;
                }
            }
        } else {
            switch (b.nup) {
                case 1: v = lds_bytes<1>(nq, true); break;
                case 7: v = lds_bytes<7>(nq, true); break;
                case 19: v = lds_bytes<19>(nq, true); break;
                case 33: v = lds_bytes<33>(nq, true); break;
            }
        }
        m = v > m ? v : m;
    }
    return (m + 255) / 256 * 256;  // keep every wave's area 256-byte aligned
}

template <int WPG>
static const void *phase_kernel_w(bool fp32, BatchPhase ph) {
    switch (ph) {
        case BatchPhase::Moves:
            return fp32 ? reinterpret_cast<const void *>(k_batch_moves<float, WPG>)
                        : reinterpret_cast<const void *>(k_batch_moves<double, WPG>);
        case BatchPhase::Refresh: return reinterpret_cast<const void *>(k_batch_refresh<WPG>);
        default: return reinterpret_cast<const void *>(k_batch_end<WPG>);
    }
}

static const void *phase_kernel(bool fp32, BatchPhase ph, int wpg) {
    switch (wpg) {
        case 1: return phase_kernel_w<1>(fp32, ph);
        case 2: return phase_kernel_w<2>(fp32, ph);
        case 4: return phase_kernel_w<4>(fp32, ph);
        case 8: return phase_kernel_w<8>(fp32, ph);
    }
    return nullptr;
}

// resident workgroups of wpg waves per CU for a phase (0: the configuration does not fit)
int batch_blocks_per_cu(const BatchHost &h, BatchPhase ph, int wpg) {
    int nb = 0;
    const size_t lds = batch_wave_lds(h, ph) * wpg;
    const void *k = phase_kernel(h.fp32, ph, wpg);
    if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, 64 * wpg, lds) != hipSuccess)
        return 0;
    return nb;
}

int launch_batch_begin(const BatchArgs &a, int *counters, int ncounters, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, st, a, counters, ncounters);
    return hipGetLastError();
}

template <int WPG>
static void batch_launch_w(const BatchHost &h, const BatchArgs &a, BatchPhase ph, int chunk,
                           int opcomp, int *counter, dim3 grid, size_t lds, int stride,
                           hipStream_t st) {
    const dim3 block(64 * WPG);
    switch (ph) {
        case BatchPhase::Moves:
            if (h.fp32)
                hipLaunchKernelGGL((k_batch_moves<float, WPG>), grid, block, lds, st, a, chunk,
                                   counter, stride);
            else
                hipLaunchKernelGGL((k_batch_moves<double, WPG>), grid, block, lds, st, a, chunk,
                                   counter, stride);
            break;
        case BatchPhase::Refresh:
            hipLaunchKernelGGL(k_batch_refresh<WPG>, grid, block, lds, st, a, chunk, counter,
                               stride);
            break;
        case BatchPhase::End:
            hipLaunchKernelGGL(k_batch_end<WPG>, grid, block, lds, st, a, opcomp, counter, stride);
            break;
    }
}

int launch_batch_phase(const BatchHost &h, const BatchArgs &a, BatchPhase ph, int chunk,
                       int opcomp, int *counter, int wpg, int ncu, hipStream_t st) {
    const size_t wl = batch_wave_lds(h, ph);
    const int nb = batch_blocks_per_cu(h, ph, wpg);
    if (nb <= 0) return hipErrorInvalidConfiguration;
    const dim3 grid(nb * ncu);
    const size_t lds = wl * wpg;
    const int stride = int(wl / sizeof(double));
    switch (wpg) {
        case 1: batch_launch_w<1>(h, a, ph, chunk, opcomp, counter, grid, lds, stride, st); break;
        case 2: batch_launch_w<2>(h, a, ph, chunk, opcomp, counter, grid, lds, stride, st); break;
        case 4: batch_launch_w<4>(h, a, ph, chunk, opcomp, counter, grid, lds, stride, st); break;
        case 8: batch_launch_w<8>(h, a, ph, chunk, opcomp, counter, grid, lds, stride, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dmc
