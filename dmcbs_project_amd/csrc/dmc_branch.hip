// Branching as a device prefix-sum compaction (branch, step.f90:287-315), and the
// fixed-order reduction of the per-walker estimator contributions.
//
// Each walker w gets a multiplicity iwt = int(weight * factor + u) with u = ran1 of its own
// RNG state; copy c (c = 1..iwt) carries that state advanced c times by ran2.  The walkers'
// copies are laid out by an exclusive scan of the multiplicities, so the new population is
// ordered by source walker -- a deterministic order; the reference's LIFO stack order is
// different, but no walker's evolution depends on order.
#include "dmc_wave.h"

namespace dmc {

constexpr int kScanThreads = 1024;
constexpr int kEstBlocks = 64;  // blocks of the estimator sums

// Device population words of DevStep::nt: [0] walkers a kernel processes, [1] total after a
// scan, [2] resident population (device-driven steps), [3] overflow flag, [4] population
// control active.  A non-null gate makes a kernel return at once unless *gate is set, so
// that a captured graph can hold the population-control pass unconditionally.
__device__ __forceinline__ bool gated_off(const int *gate) { return gate && !*gate; }

// multiplicities of state A with a factor (population control, dmcbcs.f90:445)
__global__ void k_branch_count(DevState A, DevStep T, double factor, const int *gate, int nwalk) {
    if (gated_off(gate)) return;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= T.nt[0]) return;
    // gated (device-driven) pass: the factor nwalk/nt of dmcbcs.f90:445 from the device count
    if (gate) factor = 1.0 * nwalk / T.nt[2];
    const double wt = A.weight[w] * factor;
    uint64_t s = A.irn[w];
    s = (s * kM1 + kA1) & kMask48;
    const double u = double(s | 1ull) * 0x1p-48;
    const int iwt = int(wt + u);
    T.count[w] = iwt;
    A.irn[w] = s;
    if (iwt >= 0 && iwt < 500) atomicAdd(reinterpret_cast<unsigned long long *>(&T.hist[iwt]), 1ull);
}

// exclusive scan of count[0..nt[0]) into offset, total into nt[1]; one block.  Thread t
// owns a contiguous chunk, read and written in 16-byte vectors where it is aligned; the
// chunk sums are scanned by wave shuffles and the 16 wave totals in LDS.
__global__ __launch_bounds__(kScanThreads) void k_scan(DevStep T, const int *gate) {
    if (gated_off(gate)) return;
    __shared__ int wtot[kScanThreads / 64];
    const int n = T.nt[0];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int chunk = ((n + kScanThreads - 1) / kScanThreads + 3) & ~3;  // multiple of 4
    const int lo = min(n, tid * chunk), hi = min(n, lo + chunk);
    int s = 0;
    {
        int i = lo;
        for (; i + 4 <= hi; i += 4) {
            const int4 c = *reinterpret_cast<const int4 *>(T.count + i);
            s += c.x + c.y + c.z + c.w;
        }
        for (; i < hi; ++i) s += T.count[i];
    }
    // inclusive scan of the chunk sums over the wave, then over the waves
    int incl = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wtot[wv] = incl;
    __syncthreads();
    if (wv == 0) {
        int t = lane < kScanThreads / 64 ? wtot[lane] : 0;
#pragma unroll
        for (int off = 1; off < kScanThreads / 64; off <<= 1) {
            const int y = __shfl_up(t, off, 64);
            if (lane >= off) t += y;
        }
        if (lane < kScanThreads / 64) wtot[lane] = t;
    }
    __syncthreads();
    int run = incl - s + (wv ? wtot[wv - 1] : 0);
    {
        int i = lo;
        for (; i + 4 <= hi; i += 4) {
            const int4 c = *reinterpret_cast<const int4 *>(T.count + i);
            int4 o;
            o.x = run;
            o.y = o.x + c.x;
            o.z = o.y + c.y;
            o.w = o.z + c.z;
            run = o.w + c.w;
            *reinterpret_cast<int4 *>(T.offset + i) = o;
        }
        for (; i < hi; ++i) {
            T.offset[i] = run;
            run += T.count[i];
        }
    }
    if (tid == kScanThreads - 1) T.nt[1] = wtot[kScanThreads / 64 - 1];
}

// destination map and the RNG state of every copy
__global__ void k_branch_map(DevState src, DevStep T, int capacity, const int *gate) {
    if (gated_off(gate)) return;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= T.nt[0]) return;
    const int off = T.offset[w], cnt = T.count[w];
    uint64_t s = src.irn[w];
    for (int c = 0; c < cnt; ++c) {
        const int d = off + c;
        s = (s * kM2 + kA2) & kMask48;  // ran2 before each push (step.f90:306-308)
        if (d >= capacity) break;
        T.map[d] = w;
        T.copy_irn[d] = s;
    }
}

// copy of one walker's slot of nbytes (a multiple of 16, both ends 16-byte aligned: positions
// are 24 N bytes with N even, inverse slots whole 128-byte lines) by the 64 lanes of a wave,
// in 16-byte vectors with several loads in flight per lane
__device__ __forceinline__ void copy_slot(const void *src, void *dst, size_t nbytes, int tid) {
    const uint4 *a = static_cast<const uint4 *>(src);
    uint4 *b = static_cast<uint4 *>(dst);
    const int n = int(nbytes / 16);
#pragma unroll 4
    for (int t = tid; t < n; t += 64) b[t] = a[t];
}

// copy every destination walker from its source; weight becomes sign(1, weight)
__global__ __launch_bounds__(64) void k_gather(DevSys S, DevState src, DevState dst, DevStep T,
                                               int m3, const int *gate, int capacity) {
    if (gated_off(gate)) return;
    const int d = blockIdx.x;
    if (d >= T.nt[1] || d >= capacity) return;
    const int w = T.map[d];
    copy_slot(src.x + size_t(w) * m3, dst.x + size_t(d) * m3, size_t(m3) * 8, threadIdx.x);
    copy_slot(src.dpsi + size_t(w) * m3, dst.dpsi + size_t(d) * m3, size_t(m3) * 8, threadIdx.x);
    if (S.keep_sinv) {
        const size_t sb = sinv_stride(S);
        copy_slot(src.sinv + size_t(w) * sb, dst.sinv + size_t(d) * sb, sb, threadIdx.x);
    }
    if (threadIdx.x == 0) {
        dst.psil[d] = src.psil[w];
        dst.d2psi[d] = src.d2psi[w];
        dst.v[d] = src.v[w];
        dst.is[d] = src.is[w];
        dst.wt0[d] = src.wt0[w];
        dst.weight[d] = copysign(1.0, src.weight[w]);
        dst.irn[d] = T.copy_irn[d];
    }
}

// Estimator sums over walkers 0..nw-1 (nw < 0: nt[0]) added to acc[0..41]: block b sums a
// contiguous range of walkers, reading rows of 42 values with consecutive threads, and the
// last block to finish adds the block partials in block order.  Both orders are fixed, so
// the sums do not change from run to run.  nt[5] counts finished blocks (reset by the last).
__global__ __launch_bounds__(256) void k_est_reduce(DevStep T, int nw, double *acc) {
    __shared__ double part[6][42];
    __shared__ int last;
    if (nw < 0) nw = T.nt[0];
    const int tid = threadIdx.x, b = blockIdx.x;
    const int chunk = (nw + kEstBlocks - 1) / kEstBlocks;
    const int w0 = min(nw, b * chunk), w1 = min(nw, w0 + chunk);
    const int g = tid / 42, k = tid - 42 * g;
    if (g < 6) {
        double s = 0.0;
        // unrolled: several loads in flight, the sum still taken in walker order
#pragma unroll 8
        for (int w = w0 + g; w < w1; w += 6) s += T.est[size_t(w) * 42 + k];
        part[g][k] = s;
    }
    __syncthreads();
    if (tid < 42) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) v += part[j][tid];
        T.estp[b * 42 + tid] = v;
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) last = atomicAdd(T.nt + 5, 1) == kEstBlocks - 1;
    __syncthreads();
    if (last) {
        __threadfence();
        if (tid < 42) {
            double v = 0.0;
            for (int j = 0; j < kEstBlocks; ++j) v += T.estp[j * 42 + tid];
            acc[tid] += v;
        }
        if (tid == 0) T.nt[5] = 0;
    }
}

// calder (optimizer.f90:90-112) over walkers 0..nw-1: acc[k] += sum_w term_k(w) in a fixed
// order, one block per term: k = 0 count, 1 E, 2 E^2, 3.. O_p, 3+P.. O_p E, 3+2P.. O_p O_q,
// with E = E_L / (N eunit) of the stored state and O = d log psi / d theta [nw][P]
__global__ __launch_bounds__(256) void k_opt_accumulate(DevSys S, DevState A, const double *O,
                                                        int nw, int P, double *acc) {
    __shared__ double part[256];
    const int k = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int w = tid; w < nw; w += 256) {
        const double e = (-S.hbar * A.d2psi[w] + A.v[w]) * S.nfac;
        const double *o = O + size_t(w) * P;
        double t;
        if (k == 0) t = 1.0;
        else if (k == 1) t = e;
        else if (k == 2) t = e * e;
        else if (k < 3 + P) t = o[k - 3];
        else if (k < 3 + 2 * P) t = o[k - 3 - P] * e;
        else {
            const int pq = k - 3 - 2 * P;
            t = o[pq / P] * o[pq % P];
        }
        s += t;
    }
    part[tid] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) part[tid] += part[tid + off];
        __syncthreads();
    }
    if (tid == 0) acc[k] += part[0];
}

int launch_opt_accumulate(const DevSys &S, const DevState &A, const double *O, int nw, int P,
                          double *acc, hipStream_t st) {
    if (nw <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_opt_accumulate, dim3(3 + 2 * P + P * P), dim3(256), 0, st, S, A, O, nw,
                       P, acc);
    return hipGetLastError();
}

// device-driven population bookkeeping after a branch pass: resident population
// nt[2] = min(total, capacity), overflow flag nt[3]; then (decide != 0) the population-
// control test of dmcbcs.f90:444-445, setting nt[4] and, when it fires, nt[0] = nt[2]
__global__ void k_pop_update(DevStep T, int capacity, double lo, double hi, int decide,
                             const int *gate) {
    if (gated_off(gate)) return;
    const int total = T.nt[1];
    const int n = total < capacity ? total : capacity;
    T.nt[2] = n;
    if (total > capacity) T.nt[3] = 1;
    if (decide) {
        const int on = (n > hi || n < lo) && n > 0;
        T.nt[4] = on;
        if (on) T.nt[0] = n;
    }
}

// start of a device-driven step: the kernels process the resident population
__global__ void k_pop_begin(DevStep T) {
    T.nt[0] = T.nt[2];
}

// copy walkers 0..nt[1]-1 of src to dst (the population-control result back to A)
__global__ __launch_bounds__(64) void k_copy_back(DevSys S, DevState src, DevState dst,
                                                  DevStep T, int m3, const int *gate,
                                                  int capacity) {
    if (gated_off(gate)) return;
    const int d = blockIdx.x;
    if (d >= T.nt[1] || d >= capacity) return;
    copy_slot(src.x + size_t(d) * m3, dst.x + size_t(d) * m3, size_t(m3) * 8, threadIdx.x);
    copy_slot(src.dpsi + size_t(d) * m3, dst.dpsi + size_t(d) * m3, size_t(m3) * 8, threadIdx.x);
    if (S.keep_sinv) {
        const size_t sb = sinv_stride(S);
        copy_slot(src.sinv + size_t(d) * sb, dst.sinv + size_t(d) * sb, sb, threadIdx.x);
    }
    if (threadIdx.x == 0) {
        dst.psil[d] = src.psil[d];
        dst.d2psi[d] = src.d2psi[d];
        dst.v[d] = src.v[d];
        dst.is[d] = src.is[d];
        dst.wt0[d] = src.wt0[d];
        dst.weight[d] = src.weight[d];
        dst.irn[d] = src.irn[d];
    }
}

int launch_branch_count(const DevSys &, const DevState &A, const DevStep &T, double factor,
                        int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_branch_count, dim3((grid + 255) / 256), dim3(256), 0, st, A, T, factor,
                       nullptr, 0);
    return hipGetLastError();
}

// scan, then (after the caller has read nt[1]) map + gather: split in two launchers
int launch_scan(const DevStep &T, hipStream_t st) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, st, T, nullptr);
    return hipGetLastError();
}

int launch_map_gather(const DevSys &S, const DevState &src, const DevState &dst,
                      const DevStep &T, int nsrc, int ndst, int capacity, hipStream_t st) {
    hipLaunchKernelGGL(k_branch_map, dim3((nsrc + 255) / 256), dim3(256), 0, st, src, T, capacity,
                       nullptr);
    int e = hipGetLastError();
    if (e) return e;
    if (ndst > 0)
        hipLaunchKernelGGL(k_gather, dim3(ndst), dim3(64), 0, st, S, src, dst, T, 3 * S.npart,
                           nullptr, capacity);
    return hipGetLastError();
}

// The branch and population control of one system with every size on the device (no host
// synchronisation; capturable in a graph).  The population enters in B (after the step
// kernels) with multiplicities in T.count and leaves in A; population control, when it
// fires, runs A -> B and copies back.
int launch_branch_device(const DevSys &S, const DevState &A, const DevState &B, const DevStep &T,
                         int capacity, hipStream_t st) {
    const int m3 = 3 * S.npart, g256 = (capacity + 255) / 256;
    const int *gate = T.nt + 4;
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, st, T, nullptr);
    hipLaunchKernelGGL(k_branch_map, dim3(g256), dim3(256), 0, st, B, T, capacity, nullptr);
    hipLaunchKernelGGL(k_gather, dim3(capacity), dim3(64), 0, st, S, B, A, T, m3, nullptr,
                       capacity);
    hipLaunchKernelGGL(k_pop_update, dim3(1), dim3(1), 0, st, T, capacity,
                       S.popdn * S.nwalk, S.popup * S.nwalk, 1, nullptr);
    hipLaunchKernelGGL(k_branch_count, dim3(g256), dim3(256), 0, st, A, T, 1.0, gate, S.nwalk);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, st, T, gate);
    hipLaunchKernelGGL(k_branch_map, dim3(g256), dim3(256), 0, st, A, T, capacity, gate);
    hipLaunchKernelGGL(k_gather, dim3(capacity), dim3(64), 0, st, S, A, B, T, m3, gate,
                       capacity);
    hipLaunchKernelGGL(k_copy_back, dim3(capacity), dim3(64), 0, st, S, B, A, T, m3, gate,
                       capacity);
    hipLaunchKernelGGL(k_pop_update, dim3(1), dim3(1), 0, st, T, capacity, 0.0, 0.0, 0, gate);
    return hipGetLastError();
}

int launch_pop_begin(const DevStep &T, hipStream_t st) {
    hipLaunchKernelGGL(k_pop_begin, dim3(1), dim3(1), 0, st, T);
    return hipGetLastError();
}

int launch_est_reduce(const DevStep &T, int grid, double *acc, hipStream_t st) {
    hipLaunchKernelGGL(k_est_reduce, dim3(kEstBlocks), dim3(256), 0, st, T, grid, acc);
    return hipGetLastError();
}

}  // namespace dmc
