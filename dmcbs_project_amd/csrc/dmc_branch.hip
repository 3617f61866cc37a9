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

// multiplicities of state A with a factor (population control, dmcbcs.f90:445)
__global__ void k_branch_count(DevState A, DevStep T, double factor) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= T.nt[0]) return;
    const double wt = A.weight[w] * factor;
    uint64_t s = A.irn[w];
    s = (s * kM1 + kA1) & kMask48;
    const double u = double(s | 1ull) * 0x1p-48;
    const int iwt = int(wt + u);
    T.count[w] = iwt;
    A.irn[w] = s;
    if (iwt >= 0 && iwt < 500) atomicAdd(reinterpret_cast<unsigned long long *>(&T.hist[iwt]), 1ull);
}

// exclusive scan of count[0..nt[0]) into offset, total into nt[1]; one block
__global__ __launch_bounds__(kScanThreads) void k_scan(DevStep T) {
    __shared__ int part[kScanThreads];
    const int n = T.nt[0];
    const int tid = threadIdx.x;
    const int chunk = (n + kScanThreads - 1) / kScanThreads;
    const int lo = tid * chunk, hi = min(n, lo + chunk);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += T.count[i];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {  // Hillis-Steele inclusive scan
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = tid ? part[tid - 1] : 0;
    for (int i = lo; i < hi; ++i) {
        T.offset[i] = run;
        run += T.count[i];
    }
    if (tid == kScanThreads - 1) T.nt[1] = part[kScanThreads - 1];
}

// destination map and the RNG state of every copy
__global__ void k_branch_map(DevState src, DevStep T, int capacity) {
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

// copy every destination walker from its source; weight becomes sign(1, weight)
__global__ __launch_bounds__(64) void k_gather(DevSys S, DevState src, DevState dst, DevStep T,
                                               int m3) {
    const int d = blockIdx.x;
    if (d >= T.nt[1]) return;
    const int w = T.map[d];
    const double *xs = src.x + size_t(w) * m3, *gs = src.dpsi + size_t(w) * m3;
    double *xd = dst.x + size_t(d) * m3, *gd = dst.dpsi + size_t(d) * m3;
    for (int t = threadIdx.x; t < m3; t += 64) {
        xd[t] = xs[t];
        gd[t] = gs[t];
    }
    if (S.keep_sinv) {
        const size_t nb = sinv_bytes(S) / 4;
        const uint32_t *a = reinterpret_cast<const uint32_t *>(src.sinv + size_t(w) * nb * 4);
        uint32_t *b = reinterpret_cast<uint32_t *>(dst.sinv + size_t(d) * nb * 4);
        for (size_t t = threadIdx.x; t < nb; t += 64) b[t] = a[t];
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

// est_acc[k] += sum_w est[w][k], k < 42, in a fixed order (one block per component)
__global__ __launch_bounds__(256) void k_est_reduce(DevStep T, int nw, double *acc) {
    __shared__ double part[256];
    const int k = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int w = tid; w < nw; w += 256) s += T.est[size_t(w) * 42 + k];
    part[tid] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) part[tid] += part[tid + off];
        __syncthreads();
    }
    if (tid == 0) acc[k] += part[0];
}

int launch_branch_count(const DevSys &, const DevState &A, const DevStep &T, double factor,
                        int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_branch_count, dim3((grid + 255) / 256), dim3(256), 0, st, A, T, factor);
    return hipGetLastError();
}

// scan, then (after the caller has read nt[1]) map + gather: split in two launchers
int launch_scan(const DevStep &T, hipStream_t st) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, st, T);
    return hipGetLastError();
}

int launch_map_gather(const DevSys &S, const DevState &src, const DevState &dst,
                      const DevStep &T, int nsrc, int ndst, int capacity, hipStream_t st) {
    hipLaunchKernelGGL(k_branch_map, dim3((nsrc + 255) / 256), dim3(256), 0, st, src, T, capacity);
    int e = hipGetLastError();
    if (e) return e;
    if (ndst > 0) hipLaunchKernelGGL(k_gather, dim3(ndst), dim3(64), 0, st, S, src, dst, T, 3 * S.npart);
    return hipGetLastError();
}

int launch_est_reduce(const DevStep &T, int grid, double *acc, hipStream_t st) {
    hipLaunchKernelGGL(k_est_reduce, dim3(42), dim3(256), 0, st, T, grid, acc);
    return hipGetLastError();
}

}  // namespace dmc
