// The C-ABI of include/dmc_gpu.h: context, device memory, uploads and the per-step
// orchestration (fused step kernel -> estimator reduction -> branch compaction ->
// population control).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "dmc_batch.h"
#include "dmc_internal.h"
#include "dmc_wave.h"

namespace dmc {

__global__ void k_init_weights(DevState A, int nw) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    A.wt0[w] = double(A.is[w]);
    A.weight[w] = 1.0;
}

}  // namespace dmc

using namespace dmc;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(DMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCHK(expr)                                                                          \
    do {                                                                                         \
        int e_ = (expr);                                                                         \
        if (e_ != 0)                                                                             \
            return fail(DMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(hipError_t(e_))); \
    } while (0)

template <class T>
hipError_t dalloc(T **p, size_t n) {
    return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (n ? n : 1));
}

// Device buffer of dmc_opt_accumulate_noalloc: the parameter derivatives O [nw][P] of the
// adjoint kernel, in a file-scope array instead of the context's scratch (8 MiB per device,
// held until the process ends; every call on a device shares it).  Host code must not pass
// the array's name to a kernel: it uses the device address that hipGetSymbolAddress gives
// for the current device, resolved once per device by dmc_create.
constexpr size_t kNoAllocDoubles = size_t(1) << 20;
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[kMaxDevices] = {};
std::string g_noalloc_err[kMaxDevices];

// the address of the array on device dev (current), resolved on the first call per device
double *noalloc_scratch_address(int dev);
}  // namespace

static __device__ __align__(256) double noalloc_scratch[kNoAllocDoubles];

namespace {
double *noalloc_scratch_address(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev] && g_noalloc_err[dev].empty()) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
        if (e != hipSuccess || !p)
            g_noalloc_err[dev] = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
        else
            g_noalloc_ptr[dev] = static_cast<double *>(p);
    }
    return g_noalloc_ptr[dev];
}
}  // namespace

struct dmc_ctx {
    System sys;
    DevSys ds{};
    hipStream_t st = nullptr;
    int capacity = 0, nw = 0;
    std::vector<void *> allocs;
    DevState A{}, B{};
    DevStep T{};
    double *est_acc = nullptr;
    double *opt_acc = nullptr;  // optimiser sums, kOptSums
    static constexpr int kOptSums = 3 + 2 * (DMC_MAXORB + 2) + (DMC_MAXORB + 2) * (DMC_MAXORB + 2);
    int *h_nt = nullptr;  // pinned
    bool timing = false;
    double time_ms = 0.0;
    int64_t launches = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    static constexpr int kEv = 16;  // event pairs around the move-kernel chunks of one step
    hipEvent_t evs[2 * kEv] = {};
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    int64_t acc_moves = 0, prop_moves = 0;
    // q vectors of S(q) for the last shell count asked for (sofq_ntabq = 0: none yet)
    int sofq_ntabq = 0;
    std::vector<std::array<int, 3>> sofq_ik;
    std::vector<int> sofq_numsh;
    bool pending = false;  // a propagation whose branch (factor 1) has not run yet
    int nev_pending = 0;   // timed move-kernel chunks of that propagation
    int device = 0;        // the device current at dmc_create
    double *trace = nullptr;  // move records (dmc_move_trace), [capacity][npart][kTraceRec]
    int trace_nw = 0;         // walkers of the last recorded propagation
    double *noalloc = nullptr;  // that device's address of noalloc_scratch

    template <class T>
    int alloc(T **p, size_t n) {
        hipError_t e = dalloc(p, n);
        if (e != hipSuccess) return fail(DMC_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(*p);
        return 0;
    }
    int alloc_state(DevState &S) {
        const size_t m = size_t(capacity) * 3 * sys.npart;
        const size_t sb = ds.keep_sinv ? size_t(capacity) * sinv_stride(ds) : 1;
        if (alloc(&S.x, m) || alloc(&S.dpsi, m) || alloc(&S.psil, capacity) ||
            alloc(&S.d2psi, capacity) || alloc(&S.v, capacity) || alloc(&S.weight, capacity) ||
            alloc(&S.wt0, capacity) || alloc(&S.is, capacity) || alloc(&S.irn, capacity) ||
            alloc(&S.sinv, sb))
            return DMC_ERR_HIP;
        return 0;
    }
    int ensure_scratch(size_t bytes) {
        if (bytes <= scratch_bytes) return 0;
        if (scratch) (void)hipFree(scratch);
        scratch = nullptr;
        scratch_bytes = 0;
        hipError_t e = hipMalloc(&scratch, bytes);
        if (e != hipSuccess) return fail(DMC_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        scratch_bytes = bytes;
        return 0;
    }
};

namespace {

int upload_tables(dmc_ctx *c) {
    const System &s = c->sys;
    const int n1 = s.ntab + 1;
    std::vector<double4> jt(n1), bt(n1), btb(n1), btd(n1);
    std::vector<double> bt1(n1), vt1(n1);
    std::vector<double2> bt2(n1);
    int last = -1;  // last node with a nonzero u, u' or u''
    std::vector<double4> mt(n1);
    std::vector<float4> mt32(n1);
    for (int i = 0; i < n1; ++i) {
        jt[i] = make_double4(s.vt[i], s.ut[i], s.dut[i], s.d2ut[i]);
        vt1[i] = s.vt[i];
        if (s.ut[i] != 0.0 || s.dut[i] != 0.0 || s.d2ut[i] != 0.0) last = i;
        bt[i] = make_double4(s.bt[i], s.dbt[i], s.d2bt[i], 0.0);
        bt1[i] = s.bt[i];
        bt2[i] = make_double2(s.dbt[i], s.d2bt[i]);
        btb[i] = make_double4(s.bt_b[i], s.dbt_b[i], s.d2bt_b[i], 0.0);
        btd[i] = make_double4(s.bt_d[i], s.dbt_d[i], s.d2bt_d[i], 0.0);
        // the moves' pair records: beta and the Jastrow on their common grid (both tables
        // span [0, L/2] in ntab intervals)
        mt[i] = make_double4(s.bt[i], s.dbt[i], s.ut[i], s.dut[i]);
        mt32[i] = make_float4(float(s.bt[i]), float(s.dbt[i]), float(s.ut[i]), float(s.dut[i]));
    }
    HIPCHK(hipMemcpyAsync(const_cast<double4 *>(c->ds.btb), btb.data(), sizeof(double4) * n1,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(const_cast<double4 *>(c->ds.btd), btd.data(), sizeof(double4) * n1,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(const_cast<double4 *>(c->ds.jt), jt.data(), sizeof(double4) * n1,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(const_cast<double4 *>(c->ds.bt), bt.data(), sizeof(double4) * n1,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(const_cast<double *>(c->ds.bt1), bt1.data(), sizeof(double) * n1,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(const_cast<double *>(c->ds.vt1), vt1.data(), sizeof(double) * n1,
                          hipMemcpyHostToDevice, c->st));
    c->ds.jfar = last + 2;
    HIPCHK(hipMemcpyAsync(const_cast<double2 *>(c->ds.bt2), bt2.data(), sizeof(double2) * n1,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(const_cast<double4 *>(c->ds.mt64), mt.data(), sizeof(double4) * n1,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(const_cast<float4 *>(c->ds.mt32), mt32.data(), sizeof(float4) * n1,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    // The end of a single-particle sweep reads these tables from global memory, one walker per
    // workgroup (1.52 ms at N = 66 with 32k walkers).  DMC_HPSI_TABLES=lds selects the kernel
    // that stages them in the LDS of workgroups of eight walkers (1.76 ms), where such a
    // workgroup fits the device.
    c->ds.tab_lds = 0;
    const char *tv = getenv("DMC_HPSI_TABLES");
    if (c->ds.keep_sinv && tv && !strcmp(tv, "lds")) c->ds.tab_lds = sm_end_tab_fits(s.n, c->ds);
    c->ds.mv_lds = 0;
    const char *mv = getenv("DMC_MOVE_TABLES");
    if (c->ds.keep_sinv && c->ds.sinv_f32 && mv && !strcmp(mv, "lds"))
        c->ds.mv_lds = sm_moves_tab_fits(s.n, c->ds);
    return 0;
}

int upload_q(dmc_ctx *c) {
    const System &s = c->sys;
    const int nq = s.nk - 1;
    std::vector<int4> qm(nq > 0 ? nq : 1);
    std::vector<double> wq(nq > 0 ? nq : 1);
    for (int q = 0; q < nq; ++q) {
        const auto &v = s.ik[q + 1];
        qm[q] = make_int4(v[0], v[1], v[2], s.kshell[q + 1]);
        wq[q] = 2.0 * s.vorb[s.kshell[q + 1]];
    }
    c->ds.alpha0 = s.vorb[0];
    if (nq > 0) {
        HIPCHK(hipMemcpyAsync(const_cast<int4 *>(c->ds.qm), qm.data(), sizeof(int4) * nq,
                              hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(const_cast<double *>(c->ds.wq), wq.data(), sizeof(double) * nq,
                              hipMemcpyHostToDevice, c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    return 0;
}

void fill_devsys(dmc_ctx *c) {
    const System &s = c->sys;
    DevSys &d = c->ds;
    d.npart = s.npart;
    d.n = s.n;
    d.nq = s.nk - 1;
    d.ntab = s.ntab;
    d.nobeta = s.nobeta;
    d.cutdrift = s.cutdrift;
    d.fn = s.p.fn;
    d.nwalk = s.p.nwalk;
    d.el = s.el;
    d.eli = s.eli;
    d.range = s.range;
    d.scale = s.scale;
    d.rangeb = s.rangeb;
    d.scaleb = s.scaleb;
    d.sigma = s.sigma;
    d.sig2 = s.sig2;
    d.sqrtm = s.sqrtm;
    d.invsqrtm = 1.0 / s.sqrtm;
    d.invm = 1.0 / s.mass;
    d.invmu = 1.0 / (s.mass * s.mass / (s.mass + s.mass));
    d.hbar = s.hbar;
    d.dt = s.p.dt;
    d.etrial = s.etrial;
    d.wtcut = s.p.wtcut;
    d.dcut = s.dcut_abs;
    d.nfac = s.nfac;
    d.twopi_l = 2.0 * 3.14159265358979323846 / s.el;
    d.popup = s.p.popup;
    d.popdn = s.p.popdn;
    d.keep_sinv = s.p.movetype == DMC_SINGLEPARTICLE;
    d.sinv_f32 = s.p.precision == DMC_MIXED;
}

int set_count(dmc_ctx *c, int n) {
    c->h_nt[0] = n;
    HIPCHK(hipMemcpyAsync(c->T.nt, c->h_nt, sizeof(int), hipMemcpyHostToDevice, c->st));
    return 0;
}

int read_total(dmc_ctx *c, int &total) {
    HIPCHK(hipMemcpyAsync(c->h_nt + 1, c->T.nt + 1, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    total = c->h_nt[1];
    return 0;
}

// one branch pass: multiplicities already in T.count for the n walkers of src; compacts
// into dst and returns the new count
int compact(dmc_ctx *c, const DevState &src, const DevState &dst, int n, int &total) {
    LAUNCHCHK(launch_scan(c->T, c->st));
    int st = read_total(c, total);
    if (st) return st;
    if (total > c->capacity)
        return fail(DMC_ERR_OVERFLOW, "stack overflow: population " + std::to_string(total) +
                                          " exceeds capacity " + std::to_string(c->capacity));
    LAUNCHCHK(launch_map_gather(c->ds, src, dst, c->T, n, total, c->capacity, c->st));
    return 0;
}

// q vectors of S(q) for ntabq shells (cached per shell count)
int sofq_vectors(dmc_ctx *c, int ntabq) {
    if (ntabq < 1) return fail(DMC_ERR_ARG, "ntabq must be positive");
    if (c->sofq_ntabq == ntabq) return DMC_OK;
    std::string err;
    c->sofq_ntabq = 0;
    int st = enumerate_kshells(ntabq, c->sofq_ik, c->sofq_numsh, err);
    if (st) return fail(st, err);
    c->sofq_ntabq = ntabq;
    return DMC_OK;
}

// Largest dynamic LDS of one observables workgroup.
constexpr size_t kObsLdsMax = 64 * 1024;
// n(k) shells up to |m|^2 = 63: every component <= 7, the kernel's cosine table
constexpr int kNofkMaxTab = 63;
// walkers per n(k) batch: bounds the per-walker buffers (37 KB per walker at N = 66)
constexpr int kNofkBatch = 8192;

// g(r)/density (ntab > 0) or S(q) (ntabq > 0) sums of the resident population, in the
// entry order of ObsArgs
int run_obs(dmc_ctx *c, int ntab, int ntabq, std::vector<double> &sums) {
    if (c->pending) return fail(DMC_ERR_ARG, "a propagation awaits dmc_branch");
    const int nw = c->nw, np = c->sys.npart, nup = c->sys.n;
    int nq = 0, mmax = 0;
    std::vector<int4> qm;
    if (ntabq > 0) {
        int st = sofq_vectors(c, ntabq);
        if (st) return st;
        nq = int(c->sofq_ik.size());
        qm.resize(nq);
        for (int q = 0; q < nq; ++q) {
            const auto &m = c->sofq_ik[q];
            qm[q] = make_int4(m[0], m[1], m[2], 0);
            for (int d = 0; d < 3; ++d) mmax = std::max(mmax, std::abs(m[d]));
        }
    }
    const int ng = ntab > 0 ? 3 * (ntab + 1) + 6 * (2 * ntab + 1) : 0;
    const int nsum = 1 + ng + 6 * nq;
    sums.assign(nsum, 0.0);
    if (nw == 0) return DMC_OK;
    const bool part_global = obs_lds_bytes(np, ntab, nq, mmax, false) > kObsLdsMax;
    if (obs_lds_bytes(np, ntab, nq, mmax, true) > kObsLdsMax)
        return fail(DMC_ERR_ARG, "too many bins or q vectors for the observables kernel");
    const int nblk = obs_blocks(nw), nseg = (nblk + 31) / 32;
    const size_t nd = (size_t(nblk + nseg + 1) * nsum + 1) & ~size_t(1);
    int e = c->ensure_scratch(sizeof(double) * nd + sizeof(int4) * size_t(nq) + 256);
    if (e) return e;
    double *part = static_cast<double *>(c->scratch), *tmp = part + size_t(nblk) * nsum,
           *out = tmp + size_t(nseg) * nsum;
    int4 *dq = reinterpret_cast<int4 *>(part + nd);
    if (nq > 0)
        HIPCHK(hipMemcpyAsync(dq, qm.data(), sizeof(int4) * nq, hipMemcpyHostToDevice, c->st));
    ObsArgs a{};
    a.x = c->A.x;
    a.weight = c->A.weight;
    a.wt0 = c->A.wt0;
    a.is = c->A.is;
    a.nw = nw;
    a.npart = np;
    a.nup = nup;
    a.el = c->sys.el;
    a.eli = c->sys.eli;
    a.half = 0.5 * c->sys.el;
    a.scale = ntab > 0 ? 1.0 / (0.5 * c->sys.el / ntab) : 0.0;  // gofr.f90:35-36
    a.ntab = ntab;
    a.qm = dq;
    a.nq = nq;
    a.mmax = mmax;
    a.twopi_l = 2.0 * (4.0 * std::atan(1.0)) / c->sys.el;
    a.nsum = nsum;
    a.nblk = nblk;
    a.part_global = part_global ? 1 : 0;
    a.part = part;
    LAUNCHCHK(launch_obs(a, tmp, out, c->st));
    HIPCHK(hipMemcpyAsync(sums.data(), out, sizeof(double) * nsum, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

}  // namespace

extern "C" {

const char *dmc_last_error(void) { return g_err.c_str(); }

int dmc_create(const dmc_params *p, dmc_ctx **out) {
    if (!p || !out) return fail(DMC_ERR_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(DMC_ERR_NODEVICE, "no HIP device");
    dmc_ctx *c = new dmc_ctx();
    std::string err;
    int st = build_system(*p, c->sys, err);
    if (st) {
        delete c;
        return fail(st, err);
    }
    const System &s = c->sys;
    if (!nup_supported(s.n)) {
        delete c;
        return fail(DMC_ERR_ARG, "npart/2 = " + std::to_string(s.n) +
                                     " has no compiled kernel (supported: 1, 7, 19, 33)");
    }
    for (int q = 1; q < s.nk; ++q)
        for (int d = 0; d < 3; ++d)
            if (std::abs(s.ik[q][d]) > 4) {
                delete c;
                return fail(DMC_ERR_ARG, "k vector component beyond 4 (too many shells)");
            }
    if (p->precision != DMC_FP64 && p->precision != DMC_MIXED) {
        delete c;
        return fail(DMC_ERR_ARG, "precision must be DMC_FP64 or DMC_MIXED");
    }
    if (p->movetype != DMC_ALLPARTICLE && p->movetype != DMC_SINGLEPARTICLE &&
        p->movetype != DMC_VMC && p->movetype != DMC_METROPOLIS) {
        delete c;
        return fail(DMC_ERR_ARG, "movetype must be DMC_ALLPARTICLE, DMC_SINGLEPARTICLE, DMC_VMC "
                                 "or DMC_METROPOLIS");
    }
    if (p->krefresh < 0) {
        delete c;
        return fail(DMC_ERR_ARG, "krefresh must be >= 0");
    }
    if (p->movetype != DMC_SINGLEPARTICLE && p->precision != DMC_FP64) {
        delete c;
        return fail(DMC_ERR_ARG, "all-particle, Metropolis and VMC moves run in FP64 only");
    }
    c->capacity = p->capacity > 0 ? p->capacity : 4 * (p->nwalk > 0 ? p->nwalk : 1);
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess) {
        delete c;
        return fail(DMC_ERR_HIP, "hipStreamCreate failed");
    }
    // the no-alloc entry point's device buffer on this device (a failure here only disables
    // that entry point, which then reports it)
    c->noalloc = noalloc_scratch_address(c->device);
    fill_devsys(c);
    const int n1 = s.ntab + 1, nq = s.nk - 1;
    double4 *jt, *bt, *btb, *btd, *mt64;
    double *bt1, *vt1;
    double2 *bt2;
    float4 *mt32;
    int4 *qm;
    double *wq;
    uint64_t *ja1, *jc1, *ja2, *jc2;
    if (c->alloc(&jt, n1) || c->alloc(&bt, n1) || c->alloc(&btb, n1) || c->alloc(&btd, n1) ||
        c->alloc(&bt1, n1) || c->alloc(&bt2, n1) || c->alloc(&vt1, n1) ||
        c->alloc(&mt64, n1) || c->alloc(&mt32, n1) ||
        c->alloc(&qm, nq) || c->alloc(&wq, nq) || c->alloc(&ja1, kMaxDraw) ||
        c->alloc(&jc1, kMaxDraw) || c->alloc(&ja2, kMaxDraw) || c->alloc(&jc2, kMaxDraw) ||
        c->alloc_state(c->A) || c->alloc_state(c->B) || c->alloc(&c->T.count, c->capacity) ||
        c->alloc(&c->T.offset, c->capacity) || c->alloc(&c->T.map, c->capacity) ||
        c->alloc(&c->T.copy_irn, c->capacity) || c->alloc(&c->T.est, size_t(c->capacity) * 42) ||
        c->alloc(&c->T.nt, 8) || c->alloc(&c->T.estp, 64 * 42) || c->alloc(&c->T.hist, 500) || c->alloc(&c->T.nacc, c->capacity) ||
        c->alloc(&c->T.moves, 2) ||
        c->alloc(&c->est_acc, 42) || c->alloc(&c->opt_acc, dmc_ctx::kOptSums)) {
        std::string e = g_err;
        dmc_destroy(c);
        return fail(DMC_ERR_HIP, e);
    }
    c->ds.jt = jt;
    c->ds.bt = bt;
    c->ds.bt1 = bt1;
    c->ds.vt1 = vt1;
    c->ds.bt2 = bt2;
    c->ds.btb = btb;
    c->ds.btd = btd;
    c->ds.mt64 = mt64;
    c->ds.mt32 = mt32;
    c->ds.qm = qm;
    c->ds.wq = wq;
    c->ds.ja1 = ja1;
    c->ds.jc1 = jc1;
    c->ds.ja2 = ja2;
    c->ds.jc2 = jc2;
    bool evok = true;
    for (hipEvent_t &e : c->evs) evok = evok && hipEventCreate(&e) == hipSuccess;
    if (!evok || hipHostMalloc(reinterpret_cast<void **>(&c->h_nt), 4 * sizeof(int)) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        dmc_destroy(c);
        return fail(DMC_ERR_HIP, "host/event allocation failed");
    }
    std::vector<uint64_t> a1, c1, a2, c2;
    ran1_jump_table(kMaxDraw, a1, c1);
    ran2_jump_table(kMaxDraw, a2, c2);
    int e = 0;
    if (hipMemcpy(ja1, a1.data(), 8 * kMaxDraw, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(jc1, c1.data(), 8 * kMaxDraw, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ja2, a2.data(), 8 * kMaxDraw, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(jc2, c2.data(), 8 * kMaxDraw, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->est_acc, 0, 42 * sizeof(double)) != hipSuccess ||
        hipMemset(c->T.nt, 0, 8 * sizeof(int)) != hipSuccess ||
        hipMemset(c->opt_acc, 0, dmc_ctx::kOptSums * sizeof(double)) != hipSuccess ||
        hipMemset(c->T.hist, 0, 500 * sizeof(int64_t)) != hipSuccess ||
        hipMemset(c->T.moves, 0, 2 * sizeof(int64_t)) != hipSuccess)
        e = DMC_ERR_HIP;
    if (!e) e = upload_tables(c);
    if (!e) e = upload_q(c);
    if (e) {
        std::string m = g_err.empty() ? "upload failed" : g_err;
        dmc_destroy(c);
        return fail(DMC_ERR_HIP, m);
    }
    *out = c;
    return DMC_OK;
}

int dmc_destroy(dmc_ctx *c) {
    if (!c) return DMC_OK;
    // best-effort teardown: errors here have nowhere to go
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->trace) (void)hipFree(c->trace);
    if (c->h_nt) (void)hipHostFree(c->h_nt);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t &e : c->evs)
        if (e) (void)hipEventDestroy(e);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return DMC_OK;
}

int dmc_info(const dmc_ctx *c, double out[8]) {
    if (!c || !out) return fail(DMC_ERR_ARG, "null argument");
    const System &s = c->sys;
    out[0] = s.el;
    out[1] = s.eunit;
    out[2] = s.etrial;
    out[3] = s.nfac;
    out[4] = s.nk;
    out[5] = s.ntab0;
    out[6] = s.sigma;
    out[7] = c->capacity;
    return DMC_OK;
}

int dmc_set_tables(dmc_ctx *c, const double *vt, const double *ut, const double *dut,
                   const double *d2ut, const double *bt, const double *dbt, const double *d2bt) {
    if (!c || !vt || !ut || !dut || !d2ut || !bt || !dbt || !d2bt)
        return fail(DMC_ERR_ARG, "null argument");
    System &s = c->sys;
    const int n1 = s.ntab + 1;
    s.vt.assign(vt, vt + n1);
    s.ut.assign(ut, ut + n1);
    s.dut.assign(dut, dut + n1);
    s.d2ut.assign(d2ut, d2ut + n1);
    s.bt.assign(bt, bt + n1);
    s.dbt.assign(dbt, dbt + n1);
    s.d2bt.assign(d2bt, d2bt + n1);
    // a table from the host carries no (b, d) dependence: those derivatives become NaN
    const double nan = std::nan("");
    for (auto *t : {&s.bt_b, &s.dbt_b, &s.d2bt_b, &s.bt_d, &s.dbt_d, &s.d2bt_d}) t->assign(n1, nan);
    return upload_tables(c);
}

int dmc_set_params(dmc_ctx *c, const double *vorb, double bbeta, double dbeta) {
    if (!c || !vorb) return fail(DMC_ERR_ARG, "null argument");
    System &s = c->sys;
    s.vorb.assign(vorb, vorb + s.p.norb);
    if (!s.nobeta && (bbeta != s.b || dbeta != s.d)) {
        build_beta(s, bbeta, dbeta);
        int e = upload_tables(c);
        if (e) return e;
    }
    return upload_q(c);
}

int dmc_load_walkers(dmc_ctx *c, int32_t nw, const double *x, const int64_t *irn) {
    if (!c || !x || !irn || nw < 0) return fail(DMC_ERR_ARG, "invalid argument");
    if (nw > c->capacity) return fail(DMC_ERR_OVERFLOW, "more walkers than capacity");
    const size_t m = size_t(3) * c->sys.npart;
    HIPCHK(hipMemcpyAsync(c->A.x, x, sizeof(double) * m * nw, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->A.irn, irn, sizeof(int64_t) * nw, hipMemcpyHostToDevice, c->st));
    LAUNCHCHK(launch_eval(c->sys.n, c->ds, c->A.x, nw, c->A.psil, c->A.is, c->A.dpsi, c->A.d2psi,
                          c->A.v, c->ds.keep_sinv ? c->A.sinv : nullptr, c->st));
    if (nw > 0) hipLaunchKernelGGL(k_init_weights, dim3((nw + 255) / 256), dim3(256), 0, c->st, c->A, nw);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    c->nw = nw;
    c->pending = false;
    return DMC_OK;
}

int dmc_init_walkers(dmc_ctx *c, int64_t seed, int32_t nwalk) {
    if (!c || nwalk < 0) return fail(DMC_ERR_ARG, "invalid argument");
    std::vector<double> x;
    std::vector<int64_t> irn;
    initial_walkers(c->sys, seed, nwalk, x, irn);
    return dmc_load_walkers(c, nwalk, x.data(), irn.data());
}

int dmc_population(const dmc_ctx *c, int32_t *nw) {
    if (!c || !nw) return fail(DMC_ERR_ARG, "null argument");
    *nw = c->nw;
    return DMC_OK;
}

int dmc_read_walkers(dmc_ctx *c, int32_t cap, int32_t *nw, double *x, double *dpsi, double *sc,
                     int64_t *irn) {
    if (!c || !nw) return fail(DMC_ERR_ARG, "null argument");
    if (c->pending) return fail(DMC_ERR_ARG, "a propagation awaits dmc_branch");
    const int n = c->nw;
    *nw = n;
    if (cap < n) return fail(DMC_ERR_ARG, "output capacity smaller than the population");
    const size_t m = size_t(3) * c->sys.npart;
    if (x) HIPCHK(hipMemcpyAsync(x, c->A.x, sizeof(double) * m * n, hipMemcpyDeviceToHost, c->st));
    if (dpsi)
        HIPCHK(hipMemcpyAsync(dpsi, c->A.dpsi, sizeof(double) * m * n, hipMemcpyDeviceToHost, c->st));
    if (irn) HIPCHK(hipMemcpyAsync(irn, c->A.irn, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->st));
    if (sc) {
        std::vector<double> wgt(n), psil(n), d2(n), v(n), wt0(n);
        std::vector<int> is(n);
        HIPCHK(hipMemcpyAsync(wgt.data(), c->A.weight, 8 * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(psil.data(), c->A.psil, 8 * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(d2.data(), c->A.d2psi, 8 * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(v.data(), c->A.v, 8 * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(wt0.data(), c->A.wt0, 8 * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(is.data(), c->A.is, 4 * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (int w = 0; w < n; ++w) {
            double *o = sc + 6 * size_t(w);
            o[0] = wgt[w];
            o[1] = psil[w];
            o[2] = d2[w];
            o[3] = v[w];
            o[4] = is[w];
            o[5] = wt0[w];
        }
    }
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

int dmc_eval(dmc_ctx *c, int32_t nw, const double *x, double *psil, int32_t *is, double *dpsi,
             double *d2psi, double *v) {
    if (!c || !x || nw < 0) return fail(DMC_ERR_ARG, "invalid argument");
    if (nw == 0) return DMC_OK;
    const size_t m = size_t(3) * c->sys.npart;
    const size_t bytes = sizeof(double) * (2 * m * nw + 3 * size_t(nw)) + sizeof(int) * nw + 256;
    int e = c->ensure_scratch(bytes);
    if (e) return e;
    double *dx = static_cast<double *>(c->scratch), *dg = dx + m * nw, *dp = dg + m * nw,
           *dd = dp + nw, *dv = dd + nw;
    int *di = reinterpret_cast<int *>(dv + nw);
    HIPCHK(hipMemcpyAsync(dx, x, sizeof(double) * m * nw, hipMemcpyHostToDevice, c->st));
    LAUNCHCHK(launch_eval(c->sys.n, c->ds, dx, nw, dp, di, dg, dd, dv, nullptr, c->st));
    if (psil) HIPCHK(hipMemcpyAsync(psil, dp, 8 * nw, hipMemcpyDeviceToHost, c->st));
    if (is) HIPCHK(hipMemcpyAsync(is, di, 4 * nw, hipMemcpyDeviceToHost, c->st));
    if (dpsi) HIPCHK(hipMemcpyAsync(dpsi, dg, sizeof(double) * m * nw, hipMemcpyDeviceToHost, c->st));
    if (d2psi) HIPCHK(hipMemcpyAsync(d2psi, dd, 8 * nw, hipMemcpyDeviceToHost, c->st));
    if (v) HIPCHK(hipMemcpyAsync(v, dv, 8 * nw, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

int dmc_ratio1(dmc_ctx *c, int32_t nw, const double *x, const int32_t *ip, const double *xnew,
               double *ratio) {
    if (!c || !x || !ip || !xnew || !ratio || nw < 0) return fail(DMC_ERR_ARG, "invalid argument");
    if (nw == 0) return DMC_OK;
    for (int w = 0; w < nw; ++w)
        if (ip[w] < 0 || ip[w] >= c->sys.npart) return fail(DMC_ERR_ARG, "particle index out of range");
    const size_t m = size_t(3) * c->sys.npart;
    const size_t bytes = sizeof(double) * (m * nw + 4 * size_t(nw)) + sizeof(int) * nw + 256;
    int e = c->ensure_scratch(bytes);
    if (e) return e;
    double *dx = static_cast<double *>(c->scratch), *dn = dx + m * nw, *dr = dn + 3 * nw;
    int *di = reinterpret_cast<int *>(dr + nw);
    HIPCHK(hipMemcpyAsync(dx, x, sizeof(double) * m * nw, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dn, xnew, sizeof(double) * 3 * nw, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(di, ip, sizeof(int) * nw, hipMemcpyHostToDevice, c->st));
    LAUNCHCHK(launch_ratio(c->sys.n, c->ds, dx, di, dn, nw, dr, c->st));
    HIPCHK(hipMemcpyAsync(ratio, dr, 8 * nw, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

int dmc_propagate(dmc_ctx *c, int32_t opcomp) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    const int n = c->nw;
    if (n <= 0) return fail(DMC_ERR_ARG, "no walkers loaded");
    if (c->pending) return fail(DMC_ERR_ARG, "dmc_propagate called twice without dmc_branch");
    int e = set_count(c, n);
    if (e) return e;
    // the dominant kernel is timed with events on the stream it runs on: the move kernel
    // (every chunk) for single-particle moves, the fused step kernel otherwise
    int nev = 0;
    if (c->sys.p.movetype == DMC_SINGLEPARTICLE) {
        // moves in chunks of krefresh, an FP64 refresh of the inverse between chunks, then
        // the end-of-sweep FP64 evaluation; A.sinv serves as the working inverse
        const int np = c->sys.npart;
        if (c->trace) c->trace_nw = n;
        const int k = c->sys.p.krefresh > 0 ? c->sys.p.krefresh : np;
        for (int m0 = 0; m0 < np; m0 += k) {
            const int m1 = m0 + k < np ? m0 + k : np;
            if (m0 > 0) LAUNCHCHK(launch_sm_refresh(c->sys.n, c->ds, c->B.x, c->A.sinv, c->T, n, c->st));
            const bool ev = c->timing && nev < dmc_ctx::kEv;
            if (ev) HIPCHK(hipEventRecord(c->evs[2 * nev], c->st));
            LAUNCHCHK(launch_sm_moves(c->sys.n, c->ds, m0 ? c->B.x : c->A.x, c->B.x,
                                      m0 ? c->B.irn : c->A.irn, c->B.irn, c->A.sinv, m0, m1,
                                      c->T, n, c->st));
            if (ev) HIPCHK(hipEventRecord(c->evs[2 * nev++ + 1], c->st));
        }
        LAUNCHCHK(launch_sm_end(c->sys.n, c->ds, c->A, c->B, c->T, opcomp, n, c->st));
    } else if (c->sys.p.movetype == DMC_VMC) {
        LAUNCHCHK(launch_vmc(c->sys.n, c->ds, c->A, c->B, c->T, n, c->st));
        LAUNCHCHK(launch_est_reduce(c->T, n, c->est_acc, c->st));
        std::swap(c->A, c->B);
        return DMC_OK;
    } else if (c->sys.p.movetype == DMC_METROPOLIS) {
        if (c->timing) HIPCHK(hipEventRecord(c->evs[0], c->st));
        LAUNCHCHK(launch_mdmc(c->sys.n, c->ds, c->A, c->B, c->T, opcomp, n, c->st));
        if (c->timing) HIPCHK(hipEventRecord(c->evs[1], c->st));
        nev = c->timing ? 1 : 0;
    } else {
        if (c->timing) HIPCHK(hipEventRecord(c->evs[0], c->st));
        LAUNCHCHK(launch_step(c->sys.n, c->ds, c->A, c->B, c->T, opcomp, n, c->st));
        if (c->timing) HIPCHK(hipEventRecord(c->evs[1], c->st));
        nev = c->timing ? 1 : 0;
    }
    LAUNCHCHK(launch_est_reduce(c->T, n, c->est_acc, c->st));
    c->pending = true;
    c->nev_pending = nev;
    return DMC_OK;
}

int dmc_branch(dmc_ctx *c, double factor, int32_t *nt) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    const int n = c->nw;
    int total = 0, e = 0;
    if (c->pending) {
        // the multiplicities int(weight + u) were drawn by the propagation (factor 1)
        if (factor != 1.0)
            return fail(DMC_ERR_ARG, "the branch that follows dmc_propagate takes factor 1");
        c->pending = false;
        e = compact(c, c->B, c->A, n, total);
        if (e == DMC_ERR_OVERFLOW && c->ds.keep_sinv) {
            // the population stays the one before the propagation, but its moves updated
            // A.sinv in place: rebuild that inverse from A.x (B's arrays, which the failed
            // step leaves unused, take the other outputs) so that stepping can go on
            std::string msg = g_err;
            LAUNCHCHK(launch_eval(c->sys.n, c->ds, c->A.x, n, c->B.psil, c->B.is, c->B.dpsi,
                                  c->B.d2psi, c->B.v, c->A.sinv, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            return fail(e, msg);
        }
        if (e) return e;
        for (int i = 0; i < c->nev_pending; ++i) {  // the stream was synchronised by compact()
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->evs[2 * i], c->evs[2 * i + 1]));
            c->time_ms += ms;
            c->launches += 1;
        }
        c->nev_pending = 0;
    } else {
        if (c->sys.p.movetype == DMC_VMC && factor == 1.0)
            return fail(DMC_ERR_ARG, "VMC populations are not branched");
        // population control: int(weight * factor + u) copies of every walker
        if (n <= 0) return fail(DMC_ERR_ARG, "population died out");
        if (!(factor > 0.0)) return fail(DMC_ERR_ARG, "branch factor must be positive");
        e = set_count(c, n);
        if (e) return e;
        LAUNCHCHK(launch_branch_count(c->ds, c->A, c->T, factor, n, c->st));
        e = compact(c, c->A, c->B, n, total);
        if (e) return e;
        std::swap(c->A, c->B);
    }
    c->nw = total;
    if (nt) *nt = total;
    return DMC_OK;
}

int dmc_step(dmc_ctx *c, int32_t opcomp, int32_t *nt) {
    int e = dmc_propagate(c, opcomp);
    if (e) return e;
    if (c->sys.p.movetype == DMC_VMC) {
        if (nt) nt[0] = nt[1] = c->nw;
        return DMC_OK;
    }
    int total = 0, final_n = 0;
    e = dmc_branch(c, 1.0, &total);
    if (e) return e;
    final_n = total;
    const double nwalk = c->sys.p.nwalk;
    if (total > c->sys.p.popup * nwalk || total < c->sys.p.popdn * nwalk) {
        if (total == 0) return fail(DMC_ERR_ARG, "population died out");
        e = dmc_branch(c, nwalk / total, &final_n);
        if (e) return e;
    }
    if (nt) {
        nt[0] = total;
        nt[1] = final_n;
    }
    return DMC_OK;
}

int dmc_run(dmc_ctx *c, int32_t nsteps, int32_t *nt) {
    int32_t tmp[2] = {0, 0};
    for (int s = 0; s < nsteps; ++s) {
        int e = dmc_step(c, s == nsteps - 1, tmp);
        if (e) return e;
    }
    if (nt) *nt = c ? c->nw : 0;
    return DMC_OK;
}

int dmc_estimators(dmc_ctx *c, double *est, int32_t reset) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    if (est) HIPCHK(hipMemcpyAsync(est, c->est_acc, 42 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    if (reset) HIPCHK(hipMemsetAsync(c->est_acc, 0, 42 * sizeof(double), c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

int dmc_wtisto(dmc_ctx *c, int64_t *hist) {
    if (!c || !hist) return fail(DMC_ERR_ARG, "null argument");
    HIPCHK(hipMemcpyAsync(hist, c->T.hist, 500 * sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

int dmc_acceptance(dmc_ctx *c, int64_t *accepted, int64_t *proposed) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    int64_t m[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(m, c->T.moves, sizeof(m), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemsetAsync(c->T.moves, 0, sizeof(m), c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (accepted) *accepted = m[0];
    if (proposed) *proposed = m[1];
    return DMC_OK;
}

int dmc_gofr(dmc_ctx *c, int32_t ntab, double *gnow, double *gerr, double *dnow, double *derr,
             double *wsblk) {
    if (!c || !gnow || !gerr || !dnow || !derr || !wsblk) return fail(DMC_ERR_ARG, "null argument");
    if (ntab < 1) return fail(DMC_ERR_ARG, "ntab must be positive");
    std::vector<double> sums;
    int st = run_obs(c, ntab, 0, sums);
    if (st) return st;
    *wsblk += sums[0];
    const double *g = sums.data() + 1, *d = g + 3 * (ntab + 1);
    for (int e = 0; e < 3 * (ntab + 1); ++e) {
        gnow[e] += 2.0 * g[e];
        gerr[e] += 4.0 * g[e];
    }
    for (int b = 0; b <= 2 * ntab; ++b)
        for (int ch = 0; ch < 6; ++ch) {
            const int e = 6 * b + ch;
            dnow[e] += (b == ntab ? 0.5 : 1.0) * d[e];
            derr[e] += (b == ntab ? 0.25 : 1.0) * d[e];
        }
    return DMC_OK;
}

int dmc_sofq_kvec(dmc_ctx *c, int32_t ntabq, int32_t cap, int32_t *nqtot, int32_t *numsh,
                  double *ak) {
    if (!c || !nqtot) return fail(DMC_ERR_ARG, "null argument");
    int st = sofq_vectors(c, ntabq);
    if (st) return st;
    const int nq = int(c->sofq_ik.size());
    *nqtot = nq;
    if (numsh)
        for (int i = 0; i < ntabq; ++i) numsh[i] = c->sofq_numsh[i];
    if (ak) {
        if (cap < nq) return fail(DMC_ERR_ARG, "output capacity smaller than the q vectors");
        const double piel = 2.0 * (4.0 * std::atan(1.0)) / c->sys.el;  // kshell.f90:137-138
        for (int q = 0; q < nq; ++q)
            for (int d = 0; d < 3; ++d) ak[3 * q + d] = piel * c->sofq_ik[q][d];
    }
    return DMC_OK;
}

int dmc_sofq(dmc_ctx *c, int32_t ntabq, double *snow, double *serr, double *wsblk) {
    if (!c || !snow || !serr || !wsblk) return fail(DMC_ERR_ARG, "null argument");
    std::vector<double> sums;
    int st = run_obs(c, 0, ntabq, sums);
    if (st) return st;
    *wsblk += sums[0];
    const int nq = int(c->sofq_ik.size());
    for (int q = 0; q < nq; ++q)
        for (int k = 0; k < 3; ++k) {
            snow[3 * q + k] += sums[1 + 6 * q + k];
            serr[3 * q + k] += sums[1 + 6 * q + 3 + k];
        }
    return DMC_OK;
}

int dmc_nofk_shells(dmc_ctx *c, int32_t ntab, int32_t cap, int32_t *nsh, double *kmag,
                    int32_t *nk) {
    if (!c || !nsh) return fail(DMC_ERR_ARG, "null argument");
    if (ntab < 0 || ntab > kNofkMaxTab) return fail(DMC_ERR_ARG, "ntab of n(k) must be 0..63");
    std::vector<std::array<int, 3>> ik;
    std::vector<int> first, count;
    enk_shells(ntab, ik, first, count);
    const int n = int(count.size());
    *nsh = n;
    if ((kmag || nk) && cap < n) return fail(DMC_ERR_ARG, "output capacity smaller than the shells");
    const double dk = 2.0 * (4.0 * std::atan(1.0)) / c->sys.el;
    for (int k = 0; k < n; ++k) {
        const auto &m = ik[first[k]];
        if (kmag) kmag[k] = std::sqrt(dk * dk * double(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]));
        if (nk) nk[k] = count[k];
    }
    return DMC_OK;
}

int dmc_nofk(dmc_ctx *c, int32_t ntab, double *enknow, double *enkerr, double *wsblk) {
    if (!c || !enknow || !enkerr || !wsblk) return fail(DMC_ERR_ARG, "null argument");
    if (ntab < 0 || ntab > kNofkMaxTab) return fail(DMC_ERR_ARG, "ntab of n(k) must be 0..63");
    if (c->pending) return fail(DMC_ERR_ARG, "a propagation awaits dmc_branch");
    std::vector<std::array<int, 3>> ik;
    std::vector<int> first, count;
    enk_shells(ntab, ik, first, count);
    const int nsh = int(count.size()), ntot = int(ik.size()), nw = c->nw, np = c->sys.npart;
    const int nup = c->sys.n, nq = c->ds.nq;
    const int nsamp = 10;  // momentum3d.f90:22
    const int ncol = 1 + 2 * nsh, nmove = np * nsamp, nchunk = (nmove + 63) / 64;
    if (nw == 0) return DMC_OK;
    if (nofk_prep_lds_bytes(nup, nq) > kObsLdsMax)
        return fail(DMC_ERR_ARG, "too many plane waves for the n(k) kernel's LDS");
    // state after 3 m draws of ran1, m < npart nsamp
    std::vector<uint64_t> ja(nmove), jc(nmove);
    {
        std::vector<uint64_t> a1, c1;
        ran1_jump_table(3 * nmove, a1, c1);
        ja[0] = 1;
        jc[0] = 0;
        for (int m = 1; m < nmove; ++m) {
            ja[m] = a1[3 * m - 1];
            jc[m] = c1[3 * m - 1];
        }
    }
    std::vector<int4> hik(ntot);
    for (int v = 0; v < ntot; ++v) hik[v] = make_int4(ik[v][0], ik[v][1], ik[v][2], 0);
    std::vector<int2> hsh(nsh);
    for (int k = 0; k < nsh; ++k) hsh[k] = make_int2(first[k], count[k]);
    // per-walker buffers for batches of at most kNofkBatch walkers
    const int batch = std::min(nw, kNofkBatch);
    const int seg = 256, nseg = (nw + seg - 1) / seg;
    const size_t nrows = (size_t(nw + nseg + 1) * ncol + 1) & ~size_t(1);
    const size_t nproj = size_t(batch) * np * nq * 2, ncoef = size_t(batch) * np * nup,
                 nuold = size_t(batch) * np,
                 npart = (size_t(batch) * nchunk * nsh + 1) & ~size_t(1);  // keeps int4 aligned
    const size_t bytes = sizeof(double) * (nrows + nproj + ncoef + nuold + npart) +
                         sizeof(int) * (size_t(batch) + 4) + 2 * sizeof(uint64_t) * nmove +
                         sizeof(int4) * ntot + sizeof(int2) * nsh + 512;
    int e = c->ensure_scratch(bytes);
    if (e) return e;
    double *rows = static_cast<double *>(c->scratch), *tmp = rows + size_t(nw) * ncol,
           *out = tmp + size_t(nseg) * ncol;
    double2 *proj = reinterpret_cast<double2 *>(rows + nrows);
    double *coef = reinterpret_cast<double *>(proj) + nproj, *uold = coef + ncoef,
           *prt = uold + nuold;
    uint64_t *dja = reinterpret_cast<uint64_t *>(prt + npart), *djc = dja + nmove;
    int4 *dik = reinterpret_cast<int4 *>(djc + nmove);
    int2 *dsh = reinterpret_cast<int2 *>(dik + ntot);
    int *sing = reinterpret_cast<int *>(dsh + nsh);
    HIPCHK(hipMemcpyAsync(dja, ja.data(), 8 * nmove, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(djc, jc.data(), 8 * nmove, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dik, hik.data(), sizeof(int4) * ntot, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dsh, hsh.data(), sizeof(int2) * nsh, hipMemcpyHostToDevice, c->st));
    NofkArgs a{};
    a.x = c->A.x;
    a.weight = c->A.weight;
    a.wt0 = c->A.wt0;
    a.is = c->A.is;
    a.irn = c->A.irn;
    a.nsamp = nsamp;
    a.nsh = nsh;
    a.nchunk = nchunk;
    a.ik = dik;
    a.shell = dsh;
    a.ja = dja;
    a.jc = djc;
    a.dk = 2.0 * (4.0 * std::atan(1.0)) / c->sys.el;
    a.proj = proj;
    a.coef = coef;
    a.uold = uold;
    a.sing = sing;
    a.part = prt;
    a.rows = rows;
    for (int w0 = 0; w0 < nw; w0 += batch) {
        a.w0 = w0;
        a.nw = std::min(batch, nw - w0);
        LAUNCHCHK(launch_nofk(nup, c->ds, a, c->st));
    }
    LAUNCHCHK(launch_rows_reduce(rows, nw, ncol, seg, tmp, out, c->st));
    std::vector<double> sums(ncol);
    HIPCHK(hipMemcpyAsync(sums.data(), out, sizeof(double) * ncol, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    *wsblk += sums[0];
    for (int k = 0; k < nsh; ++k) {
        enknow[k] += sums[1 + k];
        enkerr[k] += sums[1 + nsh + k];
    }
    return DMC_OK;
}

int dmc_move_trace(dmc_ctx *c, int32_t enable) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    if (c->sys.p.movetype != DMC_SINGLEPARTICLE)
        return fail(DMC_ERR_ARG, "move records need single-particle moves");
    if (c->pending) return fail(DMC_ERR_ARG, "a propagation awaits dmc_branch");
    HIPCHK(hipStreamSynchronize(c->st));
    if (!enable) {
        if (c->trace) HIPCHK(hipFree(c->trace));
        c->trace = nullptr;
    } else if (!c->trace) {
        const size_t n = size_t(c->capacity) * c->sys.npart * kTraceRec;
        HIPCHK(hipMalloc(reinterpret_cast<void **>(&c->trace), sizeof(double) * n));
    }
    c->T.trace = c->trace;
    c->trace_nw = 0;
    return DMC_OK;
}

int dmc_read_move_trace(dmc_ctx *c, int32_t cap, int32_t *nw, double *rec) {
    if (!c || !nw) return fail(DMC_ERR_ARG, "null argument");
    if (!c->trace) return fail(DMC_ERR_ARG, "move records are off (dmc_move_trace)");
    *nw = c->trace_nw;
    if (!rec) return DMC_OK;
    if (cap < c->trace_nw) return fail(DMC_ERR_ARG, "output capacity smaller than the population");
    const size_t n = size_t(c->trace_nw) * c->sys.npart * kTraceRec;
    HIPCHK(hipMemcpyAsync(rec, c->trace, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

int dmc_kernel_timing(dmc_ctx *c, int32_t enable) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    c->timing = enable != 0;
    return DMC_OK;
}

int dmc_kernel_times(dmc_ctx *c, double *ms, int64_t *launches, int32_t reset) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    if (ms) *ms = c->time_ms;
    if (launches) *launches = c->launches;
    if (reset) {
        c->time_ms = 0.0;
        c->launches = 0;
    }
    return DMC_OK;
}

int dmc_sync(dmc_ctx *c) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

int dmc_opt_accumulate(dmc_ctx *c) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    if (c->pending) return fail(DMC_ERR_ARG, "a propagation awaits dmc_branch");
    const int nw = c->nw, P = c->sys.p.norb + 2;
    if (nw <= 0) return DMC_OK;
    int e = c->ensure_scratch(sizeof(double) * size_t(nw) * P + 256);
    if (e) return e;
    double *O = static_cast<double *>(c->scratch);
    LAUNCHCHK(launch_adjoint(c->sys.n, c->ds, c->A.x, nw, c->sys.p.norb, O, nullptr, c->st));
    LAUNCHCHK(launch_opt_accumulate(c->ds, c->A, O, nw, P, c->opt_acc, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

int dmc_opt_accumulate_noalloc(dmc_ctx *c) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    if (c->pending) return fail(DMC_ERR_ARG, "a propagation awaits dmc_branch");
    const int nw = c->nw, P = c->sys.p.norb + 2;
    if (nw <= 0) return DMC_OK;
    if (size_t(nw) * P > kNoAllocDoubles)
        return fail(DMC_ERR_ARG, "dmc_opt_accumulate_noalloc: " + std::to_string(nw) +
                                     " walkers x " + std::to_string(P) + " parameters exceed its " +
                                     std::to_string(kNoAllocDoubles) + "-double device buffer");
    if (!c->noalloc) {
        const std::string why = c->device >= 0 && c->device < kMaxDevices
                                    ? g_noalloc_err[c->device] : std::string("device index");
        return fail(DMC_ERR_HIP, "dmc_opt_accumulate_noalloc: no device buffer (" + why + ")");
    }
    // O is written for every walker by the adjoint kernel before the sums read it: nothing
    // in the buffer needs clearing
    double *O = c->noalloc;
    LAUNCHCHK(launch_adjoint(c->sys.n, c->ds, c->A.x, nw, c->sys.p.norb, O, nullptr, c->st));
    LAUNCHCHK(launch_opt_accumulate(c->ds, c->A, O, nw, P, c->opt_acc, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

int dmc_opt_sums(dmc_ctx *c, double *sums, int32_t reset) {
    if (!c) return fail(DMC_ERR_ARG, "null context");
    const int P = c->sys.p.norb + 2, n = 3 + 2 * P + P * P;
    if (sums) HIPCHK(hipMemcpyAsync(sums, c->opt_acc, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    if (reset) HIPCHK(hipMemsetAsync(c->opt_acc, 0, sizeof(double) * dmc_ctx::kOptSums, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

int dmc_adjoint(dmc_ctx *c, int32_t nw, const double *x, double *dlogpsi, double *delocal) {
    if (!c || !x || !dlogpsi || nw < 0) return fail(DMC_ERR_ARG, "invalid argument");
    if (nw == 0) return DMC_OK;
    const size_t m = size_t(3) * c->sys.npart, np = size_t(c->sys.p.norb) + 2;
    const size_t bytes = sizeof(double) * (m * nw + 2 * np * nw) + 256;
    int e = c->ensure_scratch(bytes);
    if (e) return e;
    double *dx = static_cast<double *>(c->scratch), *dl = dx + m * nw, *de = dl + np * nw;
    HIPCHK(hipMemcpyAsync(dx, x, sizeof(double) * m * nw, hipMemcpyHostToDevice, c->st));
    LAUNCHCHK(launch_adjoint(c->sys.n, c->ds, dx, nw, c->sys.p.norb, dl, delocal ? de : nullptr,
                             c->st));
    HIPCHK(hipMemcpyAsync(dlogpsi, dl, sizeof(double) * np * nw, hipMemcpyDeviceToHost, c->st));
    if (delocal)
        HIPCHK(hipMemcpyAsync(delocal, de, sizeof(double) * np * nw, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DMC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// Batches: several contexts stepped together, one captured graph per step

struct dmc_batch {
    std::vector<dmc_ctx *> ctx;     // in the caller's order
    std::vector<int> order;         // launch order (largest particle number first)
    BatchHost host;
    BatchArgs args{};
    hipStream_t st = nullptr;
    int *counters = nullptr;        // work counters, kBatchCounters
    int ncu = 1, nchunk = 1;
    int wpg[3] = {1, 1, 1};         // walkers per workgroup: moves, refresh, end
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    bool dirty = true, host_current = true;
    std::vector<char *> captured_a;  // A.x of every system when the graphs were captured
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // one stream per system (default priority) for steps whose phases are all per-system
    // launches: forked from st after the step's first launch and joined back at its end
    hipStream_t side[kMaxBatch] = {};
    hipEvent_t fork = nullptr, join[kMaxBatch] = {};
    static constexpr int kBatchCounters = 64 * kMaxBatch;
};

namespace {

void batch_drop_graphs(dmc_batch *b) {
    for (int i = 0; i < 2; ++i) {
        if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
        if (b->graph[i]) (void)hipGraphDestroy(b->graph[i]);
        b->exec[i] = nullptr;
        b->graph[i] = nullptr;
    }
    b->dirty = true;
}

// The system as the per-system launches of a phase see it: setting 0 keeps the context's
// kernels, setting -1 selects the kernels that stage their tables in LDS (the two-walker
// moves, the end of the sweep) where their workgroups fit
DevSys phase_sys(const dmc_ctx *c, int nup, BatchPhase ph, int setting) {
    DevSys d = c->ds;
    if (setting < 0 && ph == BatchPhase::Moves && d.sinv_f32) d.mv_lds = sm_moves_tab_fits(nup, d);
    if (setting < 0 && ph == BatchPhase::End) d.tab_lds = sm_end_tab_fits(nup, d);
    return d;
}

// One phase of a batch step.  setting = walkers per workgroup of the one persistent launch
// over all systems, or one launch of the single-system kernel per system (grid = its
// capacity; workgroups past the device population return at once): 0 with the context's
// kernels, -1 with the kernels that stage their tables in LDS.
int batch_phase(dmc_batch *b, BatchPhase ph, int chunk, int opcomp, int *counter, int setting) {
    if (setting > 0)
        return launch_batch_phase(b->host, b->args, ph, chunk, opcomp, counter, setting, b->ncu,
                                  b->st);
    for (size_t i = 0; i < b->order.size(); ++i) {
        dmc_ctx *c = b->ctx[b->order[i]];
        const BatchSys &q = b->host.sys[i];
        const DevSys d = phase_sys(c, q.nup, ph, setting);
        int e = 0;
        if (ph == BatchPhase::End) {
            e = launch_sm_end(q.nup, d, c->A, c->B, c->T, opcomp, c->capacity, b->st);
        } else {
            const int m0 = chunk * q.k;
            if (m0 >= q.npart) continue;
            const int m1 = m0 + q.k < q.npart ? m0 + q.k : q.npart;
            if (ph == BatchPhase::Refresh)
                e = launch_sm_refresh(q.nup, d, c->B.x, c->A.sinv, c->T, c->capacity, b->st);
            else
                e = launch_sm_moves(q.nup, d, m0 ? c->B.x : c->A.x, c->B.x,
                                    m0 ? c->B.irn : c->A.irn, c->B.irn, c->A.sinv, m0, m1, c->T,
                                    c->capacity, b->st);
        }
        if (e) return e;
    }
    return hipSuccess;
}

// the launches of one system's step on stream s when every phase is a per-system launch:
// the moves of each chunk (with the FP64 refresh between chunks), the end of the sweep, the
// estimator sums and the device branch
int batch_system_chain(dmc_batch *b, int i, int opcomp, hipStream_t s) {
    dmc_ctx *c = b->ctx[b->order[i]];
    const BatchSys &q = b->host.sys[i];
    const DevSys dm = phase_sys(c, q.nup, BatchPhase::Moves, b->wpg[0]);
    const DevSys de = phase_sys(c, q.nup, BatchPhase::End, b->wpg[2]);
    for (int ch = 0; ch < b->nchunk; ++ch) {
        const int m0 = ch * q.k;
        if (m0 >= q.npart) break;
        const int m1 = m0 + q.k < q.npart ? m0 + q.k : q.npart;
        if (ch > 0)
            LAUNCHCHK(launch_sm_refresh(q.nup, c->ds, c->B.x, c->A.sinv, c->T, c->capacity, s));
        LAUNCHCHK(launch_sm_moves(q.nup, dm, m0 ? c->B.x : c->A.x, c->B.x,
                                  m0 ? c->B.irn : c->A.irn, c->B.irn, c->A.sinv, m0, m1, c->T,
                                  c->capacity, s));
    }
    LAUNCHCHK(launch_sm_end(q.nup, de, c->A, c->B, c->T, opcomp, c->capacity, s));
    LAUNCHCHK(launch_est_reduce(c->T, -1, c->est_acc, s));
    LAUNCHCHK(launch_branch_device(c->ds, c->A, c->B, c->T, c->capacity, s));
    return DMC_OK;
}

// the launches of one batch step on b->st (captured, or run directly by the tuner)
int batch_enqueue(dmc_batch *b, int opcomp) {
    const int ns = int(b->order.size());
    LAUNCHCHK(launch_batch_begin(b->args, b->counters, dmc_batch::kBatchCounters, b->st));
    if (b->wpg[0] <= 0 && b->wpg[2] <= 0 && (b->nchunk == 1 || b->wpg[1] <= 0)) {
        // no fused launch: the systems share nothing, so their chains run side by side
        HIPCHK(hipEventRecord(b->fork, b->st));
        for (int i = 0; i < ns; ++i) {
            HIPCHK(hipStreamWaitEvent(b->side[i], b->fork, 0));
            int e = batch_system_chain(b, i, opcomp, b->side[i]);
            if (e) return e;
            HIPCHK(hipEventRecord(b->join[i], b->side[i]));
            HIPCHK(hipStreamWaitEvent(b->st, b->join[i], 0));
        }
        return DMC_OK;
    }
    int slot = 0;
    for (int c = 0; c < b->nchunk; ++c) {
        if (c > 0)
            LAUNCHCHK(batch_phase(b, BatchPhase::Refresh, c, 0, b->counters + kMaxBatch * slot++,
                                  b->wpg[1]));
        LAUNCHCHK(batch_phase(b, BatchPhase::Moves, c, 0, b->counters + kMaxBatch * slot++,
                              b->wpg[0]));
    }
    LAUNCHCHK(batch_phase(b, BatchPhase::End, -1, opcomp, b->counters + kMaxBatch * slot++,
                          b->wpg[2]));
    for (int i = 0; i < ns; ++i) {
        dmc_ctx *c = b->ctx[b->order[i]];
        LAUNCHCHK(launch_est_reduce(c->T, -1, c->est_acc, b->st));
        LAUNCHCHK(launch_branch_device(c->ds, c->A, c->B, c->T, c->capacity, b->st));
    }
    return DMC_OK;
}

// refresh the system records from the contexts (their A/B may have been swapped by steps
// outside the batch) and capture the two step graphs (opcomp 0 and 1)
int batch_capture(dmc_batch *b) {
    batch_drop_graphs(b);
    const int ns = int(b->order.size());
    b->args.nsys = ns;
    b->captured_a.assign(ns, nullptr);
    for (int i = 0; i < ns; ++i) {
        dmc_ctx *c = b->ctx[b->order[i]];
        BatchSys &q = b->host.sys[i];
        q.S = c->ds;
        q.A = c->A;
        q.B = c->B;
        q.T = c->T;
        b->args.sys[i] = q;
        b->captured_a[i] = reinterpret_cast<char *>(c->A.x);
    }
    for (int oc = 0; oc < 2; ++oc) {
        HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        int e = batch_enqueue(b, oc);
        hipGraph_t g = nullptr;
        hipError_t ee = hipStreamEndCapture(b->st, &g);
        if (e) {
            if (g) (void)hipGraphDestroy(g);
            return e;
        }
        if (ee != hipSuccess) return fail(DMC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
        b->graph[oc] = g;
        HIPCHK(hipGraphInstantiate(&b->exec[oc], g, nullptr, nullptr, 0));
    }
    b->dirty = false;
    return DMC_OK;
}

// host populations into the device words (nt[2]) of every system
int batch_upload_populations(dmc_batch *b) {
    for (dmc_ctx *c : b->ctx) {
        if (c->pending) return fail(DMC_ERR_ARG, "a member context awaits dmc_branch");
        int h[5] = {c->nw, 0, c->nw, 0, 0};
        HIPCHK(hipMemcpyAsync(c->T.nt, h, sizeof(h), hipMemcpyHostToDevice, b->st));
    }
    HIPCHK(hipStreamSynchronize(b->st));
    return DMC_OK;
}

}  // namespace

extern "C" {

int dmc_batch_create(dmc_ctx *const *ctxs, int32_t n, dmc_batch **out) {
    if (!ctxs || !out || n < 1) return fail(DMC_ERR_ARG, "invalid argument");
    *out = nullptr;
    if (n > kMaxBatch) return fail(DMC_ERR_ARG, "a batch holds at most " + std::to_string(kMaxBatch) + " systems");
    for (int i = 0; i < n; ++i) {
        const dmc_ctx *c = ctxs[i];
        if (!c) return fail(DMC_ERR_ARG, "null context");
        if (c->sys.p.movetype != DMC_SINGLEPARTICLE)
            return fail(DMC_ERR_ARG, "batches run single-particle moves");
        if (c->sys.p.precision != ctxs[0]->sys.p.precision)
            return fail(DMC_ERR_ARG, "the systems of a batch share one precision");
        if (c->nw <= 0) return fail(DMC_ERR_ARG, "every system needs walkers");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == c) return fail(DMC_ERR_ARG, "a context appears twice");
    }
    dmc_batch *b = new dmc_batch();
    b->ctx.assign(ctxs, ctxs + n);
    b->order.resize(n);
    for (int i = 0; i < n; ++i) b->order[i] = i;
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
        return b->ctx[x]->sys.npart > b->ctx[y]->sys.npart;
    });
    b->host.fp32 = ctxs[0]->sys.p.precision == DMC_MIXED;
    b->host.sys.resize(n);
    b->nchunk = 1;
    for (int i = 0; i < n; ++i) {
        const dmc_ctx *c = b->ctx[b->order[i]];
        BatchSys &q = b->host.sys[i];
        q.nup = c->sys.n;
        q.npart = c->sys.npart;
        q.k = c->sys.p.krefresh > 0 ? c->sys.p.krefresh : c->sys.npart;
        q.capacity = c->capacity;
        q.S = c->ds;
        const int nc = (q.npart + q.k - 1) / q.k;
        b->nchunk = nc > b->nchunk ? nc : b->nchunk;
    }
    if (2 * b->nchunk + 1 > 64) {
        delete b;
        return fail(DMC_ERR_ARG, "too many refresh chunks for a batch");
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipStreamCreate(&b->st) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&b->counters), sizeof(int) * dmc_batch::kBatchCounters) != hipSuccess ||
        hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
        dmc_batch_destroy(b);
        return fail(DMC_ERR_HIP, "batch set-up failed");
    }
    for (size_t i = 0; i < b->ctx.size(); ++i)
        if (hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) != hipSuccess) {
            dmc_batch_destroy(b);
            return fail(DMC_ERR_HIP, "batch set-up failed");
        }
    for (dmc_ctx *c : b->ctx) (void)hipStreamSynchronize(c->st);
    int e = batch_upload_populations(b);
    if (e) {
        std::string m = g_err;
        dmc_batch_destroy(b);
        return fail(e, m);
    }
    *out = b;
    return DMC_OK;
}

int dmc_batch_destroy(dmc_batch *b) {
    if (!b) return DMC_OK;
    if (b->st) (void)hipStreamSynchronize(b->st);
    batch_drop_graphs(b);
    if (b->counters) (void)hipFree(b->counters);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    for (int i = 0; i < kMaxBatch; ++i) {
        if (b->side[i]) {
            (void)hipStreamSynchronize(b->side[i]);
            (void)hipStreamDestroy(b->side[i]);
        }
        if (b->join[i]) (void)hipEventDestroy(b->join[i]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
    return DMC_OK;
}

int dmc_batch_set_wpg(dmc_batch *b, int32_t moves, int32_t refresh, int32_t end) {
    if (!b) return fail(DMC_ERR_ARG, "null batch");
    const int32_t w[3] = {moves, refresh, end};
    const BatchPhase ph[3] = {BatchPhase::Moves, BatchPhase::Refresh, BatchPhase::End};
    for (int i = 0; i < 3; ++i) {
        if (w[i] != -1 && w[i] != 0 && w[i] != 1 && w[i] != 2 && w[i] != 4 && w[i] != 8)
            return fail(DMC_ERR_ARG, "walkers per workgroup must be 1, 2, 4 or 8 (0: per "
                                     "system, -1: per system with LDS-staged tables)");
        if (w[i] == -1 && ph[i] == BatchPhase::Refresh)
            return fail(DMC_ERR_ARG, "the refresh has no LDS-table kernel (-1)");
        if (w[i] > 0 && batch_blocks_per_cu(b->host, ph[i], w[i]) <= 0)
            return fail(DMC_ERR_ARG, "walkers per workgroup " + std::to_string(w[i]) +
                                         " does not fit the LDS or registers of a CU");
    }
    for (int i = 0; i < 3; ++i) b->wpg[i] = w[i];
    b->dirty = true;
    return DMC_OK;
}

int dmc_batch_step(dmc_batch *b, int32_t opcomp) {
    if (!b) return fail(DMC_ERR_ARG, "null batch");
    if (b->host_current) {
        int e = batch_upload_populations(b);
        if (e) return e;
        b->host_current = false;
    }
    for (size_t i = 0; i < b->order.size() && !b->dirty; ++i)
        if (reinterpret_cast<char *>(b->ctx[b->order[i]]->A.x) != b->captured_a[i]) b->dirty = true;
    if (b->dirty) {
        int e = batch_capture(b);
        if (e) return e;
    }
    HIPCHK(hipGraphLaunch(b->exec[opcomp ? 1 : 0], b->st));
    return DMC_OK;
}

int dmc_batch_run(dmc_batch *b, int32_t nsteps) {
    for (int s = 0; s < nsteps; ++s) {
        int e = dmc_batch_step(b, s == nsteps - 1);
        if (e) return e;
    }
    return DMC_OK;
}

int dmc_batch_sync(dmc_batch *b, int32_t *nt) {
    if (!b) return fail(DMC_ERR_ARG, "null batch");
    HIPCHK(hipStreamSynchronize(b->st));
    const int n = int(b->ctx.size());
    int err = DMC_OK;
    std::string msg;
    for (int i = 0; i < n; ++i) {
        dmc_ctx *c = b->ctx[i];
        int h[5];
        HIPCHK(hipMemcpy(h, c->T.nt, sizeof(h), hipMemcpyDeviceToHost));
        c->nw = h[2];
        if (nt) nt[i] = h[2];
        if (h[3] && !err) {
            err = DMC_ERR_OVERFLOW;
            msg = "stack overflow: a population exceeded capacity " + std::to_string(c->capacity);
        } else if (h[2] == 0 && !err) {
            err = DMC_ERR_ARG;
            msg = "population died out";
        }
    }
    b->host_current = true;
    if (err) return fail(err, msg);
    return DMC_OK;
}

int dmc_batch_tune(dmc_batch *b, int32_t reps, int32_t *wpg_out, double *ms_out) {
    if (!b || reps < 1) return fail(DMC_ERR_ARG, "invalid argument");
    int e = DMC_OK;
    if (b->host_current) {
        e = batch_upload_populations(b);
        if (e) return e;
        b->host_current = false;
    }
    HIPCHK(hipStreamSynchronize(b->st));
    // the tuning launches leave the population alone (they read A and write B and scratch);
    // what they do touch is saved and restored: acceptance counters, the histogram and, for
    // the refresh phase, the working inverse
    const int ns = int(b->order.size());
    std::vector<std::vector<char>> save(ns * 3);
    auto snap = [&](bool restore) -> int {
        for (int i = 0; i < ns; ++i) {
            dmc_ctx *c = b->ctx[b->order[i]];
            struct { void *p; size_t n; } r[3] = {
                {c->T.moves, 2 * sizeof(int64_t)}, {c->T.hist, 500 * sizeof(int64_t)},
                {c->A.sinv, size_t(c->nw) * sinv_stride(c->ds)}};
            for (int k = 0; k < 3; ++k) {
                auto &v = save[3 * i + k];
                if (!restore) v.resize(r[k].n);
                HIPCHK(hipMemcpy(restore ? r[k].p : v.data(), restore ? v.data() : r[k].p, r[k].n,
                                 restore ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
            }
        }
        return DMC_OK;
    };
    e = snap(false);
    if (e) return e;
    // the args of the current populations (no capture needed for direct launches)
    b->args.nsys = ns;
    for (int i = 0; i < ns; ++i) {
        dmc_ctx *c = b->ctx[b->order[i]];
        b->host.sys[i].S = c->ds;
        b->host.sys[i].A = c->A;
        b->host.sys[i].B = c->B;
        b->host.sys[i].T = c->T;
        b->args.sys[i] = b->host.sys[i];
    }
    const BatchPhase ph[3] = {BatchPhase::Moves, BatchPhase::Refresh, BatchPhase::End};
    constexpr int kCand = 6;
    const int cand[kCand] = {1, 2, 4, 8, 0, -1};
    for (int p = 0; p < 3; ++p) {
        if (p == 1 && b->nchunk < 2) {
            if (ms_out) for (int k = 0; k < kCand; ++k) ms_out[kCand * p + k] = -1.0;
            continue;
        }
        double best = 1e300;
        int bw = b->wpg[p];
        for (int k = 0; k < kCand; ++k) {
            double t = -1.0;
            // the LDS-table setting exists for the moves and the end of the sweep
            const bool run = cand[k] == 0 || (cand[k] < 0 ? ph[p] != BatchPhase::Refresh
                                                          : batch_blocks_per_cu(b->host, ph[p],
                                                                                cand[k]) > 0);
            if (run) {
                const int chunk = p == 2 ? -1 : (p == 1 ? 1 : 0);
                float ms = 0.f;
                for (int r = 0; r <= reps && !e; ++r) {  // the first launch warms up
                    if (r == 1) e = hipEventRecord(b->e0, b->st) ? DMC_ERR_HIP : 0;
                    if (!e) e = launch_batch_begin(b->args, b->counters, dmc_batch::kBatchCounters, b->st) ? DMC_ERR_HIP : 0;
                    if (!e) e = batch_phase(b, ph[p], chunk, 0, b->counters, cand[k]) ? DMC_ERR_HIP : 0;
                }
                if (!e) e = hipEventRecord(b->e1, b->st) ? DMC_ERR_HIP : 0;
                if (!e) e = hipEventSynchronize(b->e1) ? DMC_ERR_HIP : 0;
                if (!e) e = hipEventElapsedTime(&ms, b->e0, b->e1) ? DMC_ERR_HIP : 0;
                if (e) return fail(DMC_ERR_HIP, "tuning launch failed");
                t = ms / reps;
                if (t < best) {
                    best = t;
                    bw = cand[k];
                }
            }
            if (ms_out) ms_out[kCand * p + k] = t;
        }
        b->wpg[p] = bw;
    }
    e = snap(true);
    if (e) return e;
    b->dirty = true;
    if (wpg_out)
        for (int p = 0; p < 3; ++p) wpg_out[p] = b->wpg[p];
    return DMC_OK;
}

}  // extern "C"
