// Device-side data structures shared by the kernels and the C-ABI implementation.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dmc {

// One system as seen by the kernels (passed by value).
struct DevSys {
    int npart, n, nq, ntab;
    int nobeta, cutdrift, fn, nwalk;
    int keep_sinv, sinv_f32;  // walkers carry their inverse (single-particle moves), in FP32?
    double el, eli, range, scale, rangeb, scaleb;
    double alpha0, sigma, sig2, sqrtm, invsqrtm, invm, invmu, hbar, dt, etrial, wtcut, dcut,
        nfac, twopi_l, popup, popdn;
    const double4 *jt;    // Jastrow/potential records {v, u, du, d2u}, [ntab+1]
    // the potential alone, [ntab+1], and the first interval whose 4-point stencil
    // (jfar - 1 .. jfar + 2 and beyond) holds u = u' = u'' = 0 exactly: pairs past the
    // Jastrow's healing distance read 8 bytes per node instead of the 32-byte record
    const double *vt1;
    int jfar;
    // the end-of-sweep kernel stages the radial tables in the LDS of workgroups of several
    // walkers (1), or reads them from global memory with one walker per workgroup (0)
    int tab_lds;
    // the two-walker FP32 move kernel stages its pair records in the LDS of workgroups of
    // several waves (1), or reads them from global memory, one wave per workgroup (0)
    int mv_lds;
    const double4 *bt;    // beta records {b, db, lap b, 0}, [ntab+1]
    // the same beta values in compact arrays, so that a pair loop touches only the fields it
    // reads (a smaller working set in the vector L1): b, [ntab+1]; {db, lap b}, [ntab+1]
    const double *bt1;
    const double2 *bt2;
    const double4 *mt64;  // pair records of the single-particle moves {b, db, u, du}, [ntab+1]
    const float4 *mt32;   // and their FP32 copy (mixed-precision moves)
    const double4 *btb;   // d/db and d/dd of the beta records (parameter adjoint)
    const double4 *btd;
    const int4 *qm;       // integer q vectors (mx, my, mz, shell), [nq]
    const double *wq;     // 2 alpha_shell(q), [nq]
    const uint64_t *ja1;  // ran1 jump-ahead: state after t+1 draws = ja1[t]*s + jc1[t]
    const uint64_t *jc1;
    const uint64_t *ja2;  // ran2 jump-ahead
    const uint64_t *jc2;
};

// Structure-of-arrays walker population.  Positions and gradients are walker-major,
// [w][npart][3].
struct DevState {
    double *x, *dpsi;
    char *sinv;  // inverse of the pairing matrix, dense [j][i], FP64 or FP32 (keep_sinv)
    double *psil, *d2psi, *v, *weight, *wt0;
    int *is;
    uint64_t *irn;
};

// Per-step scratch.
struct DevStep {
    int *count;           // per-walker multiplicity from branch
    int *offset;          // exclusive scan of count
    int *map;             // destination -> source walker
    uint64_t *copy_irn;   // destination RNG state
    double *est;          // per-walker estimator contributions [w][14][3]
    double *estp;         // block partials of the estimator sums [64][42]
    int *nt;              // device population counters: nt[0] current, nt[1] new
    int64_t *hist;        // branch multiplicity histogram [500]
    int *nacc;            // accepted single-particle moves per walker
    int64_t *moves;       // accepted and proposed single-particle moves (running totals)
    // record of every single-particle move, [w][npart][kTraceRec], or null (not recorded)
    double *trace;
};

// fields of a move record: determinant ratio, Jastrow log-ratio, accepted (0/1), the
// acceptance draw, the proposed position (3), the drift after the cut (3), the log of the
// Metropolis ratio (NaN for a move the node test refuses), 0
constexpr int kTraceRec = 12;

}  // namespace dmc
