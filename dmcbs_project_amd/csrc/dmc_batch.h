// Several systems stepped together by persistent launches (dmc_batch.hip).
#pragma once
#include <vector>

#include "dmc_kernels.h"

namespace dmc {

// one resident system of a batch, as the persistent kernels read it (device memory)
struct BatchSys {
    DevSys S;
    DevState A, B;
    DevStep T;
    int nup, npart, k;  // k: moves between FP64 refreshes
    int capacity;
};

constexpr int kMaxBatch = 4;

// the systems of a batch, passed by value as a kernel argument (read through the scalar
// cache)
struct BatchArgs {
    BatchSys sys[kMaxBatch];
    int nsys;
};

// host view of a batch: the systems in launch order (largest first) and the precision
struct BatchHost {
    std::vector<BatchSys> sys;
    bool fp32 = true;
};

enum class BatchPhase { Moves, Refresh, End };

size_t batch_wave_lds(const BatchHost &h, BatchPhase ph);
int batch_blocks_per_cu(const BatchHost &h, BatchPhase ph, int wpg);
int launch_batch_begin(const BatchArgs &a, int *counters, int ncounters, hipStream_t st);
int launch_batch_phase(const BatchHost &h, const BatchArgs &a, BatchPhase ph, int chunk,
                       int opcomp, int *counter, int wpg, int ncu, hipStream_t st);

}  // namespace dmc
