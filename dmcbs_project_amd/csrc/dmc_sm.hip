// Single-particle-move DMC (placeholder until the kernels land)
#include "dmc_kernels.h"
namespace dmc {
int launch_sm_step(int, int, const DevSys &, const DevState &, const DevState &, const DevStep &,
                   int, int, hipStream_t) {
    return hipErrorNotSupported;
}
int launch_adjoint(int, const DevSys &, const double *, int, const double *, double *, double *,
                   hipStream_t) {
    return hipErrorNotSupported;
}
}  // namespace dmc
