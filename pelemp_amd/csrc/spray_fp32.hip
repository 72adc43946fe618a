// spray_fp32.hip -- FP32 form of the droplet update: droplet state and the
// per-droplet ODE in single precision (hardware elementary functions,
// contraction allowed), positions stored inside the droplet's cell, source
// sums in double.
#include "update_launch.h"

namespace pelemp {
int launch_update_fp32(Ctx& c, const PassParams& p, hipStream_t s) {
  return dispatch_update<float, false>(c, p, s);
}
int launch_update_surr(Ctx& c, const PassParams& p, hipStream_t s) {
  return dispatch_update<float, false, true>(c, p, s);
}
}  // namespace pelemp
