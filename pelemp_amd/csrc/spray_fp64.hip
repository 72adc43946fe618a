// spray_fp64.hip -- FP64 form of the droplet update (reference method in
// double precision).  Compiled with -ffp-contract=off so that every a*b+c
// rounds twice, as in the CPU oracle; fma only where pm_math.h calls it.
#include "update_launch.h"

namespace pelemp {
int launch_update_fp64(Ctx& c, const PassParams& p, hipStream_t s) {
  return dispatch_update<double, true>(c, p, s);
}
}  // namespace pelemp
