// spray_fp64.hip -- FP64 form of the droplet update (reference method in
// double precision).  Compiled with -ffp-contract=off so that every a*b+c
// rounds twice, as in the CPU oracle; fma only where pm_math.h calls it.
#include <algorithm>

#include "update_launch.h"

namespace pelemp {
int launch_update_fp64(Ctx& c, const PassParams& p, hipStream_t s) {
  return dispatch_update<double, true>(c, p, s);
}
int launch_pelec_convert(Ctx& c, const pelemp_fab_view& st, hipStream_t s) {
  switch (c.mech) {
    case PELEMP_MECH_DODECANE_AIR: return launch_pelec_convert_t<Mech_dodecane_air>(c, st, s);
    case PELEMP_MECH_DECANE_DODECANE_AIR: return launch_pelec_convert_t<Mech_decane_dodecane_air>(c, st, s);
    case PELEMP_MECH_ALKANE6_AIR: return launch_pelec_convert_t<Mech_alkane6_air>(c, st, s);
    case PELEMP_MECH_DECANE_AIR: return launch_pelec_convert_t<Mech_decane_air>(c, st, s);
    default: set_last_error("unsupported mechanism"); return PELEMP_E_BAD_CONFIG;
  }
}
}  // namespace pelemp
