// k2s_rician2d.hip — row-streaming 2D kernel instantiations for OT_RICIAN2D (k2s_impl.hpp).
#include "k2s_impl.hpp"

namespace ot {
ot_status launch_2ds_rician2d(int t_eff, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  return k2s::launch_st<OT_RICIAN2D>(t_eff, g, f, stream, prepare);
}
}  // namespace ot
