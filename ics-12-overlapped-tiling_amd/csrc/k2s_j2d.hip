// k2s_j2d.hip — row-streaming 2D kernel instantiations for OT_J2D (k2s_impl.hpp).
#include "k2s_impl.hpp"

namespace ot {
ot_status launch_2ds_j2d(int t_eff, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  return k2s::launch_st<OT_J2D>(t_eff, g, f, stream, prepare);
}
}  // namespace ot
