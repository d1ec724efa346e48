// k3d_p3d7.hip — instantiations of the z-streaming 3D kernel for OT_P3D7
// (kept in their own translation unit so the build compiles them in parallel).
#include "k3d_impl.hpp"

namespace ot {
ot_status launch_3d_p3d7(int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  return k3::launch_st<OT_P3D7>(t_eff, ept, g, f, stream, prepare);
}
}  // namespace ot
