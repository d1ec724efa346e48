// k3d_g3d.hip — instantiations of the z-streaming 3D kernel for OT_G3D
// (kept in their own translation unit so the build compiles them in parallel).
#include "k3d_impl.hpp"

namespace ot {
ot_status launch_3d_g3d(int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  return k3::launch_st<OT_G3D>(t_eff, ept, g, f, stream, prepare);
}
}  // namespace ot
