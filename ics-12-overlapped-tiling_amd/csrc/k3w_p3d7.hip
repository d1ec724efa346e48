// k3w_p3d7.hip — instantiations of the wave-pipelined 3D kernel for OT_P3D7
// (own translation unit so the build compiles them in parallel).
#include "k3w_impl.hpp"

namespace ot {
ot_status launch_3dw_p3d7(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  return k3w::launch_st<OT_P3D7>(t_eff, rows, g, f, stream, prepare);
}
}  // namespace ot
