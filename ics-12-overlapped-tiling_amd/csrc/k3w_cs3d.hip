// k3w_cs3d.hip — instantiations of the wave-pipelined 3D kernel for OT_CS3D
// (own translation unit so the build compiles them in parallel). The 13-point
// cs-3d update is built only on this kernel family: its diagonal taps are
// neighbouring rows and DPP shifts of the three register planes of a level.
#include "k3w_impl.hpp"

namespace ot {
ot_status launch_3dw_cs3d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  return k3w::launch_st<OT_CS3D>(t_eff, rows, g, f, stream, prepare);
}
}  // namespace ot
