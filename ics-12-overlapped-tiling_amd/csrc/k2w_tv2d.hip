// k2w_tv2d.hip — wave-per-tile 2D kernel instantiations for OT_TV2D (k2w_impl.hpp).
#include "k2w_impl.hpp"

namespace ot {
ot_status launch_2dw_tv2d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  return k2w::launch_st<OT_TV2D>(t_eff, rows, g, f, stream, prepare);
}
BatchKernel batch_wave_kernel_tv2d(int t_eff, int rows) { return k2w::pick_batch<OT_TV2D>(t_eff, rows); }
}  // namespace ot
