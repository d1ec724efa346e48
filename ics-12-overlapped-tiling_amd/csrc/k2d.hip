// k2d.hip — per-grid launches of the 2D overlapped time-tiled kernels:
// one workgroup per output tile (tile update in k2d_impl.hpp).
#include "k2d_impl.hpp"

namespace ot {
namespace {

template <int ST, int E>
__global__ void __launch_bounds__(1024) k2d(LaunchGeom g, FieldPtrs f, int t_eff) {
  extern __shared__ float lds[];
  const int nblk = g.grid[0] * g.grid[1];
  const int lid = dev::xcd_remap((int)blockIdx.x, nblk);
  tile2d<ST, E>(g, f, t_eff, lid % g.grid[0], lid / g.grid[0], lds);
}

template <int ST, int E>
ot_status launch_t(const LaunchGeom& g, const FieldPtrs& f, int t_eff, void* stream, bool prepare) {
  int bx, by, lds;
  shape_2d(ST, t_eff, E, g.tile, &bx, &by, &lds);
  if (prepare) return prepare_lds((const void*)&k2d<ST, E>, lds);
  const int nblk = g.grid[0] * g.grid[1];
  hipLaunchKernelGGL((k2d<ST, E>), dim3(nblk), dim3(bx, by), lds, (hipStream_t)stream, g, f, t_eff);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return set_error(OT_EHIP, "2D kernel launch failed: %s", hipGetErrorString(err));
  return OT_OK;
}

template <int ST>
ot_status launch_e(int ept, const LaunchGeom& g, const FieldPtrs& f, int t_eff, void* stream, bool prepare) {
  switch (ept) {
    case 1: return launch_t<ST, 1>(g, f, t_eff, stream, prepare);
    case 2: return launch_t<ST, 2>(g, f, t_eff, stream, prepare);
    case 3: return launch_t<ST, 3>(g, f, t_eff, stream, prepare);
    case 4: return launch_t<ST, 4>(g, f, t_eff, stream, prepare);
    case 5: return launch_t<ST, 5>(g, f, t_eff, stream, prepare);
    case 6: return launch_t<ST, 6>(g, f, t_eff, stream, prepare);
    case 7: return launch_t<ST, 7>(g, f, t_eff, stream, prepare);
    case 8: return launch_t<ST, 8>(g, f, t_eff, stream, prepare);
    case 9: return launch_t<ST, 9>(g, f, t_eff, stream, prepare);
    case 10: return launch_t<ST, 10>(g, f, t_eff, stream, prepare);
    case 11: return launch_t<ST, 11>(g, f, t_eff, stream, prepare);
    case 12: return launch_t<ST, 12>(g, f, t_eff, stream, prepare);
    default: return set_error(OT_EUNSUPPORTED, "2D kernels are compiled for 1..12 elements per thread (got %d)", ept);
  }
}

}  // namespace

void shape_2d(int stencil, int T, int ept, const int32_t tile[3], int* bx, int* by, int* lds) {
  const int R = (stencil == OT_RICIAN2D) ? 2 : 1;
  const int planes = (stencil == OT_FDTD2D || stencil == OT_RICIAN2D) ? 3 : 2;
  const int H = R * T;
  const int FX = tile[0] + 2 * H, FY = tile[1] + 2 * H;
  *bx = FX;
  *by = (FY + ept - 1) / ept;
  *lds = planes * (*by) * ept * FX * (int)sizeof(float);
}

ot_status launch_2d(int stencil, int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f,
                    void* stream, bool prepare) {
  if (g.variant & kProbeBits) return launch_2d_probe(stencil, t_eff, ept, g, f, stream, prepare);
  if (g.wave_rows < 0) {
    switch (stencil) {
      case OT_J2D: return launch_2ds_j2d(t_eff, g, f, stream, prepare);
      case OT_P2D: return launch_2ds_p2d(t_eff, g, f, stream, prepare);
      case OT_G2D: return launch_2ds_g2d(t_eff, g, f, stream, prepare);
      case OT_TV2D: return launch_2ds_tv2d(t_eff, g, f, stream, prepare);
      case OT_RICIAN2D: return launch_2ds_rician2d(t_eff, g, f, stream, prepare);
      case OT_FDTD2D: return launch_2ds_fdtd2d(t_eff, g, f, stream, prepare);
      default: return set_error(OT_EINVAL, "stencil %d is not a 2D stencil", stencil);
    }
  }
  if (g.wave_rows > 0) {
    switch (stencil) {
      case OT_J2D: return launch_2dw_j2d(t_eff, g.wave_rows, g, f, stream, prepare);
      case OT_P2D: return launch_2dw_p2d(t_eff, g.wave_rows, g, f, stream, prepare);
      case OT_G2D: return launch_2dw_g2d(t_eff, g.wave_rows, g, f, stream, prepare);
      case OT_TV2D: return launch_2dw_tv2d(t_eff, g.wave_rows, g, f, stream, prepare);
      case OT_RICIAN2D: return launch_2dw_rician2d(t_eff, g.wave_rows, g, f, stream, prepare);
      case OT_FDTD2D: return launch_2dw_fdtd2d(t_eff, g.wave_rows, g, f, stream, prepare);
      default: return set_error(OT_EINVAL, "stencil %d is not a 2D stencil", stencil);
    }
  }
  switch (stencil) {
    case OT_J2D: return launch_e<OT_J2D>(ept, g, f, t_eff, stream, prepare);
    case OT_P2D: return launch_e<OT_P2D>(ept, g, f, t_eff, stream, prepare);
    case OT_G2D: return launch_e<OT_G2D>(ept, g, f, t_eff, stream, prepare);
    case OT_TV2D: return launch_e<OT_TV2D>(ept, g, f, t_eff, stream, prepare);
    case OT_RICIAN2D: return launch_e<OT_RICIAN2D>(ept, g, f, t_eff, stream, prepare);
    case OT_FDTD2D: return launch_e<OT_FDTD2D>(ept, g, f, t_eff, stream, prepare);
    default: return set_error(OT_EINVAL, "stencil %d is not a 2D stencil", stencil);
  }
}

}  // namespace ot
