// k2p.hip — timing probes of the 2D LDS kernel (jacobi-2d, gradient-2d).
//
// The reference's micro-benchmark modes of its 2D generators: --time-only
// gmem / instr / smem (jacobi-2d.cpp:258-266, :296-341, :368-433, :462-481;
// gradient-2d.cpp:225-290) and the per-warp %clock stamps of --dump-clocks
// (jacobi-2d.cpp:202-205, :453-509, :957-997). Here they are separate kernel
// instantiations of the same tile update (tile2d with a probe mode), so the
// production kernels carry none of their code. The stamps are s_memtime (the
// shader clock, as %clock) at the phase boundaries, s_memrealtime (constant
// 100 MHz) at entry and exit, and the HW_ID / XCC_ID registers that name the
// wave's SIMD, CU, shader engine and die (the reference keeps %smid). Lane 0
// of every wave writes its record with an ordinary vector store.
#include "k2d_impl.hpp"

namespace ot {
namespace {

template <int ST, int E, int PM>
__global__ void __launch_bounds__(1024) k2p(LaunchGeom g, FieldPtrs f, int t_eff) {
  extern __shared__ float lds[];
  uint64_t clk[5];
  uint64_t rt0 = 0;
  if (g.clocks) {
    clk[0] = __builtin_amdgcn_s_memtime();
    rt0 = __builtin_amdgcn_s_memrealtime();
  }
  const int nblk = g.grid[0] * g.grid[1];
  const int lid = dev::xcd_remap((int)blockIdx.x, nblk);
  const int bxi = lid % g.grid[0], byi = lid / g.grid[0];
  tile2d<ST, E, PM>(g, f, t_eff, bxi, byi, lds, clk);
  if (g.clocks) {
    clk[4] = __builtin_amdgcn_s_memtime();
    const uint64_t rt1 = __builtin_amdgcn_s_memrealtime();
    uint32_t hw, xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)" : "=s"(hw));
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    if ((tid & 63) == 0) {
      // a fixed slot per (workgroup, wave): a shortened last launch has
      // smaller workgroups than the launches before it
      ot_clock_record& r = g.clocks[(int64_t)lid * kClockSlotsPerBlock + (tid >> 6)];
      r.block_x = bxi;
      r.block_y = byi;
      r.wave = tid >> 6;
      r.hw_id = hw;
      r.xcc_id = xcc;
      r.valid = 1;
      for (int k = 0; k < 5; ++k) r.clock[k] = clk[k];
      r.realtime[0] = rt0;
      r.realtime[1] = rt1;
    }
  }
}

template <int ST, int E, int PM>
ot_status launch_t(const LaunchGeom& g, const FieldPtrs& f, int t_eff, void* stream, bool prepare) {
  int bx, by, lds;
  shape_2d(ST, t_eff, E, g.tile, &bx, &by, &lds);
  if (prepare) return prepare_lds((const void*)&k2p<ST, E, PM>, lds);
  const int nblk = g.grid[0] * g.grid[1];
  hipLaunchKernelGGL((k2p<ST, E, PM>), dim3(nblk), dim3(bx, by), lds, (hipStream_t)stream, g, f, t_eff);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return set_error(OT_EHIP, "2D probe kernel launch failed: %s", hipGetErrorString(err));
  return OT_OK;
}

template <int ST, int PM>
ot_status launch_e(int ept, const LaunchGeom& g, const FieldPtrs& f, int t_eff, void* stream, bool prepare) {
  switch (ept) {
    case 1: return launch_t<ST, 1, PM>(g, f, t_eff, stream, prepare);
    case 2: return launch_t<ST, 2, PM>(g, f, t_eff, stream, prepare);
    case 3: return launch_t<ST, 3, PM>(g, f, t_eff, stream, prepare);
    case 4: return launch_t<ST, 4, PM>(g, f, t_eff, stream, prepare);
    case 5: return launch_t<ST, 5, PM>(g, f, t_eff, stream, prepare);
    case 6: return launch_t<ST, 6, PM>(g, f, t_eff, stream, prepare);
    case 7: return launch_t<ST, 7, PM>(g, f, t_eff, stream, prepare);
    case 8: return launch_t<ST, 8, PM>(g, f, t_eff, stream, prepare);
    default:
      return set_error(OT_EUNSUPPORTED, "the 2D timing probes are compiled for 1..8 elements per thread (got %d)",
                       ept);
  }
}

template <int ST>
ot_status launch_mode(int ept, const LaunchGeom& g, const FieldPtrs& f, int t_eff, void* stream, bool prepare) {
  const int v = g.variant;
  if (v & OT_VARIANT_TIME_ONLY_GMEM) return launch_e<ST, kProbeGmem>(ept, g, f, t_eff, stream, prepare);
  if (v & OT_VARIANT_TIME_ONLY_INSTR) return launch_e<ST, kProbeInstr>(ept, g, f, t_eff, stream, prepare);
  if (v & OT_VARIANT_TIME_ONLY_SMEM) return launch_e<ST, kProbeSmem>(ept, g, f, t_eff, stream, prepare);
  return launch_e<ST, kProbeClocks>(ept, g, f, t_eff, stream, prepare);
}

}  // namespace

ot_status launch_2d_probe(int stencil, int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f,
                          void* stream, bool prepare) {
  const int modes = g.variant & (OT_VARIANT_TIME_ONLY_GMEM | OT_VARIANT_TIME_ONLY_INSTR | OT_VARIANT_TIME_ONLY_SMEM);
  if (modes & (modes - 1)) return set_error(OT_EINVAL, "choose one --time-only mode");
  if (g.wave_rows != 0)
    return set_error(OT_EUNSUPPORTED, "the timing probes run the 2D LDS kernel; this tile selects a wave kernel");
  switch (stencil) {
    case OT_J2D: return launch_mode<OT_J2D>(ept, g, f, t_eff, stream, prepare);
    case OT_G2D: return launch_mode<OT_G2D>(ept, g, f, t_eff, stream, prepare);
    default:
      return set_error(OT_EUNSUPPORTED,
                       "--time-only and --dump-clocks exist for jacobi-2d and gradient-2d (as in the reference)");
  }
}

}  // namespace ot
