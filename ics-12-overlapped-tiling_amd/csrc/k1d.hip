// k1d.hip — overlapped time-tiled kernels for the 1D Jacobi stencils.
//
// Replaces the kernel text of jacobi-1d.cpp:111-355, jacobi-1d-5pt.cpp:88-258
// and jacobi-1d-7pt.cpp. As in the reference, element k of a thread sits at
// footprint position tid + k*blockDim (jacobi-1d.cpp:201-205), so global loads,
// stores and LDS accesses are unit-stride across the wavefront. The footprint is
// tile + 2H (H = radius*T); out-of-domain points are 0 after every step.
// Unlike the reference's 5/7-point kernel text (jacobi-1d-5pt.cpp:199-206,
// jacobi-1d-7pt.cpp:238-242), every tap is loaded, and the time loop runs
// ceil(steps/T) launches with the last one shortened, so the result matches the
// CPU loop's `steps` iterations (jacobi-1d-5pt.cpp:505-515).
#include "internal.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace ot {
namespace {

using namespace dev;

template <int ST>
struct K1 {
  static constexpr int R = ST == OT_J1D3 ? 1 : (ST == OT_J1D5 ? 2 : 3);
};

template <int ST, int E>
__global__ void __launch_bounds__(1024) k1d(LaunchGeom g, FieldPtrs f, int t_eff) {
  extern __shared__ float lds[];
  constexpr int R = K1<ST>::R;
  const int H = R * t_eff;
  const int FX = g.tile[0] + 2 * H;
  const int BX = (int)blockDim.x;
  const int plane = BX * E;
  const int lid = xcd_remap((int)blockIdx.x, g.grid[0]);
  const int x0 = lid * g.tile[0] - H;
  const int tid = threadIdx.x;

  bool inside[E];
  int64_t gidx[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int gx = x0 + tid + k * BX;
    inside[k] = gx >= 0 && gx < g.n[0];
    gidx[k] = g.origin + gx;
  }
  if (g.phase_limit == 1) return;

  float u[E];
#pragma unroll
  for (int k = 0; k < E; ++k) u[k] = (g.phase_limit != 3 && inside[k]) ? f.src[0][gidx[k]] : 0.0f;
#pragma unroll
  for (int k = 0; k < E; ++k) lds[tid + k * BX] = u[k];
  __syncthreads();

  int cur = 0;
  const int t_begin = (g.phase_limit == 3) ? 1 : 0;
  const int t_end = (g.phase_limit == 2) ? 1 : t_eff;
  for (int t = t_begin; t < t_end; ++t) {
    const float* S = lds + cur * plane;
    float nu[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int p = tid + k * BX;
      nu[k] = u[k];
      if (p >= R && p <= FX - 1 - R) {
        float v;
        if constexpr (ST == OT_J1D3) v = j1d3(S[p - 1], u[k], S[p + 1]);
        else if constexpr (ST == OT_J1D5) v = j1d5(S[p - 2], S[p - 1], u[k], S[p + 1], S[p + 2]);
        else v = j1d7(S[p - 3], S[p - 2], S[p - 1], u[k], S[p + 1], S[p + 2], S[p + 3]);
        nu[k] = inside[k] ? v : 0.0f;
      }
    }
#pragma unroll
    for (int k = 0; k < E; ++k) u[k] = nu[k];
    if (t != t_end - 1) {
      float* D = lds + (cur ^ 1) * plane;
#pragma unroll
      for (int k = 0; k < E; ++k) D[tid + k * BX] = nu[k];
      __syncthreads();
      cur ^= 1;
    }
  }
  if (g.phase_limit != 0) return;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int p = tid + k * BX;
    if (p >= H && p < H + g.tile[0] && inside[k]) f.dst[0][gidx[k]] = u[k];
  }
}

template <int ST, int E>
ot_status launch_t(const LaunchGeom& g, const FieldPtrs& f, int t_eff, void* stream, bool prepare) {
  int threads, lds;
  shape_1d(ST, t_eff, E, g.tile, &threads, &lds);
  if (prepare) return prepare_lds((const void*)&k1d<ST, E>, lds);
  hipLaunchKernelGGL((k1d<ST, E>), dim3(g.grid[0]), dim3(threads), lds, (hipStream_t)stream, g, f, t_eff);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return set_error(OT_EHIP, "1D kernel launch failed: %s", hipGetErrorString(err));
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
    default: return set_error(OT_EUNSUPPORTED, "1D kernels are compiled for 1..8 elements per thread (got %d)", ept);
  }
}

}  // namespace

void shape_1d(int stencil, int T, int ept, const int32_t tile[3], int* threads, int* lds) {
  const int R = stencil == OT_J1D3 ? 1 : (stencil == OT_J1D5 ? 2 : 3);
  const int FX = tile[0] + 2 * R * T;
  *threads = (FX + ept - 1) / ept;
  *lds = 2 * (*threads) * ept * (int)sizeof(float);
}

ot_status launch_1d(int stencil, int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f,
                    void* stream, bool prepare) {
  switch (stencil) {
    case OT_J1D3: return launch_e<OT_J1D3>(ept, g, f, t_eff, stream, prepare);
    case OT_J1D5: return launch_e<OT_J1D5>(ept, g, f, t_eff, stream, prepare);
    case OT_J1D7: return launch_e<OT_J1D7>(ept, g, f, t_eff, stream, prepare);
    default: return set_error(OT_EINVAL, "stencil %d is not a 1D stencil", stencil);
  }
}

}  // namespace ot
