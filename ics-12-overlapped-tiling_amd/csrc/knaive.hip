// knaive.hip — straightforward one-time-step-per-launch kernels (no tiling).
//
// The counterpart of the reference CUDA prototypes' non-time-tiled kernels
// (cuda/jacobi-1d/jacobi-1d.cu:87-101, cuda/jacobi-2d/jacobi-2d.cu:88-105):
// one thread per point, every tap read from global memory, out-of-domain taps
// read as 0. The ot-<stencil> drivers use this path for --verify (the
// overlapped-tiled result compared with an independent evaluation of the same
// formulas), and it is the no-tiling baseline of the measurements.
#include "internal.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace ot {
namespace {

using namespace dev;

struct NaiveGeom {
  int n[3];
  int64_t sy, sz, origin;
  int variant;
  float coef[4];
};

__device__ __forceinline__ float tap(const float* a, const NaiveGeom& g, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) return 0.0f;
  return a[g.origin + (int64_t)z * g.sz + (int64_t)y * g.sy + x];
}

template <int ST>
__global__ void __launch_bounds__(256) naive_step(NaiveGeom g, const float* __restrict__ u,
                                                  const float* __restrict__ r0,
                                                  const float* __restrict__ r1, float* __restrict__ out) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (id >= total) return;
  const int x = (int)(id % g.n[0]);
  const int y = (int)((id / g.n[0]) % g.n[1]);
  const int z = (int)(id / ((int64_t)g.n[0] * g.n[1]));
  const int64_t i = g.origin + (int64_t)z * g.sz + (int64_t)y * g.sy + x;
  auto U = [&](int dx, int dy, int dz) { return tap(u, g, x + dx, y + dy, z + dz); };
  float v;
  if constexpr (ST == OT_J1D3) v = j1d3(U(-1, 0, 0), U(0, 0, 0), U(1, 0, 0));
  else if constexpr (ST == OT_J1D5) v = j1d5(U(-2, 0, 0), U(-1, 0, 0), U(0, 0, 0), U(1, 0, 0), U(2, 0, 0));
  else if constexpr (ST == OT_J1D7)
    v = j1d7(U(-3, 0, 0), U(-2, 0, 0), U(-1, 0, 0), U(0, 0, 0), U(1, 0, 0), U(2, 0, 0), U(3, 0, 0));
  else if constexpr (ST == OT_J2D) v = j2d(U(-1, 0, 0), U(0, 0, 0), U(1, 0, 0), U(0, -1, 0), U(0, 1, 0));
  else if constexpr (ST == OT_G2D) v = g2d(U(-1, 0, 0), U(0, 0, 0), U(1, 0, 0), U(0, -1, 0), U(0, 1, 0));
  else if constexpr (ST == OT_P2D)
    v = p2d(U(-1, -1, 0), U(0, -1, 0), U(1, -1, 0), U(-1, 0, 0), U(0, 0, 0), U(1, 0, 0), U(-1, 1, 0),
            U(0, 1, 0), U(1, 1, 0));
  else if constexpr (ST == OT_TV2D)
    v = tv2d(U(0, -1, 0), U(1, -1, 0), U(-1, 0, 0), U(0, 0, 0), U(1, 0, 0), U(-1, 1, 0), U(0, 1, 0), r0[i], r1[i],
             g.coef[0] != 0.0f ? g.coef[0] : 1.1f);
  else if constexpr (ST == OT_J3D)
    v = j3d(U(-1, 0, 0), U(1, 0, 0), U(0, -1, 0), U(0, 1, 0), U(0, 0, -1), U(0, 0, 1), U(0, 0, 0));
  else if constexpr (ST == OT_G3D)
    v = g3d(U(-1, 0, 0), U(1, 0, 0), U(0, -1, 0), U(0, 1, 0), U(0, 0, -1), U(0, 0, 1), U(0, 0, 0));
  else if constexpr (ST == OT_P3D7)
    v = p3d7(U(-1, 0, 0), U(1, 0, 0), U(0, -1, 0), U(0, 1, 0), U(0, 0, -1), U(0, 0, 1), r0[i]);
  else if constexpr (ST == OT_CS3D)
    v = cs3d(U(0, 0, 0), U(0, 0, -1), U(0, 0, 1), U(0, -1, 0), U(0, 1, 0), U(-1, 0, 0), U(1, 0, 0), U(0, 1, -1),
             U(1, 0, -1), U(0, -1, 1), U(1, -1, 0), U(-1, 0, 1), U(-1, 1, 0));
  else {  // P3D27
    float sum = 0.0f;
    bool first = true;
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          if (dx == 0 && dy == 0 && dz == 0) continue;
          const float t = U(dx, dy, dz);
          sum = first ? t : sum + t;
          first = false;
        }
    v = 26.0f * U(0, 0, 0) - sum;
  }
  out[i] = v;
}

// Rician sub-stage 1: g over the domain.
__global__ void __launch_bounds__(256) naive_rician_g(NaiveGeom g, const float* __restrict__ u,
                                                      float* __restrict__ gout) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)g.n[0] * g.n[1]) return;
  const int x = (int)(id % g.n[0]), y = (int)(id / g.n[0]);
  const int64_t i = g.origin + (int64_t)y * g.sy + x;
  gout[i] = rician_g(tap(u, g, x, y, 0), tap(u, g, x - 1, y, 0), tap(u, g, x + 1, y, 0), tap(u, g, x, y - 1, 0),
                     tap(u, g, x, y + 1, 0));
}

__global__ void __launch_bounds__(256) naive_rician_u(NaiveGeom g, const float* __restrict__ u,
                                                      const float* __restrict__ gg, const float* __restrict__ f,
                                                      float* __restrict__ out) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)g.n[0] * g.n[1]) return;
  const int x = (int)(id % g.n[0]), y = (int)(id / g.n[0]);
  const int64_t i = g.origin + (int64_t)y * g.sy + x;
  RicianK rk;
  rk.sigma2 = g.coef[0] != 0.0f ? g.coef[0] : 0.0025f;  // SIGMA2 (rician-denoise-2d.cpp:159)
  rk.gamma = (g.coef[1] != 0.0f ? g.coef[1] : 0.065f) / rk.sigma2;
  rk.dt = g.coef[2] != 0.0f ? g.coef[2] : 5.0f;
  out[i] = rician_u(u[i], f[i], tap(u, g, x - 1, y, 0), tap(u, g, x + 1, y, 0), tap(u, g, x, y - 1, 0),
                    tap(u, g, x, y + 1, 0), tap(gg, g, x - 1, y, 0), tap(gg, g, x + 1, y, 0),
                    tap(gg, g, x, y - 1, 0), tap(gg, g, x, y + 1, 0), rk);
}

// FDTD: E update in place (reads Hz only), then H update in place.
__global__ void __launch_bounds__(256) naive_fdtd_e(NaiveGeom g, float* __restrict__ ex, float* __restrict__ ey,
                                                    const float* __restrict__ hz) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)g.n[0] * g.n[1]) return;
  const int x = (int)(id % g.n[0]), y = (int)(id / g.n[0]);
  const int64_t i = g.origin + (int64_t)y * g.sy + x;
  const float h = hz[i];
  ey[i] = fdtd_ey(ey[i], h, tap(hz, g, x, y - 1, 0));
  ex[i] = fdtd_ex(ex[i], h, tap(hz, g, x - 1, y, 0));
}

__global__ void __launch_bounds__(256) naive_fdtd_h(NaiveGeom g, const float* __restrict__ ex,
                                                    const float* __restrict__ ey, float* __restrict__ hz) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)g.n[0] * g.n[1]) return;
  const int x = (int)(id % g.n[0]), y = (int)(id / g.n[0]);
  const int64_t i = g.origin + (int64_t)y * g.sy + x;
  hz[i] = fdtd_hz(hz[i], ex[i], tap(ex, g, x + 1, y, 0), ey[i], tap(ey, g, x, y + 1, 0),
                  (g.variant & OT_VARIANT_FDTD_PHYSICAL) != 0);
}

#define OT_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return set_error(OT_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

}  // namespace
}  // namespace ot

using namespace ot;

extern "C" ot_status ot_naive_run(const ot_params* params, float* const* rw_a, float* const* rw_b,
                                  const float* const* ro, void* stream, ot_timing* timing) {
  if (!params || !rw_a || !rw_b) return set_error(OT_EINVAL, "null argument");
  const ot_params& p = *params;
  if (p.phase_limit != 0) return set_error(OT_EINVAL, "phase limits apply to the tiled kernels only");
  // The layout (padding, realSize, strides) depends on T and the blocks in
  // reference geometry, so derive it from the caller's parameters; the tiled
  // kernels' resource limits do not concern this path.
  ot_derived d;
  ot_status st = derive(p, d);
  if (st != OT_OK && st != OT_ERESOURCE) return st;
  for (int k = 0; k < d.num_rw; ++k)
    if (!rw_a[k] || !rw_b[k]) return set_error(OT_EINVAL, "null rw buffer %d", k);
  for (int k = 0; k < d.num_ro; ++k)
    if (!ro || !ro[k]) return set_error(OT_EINVAL, "null read-only buffer %d", k);
  NaiveGeom g;
  for (int a = 0; a < 3; ++a) g.n[a] = d.real_size[a];
  g.sy = d.stride_y;
  g.sz = d.stride_z;
  g.origin = d.origin;
  g.variant = p.variant;
  for (int i = 0; i < 4; ++i) g.coef[i] = p.coef[i];
  hipStream_t s = (hipStream_t)stream;
  const int64_t total = (int64_t)g.n[0] * g.n[1] * g.n[2];
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  // Rician keeps g of every point between its two sub-stages; the scratch
  // buffer and the events are released on every return path
  struct Scratch {
    float* gbuf = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Scratch() {
      if (gbuf) (void)hipFree(gbuf);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } sc;
  if (p.stencil == OT_RICIAN2D) {
    OT_HIP(hipMalloc(&sc.gbuf, (size_t)d.field_elems * sizeof(float)));
    OT_HIP(hipMemset(sc.gbuf, 0, (size_t)d.field_elems * sizeof(float)));
  }
  float* const gbuf = sc.gbuf;
  OT_HIP(hipEventCreate(&sc.e0));
  OT_HIP(hipEventCreate(&sc.e1));
  const hipEvent_t e0 = sc.e0, e1 = sc.e1;
  OT_HIP(hipEventRecord(e0, s));
  int final_buffer = 0;
  for (int t = 0; t < p.time_steps; ++t) {
    float* const* src = (t % 2 == 0) ? rw_a : rw_b;
    float* const* dst = (t % 2 == 0) ? rw_b : rw_a;
    switch (p.stencil) {
#define NAIVE(ST)                                                                                             \
  case ST:                                                                                                    \
    hipLaunchKernelGGL(naive_step<ST>, grid, block, 0, s, g, src[0], d.num_ro > 0 ? ro[0] : nullptr,          \
                       d.num_ro > 1 ? ro[1] : nullptr, dst[0]);                                               \
    final_buffer = (t % 2 == 0) ? 1 : 0;                                                                      \
    break;
      NAIVE(OT_J1D3)
      NAIVE(OT_J1D5)
      NAIVE(OT_J1D7)
      NAIVE(OT_J2D)
      NAIVE(OT_J3D)
      NAIVE(OT_P2D)
      NAIVE(OT_P3D27)
      NAIVE(OT_P3D7)
      NAIVE(OT_G2D)
      NAIVE(OT_G3D)
      NAIVE(OT_TV2D)
      NAIVE(OT_CS3D)
#undef NAIVE
      case OT_RICIAN2D:
        hipLaunchKernelGGL(naive_rician_g, grid, block, 0, s, g, src[0], gbuf);
        hipLaunchKernelGGL(naive_rician_u, grid, block, 0, s, g, src[0], gbuf, ro[0], dst[0]);
        final_buffer = (t % 2 == 0) ? 1 : 0;
        break;
      case OT_FDTD2D:
        hipLaunchKernelGGL(naive_fdtd_e, grid, block, 0, s, g, rw_a[0], rw_a[1], rw_a[2]);
        hipLaunchKernelGGL(naive_fdtd_h, grid, block, 0, s, g, rw_a[0], rw_a[1], rw_a[2]);
        final_buffer = 0;
        break;
      default:
        return set_error(OT_EINVAL, "unknown stencil %d", p.stencil);
    }
    OT_HIP(hipGetLastError());
  }
  OT_HIP(hipEventRecord(e1, s));
  OT_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  OT_HIP(hipEventElapsedTime(&ms, e0, e1));
  if (timing) {
    timing->wall_s = 0.0;
    timing->event_s = ms * 1e-3;
    timing->kernel_s = ms * 1e-3;
    timing->kernel_full_s = ms * 1e-3;
    timing->launches = p.time_steps * ((p.stencil == OT_RICIAN2D || p.stencil == OT_FDTD2D) ? 2 : 1);
    timing->full_launches = timing->launches;
    timing->final_buffer = final_buffer;
    timing->reserved = 0;
  }
  return OT_OK;
}
