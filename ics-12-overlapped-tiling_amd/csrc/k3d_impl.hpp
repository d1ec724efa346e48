// k3d_impl.hpp — overlapped time-tiled 3D kernels as a z-streaming pipeline.
//
// Replaces the kernel text of jacobi-3d.cpp:120-432, poisson-3d.cpp:120-398 and
// gradient-3d.cpp:120-434, which hold a full (bx+2)(by*ept+2)(bz+2) tile in
// local memory and shrink it in all three axes. Here a workgroup owns an
// xy output tile (tile[0] x tile[1]) and a z-chunk of tile[2] planes and
// streams the chunk plane by plane:
//   * iteration k loads plane k of the footprint (tile + 2T in x and y) from
//     HBM into LDS and registers (level 0);
//   * stage s (1..T) computes time level s of plane k - L*s from level s-1,
//     so the T stages form a pipeline that lags L planes per stage (L = 1 for
//     the 7-point stencils, whose z-neighbours live in a 3-deep register queue
//     per level; L = 2 for the 27-point stencil, which reads three LDS planes);
//   * every stage reads only LDS planes written in an earlier iteration, so each
//     level keeps a ring of NS planes and one barrier per iteration suffices;
//   * stage T's plane is written straight to HBM for the points inside the tile.
// The redundant halo work is in x and y only (the z halo of T planes per chunk
// end is the only z redundancy), out-of-domain points are 0 at every level.
// Threads: blockDim = (FX, ceil(FY/E)); thread (tx, ty) owns the column strip
// x = tx, y = ty*E .. ty*E+E-1 of the footprint.
#pragma once
#include "internal.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace ot {
namespace k3 {

using namespace dev;

template <int ST>
struct K3 {
  static constexpr int L = (ST == OT_P3D27) ? 2 : 1;
  static constexpr int NS = (L == 2) ? 4 : 2;
  static constexpr bool RHS = (ST == OT_P3D7);
};

template <int ST, int T, int E>
__global__ void __launch_bounds__(1024) k3d(LaunchGeom g, FieldPtrs f) {
  extern __shared__ float lds[];
  constexpr int L = K3<ST>::L;
  constexpr int NS = K3<ST>::NS;
  constexpr bool RHS = K3<ST>::RHS;
  constexpr int H = T;
  const int FX = g.tile[0] + 2 * H;
  const int FY = g.tile[1] + 2 * H;
  const int FYP = (int)blockDim.y * E;
  const int plane = FYP * FX;
  const int nblk = g.grid[0] * g.grid[1] * g.grid[2];
  const int lid = xcd_remap((int)blockIdx.x, nblk);
  const int bxi = lid % g.grid[0];
  const int byi = (lid / g.grid[0]) % g.grid[1];
  const int bzi = lid / (g.grid[0] * g.grid[1]);
  const int x0 = bxi * g.tile[0] - H, y0 = byi * g.tile[1] - H;
  const int TZ = g.tile[2];
  const int z0 = bzi * TZ;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int gx = x0 + tx;
  const bool in_x = gx >= 0 && gx < g.n[0];
  const bool edge_x = tx < 1 || tx > FX - 2;
  const bool out_x = tx >= H && tx < H + g.tile[0];

  bool inxy[E], work[E], outxy[E];
  int64_t col[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int y = ty * E + i;
    const int gy = y0 + y;
    inxy[i] = in_x && gy >= 0 && gy < g.n[1];
    work[i] = !edge_x && y >= 1 && y <= FY - 2;
    outxy[i] = inxy[i] && out_x && y >= H && y < H + g.tile[1];
    col[i] = g.origin + (int64_t)gy * g.sy + gx;
  }
  if (g.phase_limit == 1) return;

  auto slot = [&](int level, int k) -> float* {
    return lds + (level * NS + (((k % NS) + NS) % NS)) * plane;
  };

  float q[T][3][E];     // L == 1: level l values at planes (k-l-2, k-l-1, k-l)
  float fq[T + 1][E];   // RHS: f at plane k - j
#pragma unroll
  for (int l = 0; l < T; ++l)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int i = 0; i < E; ++i) q[l][a][i] = 0.0f;
#pragma unroll
  for (int j = 0; j <= T; ++j)
#pragma unroll
    for (int i = 0; i < E; ++i) fq[j][i] = 0.0f;

  const int kstart = z0 - T;
  const int kload_end = z0 + TZ - 1 + T;
  const int kend = z0 + TZ - 1 + L * T;
  const int zout_end = min(z0 + TZ, g.n[2]);
  const bool do_loads = g.phase_limit != 3;
  const bool do_store = g.phase_limit == 0;
  const int s_end = (g.phase_limit == 2) ? 1 : T;

  // Level-0 planes are prefetched one iteration ahead into registers, so the
  // HBM latency of plane k+1 overlaps the stage computations of plane k.
  float pf[E], pfr[E];
  auto fetch = [&](int k) {
    const bool inz = k >= 0 && k < g.n[2];
    const int64_t zoff = (int64_t)k * g.sz;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      pf[i] = (do_loads && inz && inxy[i]) ? f.src[0][col[i] + zoff] : 0.0f;
      if constexpr (RHS) pfr[i] = (inz && inxy[i]) ? f.ro[0][col[i] + zoff] : 0.0f;
    }
  };
  fetch(kstart);

  for (int k = kstart; k <= kend; ++k) {
#pragma unroll
    for (int l = 0; l < T; ++l)
#pragma unroll
      for (int i = 0; i < E; ++i) {
        q[l][0][i] = q[l][1][i];
        q[l][1][i] = q[l][2][i];
      }
    if constexpr (RHS) {
#pragma unroll
      for (int j = T; j >= 1; --j)
#pragma unroll
        for (int i = 0; i < E; ++i) fq[j][i] = fq[j - 1][i];
    }
    __syncthreads();

    // ---- level 0: plane k (prefetched), then start fetching plane k+1
    if (k <= kload_end) {
      float* D = slot(0, k);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        q[0][2][i] = pf[i];
        D[(ty * E + i) * FX + tx] = pf[i];
        if constexpr (RHS) fq[0][i] = pfr[i];
      }
      if (k + 1 <= kload_end) fetch(k + 1);
    }

    // ---- stages 1..T
#pragma unroll
    for (int s = 1; s <= T; ++s) {
      if (s > s_end) break;
      const int p = k - L * s;
      const int lo = z0 - (T - s), hi = z0 + TZ - 1 + (T - s);
      if (p < lo || p > hi) continue;
      const bool inz = p >= 0 && p < g.n[2];
      float nv[E];
      if constexpr (L == 1) {
        const float* S = slot(s - 1, k - 1);
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const int y = ty * E + i;
          float v = 0.0f;
          if (work[i]) {
            const float c = q[s - 1][1][i];
            const float zm = q[s - 1][0][i];
            const float zp = q[s - 1][2][i];
            const float ym = (i > 0) ? q[s - 1][1][i - 1] : S[(y - 1) * FX + tx];
            const float yp = (i < E - 1) ? q[s - 1][1][i + 1] : S[(y + 1) * FX + tx];
            const float xm = S[y * FX + tx - 1];
            const float xp = S[y * FX + tx + 1];
            if constexpr (ST == OT_J3D) v = j3d(xm, xp, ym, yp, zm, zp, c);
            else if constexpr (ST == OT_G3D) v = g3d(xm, xp, ym, yp, zm, zp, c);
            else v = p3d7(xm, xp, ym, yp, zm, zp, fq[s][i]);
          }
          nv[i] = (inz && inxy[i]) ? v : 0.0f;
        }
      } else {
        const float* Pm = slot(s - 1, k - 3);
        const float* P0 = slot(s - 1, k - 2);
        const float* Pp = slot(s - 1, k - 1);
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const int y = ty * E + i;
          float v = 0.0f;
          if (work[i]) {
            // poisson-3d.cpp:282-293: 26*c - sum of the 26 neighbours in
            // (dx, dy, dz) lexicographic order
            float sum = 0.0f;
            bool first = true;
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
              for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz) {
                  if (dx == 0 && dy == 0 && dz == 0) continue;
                  const float* Pz = dz < 0 ? Pm : (dz == 0 ? P0 : Pp);
                  const float t = Pz[(y + dy) * FX + tx + dx];
                  sum = first ? t : sum + t;
                  first = false;
                }
            v = 26.0f * P0[y * FX + tx] - sum;
          }
          nv[i] = (inz && inxy[i]) ? v : 0.0f;
        }
      }
      if (s < T) {
        float* D = slot(s, k);
#pragma unroll
        for (int i = 0; i < E; ++i) {
          if constexpr (L == 1) q[s][2][i] = nv[i];
          D[(ty * E + i) * FX + tx] = nv[i];
        }
      } else if (do_store && p >= z0 && p < zout_end) {
        const int64_t zoff = (int64_t)p * g.sz;
#pragma unroll
        for (int i = 0; i < E; ++i)
          if (outxy[i]) f.dst[0][col[i] + zoff] = nv[i];
      }
    }
  }
}

template <int ST, int T, int E>
ot_status launch_t(const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  int bx, by, lds;
  shape_3d(ST, T, E, g.tile, &bx, &by, &lds);
  if (prepare) return prepare_lds((const void*)&k3d<ST, T, E>, lds);
  const int nblk = g.grid[0] * g.grid[1] * g.grid[2];
  hipLaunchKernelGGL((k3d<ST, T, E>), dim3(nblk), dim3(bx, by), lds, (hipStream_t)stream, g, f);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return set_error(OT_EHIP, "3D kernel launch failed: %s", hipGetErrorString(err));
  return OT_OK;
}

template <int ST, int T>
ot_status launch_e(int ept, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  switch (ept) {
    case 1: return launch_t<ST, T, 1>(g, f, stream, prepare);
    case 2: return launch_t<ST, T, 2>(g, f, stream, prepare);
    case 4: return launch_t<ST, T, 4>(g, f, stream, prepare);
    case 8:
      if constexpr (T <= 2) return launch_t<ST, T, 8>(g, f, stream, prepare);
      [[fallthrough]];
    default:
      return set_error(OT_EUNSUPPORTED,
                       "3D kernels are compiled for 1, 2, 4 elements per thread (8 for T <= 2); got %d at T=%d",
                       ept, T);
  }
}

template <int ST>
ot_status launch_st(int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  switch (t_eff) {
    case 1: return launch_e<ST, 1>(ept, g, f, stream, prepare);
    case 2: return launch_e<ST, 2>(ept, g, f, stream, prepare);
    case 3: return launch_e<ST, 3>(ept, g, f, stream, prepare);
    case 4: return launch_e<ST, 4>(ept, g, f, stream, prepare);
    case 5: return launch_e<ST, 5>(ept, g, f, stream, prepare);
    case 6: return launch_e<ST, 6>(ept, g, f, stream, prepare);
    case 7: return launch_e<ST, 7>(ept, g, f, stream, prepare);
    case 8: return launch_e<ST, 8>(ept, g, f, stream, prepare);
    default: return set_error(OT_EUNSUPPORTED, "3D kernels are compiled for T = 1..8 (got %d)", t_eff);
  }
}

}  // namespace k3

}  // namespace ot
