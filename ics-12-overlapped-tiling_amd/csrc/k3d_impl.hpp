// k3d_impl.hpp — overlapped time-tiled 3D kernels as a z-streaming pipeline.
//
// Replaces the kernel text of jacobi-3d.cpp:120-432, poisson-3d.cpp:120-398 and
// gradient-3d.cpp:120-434, which hold a full (bx+2)(by*ept+2)(bz+2) tile in
// local memory and shrink it in all three axes. Here a workgroup owns an
// xy output tile (tile[0] x tile[1]) and a z-chunk of tile[2] planes and
// streams the chunk plane by plane:
//   * iteration k stores plane k of the footprint (tile + 2T in x and y) into
//     LDS and registers as time level 0 — the plane was fetched from HBM one
//     iteration earlier, so HBM latency overlaps the previous plane's compute;
//   * stage s (1..T) computes time level s of plane k - L*s from level s-1,
//     so the T stages form a pipeline that lags L planes per stage (L = 1 for
//     the 7-point stencils, whose z-neighbours live in a 3-deep register queue
//     per level; L = 2 for the 27-point stencil, which reads three LDS planes);
//   * every stage reads only LDS planes written in an earlier iteration, so each
//     level keeps a ring of NS planes and one barrier per iteration suffices;
//   * stage T's plane is written straight to HBM for the points inside the tile.
// The z-loop is unrolled by U (a multiple of 3 and NS) so the register-queue
// rotation and the LDS ring slot of every iteration are compile-time
// constants: no register moves and no slot arithmetic in the loop.
// The LDS planes carry one guard row (+1 element) before the first and after
// the last plane, so footprint-edge threads read in-bounds garbage instead of
// branching; their results lie outside the region later stages consume.
// The redundant halo work is in x and y only (plus T planes at each z-chunk
// end); out-of-domain points are forced to 0 at every level (Dirichlet 0).
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
  static constexpr int U = (L == 2) ? 4 : 6;   // z-loop unroll: multiple of NS (and of 3 for L == 1)
  static constexpr bool RHS = (ST == OT_P3D7);
};

template <int V>
struct IC {
  static constexpr int value = V;
};

// LDS floats needed by one workgroup (planes + the two guard regions).
__host__ __device__ inline int k3_lds_floats(int T, int NS, int FX, int FYP) {
  return T * NS * FYP * FX + 2 * (FX + 1);
}

// MAXT bounds the workgroup size the instantiation is compiled for: 1024
// leaves 128 VGPRs per lane, 512 leaves 256, which the deep pipelines (large
// T*E register queues) need to stay out of scratch.
template <int ST, int T, int E, int MAXT>
__global__ void __launch_bounds__(MAXT) k3d(LaunchGeom g, FieldPtrs f) {
  extern __shared__ float lds_raw[];
  constexpr int L = K3<ST>::L;
  constexpr int NS = K3<ST>::NS;
  constexpr int U = K3<ST>::U;
  constexpr bool RHS = K3<ST>::RHS;
  constexpr int H = T;
  const int FX = g.tile[0] + 2 * H;
  const int FYP = (int)blockDim.y * E;
  const int plane = FYP * FX;
  float* const lds = lds_raw + FX + 1;  // guard before the first plane
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
  const bool out_x = tx >= H && tx < H + g.tile[0];
  const int my = ty * E * FX + tx;  // this thread's row-0 offset inside a plane

  // HBM traffic goes through raw buffer descriptors whose base is this
  // workgroup's first plane (z0 - T): the plane is selected with the scalar
  // offset, the point inside the plane with a per-lane 32-bit byte offset.
  // Lanes outside the domain (or outside the tile, for stores) use an offset
  // beyond the descriptor's range, which the hardware turns into a 0 load or a
  // dropped store, so no per-lane branches or 64-bit address math remain.
  // The descriptor range covers the whole chunk (TZ + 2T planes), so a valid
  // point is in range whether or not the range check adds the scalar offset.
  constexpr uint32_t kOOB = 0x80000000u;
  const int64_t zbase = (int64_t)(z0 - T) * g.sz;
  const uint32_t plane_bytes = (uint32_t)(g.sz * 4);
  const int range = (int)((uint32_t)(TZ + 2 * T + 1) * plane_bytes);
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)(f.src[0] + g.origin + zbase), (short)0, range, 0x00020000);
  const __amdgpu_buffer_rsrc_t rdst =
      __builtin_amdgcn_make_buffer_rsrc((void*)(f.dst[0] + g.origin + zbase), (short)0, range, 0x00020000);
  __amdgpu_buffer_rsrc_t rro = rsrc;
  if constexpr (RHS)
    rro = __builtin_amdgcn_make_buffer_rsrc((void*)(f.ro[0] + g.origin + zbase), (short)0, range, 0x00020000);

  uint32_t mxy[E];  // all-ones inside the domain: results are ANDed with it (+0.0 outside)
  uint32_t voff[E], soff_out[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int y = ty * E + i;
    const int gy = y0 + y;
    const bool inxy = in_x && gy >= 0 && gy < g.n[1];
    const bool outxy = inxy && out_x && y >= H && y < H + g.tile[1];
    mxy[i] = inxy ? 0xFFFFFFFFu : 0u;
    const uint32_t o = (uint32_t)(((int64_t)gy * g.sy + gx) * 4);
    voff[i] = inxy ? o : kOOB;
    soff_out[i] = outxy ? o : kOOB;
  }
  if (g.phase_limit == 1) return;

  float q[T][3][E];    // L == 1: level l at the planes of the last three iterations
  float fq[T + 1][E];  // RHS: f at plane k - j
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

  // Level-0 planes are prefetched PD iterations ahead into a register ring.
  // Every load and store is issued unconditionally — a plane outside the
  // domain (or no output plane) uses the out-of-range lane offset, which reads
  // 0 and drops stores without touching memory — so the number of memory
  // operations between a load and its use is fixed and the wait before the use
  // only covers that load, not the stores issued after it.
  constexpr int PD = 2;
  static_assert(U % PD == 0, "prefetch ring must divide the unroll");
  float pf[PD][E], pfr[PD][E];
  auto fetch = [&](int k, auto slotc) {
    constexpr int SL = decltype(slotc)::value;
    const bool live = do_loads && k >= 0 && k < g.n[2] && k <= kload_end;
    const int soff = (int)((uint32_t)(live ? k - (z0 - T) : 0) * plane_bytes);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int vo = (int)(live ? voff[i] : kOOB);
      pf[SL][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, vo, soff, 0));
      if constexpr (RHS)
        pfr[SL][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rro, vo, soff, 0));
    }
  };
  fetch(kstart, IC<0>{});
  fetch(kstart + 1, IC<1>{});

  // Stage T's plane is stored at the top of the next iteration, next to the
  // prefetch: on gfx9 stores count in vmcnt, so a store issued at the end of an
  // iteration would make the next iteration's wait for its prefetched plane
  // also wait for the store.
  float outv[E];
  int out_p = -1;
  auto flush = [&]() {
    const bool live = out_p >= 0;
    const int soff = (int)((uint32_t)(live ? out_p - (z0 - T) : 0) * plane_bytes);
#pragma unroll
    for (int i = 0; i < E; ++i)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, outv[i]), rdst,
                                            (int)(live ? soff_out[i] : kOOB), soff, 0);
  };

  // One iteration at compile-time phase PH (= (k - kstart) mod U).
  auto iter = [&](auto phc, const int k) {
    constexpr int PH = decltype(phc)::value;
    constexpr int QN = PH % 3;        // register slot written this iteration
    constexpr int QP = (PH + 2) % 3;  // written one iteration ago
    constexpr int QPP = (PH + 1) % 3; // written two iterations ago
    constexpr int SW = PH % NS;       // LDS ring slot written this iteration
    if constexpr (RHS) {
#pragma unroll
      for (int j = T; j >= 1; --j)
#pragma unroll
        for (int i = 0; i < E; ++i) fq[j][i] = fq[j - 1][i];
    }
    __syncthreads();

    // ---- level 0: plane k (prefetched); store the previous output plane, then
    // start fetching plane k+PD into the ring slot just consumed
    constexpr int PS = PH % PD;
    if (k <= kload_end) {
      float* D = lds + SW * plane + my;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        q[0][QN][i] = pf[PS][i];
        D[i * FX] = pf[PS][i];
        if constexpr (RHS) fq[0][i] = pfr[PS][i];
      }
    }
    flush();
    out_p = -1;
    fetch(k + PD, IC<PS>{});

    if constexpr (L == 1) {
      // ---- 7-point stages. Every stage reads only LDS planes written in an
      // earlier iteration, so all LDS reads of all stages are issued first
      // (one exposed LDS latency per iteration), then the stages are computed
      // in order (stage s takes its z+1 neighbour from stage s-1's result of
      // this iteration), then the new planes are written.
      constexpr int SR = (PH + NS - 1) % NS;
      float xm[T][E], xp[T][E], ym0[T], ypE[T];
#pragma unroll
      for (int s = 1; s <= T; ++s) {
        const float* S = lds + ((s - 1) * NS + SR) * plane + my;
#pragma unroll
        for (int i = 0; i < E; ++i) {
          xm[s - 1][i] = S[i * FX - 1];
          xp[s - 1][i] = S[i * FX + 1];
        }
        ym0[s - 1] = S[-FX];
        ypE[s - 1] = S[E * FX];
      }
#pragma unroll
      for (int s = 1; s <= T; ++s) {
        if (s > s_end) break;
        const int p = k - s;
        const int lo = z0 - (T - s), hi = z0 + TZ - 1 + (T - s);
        if (p < lo || p > hi) continue;
        const uint32_t mz = (p >= 0 && p < g.n[2]) ? 0xFFFFFFFFu : 0u;
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const float c = q[s - 1][QP][i];
          const float zm = q[s - 1][QPP][i];
          const float zp = q[s - 1][QN][i];
          const float ym = (i > 0) ? q[s - 1][QP][i - 1] : ym0[s - 1];
          const float yp = (i < E - 1) ? q[s - 1][QP][i + 1] : ypE[s - 1];
          float v;
          if constexpr (ST == OT_J3D) v = j3d(xm[s - 1][i], xp[s - 1][i], ym, yp, zm, zp, c);
          else if constexpr (ST == OT_G3D) v = g3d(xm[s - 1][i], xp[s - 1][i], ym, yp, zm, zp, c);
          else v = p3d7(xm[s - 1][i], xp[s - 1][i], ym, yp, zm, zp, fq[s][i]);
          const float nv = __uint_as_float(__float_as_uint(v) & (mxy[i] & mz));
          if (s < T) q[s][QN][i] = nv;
          else outv[i] = nv;
        }
        if (s == T && do_store && p >= z0 && p < zout_end) out_p = p;
      }
#pragma unroll
      for (int s = 1; s < T; ++s) {
        if (s > s_end) break;
        const int p = k - s;
        const int lo = z0 - (T - s), hi = z0 + TZ - 1 + (T - s);
        if (p < lo || p > hi) continue;
        float* D = lds + (s * NS + SW) * plane + my;
#pragma unroll
        for (int i = 0; i < E; ++i) D[i * FX] = q[s][QN][i];
      }
    } else {
    // ---- 27-point stages (three LDS planes per stage)
#pragma unroll
    for (int s = 1; s <= T; ++s) {
      if (s > s_end) break;
      const int p = k - L * s;
      const int lo = z0 - (T - s), hi = z0 + TZ - 1 + (T - s);
      if (p < lo || p > hi) continue;
      const uint32_t mz = (p >= 0 && p < g.n[2]) ? 0xFFFFFFFFu : 0u;
      float nv[E];
      {
        const float* Pm = lds + ((s - 1) * NS + (PH + NS - 3) % NS) * plane + my;
        const float* P0 = lds + ((s - 1) * NS + (PH + NS - 2) % NS) * plane + my;
        const float* Pp = lds + ((s - 1) * NS + (PH + NS - 1) % NS) * plane + my;
#pragma unroll
        for (int i = 0; i < E; ++i) {
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
                const float t = Pz[(i + dy) * FX + dx];
                sum = first ? t : sum + t;
                first = false;
              }
          const float v = 26.0f * P0[i * FX] - sum;
          nv[i] = __uint_as_float(__float_as_uint(v) & (mxy[i] & mz));
        }
      }
      if (s < T) {
        float* D = lds + (s * NS + SW) * plane + my;
#pragma unroll
        for (int i = 0; i < E; ++i) D[i * FX] = nv[i];
      } else if (do_store && p >= z0 && p < zout_end) {
#pragma unroll
        for (int i = 0; i < E; ++i) outv[i] = nv[i];
        out_p = p;
      }
    }
    }
  };

  int k = kstart;
  for (; k + U - 1 <= kend; k += U) {
    iter(IC<0>{}, k);
    iter(IC<1>{}, k + 1);
    iter(IC<2>{}, k + 2);
    iter(IC<3>{}, k + 3);
    if constexpr (U > 4) {
      iter(IC<4>{}, k + 4);
      iter(IC<5>{}, k + 5);
    }
  }
  if (k <= kend) iter(IC<0>{}, k);
  if (k + 1 <= kend) iter(IC<1>{}, k + 1);
  if (k + 2 <= kend) iter(IC<2>{}, k + 2);
  if constexpr (U > 4) {
    if (k + 3 <= kend) iter(IC<3>{}, k + 3);
    if (k + 4 <= kend) iter(IC<4>{}, k + 4);
  }
  flush();
}

template <int ST, int T, int E>
ot_status launch_t(const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  int bx, by, lds;
  shape_3d(ST, T, E, g.tile, &bx, &by, &lds);
  // buffer descriptors: in-plane byte offsets must stay below the 2^31
  // out-of-range marker and a chunk's plane offsets must fit 32 bits
  const int64_t plane_bytes = g.sz * 4;
  if ((int64_t)(g.tile[2] + 2 * T + 3) * plane_bytes >= (1LL << 31))
    return set_error(OT_EUNSUPPORTED, "3D plane of %lld bytes too large for a z-chunk of %d planes",
                     (long long)plane_bytes, g.tile[2]);
  const bool small = bx * by <= 512;
  const void* kern = small ? (const void*)&k3d<ST, T, E, 512> : (const void*)&k3d<ST, T, E, 1024>;
  if (prepare) return prepare_lds(kern, lds);
  const int nblk = g.grid[0] * g.grid[1] * g.grid[2];
  if (small)
    hipLaunchKernelGGL((k3d<ST, T, E, 512>), dim3(nblk), dim3(bx, by), lds, (hipStream_t)stream, g, f);
  else
    hipLaunchKernelGGL((k3d<ST, T, E, 1024>), dim3(nblk), dim3(bx, by), lds, (hipStream_t)stream, g, f);
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
