// k2w_impl.hpp — wave-per-tile overlapped time-tiled kernels for the 2D stencils.
//
// A second kernel family for the 2D stencils (jacobi-2d.cpp:175-514,
// poisson-2d.cpp:127-364, gradient-2d.cpp:114-393, fdtd-2d.cpp:107-443,
// rician-denoise-2d.cpp:108-447, cdsc-tv-update-2d.cpp:116-412) with no LDS
// and no barriers: every wavefront is an independent tile.
//   * The 64 lanes are 64 consecutive x columns of the footprint; every lane
//     holds its whole column strip of FR footprint rows of every field in
//     registers. x-neighbours come from the adjacent lanes through DPP wave
//     shifts, y-neighbours are the lane's own neighbouring rows.
//   * All t_eff steps are advanced in registers. Step s recomputes only the
//     rows still needed by later steps (the valid region shrinks by the
//     stencil's radius per step), so the redundant work is the x-halo of 2H
//     lanes plus a shrinking y-halo (H = radius * t_eff).
//   * Output tile = (64 - 2H) x (FR - 2H) points; every point outside the
//     domain is forced to +0 after every step (Dirichlet 0).
// Multi-stage stencils keep the reference's per-step order: FDTD updates Ey, Ex
// from Hz, then Hz from the new Ex, Ey (fdtd-2d.cpp:269-330); Rician computes g
// from u, then u from g and u (rician-denoise-2d.cpp:271-327).
#pragma once
#include "internal.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace ot {
namespace k2w {

using namespace dev;

template <int ST>
struct KW2 {
  static constexpr int R = (ST == OT_RICIAN2D) ? 2 : 1;
  static constexpr int NRW = (ST == OT_FDTD2D) ? 3 : 1;
  static constexpr int NRO = (ST == OT_TV2D) ? 2 : ((ST == OT_RICIAN2D) ? 1 : 0);
};

constexpr int kWaves = 4;

template <int V>
struct IC {
  static constexpr int value = V;
};

// One wavefront advances output tile (bxi, byi) of the grid described by g by T
// steps (bxi, byi wave-uniform).
template <int ST, int T, int FR>
__device__ __forceinline__ void tile(const LaunchGeom& g, const FieldPtrs& f, const int bxi, const int byi) {
  constexpr int R = KW2<ST>::R;
  constexpr int NRW = KW2<ST>::NRW;
  constexpr int NRO = KW2<ST>::NRO;
  constexpr int H = R * T;
  static_assert(FR > 2 * H, "footprint rows must exceed the halo");
  const int lane = threadIdx.x & 63;
  // tile (TX, TY) with TX + 2H <= 64 lanes and TY + 2H <= FR rows; a plan uses
  // TX = 64 - 2H and FR = TY + 2H, a shortened last launch (smaller t_eff)
  // leaves the spare lanes and rows as dead halo
  const int TX = g.tile[0], TY = g.tile[1];
  const int x0 = bxi * TX - H, y0 = byi * TY - H;
  const int gx = x0 + lane;
  const bool in_x = gx >= 0 && gx < g.n[0];
  const bool out_x = lane >= H && lane < H + TX;

  // Buffer addressing: every field's descriptor starts at footprint row 0
  // (x = 0), and the byte offset of row i = lane term + row term. A lane or a
  // row outside the domain has term = range, so its loads return 0 and its
  // stores are dropped: no branch and no 64-bit address arithmetic per row.
  // The host keeps FR + 1 rows under 1 GiB, so no offset sum wraps.
  const uint32_t row_bytes = (uint32_t)(g.sy * 4);
  const uint32_t range = (uint32_t)(FR + 1) * row_bytes;
  const int64_t rbase = g.origin + (int64_t)y0 * g.sy;
  auto rsrc = [&](const float* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p + rbase), (short)0, (int)range, 0x00020000);
  };
  const uint32_t lane_ld = in_x ? (uint32_t)gx * 4u : range;
  const uint32_t lane_st = (in_x && out_x) ? (uint32_t)gx * 4u : range;
  uint32_t row_ld[FR], mxy[FR];
#pragma unroll
  for (int i = 0; i < FR; ++i) {
    const int gy = y0 + i;
    const bool row_in = gy >= 0 && gy < g.n[1];
    row_ld[i] = row_in ? (uint32_t)i * row_bytes : range;
    mxy[i] = (in_x && row_in) ? 0xFFFFFFFFu : 0u;
  }
  if (g.phase_limit == 1) return;
  // gradient-2d: 1e-4f held in a VGPR (the empty asm hides its value), so the
  // first add of the update can take the left lane's difference as a DPP
  // operand; set once, ahead of the loads, so the asm orders nothing else
  float eps4 = 0.0001f;
  if constexpr (ST == OT_G2D) asm volatile("" : "+v"(eps4));

  // A tile whose whole footprint lies in the domain needs no Dirichlet mask;
  // each variant holds its own registers from load to store.
  const bool interior = x0 >= 0 && x0 + 63 < g.n[0] && y0 >= 0 && y0 + FR - 1 < g.n[1];
  // PHASE: honour the phase-limit split (only the masked variant does, so the
  // mask-free variant carries no phase checks)
  auto run = [&](auto maskc, auto phasec) __attribute__((always_inline)) {
    constexpr bool MASK = decltype(maskc)::value != 0;
    constexpr bool PHASE = decltype(phasec)::value != 0;
    auto mask = [&](float v, int i) { return MASK ? __uint_as_float(__float_as_uint(v) & mxy[i]) : v; };
    float u[NRW][FR];
    float ro[NRO > 0 ? NRO : 1][FR];
#pragma unroll
    for (int k = 0; k < NRW; ++k) {
      const auto r = rsrc(f.src[k]);
#pragma unroll
      for (int i = 0; i < FR; ++i)
        u[k][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(lane_ld + row_ld[i]), 0, 0));
    }
    if (PHASE && g.phase_limit == 3) {  // timing split: synthetic input instead of the loads
#pragma unroll
      for (int k = 0; k < NRW; ++k)
#pragma unroll
        for (int i = 0; i < FR; ++i) u[k][i] = (float)lane;
    }
#pragma unroll
    for (int k = 0; k < NRO; ++k) {
      const auto r = rsrc(f.ro[k]);
#pragma unroll
      for (int i = 0; i < FR; ++i)
        ro[k][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(lane_ld + row_ld[i]), 0, 0));
    }

    // steps s = s_begin..s_end of this launch (phase limits: 2 = first step only,
    // 3 = all but the first step)
    const int s_begin = (PHASE && g.phase_limit == 3) ? 2 : 1;
    const int s_end = (PHASE && g.phase_limit == 2) ? 1 : T;

#pragma unroll
    for (int s = 1; s <= T; ++s) {
      if constexpr (PHASE) {
        if (s > s_end) break;
        if (s < s_begin) continue;
      }
      if constexpr (ST == OT_J2D || ST == OT_P2D || ST == OT_G2D || ST == OT_TV2D) {
        float nu[FR];
        // TV: rsqrt of the gradient sum (1e-8 + (y+1 diff)^2) + (x+1 diff)^2 once
        // per row; c1 (= c3), c2 and c4 are it at the row, the row above and the
        // lane to the left (cdsc-tv-update-2d.cpp:294-300, same order of terms)
        float gr[ST == OT_TV2D ? FR : 1];
        if constexpr (ST == OT_TV2D) {
#pragma unroll
          for (int i = s - 1; i < FR - s; ++i) {
            const float c = u[0][i];
            gr[i] = rsq((1e-8f + sq(u[0][i + 1] - c)) + sq(from_right(c) - c));
          }
        }
        // gradient-2d: each squared face difference (c - n)^2 equals (n - c)^2
        // of the neighbour bit for bit (IEEE subtraction is antisymmetric), so
        // a point computes its +x and +y differences and takes the -x one from
        // the lane to its left (DPP) and the -y one from the row above (dyt,
        // carried down the rows), summed in the update's order
        // (gradient-2d.cpp:691-699)
        float dyt = 0.0f;
        if constexpr (ST == OT_G2D) dyt = sq(u[0][s - 1] - u[0][s]);
#pragma unroll
        for (int i = s; i < FR - s; ++i) {
          const float c = u[0][i], tc = u[0][i - 1], bc = u[0][i + 1];
          float v;
          if constexpr (ST == OT_J2D) {
            v = j2d(from_left(c), c, from_right(c), tc, bc);
          } else if constexpr (ST == OT_G2D) {
            const float dxr = sq(c - from_right(c));
            const float dyb = sq(c - bc);
            v = c + rsq((((from_left(dxr) + eps4) + dxr) + dyt) + dyb);
            dyt = dyb;
          } else if constexpr (ST == OT_P2D) {
            v = p2d(from_left(tc), tc, from_right(tc), from_left(c), c, from_right(c), from_left(bc), bc,
                    from_right(bc));
          } else {
            const float alpha = g.coef[0] != 0.0f ? g.coef[0] : 1.1f;
            v = tv2d_g(tc, from_left(c), c, from_right(c), bc, ro[0][i], ro[1][i], alpha, gr[i], gr[i - 1],
                       from_left(gr[i]));
          }
          nu[i] = mask(v, i);
        }
#pragma unroll
        for (int i = s; i < FR - s; ++i) u[0][i] = nu[i];
      } else if constexpr (ST == OT_RICIAN2D) {
        RicianK rk;
        rk.sigma2 = g.coef[0] != 0.0f ? g.coef[0] : 0.0025f;  // SIGMA2 (rician-denoise-2d.cpp:159)
        rk.gamma = (g.coef[1] != 0.0f ? g.coef[1] : 0.065f) / rk.sigma2;
        rk.dt = g.coef[2] != 0.0f ? g.coef[2] : 5.0f;
        // g on rows [2s-1, FR-2s], then u on rows [2s, FR-1-2s]. g's squared
        // differences are shared as gradient-2d's: the point's +x and +y ones,
        // the -x one from the lane to its left, the -y one from the row above,
        // in g's order (right, left, bottom, top; rician-denoise-2d.cpp:271-281)
        float dyb[FR];
#pragma unroll
        for (int i = 2 * s - 2; i <= FR - 2 * s; ++i) dyb[i] = sq(u[0][i] - u[0][i + 1]);
        float gg[FR];
#pragma unroll
        for (int i = 2 * s - 1; i <= FR - 2 * s; ++i) {
          const float c = u[0][i];
          const float dxr = sq(c - from_right(c));
          gg[i] = mask(rsq((((1.0e-20f + dxr) + from_left(dxr)) + dyb[i]) + dyb[i - 1]), i);
        }
        float nu[FR];
#pragma unroll
        for (int i = 2 * s; i <= FR - 1 - 2 * s; ++i) {
          const float c = u[0][i];
          const float v = rician_u(c, ro[0][i], from_left(c), from_right(c), u[0][i - 1], u[0][i + 1],
                                   from_left(gg[i]), from_right(gg[i]), gg[i - 1], gg[i + 1], rk);
          nu[i] = mask(v, i);
        }
#pragma unroll
        for (int i = 2 * s; i <= FR - 1 - 2 * s; ++i) u[0][i] = nu[i];
      } else {  // FDTD: u[0] = ex, u[1] = ey, u[2] = hz
        const bool physical = (g.variant & OT_VARIANT_FDTD_PHYSICAL) != 0;
        // Ey, Ex need Hz at (x-1, y-1): rows [s, FR-s]
#pragma unroll
        for (int i = s; i <= FR - s; ++i) {
          const float hz = u[2][i];
          const float ey = fdtd_ey(u[1][i], hz, u[2][i - 1]);
          const float ex = fdtd_ex(u[0][i], hz, from_left(hz));
          u[1][i] = mask(ey, i);
          u[0][i] = mask(ex, i);
        }
        // Hz needs the new Ex, Ey at (x+1, y+1): rows [s, FR-1-s]
#pragma unroll
        for (int i = s; i <= FR - 1 - s; ++i) {
          const float ex = u[0][i];
          const float v = fdtd_hz(u[2][i], ex, from_right(ex), u[1][i], u[1][i + 1], physical);
          u[2][i] = mask(v, i);
        }
      }
    }
    if (PHASE && g.phase_limit != 0) return;

    // write back the output tile's interior (rows H .. H + TY - 1)
#pragma unroll
    for (int k = 0; k < NRW; ++k) {
      const auto r = rsrc(f.dst[k]);
#pragma unroll
      for (int i = H; i < FR - H; ++i)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, u[k][i]), r,
                                              (int)(lane_st + (i < H + TY ? row_ld[i] : range)), 0, 0);
    }
  };
  if (interior && g.phase_limit == 0) run(IC<0>{}, IC<0>{});
  else run(IC<1>{}, IC<1>{});
}

template <int ST, int T, int FR>
__global__ void __launch_bounds__(kWaves * 64) k2w(LaunchGeom g, FieldPtrs f) {
  // wave-uniform by construction; readfirstlane keeps the tile origin in SGPRs
  const int wid = __builtin_amdgcn_readfirstlane(dev::xcd_remap((int)blockIdx.x, (int)gridDim.x) * kWaves +
                                                 (int)(threadIdx.x >> 6));
  const int nw = g.grid[0] * g.grid[1];
  if (wid >= nw) return;  // whole wave: no barrier anywhere in this kernel
  tile<ST, T, FR>(g, f, wid % g.grid[0], wid / g.grid[0]);
}

// Batched variant: every wavefront pulls (image, tile) items from the launch's
// queue counter until the queue is empty (each wave leaves once its ticket is
// past the last item), and advances each item's tile with the same update.
template <int ST, int T, int FR>
__global__ void __launch_bounds__(kWaves * 64) k2w_batch(BArgs a, int /*t_eff*/, int launch) {
  for (;;) {
    int it = 0;
    if ((threadIdx.x & 63) == 0) it = atomicAdd(&a.counters[launch], 1);
    it = __builtin_amdgcn_readfirstlane(it);  // lane 0's ticket, for the whole wave
    if (it >= a.n_items) break;
    const int img = a.items[3 * it], tx = a.items[3 * it + 1], ty = a.items[3 * it + 2];
    const BImage im = a.images[img];
    LaunchGeom g;
    g.n[0] = im.w;
    g.n[1] = im.h;
    g.n[2] = 1;
    g.tile[0] = a.tile[0];
    g.tile[1] = a.tile[1];
    g.tile[2] = 1;
    g.grid[0] = g.grid[1] = g.grid[2] = 1;  // unused by the tile update
    g.sy = im.pitch;
    g.sz = (int64_t)im.pitch * im.h;
    g.origin = im.offset;
    g.phase_limit = a.phase_limit;
    g.variant = a.variant;
    for (int i = 0; i < 4; ++i) g.coef[i] = a.coef[i];
    g.wave_rows = FR;
    g.probe_sink = 0;
    g.clocks = nullptr;
    tile<ST, T, FR>(g, a.base, tx, ty);
  }
}

// Register budget: fields held per lane (rw + ro + Rician's g) x FR.
template <int ST>
constexpr int kFields = KW2<ST>::NRW + KW2<ST>::NRO + (ST == OT_RICIAN2D ? 1 : 0);

template <int ST, int T, int FR>
ot_status launch_t(const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  if ((int64_t)(FR + 1) * g.sy * 4 >= (1LL << 30))
    return set_error(OT_EUNSUPPORTED, "2D row of %lld bytes too large for a %d-row footprint", (long long)(g.sy * 4),
                     FR);
  if (prepare) return prepare_lds((const void*)&k2w<ST, T, FR>, 0);
  const int nw = g.grid[0] * g.grid[1];
  const int nblk = (nw + kWaves - 1) / kWaves;
  hipLaunchKernelGGL((k2w<ST, T, FR>), dim3(nblk), dim3(kWaves * 64), 0, (hipStream_t)stream, g, f);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return set_error(OT_EHIP, "2D wave kernel launch failed: %s", hipGetErrorString(err));
  return OT_OK;
}

// Footprint row counts compiled per time step count; k2w_rows() in runtime.hip
// applies the same rule.
template <int ST, int T>
ot_status launch_fr(int fr, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  constexpr int H = KW2<ST>::R * T;
  switch (fr) {
#define FRC(V)                                                                                   \
  case V:                                                                                        \
    if constexpr (V > 2 * H && kFields<ST> * V <= kW2RegBudget) return launch_t<ST, T, V>(g, f, stream, prepare); \
    break;
    FRC(16)
    FRC(24)
    FRC(32)
    FRC(64)
#undef FRC
    default:
      break;
  }
  return set_error(OT_EUNSUPPORTED, "2D wave kernel not compiled for T=%d with %d footprint rows", T, fr);
}

template <int ST>
ot_status launch_st(int t_eff, int fr, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  switch (t_eff) {
    case 1: return launch_fr<ST, 1>(fr, g, f, stream, prepare);
    case 2: return launch_fr<ST, 2>(fr, g, f, stream, prepare);
    case 3: return launch_fr<ST, 3>(fr, g, f, stream, prepare);
    case 4: return launch_fr<ST, 4>(fr, g, f, stream, prepare);
    case 5: return launch_fr<ST, 5>(fr, g, f, stream, prepare);
    case 6: return launch_fr<ST, 6>(fr, g, f, stream, prepare);
    case 7: return launch_fr<ST, 7>(fr, g, f, stream, prepare);
    case 8: return launch_fr<ST, 8>(fr, g, f, stream, prepare);
    default: return set_error(OT_EUNSUPPORTED, "2D wave kernels are compiled for T = 1..8 (got %d)", t_eff);
  }
}

// Batch kernels, compiled for the same (T, rows) pairs as the grid kernels.
template <int ST, int T>
BatchKernel pick_batch_fr(int fr) {
  constexpr int H = KW2<ST>::R * T;
  switch (fr) {
#define FRB(V)                                                                      \
  case V:                                                                           \
    if constexpr (V > 2 * H && kFields<ST> * V <= kW2RegBudget) return &k2w_batch<ST, T, V>; \
    break;
    FRB(16)
    FRB(24)
    FRB(32)
    FRB(64)
#undef FRB
    default:
      break;
  }
  return nullptr;
}

template <int ST>
BatchKernel pick_batch(int t_eff, int fr) {
  switch (t_eff) {
    case 1: return pick_batch_fr<ST, 1>(fr);
    case 2: return pick_batch_fr<ST, 2>(fr);
    case 3: return pick_batch_fr<ST, 3>(fr);
    case 4: return pick_batch_fr<ST, 4>(fr);
    case 5: return pick_batch_fr<ST, 5>(fr);
    case 6: return pick_batch_fr<ST, 6>(fr);
    case 7: return pick_batch_fr<ST, 7>(fr);
    case 8: return pick_batch_fr<ST, 8>(fr);
    default: return nullptr;
  }
}

}  // namespace k2w
}  // namespace ot
