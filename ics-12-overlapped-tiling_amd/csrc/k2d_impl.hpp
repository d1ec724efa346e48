// k2d_impl.hpp — overlapped time-tiled tile update for the 2D stencils.
//
// Replaces the kernel text emitted by the reference's 2D generators
// (jacobi-2d.cpp:175-514, poisson-2d.cpp:127-364, gradient-2d.cpp:114-393,
// fdtd-2d.cpp:107-443, rician-denoise-2d.cpp:108-447, cdsc-tv-update-2d.cpp:116-412).
//
// One workgroup owns an output tile (tile[0] x tile[1]) and works on the
// footprint tile + 2H (H = radius*T) held in LDS:
//   1. every thread loads its column strip of E points per field from HBM
//      (lanes consecutive in x -> coalesced), masking out-of-domain points to 0;
//   2. t_eff time steps are advanced on chip; the centre and the y-neighbours
//      inside a thread's own strip come from registers, x-neighbours and strip
//      ends from LDS; ping-pong LDS planes give one barrier per step for the
//      single-field stencils (Rician and FDTD need two sub-stages per step);
//      halo points are recomputed redundantly and every point outside the
//      domain is forced to 0 after every step (Dirichlet 0, as the reference's
//      valid* guards do, jacobi-2d.cpp:224-234);
//   3. only the tile interior is written back.
// blockDim = (FX, ceil(FY/E)) where FX/FY are the footprint extents.
#pragma once
#include "internal.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace ot {
namespace {

using namespace dev;

struct Ctx {
  int FX, FY, FYP, H;
  int x0, y0;         // domain coordinates of footprint (0,0)
  int tx, ty;
};

template <int ST>
struct K2 {
  static constexpr int R = (ST == OT_RICIAN2D) ? 2 : 1;
  static constexpr int NRW = (ST == OT_FDTD2D) ? 3 : 1;
  static constexpr int NRO = (ST == OT_TV2D) ? 2 : ((ST == OT_RICIAN2D) ? 1 : 0);
  // LDS planes: single-field = ping-pong (2); Rician = u ping-pong + g (3); FDTD = ex, ey, hz (3)
  static constexpr int NPLANES = (ST == OT_FDTD2D || ST == OT_RICIAN2D) ? 3 : 2;
};

// Timing-probe modes of tile2d (PM): 0 the stencil itself; 1 the same with clock
// stamps; 2 global-memory phases only; 3 arithmetic only; 4 on-chip phase only
// (OT_VARIANT_TIME_ONLY_*, OT_VARIANT_DUMP_CLOCKS).
enum { kProbeNone = 0, kProbeClocks = 1, kProbeGmem = 2, kProbeInstr = 3, kProbeSmem = 4 };

// Advance output tile (bxi, byi) of the domain described by g by t_eff steps.
// Called by the per-grid kernel below, by the batched work-queue kernel and,
// with PM != 0, by the timing-probe kernel, which passes clk[5] for the
// s_memtime stamps P1..P3 (the caller takes P0 and P4).
template <int ST, int E, int PM = kProbeNone>
__device__ __forceinline__ void tile2d(const LaunchGeom& g, const FieldPtrs& f, int t_eff, int bxi, int byi,
                                       float* lds, uint64_t* clk = nullptr) {
  static_assert(PM == kProbeNone || ST == OT_J2D || ST == OT_G2D,
                "the timing probes exist for jacobi-2d and gradient-2d");
  constexpr bool kGmem = PM == kProbeGmem, kInstr = PM == kProbeInstr, kSmem = PM == kProbeSmem;
  auto stamp = [&](int k) __attribute__((always_inline)) {
    if constexpr (PM != kProbeNone) clk[k] = __builtin_amdgcn_s_memtime();
  };
  constexpr int R = K2<ST>::R;
  const int H = R * t_eff;
  const int FX = g.tile[0] + 2 * H;
  const int FY = g.tile[1] + 2 * H;
  const int FYP = (int)blockDim.y * E;
  const int plane = FYP * FX;
  const int x0 = bxi * g.tile[0] - H, y0 = byi * g.tile[1] - H;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int gx = x0 + tx;
  const bool in_x = gx >= 0 && gx < g.n[0];
  const bool edge_x = tx < 1 || tx > FX - 2;  // no x neighbours in LDS

  bool inside[E];
  int64_t gidx[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int gy = y0 + ty * E + i;
    inside[i] = in_x && gy >= 0 && gy < g.n[1];
    gidx[i] = g.origin + (int64_t)gy * g.sy + gx;
  }
  if (g.phase_limit == 1) return;
  stamp(1);

  float u[K2<ST>::NRW][E];
  float ro[K2<ST>::NRO > 0 ? K2<ST>::NRO : 1][E];
  if (g.phase_limit != 3 && !kSmem) {
#pragma unroll
    for (int k = 0; k < K2<ST>::NRW; ++k)
#pragma unroll
      for (int i = 0; i < E; ++i) u[k][i] = inside[i] ? f.src[k][gidx[i]] : 0.0f;
  } else {
#pragma unroll
    for (int k = 0; k < K2<ST>::NRW; ++k)
#pragma unroll
      for (int i = 0; i < E; ++i) u[k][i] = (float)tx;
  }
#pragma unroll
  for (int k = 0; k < K2<ST>::NRO; ++k)
#pragma unroll
    for (int i = 0; i < E; ++i) ro[k][i] = inside[i] ? f.ro[k][gidx[i]] : 0.0f;

  // LDS addressing helpers
  auto at = [&](int p, int y, int x) -> float& { return lds[p * plane + y * FX + x]; };
  if constexpr (kGmem) {
    // global-memory phases only: the loaded footprint goes straight back out
    stamp(2);
    stamp(3);
    if (g.phase_limit != 0) return;
    const bool wx = tx >= H && tx < H + g.tile[0];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int y = ty * E + i;
      if (wx && y >= H && y < H + g.tile[1] && inside[i]) f.dst[0][gidx[i]] = u[0][i];
    }
    return;
  }

  int cur = 0;
  // publish the initial state
  if constexpr (kInstr) {
  } else if constexpr (ST == OT_FDTD2D) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      at(0, ty * E + i, tx) = u[0][i];
      at(1, ty * E + i, tx) = u[1][i];
      at(2, ty * E + i, tx) = u[2][i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) at(0, ty * E + i, tx) = u[0][i];
  }
  if constexpr (!kInstr) __syncthreads();
  stamp(2);

  const int t_begin = (g.phase_limit == 3) ? 1 : 0;
  const int t_end = (g.phase_limit == 2) ? 1 : t_eff;
  for (int t = t_begin; t < t_end; ++t) {
    const bool last = (t == t_end - 1);
    if constexpr (ST == OT_J2D || ST == OT_P2D || ST == OT_G2D || ST == OT_TV2D) {
      const int nxt = cur ^ 1;
      float nu[E];
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int y = ty * E + i;
        nu[i] = u[0][i];
        if (!edge_x && y >= 1 && y <= FY - 2) {
          const float c = u[0][i];
          // arithmetic-only probe: every neighbour outside the strip is the centre
          const float tc = (i > 0) ? u[0][i - 1] : (kInstr ? c : at(cur, y - 1, tx));
          const float bc = (i < E - 1) ? u[0][i + 1] : (kInstr ? c : at(cur, y + 1, tx));
          const float ml = kInstr ? c : at(cur, y, tx - 1);
          const float mr = kInstr ? c : at(cur, y, tx + 1);
          float v;
          if constexpr (ST == OT_J2D) {
            v = j2d(ml, c, mr, tc, bc);
          } else if constexpr (ST == OT_G2D) {
            v = g2d(ml, c, mr, tc, bc);
          } else if constexpr (ST == OT_P2D) {
            v = p2d(at(cur, y - 1, tx - 1), tc, at(cur, y - 1, tx + 1), ml, c, mr,
                    at(cur, y + 1, tx - 1), bc, at(cur, y + 1, tx + 1));
          } else {  // TV
            const float alpha = g.coef[0] != 0.0f ? g.coef[0] : 1.1f;
            v = tv2d(tc, at(cur, y - 1, tx + 1), ml, c, mr, at(cur, y + 1, tx - 1), bc, ro[0][i],
                     ro[1][i], alpha);
          }
          nu[i] = inside[i] ? v : 0.0f;
        }
      }
#pragma unroll
      for (int i = 0; i < E; ++i) u[0][i] = nu[i];
      if (!last && !kInstr) {
#pragma unroll
        for (int i = 0; i < E; ++i) at(nxt, ty * E + i, tx) = nu[i];
        __syncthreads();
      }
      cur = nxt;
    } else if constexpr (ST == OT_RICIAN2D) {
      // plane 2 holds g; planes 0/1 ping-pong u
      RicianK rk;
      rk.sigma2 = g.coef[0] != 0.0f ? g.coef[0] : 0.0025f;  // SIGMA2 (rician-denoise-2d.cpp:159)
      rk.gamma = (g.coef[1] != 0.0f ? g.coef[1] : 0.065f) / rk.sigma2;
      rk.dt = g.coef[2] != 0.0f ? g.coef[2] : 5.0f;
      float gg[E];
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int y = ty * E + i;
        gg[i] = 0.0f;
        if (!edge_x && y >= 1 && y <= FY - 2) {
          const float c = u[0][i];
          const float tc = (i > 0) ? u[0][i - 1] : at(cur, y - 1, tx);
          const float bc = (i < E - 1) ? u[0][i + 1] : at(cur, y + 1, tx);
          const float v = rician_g(c, at(cur, y, tx - 1), at(cur, y, tx + 1), tc, bc);
          gg[i] = inside[i] ? v : 0.0f;
        }
        at(2, y, tx) = gg[i];
      }
      __syncthreads();
      const int nxt = cur ^ 1;
      float nu[E];
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int y = ty * E + i;
        nu[i] = u[0][i];
        if (!edge_x && y >= 1 && y <= FY - 2) {
          const float c = u[0][i];
          const float ut = (i > 0) ? u[0][i - 1] : at(cur, y - 1, tx);
          const float ub = (i < E - 1) ? u[0][i + 1] : at(cur, y + 1, tx);
          const float gt = (i > 0) ? gg[i - 1] : at(2, y - 1, tx);
          const float gb = (i < E - 1) ? gg[i + 1] : at(2, y + 1, tx);
          const float v = rician_u(c, ro[0][i], at(cur, y, tx - 1), at(cur, y, tx + 1), ut, ub,
                                   at(2, y, tx - 1), at(2, y, tx + 1), gt, gb, rk);
          nu[i] = inside[i] ? v : 0.0f;
        }
      }
#pragma unroll
      for (int i = 0; i < E; ++i) u[0][i] = nu[i];
      if (!last) {
#pragma unroll
        for (int i = 0; i < E; ++i) at(nxt, ty * E + i, tx) = nu[i];
        __syncthreads();
      }
      cur = nxt;
    } else {  // FDTD: planes 0 = ex, 1 = ey, 2 = hz (single buffers, two barriers per step)
      // Ex/Ey only need the (x-1, y-1) neighbours of Hz and Hz only the
      // (x+1, y+1) neighbours of the new Ex/Ey, so the first sub-stage runs up
      // to the last footprint column/row; each full step still shrinks the
      // correct region by one point per side.
      const bool physical = (g.variant & OT_VARIANT_FDTD_PHYSICAL) != 0;
      float nex[E], ney[E];
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int y = ty * E + i;
        nex[i] = u[0][i];
        ney[i] = u[1][i];
        if (tx >= 1 && y >= 1 && y <= FY - 1) {
          const float hz = u[2][i];
          const float hz_top = (i > 0) ? u[2][i - 1] : at(2, y - 1, tx);
          const float hz_left = at(2, y, tx - 1);
          const float ey = fdtd_ey(u[1][i], hz, hz_top);
          const float ex = fdtd_ex(u[0][i], hz, hz_left);
          ney[i] = inside[i] ? ey : 0.0f;
          nex[i] = inside[i] ? ex : 0.0f;
        }
      }
      // ex/ey planes are last read by the previous step's Hz update, which a
      // barrier already separates from these writes.
#pragma unroll
      for (int i = 0; i < E; ++i) {
        at(0, ty * E + i, tx) = nex[i];
        at(1, ty * E + i, tx) = ney[i];
        u[0][i] = nex[i];
        u[1][i] = ney[i];
      }
      __syncthreads();
      float nhz[E];
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int y = ty * E + i;
        nhz[i] = u[2][i];
        if (!edge_x && y >= 1 && y <= FY - 2) {
          const float ex_r = at(0, y, tx + 1);
          const float ey_b = (i < E - 1) ? u[1][i + 1] : at(1, y + 1, tx);
          const float v = fdtd_hz(u[2][i], u[0][i], ex_r, u[1][i], ey_b, physical);
          nhz[i] = inside[i] ? v : 0.0f;
        }
      }
#pragma unroll
      for (int i = 0; i < E; ++i) u[2][i] = nhz[i];
      if (!last) {
#pragma unroll
        for (int i = 0; i < E; ++i) at(2, ty * E + i, tx) = nhz[i];
        __syncthreads();
      }
    }
  }
  stamp(3);
  if (g.phase_limit != 0) return;
  // the arithmetic-only and on-chip-only probes store nothing (probe_sink is 0)
  if constexpr (kInstr || kSmem) {
    if (g.probe_sink == 0) return;
  }

  // write back the valid interior of the tile
  const bool wx = tx >= H && tx < H + g.tile[0];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int y = ty * E + i;
    if (wx && y >= H && y < H + g.tile[1] && inside[i]) {
#pragma unroll
      for (int k = 0; k < K2<ST>::NRW; ++k) f.dst[k][gidx[i]] = u[k][i];
    }
  }
}

}  // namespace
}  // namespace ot
