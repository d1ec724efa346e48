// k2s_impl.hpp — row-streaming wave kernels for the 2D stencils.
//
// A third kernel family for the 2D stencils (jacobi-2d.cpp:175-514,
// poisson-2d.cpp:127-364, gradient-2d.cpp:114-393, fdtd-2d.cpp:107-443,
// rician-denoise-2d.cpp:108-447, cdsc-tv-update-2d.cpp:116-412), selected for
// tall tiles: every wavefront is an independent (64 - 2H)-column strip that
// streams a chunk of TY rows (H = radius * T), the 2D counterpart of the 3D
// kernels' z-streaming.
//   * Lane = x column; x-neighbours through DPP wave shifts.
//   * Iteration k loads row k of every field (prefetched PD rows ahead through
//     buffer loads whose out-of-range offsets read 0) and then advances every
//     time level by one row: level s is computed lag(s) rows behind the load
//     front from a short register queue of level s-1 rows, so the per-level
//     state is a few registers and loads overlap the arithmetic.
//   * Lags: single-field stencils compute level s at row k - s; FDTD updates
//     Ey, Ex of level s at row k - s + 1 and Hz at row k - s (fdtd-2d.cpp:269-330
//     order: E from the previous Hz, then Hz from the new E); Rician computes g
//     of level s at row k - 2s + 1 and u at row k - 2s (rician-denoise-2d.cpp:
//     271-327: g from u, then u from g and u).
//   * Every value at a point outside the domain is forced to +0 (Dirichlet 0);
//     rows outside the chunk's valid region are computed anyway (their values
//     never reach an output row), which keeps the queues uniform.
//   * Output rows are stored one iteration later (so load waits never wait for
//     stores).
//   * The queues are rings whose slots follow the (unrolled) iteration phase,
//     so advancing a level moves no registers.
//   * A chunk whose whole dependence window lies inside the domain runs a copy
//     of the loop without the per-point Dirichlet masks.
#pragma once
#include "internal.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace ot {
namespace k2s {

using namespace dev;

template <int V>
struct IC {
  static constexpr int value = V;
};

template <int ST>
struct KS {
  static constexpr int R = (ST == OT_RICIAN2D) ? 2 : 1;
  static constexpr int NRW = (ST == OT_FDTD2D) ? 3 : 1;
  static constexpr int NRO = (ST == OT_TV2D) ? 2 : ((ST == OT_RICIAN2D) ? 1 : 0);
  static constexpr int NF = NRW + NRO;  // fields streamed per row
};

constexpr int kWaves = 4;
constexpr int PD = 4;   // rows prefetched ahead
constexpr int U = 12;   // iteration unroll: a multiple of every ring depth below
constexpr int RQR = 12; // ring depth of the read-only rows (TV: T+1, Rician: 2T+1 rows needed)

// Ring slot holding the row produced j iterations before the iteration of phase PH.
template <int D>
__device__ constexpr int slot(int PH, int j) {
  return (PH + 4 * D - j) % D;
}

template <int ST>
constexpr int max_t() {
  return ST == OT_RICIAN2D ? (RQR - 1) / 2 : 8;
}

template <int ST, int T, bool CHK>
__device__ __forceinline__ void run(const LaunchGeom& g, const FieldPtrs& f, const int bxi, const int byi) {
  constexpr int R = KS<ST>::R;
  constexpr int NRW = KS<ST>::NRW;
  constexpr int NRO = KS<ST>::NRO;
  constexpr int NF = KS<ST>::NF;
  constexpr int H = R * T;
  constexpr int LAG = R * T;  // row lag of the last level behind the load front
  const int lane = threadIdx.x & 63;
  const int TX = g.tile[0], TY = g.tile[1];
  const int x0 = bxi * TX - H, y0 = byi * TY;
  const int gx = x0 + lane;
  const bool in_x = gx >= 0 && gx < g.n[0];
  const bool out_x = lane >= H && lane < H + TX;

  // chunk = rows [y0 - H, y0 + TY + H) (+PD spare); the host keeps it under 1 GiB
  const int kfirst = y0 - H;
  const int kload_last = y0 + TY - 1 + H;
  const int klast = y0 + TY - 1 + LAG;
  const uint32_t row_bytes = (uint32_t)(g.sy * 4);
  const uint32_t range = (uint32_t)(TY + 2 * H + PD + 1) * row_bytes;
  const int64_t cbase = g.origin + (int64_t)kfirst * g.sy;
  __amdgpu_buffer_rsrc_t rs[NF > 0 ? NF : 1];
#pragma unroll
  for (int k = 0; k < NRW; ++k)
    rs[k] = __builtin_amdgcn_make_buffer_rsrc((void*)(f.src[k] + cbase), (short)0, (int)range, 0x00020000);
#pragma unroll
  for (int k = 0; k < NRO; ++k)
    rs[NRW + k] = __builtin_amdgcn_make_buffer_rsrc((void*)(f.ro[k] + cbase), (short)0, (int)range, 0x00020000);
  __amdgpu_buffer_rsrc_t rd[NRW];
#pragma unroll
  for (int k = 0; k < NRW; ++k)
    rd[k] = __builtin_amdgcn_make_buffer_rsrc((void*)(f.dst[k] + cbase), (short)0, (int)range, 0x00020000);
  // lane term of a byte offset: out-of-domain lanes get `range` (read 0, store dropped)
  const uint32_t lane_ld = in_x ? (uint32_t)gx * 4u : range;
  const uint32_t lane_st = (in_x && out_x) ? (uint32_t)gx * 4u : range;
  const bool do_loads = g.phase_limit != 3;
  const bool do_store = g.phase_limit == 0;
  const int s_end = (g.phase_limit == 2) ? 1 : T;

  auto row_term = [&](int k, bool live) -> uint32_t {
    return live ? (uint32_t)(k - kfirst) * row_bytes : range;
  };
  float pf[PD][NF];
  auto fetch = [&](int k, auto slotc) {
    constexpr int SL = decltype(slotc)::value;
    const bool live = do_loads && k >= 0 && k < g.n[1] && k <= kload_last;
    const int vo = (int)(lane_ld + row_term(k, live));
#pragma unroll
    for (int q = 0; q < NF; ++q)
      pf[SL][q] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs[q], vo, 0, 0));
  };
  // Dirichlet 0: values at points outside the domain are forced to +0 (only
  // chunks whose dependence window leaves the domain need it)
  auto masked = [&](float v, int p) {
    if constexpr (CHK) return (p >= 0 && p < g.n[1] && in_x) ? v : 0.0f;
    else return v;
  };

  // deferred output row
  float outv[NRW];
#pragma unroll
  for (int q = 0; q < NRW; ++q) outv[q] = 0.0f;
  int out_p = -1;
  auto flush = [&]() {
    const bool live = out_p >= 0;
    const int vo = (int)(lane_st + row_term(out_p, live));
#pragma unroll
    for (int q = 0; q < NRW; ++q)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, outv[q]), rd[q], vo, 0, 0);
  };

  // ---- per-stencil register rings ----
  float lv[T + 1][3];                       // single-field levels
  float rq[NRO > 0 ? NRO : 1][RQR];         // TV / Rician read-only rows
  float fe[T + 1][3][2];                    // FDTD ex, ey, hz per level
  float uq[T + 1][4];                       // Rician u per level
  float gq[T + 1][3];                       // Rician g per level
#pragma unroll
  for (int s = 0; s <= T; ++s) {
#pragma unroll
    for (int j = 0; j < 3; ++j) lv[s][j] = 0.0f, gq[s][j] = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) uq[s][j] = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) fe[s][a][0] = fe[s][a][1] = 0.0f;
  }
#pragma unroll
  for (int q = 0; q < (NRO > 0 ? NRO : 1); ++q)
#pragma unroll
    for (int j = 0; j < RQR; ++j) rq[q][j] = 0.0f;

  fetch(kfirst, IC<0>{});
  fetch(kfirst + 1, IC<1>{});
  fetch(kfirst + 2, IC<2>{});
  fetch(kfirst + 3, IC<3>{});

  RicianK rk;
  if constexpr (ST == OT_RICIAN2D) {
    rk.sigma2 = g.coef[0] != 0.0f ? g.coef[0] : 0.0025f;  // SIGMA2 (rician-denoise-2d.cpp:159)
    rk.gamma = (g.coef[1] != 0.0f ? g.coef[1] : 0.065f) / rk.sigma2;
    rk.dt = g.coef[2] != 0.0f ? g.coef[2] : 5.0f;
  }
  const float alpha = g.coef[0] != 0.0f ? g.coef[0] : 1.1f;
  const bool physical = (g.variant & OT_VARIANT_FDTD_PHYSICAL) != 0;

  auto iter = [&](auto phc, const int k) {
    constexpr int PH = decltype(phc)::value;
    constexpr int PS = PH % PD;
    flush();
    out_p = -1;
    // level 0 <- row k
    float row[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) row[q] = pf[PS][q];
    fetch(k + PD, IC<PS>{});
#pragma unroll
    for (int q = 0; q < NRO; ++q) rq[q][slot<RQR>(PH, 0)] = row[NRW + q];

    if constexpr (ST == OT_J2D || ST == OT_P2D || ST == OT_G2D || ST == OT_TV2D) {
      constexpr int N0 = slot<3>(PH, 0), N1 = slot<3>(PH, 1), N2 = slot<3>(PH, 2);
      lv[0][N0] = row[0];
#pragma unroll
      for (int s = 1; s <= T; ++s) {
        if (s > s_end) break;
        // level s at row p = k - s from level s-1 rows p+1 (newest), p, p-1
        const int p = k - s;
        const float c = lv[s - 1][N1], tc = lv[s - 1][N2], bc = lv[s - 1][N0];
        float v;
        if constexpr (ST == OT_J2D) {
          v = j2d(from_left(c), c, from_right(c), tc, bc);
        } else if constexpr (ST == OT_G2D) {
          v = g2d(from_left(c), c, from_right(c), tc, bc);
        } else if constexpr (ST == OT_P2D) {
          v = p2d(from_left(tc), tc, from_right(tc), from_left(c), c, from_right(c), from_left(bc), bc,
                  from_right(bc));
        } else {
          v = tv2d(tc, from_right(tc), from_left(c), c, from_right(c), from_left(bc), bc,
                   rq[0][slot<RQR>(PH, s)], rq[1][slot<RQR>(PH, s)], alpha);
        }
        v = masked(v, p);
        lv[s][N0] = v;
        if (s == T && do_store && p >= y0 && p < y0 + TY && p < g.n[1]) {
          outv[0] = v;
          out_p = p;
        }
      }
    } else if constexpr (ST == OT_RICIAN2D) {
      uq[0][slot<4>(PH, 0)] = row[0];
#pragma unroll
      for (int s = 1; s <= T; ++s) {
        if (s > s_end) break;
        // g of level s at row k - 2s + 1 from u of level s-1 rows +1, 0, -1 (ring ages 0, 1, 2)
        const int pg = k - 2 * s + 1;
        {
          const float c = uq[s - 1][slot<4>(PH, 1)];
          const float gv = rician_g(c, from_left(c), from_right(c), uq[s - 1][slot<4>(PH, 2)],
                                    uq[s - 1][slot<4>(PH, 0)]);
          gq[s][slot<3>(PH, 0)] = masked(gv, pg);
        }
        // u of level s at row k - 2s from u of level s-1 rows +1, 0, -1 (ages 1, 2, 3) and g
        const int pu = k - 2 * s;
        const float c = uq[s - 1][slot<4>(PH, 2)];
        const float gc = gq[s][slot<3>(PH, 1)];
        const float v = masked(rician_u(c, rq[0][slot<RQR>(PH, 2 * s)], from_left(c), from_right(c),
                                        uq[s - 1][slot<4>(PH, 3)], uq[s - 1][slot<4>(PH, 1)], from_left(gc),
                                        from_right(gc), gq[s][slot<3>(PH, 2)], gq[s][slot<3>(PH, 0)], rk),
                               pu);
        uq[s][slot<4>(PH, 0)] = v;
        if (s == T && do_store && pu >= y0 && pu < y0 + TY && pu < g.n[1]) {
          outv[0] = v;
          out_p = pu;
        }
      }
    } else {  // FDTD
      constexpr int E0 = slot<2>(PH, 0), E1 = slot<2>(PH, 1);
#pragma unroll
      for (int a = 0; a < 3; ++a) fe[0][a][E0] = row[a];
#pragma unroll
      for (int s = 1; s <= T; ++s) {
        if (s > s_end) break;
        // Ey, Ex of level s at row r = k - s + 1: level s-1 Hz rows r (newest), r - 1;
        // level s-1 Ex, Ey at row r: the newest row of the loaded level, one older above it
        const int r = k - s + 1;
        const int ie = (s == 1) ? E0 : E1;
        const float hz = fe[s - 1][2][E0];
        const float ey = masked(fdtd_ey(fe[s - 1][1][ie], hz, fe[s - 1][2][E1]), r);
        const float ex = masked(fdtd_ex(fe[s - 1][0][ie], hz, from_left(hz)), r);
        fe[s][0][E0] = ex;
        fe[s][1][E0] = ey;
        // Hz of level s at row r - 1 from the new Ex, Ey (rows r - 1 = older slot, r = newest)
        const float exc = fe[s][0][E1];
        const float hzv =
            masked(fdtd_hz(fe[s - 1][2][E1], exc, from_right(exc), fe[s][1][E1], ey, physical), r - 1);
        fe[s][2][E0] = hzv;
        if (s == T && do_store && r - 1 >= y0 && r - 1 < y0 + TY && r - 1 < g.n[1]) {
          outv[0] = exc;
          outv[1] = fe[s][1][E1];
          outv[2] = hzv;
          out_p = r - 1;
        }
      }
    }
  };

  int k = kfirst;
  for (; k + U - 1 <= klast; k += U) {
    iter(IC<0>{}, k);
    iter(IC<1>{}, k + 1);
    iter(IC<2>{}, k + 2);
    iter(IC<3>{}, k + 3);
    iter(IC<4>{}, k + 4);
    iter(IC<5>{}, k + 5);
    iter(IC<6>{}, k + 6);
    iter(IC<7>{}, k + 7);
    iter(IC<8>{}, k + 8);
    iter(IC<9>{}, k + 9);
    iter(IC<10>{}, k + 10);
    iter(IC<11>{}, k + 11);
  }
  if (k <= klast) iter(IC<0>{}, k);
  if (k + 1 <= klast) iter(IC<1>{}, k + 1);
  if (k + 2 <= klast) iter(IC<2>{}, k + 2);
  if (k + 3 <= klast) iter(IC<3>{}, k + 3);
  if (k + 4 <= klast) iter(IC<4>{}, k + 4);
  if (k + 5 <= klast) iter(IC<5>{}, k + 5);
  if (k + 6 <= klast) iter(IC<6>{}, k + 6);
  if (k + 7 <= klast) iter(IC<7>{}, k + 7);
  if (k + 8 <= klast) iter(IC<8>{}, k + 8);
  if (k + 9 <= klast) iter(IC<9>{}, k + 9);
  if (k + 10 <= klast) iter(IC<10>{}, k + 10);
  flush();
}

template <int ST, int T>
__global__ void __launch_bounds__(kWaves * 64) k2s(LaunchGeom g, FieldPtrs f) {
  constexpr int H = KS<ST>::R * T;
  const int wid = __builtin_amdgcn_readfirstlane(dev::xcd_remap((int)blockIdx.x, (int)gridDim.x) * kWaves +
                                                 (int)(threadIdx.x >> 6));
  const int nw = g.grid[0] * g.grid[1];
  if (wid >= nw) return;  // whole wave: no barrier anywhere in this kernel
  if (g.phase_limit == 1) return;
  const int bxi = wid % g.grid[0], byi = wid / g.grid[0];
  const int x0 = bxi * g.tile[0] - H, y0 = byi * g.tile[1];
  // every point of the chunk's dependence window inside the domain: no masks
  const bool interior = x0 >= 0 && x0 + 63 < g.n[0] && y0 - H >= 0 && y0 + g.tile[1] - 1 + H < g.n[1];
  if (interior)
    run<ST, T, false>(g, f, bxi, byi);
  else
    run<ST, T, true>(g, f, bxi, byi);
}

template <int ST, int T>
ot_status launch_t(const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  constexpr int H = KS<ST>::R * T;
  if constexpr (T > max_t<ST>()) {
    return set_error(OT_EUNSUPPORTED, "2D row-streaming kernel not compiled for T=%d of this stencil", T);
  }
  if ((int64_t)(g.tile[1] + 2 * H + PD + 1) * g.sy * 4 >= (1LL << 30))
    return set_error(OT_EUNSUPPORTED, "2D row chunk of %d rows x %lld bytes too large", g.tile[1],
                     (long long)(g.sy * 4));
  if constexpr (T <= max_t<ST>()) {
    if (prepare) return prepare_lds((const void*)&k2s<ST, T>, 0);
    const int nw = g.grid[0] * g.grid[1];
    const int nblk = (nw + kWaves - 1) / kWaves;
    hipLaunchKernelGGL((k2s<ST, T>), dim3(nblk), dim3(kWaves * 64), 0, (hipStream_t)stream, g, f);
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess)
      return set_error(OT_EHIP, "2D row-streaming kernel launch failed: %s", hipGetErrorString(err));
    return OT_OK;
  }
}

template <int ST>
ot_status launch_st(int t_eff, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  switch (t_eff) {
    case 1: return launch_t<ST, 1>(g, f, stream, prepare);
    case 2: return launch_t<ST, 2>(g, f, stream, prepare);
    case 3: return launch_t<ST, 3>(g, f, stream, prepare);
    case 4: return launch_t<ST, 4>(g, f, stream, prepare);
    case 5: return launch_t<ST, 5>(g, f, stream, prepare);
    case 6: return launch_t<ST, 6>(g, f, stream, prepare);
    case 7: return launch_t<ST, 7>(g, f, stream, prepare);
    case 8: return launch_t<ST, 8>(g, f, stream, prepare);
    default: return set_error(OT_EUNSUPPORTED, "2D row-streaming kernels are compiled for T = 1..8 (got %d)", t_eff);
  }
}

}  // namespace k2s
}  // namespace ot
