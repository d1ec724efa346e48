// k3w_impl.hpp — wave-pipelined overlapped time-tiled 3D kernels.
//
// A second kernel family for the 3D stencils (jacobi-3d.cpp:120-432,
// poisson-3d.cpp:120-398, gradient-3d.cpp:120-434, cs-3d.cpp:156-430) with no
// LDS and no barriers: every wavefront is an independent tile.
//   * The 64 lanes of a wavefront are 64 consecutive x columns of the
//     footprint; every lane holds a whole column strip of FR footprint rows in
//     registers. x-neighbours come from the adjacent lanes through DPP
//     wave-shift operands (wave_shr:1 / wave_shl:1, folded into the adds),
//     y-neighbours are the lane's own neighbouring rows.
//   * The wave streams its z-chunk plane by plane; stage s (1..T) computes time
//     level s of plane k - s (k - 2s for the 27-point stencil) from a 3-deep
//     register queue of level s-1 planes. Stage s only evaluates the rows
//     s..FR-1-s still needed by later stages, so the redundant halo work is
//     the x-halo of 2T lanes plus a shrinking y-halo.
//   * Output tile = (64 - 2T) x (FR - 2T) points per plane; the level-0 plane
//     is prefetched PD planes ahead through buffer loads whose out-of-range
//     lane offset reads 0 outside the domain; results are stored through
//     buffer stores one iteration later (so load waits never wait for stores).
//   * Chunks away from the z faces run their middle iterations as unrolled
//     blocks without the per-stage range and domain checks (the checks are
//     scalar work that each wave issues in order with its vector work).
//   * When the whole grid is resident at once with at most three waves per
//     SIMD, each wave lowers its issue priority as it completes blocks, so the
//     waves of a SIMD advance together instead of oldest first (which leaves
//     the youngest wave alone, unable to hide memory latency, for the last
//     third of the launch).
// Out-of-domain points are forced to +0 at every level (Dirichlet 0).
#pragma once
#include "internal.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace ot {
namespace k3w {

using namespace dev;

template <int V>
struct IC {
  static constexpr int value = V;
};

template <int ST>
struct KW {
  static constexpr int L = (ST == OT_P3D27) ? 2 : 1;
  static constexpr bool RHS = (ST == OT_P3D7);
  static constexpr int NQ = (L == 2) ? 4 : 3;  // register planes kept per level
  static constexpr int U = (L == 2) ? 4 : 6;   // z-loop unroll (multiple of NQ and of PD)
  static constexpr int PD = (L == 2) ? 2 : 2;  // planes prefetched ahead
};

// Waves per workgroup (the workgroup only bundles independent waves).
constexpr int kWaves = 4;

// CLK: per-wave clock records (OT_VARIANT_DUMP_CLOCKS), a separate
// instantiation so the production kernels carry none of its code.
template <int ST, int T, int FR, bool CLK = false>
__global__ void __launch_bounds__(kWaves * 64) k3w(LaunchGeom g, FieldPtrs f, int balance) {
  uint64_t clk[5] = {0, 0, 0, 0, 0};
  uint64_t rt0 = 0;
  if constexpr (CLK) {
    clk[0] = __builtin_amdgcn_s_memtime();
    rt0 = __builtin_amdgcn_s_memrealtime();
  }
  constexpr int L = KW<ST>::L;
  constexpr bool RHS = KW<ST>::RHS;
  constexpr int NQ = KW<ST>::NQ;
  constexpr int U = KW<ST>::U;
  constexpr int PD = KW<ST>::PD;
  constexpr bool kSteady = L == 1;  // see the steady state below
  static_assert(U % PD == 0 && U % NQ == 0, "unroll must cycle the prefetch ring and the queue");
  static_assert(FR > 2 * T, "footprint rows must exceed the halo");
  const int lane = threadIdx.x & 63;
  const int nblk = (int)gridDim.x;
  // wave-uniform by construction; readfirstlane tells the compiler so, which
  // keeps the tile origin and the buffer descriptors in SGPRs
  const int wid = __builtin_amdgcn_readfirstlane(dev::xcd_remap((int)blockIdx.x, nblk) * kWaves +
                                                 (int)(threadIdx.x >> 6));
  const int nw = g.grid[0] * g.grid[1] * g.grid[2];
  if (wid >= nw) return;  // whole wave: no barrier anywhere in this kernel
  const int bxi = wid % g.grid[0];
  const int byi = (wid / g.grid[0]) % g.grid[1];
  const int bzi = wid / (g.grid[0] * g.grid[1]);
  // tile (TX, TY) with TX + 2T <= 64 lanes and TY + 2T <= FR rows; a plan
  // uses TX = 64 - 2T and FR = TY + 2T, a shortened last launch (smaller T)
  // leaves the spare lanes and rows as dead halo
  const int TX = g.tile[0], TY = g.tile[1], TZ = g.tile[2];
  const int x0 = bxi * TX - T, y0 = byi * TY - T, z0 = bzi * TZ;
  const int gx = x0 + lane;
  const bool in_x = gx >= 0 && gx < g.n[0];
  const bool out_x = lane >= T && lane < T + TX;

  const int64_t zbase = (int64_t)(z0 - T) * g.sz;
  const uint32_t plane_bytes = (uint32_t)(g.sz * 4);
  // chunk = the TZ + 2T planes this wave loads (+3 spare); the host keeps it
  // under 1 GiB, so no offset sum below wraps 32 bits
  const uint32_t range = (uint32_t)(TZ + 2 * T + 3) * plane_bytes;
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)(f.src[0] + g.origin + zbase), (short)0, (int)range, 0x00020000);
  const __amdgpu_buffer_rsrc_t rdst =
      __builtin_amdgcn_make_buffer_rsrc((void*)(f.dst[0] + g.origin + zbase), (short)0, (int)range, 0x00020000);
  __amdgpu_buffer_rsrc_t rro = rsrc;
  if constexpr (RHS)
    rro = __builtin_amdgcn_make_buffer_rsrc((void*)(f.ro[0] + g.origin + zbase), (short)0, (int)range, 0x00020000);

  // Byte offset of row i of plane k = lane term + (row term + plane term).
  // A lane outside the domain has lane term = range, a row or plane outside it
  // row term = range: the offset then reaches the descriptor's range, so the
  // load returns 0 and the store is dropped. The row terms are
  // wave-uniform and live in SGPRs; only the two lane terms take VGPRs.
  const uint32_t lane_ld = in_x ? (uint32_t)gx * 4u : range;
  const uint32_t lane_st = (in_x && out_x) ? (uint32_t)gx * 4u : range;
  uint32_t row_ld[FR], row_st[FR], mxy[FR];
#pragma unroll
  for (int i = 0; i < FR; ++i) {
    const int gy = y0 + i;
    const bool row_in = gy >= 0 && gy < g.n[1];
    const uint32_t o = row_in ? (uint32_t)((int64_t)gy * g.sy * 4) : 0u;
    row_ld[i] = row_in ? o : range;
    row_st[i] = (row_in && i >= T && i < T + TY) ? o : range;
    mxy[i] = (in_x && row_in) ? 0xFFFFFFFFu : 0u;
  }
  // jacobi-3d: the Dirichlet mask rides on the update's multiply. A point in
  // the domain takes sum * 0.143f, the very product of the update; a point
  // outside it takes sum * 0 under the DX9 rule (mul_legacy), +0 whatever the
  // sum (negative, infinite or NaN), as on the reference's ring, where the
  // other stencils AND their result with the mask. This saves one VALU
  // instruction of the eight per evaluation. (The one input on which the
  // rule departs from the reference: a point whose seven terms are all -0.0
  // has the sum -0 and gets +0 where the reference's product is -0.)
  constexpr bool kMulMask = ST == OT_J3D;
  float cmul[kMulMask ? FR : 1];
#pragma unroll
  for (int i = 0; i < (kMulMask ? FR : 1); ++i) cmul[i] = mxy[i] ? 0.143f : 0.0f;
  if (g.phase_limit == 1) return;
  // lane + row term of every footprint row, for the steady-state loads
  constexpr bool kSoff = kSteady && T >= 2;
  uint32_t vld[kSoff ? FR : 1];
#pragma unroll
  for (int i = 0; i < (kSoff ? FR : 1); ++i) vld[i] = lane_ld + row_ld[i];
  // lane + row term of every output row, for the steady-state stores
  uint32_t vst[kSoff ? FR : 1];
#pragma unroll
  for (int i = 0; i < (kSoff ? FR : 1); ++i) vst[i] = (i >= T && i < FR - T) ? lane_st + row_st[i] : 0u;

  float q[T][NQ][FR];
  float fq[T + 1][FR];
  // cs-3d: rsqrt of the forward-difference gradient sum G of level s-1 at the
  // plane stage s evaluates (slot PH % 2) and at the plane before it (the other
  // slot). The four rsqrt arguments of the update (cs-3d.cpp:286-300) are G at
  // the centre and at its z-1, y-1 and x-1 neighbours, each summed in the same
  // order, so one rsqrt per point and level serves all four.
  constexpr bool kShareG = ST == OT_CS3D;
  float rg[kShareG ? T : 1][2][kShareG ? FR : 1];
#pragma unroll
  for (int l = 0; l < (kShareG ? T : 1); ++l)
#pragma unroll
    for (int i = 0; i < (kShareG ? FR : 1); ++i) rg[l][0][i] = rg[l][1][i] = 0.0f;
  // gradient-3d: each squared face difference (c - n)^2 equals (n - c)^2 of the
  // neighbour, bit for bit (IEEE subtraction is antisymmetric), so a point
  // computes only its +x, +y and +z differences and takes the -x one from the
  // lane to its left (DPP), the -y one from the row above and the -z one from
  // the previous plane, summed in the update's order: 3 subtractions and 3
  // products per point instead of 6 and 6. dz keeps, per level, the +z
  // difference of the plane stage s evaluates (slot PH % 2) and of the plane
  // before it (the other slot); it is computed before the plane checks, since
  // plane p+1 needs it even when p itself is not evaluated.
  constexpr bool kShareD = ST == OT_G3D;
  // 1e-4f held in a VGPR (the empty asm hides its value), so that the first
  // add can take the left lane's difference as a DPP operand
  float eps4 = 0.0001f;
  if constexpr (kShareD) asm volatile("" : "+v"(eps4));
  float dz[kShareD ? T : 1][2][kShareD ? FR : 1];
#pragma unroll
  for (int l = 0; l < (kShareD ? T : 1); ++l)
#pragma unroll
    for (int i = 0; i < (kShareD ? FR : 1); ++i) dz[l][0][i] = dz[l][1][i] = 0.0f;
#pragma unroll
  for (int l = 0; l < T; ++l)
#pragma unroll
    for (int a = 0; a < NQ; ++a)
#pragma unroll
      for (int i = 0; i < FR; ++i) q[l][a][i] = 0.0f;
#pragma unroll
  for (int j = 0; j <= T; ++j)
#pragma unroll
    for (int i = 0; i < FR; ++i) fq[j][i] = 0.0f;

  const int kstart = z0 - T;
  const int kload_end = z0 + TZ - 1 + T;
  const int kend = z0 + TZ - 1 + L * T;
  const int zout_end = min(z0 + TZ, g.n[2]);
  const bool do_loads = g.phase_limit != 3;
  const bool do_store = g.phase_limit == 0;
  const int s_end = (g.phase_limit == 2) ? 1 : T;

  float pf[PD][FR], pfr[PD][FR];
  // FULL = 0: a steady-state iteration of an interior chunk (see below). No
  // plane it fetches lies outside the chunk's domain planes, and a plane past
  // kload_end is never read by a steady or later iteration, so its fetch may
  // read plane kload_end instead. The plane offset then goes in the scalar
  // offset, and the vector offset is the lane + row term computed once per
  // wave: no vector instruction per load. A lane or row outside the domain has
  // term >= range, which is out of range with or without the scalar offset.
  auto fetch = [&](int k, auto slotc, auto fullc) __attribute__((always_inline)) {
    constexpr int SL = decltype(slotc)::value;
    constexpr bool FULL = decltype(fullc)::value != 0;
    if constexpr (!FULL && kSoff) {
      const int soff = (int)((uint32_t)(min(k, kload_end) - (z0 - T)) * plane_bytes);
#pragma unroll
      for (int i = 0; i < FR; ++i) {
        pf[SL][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)vld[i], soff, 0));
        if constexpr (RHS)
          pfr[SL][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rro, (int)vld[i], soff, 0));
      }
      return;
    }
    const bool live = FULL ? (do_loads && k >= 0 && k < g.n[2] && k <= kload_end) : (k <= kload_end);
    // plane term = range for a plane that is not loaded: every offset below
    // then reaches the range (the three terms stay under 3 GiB, no wrap)
    const uint32_t pterm = live ? (uint32_t)(k - (z0 - T)) * plane_bytes : range;
#pragma unroll
    for (int i = 0; i < FR; ++i) {
      // the plane term changes every iteration, which keeps the compiler from
      // hoisting FR per-row sums into VGPRs
      const int vo = (int)(lane_ld + (row_ld[i] + pterm));
      pf[SL][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, vo, 0, 0));
      if constexpr (RHS)
        pfr[SL][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rro, vo, 0, 0));
    }
  };
  fetch(kstart, IC<0>{}, IC<1>{});
  if constexpr (PD > 1) fetch(kstart + 1, IC<1>{}, IC<1>{});
  if constexpr (PD > 2) fetch(kstart + 2, IC<2>{}, IC<1>{});
  if constexpr (PD > 3) fetch(kstart + 3, IC<3>{}, IC<1>{});
  if constexpr (PD > 4) fetch(kstart + 4, IC<4>{}, IC<1>{});
  if constexpr (PD > 5) fetch(kstart + 5, IC<5>{}, IC<1>{});

  float outv[FR];
#pragma unroll
  for (int i = 0; i < FR; ++i) outv[i] = 0.0f;
  int out_p = -1;
  auto flush = [&]() __attribute__((always_inline)) {
    const uint32_t pterm = out_p >= 0 ? (uint32_t)(out_p - (z0 - T)) * plane_bytes : range;
#pragma unroll
    for (int i = T; i < FR - T; ++i)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, outv[i]), rdst,
                                            (int)(lane_st + (row_st[i] + pterm)), 0, 0);
  };

  // flush() in a steady-state iteration k: the pending plane is k - 1 - T,
  // whose offset goes in the scalar offset; the vector offset is the lane +
  // row term computed once per wave, so no vector instruction per store. A
  // lane or row outside the domain has vector offset >= range and its store
  // is dropped with or without the scalar offset. rd_steady is the output
  // descriptor, except in the first steady iteration when the checked
  // iteration before it produced no output plane: it is then a descriptor of
  // range 0, through which every store is dropped.
  const __amdgpu_buffer_rsrc_t rnull = __builtin_amdgcn_make_buffer_rsrc((void*)f.dst[0], (short)0, 0, 0x00020000);
  __amdgpu_buffer_rsrc_t rd_steady = rdst;
  auto flush_steady = [&](int k) __attribute__((always_inline)) {
    const int soff = (int)((uint32_t)(k - 1 - z0) * plane_bytes);
#pragma unroll
    for (int i = T; i < FR - T; ++i)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, outv[i]), rd_steady, (int)vst[i], soff, 0);
    rd_steady = rdst;
  };

  auto iter = [&](auto phc, auto fullc, const int k) __attribute__((always_inline)) {
    constexpr int PH = decltype(phc)::value;
    constexpr bool FULL = decltype(fullc)::value != 0;
    constexpr int PS = PH % PD;
    // register queue slots: the plane produced at iteration j sits in slot j mod NQ
    constexpr int QN = PH % NQ;
    constexpr int Q1 = (PH + NQ - 1) % NQ;
    constexpr int Q2 = (PH + NQ - 2) % NQ;
    constexpr int Q3 = (PH + NQ - 3) % NQ;
    if constexpr (RHS) {
#pragma unroll
      for (int j = T; j >= 1; --j)
#pragma unroll
        for (int i = 0; i < FR; ++i) fq[j][i] = fq[j - 1][i];
    }
#pragma unroll
    for (int i = 0; i < FR; ++i) {
      q[0][QN][i] = pf[PS][i];
      if constexpr (RHS) fq[0][i] = pfr[PS][i];
    }
    if constexpr (FULL || !kSoff) flush();
    else flush_steady(k);
    out_p = -1;
    fetch(k + PD, IC<PS>{}, fullc);

#pragma unroll
    for (int s = 1; s <= T; ++s) {
      const int p = k - L * s;
      if constexpr (FULL) {
        if (s > s_end) break;
      }
      if constexpr (kShareG) {
        // G of level s-1 at plane p, rows s-1 .. FR-1-s, before the plane
        // checks: plane p+1's stage needs it as its z-1 term even when p
        // itself is not evaluated (p = lo - 1 reads valid level s-1 planes)
#pragma unroll
        for (int i = s - 1; i <= FR - 1 - s; ++i) {
          const float c = q[s - 1][Q1][i];
          rg[s - 1][PH % 2][i] =
              rsq(((1e-9f + sq(q[s - 1][QN][i] - c)) + sq(q[s - 1][Q1][i + 1] - c)) + sq(from_right(c) - c));
        }
      }
      if constexpr (kShareD) {
#pragma unroll
        for (int i = s; i <= FR - 1 - s; ++i) dz[s - 1][PH % 2][i] = sq(q[s - 1][Q1][i] - q[s - 1][QN][i]);
      }
      if constexpr (FULL) {
        const int lo = z0 - (T - s), hi = z0 + TZ - 1 + (T - s);
        if (p < lo || p > hi) continue;
        if (p < 0 || p >= g.n[2]) {
          // plane outside the domain: level s is 0 there (wave-uniform branch,
          // so the per-point mask below only carries the x/y validity)
          if (s < T) {
#pragma unroll
            for (int i = s; i < FR - s; ++i) q[s][QN][i] = 0.0f;
          }
          continue;
        }
      }
      // gradient-3d: the +y difference of rows s-1 .. FR-1-s (row i-1's is row i's -y)
      float dyb[kShareD ? FR : 1];
      if constexpr (kShareD) {
#pragma unroll
        for (int i = s - 1; i <= FR - 1 - s; ++i) dyb[i] = sq(q[s - 1][Q1][i] - q[s - 1][Q1][i + 1]);
      }
#pragma unroll
      for (int i = s; i < FR - s; ++i) {
        float v;
        if constexpr (kShareD) {
          const float c = q[s - 1][Q1][i];
          const float dxr = sq(c - from_right(c));
          v = c + rsq((((((from_left(dxr) + eps4) + dxr) + dyb[i - 1]) + dyb[i]) + dz[s - 1][(PH + 1) % 2][i]) +
                      dz[s - 1][PH % 2][i]);
        } else if constexpr (L == 1) {
          const float c = q[s - 1][Q1][i];
          const float zm = q[s - 1][Q2][i];
          const float zp = q[s - 1][QN][i];
          const float ym = q[s - 1][Q1][i - 1];
          const float yp = q[s - 1][Q1][i + 1];
          const float xm = from_left(c), xp = from_right(c);
          if constexpr (kMulMask) v = mul_legacy(j3d_sum(xm, xp, ym, yp, zm, zp, c), cmul[i]);
          else if constexpr (ST == OT_G3D) v = g3d(xm, xp, ym, yp, zm, zp, c);
          else if constexpr (ST == OT_CS3D)
            // the rsqrt of G at the centre, at z-1 (the previous plane's
            // slot), at y-1 (row i-1) and at x-1 (the lane to the left)
            v = cs3d_g(c, zm, zp, ym, yp, xm, xp, rg[s - 1][PH % 2][i], rg[s - 1][(PH + 1) % 2][i],
                       rg[s - 1][PH % 2][i - 1], from_left(rg[s - 1][PH % 2][i]));
          else v = p3d7(xm, xp, ym, yp, zm, zp, fq[s][i]);
        } else {
          // poisson-3d.cpp:282-293: planes p-1, p, p+1 of level s-1 were
          // produced 3, 2 and 1 iterations ago
          const float* Pz[3] = {q[s - 1][Q3], q[s - 1][Q2], q[s - 1][Q1]};
          float sum = 0.0f;
          bool first = true;
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
              for (int dz = -1; dz <= 1; ++dz) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                const float t0 = Pz[dz + 1][i + dy];
                const float t = dx < 0 ? from_left(t0) : (dx > 0 ? from_right(t0) : t0);
                sum = first ? t : sum + t;
                first = false;
              }
          v = 26.0f * Pz[1][i] - sum;
        }
        const float nv = kMulMask ? v : __uint_as_float(__float_as_uint(v) & mxy[i]);
        if (s < T) q[s][QN][i] = nv;
        else outv[i] = nv;
      }
      if constexpr (FULL) {
        if (s == T && do_store && p >= z0 && p < zout_end) out_p = p;
      } else if (s == T) {
        out_p = p;
      }
    }
    // keep the scheduler from interleaving consecutive steady iterations (it
    // would hold several iterations' values live at once)
    if constexpr (!FULL) __builtin_amdgcn_sched_barrier(0);
  };

  // Steady state (7-point stencils; the 27-point one keeps one checked loop,
  // which holds fewer registers). In a chunk whose loaded
  // planes z0 - T .. z0 + TZ - 1 + T all lie in the domain, and in a launch
  // with no phase limit, every iteration k in z0 + T .. kend runs all T stages
  // (p = k - s lies in [z0 - T + s, z0 + TZ - 1 + T - s] and in the domain),
  // stores plane k - T (in [z0, z0 + TZ), all inside the domain) and loads
  // plane k + PD exactly when k + PD <= kload_end. Whole unrolled blocks in
  // that range run without the per-stage checks.
  const bool interior = kSteady && g.phase_limit == 0 && z0 - T >= 0 && z0 + TZ - 1 + T < g.n[2];
  const int k_lo = z0 + T;
  const int k_fast_end = interior ? kend : kstart - 1;
  // issue priority 3, 2, 1, 0 for the first, second, third and later blocks
  int blocks_done = 0;
  auto block = [&](auto fullc, const int k) __attribute__((always_inline)) {
    if (balance) {
      const int b = blocks_done++;
      if (b == 0) __builtin_amdgcn_s_setprio(3);
      else if (b == 1) __builtin_amdgcn_s_setprio(2);
      else if (b == 2) __builtin_amdgcn_s_setprio(1);
      else __builtin_amdgcn_s_setprio(0);
    }
    iter(IC<0>{}, fullc, k);
    iter(IC<1>{}, fullc, k + 1);
    iter(IC<2>{}, fullc, k + 2);
    iter(IC<3>{}, fullc, k + 3);
    if constexpr (U > 4) {
      iter(IC<4>{}, fullc, k + 4);
      iter(IC<5>{}, fullc, k + 5);
    }
  };
  int k = kstart;
  // checked blocks until one fits the steady range, then steady blocks; the
  // steady range ends at kend, so only the remainder below is checked again
  for (; k + U - 1 <= kend && !(k >= k_lo && k + U - 1 <= k_fast_end); k += U) block(IC<1>{}, k);
  if constexpr (CLK) clk[1] = __builtin_amdgcn_s_memtime();
  if constexpr (kSoff) {
    // a checked iteration k - 1 >= z0 + T - 1 leaves plane k - 1 - T pending, or none
    rd_steady = out_p >= 0 ? rdst : rnull;
  }
  if constexpr (kSteady)
    for (; k + U - 1 <= k_fast_end; k += U) block(IC<0>{}, k);
  if constexpr (CLK) clk[2] = __builtin_amdgcn_s_memtime();
  if (k <= kend) iter(IC<0>{}, IC<1>{}, k);
  if (k + 1 <= kend) iter(IC<1>{}, IC<1>{}, k + 1);
  if (k + 2 <= kend) iter(IC<2>{}, IC<1>{}, k + 2);
  if constexpr (U > 4) {
    if (k + 3 <= kend) iter(IC<3>{}, IC<1>{}, k + 3);
    if (k + 4 <= kend) iter(IC<4>{}, IC<1>{}, k + 4);
  }
  flush();
  if constexpr (CLK) {
    clk[3] = __builtin_amdgcn_s_memtime();
    uint32_t hw, xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)" : "=s"(hw));
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    clk[4] = __builtin_amdgcn_s_memtime();
    const uint64_t rt1 = __builtin_amdgcn_s_memrealtime();
    if (lane == 0) {  // one vector store per field of the wave's record
      ot_clock_record& r = g.clocks[wid];
      r.block_x = bxi;
      r.block_y = byi;
      r.wave = bzi;
      r.hw_id = hw;
      r.xcc_id = xcc;
      r.valid = 1;
      for (int j = 0; j < 5; ++j) r.clock[j] = clk[j];
      r.realtime[0] = rt0;
      r.realtime[1] = rt1;
    }
  }
}

template <int ST, int T, int FR>
ot_status launch_t(const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  const int64_t plane_bytes = g.sz * 4;
  if ((int64_t)(g.tile[2] + 2 * T + 3) * plane_bytes >= (1LL << 30))
    return set_error(OT_EUNSUPPORTED, "3D plane of %lld bytes too large for a z-chunk of %d planes",
                     (long long)plane_bytes, g.tile[2]);
  // clock records exist for jacobi-3d (the plan refuses them elsewhere)
  constexpr bool kClk = ST == OT_J3D;
  if (g.clocks && !kClk) return set_error(OT_EUNSUPPORTED, "3D clock records exist for jacobi-3d only");
  if (prepare) {
    if constexpr (kClk)
      if (g.clocks) return prepare_lds((const void*)&k3w<ST, T, FR, kClk>, 0);
    return prepare_lds((const void*)&k3w<ST, T, FR>, 0);
  }
  const int nw = g.grid[0] * g.grid[1] * g.grid[2];
  const int nblk = (nw + kWaves - 1) / kWaves;
  // progress balancing pays only when every wave is resident from the start
  // and a SIMD holds at most three of them (measured: with a second round of
  // waves, or five memory-bound waves per SIMD, oldest-first issue is faster)
  const int balance = (!(g.variant & OT_VARIANT_NO_BALANCE) && nw <= 3 * simd_count()) ? 1 : 0;
  if constexpr (kClk) {
    if (g.clocks) {
      hipLaunchKernelGGL((k3w<ST, T, FR, kClk>), dim3(nblk), dim3(kWaves * 64), 0, (hipStream_t)stream, g, f,
                         balance);
      const hipError_t err = hipGetLastError();
      if (err != hipSuccess)
        return set_error(OT_EHIP, "3D wave kernel launch failed: %s", hipGetErrorString(err));
      return OT_OK;
    }
  }
  hipLaunchKernelGGL((k3w<ST, T, FR>), dim3(nblk), dim3(kWaves * 64), 0, (hipStream_t)stream, g, f, balance);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return set_error(OT_EHIP, "3D wave kernel launch failed: %s", hipGetErrorString(err));
  return OT_OK;
}

// Footprint row counts compiled per time tile (registers: 3*T*FR per lane);
// k3w_rows() in runtime.hip applies the same rule.
template <int ST, int T>
ot_status launch_fr(int fr, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  switch (fr) {
#define FRC(V)                                                      \
  case V:                                                           \
    if constexpr (V > 2 * T && 3 * T * V <= kW3RegBudget) return launch_t<ST, T, V>(g, f, stream, prepare); \
    break;
    FRC(8)
    FRC(12)
    FRC(16)
    FRC(20)
    FRC(24)
    FRC(32)
#undef FRC
    default:
      break;
  }
  return set_error(OT_EUNSUPPORTED, "wave-pipelined 3D kernel not compiled for T=%d with %d footprint rows", T, fr);
}

template <int ST>
ot_status launch_st(int t_eff, int fr, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare) {
  switch (t_eff) {
    case 1: return launch_fr<ST, 1>(fr, g, f, stream, prepare);
    case 2: return launch_fr<ST, 2>(fr, g, f, stream, prepare);
    case 3: return launch_fr<ST, 3>(fr, g, f, stream, prepare);
    case 4: return launch_fr<ST, 4>(fr, g, f, stream, prepare);
    case 5: return launch_fr<ST, 5>(fr, g, f, stream, prepare);
    default: return set_error(OT_EUNSUPPORTED, "wave-pipelined 3D kernels are compiled for T = 1..5 (got %d)", t_eff);
  }
}

}  // namespace k3w
}  // namespace ot
