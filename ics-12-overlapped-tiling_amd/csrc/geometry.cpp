// geometry.cpp — tile geometry and the reference's derived parameters.
//
// Host-only code (no HIP calls): ot_derive can run on a machine without a GPU.
//
// Two geometries:
//  * OT_GEOM_REFERENCE restates GeneratorParams::computeDerived of each driver:
//      padding P = mult*T                       (jacobi-2d.cpp:86, jacobi-1d-5pt.cpp:61)
//      compsPerBlock = (bx, by*ept, bz)         (jacobi-2d.cpp:87-88; 1D: bx*ept, jacobi-1d.cpp:78)
//      realPerBlock = comps - 2*r*(T-1)         (jacobi-2d.cpp:89-90, jacobi-1d-5pt.cpp:63)
//      sizeLCM = gcd (jacobi-2d.cpp:91, fdtd-2d.cpp:79, gradient-2d.cpp:86,
//                rician-denoise-2d.cpp:80, cdsc-tv-update-2d.cpp:79),
//                lcm (poisson-2d.cpp:80, jacobi-3d.cpp:87-89), rpb (1D)
//      realSize = (n / sizeLCM) * sizeLCM       (truncation kept)
//      numBlocks = realSize / rpb               (floor, as printed)
//      paddedSize = realSize + 2P (+109 in jacobi-1d, jacobi-1d.cpp:86)
//    The domain this engine computes is the full realSize^d interior — what the
//    reference's CPU loop computes — so the workgroup grid is ceil(realSize/rpb)
//    (the reference's floor leaves up to rpb-1 rows unwritten with gcd, a defect
//    that is not reproduced).
//  * OT_GEOM_TILE (default): (bx, by*ept, bz) is the output tile of a workgroup,
//    the halo radius*T is added around it, the grid is ceil(n/tile) with partial
//    edge tiles, and the field layout is a dense interior with no padding ring
//    and the x pitch rounded up to 16 elements (64 B) so rows start aligned.
#include "ot/ot.h"
#include "traits.hpp"
#include "internal.hpp"

#include <cstdio>
#include <cstring>
#include <numeric>

namespace ot {

static int64_t gcd64(int64_t a, int64_t b) { return std::gcd(a, b); }
static int64_t lcm64(int64_t a, int64_t b) { return a / std::gcd(a, b) * b; }

// Reference per-block report keys (ProgramGenerator.hpp:19-67 for 2D/1D-3pt,
// the 3D drivers' own blocks e.g. jacobi-3d.cpp:596-622, 1D-5/7pt
// jacobi-1d-5pt.cpp:400-429). Uses the reference's int32 arithmetic widths
// only where the values stay small; the sweep configurations never overflow.
static void fill_report_keys(const StencilTraits& tr, const ot_params& p, ot_derived& d) {
  const int64_t bx = p.block_x, by = p.block_y, bz = p.block_z, e = p.elems_per_thread;
  const int64_t T = p.time_tile;
  const int64_t rx = d.real_per_block[0], ry = d.real_per_block[1], rz = d.real_per_block[2];
  if (tr.dims == 3) {
    const int64_t threads = bx * by * bz * e;
    d.shared_size = (int32_t)((bx + 2) * (by * e + 2) * (bz + 2) * 4);
    d.total_fp_per_block = (double)(threads * p.time_steps * 7);
    d.useful_fp_per_block = (double)(7 * rx * ry * rz * p.time_steps);
    d.global_loads_per_block = threads * 7;
    d.global_stores_per_block = threads;
    d.shared_loads_per_block = threads * 7 * (T - 1);
    d.shared_stores_per_block = threads * (T - 1);
  } else if (tr.pp_arrays == 0) {  // jacobi-1d-5pt / -7pt
    const int64_t threads = bx * e;
    d.shared_size = (int32_t)((bx * e + 2 * tr.radius) * 4);
    d.total_fp_per_block = (double)(threads * p.time_steps * tr.pp_flops);
    d.useful_fp_per_block = (double)(tr.pp_flops * rx * p.time_steps);
    d.global_loads_per_block = threads * tr.pp_loads;
    d.global_stores_per_block = threads;
    d.shared_loads_per_block = threads * tr.pp_loads * (T - 1);
    d.shared_stores_per_block = threads * (T - 1);
  } else {  // printProgramParameters (2D drivers and jacobi-1d)
    const int64_t sx = (tr.dims == 1) ? bx * e + 2 : bx + 2;
    const int64_t sy = (tr.dims == 1) ? 1 : by * e + 2;
    const int64_t byy = (tr.dims == 1) ? 1 : by;
    const int64_t threads = bx * byy * e;
    d.shared_size = (int32_t)(sx * sy * tr.pp_arrays * 4);
    d.total_fp_per_block = (double)(threads * T * tr.pp_flops);
    d.useful_fp_per_block = (double)(rx * ((tr.dims == 1) ? 1 : ry) * T * tr.pp_flops);
    d.global_loads_per_block = threads * tr.pp_loads;
    d.global_stores_per_block = threads * tr.pp_stores;
    d.shared_loads_per_block = threads * tr.pp_loads * (T - 1);
    d.shared_stores_per_block = threads * tr.pp_stores * (T - 1);
  }
  {
    const int64_t threads = bx * (tr.dims > 1 ? by : 1) * (tr.dims > 2 ? bz : 1) * e;
    const int64_t fl = tr.dims == 3 ? 7 : tr.pp_flops;
    d.total_fp_steps = (double)(threads * p.time_steps * fl);
    d.useful_fp_steps = (double)(fl * rx * (tr.dims > 1 ? ry : 1) * (tr.dims > 2 ? rz : 1) * p.time_steps);
  }
  d.useful_ratio = d.total_fp_per_block > 0 ? d.useful_fp_per_block / d.total_fp_per_block : 0.0;
}

ot_status derive(const ot_params& p, ot_derived& d) {
  std::memset(&d, 0, sizeof(d));
  if (p.stencil < 0 || p.stencil >= OT_NUM_STENCILS)
    return set_error(OT_EINVAL, "unknown stencil id %d", p.stencil);
  const StencilTraits& tr = kTraits[p.stencil];
  const int dims = tr.dims;
  if (p.time_tile < 1) return set_error(OT_EINVAL, "time tile size must be >= 1 (got %d)", p.time_tile);
  if (p.time_tile > kMaxT) return set_error(OT_EUNSUPPORTED, "time tile size %d exceeds %d", p.time_tile, kMaxT);
  if (p.time_steps < 0) return set_error(OT_EINVAL, "time steps must be >= 0");
  if (p.elems_per_thread < 1) return set_error(OT_EINVAL, "elements per thread must be >= 1");
  if (p.block_x < 1 || (dims > 1 && p.block_y < 1) || (dims > 2 && p.block_z < 1))
    return set_error(OT_EINVAL, "block sizes must be >= 1");
  if (p.geometry != OT_GEOM_TILE && p.geometry != OT_GEOM_REFERENCE)
    return set_error(OT_EINVAL, "unknown geometry %d", p.geometry);
  if (p.phase_limit < 0 || p.phase_limit > 3) return set_error(OT_EINVAL, "phase limit must be 0..3");

  const int T = p.time_tile, r = tr.radius, e = p.elems_per_thread;
  d.dims = dims;
  d.radius = r;
  d.halo = r * T;
  d.num_rw = tr.num_rw;
  d.num_ro = tr.num_ro;
  d.launches = (int32_t)ceil_div(p.time_steps, T);

  int64_t comps[3] = {p.block_x, dims > 1 ? (int64_t)p.block_y * e : 1, dims > 2 ? p.block_z : 1};
  if (dims == 1) comps[0] = (int64_t)p.block_x * e;
  for (int a = 0; a < 3; ++a) d.comps_per_block[a] = (int32_t)comps[a];

  if (p.geometry == OT_GEOM_REFERENCE) {
    // The reference drivers are cubic: one problem size for every axis.
    if (p.problem_size < 1) return set_error(OT_EINVAL, "problem size must be >= 1");
    const int P = tr.ref_pad_mult * T;
    // realPerBlock = comps - 2*mult*(T-1): the reference shrinks by its padding
    // multiplier (1D-5pt: 2*2*(T-1), jacobi-1d-5pt.cpp:63), i.e. by its own
    // notion of radius, which for Rician is 1 (a defect: the real dependency
    // radius is 2). Derived keys follow the reference exactly.
    int64_t rpb[3] = {1, 1, 1};
    for (int a = 0; a < dims; ++a) rpb[a] = comps[a] - 2LL * tr.ref_pad_mult * (T - 1);
    for (int a = 0; a < dims; ++a)
      if (rpb[a] < 1)
        return set_error(OT_EINVAL,
                         "Consistency error! real per block %lld < 1 on axis %d (block %lld, T %d)",
                         (long long)rpb[a], a, (long long)comps[a], T);
    int64_t lcm = rpb[0];
    if (tr.ref_size_rule == 1) lcm = gcd64(rpb[0], rpb[1]);
    else if (tr.ref_size_rule == 2) {
      lcm = lcm64(rpb[0], rpb[1]);
      if (dims == 3) lcm = lcm64(rpb[2], lcm);
    }
    const int64_t real = (p.problem_size / lcm) * lcm;
    if (real < 1) return set_error(OT_EINVAL, "Consistency error! realSize < 1 (n %d, sizeLCM %lld)",
                                   p.problem_size, (long long)lcm);
    d.padding = P;
    d.size_lcm = (int32_t)lcm;
    for (int a = 0; a < 3; ++a) {
      const bool used = a < dims;
      d.real_per_block[a] = (int32_t)rpb[a];
      d.real_size[a] = used ? (int32_t)real : 1;
      d.padded_size[a] = used ? (int32_t)(real + 2 * P) : 1;
      d.num_blocks[a] = used ? (int32_t)(real / rpb[a]) : 1;
      d.grid_blocks[a] = used ? (int32_t)ceil_div(real, rpb[a]) : 1;
      d.tile[a] = used ? (int32_t)rpb[a] : 1;
    }
    if (dims == 1) d.padded_size[0] += tr.ref_extra_pad;
    d.stride_y = d.padded_size[0];
    d.stride_z = (int64_t)d.padded_size[0] * d.padded_size[1];
    d.origin = (int64_t)P + (dims > 1 ? (int64_t)P * d.stride_y : 0) + (dims > 2 ? (int64_t)P * d.stride_z : 0);
    d.field_elems = d.stride_z * d.padded_size[2];
  } else {
    int64_t n[3] = {p.nx > 0 ? p.nx : p.problem_size, p.ny > 0 ? p.ny : p.problem_size,
                    p.nz > 0 ? p.nz : p.problem_size};
    for (int a = dims; a < 3; ++a) n[a] = 1;
    for (int a = 0; a < dims; ++a)
      if (n[a] < 1) return set_error(OT_EINVAL, "grid extent on axis %d must be >= 1", a);
    d.padding = 0;
    d.size_lcm = 1;
    for (int a = 0; a < 3; ++a) {
      d.real_per_block[a] = (int32_t)comps[a];
      d.real_size[a] = (int32_t)n[a];
      d.tile[a] = (int32_t)comps[a];
      d.num_blocks[a] = (int32_t)ceil_div(n[a], comps[a]);
      d.grid_blocks[a] = d.num_blocks[a];
    }
    d.padded_size[0] = dims == 1 ? (int32_t)n[0] : (int32_t)round_up(n[0], 16);
    d.padded_size[1] = (int32_t)n[1];
    d.padded_size[2] = (int32_t)n[2];
    d.stride_y = d.padded_size[0];
    d.stride_z = (int64_t)d.padded_size[0] * d.padded_size[1];
    d.origin = 0;
    d.field_elems = d.stride_z * d.padded_size[2];
  }
  fill_report_keys(tr, p, d);
  // Kernel-specific launch shape (threads, LDS) of this engine.
  return kernel_shape(p, d);
}

}  // namespace ot

extern "C" {

ot_status ot_stencil_query(int32_t stencil, ot_stencil_info* info) {
  if (!info) return ot::set_error(OT_EINVAL, "null info");
  if (stencil < 0 || stencil >= OT_NUM_STENCILS) return ot::set_error(OT_EINVAL, "unknown stencil id %d", stencil);
  const ot::StencilTraits& t = ot::kTraits[stencil];
  info->stencil = stencil;
  info->dims = t.dims;
  info->radius = t.radius;
  info->num_rw = t.num_rw;
  info->num_ro = t.num_ro;
  info->flops_per_point = t.flops;
  info->loads_per_point = (int32_t)t.f_p2_gl;
  info->ref_padding_mult = t.ref_pad_mult;
  info->name = t.name;
  return OT_OK;
}

ot_status ot_stencil_from_name(const char* name, int32_t* stencil) {
  if (!name || !stencil) return ot::set_error(OT_EINVAL, "null argument");
  for (int i = 0; i < OT_NUM_STENCILS; ++i)
    if (std::strcmp(name, ot::kTraits[i].name) == 0) {
      *stencil = i;
      return OT_OK;
    }
  return ot::set_error(OT_EINVAL, "unknown stencil name '%s'", name);
}

ot_status ot_params_default(int32_t stencil, ot_params* p) {
  if (!p) return ot::set_error(OT_EINVAL, "null params");
  if (stencil < 0 || stencil >= OT_NUM_STENCILS) return ot::set_error(OT_EINVAL, "unknown stencil id %d", stencil);
  const ot::StencilTraits& t = ot::kTraits[stencil];
  std::memset(p, 0, sizeof(*p));
  // Defaults of each driver's boost::program_options block (e.g.
  // jacobi-2d.cpp:566-607, jacobi-3d.cpp:481-515).
  p->stencil = stencil;
  p->problem_size = t.default_n;
  p->time_steps = 64;
  p->time_tile = 1;
  p->block_x = t.default_bx;
  p->block_y = t.default_by;
  p->block_z = t.default_bz;
  p->elems_per_thread = 1;
  p->geometry = OT_GEOM_TILE;
  return OT_OK;
}

ot_status ot_derive(const ot_params* params, ot_derived* derived) {
  if (!params || !derived) return ot::set_error(OT_EINVAL, "null argument");
  return ot::derive(*params, *derived);
}

}  // extern "C"
