// traits.hpp — per-stencil constants shared by host and device code.
//
// One row per reference driver (opencl/<name>/<name>.cpp). The reporting
// constants reproduce what each driver prints, so the ot-<stencil> drivers keep
// the reference's stdout wire format:
//   * flops: the constant in "Actual GFlop/s" (e.g. jacobi-2d.cpp:1001-1002,
//     jacobi-3d.cpp:915-917, poisson-3d.cpp:883, gradient-3d.cpp:919,
//     fdtd-2d.cpp:885-886, rician-denoise-2d.cpp:844-847, cdsc-tv-update-2d.cpp:569,
//     cs-3d.cpp:795-798)
//   * pp_*: the (numArrays, loads, stores, flops) arguments of
//     ProgramGenerator::printProgramParameters (jacobi-2d.cpp:737,
//     poisson-2d.cpp:564, gradient-2d.cpp:606, fdtd-2d.cpp:643,
//     rician-denoise-2d.cpp:647, cdsc-tv-update-2d.cpp:615, jacobi-1d.cpp:554);
//     3D drivers print their own per-block keys with the constant 7
//     (jacobi-3d.cpp:599-622, poisson-3d.cpp:565-603, gradient-3d.cpp:601-639,
//     cs-3d.cpp:595-630)
//   * model features printed at the end of each driver (phase2_global_loads,
//     compute_per_point, ... e.g. jacobi-2d.cpp:1029-1037); cs-3d prints none,
//     so its row counts its 13 taps and 65 flops.
#pragma once
#include <stdint.h>

namespace ot {

struct StencilTraits {
  const char* name;        // reference driver name
  int dims;
  int radius;              // per-step dependency radius
  int num_rw, num_ro;
  int ref_pad_mult;        // reference padding = mult * T
  int ref_size_rule;       // 0: 1D (rpb), 1: gcd, 2: lcm
  int ref_extra_pad;       // jacobi-1d adds 109 to paddedSize (jacobi-1d.cpp:86)
  int flops;               // "Actual GFlop/s" flops per point
  int pp_arrays, pp_loads, pp_stores, pp_flops;   // printProgramParameters args (0 = 3D/1D style keys)
  double f_p2_gl, f_p2_sh, f_compute, f_p3_sh, f_p4_gs, f_sh_st, f_fields;  // model feature keys
  int default_n;           // reference default -n
  int default_bx, default_by, default_bz;
};

// Indexed by ot_stencil.
static const StencilTraits kTraits[14] = {
  // name               dims r rw ro pad rule xpad fl  arrays/ld/st/fl   p2gl p2sh comp p3sh p4gs shst fld   n    bx  by bz
  {"jacobi-1d",          1, 1, 1, 0, 1, 0, 109,  3,  1, 3, 1, 3,          3, 0, 3, 3, 1, 1, 1,       1024, 16, 1, 1},
  {"jacobi-1d-5pt",      1, 2, 1, 0, 2, 0, 0,    5,  0, 5, 1, 5,          5, 0, 5, 5, 1, 1, 1,       1024, 16, 1, 1},
  {"jacobi-1d-7pt",      1, 3, 1, 0, 3, 0, 0,    7,  0, 7, 1, 7,          7, 0, 7, 7, 1, 1, 1,       1024, 16, 1, 1},
  {"jacobi-2d",          2, 1, 1, 0, 1, 1, 0,    5,  1, 5, 1, 5,          5, 0, 5, 5, 1, 1, 1,       1024, 16, 16, 1},
  {"jacobi-3d",          3, 1, 1, 0, 1, 2, 0,    7,  0, 7, 1, 7,          7, 0, 7, 5, 1, 1, 1,       128, 8, 8, 8},
  {"poisson-2d",         2, 1, 1, 0, 1, 2, 0,    9,  1, 9, 1, 9,          9, 0, 9, 9, 1, 1, 1,       1024, 16, 16, 1},
  {"poisson-3d",         3, 1, 1, 0, 1, 2, 0,    27, 0, 7, 1, 7,          27, 0, 27, 27, 1, 1, 1,    128, 8, 8, 8},
  {"poisson-3d-rhs",     3, 1, 1, 1, 1, 2, 0,    8,  0, 8, 1, 8,          8, 0, 8, 6, 1, 1, 2,       128, 8, 8, 8},
  {"fdtd-2d",            2, 1, 3, 0, 1, 1, 0,    11, 3, 8, 3, 11,         7, 4, 11, 5, 3, 3, 3,      1024, 16, 16, 1},
  {"gradient-2d",        2, 1, 1, 0, 1, 1, 0,    15, 1, 5, 1, 18,         5, 0, 12, 5, 1, 1, 1,      1024, 16, 16, 1},
  {"gradient-3d",        3, 1, 1, 0, 1, 2, 0,    20, 0, 7, 1, 7,          7, 0, 20, 7, 1, 1, 1,      128, 8, 8, 8},
  {"rician-denoise-2d",  2, 2, 1, 1, 1, 1, 0,    42, 3, 6, 3, 44,         13, 0, 42, 13, 1, 1, 1,    1024, 16, 16, 1},
  {"cdsc-tv-update-2d",  2, 1, 1, 2, 1, 1, 0,    59, 1, 7, 1, 59,         7, 0, 59, 7, 1, 1, 1,      1024, 16, 16, 1},
  {"cs-3d",              3, 1, 1, 0, 1, 2, 0,    65, 0, 7, 1, 7,          13, 0, 65, 13, 1, 1, 1,    128, 8, 8, 8},
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

}  // namespace ot
