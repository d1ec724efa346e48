// internal.hpp — declarations shared by the engine's translation units.
#pragma once
#include "ot/ot.h"

#include <stdint.h>

namespace ot {

constexpr int kMaxT = 8;          // largest time tile with compiled kernels
constexpr int kMaxThreads = 1024;
constexpr int kMaxLdsBytes = 160 * 1024;

// Record an error message for ot_last_error() and return the status.
ot_status set_error(ot_status st, const char* fmt, ...);

ot_status derive(const ot_params& p, ot_derived& d);

// Fill threads_per_block / lds_bytes for the kernel family that runs p.
ot_status kernel_shape(const ot_params& p, ot_derived& d);

// Geometry of one launch, passed by value to every kernel.
struct LaunchGeom {
  int32_t n[3];          // domain extents
  int32_t tile[3];       // output tile per workgroup
  int32_t grid[3];       // workgroups per axis
  int64_t sy, sz;        // strides (elements)
  int64_t origin;        // element index of interior (0,0,0)
  int32_t phase_limit;
  int32_t variant;
  float coef[4];
  int32_t wave_rows;     // footprint rows of the wave-per-tile kernel (2D, 3D), 0 = LDS kernel,
                         // -1 = 2D row-streaming kernel
  int32_t probe_sink;    // always 0: the timing probes store only when it is not, so their
                         // work cannot be removed as dead code
  ot_clock_record* clocks;  // per-wave clock records (OT_VARIANT_DUMP_CLOCKS), else null
};

// Clock-record slots per workgroup of the timing probes (1024 threads / 64).
constexpr int kClockSlotsPerBlock = 16;

// Variant bits that select the 2D timing-probe kernels (k2p.hip).
constexpr int kProbeBits = OT_VARIANT_TIME_ONLY_GMEM | OT_VARIANT_TIME_ONLY_INSTR | OT_VARIANT_TIME_ONLY_SMEM |
                           OT_VARIANT_DUMP_CLOCKS;


// Register budget (3*T*rows per lane) of the wave-pipelined 3D kernels.
constexpr int kW3RegBudget = 192;

// Footprint rows of the wave-pipelined 3D kernel for this tile, or 0 when the
// tile does not fill a wavefront (x tile + 2T == 64) or has no compiled row count.
int k3w_rows(int T, const int32_t tile[3]);

// Register budget (fields x rows per lane) of the wave-per-tile 2D kernels.
constexpr int kW2RegBudget = 128;

// Footprint rows of the wave-per-tile 2D kernel for this stencil and tile, or 0
// when the tile does not fill a wavefront (x tile + 2*radius*T == 64) or has no
// compiled row count.
int k2w_rows(int stencil, int T, const int32_t tile[3]);

// True when the row-streaming 2D kernel runs this tile: x tile + 2*radius*T
// fills a wavefront and the y tile (the streamed row chunk) is 64 rows or more.
bool k2s_tile(int stencil, int T, const int32_t tile[3]);

// Per-launch field pointers (up to 3 rw fields, 2 ro fields).
struct FieldPtrs {
  const float* src[3];
  float* dst[3];
  const float* ro[2];
};

// Timing-probe launches of the 2D LDS kernel (jacobi-2d, gradient-2d).
ot_status launch_2d_probe(int stencil, int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f,
                          void* stream, bool prepare);

// Batched 2D launches (kbatch.hip): one image of the arena, and the arguments
// of one persistent launch over the (image, tile) work queue.
struct BImage {
  int32_t w, h, pitch, reserved;
  int64_t offset;
};

struct BArgs {
  const BImage* images;
  const int32_t* items;  // 3 ints per item: image, tile x, tile y
  int32_t* counters;     // one per launch
  int32_t n_items;
  int32_t tile[2];
  int32_t phase_limit, variant;
  float coef[4];
  FieldPtrs base;        // arena pointers
};

// (arguments, t_eff, launch index)
using BatchKernel = void (*)(BArgs, int, int);

// Wave-per-tile batch kernel advancing t_eff steps with `rows` footprint rows,
// or nullptr when that instantiation is not compiled (k2w_<stencil>.hip).
BatchKernel batch_wave_kernel(int stencil, int t_eff, int rows);
BatchKernel batch_wave_kernel_j2d(int t_eff, int rows);
BatchKernel batch_wave_kernel_p2d(int t_eff, int rows);
BatchKernel batch_wave_kernel_g2d(int t_eff, int rows);
BatchKernel batch_wave_kernel_tv2d(int t_eff, int rows);
BatchKernel batch_wave_kernel_fdtd2d(int t_eff, int rows);
BatchKernel batch_wave_kernel_rician2d(int t_eff, int rows);

// Launch one overlapped-tile kernel advancing t_eff steps. Implemented per
// dimensionality in k1d.hip / k2d.hip / k3d.hip.
// With prepare=true nothing is launched: the kernel instantiation is looked up
// and its dynamic-LDS limit raised when it needs more than 64 KiB, so that a
// later launch (possibly inside a hipGraph capture) needs no attribute call.
ot_status launch_1d(int stencil, int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f,
                    void* stream, bool prepare);
ot_status launch_2d(int stencil, int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f,
                    void* stream, bool prepare);
ot_status launch_3d(int stencil, int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f,
                    void* stream, bool prepare);

// Per-stencil 3D launchers (k3d_<name>.hip).
ot_status launch_3d_j3d(int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_3d_g3d(int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_3d_p3d7(int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_3d_p3d27(int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_3dw_j3d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2dw_j2d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2ds_j2d(int t_eff, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2ds_p2d(int t_eff, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2ds_g2d(int t_eff, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2ds_tv2d(int t_eff, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2ds_fdtd2d(int t_eff, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2ds_rician2d(int t_eff, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2dw_p2d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2dw_g2d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2dw_tv2d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2dw_fdtd2d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_2dw_rician2d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_3dw_g3d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_3dw_cs3d(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_3dw_p3d7(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);
ot_status launch_3dw_p3d27(int t_eff, int rows, const LaunchGeom& g, const FieldPtrs& f, void* stream, bool prepare);

// Raise the dynamic-LDS limit of a kernel that needs more than 64 KiB.
ot_status prepare_lds(const void* kernel, int lds_bytes);

// SIMDs of the current device (4 per compute unit).
int simd_count();

// Launch shapes (host-side, also used by ot_derive).
void shape_1d(int stencil, int T, int ept, const int32_t tile[3], int* threads, int* lds);
void shape_2d(int stencil, int T, int ept, const int32_t tile[3], int* bx, int* by, int* lds);
void shape_3d(int stencil, int T, int ept, const int32_t tile[3], int* bx, int* by, int* lds);

}  // namespace ot
