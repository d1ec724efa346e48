// runtime.hip — plans, the host time loop, hipGraph capture, HIP helpers.
//
// Replaces the reference's OpenCL runtime shim and host loop:
//   CLContext (opencl/common/src/CLContext.cpp:11-56)      -> HIP runtime on the current device
//   XGenerator::generate + Program::build                   -> ot_plan_create (kernels are
//     (jacobi-2d.cpp:162-173, :741-760)                        precompiled for gfx950)
//   time loop + ping-pong + events (jacobi-2d.cpp:888-948)  -> ot_run (optionally one hipGraph)
//   throwOnError                                            -> ot_status + ot_last_error
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "internal.hpp"
#include "traits.hpp"

namespace ot {

static thread_local char g_err[512] = "";

ot_status set_error(ot_status st, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return st;
}

#define OT_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return set_error(OT_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// VGPR count of the kernel most recently prepared on this thread.
static thread_local int g_prepared_regs = 0;

ot_status prepare_lds(const void* kernel, int lds_bytes) {
  if (lds_bytes > 65536)
    OT_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes));
  hipFuncAttributes attr;
  if (hipFuncGetAttributes(&attr, kernel) == hipSuccess) g_prepared_regs = attr.numRegs;
  return OT_OK;
}

int simd_count() {
  static std::atomic<int> cached[64];  // per device ordinal, 0 = not queried yet
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 0;
  if (dev < 64) {
    const int c = cached[dev].load(std::memory_order_relaxed);
    if (c > 0) return c;
  }
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return 0;  // unknown: callers then take no decision that depends on it
  if (dev < 64) cached[dev].store(4 * cus, std::memory_order_relaxed);
  return 4 * cus;
}

void shape_3d(int stencil, int T, int ept, const int32_t tile[3], int* bx, int* by, int* lds) {
  const int NS = (stencil == OT_P3D27) ? 4 : 2;
  const int FX = tile[0] + 2 * T, FY = tile[1] + 2 * T;
  *bx = FX;
  *by = (FY + ept - 1) / ept;
  // T levels x NS ring planes + one guard row (+1) before and after
  *lds = (T * NS * (*by) * ept * FX + 2 * (FX + 1)) * (int)sizeof(float);
}

int k3w_rows(int T, const int32_t tile[3]) {
  if (tile[0] + 2 * T != 64) return 0;
  const int fr = tile[1] + 2 * T;
  static const int kRows[] = {8, 12, 16, 20, 24, 32};
  for (int r : kRows)
    if (r == fr && fr > 2 * T && 3 * T * fr <= kW3RegBudget) return fr;
  return 0;
}

int k2w_rows(int stencil, int T, const int32_t tile[3]) {
  const int R = (stencil == OT_RICIAN2D) ? 2 : 1;
  const int fields = (stencil == OT_FDTD2D || stencil == OT_RICIAN2D || stencil == OT_TV2D) ? 3 : 1;
  if (tile[0] + 2 * R * T != 64) return 0;
  const int fr = tile[1] + 2 * R * T;
  static const int kRows[] = {16, 24, 32, 64};
  for (int r : kRows)
    if (r == fr && fr > 2 * R * T && fields * fr <= kW2RegBudget) return fr;
  return 0;
}

bool k2s_tile(int stencil, int T, const int32_t tile[3]) {
  const int R = (stencil == OT_RICIAN2D) ? 2 : 1;
  // Rician keeps 2T+1 rows of f in a 12-slot ring: compiled for T <= 5
  if (stencil == OT_RICIAN2D && T > 5) return false;
  return tile[0] + 2 * R * T == 64 && tile[1] >= 64;
}

ot_status launch_3d(int stencil, int t_eff, int ept, const LaunchGeom& g, const FieldPtrs& f,
                    void* stream, bool prepare) {
  if (g.wave_rows > 0) {
    switch (stencil) {
      case OT_J3D: return launch_3dw_j3d(t_eff, g.wave_rows, g, f, stream, prepare);
      case OT_G3D: return launch_3dw_g3d(t_eff, g.wave_rows, g, f, stream, prepare);
      case OT_P3D7: return launch_3dw_p3d7(t_eff, g.wave_rows, g, f, stream, prepare);
      case OT_P3D27: return launch_3dw_p3d27(t_eff, g.wave_rows, g, f, stream, prepare);
      case OT_CS3D: return launch_3dw_cs3d(t_eff, g.wave_rows, g, f, stream, prepare);
      default: return set_error(OT_EINVAL, "stencil %d is not a 3D stencil", stencil);
    }
  }
  switch (stencil) {
    case OT_J3D: return launch_3d_j3d(t_eff, ept, g, f, stream, prepare);
    case OT_G3D: return launch_3d_g3d(t_eff, ept, g, f, stream, prepare);
    case OT_P3D7: return launch_3d_p3d7(t_eff, ept, g, f, stream, prepare);
    case OT_P3D27: return launch_3d_p3d27(t_eff, ept, g, f, stream, prepare);
    case OT_CS3D: return set_error(OT_EUNSUPPORTED, "cs-3d has no LDS kernel; use a wave-pipelined tile");
    default: return set_error(OT_EINVAL, "stencil %d is not a 3D stencil", stencil);
  }
}

ot_status kernel_shape(const ot_params& p, ot_derived& d) {
  const int dims = d.dims;
  int threads = 0, lds = 0;
  if (dims == 1) {
    shape_1d(p.stencil, p.time_tile, p.elems_per_thread, d.tile, &threads, &lds);
  } else {
    int bx = 0, by = 0;
    if (dims == 2 && (k2w_rows(p.stencil, p.time_tile, d.tile) > 0 || k2s_tile(p.stencil, p.time_tile, d.tile))) {
      bx = 256;  // four independent wavefronts per workgroup, no LDS
      by = 1;
      lds = 0;
    } else if (dims == 2) {
      shape_2d(p.stencil, p.time_tile, p.elems_per_thread, d.tile, &bx, &by, &lds);
    } else if (k3w_rows(p.time_tile, d.tile) > 0) {
      bx = 256;  // four independent wavefronts per workgroup, no LDS
      by = 1;
      lds = 0;
    } else if (p.stencil == OT_CS3D) {
      return set_error(OT_EUNSUPPORTED,
                       "cs-3d runs on the wave-pipelined 3D kernel only: tile x must be 64 - 2T (got %d at T=%d) "
                       "and tile y + 2T one of 8, 12, 16, 20, 24, 32 rows within the register budget (got %d)",
                       d.tile[0], p.time_tile, d.tile[1] + 2 * p.time_tile);
    } else {
      shape_3d(p.stencil, p.time_tile, p.elems_per_thread, d.tile, &bx, &by, &lds);
    }
    threads = bx * by;
  }
  d.threads_per_block = threads;
  d.lds_bytes = lds;
  if (threads > kMaxThreads)
    return set_error(OT_ERESOURCE, "ERROR: Block dimensions are too large! (%d threads > %d)", threads, kMaxThreads);
  if (lds > kMaxLdsBytes)
    return set_error(OT_ERESOURCE, "ERROR: Not enough local memory for even one block! (%d B > %d B)", lds,
                     kMaxLdsBytes);
  return OT_OK;
}

static ot_status launch_any(int stencil, int dims, int t_eff, int ept, const LaunchGeom& g,
                            const FieldPtrs& f, void* stream, bool prepare) {
  if (dims == 1) return launch_1d(stencil, t_eff, ept, g, f, stream, prepare);
  if (dims == 2) return launch_2d(stencil, t_eff, ept, g, f, stream, prepare);
  return launch_3d(stencil, t_eff, ept, g, f, stream, prepare);
}

}  // namespace ot

using namespace ot;

struct ot_plan {
  ot_params p;
  ot_derived d;
  LaunchGeom g;
  hipStream_t own_stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // cached graph of the whole time loop
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
  const float* graph_key[8] = {};
  // per-launch events (OT_RUN_TIME_LAUNCHES)
  hipEvent_t* lev = nullptr;
  int nlev = 0;
  // clock records of the last launch (OT_VARIANT_DUMP_CLOCKS)
  ot_clock_record* clocks = nullptr;
  int64_t nclocks = 0;
};

static void fill_ptrs(const ot_plan* pl, int launch, float* const* rw_a, float* const* rw_b,
                      const float* const* ro, FieldPtrs& f) {
  std::memset(&f, 0, sizeof(f));
  const bool even = (launch % 2) == 0;
  for (int k = 0; k < pl->d.num_rw; ++k) {
    f.src[k] = even ? rw_a[k] : rw_b[k];
    f.dst[k] = even ? rw_b[k] : rw_a[k];
  }
  for (int k = 0; k < pl->d.num_ro; ++k) f.ro[k] = ro[k];
}

// Enqueue the whole time loop. With lev (2 events per launch) every launch is
// bracketed by its own event pair.
// Clears the clock records before a probe launch (a kernel rather than
// hipMemsetAsync, so that a captured graph replays it correctly).
__global__ void zero_words(uint32_t* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

static ot_status enqueue_loop(ot_plan* pl, float* const* rw_a, float* const* rw_b, const float* const* ro,
                              hipStream_t s, hipEvent_t* lev) {
  const int T = pl->p.time_tile;
  for (int l = 0; l < pl->d.launches; ++l) {
    const int base = l * T;
    const int t_eff = (pl->p.time_steps - base) < T ? (pl->p.time_steps - base) : T;
    FieldPtrs f;
    fill_ptrs(pl, l, rw_a, rw_b, ro, f);
    if (pl->clocks) {
      // the records keep only the last launch's waves
      const int64_t words = pl->nclocks * (int64_t)(sizeof(ot_clock_record) / 4);
      hipLaunchKernelGGL(zero_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s,
                         (uint32_t*)pl->clocks, words);
      OT_HIP(hipGetLastError());
    }
    if (lev) OT_HIP(hipEventRecord(lev[2 * l], s));
    ot_status st = launch_any(pl->p.stencil, pl->d.dims, t_eff, pl->p.elems_per_thread, pl->g, f, s, false);
    if (st != OT_OK) return st;
    if (lev) OT_HIP(hipEventRecord(lev[2 * l + 1], s));
  }
  return OT_OK;
}

static ot_status check_buffers(const ot_plan* pl, float* const* rw_a, float* const* rw_b,
                               const float* const* ro) {
  if (!rw_a || !rw_b) return set_error(OT_EINVAL, "null rw buffer arrays");
  for (int k = 0; k < pl->d.num_rw; ++k)
    if (!rw_a[k] || !rw_b[k]) return set_error(OT_EINVAL, "null rw buffer for field %d", k);
  if (pl->d.num_ro > 0) {
    if (!ro) return set_error(OT_EINVAL, "stencil needs %d read-only fields", pl->d.num_ro);
    for (int k = 0; k < pl->d.num_ro; ++k)
      if (!ro[k]) return set_error(OT_EINVAL, "null read-only buffer %d", k);
  }
  return OT_OK;
}

extern "C" {

const char* ot_version(void) { return "ot-mi355x 0.1 (gfx950, ABI " "1" ")"; }

const char* ot_last_error(void) { return g_err; }

ot_status ot_device_query(int32_t* num_cus, int32_t* lds_per_block, int32_t* max_threads, char* name,
                          int32_t name_len) {
  int dev = 0;
  OT_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  OT_HIP(hipGetDeviceProperties(&prop, dev));
  if (num_cus) *num_cus = prop.multiProcessorCount;
  if (lds_per_block) *lds_per_block = (int32_t)prop.sharedMemPerBlock;
  if (max_threads) *max_threads = prop.maxThreadsPerBlock;
  if (name && name_len > 0) std::snprintf(name, name_len, "%s", prop.gcnArchName);
  return OT_OK;
}

ot_status ot_plan_create(const ot_params* params, ot_plan** plan) {
  if (!params || !plan) return set_error(OT_EINVAL, "null argument");
  *plan = nullptr;
  ot_derived d;
  ot_status st = derive(*params, d);
  if (st != OT_OK) return st;
  ot_plan* pl = new (std::nothrow) ot_plan();
  if (!pl) return set_error(OT_ERESOURCE, "out of host memory");
  pl->p = *params;
  pl->d = d;
  LaunchGeom& g = pl->g;
  std::memset(&g, 0, sizeof(g));
  for (int a = 0; a < 3; ++a) {
    g.n[a] = d.real_size[a];
    g.tile[a] = d.tile[a];
    g.grid[a] = d.grid_blocks[a];
  }
  g.sy = d.stride_y;
  g.sz = d.stride_z;
  g.origin = d.origin;
  g.phase_limit = params->phase_limit;
  g.variant = params->variant;
  for (int i = 0; i < 4; ++i) g.coef[i] = params->coef[i];
  g.wave_rows = (d.dims == 3)   ? k3w_rows(params->time_tile, d.tile)
                : (d.dims == 2) ? (k2s_tile(params->stencil, params->time_tile, d.tile)
                                       ? -1
                                       : k2w_rows(params->stencil, params->time_tile, d.tile))
                                : 0;
  // timing probes: 2D LDS kernel (checked at launch); clock records also for
  // the wave-pipelined jacobi-3d kernel
  const bool clocks_3dw = d.dims == 3 && params->stencil == OT_J3D && g.wave_rows > 0 &&
                          (params->variant & kProbeBits) == OT_VARIANT_DUMP_CLOCKS;
  if ((params->variant & kProbeBits) && d.dims != 2 && !clocks_3dw) {
    delete pl;
    return set_error(OT_EUNSUPPORTED,
                     "--time-only and --dump-clocks exist for jacobi-2d and gradient-2d (as in the reference); "
                     "--dump-clocks also for jacobi-3d tiles of the wave-pipelined kernel");
  }
  // Check the instantiations this plan will launch (full T and the shortened
  // last launch) exist, and raise their LDS limits now.
  FieldPtrs f;
  std::memset(&f, 0, sizeof(f));
  const int T = params->time_tile;
  const int last = params->time_steps % T;
  g_prepared_regs = 0;
  st = launch_any(params->stencil, d.dims, T, params->elems_per_thread, g, f, nullptr, true);
  pl->d.registers = g_prepared_regs;
  if (st == OT_OK && last > 0)
    st = launch_any(params->stencil, d.dims, last, params->elems_per_thread, g, f, nullptr, true);
  if (st == OT_OK && (params->variant & OT_VARIANT_DUMP_CLOCKS)) {
    // 2D: a fixed slot per (workgroup, wave); 3D wave kernel: one per wave
    pl->nclocks = clocks_3dw ? (int64_t)d.grid_blocks[0] * d.grid_blocks[1] * d.grid_blocks[2]
                             : (int64_t)d.grid_blocks[0] * d.grid_blocks[1] * kClockSlotsPerBlock;
    const size_t bytes = (size_t)pl->nclocks * sizeof(ot_clock_record);
    hipError_t e = hipMalloc((void**)&pl->clocks, bytes);
    if (e == hipSuccess) e = hipMemset(pl->clocks, 0, bytes);
    if (e != hipSuccess) st = set_error(OT_EHIP, "clock records: %s", hipGetErrorString(e));
    pl->g.clocks = pl->clocks;
  }
  if (st != OT_OK) {
    if (pl->clocks) (void)hipFree(pl->clocks);
    delete pl;
    return st;
  }
  *plan = pl;
  return OT_OK;
}

ot_status ot_plan_clocks(ot_plan* plan, ot_clock_record* out, int32_t capacity, int32_t* count) {
  if (!plan || !count || capacity < 0 || (capacity > 0 && !out)) return set_error(OT_EINVAL, "null argument");
  if (!plan->clocks) return set_error(OT_EINVAL, "the plan was not created with OT_VARIANT_DUMP_CLOCKS");
  *count = (int32_t)plan->nclocks;
  const int64_t n = capacity < plan->nclocks ? capacity : plan->nclocks;
  if (n > 0) {
    // the plan's launches may have run on any stream
    OT_HIP(hipDeviceSynchronize());
    OT_HIP(hipMemcpy(out, plan->clocks, (size_t)n * sizeof(ot_clock_record), hipMemcpyDeviceToHost));
  }
  return OT_OK;
}

ot_status ot_plan_derived(const ot_plan* plan, ot_derived* derived) {
  if (!plan || !derived) return set_error(OT_EINVAL, "null argument");
  *derived = plan->d;
  return OT_OK;
}

void ot_plan_destroy(ot_plan* plan) {
  if (!plan) return;
  // solves enqueued with OT_RUN_ASYNC may still run on the plan's stream
  if (plan->own_stream) (void)hipStreamSynchronize(plan->own_stream);
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->ev0) (void)hipEventDestroy(plan->ev0);
  if (plan->ev1) (void)hipEventDestroy(plan->ev1);
  if (plan->own_stream) (void)hipStreamDestroy(plan->own_stream);
  for (int i = 0; i < plan->nlev; ++i) (void)hipEventDestroy(plan->lev[i]);
  delete[] plan->lev;
  if (plan->clocks) (void)hipFree(plan->clocks);
  delete plan;
}

ot_status ot_plan_stream(ot_plan* plan, void** stream) {
  if (!plan || !stream) return set_error(OT_EINVAL, "null argument");
  if (!plan->own_stream) OT_HIP(hipStreamCreateWithFlags(&plan->own_stream, hipStreamNonBlocking));
  *stream = (void*)plan->own_stream;
  return OT_OK;
}

ot_status ot_launch(ot_plan* plan, const float* const* rw_src, float* const* rw_dst,
                    const float* const* ro, int32_t base_time, void* stream) {
  if (!plan) return set_error(OT_EINVAL, "null plan");
  if (!rw_src || !rw_dst) return set_error(OT_EINVAL, "null buffers");
  const int T = plan->p.time_tile;
  if (base_time < 0 || base_time >= plan->p.time_steps)
    return set_error(OT_EINVAL, "base time %d outside [0, %d)", base_time, plan->p.time_steps);
  const int t_eff = (plan->p.time_steps - base_time) < T ? (plan->p.time_steps - base_time) : T;
  FieldPtrs f;
  std::memset(&f, 0, sizeof(f));
  for (int k = 0; k < plan->d.num_rw; ++k) {
    if (!rw_src[k] || !rw_dst[k]) return set_error(OT_EINVAL, "null rw buffer %d", k);
    f.src[k] = rw_src[k];
    f.dst[k] = rw_dst[k];
  }
  for (int k = 0; k < plan->d.num_ro; ++k) {
    if (!ro || !ro[k]) return set_error(OT_EINVAL, "null read-only buffer %d", k);
    f.ro[k] = ro[k];
  }
  return launch_any(plan->p.stencil, plan->d.dims, t_eff, plan->p.elems_per_thread, plan->g, f, stream, false);
}

ot_status ot_run(ot_plan* pl, float* const* rw_a, float* const* rw_b, const float* const* ro,
                 void* stream, int32_t mode, ot_timing* timing) {
  if (!pl) return set_error(OT_EINVAL, "null plan");
  ot_status st = check_buffers(pl, rw_a, rw_b, ro);
  if (st != OT_OK) return st;
  const bool use_graph = (mode & OT_RUN_GRAPH) != 0;
  const bool time_launches = (mode & OT_RUN_TIME_LAUNCHES) != 0;
  const bool async = (mode & OT_RUN_ASYNC) != 0;
  if (use_graph && time_launches)
    return set_error(OT_EINVAL, "OT_RUN_TIME_LAUNCHES times eager launches; do not combine with OT_RUN_GRAPH");
  if (async && time_launches)
    return set_error(OT_EINVAL, "OT_RUN_TIME_LAUNCHES waits for the launches; do not combine with OT_RUN_ASYNC");
  if (time_launches && pl->nlev < 2 * pl->d.launches) {
    for (int i = 0; i < pl->nlev; ++i) (void)hipEventDestroy(pl->lev[i]);
    delete[] pl->lev;
    pl->nlev = 0;
    pl->lev = new (std::nothrow) hipEvent_t[2 * pl->d.launches];
    if (!pl->lev) return set_error(OT_ERESOURCE, "out of host memory");
    for (int i = 0; i < 2 * pl->d.launches; ++i) {
      OT_HIP(hipEventCreate(&pl->lev[i]));
      pl->nlev = i + 1;
    }
  }
  hipStream_t s = (hipStream_t)stream;
  if (!s && use_graph) {
    if (!pl->own_stream) OT_HIP(hipStreamCreateWithFlags(&pl->own_stream, hipStreamNonBlocking));
    s = pl->own_stream;
  }
  if (!pl->ev0) OT_HIP(hipEventCreate(&pl->ev0));
  if (!pl->ev1) OT_HIP(hipEventCreate(&pl->ev1));
  ot_timing tm;
  std::memset(&tm, 0, sizeof(tm));
  tm.launches = pl->d.launches;
  tm.final_buffer = pl->d.launches % 2;
  if (pl->d.launches == 0) {
    if (timing) *timing = tm;
    return OT_OK;
  }
  if (use_graph) {
    const float* key[8] = {};
    int nk = 0;
    for (int k = 0; k < pl->d.num_rw; ++k) {
      key[nk++] = rw_a[k];
      key[nk++] = rw_b[k];
    }
    for (int k = 0; k < pl->d.num_ro; ++k) key[nk++] = ro[k];
    const bool hit = pl->exec && pl->graph_stream == s && std::memcmp(key, pl->graph_key, sizeof(key)) == 0;
    if (!hit) {
      if (pl->exec) {
        (void)hipGraphExecDestroy(pl->exec);
        pl->exec = nullptr;
      }
      OT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      ot_status cst = enqueue_loop(pl, rw_a, rw_b, ro, s, nullptr);
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(s, &graph);
      if (cst != OT_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return cst;
      }
      if (e != hipSuccess) return set_error(OT_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&pl->exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return set_error(OT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      pl->graph_stream = s;
      std::memcpy(pl->graph_key, key, sizeof(key));
    }
  }
  if (async) {
    // enqueue only: no events, no wait (the caller synchronises the stream)
    if (use_graph) {
      OT_HIP(hipGraphLaunch(pl->exec, s));
    } else {
      st = enqueue_loop(pl, rw_a, rw_b, ro, s, nullptr);
      if (st != OT_OK) return st;
    }
    tm.full_launches = pl->p.time_steps / pl->p.time_tile;
    if (timing) *timing = tm;
    return OT_OK;
  }
  const auto t0 = std::chrono::steady_clock::now();
  OT_HIP(hipEventRecord(pl->ev0, s));
  if (use_graph) {
    OT_HIP(hipGraphLaunch(pl->exec, s));
  } else {
    st = enqueue_loop(pl, rw_a, rw_b, ro, s, time_launches ? pl->lev : nullptr);
    if (st != OT_OK) return st;
  }
  OT_HIP(hipEventRecord(pl->ev1, s));
  OT_HIP(hipEventSynchronize(pl->ev1));
  const auto t1 = std::chrono::steady_clock::now();
  float ms = 0.0f;
  OT_HIP(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
  tm.wall_s = std::chrono::duration<double>(t1 - t0).count();
  tm.event_s = ms * 1e-3;
  tm.full_launches = pl->p.time_steps / pl->p.time_tile;
  if (time_launches) {
    for (int l = 0; l < pl->d.launches; ++l) {
      float lms = 0.0f;
      OT_HIP(hipEventElapsedTime(&lms, pl->lev[2 * l], pl->lev[2 * l + 1]));
      tm.kernel_s += lms * 1e-3;
      if (l < tm.full_launches) tm.kernel_full_s += lms * 1e-3;
    }
  }
  if (timing) *timing = tm;
  return OT_OK;
}

// ---- device memory helpers ------------------------------------------------
ot_status ot_dev_alloc(size_t bytes, void** ptr) {
  if (!ptr) return set_error(OT_EINVAL, "null ptr");
  OT_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return OT_OK;
}
ot_status ot_dev_free(void* ptr) {
  if (ptr) OT_HIP(hipFree(ptr));
  return OT_OK;
}
ot_status ot_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  OT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return OT_OK;
}
ot_status ot_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  OT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return OT_OK;
}
ot_status ot_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  OT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return OT_OK;
}
ot_status ot_dev_memset(void* ptr, int32_t value, size_t bytes) {
  OT_HIP(hipMemset(ptr, value, bytes));
  return OT_OK;
}
ot_status ot_dev_sync(void) {
  OT_HIP(hipDeviceSynchronize());
  return OT_OK;
}

}  // extern "C"
