// kbatch.hip — one persistent launch per time tile over a batch of 2D images.
//
// No reference counterpart: the reference runs one square grid per process.
// Every (image, tile) pair of the batch is an item of a device work queue (one
// atomic counter per launch); the workgroups of a launch — as many as fit on
// the device at once — pull items until the queue is empty, and advance each
// item's tile by t_eff steps with the same tile update as the per-grid 2D
// kernels: the LDS tile update (k2d_impl.hpp), or, for a tile that fills a
// wavefront, the wave-per-tile update (k2w_impl.hpp) with every wave pulling
// its own items. Items are ordered largest image first. The queue
// counters are cleared by a small kernel at the start of every run (no memset
// node, so the loop can be captured and replayed as a hipGraph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "k2d_impl.hpp"
#include "traits.hpp"

namespace ot {
namespace {

template <int ST, int E>
__global__ void __launch_bounds__(1024) k2d_batch(BArgs a, int t_eff, int launch) {
  extern __shared__ float lds[];
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0 && threadIdx.y == 0) s_item = atomicAdd(&a.counters[launch], 1);
    __syncthreads();
    const int it = s_item;
    __syncthreads();  // every thread has the ticket before it is replaced
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
    g.wave_rows = 0;
    g.probe_sink = 0;
    g.clocks = nullptr;
    tile2d<ST, E>(g, a.base, t_eff, tx, ty, lds);
    __syncthreads();  // LDS is reused by the next item
  }
}

__global__ void zero_counters(int32_t* c, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicExch(&c[i], 0);
}

using KernelFn = BatchKernel;

template <int ST>
KernelFn pick_e(int ept) {
  switch (ept) {
    case 1: return k2d_batch<ST, 1>;
    case 2: return k2d_batch<ST, 2>;
    case 4: return k2d_batch<ST, 4>;
    default: return nullptr;
  }
}

KernelFn pick(int stencil, int ept) {
  switch (stencil) {
    case OT_J2D: return pick_e<OT_J2D>(ept);
    case OT_P2D: return pick_e<OT_P2D>(ept);
    case OT_G2D: return pick_e<OT_G2D>(ept);
    case OT_TV2D: return pick_e<OT_TV2D>(ept);
    case OT_RICIAN2D: return pick_e<OT_RICIAN2D>(ept);
    case OT_FDTD2D: return pick_e<OT_FDTD2D>(ept);
    default: return nullptr;
  }
}

}  // namespace

BatchKernel batch_wave_kernel(int stencil, int t_eff, int rows) {
  switch (stencil) {
    case OT_J2D: return batch_wave_kernel_j2d(t_eff, rows);
    case OT_P2D: return batch_wave_kernel_p2d(t_eff, rows);
    case OT_G2D: return batch_wave_kernel_g2d(t_eff, rows);
    case OT_TV2D: return batch_wave_kernel_tv2d(t_eff, rows);
    case OT_FDTD2D: return batch_wave_kernel_fdtd2d(t_eff, rows);
    case OT_RICIAN2D: return batch_wave_kernel_rician2d(t_eff, rows);
    default: return nullptr;
  }
}

namespace {

#define OT_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return set_error(OT_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

struct LaunchShape {
  KernelFn fn = nullptr;
  int bx = 0, by = 0, lds = 0, wgs = 0;
};

}  // namespace
}  // namespace ot

using namespace ot;

struct ot_batch {
  ot_params p;
  int n_images = 0, n_items = 0, launches = 0, num_rw = 0, num_ro = 0;
  LaunchShape full, last;
  BImage* d_images = nullptr;
  int32_t* d_items = nullptr;
  int32_t* d_counters = nullptr;
  hipStream_t own_stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
  const float* graph_key[8] = {};
  int graph_per_image = 0;
  std::vector<ot_image> host_images;
  bool per_image_mode = false;  // set for the duration of one ot_batch_run
};

// Launch shape of one persistent batch launch advancing t_eff steps with the
// batch's tile (time tile T). A tile that fills a wavefront (x tile + 2rT = 64
// with a compiled footprint row count, as the per-grid wave kernel takes) runs
// the wave-per-tile batch kernel, four independent waves per workgroup; other
// tiles run the LDS tile update.
static ot_status make_shape(int stencil, int T, int t_eff, int ept, const int32_t tile[3], LaunchShape& s,
                            int n_items) {
  const int rows = k2w_rows(stencil, T, tile);
  int items_per_wg = 1;
  if (rows > 0) {
    s.fn = batch_wave_kernel(stencil, t_eff, rows);
    if (!s.fn)
      return set_error(OT_EUNSUPPORTED, "batched wave kernel not compiled for T=%d with %d footprint rows", t_eff,
                       rows);
    s.bx = 256;
    s.by = 1;
    s.lds = 0;
    items_per_wg = 4;
  } else {
    s.fn = pick(stencil, ept);
    if (!s.fn) return set_error(OT_EUNSUPPORTED, "batched kernels are compiled for 2D stencils with 1, 2 or 4 "
                                                 "elements per thread");
    shape_2d(stencil, t_eff, ept, tile, &s.bx, &s.by, &s.lds);
    if (s.bx * s.by > kMaxThreads) return set_error(OT_ERESOURCE, "batch tile needs %d threads", s.bx * s.by);
    if (s.lds > kMaxLdsBytes) return set_error(OT_ERESOURCE, "batch tile needs %d B of LDS", s.lds);
  }
  const ot_status pst = prepare_lds((const void*)s.fn, s.lds);
  if (pst != OT_OK) return pst;
  int per_cu = 0;
  OT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)s.fn, s.bx * s.by, s.lds));
  if (per_cu < 1) per_cu = 1;
  int dev = 0, cus = 0;
  OT_HIP(hipGetDevice(&dev));
  OT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  s.wgs = std::min((n_items + items_per_wg - 1) / items_per_wg, per_cu * cus);
  return OT_OK;
}

// Baseline: every image advanced by its own per-grid launches.
static ot_status per_image_enqueue(ot_batch* b, float* const* rw_a, float* const* rw_b, const float* const* ro,
                                   hipStream_t s) {
  const int T = b->p.time_tile;
  for (int l = 0; l < b->launches; ++l) {
    const int t_eff = std::min(T, b->p.time_steps - l * T);
    for (int i = 0; i < b->n_images; ++i) {
      const ot_image& im = b->host_images[i];
      LaunchGeom g;
      std::memset(&g, 0, sizeof(g));
      g.n[0] = im.width;
      g.n[1] = im.height;
      g.n[2] = 1;
      g.tile[0] = b->p.block_x;
      g.tile[1] = b->p.block_y * b->p.elems_per_thread;
      g.tile[2] = 1;
      g.grid[0] = (int)ceil_div(im.width, g.tile[0]);
      g.grid[1] = (int)ceil_div(im.height, g.tile[1]);
      g.grid[2] = 1;
      g.sy = im.pitch;
      g.sz = (int64_t)im.pitch * im.height;
      g.origin = im.offset;
      g.wave_rows = k2w_rows(b->p.stencil, T, g.tile);  // the kernel family of the batched launch
      g.phase_limit = b->p.phase_limit;
      g.variant = b->p.variant;
      for (int c = 0; c < 4; ++c) g.coef[c] = b->p.coef[c];
      FieldPtrs f;
      std::memset(&f, 0, sizeof(f));
      const bool even = (l % 2) == 0;
      for (int k = 0; k < b->num_rw; ++k) {
        f.src[k] = even ? rw_a[k] : rw_b[k];
        f.dst[k] = even ? rw_b[k] : rw_a[k];
      }
      for (int k = 0; k < b->num_ro; ++k) f.ro[k] = ro[k];
      const ot_status st = launch_2d(b->p.stencil, t_eff, b->p.elems_per_thread, g, f, s, false);
      if (st != OT_OK) return st;
    }
  }
  return OT_OK;
}

// The persistent work-queue launches of one batch solve (counters already
// cleared): launch l advances every item by min(T, steps - l T) steps.
static ot_status enqueue_queue(const ot_params& p, int launches, int n_items, const LaunchShape& full,
                               const LaunchShape& last, const BImage* images, const int32_t* items,
                               int32_t* counters, float* const* rw_a, float* const* rw_b, const float* const* ro,
                               hipStream_t s) {
  const int T = p.time_tile;
  const int num_rw = kTraits[p.stencil].num_rw, num_ro = kTraits[p.stencil].num_ro;
  for (int l = 0; l < launches; ++l) {
    const int t_eff = std::min(T, p.time_steps - l * T);
    const LaunchShape& sh = (t_eff == T) ? full : last;
    BArgs a;
    std::memset(&a, 0, sizeof(a));
    a.images = images;
    a.items = items;
    a.counters = counters;
    a.n_items = n_items;
    a.tile[0] = p.block_x;
    a.tile[1] = p.block_y * p.elems_per_thread;
    a.phase_limit = p.phase_limit;
    a.variant = p.variant;
    for (int i = 0; i < 4; ++i) a.coef[i] = p.coef[i];
    const bool even = (l % 2) == 0;
    for (int k = 0; k < num_rw; ++k) {
      a.base.src[k] = even ? rw_a[k] : rw_b[k];
      a.base.dst[k] = even ? rw_b[k] : rw_a[k];
    }
    for (int k = 0; k < num_ro; ++k) a.base.ro[k] = ro[k];
    hipLaunchKernelGGL(sh.fn, dim3(sh.wgs), dim3(sh.bx, sh.by), sh.lds, s, a, t_eff, l);
    OT_HIP(hipGetLastError());
  }
  return OT_OK;
}

static ot_status batch_enqueue(ot_batch* b, float* const* rw_a, float* const* rw_b, const float* const* ro,
                               hipStream_t s) {
  if (b->per_image_mode) return per_image_enqueue(b, rw_a, rw_b, ro, s);
  hipLaunchKernelGGL(zero_counters, dim3((b->launches + 255) / 256), dim3(256), 0, s, b->d_counters, b->launches);
  OT_HIP(hipGetLastError());
  return enqueue_queue(b->p, b->launches, b->n_items, b->full, b->last, b->d_images, b->d_items, b->d_counters,
                       rw_a, rw_b, ro, s);
}

// ---- the allocation-free entry point ---------------------------------------
// Its work queue (launch counters, image table, items) lives in `scratch`, a
// file-scope device array of fixed size that every GPU the library runs on
// holds (its own copy) until the process ends. Host code never passes the
// array's name to the device: it takes each GPU's address once with
// hipGetSymbolAddress and uses that.
namespace {
constexpr size_t kStaticScratchBytes = size_t(8) << 20;  // 8 MiB per GPU
constexpr int kStaticMaxImages = 4096;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on
// that device (with the device current) and kept.
ot_status static_scratch(void** out) {
  int dev = 0;
  OT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(OT_EUNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(OT_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return OT_OK;
}

// One workgroup builds the work queue of ot_batch_create in device memory:
// clears the launch counters, ranks the images largest first (ties by index,
// the stable order of ot_batch_create), and writes the items of each image
// (tile rows outer, tile columns inner) after those of the larger ones.
__global__ void __launch_bounds__(1024) build_queue(const BImage* images, int n, int tile_x, int tile_y,
                                                    int32_t* items, int n_items, int32_t* counters, int launches) {
  __shared__ int s_img[kStaticMaxImages];     // image at rank r
  __shared__ int s_off[kStaticMaxImages + 1]; // first item of rank r
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  for (int i = tid; i < launches; i += nt) counters[i] = 0;
  for (int i = tid; i < n; i += nt) {
    const int64_t ai = (int64_t)images[i].w * images[i].h;
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const int64_t aj = (int64_t)images[j].w * images[j].h;
      r += (aj > ai || (aj == ai && j < i)) ? 1 : 0;
    }
    s_img[r] = i;
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int r = 0; r < n; ++r) {
      const BImage im = images[s_img[r]];
      s_off[r] = acc;
      acc += ((im.w + tile_x - 1) / tile_x) * ((im.h + tile_y - 1) / tile_y);
    }
    s_off[n] = acc;
  }
  __syncthreads();
  for (int it = tid; it < n_items; it += nt) {
    int lo = 0, hi = n - 1;  // the rank whose items hold `it`: s_off[lo] <= it < s_off[lo + 1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_off[mid] <= it) lo = mid;
      else hi = mid - 1;
    }
    const int img = s_img[lo], k = it - s_off[lo];
    const int tiles_x = (images[img].w + tile_x - 1) / tile_x;
    items[3 * it] = img;
    items[3 * it + 1] = k % tiles_x;
    items[3 * it + 2] = k / tiles_x;
  }
}
}  // namespace

static_assert(sizeof(ot_image) == sizeof(BImage) && offsetof(ot_image, pitch) == offsetof(BImage, pitch) &&
                  offsetof(ot_image, offset) == offsetof(BImage, offset),
              "the image table is copied to the device as it is");

// Checks shared by ot_batch_create and ot_batch_run_static.
static ot_status check_batch(const ot_params& p, int32_t n_images, const ot_image* images) {
  if (p.stencil < 0 || p.stencil >= OT_NUM_STENCILS || kTraits[p.stencil].dims != 2)
    return set_error(OT_EINVAL, "batches take a 2D stencil");
  if (p.time_tile < 1 || p.time_tile > kMaxT || p.time_steps < 0 || p.block_x < 1 || p.block_y < 1 ||
      p.elems_per_thread < 1)
    return set_error(OT_EINVAL, "invalid time tile / steps / block");
  if (p.phase_limit < 0 || p.phase_limit > 3) return set_error(OT_EINVAL, "phase limit must be 0..3");
  if (p.variant & kProbeBits)
    return set_error(OT_EUNSUPPORTED, "the timing probes run per grid (ot_plan_create), not in batches");
  for (int i = 0; i < n_images; ++i) {
    const ot_image& im = images[i];
    if (im.width < 1 || im.height < 1 || im.pitch < im.width || im.offset < 0)
      return set_error(OT_EINVAL, "image %d has an invalid layout", i);
  }
  return OT_OK;
}

extern "C" {

ot_status ot_batch_layout(int32_t n_images, const int32_t* widths, const int32_t* heights, ot_image* images,
                          int64_t* arena_elems) {
  if (n_images < 1 || !widths || !heights || !images || !arena_elems) return set_error(OT_EINVAL, "bad arguments");
  int64_t off = 0;
  for (int i = 0; i < n_images; ++i) {
    if (widths[i] < 1 || heights[i] < 1) return set_error(OT_EINVAL, "image %d has an empty extent", i);
    images[i].width = widths[i];
    images[i].height = heights[i];
    images[i].pitch = (int32_t)round_up(widths[i], 16);
    images[i].reserved = 0;
    images[i].offset = off;
    off += round_up((int64_t)images[i].pitch * heights[i], 16);
  }
  *arena_elems = off;
  return OT_OK;
}

ot_status ot_batch_create(const ot_params* params, int32_t n_images, const ot_image* images, ot_batch** out) {
  if (!params || !images || !out || n_images < 1) return set_error(OT_EINVAL, "bad arguments");
  *out = nullptr;
  const ot_params& p = *params;
  const ot_status cst = check_batch(p, n_images, images);
  if (cst != OT_OK) return cst;
  const int32_t tile[3] = {p.block_x, p.block_y * p.elems_per_thread, 1};
  // items: every tile of every image, largest image first
  std::vector<int> order(n_images);
  for (int i = 0; i < n_images; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (int64_t)images[a].width * images[a].height > (int64_t)images[b].width * images[b].height;
  });
  std::vector<int32_t> items;
  std::vector<BImage> bim(n_images);
  for (int i = 0; i < n_images; ++i) {
    const ot_image& im = images[i];
    bim[i] = {im.width, im.height, im.pitch, 0, im.offset};
  }
  for (int i : order) {
    const int tx = (int)ceil_div(images[i].width, tile[0]), ty = (int)ceil_div(images[i].height, tile[1]);
    for (int y = 0; y < ty; ++y)
      for (int x = 0; x < tx; ++x) {
        items.push_back(i);
        items.push_back(x);
        items.push_back(y);
      }
  }
  ot_batch* b = new (std::nothrow) ot_batch();
  if (!b) return set_error(OT_ERESOURCE, "out of host memory");
  b->p = p;
  b->n_images = n_images;
  b->n_items = (int)(items.size() / 3);
  b->launches = (int)ceil_div(p.time_steps, p.time_tile);
  b->num_rw = kTraits[p.stencil].num_rw;
  b->num_ro = kTraits[p.stencil].num_ro;
  b->host_images.assign(images, images + n_images);
  ot_status st = make_shape(p.stencil, p.time_tile, p.time_tile, p.elems_per_thread, tile, b->full, b->n_items);
  const int last = p.time_steps % p.time_tile;
  if (st == OT_OK)
    st = last ? make_shape(p.stencil, p.time_tile, last, p.elems_per_thread, tile, b->last, b->n_items) : OT_OK;
  if (st == OT_OK && last == 0) b->last = b->full;
  // the per-image baseline launches the per-grid kernels of the same tile
  if (st == OT_OK) {
    LaunchGeom g;
    std::memset(&g, 0, sizeof(g));
    for (int a = 0; a < 3; ++a) g.tile[a] = tile[a];
    g.wave_rows = k2w_rows(p.stencil, p.time_tile, tile);
    FieldPtrs f;
    std::memset(&f, 0, sizeof(f));
    st = launch_2d(p.stencil, p.time_tile, p.elems_per_thread, g, f, nullptr, true);
    if (st == OT_OK && last) st = launch_2d(p.stencil, last, p.elems_per_thread, g, f, nullptr, true);
  }
  hipError_t e = hipSuccess;
  if (st == OT_OK) {
    e = hipMalloc(&b->d_images, sizeof(BImage) * n_images);
    if (e == hipSuccess) e = hipMalloc(&b->d_items, sizeof(int32_t) * items.size());
    if (e == hipSuccess) e = hipMalloc(&b->d_counters, sizeof(int32_t) * std::max(1, b->launches));
    if (e == hipSuccess) e = hipMemcpy(b->d_images, bim.data(), sizeof(BImage) * n_images, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(b->d_items, items.data(), sizeof(int32_t) * items.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) st = set_error(OT_EHIP, "batch setup: %s", hipGetErrorString(e));
  }
  if (st != OT_OK) {
    ot_batch_destroy(b);
    return st;
  }
  *out = b;
  return OT_OK;
}

ot_status ot_batch_info(const ot_batch* b, int32_t* items, int32_t* workgroups, int32_t* launches) {
  if (!b) return set_error(OT_EINVAL, "null batch");
  if (items) *items = b->n_items;
  if (workgroups) *workgroups = b->full.wgs;
  if (launches) *launches = b->launches;
  return OT_OK;
}

ot_status ot_batch_run(ot_batch* b, float* const* rw_a, float* const* rw_b, const float* const* ro, void* stream,
                       int32_t mode, ot_timing* timing) {
  if (!b || !rw_a || !rw_b) return set_error(OT_EINVAL, "null argument");
  for (int k = 0; k < b->num_rw; ++k)
    if (!rw_a[k] || !rw_b[k]) return set_error(OT_EINVAL, "null rw arena %d", k);
  for (int k = 0; k < b->num_ro; ++k)
    if (!ro || !ro[k]) return set_error(OT_EINVAL, "null read-only arena %d", k);
  if (mode & OT_RUN_TIME_LAUNCHES) return set_error(OT_EINVAL, "per-launch timing is not offered for batches");
  const bool use_graph = (mode & OT_RUN_GRAPH) != 0;
  b->per_image_mode = (mode & OT_RUN_PER_IMAGE) != 0;
  hipStream_t s = (hipStream_t)stream;
  if (!s && use_graph) {
    if (!b->own_stream) OT_HIP(hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking));
    s = b->own_stream;
  }
  if (!b->ev0) OT_HIP(hipEventCreate(&b->ev0));
  if (!b->ev1) OT_HIP(hipEventCreate(&b->ev1));
  ot_timing tm;
  std::memset(&tm, 0, sizeof(tm));
  tm.launches = b->launches;
  tm.full_launches = b->p.time_steps / b->p.time_tile;
  tm.final_buffer = b->launches % 2;
  if (b->launches == 0) {
    if (timing) *timing = tm;
    return OT_OK;
  }
  if (use_graph) {
    const float* key[8] = {};
    int nk = 0;
    for (int k = 0; k < b->num_rw; ++k) {
      key[nk++] = rw_a[k];
      key[nk++] = rw_b[k];
    }
    for (int k = 0; k < b->num_ro; ++k) key[nk++] = ro[k];
    if (!(b->exec && b->graph_stream == s && b->graph_per_image == (int)b->per_image_mode &&
          std::memcmp(key, b->graph_key, sizeof(key)) == 0)) {
      if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
      }
      OT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      ot_status cst = batch_enqueue(b, rw_a, rw_b, ro, s);
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(s, &graph);
      if (cst != OT_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return cst;
      }
      if (e != hipSuccess) return set_error(OT_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return set_error(OT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      b->graph_stream = s;
      b->graph_per_image = (int)b->per_image_mode;
      std::memcpy(b->graph_key, key, sizeof(key));
    }
  }
  const auto t0 = std::chrono::steady_clock::now();
  OT_HIP(hipEventRecord(b->ev0, s));
  if (use_graph) {
    OT_HIP(hipGraphLaunch(b->exec, s));
  } else {
    ot_status st = batch_enqueue(b, rw_a, rw_b, ro, s);
    if (st != OT_OK) return st;
  }
  OT_HIP(hipEventRecord(b->ev1, s));
  OT_HIP(hipEventSynchronize(b->ev1));
  const auto t1 = std::chrono::steady_clock::now();
  float ms = 0.0f;
  OT_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  tm.wall_s = std::chrono::duration<double>(t1 - t0).count();
  tm.event_s = ms * 1e-3;
  if (timing) *timing = tm;
  return OT_OK;
}

ot_status ot_batch_run_static(const ot_params* params, int32_t n_images, const ot_image* images,
                              float* const* rw_a, float* const* rw_b, const float* const* ro, void* stream,
                              ot_timing* timing) {
  if (!params || !images || n_images < 1 || !rw_a || !rw_b) return set_error(OT_EINVAL, "bad arguments");
  const ot_params& p = *params;
  ot_status st = check_batch(p, n_images, images);
  if (st != OT_OK) return st;
  for (int k = 0; k < kTraits[p.stencil].num_rw; ++k)
    if (!rw_a[k] || !rw_b[k]) return set_error(OT_EINVAL, "null rw arena %d", k);
  for (int k = 0; k < kTraits[p.stencil].num_ro; ++k)
    if (!ro || !ro[k]) return set_error(OT_EINVAL, "null read-only arena %d", k);
  const int32_t tile[3] = {p.block_x, p.block_y * p.elems_per_thread, 1};
  int64_t n_items = 0;
  for (int i = 0; i < n_images; ++i) n_items += ceil_div(images[i].width, tile[0]) * ceil_div(images[i].height, tile[1]);
  const int launches = (int)ceil_div(p.time_steps, p.time_tile);
  // scratch layout: launch counters | image table | items (256-byte aligned parts)
  const size_t counters_b = (size_t)round_up(std::max(1, launches) * 4, 256);
  const size_t images_b = (size_t)round_up((int64_t)n_images * (int64_t)sizeof(BImage), 256);
  const size_t need = counters_b + images_b + (size_t)n_items * 12;
  if (n_images > kStaticMaxImages || n_items > INT32_MAX || need > kStaticScratchBytes)
    return set_error(OT_ERESOURCE,
                     "the batch's work queue (%d images, %lld tiles, %d launches: %zu bytes) exceeds the "
                     "allocation-free entry point's fixed scratch (%d images, %zu bytes); use ot_batch_create",
                     n_images, (long long)n_items, launches, need, kStaticMaxImages, kStaticScratchBytes);
  ot_timing tm;
  std::memset(&tm, 0, sizeof(tm));
  tm.launches = launches;
  tm.full_launches = p.time_steps / p.time_tile;
  tm.final_buffer = launches % 2;
  if (launches == 0) {
    if (timing) *timing = tm;
    return OT_OK;
  }
  LaunchShape full, last;
  st = make_shape(p.stencil, p.time_tile, p.time_tile, p.elems_per_thread, tile, full, (int)n_items);
  const int rem = p.time_steps % p.time_tile;
  if (st == OT_OK) st = rem ? make_shape(p.stencil, p.time_tile, rem, p.elems_per_thread, tile, last, (int)n_items) : OT_OK;
  if (st != OT_OK) return st;
  if (rem == 0) last = full;
  void* base = nullptr;
  st = static_scratch(&base);
  if (st != OT_OK) return st;
  int32_t* counters = (int32_t*)base;
  BImage* d_images = (BImage*)((unsigned char*)base + counters_b);
  int32_t* items = (int32_t*)((unsigned char*)base + counters_b + images_b);
  hipStream_t s = (hipStream_t)stream;
  OT_HIP(hipMemcpyAsync(d_images, images, (size_t)n_images * sizeof(BImage), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(build_queue, dim3(1), dim3(1024), 0, s, (const BImage*)d_images, (int)n_images, tile[0], tile[1],
                     items, (int)n_items, counters, launches);
  OT_HIP(hipGetLastError());
  st = enqueue_queue(p, launches, (int)n_items, full, last, d_images, items, counters, rw_a, rw_b, ro, s);
  if (st != OT_OK) return st;
  if (timing) *timing = tm;
  return OT_OK;
}

void ot_batch_destroy(ot_batch* b) {
  if (!b) return;
  if (b->own_stream) (void)hipStreamSynchronize(b->own_stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
  if (b->d_images) (void)hipFree(b->d_images);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_counters) (void)hipFree(b->d_counters);
  delete b;
}

}  // extern "C"
