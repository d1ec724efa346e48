/*
 * ot.h — C ABI of the MI355X overlapped time-tiling stencil engine.
 *
 * The reference (jholewinski/ics-12-overlapped-tiling) exposes this path in
 * three layers: an in-process generator API (GeneratorParams +
 * XGenerator::generate, opencl/jacobi-2d/jacobi-2d.cpp:21-173,
 * opencl/common/inc/ProgramGenerator.hpp:10-69), the generated kernel ABI
 * (kernel_func(input, [aux...], output, baseTime), e.g.
 * opencl/jacobi-2d/jacobi-2d.cpp:184-190) driven by a host time loop
 * (opencl/jacobi-2d/jacobi-2d.cpp:888-933), and a driver CLI whose stdout
 * "Key: value" lines feed the sweep scripts (scripts/utils.py:104-155).
 * This header replaces the first two layers; the ot-<stencil> drivers built on
 * top of it keep the third.
 *
 * Plain C: no exceptions cross this boundary, every entry point returns an
 * ot_status and leaves a message for ot_last_error() on failure.
 * Device buffers are caller-owned (hipMalloc'ed or from any allocator) and use
 * the padded row-major layout described by ot_derived (x fastest).
 * Threading: a plan is single-stream and not re-entrant; separate plans may run
 * concurrently on separate streams.
 */
#ifndef OT_OT_H
#define OT_OT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OT_ABI_VERSION 1

typedef enum ot_status {
  OT_OK = 0,
  OT_EINVAL = 1,        /* invalid parameter / inconsistent geometry
                           (reference: assert / "Consistency error!",
                           opencl/jacobi-3d/jacobi-3d.cpp:105-113) */
  OT_ERESOURCE = 2,     /* block or LDS too large for the device
                           (reference: exit 1, opencl/jacobi-2d/jacobi-2d.cpp:682-690) */
  OT_EHIP = 3,          /* HIP runtime error (reference: CLContext::throwOnError,
                           opencl/common/src/CLContext.cpp:50-56) */
  OT_EUNSUPPORTED = 4   /* configuration the engine does not implement */
} ot_status;

/* Stencils. Each one is the overlapped-tiled kernel of one reference driver
 * (opencl/<name>/<name>.cpp); P3D7 is the 7-point Poisson-Jacobi relaxation with
 * a read-only right-hand side that has no reference driver. CS3D runs on the
 * wave-pipelined 3D kernel only (tile x = 64 - 2T, tile geometry). */
typedef enum ot_stencil {
  OT_J1D3 = 0,    /* jacobi-1d        0.333f*(l+c+r)                       */
  OT_J1D5 = 1,    /* jacobi-1d-5pt    0.2f*sum5                            */
  OT_J1D7 = 2,    /* jacobi-1d-7pt    0.2f*sum7 (sic)                      */
  OT_J2D = 3,     /* jacobi-2d        0.2f*(l+c+r+t+b)                     */
  OT_J3D = 4,     /* jacobi-3d        0.143f*sum7                          */
  OT_P2D = 5,     /* poisson-2d       8c - sum8                            */
  OT_P3D27 = 6,   /* poisson-3d       26c - sum26                          */
  OT_P3D7 = 7,    /* Poisson-Jacobi   (sum6 - f)/6, f read-only            */
  OT_FDTD2D = 8,  /* fdtd-2d          Ex, Ey, Hz                           */
  OT_G2D = 9,     /* gradient-2d      c + 1/sqrt(1e-4 + sum4 (c-n)^2)      */
  OT_G3D = 10,    /* gradient-3d      c + 1/sqrt(1e-4 + sum6 (c-n)^2)      */
  OT_RICIAN2D = 11, /* rician-denoise-2d, f read-only                      */
  OT_TV2D = 12,   /* cdsc-tv-update-2d, V and F read-only                  */
  OT_CS3D = 13,   /* cs-3d            13-point gradient-weighted average   */
  OT_NUM_STENCILS = 14
} ot_stencil;

enum {
  OT_GEOM_TILE = 0,       /* block = output tile, halo T*r added, full grid   */
  OT_GEOM_REFERENCE = 1   /* block includes halo, realSize truncated as in the
                             reference's computeDerived()                     */
};

/* variant flags (ot_params.variant) */
enum {
  OT_VARIANT_FDTD_PHYSICAL = 1,  /* Hz uses -Ey(y) (curl sign); default keeps the
                                    reference's "+eyCenter" (fdtd-2d.cpp:320-322) */
  /* Timing probes of the 2D LDS kernel (jacobi-2d and gradient-2d, the drivers
   * with --time-only, jacobi-2d.cpp:258-266, :296-341, :368-433, :462-481).
   * Results are NOT the stencil's: the modes time one part of the kernel.
   *   GMEM:  the global-memory phases only: every footprint point is loaded and
   *          the tile interior stored back, no step is computed, no LDS used;
   *   INSTR: the arithmetic only: one load per point, every step computed with
   *          the neighbours taken from the thread's own registers, no LDS, no
   *          barrier and no store;
   *   SMEM:  the on-chip phase only: no load (each point starts at its x
   *          index), every step through LDS with its barriers, no store. */
  OT_VARIANT_TIME_ONLY_GMEM = 2,
  OT_VARIANT_TIME_ONLY_INSTR = 4,
  OT_VARIANT_TIME_ONLY_SMEM = 8,
  /* Per-wave clock stamps at the kernel's phase boundaries (-c/--dump-clocks,
   * jacobi-2d.cpp:202-205, :453-509, :957-997), read with ot_plan_clocks. 2D
   * LDS kernel of jacobi-2d and gradient-2d, and the wave-pipelined 3D kernel
   * of jacobi-3d (an extension: the reference stamps its 2D kernels only);
   * results are unchanged. */
  OT_VARIANT_DUMP_CLOCKS = 16,
  /* Wave-pipelined 3D kernels: no issue-priority progress balancing (the
   * hardware's oldest-first issue order); results are unchanged. */
  OT_VARIANT_NO_BALANCE = 32
};

typedef struct ot_stencil_info {
  int32_t stencil;
  int32_t dims;             /* 1, 2 or 3 */
  int32_t radius;           /* per-time-step dependency radius (Rician: 2) */
  int32_t num_rw;           /* read-write (ping-pong) fields */
  int32_t num_ro;           /* read-only fields */
  int32_t flops_per_point;  /* the driver's "Actual GFlop/s" constant */
  int32_t loads_per_point;  /* phase2_global_loads key */
  int32_t ref_padding_mult; /* reference padding = mult * T */
  const char* name;         /* reference driver name, e.g. "jacobi-2d" */
} ot_stencil_info;

typedef struct ot_params {
  int32_t stencil;          /* ot_stencil */
  int32_t problem_size;     /* -n/--problem-size: edge length of the square/cube/line */
  int32_t nx, ny, nz;       /* tile geometry only: per-axis extents, 0 = problem_size */
  int32_t time_steps;       /* -t/--time-steps */
  int32_t time_tile;        /* -s/--time-tile-size (T) */
  int32_t block_x;          /* -x/--block-size-x */
  int32_t block_y;          /* -y/--block-size-y */
  int32_t block_z;          /* -z/--block-size-z (3D: z-chunk streamed by one workgroup) */
  int32_t elems_per_thread; /* -e/--elements-per-thread */
  int32_t geometry;         /* OT_GEOM_TILE | OT_GEOM_REFERENCE */
  int32_t phase_limit;      /* -p/--phase-limit: 0 full kernel, 1 prologue only,
                               2 global phase only, 3 on-chip phase only */
  int32_t variant;          /* OT_VARIANT_* flags */
  float coef[4];            /* stencil scalars, 0 = reference default:
                               TV2D: coef[0] = alpha (1.1, cdsc-tv-update-2d.cpp:164)
                               RICIAN2D: coef[0] = sigma^2 (SIGMA2, the literal
                               0.0025), coef[1] = lambda (0.065), coef[2] = dt (5.0);
                               gamma = lambda / sigma^2 in FP32,
                               rician-denoise-2d.cpp:153-160 */
} ot_params;

typedef struct ot_derived {
  int32_t dims;
  int32_t radius;
  int32_t padding;             /* P: ring width of the padded layout */
  int32_t halo;                /* radius * T: halo loaded around each output tile */
  int32_t comps_per_block[3];  /* (bx, by*ept, bz) */
  int32_t real_per_block[3];   /* output points per workgroup per axis */
  int32_t size_lcm;            /* reference geometry: gcd/lcm of real_per_block */
  int32_t real_size[3];        /* domain extents (x, y, z); 1 for unused axes */
  int32_t padded_size[3];      /* real_size + 2P (1D adds the reference's +109) */
  int32_t num_blocks[3];       /* workgroups per axis as the reference reports them */
  int32_t tile[3];             /* output points per workgroup this engine uses */
  int32_t grid_blocks[3];      /* workgroups this engine launches per axis (covers the domain) */
  int32_t shared_size;         /* reference "Shared Size" key (bytes) */
  int32_t lds_bytes;           /* LDS this engine's kernel allocates per workgroup */
  int32_t threads_per_block;   /* this engine's workgroup size */
  int32_t launches;            /* ceil(time_steps / T) */
  int32_t num_rw, num_ro;
  int64_t field_elems;         /* elements of one padded field buffer */
  int64_t stride_y;            /* elements between rows    (= padded_size[0]) */
  int64_t stride_z;            /* elements between planes  (= padded_size[0]*padded_size[1]) */
  int64_t origin;              /* element index of interior point (0,0,0) */
  double total_fp_per_block;   /* printProgramParameters keys (ProgramGenerator.hpp:46-66) */
  double useful_fp_per_block;
  double useful_ratio;
  double total_fp_steps;        /* "Total FP" / "Useful FP" keys of drivers that scale by
                                  time_steps instead of T (poisson-2d.cpp, cdsc-tv-update-2d.cpp) */
  double useful_fp_steps;
  int64_t global_loads_per_block;
  int64_t global_stores_per_block;
  int64_t shared_loads_per_block;
  int64_t shared_stores_per_block;
  int32_t registers;           /* VGPRs per lane of the plan's main kernel ("Register
                                  Usage"); 0 from ot_derive, filled by ot_plan_create */
  int32_t reserved0;
} ot_derived;

typedef struct ot_timing {
  double wall_s;          /* host wall clock around the launch loop ("Elapsed Time") */
  double event_s;         /* HIP events first launch start -> last launch end ("EventElapsed") */
  double kernel_s;        /* OT_RUN_TIME_LAUNCHES: sum of per-launch event durations */
  double kernel_full_s;   /* ... of the launches advancing a full T steps */
  int32_t launches;       /* kernel launches issued */
  int32_t full_launches;  /* launches advancing a full T steps */
  int32_t final_buffer;   /* 0: result is in the rw_a buffers, 1: in rw_b */
  int32_t reserved;
} ot_timing;

/* ot_run mode bits */
enum {
  OT_RUN_GRAPH = 1,         /* capture the loop into a hipGraph once, replay it */
  OT_RUN_TIME_LAUNCHES = 2, /* bracket every launch with HIP events (eager only) */
  OT_RUN_PER_IMAGE = 4,     /* ot_batch_run: one per-grid launch per image and time
                               tile instead of the work-queue kernel (baseline) */
  OT_RUN_ASYNC = 8          /* ot_run: enqueue the loop (or the graph launch) and
                               return without waiting or timing; the caller
                               synchronises the stream (back-to-back solves) */
};

typedef struct ot_plan ot_plan;

/* ---- library / device ---------------------------------------------------- */
const char* ot_version(void);
const char* ot_last_error(void);
/* Device properties of the current HIP device (no-op error on a host without GPU). */
ot_status ot_device_query(int32_t* num_cus, int32_t* lds_per_block, int32_t* max_threads,
                          char* name, int32_t name_len);

/* ---- stencil catalogue --------------------------------------------------- */
ot_status ot_stencil_query(int32_t stencil, ot_stencil_info* info);
ot_status ot_stencil_from_name(const char* name, int32_t* stencil);

/* ---- geometry (pure host code; replaces GeneratorParams::computeDerived,
 *      e.g. opencl/jacobi-2d/jacobi-2d.cpp:84-126, jacobi-3d.cpp:81-115) -------- */
ot_status ot_params_default(int32_t stencil, ot_params* params);
ot_status ot_derive(const ot_params* params, ot_derived* derived);

/* ---- plans (replace XGenerator::generate + cl::Program::build,
 *      opencl/jacobi-2d/jacobi-2d.cpp:162-173, :741-760) ----------------------- */
ot_status ot_plan_create(const ot_params* params, ot_plan** plan);
ot_status ot_plan_derived(const ot_plan* plan, ot_derived* derived);
void ot_plan_destroy(ot_plan* plan);
/* The plan's own non-blocking stream (created on first use), on which
 * ot_run(..., stream = NULL, OT_RUN_GRAPH) replays its graph. */
ot_status ot_plan_stream(ot_plan* plan, void** stream);

/* Run the whole time loop (replaces opencl/jacobi-2d/jacobi-2d.cpp:903-933):
 * ceil(steps/T) launches ping-ponging between the rw_a and rw_b field sets,
 * the last launch guarded to the remaining steps. rw_a holds the initial state;
 * read-only fields (ro) are never swapped. stream is a hipStream_t (NULL = the
 * null stream). mode is a mask of OT_RUN_* bits: OT_RUN_GRAPH captures the
 * loop once into a hipGraph and replays it on later calls with the same
 * buffers and stream; OT_RUN_TIME_LAUNCHES times every launch with its own
 * event pair. On return the stream is synchronised and timing (optional) is
 * filled — except with OT_RUN_ASYNC, which only enqueues (timing then holds
 * the launch counts and final_buffer). With OT_RUN_GRAPH and a NULL stream the
 * graph runs on a stream the plan owns; ot_plan_stream returns it. */
ot_status ot_run(ot_plan* plan, float* const* rw_a, float* const* rw_b,
                 const float* const* ro, void* stream, int32_t mode,
                 ot_timing* timing);

/* Enqueue exactly one overlapped-tile launch advancing min(T, steps-base_time)
 * steps from src to dst (the generated kernel_func(input, aux..., output,
 * baseTime) of the reference). Asynchronous on stream. */
ot_status ot_launch(ot_plan* plan, const float* const* rw_src, float* const* rw_dst,
                    const float* const* ro, int32_t base_time, void* stream);

/* Clock stamps of one wavefront of the last launch of a plan created with
 * OT_VARIANT_DUMP_CLOCKS (the reference's ClockData, jacobi-2d.cpp:854-858):
 * s_memtime (shader clock) at P0 kernel entry, P1 after the index set-up, P2
 * after the global loads and the LDS publish, P3 after the on-chip time steps
 * and P4 after the stores, plus s_memrealtime (constant 100 MHz) at P0 and P4
 * and the hardware ids of the wave's slot.
 * Wave-pipelined 3D kernel (one record per wave, slot = wave index): block_x,
 * block_y = the wave's x and y tile, wave = its z-chunk; P1 after the
 * pipeline fill (the first unrolled block of planes), P2 after the steady-state
 * blocks, P3 after the last plane's stores are issued, P4 at exit. */
typedef struct ot_clock_record {
  int32_t block_x, block_y; /* output tile of the workgroup */
  int32_t wave;             /* wavefront index inside the workgroup */
  uint32_t hw_id;           /* HW_ID register: wave 3:0, SIMD 5:4, CU 11:8, SH 12, SE 15:13 */
  uint32_t xcc_id;          /* XCC_ID register: the accelerator complex die */
  uint32_t valid;           /* 1 when the wave wrote this record */
  uint64_t clock[5];
  uint64_t realtime[2];
} ot_clock_record;

/* Copy the clock records of the plan's last launch to host memory: count
 * receives the number of record slots (2D: 16 per workgroup, slot 16 *
 * workgroup + wave; 3D: one per wave), out gets min(capacity, count) of them
 * (out may be NULL with capacity 0). Slots whose valid is 0 had no wavefront
 * in the last launch. */
ot_status ot_plan_clocks(ot_plan* plan, ot_clock_record* out, int32_t capacity, int32_t* count);

/* Run the same problem with straightforward one-step-per-launch kernels (no
 * tiling, every tap from global memory): the counterpart of the reference CUDA
 * prototypes' *_notimetiling kernels (cuda/jacobi-2d/jacobi-2d.cu:88-105),
 * used by the drivers' --verify and as the no-tiling baseline. params gives
 * the grid and steps (time_tile and blocks only shape the layout). rw_a holds
 * the initial state; timing->final_buffer says where the result is. */
ot_status ot_naive_run(const ot_params* params, float* const* rw_a, float* const* rw_b,
                       const float* const* ro, void* stream, ot_timing* timing);

/* ---- batches of independent 2D images --------------------------------------
 * No reference counterpart (the reference runs one square grid per process).
 * Many images of uneven size (e.g. Rician or TV denoising of a batch of
 * differently sized images) advance together: each time-tile launch is one
 * persistent kernel whose workgroups pull (image, tile) items from a device
 * work queue, and the whole time loop can be replayed from one hipGraph.
 * Images live in arenas (one per field) with a per-image offset and pitch. */
typedef struct ot_image {
  int32_t width, height;  /* domain extents (x, y) */
  int32_t pitch;          /* elements between rows (>= width) */
  int32_t reserved;
  int64_t offset;         /* element index of pixel (0,0) inside the arena */
} ot_image;

typedef struct ot_batch ot_batch;

/* Pack images into one arena: pitch = width rounded up to 16 elements, every
 * image starting 64-byte aligned. images (out) has n_images entries. */
ot_status ot_batch_layout(int32_t n_images, const int32_t* widths, const int32_t* heights,
                          ot_image* images, int64_t* arena_elems);
/* params: a 2D stencil with time_steps, time_tile, block_x, block_y,
 * elems_per_thread, coef, variant (problem sizes are ignored). */
ot_status ot_batch_create(const ot_params* params, int32_t n_images, const ot_image* images,
                          ot_batch** batch);
ot_status ot_batch_info(const ot_batch* batch, int32_t* items, int32_t* workgroups,
                        int32_t* launches);
/* rw_a/rw_b/ro: arena pointers per field, as in ot_run; mode: OT_RUN_* bits. */
ot_status ot_batch_run(ot_batch* batch, float* const* rw_a, float* const* rw_b,
                       const float* const* ro, void* stream, int32_t mode, ot_timing* timing);
void ot_batch_destroy(ot_batch* batch);
/* The same batched solve as ot_batch_create + ot_batch_run (work-queue
 * launches, identical results) from an entry point that allocates nothing. No
 * counterpart in the reference. For callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc. The work queue (launch counters, image table, items) lives in a
 * fixed 8 MiB device array that every GPU the library runs on holds until the
 * process ends: at most 4096 images, and counters + table + 12 bytes per tile
 * within 8 MiB (about 690,000 tiles); larger batches fail with OT_ERESOURCE.
 * Enqueues on stream and returns without waiting (timing, optional, gets the
 * launch counts and final_buffer); images is the host table of
 * ot_batch_layout, copied at the call. Calls that overlap on one GPU share its
 * array and race: order them on one stream or synchronise between them. */
ot_status ot_batch_run_static(const ot_params* params, int32_t n_images, const ot_image* images,
                              float* const* rw_a, float* const* rw_b, const float* const* ro, void* stream,
                              ot_timing* timing);
/* Seeded test images for the batch workload: per image a piecewise-constant
 * phantom of 8..32 random rectangles with intensities in [0,1], corrupted by
 * Rician noise |phantom + N(0,sigma) + i N(0,sigma)|. Writes the arena. */
ot_status ot_fill_rician_phantom(float* arena, int32_t n_images, const ot_image* images,
                                 uint64_t seed, float sigma);

/* ---- device memory helpers (thin wrappers over the HIP runtime) --------- */
ot_status ot_dev_alloc(size_t bytes, void** ptr);
ot_status ot_dev_free(void* ptr);
ot_status ot_memcpy_h2d(void* dst, const void* src, size_t bytes);
ot_status ot_memcpy_d2h(void* dst, const void* src, size_t bytes);
ot_status ot_memcpy_d2d(void* dst, const void* src, size_t bytes);
ot_status ot_dev_memset(void* ptr, int32_t value, size_t bytes);
ot_status ot_dev_sync(void);

/* ---- seeded input generators (host; padded layout of ot_derived) ------------
 * ot_fill_reference_uniform: the reference drivers' init — srand(seed) then
 * (float)rand()/((float)RAND_MAX+1.0f) for every interior point in row-major
 * padded order, 0 in the padding ring (opencl/poisson-2d/poisson-2d.cpp:611-621).
 * Counter-based generators (seed, index) -> value are used for the larger
 * synthetic workloads, so any sub-block can be generated independently. */
ot_status ot_fill_reference_uniform(float* host, const ot_derived* d, uint32_t seed);
ot_status ot_fill_uniform(float* host, const ot_derived* d, uint64_t seed, float lo, float hi);
ot_status ot_fill_sinusoid_noise(float* host, const ot_derived* d, uint64_t seed, float noise_amp);
ot_status ot_fill_gaussian_pulse(float* host, const ot_derived* d, float amplitude, float sigma);

/* ---- self-test ------------------------------------------------------------
 * No reference counterpart. The kernels evaluate 1.0f/sqrtf(x) (the reference
 * CPU loop's rsqrt, gradient-2d.cpp:697-699) from hardware estimates plus
 * corrections; this entry point evaluates that sequence and the IEEE
 * expression on every float bit pattern in [lo_bits, hi_bits] on the device
 * and returns how many differ bit for bit (NaN == NaN) and the smallest
 * differing pattern (0xFFFFFFFF if none). */
ot_status ot_selftest_rsq(uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches, uint32_t* first_bad);

#ifdef __cplusplus
}
#endif

#endif /* OT_OT_H */
