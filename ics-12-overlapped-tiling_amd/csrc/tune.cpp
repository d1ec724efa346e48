// tune.cpp — configuration space, measurement, exhaustive autotuner and the
// analytic performance model (host code over the C ABI).
//
// The reference's tuner is a Python 2 harness that expands a YAML sweep and
// spawns one driver process per configuration under a watchdog, scraping the
// driver's stdout keys into a results file (scripts/utils.py:36-244,
// scripts/benchrunner/benchmarks.py:7-30). Here the sweep runs in-process:
// plans are cheap (no source generation or JIT), so a configuration costs only
// its solves, and the results file is a CSV whose columns are the driver's keys.
#include "ot/ot.h"
#include "ot/ot_tune.h"
#include "internal.hpp"
#include "traits.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <tuple>
#include <vector>

namespace {

using ot::kTraits;

struct Cand {
  int T, bx, by, bz, e;
};

// Sweep ranges per dimensionality. They cover every instantiation compiled in
// k1d.hip (ept 1..8), k2d.hip (ept 1..12), k3d_impl.hpp (T 1..8, ept 1/2/4, 8
// for T <= 2) and k3w_impl.hpp (T 1..5, footprint rows 8..32).
void candidates(const ot_params& base, int dims, std::vector<Cand>& out) {
  const int stencil = base.stencil;
  out.clear();
  if (dims == 1) {
    for (int T = 1; T <= 8; ++T)
      for (int bx : {64, 128, 256, 512, 1024})
        for (int e : {1, 2, 4, 8}) out.push_back({T, bx, 1, 1, e});
  } else if (dims == 2) {
    for (int T = 1; T <= 8; ++T)
      for (int bx : {16, 32, 64, 128})
        for (int by : {1, 2, 4, 8, 16})
          for (int e : {1, 2, 4, 8}) out.push_back({T, bx, by, 1, e});
    // wave-per-tile kernel: x tile + 2*radius*T fills the 64 lanes
    const int R = kTraits[stencil].radius;
    for (int T = 1; T <= 8; ++T)
      for (int fr : {16, 24, 32, 64})
        if (fr > 2 * R * T) out.push_back({T, 64 - 2 * R * T, fr - 2 * R * T, 1, 1});
    // row-streaming kernel: the same x tiles streaming chunks of 64..512 rows
    for (int T = 1; T <= (stencil == OT_RICIAN2D ? 5 : 8); ++T)
      for (int rows : {64, 128, 256, 512}) out.push_back({T, 64 - 2 * R * T, rows, 1, 1});
  } else {
    // LDS kernel: T 1..8 and tiles from 8 x 8 x 4 to 64 x 16 x 32 (x, y rows
    // of ept, z-chunk), which holds the reference's J3D sweep grid
    // (scripts/c2050/j3d.yml:7-13: block 8..32 x 2..8 x 2..8, T 1..4)
    for (int T = 1; T <= 8; ++T)
      for (int bx : {8, 16, 32, 64})
        for (int by : {2, 4, 8, 16})
          for (int bz : {4, 8, 16, 32})
            for (int e : {1, 2, 4, 8})
              if (e < 8 || T <= 2) out.push_back({T, bx, by, bz, e});
    // wave-pipelined kernel: x tile + 2T fills the 64 lanes, footprint rows FR.
    // Besides fixed z-chunks, the chunks that make the wave count just fill k
    // waves per SIMD (k = 1..6, 4 SIMDs per CU of the default machine), so that
    // no SIMD runs one wave more than most others.
    ot_arch arch;
    ot_arch_default(&arch);
    const int64_t simds = 4 * (int64_t)arch.num_cu;
    const int nx = base.nx > 0 ? base.nx : base.problem_size, ny = base.ny > 0 ? base.ny : base.problem_size;
    const int nz = base.nz > 0 ? base.nz : base.problem_size;
    for (int T = 1; T <= 5; ++T)
      for (int fr : {8, 12, 16, 20, 24, 32})
        if (fr > 2 * T && 3 * T * fr <= ot::kW3RegBudget) {
          const int tx = 64 - 2 * T, ty = fr - 2 * T;
          std::vector<int> bzs = {8, 16, 20, 24, 32};
          const int64_t tiles_xy = (int64_t)((nx + tx - 1) / tx) * ((ny + ty - 1) / ty);
          for (int k = 1; k <= 6 && nx > 0 && ny > 0 && nz > 0; ++k) {
            const int64_t chunks = k * simds / tiles_xy;
            if (chunks < 1) continue;
            const int bz = (int)((nz + chunks - 1) / chunks);
            if (bz >= 4 && std::find(bzs.begin(), bzs.end(), bz) == bzs.end()) bzs.push_back(bz);
          }
          for (int bz : bzs) out.push_back({T, tx, ty, bz, 1});
        }
  }
}

ot_params with_cand(const ot_params& base, const Cand& c) {
  ot_params p = base;
  p.time_tile = c.T;
  p.block_x = c.bx;
  p.block_y = c.by;
  p.block_z = c.bz;
  p.elems_per_thread = c.e;
  return p;
}

double log2d(double v) { return v > 0 ? std::log2(v) : 0.0; }

// ---- results file ---------------------------------------------------------
const char* kCsvHeader =
    "Stencil,Dimensions,Nx,Ny,Nz,Time Steps,Geometry,Time Tile Size,Block Size X,Block Size Y,"
    "Block Size Z,Elements/Thread,Num Blocks X,Num Blocks Y,Num Blocks Z,Threads,LDS Bytes,"
    "Shared Size,Register Usage,Wave Kernel,EventElapsed,Status";

using Key = std::tuple<std::string, int, int, int, int, int, int, int, int, int, int>;

Key key_of(const ot_params& p, const ot_derived& d) {
  return Key(kTraits[p.stencil].name, d.real_size[0], d.real_size[1], d.real_size[2], p.time_steps, p.geometry,
             p.time_tile, p.block_x, p.block_y, p.block_z, p.elems_per_thread);
}

std::vector<std::string> split_csv(const std::string& line) {
  std::vector<std::string> out;
  size_t s = 0;
  while (true) {
    const size_t c = line.find(',', s);
    out.push_back(line.substr(s, c == std::string::npos ? std::string::npos : c - s));
    if (c == std::string::npos) break;
    s = c + 1;
  }
  return out;
}

}  // namespace

extern "C" {

ot_status ot_tune_space(const ot_params* base, ot_params* out, int32_t capacity, int32_t* count) {
  if (!base || !count || capacity < 0 || (capacity > 0 && !out))
    return ot::set_error(OT_EINVAL, "ot_tune_space: null argument");
  if (base->stencil < 0 || base->stencil >= OT_NUM_STENCILS)
    return ot::set_error(OT_EINVAL, "ot_tune_space: unknown stencil %d", base->stencil);
  std::vector<Cand> cands;
  candidates(*base, kTraits[base->stencil].dims, cands);
  int n = 0;
  for (const Cand& c : cands) {
    ot_params p = with_cand(*base, c);
    ot_derived d;
    if (ot_derive(&p, &d) != OT_OK) continue;  // inconsistent tile or over the LDS / thread limits
    if (n < capacity) out[n] = p;
    ++n;
  }
  *count = n;
  return OT_OK;
}

ot_status ot_features(const ot_params* p, float* f) {
  if (!p || !f) return ot::set_error(OT_EINVAL, "ot_features: null argument");
  ot_derived d;
  const ot_status st = ot_derive(p, &d);
  if (st != OT_OK) return st;
  const ot::StencilTraits& tr = kTraits[p->stencil];
  for (int i = 0; i < OT_FEATURES; ++i) f[i] = 0.0f;
  // one-hot over the 13 stencils the model was trained on; cs-3d has no slot
  // and is placed by its descriptors (f[16], f[17], f[24..27]) alone
  if (p->stencil < 13) f[p->stencil] = 1.0f;
  f[13] = (float)log2d(d.real_size[0]);
  f[14] = (float)log2d(d.real_size[1]);
  f[15] = (float)log2d(d.real_size[2]);
  f[16] = (float)tr.radius;
  f[17] = (float)(tr.num_rw + tr.num_ro);
  f[18] = (float)p->time_tile;
  f[19] = (float)log2d(p->block_x);
  f[20] = (float)log2d(p->block_y);
  f[21] = (float)log2d(p->block_z);
  f[22] = (float)log2d(p->elems_per_thread);
  f[23] = ((d.dims == 3 && ot::k3w_rows(p->time_tile, d.tile) > 0) ||
           (d.dims == 2 && (ot::k2w_rows(p->stencil, p->time_tile, d.tile) > 0 ||
                            ot::k2s_tile(p->stencil, p->time_tile, d.tile))))
              ? 1.0f
              : 0.0f;
  // stencil descriptors, so that a stencil absent from the training data is
  // still placed among the others, and the analytic model as a prior
  f[24] = (float)d.dims;
  f[25] = (float)tr.flops;
  f[26] = (float)tr.f_p2_gl;
  ot_arch arch;
  ot_arch_default(&arch);
  double est = 0.0;
  if (ot_model_analytic(p, &arch, 0, &est) == OT_OK && est > 0.0) {
    const double pts = (double)d.real_size[0] * d.real_size[1] * d.real_size[2];
    f[27] = (float)std::log(est / (pts * std::max(1, p->time_steps)));
  }
  return OT_OK;
}

ot_status ot_tune_measure(const ot_params* p, float* const* rw_a, float* const* rw_b, const float* const* ro,
                          int32_t warmup, int32_t reps, double time_limit_s, double* median_s,
                          int32_t* registers, int32_t* timed_out) {
  if (!p || !median_s || reps < 1 || warmup < 0) return ot::set_error(OT_EINVAL, "ot_tune_measure: bad argument");
  if (timed_out) *timed_out = 0;
  ot_plan* plan = nullptr;
  ot_status st = ot_plan_create(p, &plan);
  if (st != OT_OK) return st;
  if (registers) {
    ot_derived d;
    ot_plan_derived(plan, &d);
    *registers = d.registers;
  }
  std::vector<double> times;
  ot_timing tm;
  for (int i = 0; i < warmup + reps; ++i) {
    st = ot_run(plan, rw_a, rw_b, ro, nullptr, 0, &tm);
    if (st != OT_OK) break;
    if (i == 0 && time_limit_s > 0 && tm.event_s > time_limit_s) {
      // watchdog: one solve already exceeds the budget of a configuration
      times.assign(1, tm.event_s);
      if (timed_out) *timed_out = 1;
      break;
    }
    if (i >= warmup) times.push_back(tm.event_s);
  }
  ot_plan_destroy(plan);
  if (st != OT_OK) return st;
  std::sort(times.begin(), times.end());
  const size_t m = times.size();
  *median_s = (m % 2) ? times[m / 2] : 0.5 * (times[m / 2 - 1] + times[m / 2]);
  return OT_OK;
}

ot_status ot_tune_exhaustive(const ot_params* base, const char* csv_path, int32_t warmup, int32_t reps,
                             double time_limit_s, ot_params* best, double* best_s, int32_t* measured) {
  if (!base || !csv_path) return ot::set_error(OT_EINVAL, "ot_tune_exhaustive: null argument");
  int32_t n = 0;
  ot_status st = ot_tune_space(base, nullptr, 0, &n);
  if (st != OT_OK) return st;
  std::vector<ot_params> space(n);
  ot_tune_space(base, space.data(), n, &n);

  // resume: configurations already in the file are not measured again
  std::set<Key> done;
  double best_time = 1e300;
  ot_params best_p = *base;
  bool file_exists = false;
  if (FILE* in = std::fopen(csv_path, "r")) {
    file_exists = true;
    char buf[4096];
    bool header = true;
    while (std::fgets(buf, sizeof(buf), in)) {
      std::string line(buf);
      while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
      if (header) {
        header = false;
        continue;
      }
      const auto c = split_csv(line);
      if (c.size() < 22) continue;
      ot_params q = *base;
      if (ot_stencil_from_name(c[0].c_str(), &q.stencil) != OT_OK) continue;
      const int nx = std::atoi(c[2].c_str()), ny = std::atoi(c[3].c_str()), nz = std::atoi(c[4].c_str());
      q.time_steps = std::atoi(c[5].c_str());
      q.geometry = std::atoi(c[6].c_str());
      q.time_tile = std::atoi(c[7].c_str());
      q.block_x = std::atoi(c[8].c_str());
      q.block_y = std::atoi(c[9].c_str());
      q.block_z = std::atoi(c[10].c_str());
      q.elems_per_thread = std::atoi(c[11].c_str());
      done.insert(Key(c[0], nx, ny, nz, q.time_steps, q.geometry, q.time_tile, q.block_x, q.block_y, q.block_z,
                      q.elems_per_thread));
      ot_derived dq;
      if (c[21] == "ok" && ot_derive(&q, &dq) == OT_OK && dq.real_size[0] == nx && dq.real_size[1] == ny &&
          dq.real_size[2] == nz && q.stencil == base->stencil) {
        const double t = std::atof(c[20].c_str());
        if (t < best_time) {
          best_time = t;
          best_p = q;
        }
      }
    }
    std::fclose(in);
  }
  FILE* out = std::fopen(csv_path, "a");
  if (!out) return ot::set_error(OT_EINVAL, "ot_tune_exhaustive: cannot open %s", csv_path);
  if (!file_exists) std::fprintf(out, "%s\n", kCsvHeader);

  const ot::StencilTraits& tr = kTraits[base->stencil];
  const int nfields = 2 * tr.num_rw + tr.num_ro;
  std::vector<float*> bufs(nfields, nullptr);
  int64_t cap_elems = 0;
  int count = 0;
  st = OT_OK;
  for (const ot_params& p : space) {
    ot_derived d;
    if (ot_derive(&p, &d) != OT_OK) continue;
    if (done.count(key_of(p, d))) continue;
    if (d.field_elems > cap_elems) {
      // (re)allocate and seed every field for this layout
      for (float*& b : bufs)
        if (b) ot_dev_free(b), b = nullptr;
      std::vector<float> host((size_t)d.field_elems);
      ot_fill_uniform(host.data(), &d, 4242, 0.0f, 1.0f);
      for (float*& b : bufs) {
        st = ot_dev_alloc((size_t)d.field_elems * 4, (void**)&b);
        if (st != OT_OK) break;
        st = ot_memcpy_h2d(b, host.data(), (size_t)d.field_elems * 4);
        if (st != OT_OK) break;
      }
      if (st != OT_OK) break;
      cap_elems = d.field_elems;
    }
    float* a[3] = {nullptr, nullptr, nullptr};
    float* b[3] = {nullptr, nullptr, nullptr};
    const float* ro[2] = {nullptr, nullptr};
    for (int i = 0; i < tr.num_rw; ++i) {
      a[i] = bufs[2 * i];
      b[i] = bufs[2 * i + 1];
    }
    for (int i = 0; i < tr.num_ro; ++i) ro[i] = bufs[2 * tr.num_rw + i];
    double t = 0.0;
    int32_t regs = 0, to = 0;
    const ot_status ms = ot_tune_measure(&p, a, b, ro, warmup, reps, time_limit_s, &t, &regs, &to);
    const char* status = ms != OT_OK ? "error" : (to ? "timeout" : "ok");
    std::fprintf(out, "%s,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%.9g,%s\n", tr.name, d.dims,
                 d.real_size[0], d.real_size[1], d.real_size[2], p.time_steps, p.geometry, p.time_tile, p.block_x,
                 p.block_y, p.block_z, p.elems_per_thread, d.grid_blocks[0], d.grid_blocks[1], d.grid_blocks[2],
                 d.threads_per_block, d.lds_bytes, d.shared_size, regs,
                 ((d.dims == 3 && ot::k3w_rows(p.time_tile, d.tile) > 0) ||
                  (d.dims == 2 && (ot::k2w_rows(p.stencil, p.time_tile, d.tile) > 0 ||
                                   ot::k2s_tile(p.stencil, p.time_tile, d.tile))))
                     ? 1
                     : 0,
                 ms == OT_OK ? t : -1.0, status);
    std::fflush(out);
    ++count;
    if (ms == OT_OK && !to && t < best_time) {
      best_time = t;
      best_p = p;
    }
  }
  std::fclose(out);
  for (float* b : bufs)
    if (b) ot_dev_free(b);
  if (st != OT_OK) return st;
  if (best) *best = best_p;
  if (best_s) *best_s = best_time < 1e300 ? best_time : -1.0;
  if (measured) *measured = count;
  return OT_OK;
}

// ---- analytic model -----------------------------------------------------------

ot_status ot_arch_default(ot_arch* a) {
  if (!a) return ot::set_error(OT_EINVAL, "ot_arch_default: null argument");
  // MI355X: 256 CUs in 8 XCDs, 4 SIMDs per CU, 2.4 GHz, 160 KiB LDS and 32 KiB
  // vector L1 per CU, 512 VGPRs per SIMD lane, 8 TB/s HBM3E.
  a->total_lds = 160.0 * 1024.0;
  a->max_waves_per_cu = 32.0;        // 8 per SIMD
  a->max_threads_per_cu = 2048.0;
  a->max_blocks_per_cu = 32.0;
  a->max_regs_per_cu = 512.0 * 64.0 * 4.0;
  a->num_cu = 256.0;
  a->clock_hz = 2.4e9;
  // a CU issues about one wave instruction per cycle across its 4 SIMDs (a
  // wave64 ALU instruction occupies its SIMD for 4 cycles)
  a->c_load = 1.0;
  a->c_op = 1.0;
  // per-CU share of 8 TB/s is 13 B/cycle: a 256-byte wave access takes ~19.7
  // cycles; LDS delivers 128 B/cycle per CU
  a->b_gmem = 256.0 / (8e12 / 256.0 / 2.4e9);
  a->b_smem = 2.0;
  a->l_gmem = 1200.0;
  a->l_smem = 64.0;
  a->launch_penalty = 2.4e9 * 4e-6;  // ~4 us between dependent launches
  a->wave_size = 64.0;
  a->max_regs_per_thread = 512.0;
  a->l1_bytes = 32.0 * 1024.0;
  a->sfu_cycles = 4.0;               // transcendentals issue at quarter rate
  return OT_OK;
}

ot_status ot_model_analytic(const ot_params* p, const ot_arch* A, int32_t registers, double* seconds) {
  if (!p || !A || !seconds) return ot::set_error(OT_EINVAL, "ot_model_analytic: null argument");
  ot_derived d;
  const ot_status st = ot_derive(p, &d);
  if (st != OT_OK) return st;
  const ot::StencilTraits& tr = kTraits[p->stencil];
  const double x = p->block_x, y = p->block_y, z = (d.dims == 3) ? p->block_z : 1.0;
  const double e = p->elems_per_thread, t = p->time_tile;
  const int dim = d.dims;

  // occupancy (run-model.py:9-58), with this engine's workgroup size and LDS
  // allocation in place of the reference kernel's estimates
  const double waves_per_block = std::ceil(d.threads_per_block / A->wave_size);
  const double shared_per_block = d.lds_bytes;
  const double regs_per_block = (double)registers * d.threads_per_block;
  const double from_waves = std::floor(A->max_waves_per_cu / waves_per_block);
  const double from_shared = shared_per_block > 0 ? std::floor(A->total_lds / shared_per_block) : 1e30;
  const double from_regs = regs_per_block > 0 ? std::floor(A->max_regs_per_cu / regs_per_block) : 1e30;
  const double total_blocks = (double)d.grid_blocks[0] * d.grid_blocks[1] * d.grid_blocks[2];
  double blocks_per_cu =
      std::min(std::min(from_waves, from_shared), std::min(from_regs, std::ceil(total_blocks / A->num_cu)));
  blocks_per_cu = std::max(std::min(blocks_per_cu, A->max_blocks_per_cu), 1.0);
  const double W = waves_per_block * blocks_per_cu;  // active waves per CU
  const double num_stages = std::ceil(total_blocks / (blocks_per_cu * A->num_cu));

  const int steps = p->time_steps;
  const double full_inv = steps / p->time_tile;
  const int extra = steps % p->time_tile;
  const double extra_global = extra > 0 ? 1.0 : 0.0;
  const double extra_shared = std::max(0, extra - 1);

  // 'fermi' cache model (run-model.py:314-433): halo loads of the global phase
  // that hit in L1, by the reference's per-stencil-shape rules
  const double p2gl = tr.f_p2_gl, p2sh = tr.f_p2_sh;
  double p2_glb = p2gl * e;
  auto hit_rule = [&](double miss_num, double footprint_bytes) {
    const double miss_rate = miss_num / (p2gl * e);
    const double cap_ratio = A->l1_bytes * 0.5 / footprint_bytes;
    double hit = 1.0 - miss_rate;
    if (cap_ratio < 1.0) hit *= cap_ratio;
    return (1.0 - hit) * p2gl * e;
  };
  if (dim == 3 && p2gl == 7.0)
    p2_glb = hit_rule(3.0 * e + 4.0, (y * e + 2) * (x + 2) * (z + 2) * 4.0);
  else if (dim == 3 && p2gl == 27.0)
    p2_glb = 9.0 * e + 18.0;
  else if (dim == 2 && p2gl == 5.0)
    p2_glb = hit_rule(e + 4.0, (y * e + 2) * (x + 2) * 4.0);
  else if (dim == 2 && p2gl == 9.0)
    p2_glb = std::min(3 * e + 6.0, p2gl * e);
  else if (dim == 2 && p2gl == 7.0)
    p2_glb = hit_rule(2.0 * e + 5.0, (y * e + 2) * (x + 2) * (z + 2) * 4.0);
  else if (dim == 2 && p2gl == 13.0)
    p2_glb = std::min(e + 4.0, p2gl * e);
  // other shapes (1D, FDTD's 11 loads): no cache hits — the reference script
  // stops with "Dont know how to apply cache model" there
  const double p2_shd = p2sh * e + (p2gl * e - p2_glb);

  const double k_op = (tr.f_compute - 1.0) * e;
  const double k_sfu = e;
  const double Bprime = A->b_gmem + (1.0 + 1.0 / (x / A->wave_size));
  const double p3 = tr.f_p3_sh * e, p4 = tr.f_p4_gs;

  const double t_glb = std::max(A->c_load * (p2_glb + p2_shd) * W + A->c_op * k_op * W + 4.0 * A->sfu_cycles * k_sfu * W,
                                A->l_gmem + A->c_load + std::max(Bprime * p2_glb * W, A->b_smem * p2_shd * W));
  const double t_shd = std::max(A->c_load * p3 * W + A->c_op * k_op * W + A->sfu_cycles * k_sfu * W,
                                A->l_smem + A->c_load + A->b_smem * p3 * W);
  const double write_back = A->l_gmem + A->c_load + A->b_gmem * p4 * W;
  const double t_phase1 = (10.0 * e + 20.0) * W;

  double t_stage, t_stage_extra;
  switch (p->phase_limit) {
    case 1:
      t_stage = t_stage_extra = 0.0;
      break;
    case 2:
      t_stage = t_glb + write_back;
      t_stage_extra = extra_global * t_glb;
      break;
    case 3:
      t_stage = (t - 1.0) * t_shd;
      t_stage_extra = extra_shared * t_shd;
      break;
    default:
      t_stage = t_glb + (t - 1.0) * t_shd + write_back;
      t_stage_extra = extra_global * t_glb + extra_shared * t_shd + extra_global * write_back;
  }
  t_stage += t_phase1;
  if (extra_global > 0) t_stage_extra += t_phase1;
  double clk = (t_stage * num_stages + A->launch_penalty) * full_inv;
  if (extra_global > 0) clk += t_stage_extra * num_stages + A->launch_penalty;
  *seconds = clk / A->clock_hz;
  return OT_OK;
}

}  // extern "C"
