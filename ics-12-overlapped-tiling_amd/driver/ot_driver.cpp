// ot_driver.cpp — the ot-<stencil> benchmark drivers.
//
// Drop-in for the reference's per-stencil driver programs (main() of
// opencl/<stencil>/<stencil>.cpp, e.g. opencl/jacobi-2d/jacobi-2d.cpp:553-1048):
// same command-line options (boost::program_options block, jacobi-2d.cpp:566-607)
// and the same "Key: value" stdout lines (CLCommon.hpp:23-27 printValue), which
// the reference's sweep scripts parse (scripts/utils.py:154-155). The stencil is
// taken from the program name (ot-jacobi-2d -> jacobi-2d) or --stencil.
//
// Differences from the reference, by design:
//   * kernels are precompiled for gfx950; --load-kernel/--save-kernel read and
//     write a plan description (the generator parameters) instead of OpenCL C;
//   * the default geometry is the reference's (truncated realSize, padded
//     layout, srand(123456) init); --geometry tile selects the engine's tiles;
//   * --verify compares the overlapped-tiled result with the non-tiled
//     one-step-per-launch kernels (ot_naive_run) using the reference's L2 test
//     (compareResults, jacobi-2d.cpp:517-551), and also prints the max relative
//     error;
//   * --time-only gmem|instr|smem and -c/--dump-clocks run timing-probe
//     instantiations of the 2D LDS kernel (jacobi-2d and gradient-2d, the
//     drivers that have --time-only); -c also of the wave-pipelined jacobi-3d
//     kernel (bx, by = the wave's x and y tile, warp = its z-chunk); the clock
//     CSV goes to stderr as in the reference (jacobi-2d.cpp:957-997), one row
//     per wave of one compute unit, stamps from s_memtime; --break-threads is
//     accepted and has no effect;
//   * a malformed option value is an error (usage, exit status 1).
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdint>
#include <stdexcept>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "ot/ot.h"

namespace {

template <typename T>
void printValue(const char* name, T value) {
  std::cout << name << ": " << value << "\n";
}

struct Options {
  std::string stencil;
  ot_params p{};
  bool verify = false, help = false, graph = false, clocks = false;
  std::string load, save, geometry = "reference", time_only;
  uint64_t seed = 123456;
};

// Option values: the whole text must be the number (std::stoi alone accepts
// "12abc"); anything else throws std::invalid_argument / std::out_of_range.
int32_t to_int(const std::string& v) {
  size_t pos = 0;
  const long x = std::stol(v, &pos);
  if (pos != v.size() || x < INT32_MIN || x > INT32_MAX) throw std::invalid_argument(v);
  return (int32_t)x;
}
float to_float(const std::string& v) {
  size_t pos = 0;
  const float x = std::stof(v, &pos);
  if (pos != v.size()) throw std::invalid_argument(v);
  return x;
}
uint64_t to_u64(const std::string& v) {
  size_t pos = 0;
  if (!v.empty() && v[0] == '-') throw std::invalid_argument(v);
  const unsigned long long x = std::stoull(v, &pos);
  if (pos != v.size()) throw std::invalid_argument(v);
  return (uint64_t)x;
}
// boost::program_options' bool syntax (the reference's -c arg)
bool parse_bool(const std::string& v) {
  if (v == "1" || v == "true" || v == "on" || v == "yes") return true;
  if (v == "0" || v == "false" || v == "off" || v == "no") return false;
  throw std::invalid_argument(v);
}

int usage(const char* prog) {
  std::cerr << "Usage: " << prog << " [options]\n"
            << "Options:\n"
               "  -h [ --help ]                      Show usage information\n"
               "  -n [ --problem-size ] arg          Set problem size\n"
               "  -t [ --time-steps ] arg (=64)      Set number of time steps\n"
               "  -x [ --block-size-x ] arg          Set block size (X)\n"
               "  -y [ --block-size-y ] arg          Set block size (Y)\n"
               "  -z [ --block-size-z ] arg          Set block size (Z)\n"
               "  -e [ --elements-per-thread ] arg (=1) Set elements per thread\n"
               "  -s [ --time-tile-size ] arg (=1)   Set time tile size\n"
               "  -p [ --phase-limit ] arg (=0)      Stop after a certain kernel phase\n"
               "  -f [ --load-kernel ] arg           Load plan description from disk\n"
               "  -w [ --save-kernel ] arg           Save plan description to disk\n"
               "  -v [ --verify ]                    Verify results\n"
               "  -c [ --dump-clocks ] arg (=0)      Dump per-wave phase clocks to stderr (2D LDS kernel,\n"
               "                                     jacobi-3d wave kernel)\n"
               "  -b [ --break-threads ]             Accepted for compatibility\n"
               "  --time-only arg                    Time one phase only: gmem | instr | smem (2D LDS kernel)\n"
               "  --stencil arg                      Stencil name (default: from program name)\n"
               "  --geometry arg (=reference)        reference | tile\n"
               "  --nx / --ny / --nz arg             Per-axis extents (tile geometry)\n"
               "  --graph                            Replay the time loop from a hipGraph\n"
               "  --alpha / --sigma2 / --lambda / --dt arg  Stencil scalars (TV alpha; Rician SIGMA2, LAMBDA, DT)\n"
               "  --fdtd-physical                    Hz update with the curl sign\n"
               "  --seed arg (=123456)               srand() seed of the input\n";
  return 1;
}

int parse_args(int argc, char** argv, Options& o);

int parse(int argc, char** argv, Options& o) {
  try {
    return parse_args(argc, argv, o);
  } catch (const std::exception& e) {
    std::cerr << "invalid option value: " << e.what() << "\n";
    return -1;
  }
}

int parse_args(int argc, char** argv, Options& o) {
  std::map<std::string, std::string> alias = {
      {"-h", "help"}, {"-n", "problem-size"}, {"-t", "time-steps"}, {"-x", "block-size-x"},
      {"-y", "block-size-y"}, {"-z", "block-size-z"}, {"-e", "elements-per-thread"},
      {"-s", "time-tile-size"}, {"-p", "phase-limit"}, {"-f", "load-kernel"}, {"-w", "save-kernel"},
      {"-v", "verify"}, {"-c", "dump-clocks"}, {"-b", "break-threads"}};
  const std::vector<std::string> flags = {"help", "verify", "break-threads", "graph", "fdtd-physical"};
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i], key, val;
    if (alias.count(a)) key = alias[a];
    else if (a.rfind("--", 0) == 0) {
      key = a.substr(2);
      const size_t eq = key.find('=');
      if (eq != std::string::npos) {
        val = key.substr(eq + 1);
        key = key.substr(0, eq);
      }
    } else {
      std::cerr << "unrecognised argument " << a << "\n";
      return -1;
    }
    const bool is_flag = std::find(flags.begin(), flags.end(), key) != flags.end();
    if (!is_flag && val.empty()) {
      if (i + 1 >= argc) {
        std::cerr << "missing value for " << a << "\n";
        return -1;
      }
      val = argv[++i];
    }
    ot_params& p = o.p;
    if (key == "help") o.help = true;
    else if (key == "problem-size") p.problem_size = to_int(val);
    else if (key == "time-steps") p.time_steps = to_int(val);
    else if (key == "block-size-x") p.block_x = to_int(val);
    else if (key == "block-size-y") p.block_y = to_int(val);
    else if (key == "block-size-z") p.block_z = to_int(val);
    else if (key == "elements-per-thread") p.elems_per_thread = to_int(val);
    else if (key == "time-tile-size") p.time_tile = to_int(val);
    else if (key == "phase-limit") p.phase_limit = to_int(val);
    else if (key == "load-kernel") o.load = val;
    else if (key == "save-kernel") o.save = val;
    else if (key == "verify") o.verify = true;
    else if (key == "dump-clocks") o.clocks = parse_bool(val);
    else if (key == "break-threads") {
    } else if (key == "time-only") o.time_only = val;
    else if (key == "stencil") o.stencil = val;
    else if (key == "geometry") o.geometry = val;
    else if (key == "nx") p.nx = to_int(val);
    else if (key == "ny") p.ny = to_int(val);
    else if (key == "nz") p.nz = to_int(val);
    else if (key == "graph") o.graph = true;
    else if (key == "alpha") p.coef[0] = to_float(val);
    else if (key == "sigma2") p.coef[0] = to_float(val);
    else if (key == "lambda") p.coef[1] = to_float(val);
    else if (key == "dt") p.coef[2] = to_float(val);
    else if (key == "fdtd-physical") p.variant |= OT_VARIANT_FDTD_PHYSICAL;
    else if (key == "seed") o.seed = to_u64(val);
    else {
      std::cerr << "unrecognised option --" << key << "\n";
      return -1;
    }
  }
  return 0;
}

// Plan description written by --save-kernel / read by --load-kernel.
void save_plan(const std::string& path, const std::string& stencil, const ot_params& p) {
  std::ofstream f(path);
  f << "# ot plan description\nstencil " << stencil << "\nproblem_size " << p.problem_size << "\nnx " << p.nx
    << "\nny " << p.ny << "\nnz " << p.nz << "\ntime_steps " << p.time_steps << "\ntime_tile " << p.time_tile
    << "\nblock_x " << p.block_x << "\nblock_y " << p.block_y << "\nblock_z " << p.block_z
    << "\nelems_per_thread " << p.elems_per_thread << "\ngeometry " << p.geometry << "\nphase_limit "
    << p.phase_limit << "\nvariant " << p.variant << "\ncoef " << p.coef[0] << " " << p.coef[1] << " " << p.coef[2]
    << " " << p.coef[3] << "\n";
}

bool load_plan(const std::string& path, std::string& stencil, ot_params& p) {
  std::ifstream f(path);
  if (!f) return false;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    std::string k;
    is >> k;
    if (k == "stencil") is >> stencil;
    else if (k == "problem_size") is >> p.problem_size;
    else if (k == "nx") is >> p.nx;
    else if (k == "ny") is >> p.ny;
    else if (k == "nz") is >> p.nz;
    else if (k == "time_steps") is >> p.time_steps;
    else if (k == "time_tile") is >> p.time_tile;
    else if (k == "block_x") is >> p.block_x;
    else if (k == "block_y") is >> p.block_y;
    else if (k == "block_z") is >> p.block_z;
    else if (k == "elems_per_thread") is >> p.elems_per_thread;
    else if (k == "geometry") is >> p.geometry;
    else if (k == "phase_limit") is >> p.phase_limit;
    else if (k == "variant") is >> p.variant;
    else if (k == "coef") is >> p.coef[0] >> p.coef[1] >> p.coef[2] >> p.coef[3];
  }
  return true;
}

int fail(const char* what) {
  std::cout << "ERROR: " << what << ": " << ot_last_error() << "\n";
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  std::string prog = argv[0];
  const size_t slash = prog.find_last_of('/');
  if (slash != std::string::npos) prog = prog.substr(slash + 1);
  // stencil from the program name ot-<stencil>, overridable by --stencil
  for (int i = 1; i < argc; ++i)
    if (std::string(argv[i]) == "--stencil" && i + 1 < argc) o.stencil = argv[i + 1];
  if (o.stencil.empty() && prog.rfind("ot-", 0) == 0) o.stencil = prog.substr(3);
  int32_t sid = -1;
  if (o.stencil.empty() || ot_stencil_from_name(o.stencil.c_str(), &sid) != OT_OK) {
    std::cerr << "unknown stencil '" << o.stencil << "' (use --stencil <name> or call as ot-<name>)\n";
    return 1;
  }
  ot_params_default(sid, &o.p);
  if (parse(argc, argv, o) != 0) return usage(prog.c_str());
  if (o.help) return usage(prog.c_str());
  if (!o.time_only.empty()) {
    // jacobi-2d.cpp:622-629
    if (o.time_only == "gmem") o.p.variant |= OT_VARIANT_TIME_ONLY_GMEM;
    else if (o.time_only == "instr") o.p.variant |= OT_VARIANT_TIME_ONLY_INSTR;
    else if (o.time_only == "smem") o.p.variant |= OT_VARIANT_TIME_ONLY_SMEM;
    else {
      std::cerr << "--time-only takes gmem, instr or smem (got '" << o.time_only << "')\n";
      return usage(prog.c_str());
    }
  }
  if (o.clocks) o.p.variant |= OT_VARIANT_DUMP_CLOCKS;
  if (!o.load.empty()) {
    std::string st = o.stencil;
    if (!load_plan(o.load, st, o.p)) {
      std::cout << "ERROR: cannot read " << o.load << "\n";
      return 1;
    }
  } else {
    o.p.geometry = (o.geometry == "tile") ? OT_GEOM_TILE : OT_GEOM_REFERENCE;
  }
  ot_stencil_info info;
  ot_stencil_query(sid, &info);
  if (!o.save.empty()) save_plan(o.save, o.stencil, o.p);

  ot_derived d;
  if (ot_derive(&o.p, &d) != OT_OK) return fail("Consistency error");
  ot_plan* plan = nullptr;
  if (ot_plan_create(&o.p, &plan) != OT_OK) return fail("plan");
  ot_plan_derived(plan, &d);

  // printProgramParameters block (ProgramGenerator.hpp:46-66) and the 3D
  // drivers' own keys (jacobi-3d.cpp:561-622)
  const int dims = d.dims;
  printValue("Problem Size", o.p.problem_size);
  printValue("Time Tile Size", o.p.time_tile);
  printValue("Block Size X", o.p.block_x);
  printValue("Block Size Y", dims > 1 ? o.p.block_y : 1);
  printValue("Block Size Z", dims > 2 ? o.p.block_z : 1);
  printValue("Num Blocks X", d.num_blocks[0]);
  printValue("Num Blocks Y", d.num_blocks[1]);
  printValue("Num Blocks Z", d.num_blocks[2]);
  printValue("Elements/Thread", o.p.elems_per_thread);
  printValue("Shared Size", d.shared_size);
  if (dims == 3) {
    printValue("Total FP", (int64_t)d.total_fp_per_block);
    printValue("Useful FP", (int64_t)d.useful_fp_per_block);
  } else {
    printValue("Total FP/Block", (int64_t)d.total_fp_per_block);
    printValue("Useful FP/Block", (int64_t)d.useful_fp_per_block);
    if (sid == OT_P2D || sid == OT_TV2D) {
      printValue("Total FP", (int64_t)d.total_fp_steps);
      printValue("Useful FP", (int64_t)d.useful_fp_steps);
    }
  }
  printValue("Useful Ratio", d.useful_ratio);
  printValue("Global Loads/Block", d.global_loads_per_block);
  printValue("Global Stores/Block", d.global_stores_per_block);
  printValue("Shared Loads/Block", d.shared_loads_per_block);
  printValue("Shared Stores/Block", d.shared_stores_per_block);
  printValue("Real Per Block X", d.real_per_block[0]);
  printValue("Real Per Block Y", d.real_per_block[1]);
  printValue("Padded Size", d.padded_size[0]);
  printValue("Padding", d.padding);
  printValue("phase_limit", o.p.phase_limit);
  printValue("Time Steps", o.p.time_steps);
  printValue("Register Usage", d.registers);
  printValue("LDS Bytes/Block", d.lds_bytes);
  printValue("Threads/Block", d.threads_per_block);

  // host fields: srand(seed) uniform [0,1) interior (the reference init), for
  // read-only fields too (F of TV/Rician; V of TV shifted to [0.5,1.5) so it is
  // never 0), FDTD: Ex = Ey = 0 and the uniform field in Hz
  const size_t elems = (size_t)d.field_elems;
  std::vector<std::vector<float>> rw(d.num_rw, std::vector<float>(elems, 0.0f)), ro(d.num_ro, std::vector<float>(elems));
  ot_fill_reference_uniform(rw[d.num_rw - 1].data(), &d, (uint32_t)o.seed);
  if (sid == OT_RICIAN2D) ro[0] = rw[0];
  if (sid == OT_TV2D) {
    ot_fill_reference_uniform(ro[0].data(), &d, (uint32_t)o.seed + 1);
    for (auto& v : ro[0]) v += 0.5f;
    ot_fill_reference_uniform(ro[1].data(), &d, (uint32_t)o.seed + 2);
  }
  if (sid == OT_P3D7) {
    ot_fill_reference_uniform(ro[0].data(), &d, (uint32_t)o.seed + 1);
    for (auto& v : ro[0]) v = 2.0f * v - 1.0f;
  }

  std::vector<void*> a(d.num_rw), b(d.num_rw), r(d.num_ro);
  const size_t bytes = elems * sizeof(float);
  for (int k = 0; k < d.num_rw; ++k) {
    if (ot_dev_alloc(bytes, &a[k]) != OT_OK || ot_dev_alloc(bytes, &b[k]) != OT_OK) return fail("allocation");
    ot_memcpy_h2d(a[k], rw[k].data(), bytes);
    ot_memcpy_h2d(b[k], rw[k].data(), bytes);
  }
  for (int k = 0; k < d.num_ro; ++k) {
    if (ot_dev_alloc(bytes, &r[k]) != OT_OK) return fail("allocation");
    ot_memcpy_h2d(r[k], ro[k].data(), bytes);
  }
  ot_timing tm;
  const int mode = o.graph ? OT_RUN_GRAPH : 0;
  if (ot_run(plan, (float* const*)a.data(), (float* const*)b.data(), (const float* const*)r.data(), nullptr, mode,
             &tm) != OT_OK)
    return fail("run");
  printValue("EventElapsed", tm.event_s);
  printValue("ProfileTimerResolution", 1);
  printValue("Elapsed Time", tm.wall_s);
  const double pts = (double)d.real_size[0] * d.real_size[1] * d.real_size[2];
  printValue("Actual GFlop/s", pts * info.flops_per_point * o.p.time_steps / tm.wall_s / 1e9);
  printValue("Real Size", d.real_size[0]);
  const double dev_pts = (double)d.threads_per_block * d.grid_blocks[0] * d.grid_blocks[1] * d.grid_blocks[2] *
                         (dims == 3 ? 1 : o.p.elems_per_thread);
  printValue("Device GFlop/s", dev_pts * info.flops_per_point * o.p.time_steps / tm.wall_s / 1e9);
  printValue("GStencil/s", pts * o.p.time_steps / tm.event_s / 1e9);

  if (o.clocks) {
    // ClockData rows of one compute unit (the reference keeps SM 1,
    // jacobi-2d.cpp:490-507, :957-997): here CU 1 of shader engine 0 on die 0;
    // clock0..clock4 are s_memtime at P0..P4
    int32_t n = 0;
    if (ot_plan_clocks(plan, nullptr, 0, &n) != OT_OK) return fail("clocks");
    std::vector<ot_clock_record> rec(n);
    if (n > 0 && ot_plan_clocks(plan, rec.data(), n, &n) != OT_OK) return fail("clocks");
    std::cerr << "bx,by,warp,sm,clock0,clock1,clock2,clock3,clock4,\n";
    for (const ot_clock_record& r : rec) {
      const unsigned cu = (r.hw_id >> 8) & 0xF, sh = (r.hw_id >> 12) & 1, se = (r.hw_id >> 13) & 0x7;
      if (!r.valid || cu != 1 || sh != 0 || se != 0 || r.xcc_id != 0) continue;
      std::cerr << r.block_x << "," << r.block_y << "," << r.wave << "," << cu << "," << r.clock[0] << ","
                << r.clock[1] << "," << r.clock[2] << "," << r.clock[3] << "," << r.clock[4] << ",\n";
    }
  }

  int rc = 0;
  if (o.verify && (o.p.variant & (OT_VARIANT_TIME_ONLY_GMEM | OT_VARIANT_TIME_ONLY_INSTR |
                                  OT_VARIANT_TIME_ONLY_SMEM))) {
    printValue("Correctness", "SKIPPED (--time-only results are not the stencil's)");
  } else if (o.verify && o.p.phase_limit == 0) {
    std::vector<std::vector<float>> dev(d.num_rw, std::vector<float>(elems)), ref(d.num_rw, std::vector<float>(elems));
    for (int k = 0; k < d.num_rw; ++k) ot_memcpy_d2h(dev[k].data(), tm.final_buffer ? b[k] : a[k], bytes);
    std::vector<void*> na(d.num_rw), nb(d.num_rw);
    for (int k = 0; k < d.num_rw; ++k) {
      ot_dev_alloc(bytes, &na[k]);
      ot_dev_alloc(bytes, &nb[k]);
      ot_memcpy_h2d(na[k], rw[k].data(), bytes);
      ot_memcpy_h2d(nb[k], rw[k].data(), bytes);
    }
    ot_timing nt;
    if (ot_naive_run(&o.p, (float* const*)na.data(), (float* const*)nb.data(), (const float* const*)r.data(), nullptr,
                     &nt) != OT_OK)
      return fail("verify");
    for (int k = 0; k < d.num_rw; ++k) ot_memcpy_d2h(ref[k].data(), nt.final_buffer ? nb[k] : na[k], bytes);
    // compareResults over the interior (jacobi-2d.cpp:517-551)
    double err = 0.0, refn = 0.0, maxrel = 0.0;
    for (int k = 0; k < d.num_rw; ++k)
      for (int z = 0; z < d.real_size[2]; ++z)
        for (int y = 0; y < d.real_size[1]; ++y)
          for (int x = 0; x < d.real_size[0]; ++x) {
            const int64_t i = d.origin + z * d.stride_z + y * d.stride_y + x;
            const float h = ref[k][i], dv = dev[k][i];
            const double diff = (double)h - (double)dv;
            err += diff * diff;
            refn += (double)h * h;
            if (std::fabs(h) > 0.0) maxrel = std::max(maxrel, std::fabs(diff) / std::fabs((double)h));
          }
    err = std::sqrt(err);
    refn = std::sqrt(refn);
    printValue("Error Norm", err);
    printValue("Ref Norm", refn);
    printValue("Max Relative Error", maxrel);
    const bool ok = std::fabs(refn) >= 1e-7 && err / refn <= 1e-2;
    printValue("Correctness", ok ? "PASSED" : "FAILED");
    for (int k = 0; k < d.num_rw; ++k) {
      ot_dev_free(na[k]);
      ot_dev_free(nb[k]);
    }
  }

  // model feature keys as each reference driver prints them (jacobi-2d.cpp:1029-1037,
  // jacobi-3d.cpp:919-926, fdtd-2d.cpp, rician-denoise-2d.cpp, cdsc-tv-update-2d.cpp:847-853)
  static const double table[14][7] = {
      {3, 0, 3, 3, 1, 1, 1},   {5, 0, 5, 5, 1, 1, 1},    {7, 0, 7, 7, 1, 1, 1},   {5, 0, 5, 5, 1, 1, 1},
      {7, 0, 7, 5, 1, 1, 1},   {9, 0, 9, 9, 1, 1, 1},    {27, 0, 27, 27, 1, 1, 1}, {8, 0, 8, 6, 1, 1, 2},
      {7, 4, 11, 5, 3, 3, 3},  {5, 0, 12, 5, 1, 1, 1},   {7, 0, 20, 7, 1, 1, 1},  {13, 0, 42, 13, 1, 1, 1},
      {7, 0, 59, 7, 1, 1, 1},  {13, 0, 65, 13, 1, 1, 1}};
  const char* names[7] = {"phase2_global_loads", "phase2_shared_loads", "compute_per_point", "phase3_shared_loads",
                          "phase4_global_stores", "shared_stores", "num_fields"};
  for (int i = 0; i < 7; ++i) printValue(names[i], table[sid][i]);
  printValue("data_size", 4.0);
  printValue("Dimensions", dims);

  for (int k = 0; k < d.num_rw; ++k) {
    ot_dev_free(a[k]);
    ot_dev_free(b[k]);
  }
  for (int k = 0; k < d.num_ro; ++k) ot_dev_free(r[k]);
  ot_plan_destroy(plan);
  return rc;
}
