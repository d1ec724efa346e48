/*
 * ot_oracle.c — CPU oracle for the overlapped time-tiled stencil engine.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline); the engine never calls it.
 *
 * It restates, in plain C compiled with -O2 -ffp-contract=off (no FMA, like the
 * reference build config.ninja:2-5), the reference's CPU loops:
 *   - per step, a Jacobi-style sweep over the interior [P, padded-P)^d of a
 *     padded array whose ring holds 0 (Dirichlet 0), written into refB, then a
 *     full-array memcpy(refA, refB) (jacobi-2d.cpp:821-840);
 *   - the exact association order of each reference CPU expression.
 * Stencils without a matching reference CPU loop restate the first-time-step
 * equations of the reference's generated kernel text (cited per function).
 *
 * Arrays use the oracle's own layout: a dense padded box of extent n+2P per
 * used axis (P >= radius), x fastest, ring = 0. Multi-field stencils take
 * separate arrays. Each function exists in float (the parity target) and
 * double (the FP64 companion used to report FP32 rounding error); the double
 * variants use the float constants widened to double so that only the
 * arithmetic precision differs.
 *
 * Parity pins (see tests/test_oracle.py): closed forms (single-delta
 * multinomial weights, constant fields, fixed points), brute-force small cases
 * against an independent numpy evaluation, mirror symmetry, and the
 * gcd/lcm-free structure of the sweep. rsqrt is evaluated as 1.0f/sqrtf(x)
 * (correctly rounded), matching the reference CPU loop of gradient-2d
 * (gradient-2d.cpp:697-699, float overload) — the double-overload variant is
 * provided as oracle_g2d_double_sqrt for the toolchain-dependent case.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_J1D3 0
#define OR_J1D5 1
#define OR_J1D7 2
#define OR_J2D 3
#define OR_J3D 4
#define OR_P2D 5
#define OR_P3D27 6
#define OR_P3D7 7
#define OR_FDTD2D 8
#define OR_G2D 9
#define OR_G3D 10
#define OR_RICIAN2D 11
#define OR_TV2D 12
#define OR_CS3D 13

typedef struct {
  int nx, ny, nz; /* interior extents (1 for unused axes) */
  int P;          /* ring width of every used axis */
  int dims;
  long sx, sy, sz; /* padded extents */
} box_t;

static box_t mkbox(int dims, int nx, int ny, int nz, int P) {
  box_t b;
  b.dims = dims;
  b.nx = nx;
  b.ny = dims > 1 ? ny : 1;
  b.nz = dims > 2 ? nz : 1;
  b.P = P;
  b.sx = nx + 2L * P;
  b.sy = dims > 1 ? ny + 2L * P : 1;
  b.sz = dims > 2 ? nz + 2L * P : 1;
  return b;
}
static long nel(const box_t* b) { return b->sx * b->sy * b->sz; }
/* index of interior point (x, y, z) (may be -P..n+P-1) */
#define IX(b, x, y, z) \
  ((((long)(z) + ((b)->dims > 2 ? (b)->P : 0)) * (b)->sy + ((long)(y) + ((b)->dims > 1 ? (b)->P : 0))) * (b)->sx + \
   ((long)(x) + (b)->P))

/* ------------------------------------------------------------------ float */

#define REAL float
#define RS_ARG(x) sqrtf(x)
#define SFX(name) name##_f32
#include "ot_oracle_body.h"
#undef REAL
#undef RS_ARG
#undef SFX

/* ------------------------------------------------------------------ double */

#define REAL double
#define RS_ARG(x) sqrt(x)
#define SFX(name) name##_f64
#include "ot_oracle_body.h"
#undef REAL
#undef RS_ARG
#undef SFX

/* Gradient-2D with the double-precision sqrt overload: the reference's
 * `1.0f/sqrt(float)` resolves to ::sqrt(double) when no float overload is in
 * scope, giving center + (float)(1.0/sqrt((double)s)) evaluated in double. */
void oracle_g2d_double_sqrt(int nx, int ny, int P, int steps, float* u) {
  box_t b = mkbox(2, nx, ny, 1, P);
  const long n = nel(&b);
  float* a = (float*)malloc(sizeof(float) * n);
  float* o = (float*)malloc(sizeof(float) * n);
  memcpy(a, u, sizeof(float) * n);
  memcpy(o, u, sizeof(float) * n);
  for (int t = 0; t < steps; ++t) {
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const float l = a[IX(&b, x - 1, y, 0)], c = a[IX(&b, x, y, 0)], r = a[IX(&b, x + 1, y, 0)];
        const float tp = a[IX(&b, x, y - 1, 0)], bt = a[IX(&b, x, y + 1, 0)];
        const float s = 0.0001f + (c - l) * (c - l) + (c - r) * (c - r) + (c - tp) * (c - tp) + (c - bt) * (c - bt);
        o[IX(&b, x, y, 0)] = (float)((double)c + 1.0f / sqrt((double)s));
      }
    memcpy(a, o, sizeof(float) * n);
  }
  memcpy(u, a, sizeof(float) * n);
  free(a);
  free(o);
}

/* glibc srand/rand init of the reference drivers, restated independently of
 * the engine's generator: row-major over the padded box of extents
 * (sx, sy, sz), ring of width P = 0, interior (float)rand()/((float)RAND_MAX+1.0f)
 * (poisson-2d.cpp:611-621). */
void oracle_reference_init(int dims, int sx, int sy, int sz, int P, int extra_tail, unsigned seed, float* out) {
  srand(seed);
  long i = 0;
  for (int z = 0; z < sz; ++z)
    for (int y = 0; y < sy; ++y)
      for (int x = 0; x < sx; ++x, ++i) {
        int ring = x < P || x >= sx - P - extra_tail;
        if (dims > 1) ring = ring || y < P || y >= sy - P;
        if (dims > 2) ring = ring || z < P || z >= sz - P;
        out[i] = ring ? 0.0f : (float)rand() / ((float)RAND_MAX + 1.0f);
      }
}
