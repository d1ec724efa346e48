/*
 * ot_oracle_body.h — stencil sweeps of the CPU oracle, instantiated twice by
 * ot_oracle.c (REAL = float and REAL = double). TEST INFRASTRUCTURE ONLY.
 *
 * Every function takes the oracle's padded box (see ot_oracle.c), runs `steps`
 * full time steps in place on the rw arrays and leaves the ring untouched.
 * Constants are written as float literals so the double variant widens the
 * very same values.
 */

#define SQ(v) ((v) * (v))

/* one Jacobi-style sweep: out = f(a) over the interior, then a = out entirely
 * (refA/refB + memcpy, jacobi-2d.cpp:821-840) */

/* jacobi-1d.cpp:631-639, jacobi-1d-5pt.cpp:505-515, jacobi-1d-7pt.cpp:513-523 */
static void SFX(sweep_1d)(int which, const box_t* b, const REAL* a, REAL* o) {
  for (int x = 0; x < b->nx; ++x) {
    const long i = IX(b, x, 0, 0);
    REAL v;
    if (which == OR_J1D3)
      v = (REAL)0.333f * (a[i - 1] + a[i] + a[i + 1]);
    else if (which == OR_J1D5)
      v = (REAL)0.2f * (a[i - 2] + a[i - 1] + a[i] + a[i + 1] + a[i + 2]);
    else
      v = (REAL)0.2f * (a[i - 3] + a[i - 2] + a[i - 1] + a[i] + a[i + 1] + a[i + 2] + a[i + 3]);
    o[i] = v;
  }
}

/* jacobi-2d.cpp:824-828 ((1.0f/5.0f) == 0.2f), poisson-2d.cpp:644-648,
 * gradient-2d.cpp:691-699 (float overload of sqrt) */
static void SFX(sweep_2d)(int which, const box_t* b, const REAL* a, REAL* o) {
  const long sx = b->sx;
  for (int y = 0; y < b->ny; ++y)
    for (int x = 0; x < b->nx; ++x) {
      const long i = IX(b, x, y, 0);
      REAL v;
      if (which == OR_J2D) {
        v = (REAL)(1.0f / 5.0f) * (a[i - 1] + a[i] + a[i + 1] + a[i - sx] + a[i + sx]);
      } else if (which == OR_P2D) {
        v = (REAL)8.0f * a[i] - (a[i - sx - 1] + a[i - sx] + a[i - sx + 1] + a[i - 1] + a[i + 1] +
                                 a[i + sx - 1] + a[i + sx] + a[i + sx + 1]);
      } else { /* G2D */
        const REAL l = a[i - 1], c = a[i], r = a[i + 1], t = a[i - sx], bt = a[i + sx];
        v = c + (REAL)1.0f / RS_ARG(((REAL)0.0001f + SQ(c - l) + SQ(c - r) + SQ(c - t) + SQ(c - bt)));
      }
      o[i] = v;
    }
}

/* jacobi-3d.cpp:729-746: order (y-1, c, y+1, x-1, x+1, z+1, z-1);
 * gradient-3d.cpp:283-307 (GPU text; its CPU loop is a copy of Jacobi-3D):
 *   c + rsqrt(1e-4 + (c-x-)^2 + (c-x+)^2 + (c-y-)^2 + (c-y+)^2 + (c-z-)^2 + (c-z+)^2);
 * poisson-3d.cpp:282-293 (GPU text; its CPU loop is a copy of Jacobi-3D):
 *   26c - sum of 26 neighbours in (dx, dy, dz) lexicographic order;
 * P3D7: the 7-point Poisson-Jacobi relaxation with right-hand side f,
 *   u' = (x- + x+ + y- + y+ + z- + z+ - f) / 6 (no reference counterpart);
 * cs-3d.cpp:286-308 with the constants of :161-163 (GPU text; its CPU loop
 *   :700-716 is a copy of Jacobi-3D). Taps as the on-chip steps address them
 *   (:325-379): val_0_m1_0 = y-1, val_0_p1_0 = y+1 (the reference's first step,
 *   :258-263, swaps those two). eps is the float 1e-9f, pow(d, 2) is d*d. */
static void SFX(sweep_3d)(int which, const box_t* b, const REAL* a, const REAL* f, REAL* o) {
  const long sx = b->sx, sxy = b->sx * b->sy;
  for (int z = 0; z < b->nz; ++z)
    for (int y = 0; y < b->ny; ++y)
      for (int x = 0; x < b->nx; ++x) {
        const long i = IX(b, x, y, z);
        REAL v;
        if (which == OR_J3D) {
          v = (REAL)0.143f * (a[i - sx] + a[i] + a[i + sx] + a[i - 1] + a[i + 1] + a[i + sxy] + a[i - sxy]);
        } else if (which == OR_G3D) {
          const REAL c = a[i];
          v = c + (REAL)1.0f / RS_ARG(((REAL)0.0001f + SQ(c - a[i - 1]) + SQ(c - a[i + 1]) + SQ(c - a[i - sx]) +
                                       SQ(c - a[i + sx]) + SQ(c - a[i - sxy]) + SQ(c - a[i + sxy])));
        } else if (which == OR_CS3D) {
          const REAL u0 = a[i], zm = a[i - sxy], zp = a[i + sxy], ym = a[i - sx], yp = a[i + sx];
          const REAL xm = a[i - 1], xp = a[i + 1];
          const REAL c1 = u0 * ((REAL)1.0f / RS_ARG((REAL)1e-9f + SQ(zp - u0) + SQ(yp - u0) + SQ(xp - u0))) /
                          (REAL)1.1f;
          const REAL c2 = u0 *
                          ((REAL)1.0f / RS_ARG((REAL)1e-9f + SQ(u0 - zm) + SQ(a[i - sxy + sx] - zm) +
                                               SQ(a[i - sxy + 1] - zm))) /
                          (REAL)1.1f;
          const REAL c3 = u0 *
                          ((REAL)1.0f / RS_ARG((REAL)1e-9f + SQ(a[i + sxy - sx] - ym) + SQ(u0 - ym) +
                                               SQ(a[i - sx + 1] - ym))) /
                          (REAL)1.1f;
          const REAL c4 = u0 *
                          ((REAL)1.0f / RS_ARG((REAL)1e-9f + SQ(a[i + sxy - 1] - xm) + SQ(a[i + sx - 1] - xm) +
                                               SQ(u0 - xm))) /
                          (REAL)1.1f;
          v = ((REAL)1.1f * u0 + c1 * (zp + yp + xp) + c2 * zm + c3 * ym + c4 * xm) /
              ((REAL)1.1f + (REAL)3.0f * c1 + c2 + c3 + c4);
        } else if (which == OR_P3D7) {
          v = (a[i - 1] + a[i + 1] + a[i - sx] + a[i + sx] + a[i - sxy] + a[i + sxy] - f[i]) / (REAL)6.0f;
        } else { /* P3D27 */
          REAL s = 0;
          int first = 1;
          for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dz = -1; dz <= 1; ++dz) {
                if (!dx && !dy && !dz) continue;
                const REAL t = a[i + dz * sxy + dy * sx + dx];
                s = first ? t : s + t;
                first = 0;
              }
          v = (REAL)26.0f * a[i] - s;
        }
        o[i] = v;
      }
}

/* cdsc-tv-update-2d.cpp:290-309 (GPU text; its CPU loop is a copy of
 * Jacobi-2D). c3 duplicates c1 as in the reference. */
static void SFX(sweep_tv)(const box_t* b, const REAL* a, const REAL* V, const REAL* F, REAL alpha, REAL* o) {
  const long sx = b->sx;
  for (int y = 0; y < b->ny; ++y)
    for (int x = 0; x < b->nx; ++x) {
      const long i = IX(b, x, y, 0);
      const REAL tc = a[i - sx], tr = a[i - sx + 1], ml = a[i - 1], mc = a[i], mr = a[i + 1];
      const REAL bl = a[i + sx - 1], bc = a[i + sx];
      const REAL v = V[i];
      const REAL c1 = mc * ((REAL)1.0f / RS_ARG((REAL)1e-8f + SQ(bc - mc) + SQ(mr - mc))) / v;
      const REAL c2 = mc * ((REAL)1.0f / RS_ARG((REAL)1e-8f + SQ(mc - tc) + SQ(tr - tc))) / v;
      const REAL c3 = mc * ((REAL)1.0f / RS_ARG((REAL)1e-8f + SQ(bc - mc) + SQ(mr - mc))) / v;
      const REAL c4 = mc * ((REAL)1.0f / RS_ARG((REAL)1e-8f + SQ(bl - ml) + SQ(mc - ml))) / v;
      o[i] = (REAL)1.0f / (alpha + c1 + c2 + c3 + c4) * (alpha * F[i] + c1 * bc + c2 * tc + c3 * mr + c4 * ml);
    }
}

/* rician-denoise-2d.cpp:154-162 (constants), :271-281 (g and r),
 * :292-327 (update); no CPU loop in the reference (:728-740 is Jacobi-2D).
 * Per step: g over the interior (0 outside), then u' from u, g and f. */
static void SFX(step_rician)(const box_t* b, const REAL* a, const REAL* F, REAL* g, REAL sigma2, REAL gamma,
                             REAL dt, REAL* o) {
  const long sx = b->sx;
  for (int y = 0; y < b->ny; ++y)
    for (int x = 0; x < b->nx; ++x) {
      const long i = IX(b, x, y, 0);
      const REAL c = a[i];
      g[i] = (REAL)1.0f / RS_ARG((REAL)1.0e-20f + SQ(c - a[i + 1]) + SQ(c - a[i - 1]) + SQ(c - a[i + sx]) +
                                 SQ(c - a[i - sx]));
    }
  for (int y = 0; y < b->ny; ++y)
    for (int x = 0; x < b->nx; ++x) {
      const long i = IX(b, x, y, 0);
      const REAL c = a[i], f = F[i];
      REAL r = c * f / sigma2;
      r = (r * ((REAL)2.38944f + r * ((REAL)0.950037f + r))) /
          ((REAL)4.65314f + r * ((REAL)2.57541f + r * ((REAL)1.48937f + r)));
      const REAL gl = g[i - 1], gr = g[i + 1], gt = g[i - sx], gb = g[i + sx];
      o[i] = (c + dt * (a[i + sx] * gb + a[i - sx] * gt + a[i + 1] * gr + a[i - 1] * gl + gamma * f * r)) /
             ((REAL)1.0f + dt * (gr + gl + gb + gt + gamma));
    }
}

/* fdtd-2d.cpp:269-330 (phase-2 equations; no CPU loop in the reference).
 * Ey' = Ey - 0.5(Hz - Hz(y-1)); Ex' = Ex - 0.5(Hz - Hz(x-1));
 * Hz' = Hz - 0.7((Ex'(x+1) - Ex') + Ey'(y+1) + Ey')  (physical: - Ey'). */
static void SFX(step_fdtd)(const box_t* b, REAL* ex, REAL* ey, REAL* hz, int physical) {
  const long sx = b->sx;
  for (int y = 0; y < b->ny; ++y)
    for (int x = 0; x < b->nx; ++x) {
      const long i = IX(b, x, y, 0);
      const REAL h = hz[i];
      ey[i] = ey[i] - (REAL)0.5f * (h - hz[i - sx]);
      ex[i] = ex[i] - (REAL)0.5f * (h - hz[i - 1]);
    }
  for (int y = 0; y < b->ny; ++y)
    for (int x = 0; x < b->nx; ++x) {
      const long i = IX(b, x, y, 0);
      if (physical)
        hz[i] = hz[i] - (REAL)0.7f * (ex[i + 1] - ex[i] + ey[i + sx] - ey[i]);
      else
        hz[i] = hz[i] - (REAL)0.7f * (ex[i + 1] - ex[i] + ey[i + sx] + ey[i]);
    }
}

/* Run `steps` time steps of `stencil` in place.
 *   rw: num_rw arrays (FDTD: ex, ey, hz), ro: read-only arrays (TV: V, F;
 *   Rician: f; P3D7: f), coef: TV alpha / Rician (sigma^2, lambda, dt), 0 = default. */
int SFX(oracle_run)(int stencil, int nx, int ny, int nz, int P, int steps, REAL** rw, const REAL** ro,
                    const float* coef, int variant) {
  int dims = (stencil <= OR_J1D7) ? 1 : ((stencil == OR_J3D || stencil == OR_P3D27 || stencil == OR_P3D7 ||
                                          stencil == OR_G3D || stencil == OR_CS3D) ? 3 : 2);
  box_t b = mkbox(dims, nx, ny, nz, P);
  const long n = nel(&b);
  if (stencil == OR_FDTD2D) {
    /* Ex/Ey/Hz updates only read values of the current or already-updated
     * field at other points, exactly as the two sub-stages are ordered. */
    for (int t = 0; t < steps; ++t) SFX(step_fdtd)(&b, rw[0], rw[1], rw[2], variant & 1);
    return 0;
  }
  REAL* o = (REAL*)malloc(sizeof(REAL) * n);
  REAL* g = NULL;
  if (!o) return -1;
  memcpy(o, rw[0], sizeof(REAL) * n);
  if (stencil == OR_RICIAN2D) {
    g = (REAL*)calloc(n, sizeof(REAL));
    if (!g) {
      free(o);
      return -1;
    }
  }
  /* Rician: coef[0] is SIGMA2 itself (the literal 0.0025f of :159 by default) */
  const REAL sigma2 = (coef && coef[0] != 0.0f) ? (REAL)coef[0] : (REAL)0.0025f;
  const REAL lambda = (coef && coef[1] != 0.0f) ? (REAL)coef[1] : (REAL)0.065f;
  const REAL gamma = (REAL)((float)lambda / (float)sigma2);
  const REAL dt = (coef && coef[2] != 0.0f) ? (REAL)coef[2] : (REAL)5.0f;
  const REAL alpha = (coef && coef[0] != 0.0f) ? (REAL)coef[0] : (REAL)1.1f;
  for (int t = 0; t < steps; ++t) {
    if (dims == 1) SFX(sweep_1d)(stencil, &b, rw[0], o);
    else if (stencil == OR_TV2D) SFX(sweep_tv)(&b, rw[0], ro[0], ro[1], alpha, o);
    else if (stencil == OR_RICIAN2D) SFX(step_rician)(&b, rw[0], ro[0], g, sigma2, gamma, dt, o);
    else if (dims == 2) SFX(sweep_2d)(stencil, &b, rw[0], o);
    else SFX(sweep_3d)(stencil, &b, rw[0], ro ? ro[0] : NULL, o);
    memcpy(rw[0], o, sizeof(REAL) * n);
  }
  free(o);
  free(g);
  return 0;
}

#undef SQ
