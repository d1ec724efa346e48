// fill.cpp — seeded synthetic inputs in the padded layout of an ot_derived.
//
// ot_fill_reference_uniform restates the reference drivers' host init
// (opencl/poisson-2d/poisson-2d.cpp:611-621; the same block is compiled out
// with #if 0 in jacobi-2d.cpp:790-802 and jacobi-3d.cpp:688-706): srand(seed),
// then for every element in row-major padded order 0 in the padding ring and
// (float)rand()/((float)RAND_MAX+1.0f) in the interior. It calls the C
// library's rand(), i.e. glibc's generator, as the reference does.
// The other generators are counter-based (splitmix64 of seed and the interior
// linear index), so any sub-block of a large grid can be produced independently.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "internal.hpp"

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// uniform in [0, 1) with 24 random bits (exactly representable in float)
inline float u01(uint64_t seed, uint64_t idx) {
  const uint64_t h = splitmix64(seed ^ splitmix64(idx));
  return (float)(h >> 40) * (1.0f / 16777216.0f);
}

template <class F>
void for_each_padded(const ot_derived* d, F&& f) {
  const int64_t nz = d->padded_size[2], ny = d->padded_size[1], nx = d->stride_y;
  const int64_t oz = d->dims > 2 ? d->origin / d->stride_z : 0;
  const int64_t oy = d->dims > 1 ? (d->origin - oz * d->stride_z) / d->stride_y : 0;
  const int64_t ox = d->origin - oz * d->stride_z - oy * d->stride_y;
  for (int64_t z = 0; z < nz; ++z)
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) {
        const int64_t ix = x - ox, iy = y - oy, iz = z - oz;
        const bool interior = ix >= 0 && ix < d->real_size[0] && iy >= 0 && iy < d->real_size[1] &&
                              iz >= 0 && iz < d->real_size[2];
        f(z * d->stride_z + y * d->stride_y + x, interior, ix, iy, iz);
      }
}

ot_status check(const float* host, const ot_derived* d) {
  if (!host || !d) return ot::set_error(OT_EINVAL, "null argument");
  if (d->field_elems <= 0) return ot::set_error(OT_EINVAL, "empty layout");
  return OT_OK;
}

}  // namespace

extern "C" {

ot_status ot_fill_reference_uniform(float* host, const ot_derived* d, uint32_t seed) {
  ot_status st = check(host, d);
  if (st != OT_OK) return st;
  std::srand(seed);
  for_each_padded(d, [&](int64_t i, bool in, int64_t, int64_t, int64_t) {
    host[i] = in ? (float)std::rand() / ((float)RAND_MAX + 1.0f) : 0.0f;
  });
  return OT_OK;
}

ot_status ot_fill_uniform(float* host, const ot_derived* d, uint64_t seed, float lo, float hi) {
  ot_status st = check(host, d);
  if (st != OT_OK) return st;
  const int64_t nx = d->real_size[0], ny = d->real_size[1];
  for_each_padded(d, [&](int64_t i, bool in, int64_t x, int64_t y, int64_t z) {
    host[i] = in ? lo + (hi - lo) * u01(seed, (uint64_t)((z * ny + y) * nx + x)) : 0.0f;
  });
  return OT_OK;
}

ot_status ot_fill_sinusoid_noise(float* host, const ot_derived* d, uint64_t seed, float amp) {
  ot_status st = check(host, d);
  if (st != OT_OK) return st;
  const int64_t nx = d->real_size[0], ny = d->real_size[1], nz = d->real_size[2];
  const double tp = 6.283185307179586;
  for_each_padded(d, [&](int64_t i, bool in, int64_t x, int64_t y, int64_t z) {
    if (!in) {
      host[i] = 0.0f;
      return;
    }
    double v = std::sin(tp * (double)x / (double)nx);
    if (d->dims > 1) v *= std::sin(tp * (double)y / (double)ny);
    if (d->dims > 2) v *= std::sin(tp * (double)z / (double)nz);
    const double noise = (double)amp * (2.0 * (double)u01(seed, (uint64_t)((z * ny + y) * nx + x)) - 1.0);
    host[i] = (float)(v + noise);
  });
  return OT_OK;
}

ot_status ot_fill_rician_phantom(float* arena, int32_t n_images, const ot_image* images, uint64_t seed,
                                 float sigma) {
  if (!arena || !images || n_images < 1) return ot::set_error(OT_EINVAL, "bad arguments");
  for (int32_t m = 0; m < n_images; ++m) {
    const ot_image& im = images[m];
    const uint64_t s = splitmix64(seed ^ (0x51ed2701ull * (uint64_t)(m + 1)));
    uint64_t ctr = 0;
    auto uni = [&]() { return u01(s, ctr++); };
    const int w = im.width, h = im.height;
    float* img = arena + im.offset;
    // background plus 8..32 rectangles, each with one intensity in [0, 1)
    const float bg = uni();
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < im.pitch; ++x) img[(int64_t)y * im.pitch + x] = (x < w) ? bg : 0.0f;
    const int nrect = 8 + (int)(uni() * 25.0f);
    for (int r = 0; r < nrect; ++r) {
      int xa = (int)(uni() * w), xb = (int)(uni() * w), ya = (int)(uni() * h), yb = (int)(uni() * h);
      if (xa > xb) std::swap(xa, xb);
      if (ya > yb) std::swap(ya, yb);
      const float v = uni();
      for (int y = ya; y <= yb && y < h; ++y)
        for (int x = xa; x <= xb && x < w; ++x) img[(int64_t)y * im.pitch + x] = v;
    }
    // Rician corruption: |p + n1 + i n2|, n1, n2 ~ N(0, sigma) (Box-Muller)
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const double u1 = std::max(1e-12, (double)uni()), u2 = (double)uni();
        const double rad = std::sqrt(-2.0 * std::log(u1));
        const double n1 = sigma * rad * std::cos(6.283185307179586 * u2);
        const double n2 = sigma * rad * std::sin(6.283185307179586 * u2);
        float& p = img[(int64_t)y * im.pitch + x];
        const double re = (double)p + n1;
        p = (float)std::sqrt(re * re + n2 * n2);
      }
  }
  return OT_OK;
}

ot_status ot_fill_gaussian_pulse(float* host, const ot_derived* d, float amplitude, float sigma) {
  ot_status st = check(host, d);
  if (st != OT_OK) return st;
  if (!(sigma > 0.0f)) return ot::set_error(OT_EINVAL, "sigma must be > 0");
  const double cx = d->real_size[0] / 2, cy = d->real_size[1] / 2, cz = d->real_size[2] / 2;
  const double inv = 1.0 / (2.0 * (double)sigma * (double)sigma);
  for_each_padded(d, [&](int64_t i, bool in, int64_t x, int64_t y, int64_t z) {
    if (!in) {
      host[i] = 0.0f;
      return;
    }
    double r2 = (x - cx) * (x - cx);
    if (d->dims > 1) r2 += (y - cy) * (y - cy);
    if (d->dims > 2) r2 += (z - cz) * (z - cz);
    host[i] = (float)((double)amplitude * std::exp(-r2 * inv));
  });
  return OT_OK;
}

}  // extern "C"
