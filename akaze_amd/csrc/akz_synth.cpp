// Seeded synthetic test images and homography warps (host code).
//
// Image model, per 640x480 of area (counts scale with area):
//   * a constant background drawn from U[80, 170];
//   * 25 filled shapes, each an ellipse or a rotated rectangle with equal
//     probability, centre uniform over the image, half-extents U[8, 60] px,
//     rotation U[0, pi), intensity U[30, 220];
//   * 40 Gaussian blobs added on top, sigma U[3, 30] px, amplitude U[-60, 60];
//   * 1/f band-limited noise of standard deviation 8: the sum of seven
//     value-noise octaves (lattice spacings 2, 4, ..., 128 px, smoothstep
//     interpolation) with equal amplitude per octave, i.e. a 1/f amplitude
//     spectrum between 1/128 and 1/2 cycles per pixel, rescaled to std 8;
//   * clipped to [0, 255] and rounded to uint8.
// The generator uses its own PCG32 stream, so an image is a pure function of
// (seed, width, height).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "akz_internal.h"

namespace {

struct Pcg32 {
  uint64_t state, inc;
  explicit Pcg32(uint64_t seed, uint64_t stream = 0x5851f42d4c957f2dULL) {
    // SplitMix64 whitening of the user seed
    uint64_t z = seed + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    state = 0;
    inc = (stream << 1) | 1u;
    next();
    state += z;
    next();
  }
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((32 - rot) & 31));
  }
  double uniform() { return (next() >> 8) * (1.0 / 16777216.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

void add_value_noise(std::vector<double>& acc, int w, int h, int spacing, Pcg32& rng) {
  int gw = w / spacing + 2, gh = h / spacing + 2;
  std::vector<double> lat((size_t)gw * gh);
  for (auto& v : lat) v = rng.normal();
  for (int y = 0; y < h; y++) {
    int gy = y / spacing;
    double ty = smooth((y - gy * spacing) / (double)spacing);
    for (int x = 0; x < w; x++) {
      int gx = x / spacing;
      double tx = smooth((x - gx * spacing) / (double)spacing);
      double a = lat[(size_t)gy * gw + gx], b = lat[(size_t)gy * gw + gx + 1];
      double c = lat[(size_t)(gy + 1) * gw + gx], d = lat[(size_t)(gy + 1) * gw + gx + 1];
      acc[(size_t)y * w + x] += (a + (b - a) * tx) * (1 - ty) + (c + (d - c) * tx) * ty;
    }
  }
}

}  // namespace

void akz_synth_render(uint64_t seed, int w, int h, uint8_t* out) {
  Pcg32 rng(seed);
  double area = (double)w * h / (640.0 * 480.0);
  int nshapes = (int)std::lround(25 * area);
  int nblobs = (int)std::lround(40 * area);
  std::vector<double> img((size_t)w * h, rng.uniform(80.0, 170.0));
  for (int s = 0; s < nshapes; s++) {
    bool ellipse = rng.uniform() < 0.5;
    double cx = rng.uniform(0, w), cy = rng.uniform(0, h);
    double a = rng.uniform(8, 60), b = rng.uniform(8, 60);
    double th = rng.uniform(0, 3.141592653589793);
    double val = rng.uniform(30, 220);
    double ct = std::cos(th), st = std::sin(th);
    double rad = std::sqrt(a * a + b * b) + 1;
    int x0 = std::max(0, (int)(cx - rad)), x1 = std::min(w - 1, (int)(cx + rad));
    int y0 = std::max(0, (int)(cy - rad)), y1 = std::min(h - 1, (int)(cy + rad));
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++) {
        double dx = x + 0.5 - cx, dy = y + 0.5 - cy;
        double u = dx * ct + dy * st, v = -dx * st + dy * ct;
        bool in = ellipse ? (u * u / (a * a) + v * v / (b * b) <= 1.0) : (std::fabs(u) <= a && std::fabs(v) <= b);
        if (in) img[(size_t)y * w + x] = val;
      }
  }
  for (int s = 0; s < nblobs; s++) {
    double cx = rng.uniform(0, w), cy = rng.uniform(0, h);
    double sg = rng.uniform(3, 30), amp = rng.uniform(-60, 60);
    double rad = 4 * sg;
    int x0 = std::max(0, (int)(cx - rad)), x1 = std::min(w - 1, (int)(cx + rad));
    int y0 = std::max(0, (int)(cy - rad)), y1 = std::min(h - 1, (int)(cy + rad));
    double inv = 1.0 / (2 * sg * sg);
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++) {
        double dx = x + 0.5 - cx, dy = y + 0.5 - cy;
        img[(size_t)y * w + x] += amp * std::exp(-(dx * dx + dy * dy) * inv);
      }
  }
  std::vector<double> noise((size_t)w * h, 0.0);
  for (int sp = 2; sp <= 128; sp *= 2) add_value_noise(noise, w, h, sp, rng);
  double mean = 0, var = 0;
  for (double v : noise) mean += v;
  mean /= (double)noise.size();
  for (double v : noise) var += (v - mean) * (v - mean);
  double scale = 8.0 / std::sqrt(var / (double)noise.size() + 1e-30);
  for (size_t i = 0; i < img.size(); i++) {
    double v = img[i] + (noise[i] - mean) * scale;
    v = v < 0 ? 0 : (v > 255 ? 255 : v);
    out[i] = (uint8_t)std::lround(v);
  }
}

void akz_synth_homography_impl(uint64_t seed, int w, int h, double max_rot_deg, double min_scale, double max_scale,
                          double max_persp, double* H) {
  Pcg32 rng(seed, 0x2545f4914f6cdd1dULL);
  double th = rng.uniform(-max_rot_deg, max_rot_deg) * 3.141592653589793 / 180.0;
  double s = rng.uniform(min_scale, max_scale);
  double p1 = rng.uniform(-max_persp, max_persp), p2 = rng.uniform(-max_persp, max_persp);
  double cx = 0.5 * w, cy = 0.5 * h;
  // H = T(c) * [[s R, 0], [p1 p2 1]] * T(-c)
  double M[9] = {s * std::cos(th), -s * std::sin(th), 0, s * std::sin(th), s * std::cos(th), 0, p1, p2, 1};
  double T1[9] = {1, 0, cx, 0, 1, cy, 0, 0, 1}, T0[9] = {1, 0, -cx, 0, 1, -cy, 0, 0, 1};
  double A[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double acc = 0;
      for (int k = 0; k < 3; k++) acc += M[i * 3 + k] * T0[k * 3 + j];
      A[i * 3 + j] = acc;
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double acc = 0;
      for (int k = 0; k < 3; k++) acc += T1[i * 3 + k] * A[k * 3 + j];
      H[i * 3 + j] = acc;
    }
}

static bool invert3(const double* m, double* r) {
  double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], hh = m[7], i = m[8];
  double A = e * i - f * hh, B = -(d * i - f * g), C = d * hh - e * g;
  double det = a * A + b * B + c * C;
  if (std::fabs(det) < 1e-300) return false;
  double id = 1.0 / det;
  r[0] = A * id; r[1] = -(b * i - c * hh) * id; r[2] = (b * f - c * e) * id;
  r[3] = B * id; r[4] = (a * i - c * g) * id; r[5] = -(a * f - c * d) * id;
  r[6] = C * id; r[7] = -(a * hh - b * g) * id; r[8] = (a * e - b * d) * id;
  return true;
}

int akz_synth_warp_u8(const uint8_t* src, int w, int h, const double* H, double noise_std, uint64_t noise_seed,
                      uint8_t* out) {
  double Hi[9];
  if (!invert3(H, Hi)) return -1;
  Pcg32 rng(noise_seed, 0x14057b7ef767814fULL);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      double u = x + 0.5, v = y + 0.5;
      double zz = Hi[6] * u + Hi[7] * v + Hi[8];
      double sx = (Hi[0] * u + Hi[1] * v + Hi[2]) / zz - 0.5;
      double sy = (Hi[3] * u + Hi[4] * v + Hi[5]) / zz - 0.5;
      double val = 0.0;
      // bilinear inside [0, w-1] x [0, h-1]; constant 0 outside
      if (sx >= 0.0 && sy >= 0.0 && sx <= w - 1 && sy <= h - 1) {
        int ix = (int)std::floor(sx), iy = (int)std::floor(sy);
        int ix1 = std::min(ix + 1, w - 1), iy1 = std::min(iy + 1, h - 1);
        double fx = sx - ix, fy = sy - iy;
        const uint8_t* r0 = src + (size_t)iy * w;
        const uint8_t* r1 = src + (size_t)iy1 * w;
        val = (r0[ix] * (1 - fx) + r0[ix1] * fx) * (1 - fy) + (r1[ix] * (1 - fx) + r1[ix1] * fx) * fy;
      }
      if (noise_std > 0) val += noise_std * rng.normal();
      val = val < 0 ? 0 : (val > 255 ? 255 : val);
      out[(size_t)y * w + x] = (uint8_t)std::lround(val);
    }
  return 0;
}
