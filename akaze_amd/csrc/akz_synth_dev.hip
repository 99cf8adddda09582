// Seeded synthetic images rendered on the device, for training and
// evaluation sets of many images.
//
// Same image model as the host generator (akz_synth.cpp): constant
// background U[80,170]; 25 shapes and 40 Gaussian blobs per 640x480 of area
// (ellipses or rotated rectangles, half-extents U[8,60], intensity
// U[30,220]; blobs sigma U[3,30], amplitude U[-60,60]); 1/f noise from
// seven value-noise octaves (lattice spacings 2..128, smoothstep) rescaled
// to standard deviation 8; clipped and rounded to uint8, then /255.  Shape
// and blob parameters come from the host PCG32 stream; the noise lattice
// values are a counter-based hash of (seed, octave, lattice index), so every
// pixel is computed independently.  An image is a pure function of (seed,
// width, height), but it is not the host generator's image for that seed.
#include <cmath>
#include <cstring>
#include <vector>

#include "akz_internal.h"
#include "akz_kernels.h"

namespace {

struct Shape { float cx, cy, a, b, ct, st, val; int ellipse; int x0, x1, y0, y1; };
struct Blob { float cx, cy, inv, amp; int x0, x1, y0, y1; };

struct Pcg {
  uint64_t state, inc;
  explicit Pcg(uint64_t seed) {
    uint64_t z = seed + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    state = 0;
    inc = (0x61c8864680b583ebULL << 1) | 1u;
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
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// standard normal lattice value (Box-Muller on two hashed uniforms)
__device__ __forceinline__ float lattice(uint64_t key) {
  const uint64_t r = mix64(key);
  const float u1 = ((uint32_t)(r >> 40) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = ((uint32_t)(r >> 8) & 0xFFFFFFu) * (1.0f / 16777216.0f);
  return sqrtf(-2.0f * logf(u1)) * cosf(6.2831853f * u2);
}

__device__ __forceinline__ float smoothf(float t) { return t * t * (3.0f - 2.0f * t); }

constexpr int SN_BLOCKS = 256;

// Noise field and per-block partial sums of v and v^2.
__global__ __launch_bounds__(256) void k_synth_noise(const uint64_t* __restrict__ seeds, int w, int h,
                                                     float* __restrict__ noise, size_t nstride,
                                                     double* __restrict__ part) {
  __shared__ double s_a[4], s_b[4];
  const uint64_t seed = seeds[blockIdx.z];
  noise += (size_t)blockIdx.z * nstride;
  double sa = 0.0, sb = 0.0;
  const size_t n = (size_t)w * h;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
    float acc = 0.0f;
    for (int o = 0, sp = 2; o < 7; o++, sp *= 2) {
      const int gw = w / sp + 2;
      const int gx = x / sp, gy = y / sp;
      const float tx = smoothf((float)(x - gx * sp) / sp), ty = smoothf((float)(y - gy * sp) / sp);
      const uint64_t base = (seed * 0x100000001b3ULL) ^ ((uint64_t)(o + 1) << 56);
      const float a = lattice(base + (uint64_t)gy * gw + gx), b = lattice(base + (uint64_t)gy * gw + gx + 1);
      const float c = lattice(base + (uint64_t)(gy + 1) * gw + gx);
      const float d = lattice(base + (uint64_t)(gy + 1) * gw + gx + 1);
      acc += (a + (b - a) * tx) * (1.0f - ty) + (c + (d - c) * tx) * ty;
    }
    noise[i] = acc;
    sa += acc;
    sb += (double)acc * acc;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sa += __shfl_xor(sa, off, 64);
    sb += __shfl_xor(sb, off, 64);
  }
  const int wv = akz_wave();
  if ((threadIdx.x & 63) == 0) { s_a[wv] = sa; s_b[wv] = sb; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* p = part + ((size_t)blockIdx.z * SN_BLOCKS + blockIdx.x) * 2;
    p[0] = s_a[0] + s_a[1] + s_a[2] + s_a[3];
    p[1] = s_b[0] + s_b[1] + s_b[2] + s_b[3];
  }
}

// Composition: background, shapes (the last covering shape wins), blobs,
// normalised noise; rounded to uint8 and written as float / 255.
__global__ __launch_bounds__(256) void k_synth_compose(const float* __restrict__ bg, const Shape* __restrict__ shapes,
                                                       int nshapes, const Blob* __restrict__ blobs, int nblobs, int w,
                                                       int h, const float* __restrict__ noise, size_t nstride,
                                                       const double* __restrict__ part, float* __restrict__ out,
                                                       int pitch, size_t ostride) {
  __shared__ float s_mean, s_scale;
  const int z = blockIdx.z;
  shapes += (size_t)z * nshapes;
  blobs += (size_t)z * nblobs;
  noise += (size_t)z * nstride;
  out = reinterpret_cast<float*>(reinterpret_cast<char*>(out) + (size_t)z * ostride);
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < SN_BLOCKS; i++) {
      sa += part[((size_t)z * SN_BLOCKS + i) * 2];
      sb += part[((size_t)z * SN_BLOCKS + i) * 2 + 1];
    }
    const double n = (double)w * h, mean = sa / n;
    const double var = sb / n - mean * mean;
    s_mean = (float)mean;
    s_scale = (float)(8.0 / sqrt((var > 0 ? var : 0.0) + 1e-30));
  }
  __syncthreads();
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + akz_wave();
  if (x >= w || y >= h) return;
  float v = bg[z];
  const float px = x + 0.5f, py = y + 0.5f;
  for (int s = 0; s < nshapes; s++) {
    const Shape S = shapes[s];
    if (x < S.x0 || x > S.x1 || y < S.y0 || y > S.y1) continue;
    const float dx = px - S.cx, dy = py - S.cy;
    const float u = dx * S.ct + dy * S.st, t = -dx * S.st + dy * S.ct;
    const bool in = S.ellipse ? (u * u / (S.a * S.a) + t * t / (S.b * S.b) <= 1.0f)
                              : (fabsf(u) <= S.a && fabsf(t) <= S.b);
    if (in) v = S.val;
  }
  for (int s = 0; s < nblobs; s++) {
    const Blob B = blobs[s];
    if (x < B.x0 || x > B.x1 || y < B.y0 || y > B.y1) continue;
    const float dx = px - B.cx, dy = py - B.cy;
    v += B.amp * expf(-(dx * dx + dy * dy) * B.inv);
  }
  v += (noise[(size_t)y * w + x] - s_mean) * s_scale;
  v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
  out[(size_t)y * pitch + x] = rintf(v) * (1.0f / 255.0f);
}

// Homography warp of a float image in [0,1]: destination pixel centre u
// maps to H^-1 u in the source, bilinear inside the source, 0 outside; plus
// Gaussian noise (hashed Box-Muller), clipped to [0,1] and quantised to
// 1/255 steps like an 8-bit frame.
struct Hinv { float m[9]; };

__global__ __launch_bounds__(256) void k_warp(const float* __restrict__ src, int w, int h, int spitch, Hinv H,
                                              float noise_std, uint64_t seed, float* __restrict__ dst, int dpitch) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + akz_wave();
  if (x >= w || y >= h) return;
  const float u = x + 0.5f, v = y + 0.5f;
  const float zz = H.m[6] * u + H.m[7] * v + H.m[8];
  const float sx = (H.m[0] * u + H.m[1] * v + H.m[2]) / zz - 0.5f;
  const float sy = (H.m[3] * u + H.m[4] * v + H.m[5]) / zz - 0.5f;
  float val = 0.0f;
  if (sx >= 0.0f && sy >= 0.0f && sx <= w - 1 && sy <= h - 1) {
    const int ix = (int)sx, iy = (int)sy;
    const int ix1 = akz_imin(ix + 1, w - 1), iy1 = akz_imin(iy + 1, h - 1);
    const float fx = sx - ix, fy = sy - iy;
    const float* r0 = src + (size_t)iy * spitch;
    const float* r1 = src + (size_t)iy1 * spitch;
    val = (r0[ix] * (1 - fx) + r0[ix1] * fx) * (1 - fy) + (r1[ix] * (1 - fx) + r1[ix1] * fx) * fy;
  }
  if (noise_std > 0.0f) val += noise_std * lattice(seed ^ ((uint64_t)y << 32 | (uint32_t)x));
  val = fminf(fmaxf(val, 0.0f), 1.0f);
  dst[(size_t)y * dpitch + x] = rintf(val * 255.0f) * (1.0f / 255.0f);
}

}  // namespace

extern "C" akz_status akz_synth_device_warp(const float* d_src, int w, int h, size_t src_pitch_bytes,
                                            const double* H9, float noise_std, uint64_t noise_seed, float* d_dst,
                                            size_t dst_pitch_bytes) {
  if (!d_src || !d_dst || !H9 || w < 1 || h < 1 || src_pitch_bytes < sizeof(float) * w ||
      dst_pitch_bytes < sizeof(float) * w || src_pitch_bytes % 4 || dst_pitch_bytes % 4 || d_src == d_dst)
    return AKZ_ERR_ARG;
  const double* m = H9;
  const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], hh = m[7], i = m[8];
  const double A = e * i - f * hh, B = -(d * i - f * g), C = d * hh - e * g;
  const double det = a * A + b * B + c * C;
  if (std::fabs(det) < 1e-300) return AKZ_ERR_ARG;
  const double id = 1.0 / det;
  const double r[9] = {A * id, -(b * i - c * hh) * id, (b * f - c * e) * id,
                       B * id, (a * i - c * g) * id,   -(a * f - c * d) * id,
                       C * id, -(a * hh - b * g) * id, (a * e - b * d) * id};
  Hinv Hi;
  for (int k = 0; k < 9; k++) Hi.m[k] = (float)r[k];
  hipLaunchKernelGGL(k_warp, dim3((w + 63) / 64, (h + 3) / 4), dim3(256), 0, nullptr, d_src, w, h,
                     (int)(src_pitch_bytes / 4), Hi, noise_std, noise_seed, d_dst, (int)(dst_pitch_bytes / 4));
  AKZ_HIP_TRY(hipGetLastError());
  AKZ_HIP_TRY(hipDeviceSynchronize());
  return AKZ_OK;
}

extern "C" akz_status akz_synth_device_batch(uint64_t seed0, int n, int w, int h, float* d_out, size_t pitch_bytes,
                                             size_t image_stride_bytes) {
  if (n < 1 || w < 16 || h < 16 || !d_out || pitch_bytes < sizeof(float) * w || pitch_bytes % sizeof(float))
    return AKZ_ERR_ARG;
  const size_t ostride = image_stride_bytes ? image_stride_bytes : pitch_bytes * h;
  if (n > 1 && ostride < pitch_bytes * h) return AKZ_ERR_ARG;
  const double area = (double)w * h / (640.0 * 480.0);
  const int nshapes = (int)std::lround(25 * area), nblobs = (int)std::lround(40 * area);
  std::vector<Shape> shapes((size_t)n * std::max(nshapes, 1));
  std::vector<Blob> blobs((size_t)n * std::max(nblobs, 1));
  std::vector<float> bg(n);
  std::vector<uint64_t> seeds(n);
  for (int z = 0; z < n; z++) {
    const uint64_t seed = seed0 + (uint64_t)z;
    seeds[z] = seed;
    Pcg rng(seed);
    bg[z] = (float)rng.uniform(80.0, 170.0);
    for (int s = 0; s < nshapes; s++) {
      Shape& S = shapes[(size_t)z * nshapes + s];
      S.ellipse = rng.uniform() < 0.5;
      S.cx = (float)rng.uniform(0, w);
      S.cy = (float)rng.uniform(0, h);
      S.a = (float)rng.uniform(8, 60);
      S.b = (float)rng.uniform(8, 60);
      const double th = rng.uniform(0, 3.141592653589793);
      S.val = (float)rng.uniform(30, 220);
      S.ct = (float)std::cos(th);
      S.st = (float)std::sin(th);
      const double rad = std::sqrt((double)S.a * S.a + (double)S.b * S.b) + 1;
      S.x0 = std::max(0, (int)(S.cx - rad)); S.x1 = std::min(w - 1, (int)(S.cx + rad));
      S.y0 = std::max(0, (int)(S.cy - rad)); S.y1 = std::min(h - 1, (int)(S.cy + rad));
    }
    for (int s = 0; s < nblobs; s++) {
      Blob& B = blobs[(size_t)z * nblobs + s];
      B.cx = (float)rng.uniform(0, w);
      B.cy = (float)rng.uniform(0, h);
      const double sg = rng.uniform(3, 30);
      B.amp = (float)rng.uniform(-60, 60);
      const double rad = 4 * sg;
      B.inv = (float)(1.0 / (2 * sg * sg));
      B.x0 = std::max(0, (int)(B.cx - rad)); B.x1 = std::min(w - 1, (int)(B.cx + rad));
      B.y0 = std::max(0, (int)(B.cy - rad)); B.y1 = std::min(h - 1, (int)(B.cy + rad));
    }
  }
  const size_t nstride = (size_t)w * h;
  char* scratch = nullptr;
  const size_t b_noise = sizeof(float) * nstride * n, b_part = sizeof(double) * 2 * SN_BLOCKS * n;
  const size_t b_sh = sizeof(Shape) * shapes.size(), b_bl = sizeof(Blob) * blobs.size();
  const size_t b_bg = sizeof(float) * n, b_seed = sizeof(uint64_t) * n;
  size_t off[7] = {0};
  const size_t sizes[6] = {b_noise, b_part, b_sh, b_bl, b_bg, b_seed};
  for (int i = 0; i < 6; i++) off[i + 1] = off[i] + ((sizes[i] + 255) / 256) * 256;
  AKZ_HIP_TRY(hipMalloc(&scratch, off[6]));
  auto fail = [&](hipError_t e) {
    hipFree(scratch);
    return akz_hip_fail(e, "akz_synth_device_batch", __FILE__, __LINE__);
  };
  hipError_t e;
  if ((e = hipMemcpy(scratch + off[2], shapes.data(), b_sh, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(scratch + off[3], blobs.data(), b_bl, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(scratch + off[4], bg.data(), b_bg, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(scratch + off[5], seeds.data(), b_seed, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  float* noise = (float*)(scratch + off[0]);
  double* part = (double*)(scratch + off[1]);
  hipLaunchKernelGGL(k_synth_noise, dim3(SN_BLOCKS, 1, n), dim3(256), 0, nullptr, (const uint64_t*)(scratch + off[5]),
                     w, h, noise, nstride, part);
  hipLaunchKernelGGL(k_synth_compose, dim3((w + 63) / 64, (h + 3) / 4, n), dim3(256), 0, nullptr,
                     (const float*)(scratch + off[4]), (const Shape*)(scratch + off[2]), nshapes,
                     (const Blob*)(scratch + off[3]), nblobs, w, h, noise, nstride, part, d_out,
                     (int)(pitch_bytes / sizeof(float)), ostride);
  if ((e = hipGetLastError()) != hipSuccess) return fail(e);
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e);
  hipFree(scratch);
  return AKZ_OK;
}
