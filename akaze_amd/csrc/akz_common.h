// Shared definitions of the HIP A-KAZE library (device and host side).
//
// Floating-point contract: all HIP code is compiled with -ffp-contract=off and
// keeps the operand order of the reference CUDA kernels, so every stage is a
// deterministic function of its inputs.  Transcendentals on the path (atan2 in
// orientation, sin/cos in MLDB) are the Cephes single-precision methods
// (S. L. Moshier, "Methods and Programs for Mathematical Functions", 1989)
// written out below; the Gaussian taps and orientation weights are computed
// once on the host.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/akaze_mi.h"

#define AKZ_MAX_LEVELS 32
#define AKZ_MAX_TAU 256
#define AKZ_NB_PER_CAND 4     // CSR neighbour-list capacity: this many entries per candidate on average
#define AKZ_DESC_STRIDE 64    // descriptor rows are 61 bytes padded to 64

#define AKZ_HD __host__ __device__ __forceinline__

// Detector candidate / keypoint, cv::KeyPoint layout.  Candidates keep the
// reference's encodings: `octave` holds float bits (|.| = octave + |dx|,
// sign = sign of dx) and `angle` holds the y subpixel offset.
struct alignas(4) akz_kp {
  float x, y, size, angle, response;
  int32_t octave, class_id;
};
static_assert(sizeof(akz_kp) == 28, "KeyPoint wire layout");

// Per-level geometry as seen by the kernels.
struct akz_level_desc {
  int w, h, pitch;          // pitch in floats
  int octave, sublevel;
  int deriv_step;
  float border, det_size;   // FindExtrema border and size (AKAZE.cpp:353-367)
  int seg_base;             // offset of this level's row-segment counts
  int nseg_x;               // row segments per row: ceil(w / 64)
  // the coordinates FindExtrema's border test accepts, (int)(c - border + 0.5f)
  // - 1 >= 0 && (int)(c + border + 0.5f) + 1 < size, as intervals [lo, hi)
  // evaluated once on the host with the same float expression
  int xlo, xhi, ylo, yhi;
};

// The interval of coordinates c in [0, size) passing FindExtrema's border
// test (cuda_akaze.cu:749-752); the test is monotone on each side.
inline void akz_border_interval(float border, int size, int* lo, int* hi) {
  auto in = [&](int c) {
    return (int)((float)c - border + 0.5f) - 1 >= 0 && (int)((float)c + border + 0.5f) + 1 < size;
  };
  int a = 0;
  while (a < size && !in(a)) a++;
  int b = a;
  while (b < size && in(b)) b++;
  *lo = a;
  *hi = b;
}

// Status block of one image (read back with one copy): kept keypoints,
// overflow flags, candidates before the capacity cut, suppression waves.
enum { AKZ_ST_KEPT = 0, AKZ_ST_FLAGS = 1, AKZ_ST_TOTAL = 2, AKZ_ST_WAVES = 3 };

// Batch of same-size images: image z's arrays sit z * zs bytes after image
// 0's, and kernels launched with grid.z = nz pick their image by blockIdx.z.
struct akz_batch {
  int nz;
  size_t zs;
};

#ifdef __HIPCC__
// The offset pointer goes through the global address space and back: the
// integer round trip would otherwise hide where it points, and every access
// through it would be a flat instruction (which also counts against the LDS
// wait counter, so an LDS wait would wait for outstanding stores too).  From
// the cast the compiler infers global accesses with a scalar base.
template <class T>
__device__ __forceinline__ T* akz_zoff(T* p, size_t zs, int z) {
  using G = __attribute__((address_space(1))) T*;
  G g = (G)(reinterpret_cast<uintptr_t>(p) + (size_t)z * zs);
  return (T*)g;
}
template <class T>
__device__ __forceinline__ T* akz_zoff(T* p, size_t zs) {
  return akz_zoff(p, zs, (int)blockIdx.z);
}

// Order-preserving key of a float (a < b for finite a, b gives key(a) <
// key(b)), so that an unsigned atomicMax keeps a float maximum; and back.
__device__ __forceinline__ uint32_t akz_det_key(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float akz_det_unkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// 1/x for x >= 1, equal to IEEE division: the hardware reciprocal (within one
// ulp) and one Newton step of fused multiply-adds round to the same float as
// 1.0f / x for every float x in [1, 2^126) (checked for all of them on
// gfx950); beyond that, where 1/x leaves the normal range or x is infinite,
// the division itself.  Three instructions instead of the division sequence.
__device__ __forceinline__ float akz_rcp_ge1(float x) {
  if (__builtin_expect(x >= 0x1p126f, 0)) return 1.0f / x;
  const float y = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, y, 1.0f), y, y);
}

// Index of the calling wavefront in its workgroup.  It is uniform across the
// wavefront; read through readfirstlane the compiler knows that, and rows,
// loop bounds and addresses derived from it become scalar work.
__device__ __forceinline__ int akz_wave() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Workgroups of a launch are dispatched round-robin over MI355X's 8 XCDs,
// each with its own L2: workgroup i (x fastest, then y, then z) runs on XCD
// i % 8.  Renumbering them so that XCD k takes the k-th eighth of the tiles
// in that raster order keeps neighbouring tiles, which read each other's
// halo and share the cache lines at their edges, in one L2.  A bijection for
// any grid; only the placement changes.
struct akz_tile3 { int x, y, z; };
__device__ __forceinline__ akz_tile3 akz_xcd_tile() {
  const uint32_t gx = gridDim.x, gy = gridDim.y;
  const uint32_t nb = gx * gy * gridDim.z;
  const uint32_t id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const uint32_t q = nb / 8, r = nb % 8, xcd = id % 8, k = id / 8;
  const uint32_t t = xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
  akz_tile3 o;
  o.x = (int)(t % gx);
  o.y = (int)((t / gx) % gy);
  o.z = (int)(t / (gx * gy));
  return o;
}
#endif

AKZ_HD int akz_imin(int a, int b) { return a < b ? a : b; }
AKZ_HD int akz_imax(int a, int b) { return a > b ? a : b; }
AKZ_HD int akz_iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
AKZ_HD int akz_reflect101(int v, int n) { return v < 0 ? -v : (v >= n ? 2 * n - 2 - v : v); }

AKZ_HD float akz_bits2f(int32_t b) { return __builtin_bit_cast(float, b); }
AKZ_HD int32_t akz_f2bits(float f) { return __builtin_bit_cast(int32_t, f); }

// ---- Cephes single-precision atan / atan2 / sin / cos -------------------
AKZ_HD float akz_atanf(float xx) {
  float x = xx, y, z;
  bool sign = false;
  if (x < 0.0f) { sign = true; x = -x; }
  if (x > 2.414213562373095f) {
    y = 1.5707963267948966f;
    x = -(1.0f / x);
  } else if (x > 0.4142135623730950f) {
    y = 0.7853981633974483f;
    x = (x - 1.0f) / (x + 1.0f);
  } else {
    y = 0.0f;
  }
  z = x * x;
  y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x + x);
  return sign ? -y : y;
}

AKZ_HD float akz_atan2f(float y, float x) {
  const float PIF = 3.14159265358979f, PIO2F = 1.5707963267948966f;
  if (x == 0.0f) {
    if (y > 0.0f) return PIO2F;
    if (y < 0.0f) return -PIO2F;
    if (__builtin_signbit(x)) return __builtin_signbit(y) ? -PIF : PIF;
    return y;
  }
  if (y == 0.0f) {
    if (x > 0.0f) return y;
    return __builtin_signbit(y) ? -PIF : PIF;
  }
  float z = akz_atanf(y / x);
  if (x < 0.0f) z = (y < 0.0f) ? z - PIF : z + PIF;
  return z;
}

AKZ_HD void akz_sincosf(float xx, float* s, float* c) {
  const float FOPI = 1.27323954473516f;
  const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
  float x = xx;
  int sgn_s = 1;
  if (x < 0.0f) { x = -x; sgn_s = -1; }
  int j = (int)(FOPI * x);
  float y = (float)j;
  if (j & 1) { j += 1; y += 1.0f; }
  j &= 7;
  int sgn_c = 1;
  if (j > 3) { j -= 4; sgn_s = -sgn_s; sgn_c = -sgn_c; }
  if (j > 1) sgn_c = -sgn_c;
  x = ((x - y * DP1) - y * DP2) - y * DP3;
  float z = x * x;
  float sp = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
  float cp = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z - 0.5f * z + 1.0f;
  float sv, cv;
  if (j == 1 || j == 2) { sv = cp; cv = sp; } else { sv = sp; cv = cp; }
  *s = sgn_s < 0 ? -sv : sv;
  *c = sgn_c < 0 ? -cv : cv;
}

// ---- error handling -------------------------------------------------------
#define AKZ_HIP_TRY(expr)                                  \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return akz_hip_fail(_e, #expr, __FILE__, __LINE__); \
  } while (0)

akz_status akz_hip_fail(hipError_t e, const char* what, const char* file, int line);
