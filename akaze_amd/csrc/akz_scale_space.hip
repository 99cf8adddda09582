// Nonlinear scale space kernels: separable Gaussian, contrast factor,
// conductance, FED step, downsampling.
//
// Semantics follow the reference CUDA kernels (src/lib/cuda_akaze.cu), with
// the operand order of every float expression kept so that results equal the
// CPU oracle bit for bit (the library is built with -ffp-contract=off).
#include "akz_kernels.h"

namespace {

constexpr int GT_X = 64;   // Gaussian tile width (one wavefront per row)
constexpr int GT_Y = 16;   // Gaussian tile height
constexpr int GT_THREADS = 256;

// Separable Gaussian, row pass then column pass, replicate border
// (ConvRowGPU/ConvColGPU, cuda_akaze.cu:56-128), fused through LDS: the tile
// plus an R-pixel halo is loaded once, the row pass is kept in LDS and the
// column pass reads it.  Sums run i = 0..2R in the reference's order.
template <int R>
__global__ __launch_bounds__(GT_THREADS) void k_gauss(const float* __restrict__ in, int in_pitch,
                                                      float* __restrict__ out, int out_pitch, int w, int h,
                                                      akz_taps taps) {
  constexpr int AW = GT_X + 2 * R, AH = GT_Y + 2 * R;
  __shared__ float A[AH][AW];
  __shared__ float B[AH][GT_X];
  const int x0 = blockIdx.x * GT_X, y0 = blockIdx.y * GT_Y;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < AH * AW; idx += GT_THREADS) {
    int r = idx / AW, c = idx - r * AW;
    int gy = akz_iclamp(y0 - R + r, 0, h - 1);
    int gx = akz_iclamp(x0 - R + c, 0, w - 1);
    A[r][c] = in[(size_t)gy * in_pitch + gx];
  }
  __syncthreads();
  for (int idx = tid; idx < AH * GT_X; idx += GT_THREADS) {
    int r = idx / GT_X, c = idx - r * GT_X;
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i <= 2 * R; i++) s += A[r][c + i] * taps.k[i];
    B[r][c] = s;
  }
  __syncthreads();
  const int c = tid & (GT_X - 1);
  for (int r = tid / GT_X; r < GT_Y; r += GT_THREADS / GT_X) {
    int gx = x0 + c, gy = y0 + r;
    if (gx < w && gy < h) {
      float s = 0.0f;
#pragma unroll
      for (int i = 0; i <= 2 * R; i++) s += B[r + i][c] * taps.k[i];
      out[(size_t)gy * out_pitch + gx] = s;
    }
  }
}

__global__ void k_contrast_init(uint32_t* maxbits, int32_t* hist, uint32_t* ticket) {
  int t = threadIdx.x;
  for (int i = t; i < 512; i += blockDim.x) hist[i] = 0;
  if (t == 0) { *maxbits = 0u; *ticket = 0u; }
}

// Gradient magnitude of the sigma=1 blur with the unnormalised Scharr pair of
// MaxContrast (cuda_akaze.cu:552-579): the 3x3 window with top-left (x, y)
// gives the value stored at (x+1, y+1); interior pixels only.  The maximum
// is reduced per wavefront and merged with one atomicMax on the float bits.
__global__ __launch_bounds__(256) void k_contrast_grad(const float* __restrict__ blur, int pitch,
                                                       float* __restrict__ grad, int w, int h, uint32_t* maxbits) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  float gv = 0.0f;
  if (x < w - 2 && y < h - 2) {
    const float* b = blur + (size_t)y * pitch + x;
    const int W = pitch;
    float dx = 3.0f * (b[0] - b[2] + b[2 * W] - b[2 * W + 2]) + 10.0f * (b[W] - b[W + 2]);
    float dy = 3.0f * (b[0] + b[2] - b[2 * W] - b[2 * W + 2]) + 10.0f * (b[1] - b[2 * W + 1]);
    gv = sqrtf(dx * dx + dy * dy);
    grad[(size_t)(y + 1) * pitch + (x + 1)] = gv;
  }
  uint32_t bits = __float_as_uint(gv);
  for (int off = 32; off > 0; off >>= 1) {
    uint32_t o = __shfl_xor(bits, off, 64);
    bits = bits > o ? bits : o;
  }
  if ((threadIdx.x & 63) == 0 && bits) atomicMax(maxbits, bits);
}

struct akz_octaves { int nlev; int oct[AKZ_MAX_LEVELS]; };

// Histogram of grad/max over interior pixels (HistContrast, :581-602) and the
// percentile search of ContrastPercentile (:637-642), done by the last block
// to finish so that no value round-trips through the host.
__global__ __launch_bounds__(256) void k_contrast_hist(const float* __restrict__ grad, int pitch, int w, int h,
                                                       const uint32_t* maxbits, int nbins, float perc,
                                                       int32_t* hist, uint32_t* ticket, float* kc,
                                                       akz_octaves octs) {
  __shared__ int32_t lh[512];
  __shared__ bool last;
  for (int i = threadIdx.x; i < 512; i += 256) lh[i] = 0;
  __syncthreads();
  const float maxv = __uint_as_float(__hip_atomic_load(maxbits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  const float imaxval = 1.0f / maxv;
  const int y0 = blockIdx.y * 16;
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  if (x > 0 && x < w - 1) {
    for (int y = y0 + (threadIdx.x >> 6); y < y0 + 16; y += 4) {
      if (y > 0 && y < h - 1) {
        float t = nbins * grad[(size_t)y * pitch + x] * imaxval;
        int idx = (t != t) ? 0 : (int)t;
        idx = akz_imin(idx, nbins - 1);
        atomicAdd(&lh[idx], 1);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += 256)
    if (lh[i] > 0) atomicAdd(&hist[i], lh[i]);
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t nb = gridDim.x * gridDim.y;
    last = (atomicAdd(ticket, 1u) == nb - 1);
  }
  __syncthreads();
  if (!last || threadIdx.x != 0) return;
  __threadfence();
  int npoints = (w - 2) * (h - 2);
  int nthreshold = (int)(npoints * perc);
  int k = 0, nelements = 0;
  for (k = 0; nelements < nthreshold && k < nbins; k++)
    nelements += __hip_atomic_load(&hist[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  float contrast = (nelements < nthreshold ? 0.03f : maxv * ((float)k / nbins));
  // options_.kcontrast decays by 0.75 at each new octave (AKAZE.cpp:254-256)
  kc[0] = contrast;
  for (int i = 1; i < octs.nlev; i++) {
    if (octs.oct[i] > octs.oct[i - 1]) contrast = contrast * 0.75;
    kc[i] = contrast;
  }
}

// Conductance of Flow (cuda_akaze.cu:240-270): unnormalised Scharr of Lsmooth
// with reflect-101 borders, then the diffusivity function.
__global__ __launch_bounds__(256) void k_flow(const float* __restrict__ L, int pitch, float* __restrict__ g, int w,
                                              int h, int type, const float* kcp) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  float kcontrast = *kcp;
  float invk = 1.0f / (kcontrast * kcontrast);
  int ym = akz_reflect101(y - 1, h), yp = akz_reflect101(y + 1, h);
  int xm = akz_reflect101(x - 1, w), xp = akz_reflect101(x + 1, w);
  const float* rm = L + (size_t)ym * pitch;
  const float* r0 = L + (size_t)y * pitch;
  const float* rp = L + (size_t)yp * pitch;
  float ul = rm[xm], ur = rm[xp], ll = rp[xm], lr = rp[xp];
  float lx = 3.0f * (lr - ll + ur - ul) + 10.0f * (r0[xp] - r0[xm]);
  float ly = 3.0f * (lr + ll - ur - ul) + 10.0f * (rp[x] - rm[x]);
  float dif2 = invk * (lx * lx + ly * ly);
  float r;
  if (type == 0)
    r = expf(-dif2);
  else if (type == 1)
    r = 1.0f / (1.0f + dif2);
  else if (type == 2)
    r = 1.0f - (float)exp(-3.315 / (dif2 * dif2 * dif2 * dif2));
  else
    r = 1.0f / sqrtf(1.0f + dif2);
  g[(size_t)y * pitch + x] = r;
}

// One FED step (NLDStep + NLDUpdate, cuda_akaze.cu:289-345), out of place:
// L_out = L_in + (tau/2) * div(g grad L_in), neighbours clamped at the border.
__global__ __launch_bounds__(256) void k_fed_step(const float* __restrict__ L, const float* __restrict__ g,
                                                  float* __restrict__ out, int pitch, int w, int h,
                                                  float stepsize) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  int ym = akz_imax(y - 1, 0), yp = akz_imin(y + 1, h - 1);
  int xm = akz_imax(x - 1, 0), xp = akz_imin(x + 1, w - 1);
  size_t c = (size_t)y * pitch + x;
  float i0 = L[c], f0 = g[c];
  size_t e = (size_t)y * pitch + xp, wv = (size_t)y * pitch + xm;
  size_t s = (size_t)yp * pitch + x, n = (size_t)ym * pitch + x;
  float xpos = (f0 + g[e]) * (L[e] - i0);
  float xneg = (f0 + g[wv]) * (i0 - L[wv]);
  float ypos = (f0 + g[s]) * (L[s] - i0);
  float yneg = (f0 + g[n]) * (i0 - L[n]);
  out[c] = i0 + stepsize * (xpos - xneg + ypos - yneg);
}

// 2x2 box mean for exact halving (HalfSample2, cuda_akaze.cu:432-440).
__global__ __launch_bounds__(256) void k_half2(const float* __restrict__ in, int ipitch, float* __restrict__ out,
                                               int ow, int oh, int opitch) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= ow || y >= oh) return;
  const float* p = in + (size_t)(2 * y) * ipitch + 2 * x;
  out[(size_t)y * opitch + x] = 0.25f * (p[0] + p[1] + p[ipitch] + p[ipitch + 1]);
}

// Area downsampling for sizes that do not halve exactly (HalfSample,
// cuda_akaze.cu:385-430), including its ptri[1] read (:413) on output rows
// y % 16 == 15 when both width and height are odd.
__device__ __forceinline__ float hs_row(const float* in, int ipitch, int iw, int r, int x, int ow, bool quirk) {
  const float* p = in + (size_t)r * ipitch + 2 * x;
  if (2 * ow == iw) return ow * (p[0] + p[1]);
  float f0 = ow - x, f2 = 1 + x;
  return f0 * p[0] + ow * p[1] + f2 * (quirk ? p[1] : p[2]);
}

__global__ __launch_bounds__(256) void k_half_general(const float* __restrict__ in, int iw, int ih, int ipitch,
                                                      float* __restrict__ out, int ow, int oh, int opitch) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= ow || y >= oh) return;
  bool oddw = (2 * ow != iw), oddh = (2 * oh != ih);
  float b0 = hs_row(in, ipitch, iw, 2 * y, x, ow, false);
  float b16 = hs_row(in, ipitch, iw, 2 * y + 1, x, ow, false);
  float r;
  if (!oddh) {
    r = oh * (b0 + b16) / (iw * ih);
  } else {
    float b32 = hs_row(in, ipitch, iw, 2 * y + 2, x, ow, oddw && (y % 16 == 15));
    float f0 = oh - y, f2 = 1 + y;
    r = (f0 * b0 + oh * b16 + f2 * b32) / (iw * ih);
  }
  out[(size_t)y * opitch + x] = r;
}

__global__ __launch_bounds__(256) void k_copy2d(const float* __restrict__ in, int ipitch, float* __restrict__ out,
                                                int opitch, int w, int h) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  out[(size_t)y * opitch + x] = in[(size_t)y * ipitch + x];
}

inline dim3 grid64x4(int w, int h) { return dim3((w + 63) / 64, (h + 3) / 4); }

}  // namespace

void akz_launch_gauss(hipStream_t s, const float* in, int in_pitch, float* out, int out_pitch, int w, int h, int R,
                      const akz_taps& taps) {
  dim3 grid((w + GT_X - 1) / GT_X, (h + GT_Y - 1) / GT_Y);
  switch (R) {
    case 2: hipLaunchKernelGGL(k_gauss<2>, grid, dim3(GT_THREADS), 0, s, in, in_pitch, out, out_pitch, w, h, taps); break;
    case 3: hipLaunchKernelGGL(k_gauss<3>, grid, dim3(GT_THREADS), 0, s, in, in_pitch, out, out_pitch, w, h, taps); break;
    case 4: hipLaunchKernelGGL(k_gauss<4>, grid, dim3(GT_THREADS), 0, s, in, in_pitch, out, out_pitch, w, h, taps); break;
    default: hipLaunchKernelGGL(k_gauss<5>, grid, dim3(GT_THREADS), 0, s, in, in_pitch, out, out_pitch, w, h, taps); break;
  }
}

void akz_launch_contrast_init(hipStream_t s, uint32_t* maxbits, int32_t* hist, uint32_t* ticket) {
  hipLaunchKernelGGL(k_contrast_init, dim3(1), dim3(256), 0, s, maxbits, hist, ticket);
}

void akz_launch_contrast_grad(hipStream_t s, const float* blur, int pitch, float* grad, int w, int h,
                              uint32_t* maxbits) {
  hipLaunchKernelGGL(k_contrast_grad, grid64x4(w, h), dim3(256), 0, s, blur, pitch, grad, w, h, maxbits);
}

void akz_launch_contrast_hist(hipStream_t s, const float* grad, int pitch, int w, int h, const uint32_t* maxbits,
                              int nbins, float perc, int32_t* hist, uint32_t* ticket, float* kc, int nlev,
                              const int* octaves_host) {
  akz_octaves o;
  o.nlev = nlev;
  for (int i = 0; i < nlev; i++) o.oct[i] = octaves_host[i];
  dim3 grid((w + 63) / 64, (h + 15) / 16);
  hipLaunchKernelGGL(k_contrast_hist, grid, dim3(256), 0, s, grad, pitch, w, h, maxbits, nbins, perc, hist, ticket,
                     kc, o);
}

void akz_launch_flow(hipStream_t s, const float* Lsm, int pitch, float* g, int w, int h, int type, const float* kc) {
  hipLaunchKernelGGL(k_flow, grid64x4(w, h), dim3(256), 0, s, Lsm, pitch, g, w, h, type, kc);
}

void akz_launch_fed_step(hipStream_t s, const float* Lin, const float* g, float* Lout, int pitch, int w, int h,
                         float half_tau) {
  hipLaunchKernelGGL(k_fed_step, grid64x4(w, h), dim3(256), 0, s, Lin, g, Lout, pitch, w, h, half_tau);
}

void akz_launch_halfsample(hipStream_t s, const float* in, int iw, int ih, int ipitch, float* out, int ow, int oh,
                           int opitch) {
  if (iw == 2 * ow && ih == 2 * oh)
    hipLaunchKernelGGL(k_half2, grid64x4(ow, oh), dim3(256), 0, s, in, ipitch, out, ow, oh, opitch);
  else
    hipLaunchKernelGGL(k_half_general, grid64x4(ow, oh), dim3(256), 0, s, in, iw, ih, ipitch, out, ow, oh, opitch);
}

void akz_launch_copy2d(hipStream_t s, const float* in, int ipitch, float* out, int opitch, int w, int h) {
  hipLaunchKernelGGL(k_copy2d, grid64x4(w, h), dim3(256), 0, s, in, ipitch, out, opitch, w, h);
}
