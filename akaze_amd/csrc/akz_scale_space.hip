// Nonlinear scale space kernels: separable Gaussian, contrast factor,
// conductance, FED step, downsampling.
//
// Semantics follow the reference CUDA kernels (src/lib/cuda_akaze.cu), with
// the operand order of every float expression kept so that results equal the
// CPU oracle bit for bit (the library is built with -ffp-contract=off).
#include "akz_kernels.h"

namespace {

// Copies an AH x AW window (top-left at image (x0, y0), coordinates clamped
// into the image) into LDS.  Each thread first issues all of its loads, then
// writes them, so a block waits for one memory latency instead of one per
// element.
template <int AH, int AW, int NT>
__device__ __forceinline__ void load_window(float* __restrict__ dst, const float* __restrict__ src, int pitch,
                                            int x0, int y0, int w, int h) {
  constexpr int PER = (AH * AW + NT - 1) / NT;
  float v[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int idx = threadIdx.x + k * NT;
    if (idx < AH * AW) {
      const int r = idx / AW, c = idx - r * AW;
      v[k] = src[(size_t)akz_iclamp(y0 + r, 0, h - 1) * pitch + akz_iclamp(x0 + c, 0, w - 1)];
    }
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int idx = threadIdx.x + k * NT;
    if (idx < AH * AW) dst[idx] = v[k];
  }
}

// load_window of the image HalfSample2 makes of `src`, a parent image of
// exactly twice the size (cuda_akaze.cu:432-440): each element is the 2x2
// box mean of the parent at twice its clamped coordinates, in k_half2's
// expression, so the window equals a load of k_half2's output bit for bit.
template <int AH, int AW, int NT>
__device__ __forceinline__ void load_window_half(float* __restrict__ dst, const float* __restrict__ src, int ppitch,
                                                 int x0, int y0, int w, int h) {
  constexpr int PER = (AH * AW + NT - 1) / NT;
  float v[PER][4];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int idx = threadIdx.x + k * NT;
    if (idx < AH * AW) {
      const int r = idx / AW, c = idx - r * AW;
      const float* p = src + (size_t)(2 * akz_iclamp(y0 + r, 0, h - 1)) * ppitch + 2 * akz_iclamp(x0 + c, 0, w - 1);
      v[k][0] = p[0];
      v[k][1] = p[1];
      v[k][2] = p[ppitch];
      v[k][3] = p[ppitch + 1];
    }
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int idx = threadIdx.x + k * NT;
    if (idx < AH * AW) dst[idx] = 0.25f * (v[k][0] + v[k][1] + v[k][2] + v[k][3]);
  }
}

struct akz_octaves { int nlev; int oct[AKZ_MAX_LEVELS]; };

// Wavefront lane shifts by DPP with zero fill (one instruction, no copy of an
// old value): the lane with no source receives 0.
__device__ __forceinline__ float from_right(float v) {  // lane i receives lane i+1 (lane 63 receives 0)
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_left(float v) {  // lane i receives lane i-1 (lane 0 receives 0)
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}


// Front end of the contrast factor and the level-0 image in one pass over the
// input (coordinates clamped: the replicate border of LowPass):
// * Lt of level 0, the Gaussian of soffset with R0-pixel taps (ConvRowGPU /
//   ConvColGPU, cuda_akaze.cu:56-128; AKAZE.cpp:234-236), row pass then
//   column pass, sums in tap order from 0.0f;
// * the sigma=1 blur of ContrastPercentile (:604-649), the same way, in
//   registers only;
// * the gradient magnitude of MaxContrast (:552-579): the unnormalised Scharr
//   pair over the blur's 3x3 window centred on the pixel, interior pixels only;
// * the strip's largest gradient (float bits; all values are >= 0).
// One wavefront per 64-column strip whose lanes M..63-M (M = max(R0, 3)) are
// output columns, TH output rows.  The TH + 2M input rows stream through the
// wavefront once; each row's horizontal taps come from DPP lane shifts,
// shared by the soffset Gaussian and the sigma=1 blur; the column passes and
// the Scharr rows are register rings.  No LDS and no barriers (an LDS-tiled
// form with 64x16 tiles took 8% longer on 16-image batches).
template <int R0, int TH>
__global__ __launch_bounds__(64) void k_contrast_front(const float* __restrict__ in, int in_pitch,
                                                           size_t in_zs, float* __restrict__ lt,
                                                           float* __restrict__ grad, int pitch, int w, int h,
                                                           akz_taps taps_c, akz_taps taps0,
                                                           uint32_t* __restrict__ blockmax,
                                                           int32_t* __restrict__ hist_tot, size_t zs) {
  constexpr int M = R0 > 3 ? R0 : 3, TW = 64 - 2 * M, NA = TH + 2 * M;
  constexpr int K0 = 2 * R0 + 1;
  const akz_tile3 tl = akz_xcd_tile();
  in = akz_zoff(in, in_zs, tl.z);
  lt = akz_zoff(lt, zs, tl.z);
  grad = akz_zoff(grad, zs, tl.z);
  blockmax = akz_zoff(blockmax, zs, tl.z);
  // the image's histogram, which k_contrast_hist adds to, starts at zero
  if (tl.x == 0 && tl.y == 0) {
    int32_t* ht = akz_zoff(hist_tot, zs, tl.z);
    for (int i = threadIdx.x; i < 512; i += 64) ht[i] = 0;
  }
  const int lane = threadIdx.x;
  const int x0 = tl.x * TW - M, y0 = tl.y * TH;
  const int gx = x0 + lane;
  const int cx = akz_iclamp(gx, 0, w - 1);
  const bool out_col = lane >= M && lane < 64 - M && gx < w;
  const bool grad_col = out_col && gx >= 1 && gx <= w - 2;
  float k0[K0], kc[5];
#pragma unroll
  for (int i = 0; i < K0; i++) k0[i] = taps0.k[i];
#pragma unroll
  for (int i = 0; i < 5; i++) kc[i] = taps_c.k[i];
  // rings: B0 (Gaussian row pass) rows i-2R0..i, B1 (blur row pass) rows
  // i-4..i, C (blur) rows with their x-1 / x+1 neighbours
  float B0[K0], B1[5], C[3], CL[3], CR[3];
  uint32_t bits = 0u;
  // every input row of the strip is loaded before the first is used: loaded
  // row by row, the loads wait one after another and the kernel is bound by
  // memory latency, not by its arithmetic
  float rin[NA];
#pragma unroll
  for (int i = 0; i < NA; i++) rin[i] = in[(size_t)akz_iclamp(y0 - M + i, 0, h - 1) * in_pitch + cx];
#pragma unroll
  for (int i = 0; i < NA; i++) {
    const float a = rin[i];  // input row y0 - M + i
    // v[d + R0]: the row at column x + d, d = -R0..R0
    float v[K0];
    v[R0] = a;
#pragma unroll
    for (int d = 1; d <= R0; d++) {
      v[R0 - d] = from_left(v[R0 - d + 1]);
      v[R0 + d] = from_right(v[R0 + d - 1]);
    }
    {
      float sum = 0.0f;
#pragma unroll
      for (int t = 0; t < K0; t++) sum += v[t] * k0[t];
      B0[i % K0] = sum;
    }
    {
      float sum = 0.0f;
#pragma unroll
      for (int t = 0; t < 5; t++) sum += v[R0 - 2 + t] * kc[t];
      B1[i % 5] = sum;
    }
    // Lt row y0 + i - R0 - M: B0 rows of that row -R0 .. +R0
    {
      const int r = i - R0 - M;
      if (r >= 0 && r < TH) {
        float sum = 0.0f;
#pragma unroll
        for (int t = 0; t < K0; t++) sum += B0[(i - 2 * R0 + t) % K0] * k0[t];
        const int y = y0 + r;
        if (out_col && y < h) lt[(size_t)y * pitch + gx] = sum;
      }
    }
    // blur row yc = y0 - M + i - 2 (B1 rows yc-2 .. yc+2), kept for rows
    // y0-1 .. y0+TH; the gradient of row yc - 1 once row yc exists
    const int jc = i - M - 2 + 1;  // blur ring index: yc = y0 - 1 + jc
    if (jc >= 0 && jc <= TH + 1) {
      float sum = 0.0f;
#pragma unroll
      for (int t = 0; t < 5; t++) sum += B1[(i - 4 + t) % 5] * kc[t];
      C[jc % 3] = sum;
      CL[jc % 3] = from_left(sum);
      CR[jc % 3] = from_right(sum);
      const int r = jc - 2;  // gradient row y0 + r: blur rows r-1, r, r+1 are jc-2, jc-1, jc
      if (r >= 0) {
        const int Y = y0 + r;
        const int u = (jc - 2) % 3, m = (jc - 1) % 3, d = jc % 3;
        if (grad_col && Y >= 1 && Y <= h - 2) {
          const float dx = 3.0f * (CL[u] - CR[u] + CL[d] - CR[d]) + 10.0f * (CL[m] - CR[m]);
          const float dy = 3.0f * (CL[u] + CR[u] - CL[d] - CR[d]) + 10.0f * (C[u] - C[d]);
          const float gv = sqrtf(dx * dx + dy * dy);
          grad[(size_t)Y * pitch + gx] = gv;
          const uint32_t gb = __float_as_uint(gv);
          bits = bits > gb ? bits : gb;
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(bits, off, 64);
    bits = bits > o ? bits : o;
  }
  if (lane == 0) blockmax[tl.y * (int)gridDim.x + tl.x] = bits;
}

constexpr int CT_HIST_BLOCKS = 512;
constexpr int CT_HIST_THREADS = 512;

// Histogram of grad/max over interior pixels (HistContrast, :581-602).  Every
// block first reduces the strip maxima of k_contrast_front to the global
// maximum (block 0 stores it for the percentile).  Bins are kept in 16
// columns (LDS hist[bin][lane % 16]): at most four lanes of a wavefront add
// to one counter, and the 19 KB table clears and sums four times faster
// than one column per lane; each wavefront walks whole rows; a block adds
// its bin totals to the image's histogram (integer counts: the sums do not
// depend on the order), which k_contrast_front cleared.
__global__ __launch_bounds__(CT_HIST_THREADS) void k_contrast_hist(const float* __restrict__ grad, int pitch,
                                                                   int w, int h, const uint32_t* blockmax, int nbm,
                                                                   int nbins, int32_t* __restrict__ hist_tot,
                                                                   uint32_t* maxbits, size_t zs) {
  // 16 columns per bin, each shared by lanes l, l+16, l+32, l+48; bins >= 300
  // (nbins <= 512) go to `tail`
  __shared__ int32_t lh[300 * 16];
  __shared__ int32_t tail[212];
  __shared__ uint32_t s_w[CT_HIST_THREADS / 64];
  grad = akz_zoff(grad, zs);
  blockmax = akz_zoff(blockmax, zs);
  hist_tot = akz_zoff(hist_tot, zs);
  maxbits = akz_zoff(maxbits, zs);
  const int lane = threadIdx.x & 63, wv = akz_wave();
  constexpr int NW = CT_HIST_THREADS / 64;
  for (int i = threadIdx.x; i < 300 * 16; i += CT_HIST_THREADS) lh[i] = 0;
  for (int i = threadIdx.x; i < 212; i += CT_HIST_THREADS) tail[i] = 0;
  // strip maxima and gradient rows: eight independent loads in flight per
  // lane before any is used (a loop that waits on each load leaves the
  // kernel bound by memory latency at one wavefront per SIMD)
  constexpr int U = 8;
  uint32_t mb = 0u;
  for (int i0 = threadIdx.x; i0 < nbm; i0 += U * CT_HIST_THREADS) {
    uint32_t b[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * CT_HIST_THREADS;
      b[u] = i < nbm ? blockmax[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; u++) mb = mb > b[u] ? mb : b[u];
  }
  for (int off = 32; off > 0; off >>= 1) {
    uint32_t o = __shfl_xor(mb, off, 64);
    mb = mb > o ? mb : o;
  }
  if (lane == 0) s_w[wv] = mb;
  __syncthreads();
  mb = s_w[0];
  for (int i = 1; i < NW; i++) mb = mb > s_w[i] ? mb : s_w[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) *maxbits = mb;
  const float maxv = __uint_as_float(mb);
  const float imaxval = 1.0f / maxv;
  const int nwaves = gridDim.x * NW;
  for (int y = 1 + blockIdx.x * NW + wv; y < h - 1; y += nwaves) {
    const float* row = grad + (size_t)y * pitch;
    for (int x0 = 1 + lane; x0 < w - 1; x0 += U * 64) {
      float g[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int x = x0 + u * 64;
        g[u] = x < w - 1 ? row[x] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (x0 + u * 64 >= w - 1) break;
        float v = nbins * g[u] * imaxval;
        int idx = (v != v) ? 0 : (int)v;
        idx = akz_imin(idx, nbins - 1);
        if (idx < 300) atomicAdd(&lh[idx * 16 + (lane & 15)], 1);
        else atomicAdd(&tail[idx - 300], 1);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += CT_HIST_THREADS) {
    int sum = 0;
    if (i < 300) {
      for (int l = 0; l < 16; l++) sum += lh[i * 16 + ((l + i) & 15)];
    } else {
      sum = tail[i - 300];
    }
    if (sum != 0) atomicAdd(&hist_tot[i], sum);
  }
}

// Percentile search of ContrastPercentile (:637-642) over the image's
// histogram, on the device so no value round-trips through the host; the
// per-level factor decays by 0.75 at each new octave (AKAZE.cpp:254-256).
__global__ __launch_bounds__(1024) void k_contrast_finish(const int32_t* __restrict__ hist_tot, int nbins,
                                                          const uint32_t* maxbits, int w, int h, float perc, float* kc,
                                                          akz_octaves octs, size_t zs) {
  __shared__ int32_t tot[512];
  hist_tot = akz_zoff(hist_tot, zs);
  maxbits = akz_zoff(maxbits, zs);
  kc = akz_zoff(kc, zs);
  const int lane = threadIdx.x & 63, wv = akz_wave();
  // the bin totals k_contrast_hist's blocks added up
  if ((int)threadIdx.x < nbins) tot[threadIdx.x] = hist_tot[threadIdx.x];
  __syncthreads();
  // percentile: the loop of ContrastPercentile stops at k = 1 + the first bin
  // whose cumulative count reaches nthreshold (k = 0 when nthreshold <= 0);
  // found here with a block prefix sum over the bins
  __shared__ int32_t s_w[16];
  __shared__ int s_k;
  const int npoints = (w - 2) * (h - 2);
  const int nthreshold = (int)(npoints * perc);
  if (threadIdx.x == 0) s_k = -1;
  const int v = (threadIdx.x < nbins) ? tot[threadIdx.x] : 0;
  int inc = v;
  for (int d = 1; d < 64; d <<= 1) {
    int o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_w[wv] = inc;
  __syncthreads();
  int before = 0;
  for (int i = 0; i < wv; i++) before += s_w[i];
  inc += before;
  if (threadIdx.x < nbins && inc >= nthreshold && inc - v < nthreshold) s_k = threadIdx.x + 1;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const float maxv = __uint_as_float(*maxbits);
  float contrast;
  if (nthreshold <= 0) contrast = maxv * ((float)0 / nbins);
  else contrast = (s_k < 0) ? 0.03f : maxv * ((float)s_k / nbins);
  kc[0] = contrast;
  for (int i = 1; i < octs.nlev; i++) {
    if (octs.oct[i] > octs.oct[i - 1]) contrast = contrast * 0.75;
    kc[i] = contrast;
  }
}

// One FED step (NLDStep + NLDUpdate, cuda_akaze.cu:289-345), out of place:
// L_out = L_in + (tau/2) * div(g grad L_in), neighbours clamped at the border.
__global__ __launch_bounds__(256) void k_fed_step(const float* __restrict__ L, const float* __restrict__ g,
                                                  float* __restrict__ out, int pitch, int w, int h,
                                                  float stepsize, size_t zs) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + akz_wave();
  if (x >= w || y >= h) return;
  L = akz_zoff(L, zs);
  g = akz_zoff(g, zs);
  out = akz_zoff(out, zs);
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

// FED steps with temporal blocking.  A block holds an RW x RH region of L
// and g in LDS: its output tile plus a halo as wide as the launch's step
// count n (at most K).  Step s updates the part of the region still exact
// after s steps (it shrinks by one pixel per side and step, except along
// image borders, where neighbours are clamped to the pixel itself as in
// NLDStep), ping-ponging between two LDS buffers.  Per pixel and step the
// arithmetic is k_fed_step's, so the result equals one launch per step bit
// for bit.  RW and RH are powers of two, so the flattened region index
// splits into row and column with shifts.  Region shapes per level size are
// chosen from measurements (akz_fed_variant_for below).
constexpr int FB_K = 8;      // steps per launch of most variants
constexpr int FB_KMAX = 16;  // the deep-halo variants
struct akz_fed_steps { float half_tau[FB_KMAX]; int n; };

// The FED steps of one launch on an RW x RH LDS region (L in sa, g in sg;
// sb the second buffer; the region's origin is image pixel (x0, y0)).
// Thread t owns region pixels t + k * NT throughout the launch: g does not
// change within it, so each owned pixel keeps its conductance pair sums
// (f0 + gE, f0 + gW, f0 + gS, f0 + gN: the operands of k_fed_step's
// fluxes, formed once) and its border clamps in registers, and a step
// reads L only.  Step s updates the pixels at least s from the region edge
// inside the image (one pixel of halo is consumed per step); at image
// borders the outward neighbour is the pixel itself, as in NLDStep.
// Returns the buffer holding the last step.
template <int RW, int RH, int NT>
__device__ __forceinline__ float* fed_lds_steps(float* sa, float* sb, const float* sg, const akz_fed_steps& st,
                                                int x0, int y0, int w, int h) {
  constexpr int LW = (RW == 128) ? 7 : (RW == 64) ? 6 : (RW == 32 ? 5 : 4);
  static_assert((1 << LW) == RW, "power-of-two region width");
  constexpr int PX = (RH * RW + NT - 1) / NT;
  const int tid = threadIdx.x;
  // per owned pixel: margin to the region edge (-1 outside the image) in
  // bits 4.., self-neighbour flags E, W, S, N in bits 0..3
  int pk[PX];
  float se[PX], sw[PX], ss[PX], sn[PX];
#pragma unroll
  for (int k = 0; k < PX; k++) {
    const int idx = tid + k * NT;
    pk[k] = -1 << 4;
    se[k] = sw[k] = ss[k] = sn[k] = 0.0f;
    if (idx < RH * RW) {
      const int r = idx >> LW, c = idx & (RW - 1);
      const int gx = x0 + c, gy = y0 + r;
      if (gx >= 0 && gy >= 0 && gx < w && gy < h) {
        const int m = akz_imin(akz_imin(r, RH - 1 - r), akz_imin(c, RW - 1 - c));
        const bool ee = gx >= w - 1, ww = gx <= 0, es = gy >= h - 1, en = gy <= 0;
        pk[k] = (m << 4) | (ee ? 1 : 0) | (ww ? 2 : 0) | (es ? 4 : 0) | (en ? 8 : 0);
        const float f0 = sg[idx];
        se[k] = f0 + sg[ee ? idx : idx + 1];
        sw[k] = f0 + sg[ww ? idx : idx - 1];
        ss[k] = f0 + sg[es ? idx : idx + RW];
        sn[k] = f0 + sg[en ? idx : idx - RW];
      }
    }
  }
  float* src = sa;
  float* dst = sb;
  for (int s = 1; s <= st.n; s++) {
    const float stepsize = st.half_tau[s - 1];
    const float* __restrict__ cs = src;
    float* __restrict__ cd = dst;
#pragma unroll
    for (int k = 0; k < PX; k++) {
      if ((pk[k] >> 4) < s) continue;
      const int idx = tid + k * NT;
      const int e = (pk[k] & 1) ? idx : idx + 1, wi = (pk[k] & 2) ? idx : idx - 1;
      const int so = (pk[k] & 4) ? idx : idx + RW, no = (pk[k] & 8) ? idx : idx - RW;
      const float i0 = cs[idx];
      const float le = cs[e], lw = cs[wi], ls = cs[so], ln = cs[no];
      float xpos = se[k] * (le - i0);
      float xneg = sw[k] * (i0 - lw);
      float ypos = ss[k] * (ls - i0);
      float yneg = sn[k] * (i0 - ln);
      cd[idx] = i0 + stepsize * (xpos - xneg + ypos - yneg);
    }
    __syncthreads();
    float* t = src;
    src = dst;
    dst = t;
  }
  return src;
}

template <int RW, int RH, int NT, int K = FB_K>
__global__ __launch_bounds__(NT) void k_fed_multi(const float* __restrict__ L, const float* __restrict__ g,
                                                  float* __restrict__ out, int pitch, int w, int h,
                                                  akz_fed_steps st, size_t zs) {
  static_assert(RW > 2 * K && RH > 2 * K, "a tile remains inside the halo");
  const akz_tile3 tl = akz_xcd_tile();
  L = akz_zoff(L, zs, tl.z);
  g = akz_zoff(g, zs, tl.z);
  out = akz_zoff(out, zs, tl.z);
  constexpr int LW = (RW == 128) ? 7 : (RW == 64) ? 6 : (RW == 32 ? 5 : 4);
  static_assert((1 << LW) == RW, "power-of-two region width");
  __shared__ float sg[RH * RW];
  __shared__ float sa[RH * RW];
  __shared__ float sb[RH * RW];
  const int tid = threadIdx.x;
  const int hn = st.n;  // halo: one pixel per step
  const int x0 = tl.x * (RW - 2 * hn) - hn, y0 = tl.y * (RH - 2 * hn) - hn;
  load_window<RH, RW, NT>(sa, L, pitch, x0, y0, w, h);
  load_window<RH, RW, NT>(sg, g, pitch, x0, y0, w, h);
  __syncthreads();
  const float* src = fed_lds_steps<RW, RH, NT>(sa, sb, sg, st, x0, y0, w, h);
  for (int idx = hn * RW + tid; idx < (RH - hn) * RW; idx += NT) {
    const int r = idx >> LW, c = idx & (RW - 1);
    const int gx = x0 + c, gy = y0 + r;
    if (c >= hn && c < RW - hn && gx < w && gy < h) out[(size_t)gy * pitch + gx] = src[idx];
  }
}

// The first FED launch of a level fused with its Lsmooth and conductance
// (the work of k_smooth_flow_reg: sigma=1 LowPass with replicate border,
// unnormalised Scharr with reflect-101 neighbours, diffusivity TYPE; same
// operands in the same order).  A block loads the level's input over its
// RW x RH region plus 3 pixels (coordinates clamped), runs the row and column
// Gaussian and the flow in LDS, writes Lsmooth (and g when later launches of
// the level read it) for its tile, then runs the launch's FED steps as
// k_fed_multi.  One launch instead of two, and g never goes through HBM when
// the level takes one launch.  LDS: input window, row pass and Lsmooth with
// its ring; the FED buffers reuse the first two once they are dead.
// With `parent` given (the previous octave's last level, exactly twice the
// size, row pitch ppitch), the input is that image's HalfSample2, formed
// while loading: the first launch of an octave needs no separate halving.
template <int RW, int RH, int NT, int K = FB_K>
__global__ __launch_bounds__(NT) void k_fed_multi_sf(const float* __restrict__ in, const float* __restrict__ parent,
                                                     int ppitch, float* __restrict__ lsm, float* __restrict__ gout,
                                                     float* __restrict__ out, int pitch, int w, int h, int type,
                                                     const float* kcp, akz_taps taps, akz_fed_steps st, size_t zs) {
  static_assert(RW > 2 * K && RH > 2 * K, "a tile remains inside the halo");
  constexpr int AW = RW + 6, AH = RH + 6;  // input window
  constexpr int BW = RW + 2;               // row pass (AH rows) and Lsmooth ring (RH + 2 rows)
  constexpr int CH = RH + 2;
  constexpr int LW = (RW == 128) ? 7 : (RW == 64) ? 6 : (RW == 32 ? 5 : 4);
  static_assert((1 << LW) == RW, "power-of-two region width");
  static_assert(RH * RW <= AH * AW && RH * RW <= AH * BW, "FED buffers fit the dead windows");
  const akz_tile3 tl = akz_xcd_tile();
  lsm = akz_zoff(lsm, zs, tl.z);
  if (gout) gout = akz_zoff(gout, zs, tl.z);
  out = akz_zoff(out, zs, tl.z);
  kcp = akz_zoff(kcp, zs, tl.z);
  __shared__ float sA[AH * AW];  // input window, then the second FED buffer
  __shared__ float sB[AH * BW];  // row pass, then g
  __shared__ float sC[CH * BW];  // Lsmooth with a one-pixel ring
  __shared__ float sa[RH * RW];  // L at step 0
  float* const sb = sA;
  float* const sg = sB;
  const int tid = threadIdx.x;
  const int hn = st.n;
  const int x0 = tl.x * (RW - 2 * hn) - hn, y0 = tl.y * (RH - 2 * hn) - hn;
  if (parent)
    load_window_half<AH, AW, NT>(sA, akz_zoff(parent, zs, tl.z), ppitch, x0 - 3, y0 - 3, w, h);
  else
    load_window<AH, AW, NT>(sA, akz_zoff(in, zs, tl.z), pitch, x0 - 3, y0 - 3, w, h);
  __syncthreads();
  // L of the region (the window without its 3-pixel ring), and the row pass
  for (int idx = tid; idx < RH * RW; idx += NT) {
    const int r = idx >> LW, c = idx & (RW - 1);
    sa[idx] = sA[(r + 3) * AW + c + 3];
  }
  for (int idx = tid; idx < AH * BW; idx += NT) {
    const int r = idx / BW, c = idx - r * BW;
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < 5; i++) sum += sA[r * AW + c + i] * taps.k[i];
    sB[idx] = sum;
  }
  __syncthreads();
  // column pass: Lsmooth at image (y0 - 1 + r, x0 - 1 + c)
  for (int idx = tid; idx < CH * BW; idx += NT) {
    const int r = idx / BW, c = idx - r * BW;
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < 5; i++) sum += sB[(r + i) * BW + c] * taps.k[i];
    sC[idx] = sum;
  }
  __syncthreads();
  // flow on the region's in-image pixels (the FED steps never read g
  // outside the image); Lsmooth and g of the tile go out
  const float kcontrast = *kcp;
  const float invk = 1.0f / (kcontrast * kcontrast);
  for (int idx = tid; idx < RH * RW; idx += NT) {
    const int r = idx >> LW, c = idx & (RW - 1);
    const int gx = x0 + c, gy = y0 + r;
    if (gx < 0 || gy < 0 || gx >= w || gy >= h) continue;
    const int cy = r + 1, cx = c + 1;
    const int cym = akz_reflect101(gy - 1, h) - y0 + 1, cyp = akz_reflect101(gy + 1, h) - y0 + 1;
    const int cxm = akz_reflect101(gx - 1, w) - x0 + 1, cxp = akz_reflect101(gx + 1, w) - x0 + 1;
    const float ul = sC[cym * BW + cxm], ur = sC[cym * BW + cxp], ll = sC[cyp * BW + cxm], lr = sC[cyp * BW + cxp];
    const float lx = 3.0f * (lr - ll + ur - ul) + 10.0f * (sC[cy * BW + cxp] - sC[cy * BW + cxm]);
    const float ly = 3.0f * (lr + ll - ur - ul) + 10.0f * (sC[cyp * BW + cx] - sC[cym * BW + cx]);
    const float dif2 = invk * (lx * lx + ly * ly);
    float res;
    if (type == 0)
      res = expf(-dif2);
    else if (type == 1)
      res = akz_rcp_ge1(1.0f + dif2);
    else if (type == 2)
      res = 1.0f - (float)exp(-3.315 / (dif2 * dif2 * dif2 * dif2));
    else
      res = akz_rcp_ge1(sqrtf(1.0f + dif2));
    sg[idx] = res;
    if (r >= hn && r < RH - hn && c >= hn && c < RW - hn) {
      const size_t o = (size_t)gy * pitch + gx;
      lsm[o] = sC[cy * BW + cx];
      if (gout) gout[o] = res;
    }
  }
  __syncthreads();
  const float* src = fed_lds_steps<RW, RH, NT>(sa, sb, sg, st, x0, y0, w, h);
  for (int idx = hn * RW + tid; idx < (RH - hn) * RW; idx += NT) {
    const int r = idx >> LW, c = idx & (RW - 1);
    const int gx = x0 + c, gy = y0 + r;
    if (c >= hn && c < RW - hn && gx < w && gy < h) out[(size_t)gy * pitch + gx] = src[idx];
  }
}

// FED steps with temporal blocking in registers: one wavefront per region of
// 64 columns (one per lane) x R rows (R registers per lane), with a halo as
// wide as the launch's step count.
// The conductance sums of each pixel pair are formed once per launch: Wh[r]
// = g + g(east) per lane, Wv[r] = g(row r) + g(row r+1); the west and north
// sums are the neighbours' east and south sums (IEEE addition commutes).
// Vertical neighbours are registers, horizontal ones arrive by DPP
// wavefront shifts; no LDS and no barriers.  Per pixel and step the
// arithmetic is k_fed_step's in the same order.  Rows 1..R-2 are updated
// every step; the rows and lanes near the region edge go stale one pixel
// per step, never reaching the output tile.  At image borders the outward
// neighbour is the pixel itself (NLDStep's clamping), which zeroes that
// flux whatever the sum next to it.
// The shifts fill the lane with no source with zero (one instruction, no
// copy of an old value); both kernels that use them read that lane's result
// only in their halo, which never reaches an output.

// The steps of one launch on a register region whose rows 0..R-1 are
// A[OFF..OFF+R-1]; rows 1..R-2 are updated.  Each flux is formed once and
// serves both pixels it joins: a pixel's west flux is its west neighbour's
// east flux, Wh(x-1) * (L(x) - L(x-1)), the same product taken by a DPP
// shift, and its north flux is the south flux of the row above, kept from
// that row (formed from the row's values before this step, like every
// operand of the step).  The update is k_fed_step's expression in its
// order, so the values are those of forming all four fluxes per pixel.  At
// image borders the outward flux is +0, what NLDStep's clamped neighbour
// (the pixel itself) gives for a finite conductance.
// The empty asm makes ytop opaque in every step, so the per-row border tests
// stay scalar compares inside the loop instead of R hoisted 64-bit masks
// (which spill).
template <int R, int OFF, bool BORDER, int NA>
__device__ __forceinline__ void fed_reg_steps(float (&A)[NA], const float (&Wh)[R], const float (&Wv)[R],
                                              const akz_fed_steps& st, int gx, int w, int y0, int h) {
  const bool east_self = BORDER && gx >= w - 1, west_self = BORDER && gx <= 0;
  for (int s = 0; s < st.n; s++) {
    const float stepsize = st.half_tau[s];
    int ytop = y0;
    if (BORDER) asm volatile("" : "+s"(ytop));
    // south flux of row 0: the north flux of row 1
    float yup = Wv[0] * (A[OFF + 1] - A[OFF]);
#pragma unroll
    for (int r = 1; r < R - 1; r++) {
      const float i0 = A[OFF + r];
      float le = from_right(i0), ls = A[OFF + r + 1];
      if (BORDER) {
        if (east_self) le = i0;
        if (ytop + r >= h - 1) ls = i0;
      }
      const float xpos = Wh[r] * (le - i0);
      float xneg = from_left(xpos);
      const float ypos = Wv[r] * (ls - i0);
      float yneg = yup;
      if (BORDER) {
        if (west_self) xneg = 0.0f;
        if (ytop + r <= 0) yneg = 0.0f;
      }
      yup = ypos;
      A[OFF + r] = i0 + stepsize * (xpos - xneg + ypos - yneg);
    }
  }
}

template <int R, bool BORDER>
__device__ __forceinline__ void fed_reg_body(const float* __restrict__ L, const float* __restrict__ g,
                                             float* __restrict__ out, int pitch, int w, int h,
                                             const akz_fed_steps& st, int x0, int y0) {
  const int lane = threadIdx.x;
  const int gx = x0 + lane;
  const int cx = BORDER ? akz_iclamp(gx, 0, w - 1) : gx;
  float Lr[R], Wv[R], Wh[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int gy = BORDER ? akz_iclamp(y0 + r, 0, h - 1) : y0 + r;
    Lr[r] = L[(size_t)gy * pitch + cx];
    Wv[r] = g[(size_t)gy * pitch + cx];  // g for now
  }
#pragma unroll
  for (int r = 0; r < R; r++) Wh[r] = Wv[r] + from_right(Wv[r]);
#pragma unroll
  for (int r = 0; r < R - 1; r++) Wv[r] = Wv[r] + Wv[r + 1];
  fed_reg_steps<R, 0, BORDER>(Lr, Wh, Wv, st, gx, w, y0, h);
  const int hn = st.n;
  if (lane < hn || lane >= 64 - hn || gx >= w) return;
#pragma unroll
  for (int r = 1; r < R - 1; r++) {
    const int gy = y0 + r;
    if (r >= hn && r < R - hn && gy < h) out[(size_t)gy * pitch + gx] = Lr[r];
  }
}

// Resident wavefronts per SIMD asked of the register allocator for 48 rows:
// three (150 VGPRs without spills; left alone it takes 169, two waves).
// 32 rows fit four by themselves, 64 rows only one.
template <int R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(R == 48 ? 3 : 1)))
void k_fed_reg(const float* __restrict__ L, const float* __restrict__ g,
                                                float* __restrict__ out, int pitch, int w, int h, akz_fed_steps st,
                                                size_t zs) {
  const akz_tile3 tl = akz_xcd_tile();
  L = akz_zoff(L, zs, tl.z);
  g = akz_zoff(g, zs, tl.z);
  out = akz_zoff(out, zs, tl.z);
  const int hn = st.n;  // halo: one pixel per step
  const int x0 = tl.x * (64 - 2 * hn) - hn, y0 = tl.y * (R - 2 * hn) - hn;
  // no pixel of the region is on or beyond an image border
  const bool interior = x0 >= 1 && x0 + 64 <= w - 1 && y0 >= 1 && y0 + R <= h - 1;
  if (interior) fed_reg_body<R, false>(L, g, out, pitch, w, h, st, x0, y0);
  else fed_reg_body<R, true>(L, g, out, pitch, w, h, st, x0, y0);
}

// FED steps with temporal blocking on a workgroup of NW wavefronts stacked
// down one region of 64 columns x NW * RPW rows: wavefront w keeps region
// rows w * RPW .. w * RPW + RPW - 1 in registers (lanes are columns, as in
// k_fed_reg), so vertical neighbours are registers and horizontal ones DPP
// shifts; per step only the first and last row of every wavefront go
// through LDS (double-buffered: one barrier per step) to the wavefronts
// above and below.  The conductance pair sums of the region are formed once
// per launch, a wavefront's sums with the rows just outside its strip from
// g itself.  The arithmetic per pixel and step is k_fed_reg's (k_fed_step's
// expression in its order); region rows 0 and NW * RPW - 1 are not updated,
// and the halo is the launch's step count.  A launch of 480x270 takes 60
// workgroups of 8 wavefronts where 32x32 LDS regions take 510 of 16.
// The steps of one launch on wavefront wv's strip, rows A[OFF..OFF+RPW-1]
// (image row ytop for the first), with the pair sums of the strip's rows
// (Wv[r]: rows r, r + 1 of the region, the last with the row below the
// strip) and wup (the pair above the strip).
template <int RPW, int NW, int OFF, bool BORDER, int NA>
__device__ __forceinline__ void fed_strip_steps(float (&A)[NA], const float (&Wh)[RPW], const float (&Wv)[RPW],
                                                float wup, const akz_fed_steps& st, int gx, int w, int ytop0,
                                                int h, float* s_x) {
  const int lane = threadIdx.x & 63, wv = akz_wave();
  const bool east_self = BORDER && gx >= w - 1, west_self = BORDER && gx <= 0;
  const bool first = wv == 0, last = wv == NW - 1;
  for (int s = 0; s < st.n; s++) {
    const float stepsize = st.half_tau[s];
    float* xb = s_x + (s & 1) * (NW * 128);
    xb[wv * 128 + lane] = A[OFF];
    xb[wv * 128 + 64 + lane] = A[OFF + RPW - 1];
    __syncthreads();
    const float above = first ? 0.0f : xb[(wv - 1) * 128 + 64 + lane];
    const float below = last ? 0.0f : xb[(wv + 1) * 128 + lane];
    int ytop = ytop0;
    if (BORDER) asm volatile("" : "+s"(ytop));
    // the north flux of the strip's first row: the south flux of the row
    // above, from the values before this step
    float yup = wup * (A[OFF] - above);
#pragma unroll
    for (int r = 0; r < RPW; r++) {
      const float i0 = A[OFF + r];
      float le = from_right(i0), ls = r + 1 < RPW ? A[OFF + r + 1] : below;
      if (BORDER) {
        if (east_self) le = i0;
        if (ytop + r >= h - 1) ls = i0;
      }
      const float xpos = Wh[r] * (le - i0);
      float xneg = from_left(xpos);
      const float ypos = Wv[r] * (ls - i0);
      float yneg = yup;
      if (BORDER) {
        if (west_self) xneg = 0.0f;
        if (ytop + r <= 0) yneg = 0.0f;
      }
      yup = ypos;
      const bool edge = (r == 0 && first) || (r == RPW - 1 && last);
      if (!edge) A[OFF + r] = i0 + stepsize * (xpos - xneg + ypos - yneg);
    }
  }
}

template <int RPW, int NW, bool BORDER>
__device__ __forceinline__ void fed_strip_body(const float* __restrict__ L, const float* __restrict__ g,
                                               float* __restrict__ out, int pitch, int w, int h,
                                               const akz_fed_steps& st, int x0, int y0, float* s_x) {
  constexpr int RH = RPW * NW;
  const int lane = threadIdx.x & 63, wv = akz_wave();
  const int gx = x0 + lane;
  const int cx = BORDER ? akz_iclamp(gx, 0, w - 1) : gx;
  const int q0 = wv * RPW;  // the wavefront's first region row
  auto row = [&](int q) {   // image row of region row q (clamped at the borders)
    const int gy = y0 + q;
    return (size_t)(BORDER ? akz_iclamp(gy, 0, h - 1) : gy) * pitch + cx;
  };
  float A[RPW], Wv[RPW], Wh[RPW];
#pragma unroll
  for (int r = 0; r < RPW; r++) {
    A[r] = L[row(q0 + r)];
    Wv[r] = g[row(q0 + r)];  // g for now
  }
  // g of the rows just above and below the strip (inside the image for an
  // interior region, clamped otherwise)
  const float gup = g[row(q0 - 1)], gdn = g[row(q0 + RPW)];
#pragma unroll
  for (int r = 0; r < RPW; r++) Wh[r] = Wv[r] + from_right(Wv[r]);
  const float wup = gup + Wv[0];  // the sum of the pair (q0 - 1, q0)
#pragma unroll
  for (int r = 0; r < RPW - 1; r++) Wv[r] = Wv[r] + Wv[r + 1];
  Wv[RPW - 1] = Wv[RPW - 1] + gdn;
  fed_strip_steps<RPW, NW, 0, BORDER>(A, Wh, Wv, wup, st, gx, w, y0 + q0, h, s_x);
  const int hn = st.n;
  if (lane < hn || lane >= 64 - hn || gx >= w) return;
#pragma unroll
  for (int r = 0; r < RPW; r++) {
    const int q = q0 + r, gy = y0 + q;
    if (q >= hn && q < RH - hn && gy < h) out[(size_t)gy * pitch + gx] = A[r];
  }
}

template <int RPW, int NW>
__global__ __launch_bounds__(64 * NW) void k_fed_strip(const float* __restrict__ L, const float* __restrict__ g,
                                                      float* __restrict__ out, int pitch, int w, int h,
                                                      akz_fed_steps st, size_t zs) {
  __shared__ float s_x[2 * NW * 128];
  const akz_tile3 tl = akz_xcd_tile();
  L = akz_zoff(L, zs, tl.z);
  g = akz_zoff(g, zs, tl.z);
  out = akz_zoff(out, zs, tl.z);
  constexpr int RH = RPW * NW;
  const int hn = st.n;  // halo: one pixel per step
  const int x0 = tl.x * (64 - 2 * hn) - hn, y0 = tl.y * (RH - 2 * hn) - hn;
  // no pixel of the region is on or beyond an image border (so the rows
  // just above and below it are inside); uniform over the workgroup
  const bool interior = x0 >= 1 && x0 + 64 <= w - 1 && y0 >= 1 && y0 + RH <= h - 1;
  if (interior) fed_strip_body<RPW, NW, false>(L, g, out, pitch, w, h, st, x0, y0, s_x);
  else fed_strip_body<RPW, NW, true>(L, g, out, pitch, w, h, st, x0, y0, s_x);
}

// Diffusivity of Flow (cuda_akaze.cu:240-287) from dif2 = (Lx^2 + Ly^2) / k^2:
// 0 Perona-Malik g1, 1 Perona-Malik g2, 2 Weickert, 3 Charbonnier.
template <int TYPE>
__device__ __forceinline__ float diffusivity(float dif2) {
  if (TYPE == 0) return expf(-dif2);
  if (TYPE == 1) return akz_rcp_ge1(1.0f + dif2);
  if (TYPE == 2) return 1.0f - (float)exp(-3.315 / (dif2 * dif2 * dif2 * dif2));
  return akz_rcp_ge1(sqrtf(1.0f + dif2));
}

// Lsmooth and conductance of one level in one pass (AKAZE.cpp:260-261): the
// sigma=1 Gaussian of LowPass (5 taps, replicate border, row pass then
// column pass, sums in tap order from 0.0f) fused with Flow (unnormalised
// Scharr of Lsmooth with reflect-101 neighbours, then the diffusivity of
// TYPE), cuda_akaze.cu:56-128, 240-287.  One wavefront per region of 64
// columns, one per lane, whose lanes 3..60 are output columns, and TH output
// rows.  The TH + 6 input rows are loaded at once (coordinates clamped: the
// replicate border), the row pass takes its horizontal taps by DPP
// wavefront shifts, the column pass and the Scharr rows are registers.  At
// the image's first and last row and column the outward Scharr neighbour is
// the one on the other side (reflect-101).
template <int TH, int TYPE, bool BORDER>
__device__ __forceinline__ void smooth_flow_reg_body(const float* __restrict__ in, int pitch, float* __restrict__ lsm,
                                                     float* __restrict__ g, int w, int h, float invk,
                                                     const akz_taps& taps, int x0, int y0) {
  constexpr int NA = TH + 6, NC = TH + 2;
  const int lane = threadIdx.x;
  const int gx = x0 + lane;
  const int cx = BORDER ? akz_iclamp(gx, 0, w - 1) : gx;
  float A[NA];
#pragma unroll
  for (int i = 0; i < NA; i++) {
    const int gy = BORDER ? akz_iclamp(y0 - 3 + i, 0, h - 1) : y0 - 3 + i;
    A[i] = in[(size_t)gy * pitch + cx];
  }
  const float k0 = taps.k[0], k1 = taps.k[1], k2 = taps.k[2], k3 = taps.k[3], k4 = taps.k[4];
  // row pass: B[i] is row y0 - 3 + i, valid on lanes 2..61
  float B[NA];
#pragma unroll
  for (int i = 0; i < NA; i++) {
    const float m1 = from_left(A[i]), p1 = from_right(A[i]);
    const float m2 = from_left(m1), p2 = from_right(p1);
    float sum = 0.0f;
    sum += m2 * k0;
    sum += m1 * k1;
    sum += A[i] * k2;
    sum += p1 * k3;
    sum += p2 * k4;
    B[i] = sum;
  }
  // column pass: C[j] is row y0 - 1 + j; CL / CR hold its Scharr column
  // neighbours x-1 / x+1 (reflected at the image's first and last column)
  float C[NC], CL[NC], CR[NC];
  const bool first_col = BORDER && gx == 0, last_col = BORDER && gx == w - 1;
#pragma unroll
  for (int j = 0; j < NC; j++) {
    float sum = 0.0f;
    sum += B[j] * k0;
    sum += B[j + 1] * k1;
    sum += B[j + 2] * k2;
    sum += B[j + 3] * k3;
    sum += B[j + 4] * k4;
    C[j] = sum;
    const float fl = from_left(sum), fr = from_right(sum);
    CL[j] = first_col ? fr : fl;
    CR[j] = last_col ? fl : fr;
  }
  const bool out_col = lane >= 3 && lane < 61 && gx < w;
#pragma unroll
  for (int r = 0; r < TH; r++) {
    const int y = y0 + r;
    if (BORDER && y >= h) break;
    // rows y-1 and y+1, reflected at the image's first and last row
    const bool top = BORDER && y == 0, bot = BORDER && y == h - 1;
    const int im = top ? r + 2 : r, ip = bot ? r : r + 2;
    const float ul = CL[im], ur = CR[im], ll = CL[ip], lr = CR[ip];
    const float lx = 3.0f * (lr - ll + ur - ul) + 10.0f * (CR[r + 1] - CL[r + 1]);
    const float ly = 3.0f * (lr + ll - ur - ul) + 10.0f * (C[ip] - C[im]);
    const float res = diffusivity<TYPE>(invk * (lx * lx + ly * ly));
    if (out_col) {
      const size_t o = (size_t)y * pitch + gx;
      lsm[o] = C[r + 1];
      g[o] = res;
    }
  }
}

template <int TH, int TYPE>
__global__ __launch_bounds__(64) void k_smooth_flow_reg(const float* __restrict__ in, int pitch,
                                                        float* __restrict__ lsm, float* __restrict__ g, int w, int h,
                                                        const float* kcp, akz_taps taps, size_t zs) {
  const akz_tile3 tl = akz_xcd_tile();
  in = akz_zoff(in, zs, tl.z);
  lsm = akz_zoff(lsm, zs, tl.z);
  g = akz_zoff(g, zs, tl.z);
  kcp = akz_zoff(kcp, zs, tl.z);
  const float kcontrast = *kcp;
  const float invk = 1.0f / (kcontrast * kcontrast);
  const int x0 = tl.x * 58 - 3, y0 = tl.y * TH;
  // every input row and column in the image, no output on its first or
  // last row or column
  const bool interior = x0 >= 0 && x0 + 64 <= w && y0 >= 3 && y0 + TH + 3 <= h;
  if (interior)
    smooth_flow_reg_body<TH, TYPE, false>(in, pitch, lsm, g, w, h, invk, taps, x0, y0);
  else
    smooth_flow_reg_body<TH, TYPE, true>(in, pitch, lsm, g, w, h, invk, taps, x0, y0);
}

// The first FED launch of a level with the register variants, fused with
// the level's smooth + flow: the work of k_smooth_flow_reg and k_fed_reg in
// one pass, with the same arithmetic in the same order.  One wavefront per
// region of 64 columns x R rows: the R + 6 input rows (3 above, 3 below)
// are loaded once and become the FED rows; the row pass, column pass and
// Scharr/diffusivity stream down the region, leaving the conductance of
// every region row in registers, and Lsmooth (and g, when later launches of
// the level read it) is written for the tile.  g is right on lanes 3..60,
// so after n steps lanes 3+n..60-n and rows n..R-1-n are the tile.  Outside
// the image the conductance is computed from clamped input: finite, and
// only ever multiplied by a zero flux difference at the border.
template <int R, int TYPE, bool BORDER>
__device__ __forceinline__ void fed_reg_sf_body(const float* __restrict__ in, const float* __restrict__ parent,
                                                int ppitch, float* __restrict__ lsm, float* __restrict__ gout,
                                                float* __restrict__ out, int pitch, int w, int h, float invk,
                                                const akz_taps& taps, const akz_fed_steps& st, int x0, int y0) {
  constexpr int NA = R + 6;
  const int lane = threadIdx.x;
  const int gx = x0 + lane;
  const int cx = BORDER ? akz_iclamp(gx, 0, w - 1) : gx;
  const int hn = st.n;
  // A[i]: input row y0 - 3 + i; rows 3..R+2 are the FED rows.  With a
  // parent (the previous level, exactly twice the size) the input is its
  // 2x2 box mean at twice the clamped coordinates, k_half2's expression, so
  // the rows equal a load of k_half2's output bit for bit.
  float A[NA];
  if (parent) {
#pragma unroll
    for (int i = 0; i < NA; i++) {
      const int gy = BORDER ? akz_iclamp(y0 - 3 + i, 0, h - 1) : y0 - 3 + i;
      const float* p = parent + (size_t)(2 * gy) * ppitch + 2 * cx;
      A[i] = 0.25f * (p[0] + p[1] + p[ppitch] + p[ppitch + 1]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < NA; i++) {
      const int gy = BORDER ? akz_iclamp(y0 - 3 + i, 0, h - 1) : y0 - 3 + i;
      A[i] = in[(size_t)gy * pitch + cx];
    }
  }
  const float k0 = taps.k[0], k1 = taps.k[1], k2 = taps.k[2], k3 = taps.k[3], k4 = taps.k[4];
  const bool first_col = BORDER && gx == 0, last_col = BORDER && gx == w - 1;
  const bool out_col = lane >= 3 + hn && lane < 61 - hn && gx < w;
  float Wv[R], Wh[R];
  // rings: B rows i-4..i, C / CL / CR rows j-2..j (C[j]: row y0 - 1 + j)
  float B[5], C[3], CL[3], CR[3];
#pragma unroll
  for (int i = 0; i < NA; i++) {
    {
      const float m1 = from_left(A[i]), p1 = from_right(A[i]);
      const float m2 = from_left(m1), p2 = from_right(p1);
      float sum = 0.0f;
      sum += m2 * k0;
      sum += m1 * k1;
      sum += A[i] * k2;
      sum += p1 * k3;
      sum += p2 * k4;
      B[i % 5] = sum;
    }
    const int j = i - 4;
    if (j < 0) continue;
    {
      float sum = 0.0f;
      sum += B[(j + 0) % 5] * k0;
      sum += B[(j + 1) % 5] * k1;
      sum += B[(j + 2) % 5] * k2;
      sum += B[(j + 3) % 5] * k3;
      sum += B[(j + 4) % 5] * k4;
      C[j % 3] = sum;
      const float fl = from_left(sum), fr = from_right(sum);
      CL[j % 3] = first_col ? fr : fl;
      CR[j % 3] = last_col ? fl : fr;
    }
    const int r = j - 2;  // region row y0 + r: C rows r, r+1, r+2 exist
    if (r < 0) continue;
    // opaque per row, so the row tests stay scalar compares here instead
    // of R hoisted masks (which spill)
    int ytop = y0, hk = hn;
    asm volatile("" : "+s"(ytop), "+s"(hk));
    const int y = ytop + r;
    const bool top = BORDER && y == 0, bot = BORDER && y == h - 1;
    const int im = (top ? r + 2 : r) % 3, ip = (bot ? r : r + 2) % 3, ic = (r + 1) % 3;
    const float ul = CL[im], ur = CR[im], ll = CL[ip], lr = CR[ip];
    const float lx = 3.0f * (lr - ll + ur - ul) + 10.0f * (CR[ic] - CL[ic]);
    const float ly = 3.0f * (lr + ll - ur - ul) + 10.0f * (C[ip] - C[im]);
    const float res = diffusivity<TYPE>(invk * (lx * lx + ly * ly));
    Wv[r] = res;  // g for now
    if (out_col && r >= hk && r < R - hk && (!BORDER || y < h)) {
      const size_t o = (size_t)y * pitch + gx;
      lsm[o] = C[ic];
      if (gout) gout[o] = res;
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) Wh[r] = Wv[r] + from_right(Wv[r]);
#pragma unroll
  for (int r = 0; r < R - 1; r++) Wv[r] = Wv[r] + Wv[r + 1];
  fed_reg_steps<R, 3, BORDER>(A, Wh, Wv, st, gx, w, y0, h);
  if (!out_col) return;
#pragma unroll
  for (int r = 1; r < R - 1; r++) {
    const int gy = y0 + r;
    if (r >= hn && r < R - hn && gy < h) out[(size_t)gy * pitch + gx] = A[r + 3];
  }
}

template <int R, int TYPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(R == 48 ? 3 : 1)))
void k_fed_reg_sf(const float* __restrict__ in, const float* __restrict__ parent, int ppitch,
                  float* __restrict__ lsm, float* __restrict__ gout, float* __restrict__ out, int pitch, int w,
                  int h, const float* kcp, akz_taps taps, akz_fed_steps st, size_t zs) {
  const akz_tile3 tl = akz_xcd_tile();
  if (parent) parent = akz_zoff(parent, zs, tl.z);
  else in = akz_zoff(in, zs, tl.z);
  lsm = akz_zoff(lsm, zs, tl.z);
  if (gout) gout = akz_zoff(gout, zs, tl.z);
  out = akz_zoff(out, zs, tl.z);
  kcp = akz_zoff(kcp, zs, tl.z);
  const float kcontrast = *kcp;
  const float invk = 1.0f / (kcontrast * kcontrast);
  const int hn = st.n;
  const int x0 = tl.x * (58 - 2 * hn) - 3 - hn, y0 = tl.y * (R - 2 * hn) - hn;
  // every input pixel inside the image, no Scharr reflection and no region
  // pixel on or beyond an image border
  const bool interior = x0 >= 1 && x0 + 65 <= w && y0 >= 3 && y0 + R + 3 <= h;
  if (interior)
    fed_reg_sf_body<R, TYPE, false>(in, parent, ppitch, lsm, gout, out, pitch, w, h, invk, taps, st, x0, y0);
  else
    fed_reg_sf_body<R, TYPE, true>(in, parent, ppitch, lsm, gout, out, pitch, w, h, invk, taps, st, x0, y0);
}

// The first FED launch of a level on wavefront strips, fused with the
// level's smooth + flow (k_fed_reg_sf's streaming passes, per wavefront):
// wavefront w loads the input rows q0 - 3 .. q0 + RPW + 2 of its strip
// (region rows q0 .. q0 + RPW - 1), streams the row pass, the column pass
// and Scharr/diffusivity down them to the conductance of its own rows, and
// takes the conductance of the rows just above and below the strip from
// the neighbouring wavefronts through LDS (the region's first and last
// rows are not updated, so the outermost wavefronts need none); its input
// rows become the FED rows.  Lsmooth (and g when later launches read it) is
// written for the tile.  g is right on lanes 3..60, so the tile is
// (58 - 2n) x (NW * RPW - 2n).  With a parent (the previous level, exactly
// twice the size) the input is its 2x2 box mean in k_half2's expression.
template <int RPW, int NW, int TYPE, bool BORDER>
__device__ __forceinline__ void fed_strip_sf_body(const float* __restrict__ in, const float* __restrict__ parent,
                                                  int ppitch, float* __restrict__ lsm, float* __restrict__ gout,
                                                  float* __restrict__ out, int pitch, int w, int h, float invk,
                                                  const akz_taps& taps, const akz_fed_steps& st, int x0, int y0,
                                                  float* s_x) {
  constexpr int RH = RPW * NW, NA = RPW + 6;
  const int lane = threadIdx.x & 63, wv = akz_wave();
  const int gx = x0 + lane;
  const int cx = BORDER ? akz_iclamp(gx, 0, w - 1) : gx;
  const int hn = st.n;
  const int q0 = wv * RPW, ys = y0 + q0;  // the strip's first row in the image
  // A[i]: input row ys - 3 + i; A[3..RPW+2] are the FED rows
  float A[NA];
  if (parent) {
#pragma unroll
    for (int i = 0; i < NA; i++) {
      const int gy = BORDER ? akz_iclamp(ys - 3 + i, 0, h - 1) : ys - 3 + i;
      const float* p = parent + (size_t)(2 * gy) * ppitch + 2 * cx;
      A[i] = 0.25f * (p[0] + p[1] + p[ppitch] + p[ppitch + 1]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < NA; i++) {
      const int gy = BORDER ? akz_iclamp(ys - 3 + i, 0, h - 1) : ys - 3 + i;
      A[i] = in[(size_t)gy * pitch + cx];
    }
  }
  const float k0 = taps.k[0], k1 = taps.k[1], k2 = taps.k[2], k3 = taps.k[3], k4 = taps.k[4];
  const bool first_col = BORDER && gx == 0, last_col = BORDER && gx == w - 1;
  const bool out_col = lane >= 3 + hn && lane < 61 - hn && gx < w;
  float Wv[RPW], Wh[RPW];
  // rings: B rows i-4..i, C / CL / CR rows j-2..j (C[j]: row ys - 1 + j)
  float B[5], C[3], CL[3], CR[3];
#pragma unroll
  for (int i = 0; i < NA; i++) {
    {
      const float m1 = from_left(A[i]), p1 = from_right(A[i]);
      const float m2 = from_left(m1), p2 = from_right(p1);
      float sum = 0.0f;
      sum += m2 * k0;
      sum += m1 * k1;
      sum += A[i] * k2;
      sum += p1 * k3;
      sum += p2 * k4;
      B[i % 5] = sum;
    }
    const int j = i - 4;
    if (j < 0) continue;
    {
      float sum = 0.0f;
      sum += B[(j + 0) % 5] * k0;
      sum += B[(j + 1) % 5] * k1;
      sum += B[(j + 2) % 5] * k2;
      sum += B[(j + 3) % 5] * k3;
      sum += B[(j + 4) % 5] * k4;
      C[j % 3] = sum;
      const float fl = from_left(sum), fr = from_right(sum);
      CL[j % 3] = first_col ? fr : fl;
      CR[j % 3] = last_col ? fl : fr;
    }
    const int r = j - 2;  // strip row r: C rows r, r+1, r+2 exist
    if (r < 0) continue;
    int ytop = ys, hk = hn;
    asm volatile("" : "+s"(ytop), "+s"(hk));
    const int y = ytop + r;
    const bool top = BORDER && y == 0, bot = BORDER && y == h - 1;
    const int im = (top ? r + 2 : r) % 3, ip = (bot ? r : r + 2) % 3, ic = (r + 1) % 3;
    const float ul = CL[im], ur = CR[im], ll = CL[ip], lr = CR[ip];
    const float lx = 3.0f * (lr - ll + ur - ul) + 10.0f * (CR[ic] - CL[ic]);
    const float ly = 3.0f * (lr + ll - ur - ul) + 10.0f * (C[ip] - C[im]);
    const float res = diffusivity<TYPE>(invk * (lx * lx + ly * ly));
    Wv[r] = res;  // g for now
    const int q = q0 + r;
    if (out_col && q >= hk && q < RH - hk && (!BORDER || y < h)) {
      const size_t o = (size_t)y * pitch + gx;
      lsm[o] = C[ic];
      if (gout) gout[o] = res;
    }
  }
  // the conductance of the rows next to the strip, from the neighbouring
  // wavefronts (s_x[0, NW * 128); the steps use the rest)
  s_x[wv * 128 + lane] = Wv[0];
  s_x[wv * 128 + 64 + lane] = Wv[RPW - 1];
  __syncthreads();
  const float gup = wv > 0 ? s_x[(wv - 1) * 128 + 64 + lane] : 0.0f;
  const float gdn = wv < NW - 1 ? s_x[(wv + 1) * 128 + lane] : 0.0f;
#pragma unroll
  for (int r = 0; r < RPW; r++) Wh[r] = Wv[r] + from_right(Wv[r]);
  const float wup = gup + Wv[0];
#pragma unroll
  for (int r = 0; r < RPW - 1; r++) Wv[r] = Wv[r] + Wv[r + 1];
  Wv[RPW - 1] = Wv[RPW - 1] + gdn;
  fed_strip_steps<RPW, NW, 3, BORDER>(A, Wh, Wv, wup, st, gx, w, ys, h, s_x + NW * 128);
  if (!out_col) return;
#pragma unroll
  for (int r = 0; r < RPW; r++) {
    const int q = q0 + r, gy = ys + r;
    if (q >= hn && q < RH - hn && gy < h) out[(size_t)gy * pitch + gx] = A[r + 3];
  }
}

template <int RPW, int NW, int TYPE>
__global__ __launch_bounds__(64 * NW) void k_fed_strip_sf(const float* __restrict__ in,
                                                         const float* __restrict__ parent, int ppitch,
                                                         float* __restrict__ lsm, float* __restrict__ gout,
                                                         float* __restrict__ out, int pitch, int w, int h,
                                                         const float* kcp, akz_taps taps, akz_fed_steps st,
                                                         size_t zs) {
  __shared__ float s_x[3 * NW * 128];  // the conductance exchange, then the steps' two buffers
  const akz_tile3 tl = akz_xcd_tile();
  if (parent) parent = akz_zoff(parent, zs, tl.z);
  else in = akz_zoff(in, zs, tl.z);
  lsm = akz_zoff(lsm, zs, tl.z);
  if (gout) gout = akz_zoff(gout, zs, tl.z);
  out = akz_zoff(out, zs, tl.z);
  kcp = akz_zoff(kcp, zs, tl.z);
  const float kcontrast = *kcp;
  const float invk = 1.0f / (kcontrast * kcontrast);
  constexpr int RH = RPW * NW;
  const int hn = st.n;
  const int x0 = tl.x * (58 - 2 * hn) - 3 - hn, y0 = tl.y * (RH - 2 * hn) - hn;
  // every input row (3 above and below the region) inside the image, no
  // Scharr reflection and no region pixel on or beyond an image border
  const bool interior = x0 >= 1 && x0 + 65 <= w && y0 >= 3 && y0 + RH + 3 <= h;
  if (interior)
    fed_strip_sf_body<RPW, NW, TYPE, false>(in, parent, ppitch, lsm, gout, out, pitch, w, h, invk, taps, st, x0, y0,
                                            s_x);
  else
    fed_strip_sf_body<RPW, NW, TYPE, true>(in, parent, ppitch, lsm, gout, out, pitch, w, h, invk, taps, st, x0, y0,
                                           s_x);
}

// One FED step, LDS-tiled: a 64x16 output tile plus a one-pixel halo of L and
// g is read once with coalesced rows; arithmetic as k_fed_step.
__global__ __launch_bounds__(256) void k_fed_tile(const float* __restrict__ L, const float* __restrict__ g,
                                                  float* __restrict__ out, int pitch, int w, int h,
                                                  float stepsize, size_t zs) {
  constexpr int TW = 64, TH = 16, SW = TW + 2, SH = TH + 2;
  L = akz_zoff(L, zs);
  g = akz_zoff(g, zs);
  out = akz_zoff(out, zs);
  __shared__ float sl[SH * SW], sg[SH * SW];
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * TW - 1, y0 = blockIdx.y * TH - 1;
  load_window<SH, SW, 256>(sl, L, pitch, x0, y0, w, h);
  load_window<SH, SW, 256>(sg, g, pitch, x0, y0, w, h);
  __syncthreads();
  const int c = 1 + (tid & 63);
  const int gx = x0 + c;
  if (gx >= w) return;
  const int de = (gx < w - 1) ? 1 : 0, dw = (gx > 0) ? 1 : 0;
  for (int r = 1 + akz_wave(); r <= TH; r += 4) {
    const int gy = y0 + r;
    if (gy >= h) break;
    const int i = r * SW + c;
    const int so = (gy < h - 1) ? i + SW : i, no = (gy > 0) ? i - SW : i;
    const float i0 = sl[i], f0 = sg[i];
    float xpos = (f0 + sg[i + de]) * (sl[i + de] - i0);
    float xneg = (f0 + sg[i - dw]) * (i0 - sl[i - dw]);
    float ypos = (f0 + sg[so]) * (sl[so] - i0);
    float yneg = (f0 + sg[no]) * (i0 - sl[no]);
    out[(size_t)gy * pitch + gx] = i0 + stepsize * (xpos - xneg + ypos - yneg);
  }
}

// 2x2 box mean for exact halving (HalfSample2, cuda_akaze.cu:432-440).
__global__ __launch_bounds__(256) void k_half2(const float* __restrict__ in, int ipitch, float* __restrict__ out,
                                               int ow, int oh, int opitch, size_t zs) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + akz_wave();
  if (x >= ow || y >= oh) return;
  in = akz_zoff(in, zs);
  out = akz_zoff(out, zs);
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
                                                      float* __restrict__ out, int ow, int oh, int opitch,
                                                      size_t zs) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + akz_wave();
  if (x >= ow || y >= oh) return;
  in = akz_zoff(in, zs);
  out = akz_zoff(out, zs);
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
                                                int opitch, int w, int h, size_t zs) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + akz_wave();
  if (x >= w || y >= h) return;
  in = akz_zoff(in, zs);
  out = akz_zoff(out, zs);
  out[(size_t)y * opitch + x] = in[(size_t)y * ipitch + x];
}

inline dim3 grid64x4(int w, int h, int nz) { return dim3((w + 63) / 64, (h + 3) / 4, nz); }

}  // namespace


// k_contrast_front: 64 - 2 max(R0, 3) columns per wavefront (54 at least:
// R0 <= 5) and TH rows: 32 for batches, 16 for one image (more strips to
// fill the GPU; 8 rows repeat more halo rows, which costs more than it gains
// with three images in flight)
constexpr int CFR_TH = 32, CFR_TH1 = 16;
template <int R0, int TH>
static dim3 cfr_grid(int w, int h, int nz) {
  constexpr int TW = 64 - 2 * (R0 > 3 ? R0 : 3);
  return dim3((w + TW - 1) / TW, (h + TH - 1) / TH, nz);
}
int akz_contrast_front_blocks(int w, int h) { return ((w + 53) / 54) * ((h + CFR_TH1 - 1) / CFR_TH1); }

void akz_launch_contrast(hipStream_t s, const float* img, int in_pitch, size_t in_zs, float* lt, float* grad,
                         int pitch, int w, int h, int R0, const akz_taps& taps0, const akz_taps& taps_c, int nbins,
                         float perc, uint32_t* maxbits, uint32_t* blockmax, int32_t* hist_tot, float* kc, int nlev,
                         const int* octaves_host, akz_batch bt) {
  akz_octaves o;
  o.nlev = nlev;
  for (int i = 0; i < nlev; i++) o.oct[i] = octaves_host[i];
  const int nz = bt.nz;
  const size_t zs = bt.zs;
  dim3 fgrid;
#define AKZ_CF(R, TH)                                                                                    \
  fgrid = cfr_grid<R, TH>(w, h, nz);                                                                     \
  hipLaunchKernelGGL((k_contrast_front<R, TH>), fgrid, dim3(64), 0, s, img, in_pitch, in_zs, lt, grad, \
                     pitch, w, h, taps_c, taps0, blockmax, hist_tot, zs)
#define AKZ_CF2(R)    \
  if (nz > 1) {       \
    AKZ_CF(R, CFR_TH);  \
  } else {            \
    AKZ_CF(R, CFR_TH1); \
  }
  switch (R0) {
    case 2: AKZ_CF2(2); break;
    case 3: AKZ_CF2(3); break;
    case 4: AKZ_CF2(4); break;
    default: AKZ_CF2(5); break;
  }
#undef AKZ_CF2
#undef AKZ_CF
  // histogram blocks, which each clear and reduce 19 KB of LDS, are spread
  // over the images of a batch; each wavefront gets a row at least
  const int per = nz > 0 ? nz : 1;
  const int hb = akz_imin((h - 2 + CT_HIST_THREADS / 64 - 1) / (CT_HIST_THREADS / 64),
                          akz_imax(CT_HIST_BLOCKS / per, 32));
  hipLaunchKernelGGL(k_contrast_hist, dim3(hb, 1, nz), dim3(CT_HIST_THREADS), 0, s, grad, pitch, w, h, blockmax,
                     (int)(fgrid.x * fgrid.y), nbins, hist_tot, maxbits, zs);
  hipLaunchKernelGGL(k_contrast_finish, dim3(1, 1, nz), dim3(1024), 0, s, hist_tot, nbins, maxbits, w, h, perc,
                     kc, o, zs);
}

template <int TH>
static void launch_smooth_flow_reg(hipStream_t s, const float* in, int pitch, float* lsm, float* g, int w, int h, int type,
                            const float* kc, const akz_taps& taps, akz_batch bt) {
  const dim3 grid((w + 57) / 58, (h + TH - 1) / TH, bt.nz);
  if (type == 0)
    hipLaunchKernelGGL((k_smooth_flow_reg<TH, 0>), grid, dim3(64), 0, s, in, pitch, lsm, g, w, h, kc, taps, bt.zs);
  else if (type == 1)
    hipLaunchKernelGGL((k_smooth_flow_reg<TH, 1>), grid, dim3(64), 0, s, in, pitch, lsm, g, w, h, kc, taps, bt.zs);
  else if (type == 2)
    hipLaunchKernelGGL((k_smooth_flow_reg<TH, 2>), grid, dim3(64), 0, s, in, pitch, lsm, g, w, h, kc, taps, bt.zs);
  else
    hipLaunchKernelGGL((k_smooth_flow_reg<TH, 3>), grid, dim3(64), 0, s, in, pitch, lsm, g, w, h, kc, taps, bt.zs);
}

void akz_launch_smooth_flow(hipStream_t s, const float* in, int pitch, float* lsm, float* g, int w, int h, int type,
                            const float* kc, const akz_taps& taps, akz_batch bt) {
  // 32-row regions where they still give a few wavefronts per SIMD
  const long waves32 = (long)((w + 57) / 58) * ((h + 31) / 32) * bt.nz;
  if (waves32 >= 4096)
    launch_smooth_flow_reg<32>(s, in, pitch, lsm, g, w, h, type, kc, taps, bt);
  else
    launch_smooth_flow_reg<16>(s, in, pitch, lsm, g, w, h, type, kc, taps, bt);
}

void akz_launch_fed_step(hipStream_t s, const float* Lin, const float* g, float* Lout, int pitch, int w, int h,
                         float half_tau, akz_batch bt) {
  hipLaunchKernelGGL(k_fed_step, grid64x4(w, h, bt.nz), dim3(256), 0, s, Lin, g, Lout, pitch, w, h, half_tau, bt.zs);
}

int akz_fed_block_steps() { return FB_KMAX; }

int akz_fed_variant_steps(int variant) {
  if (variant <= 1) return 1;
  return (variant == 14 || variant == 15 || variant == 19) ? FB_KMAX : FB_K;
}

// Kernel variant per FED launch: 0 k_fed_step, 1 k_fed_tile; blocked regions
// of up to 8 steps (tile = region minus a halo of one pixel per step of the
// launch): 2 64x64/512 threads, 3 32x32/256, 4 64x32/256, 5 32x32/1024,
// 6 64x64/1024, 7 64x48/512, 8 32x48/512, 9 128x64/1024, 10 64x128/1024,
// 11 128x32/512; registers (one wavefront per 64-column region): 12 64 rows,
// 13 48 rows, 16 32 rows; up to 16 steps per launch (half the launches of
// the small, latency-bound levels): 14 64x64/1024, 15 128x64/1024.
void akz_launch_fed_variant(hipStream_t s, int variant, const float* Lin, const float* g, float* Lout, int pitch,
                            int w, int h, const float* half_tau, int n, akz_batch bt) {
  const int nz = bt.nz;
  const size_t zs = bt.zs;
  if (variant <= 1) {
    for (int j = 0; j < n; j++) {
      const float* src = (j == 0) ? Lin : Lout;
      if (variant == 0) {
        hipLaunchKernelGGL(k_fed_step, dim3((w + 63) / 64, (h + 3) / 4, nz), dim3(256), 0, s, src, g, Lout, pitch, w,
                           h, half_tau[j], zs);
      } else {
        hipLaunchKernelGGL(k_fed_tile, dim3((w + 63) / 64, (h + 15) / 16, nz), dim3(256), 0, s, src, g, Lout, pitch,
                           w, h, half_tau[j], zs);
      }
    }
    return;
  }
  akz_fed_steps st;
  // callers split a level's steps into runs of at most akz_fed_variant_steps()
  if (n > akz_fed_variant_steps(variant)) n = akz_fed_variant_steps(variant);
  st.n = n;
  for (int i = 0; i < FB_KMAX; i++) st.half_tau[i] = i < n ? half_tau[i] : 0.0f;
  // region (rw x rh) minus an n-pixel halo on each side is a block's tile
  auto grid = [&](int rw, int rh) {
    const int tw = rw - 2 * n, th = rh - 2 * n;
    return dim3((w + tw - 1) / tw, (h + th - 1) / th, nz);
  };
  if (variant == 2) {
    hipLaunchKernelGGL((k_fed_multi<64, 64, 512>), grid(64, 64), dim3(512), 0, s, Lin, g, Lout, pitch, w, h, st, zs);
  } else if (variant == 3) {
    hipLaunchKernelGGL((k_fed_multi<32, 32, 256>), grid(32, 32), dim3(256), 0, s, Lin, g, Lout, pitch, w, h, st, zs);
  } else if (variant == 5) {
    hipLaunchKernelGGL((k_fed_multi<32, 32, 1024>), grid(32, 32), dim3(1024), 0, s, Lin, g, Lout, pitch, w, h, st,
                       zs);
  } else if (variant == 6) {
    hipLaunchKernelGGL((k_fed_multi<64, 64, 1024>), grid(64, 64), dim3(1024), 0, s, Lin, g, Lout, pitch, w, h, st,
                       zs);
  } else if (variant == 7) {
    hipLaunchKernelGGL((k_fed_multi<64, 48, 512>), grid(64, 48), dim3(512), 0, s, Lin, g, Lout, pitch, w, h, st, zs);
  } else if (variant == 8) {
    hipLaunchKernelGGL((k_fed_multi<32, 48, 512>), grid(32, 48), dim3(512), 0, s, Lin, g, Lout, pitch, w, h, st, zs);
  } else if (variant == 9) {
    hipLaunchKernelGGL((k_fed_multi<128, 64, 1024>), grid(128, 64), dim3(1024), 0, s, Lin, g, Lout, pitch, w, h, st,
                       zs);
  } else if (variant == 10) {
    hipLaunchKernelGGL((k_fed_multi<64, 128, 1024>), grid(64, 128), dim3(1024), 0, s, Lin, g, Lout, pitch, w, h, st,
                       zs);
  } else if (variant == 14) {
    hipLaunchKernelGGL((k_fed_multi<64, 64, 1024, 16>), grid(64, 64), dim3(1024), 0, s, Lin, g, Lout, pitch, w, h, st,
                       zs);
  } else if (variant == 15) {
    hipLaunchKernelGGL((k_fed_multi<128, 64, 1024, 16>), grid(128, 64), dim3(1024), 0, s, Lin, g, Lout, pitch, w, h,
                       st, zs);
  } else if (variant == 12) {
    hipLaunchKernelGGL(k_fed_reg<64>, grid(64, 64), dim3(64), 0, s, Lin, g, Lout, pitch, w, h, st, zs);
  } else if (variant == 13) {
    hipLaunchKernelGGL(k_fed_reg<48>, grid(64, 48), dim3(64), 0, s, Lin, g, Lout, pitch, w, h, st, zs);
  } else if (variant == 16) {
    hipLaunchKernelGGL(k_fed_reg<32>, grid(64, 32), dim3(64), 0, s, Lin, g, Lout, pitch, w, h, st, zs);
  } else if (variant == 11) {
    hipLaunchKernelGGL((k_fed_multi<128, 32, 512>), grid(128, 32), dim3(512), 0, s, Lin, g, Lout, pitch, w, h, st,
                       zs);
  } else if (variant == 17 || variant == 18 || variant == 19) {
    hipLaunchKernelGGL((k_fed_strip<8, 8>), grid(64, 64), dim3(512), 0, s, Lin, g, Lout, pitch, w, h, st, zs);
  } else {
    hipLaunchKernelGGL((k_fed_multi<64, 32, 256>), grid(64, 32), dim3(256), 0, s, Lin, g, Lout, pitch, w, h, st, zs);
  }
}

bool akz_fed_variant_fusable(int variant, int n, int nz) {
  if (variant < 2 || variant > AKZ_FED_VARIANT_MAX || n < 1) return false;
  if (variant != 12 && variant != 13 && variant != 16) return true;
  // a register region computes smooth + flow on all of its 64 x R pixels
  // for a tile of (58 - 2n) x (R - 2n).  With one image per launch the
  // launch saved counts for more than that extra work: always fused
  // (measured +1.8% images/s with three contexts in flight; 960x540 levels
  // of 4-7 steps included).  Batches fuse while the region is at most 1.56
  // times the tile (measured: 1920x1080 with 3-4 steps gains, 960x540 with
  // 4-7 steps on 32 rows loses).
  const int R = variant == 12 ? 64 : variant == 13 ? 48 : 32;
  if (nz == 1) return 2 * n + 2 < R;
  return 64 * R * 100 <= 156 * (58 - 2 * n) * (R - 2 * n);
}

template <int R>
static void launch_fed_reg_sf(hipStream_t s, dim3 grid, const float* in, const float* parent, int ppitch,
                              float* lsm, float* gout, float* Lout, int pitch, int w, int h, int type,
                              const float* kc, const akz_taps& taps, const akz_fed_steps& st, size_t zs) {
#define AKZ_REG_SF(TYPE_)                                                                                        \
  hipLaunchKernelGGL((k_fed_reg_sf<R, TYPE_>), grid, dim3(64), 0, s, in, parent, ppitch, lsm, gout, Lout, pitch, w, \
                     h, kc, taps, st, zs)
  if (type == 0) AKZ_REG_SF(0);
  else if (type == 1) AKZ_REG_SF(1);
  else if (type == 2) AKZ_REG_SF(2);
  else AKZ_REG_SF(3);
#undef AKZ_REG_SF
}

template <int RPW, int NW>
static void launch_fed_strip_sf(hipStream_t s, dim3 grid, const float* in, const float* parent, int ppitch,
                                float* lsm, float* gout, float* Lout, int pitch, int w, int h, int type,
                                const float* kc, const akz_taps& taps, const akz_fed_steps& st, size_t zs) {
#define AKZ_STRIP_SF(TYPE_)                                                                                  \
  hipLaunchKernelGGL((k_fed_strip_sf<RPW, NW, TYPE_>), grid, dim3(64 * NW), 0, s, in, parent, ppitch, lsm, gout, \
                     Lout, pitch, w, h, kc, taps, st, zs)
  if (type == 0) AKZ_STRIP_SF(0);
  else if (type == 1) AKZ_STRIP_SF(1);
  else if (type == 2) AKZ_STRIP_SF(2);
  else AKZ_STRIP_SF(3);
#undef AKZ_STRIP_SF
}

// First launch of a level, fused with the level's smooth + flow
// (k_fed_multi_sf for the LDS variants, k_fed_reg_sf for the register
// ones); g is written only when `gout` is given.
bool akz_fed_variant_halves(int variant) { return variant >= 2 && variant <= AKZ_FED_VARIANT_MAX; }

void akz_launch_fed_variant_sf(hipStream_t s, int variant, const float* in, const float* parent, int ppitch,
                               float* lsm, float* gout, float* Lout, int pitch, int w, int h, int type,
                               const float* kc, const akz_taps& taps, const float* half_tau, int n, akz_batch bt) {
  const int nz = bt.nz;
  const size_t zs = bt.zs;
  akz_fed_steps st;
  if (n > akz_fed_variant_steps(variant)) n = akz_fed_variant_steps(variant);
  st.n = n;
  for (int i = 0; i < FB_KMAX; i++) st.half_tau[i] = i < n ? half_tau[i] : 0.0f;
  auto grid = [&](int rw, int rh) {
    const int tw = rw - 2 * n, th = rh - 2 * n;
    return dim3((w + tw - 1) / tw, (h + th - 1) / th, nz);
  };
#define AKZ_SF(RW_, RH_, NT_, K_)                                                                               \
  hipLaunchKernelGGL((k_fed_multi_sf<RW_, RH_, NT_, K_>), grid(RW_, RH_), dim3(NT_), 0, s, in, parent, ppitch, lsm, \
                     gout, Lout, pitch, w, h, type, kc, taps, st, zs)
  switch (variant) {
    case 2: AKZ_SF(64, 64, 512, FB_K); break;
    case 3: AKZ_SF(32, 32, 256, FB_K); break;
    case 5: AKZ_SF(32, 32, 1024, FB_K); break;
    case 6: AKZ_SF(64, 64, 1024, FB_K); break;
    case 7: AKZ_SF(64, 48, 512, FB_K); break;
    case 8: AKZ_SF(32, 48, 512, FB_K); break;
    case 9: AKZ_SF(128, 64, 1024, FB_K); break;
    case 10: AKZ_SF(64, 128, 1024, FB_K); break;
    case 11: AKZ_SF(128, 32, 512, FB_K); break;
    case 14: AKZ_SF(64, 64, 1024, FB_KMAX); break;
    // the strip variants' first launch: 64x48 LDS regions, or strips
    case 17: AKZ_SF(64, 48, 512, FB_K); break;
    case 18:
    case 19: {
      const dim3 g((w + 57 - 2 * n) / (58 - 2 * n), (h + 63 - 2 * n) / (64 - 2 * n), nz);
      launch_fed_strip_sf<8, 8>(s, g, in, parent, ppitch, lsm, gout, Lout, pitch, w, h, type, kc, taps, st, zs);
      break;
    }
    case 15: AKZ_SF(128, 64, 1024, FB_KMAX); break;
    case 12:
    case 13:
    case 16: {
      // 58 columns carry g; the tile is those minus the halo of n steps
      const int R = variant == 12 ? 64 : variant == 13 ? 48 : 32;
      const dim3 g((w + 57 - 2 * n) / (58 - 2 * n), (h + R - 1 - 2 * n) / (R - 2 * n), nz);
      if (R == 64) launch_fed_reg_sf<64>(s, g, in, parent, ppitch, lsm, gout, Lout, pitch, w, h, type, kc, taps, st, zs);
      else if (R == 48)
        launch_fed_reg_sf<48>(s, g, in, parent, ppitch, lsm, gout, Lout, pitch, w, h, type, kc, taps, st, zs);
      else launch_fed_reg_sf<32>(s, g, in, parent, ppitch, lsm, gout, Lout, pitch, w, h, type, kc, taps, st, zs);
      break;
    }
    default: AKZ_SF(64, 32, 256, FB_K); break;
  }
#undef AKZ_SF
}

// Region shape by level size and images per launch (halo = steps of the
// launch, XCD-aware tile placement), measured on MI355X in the whole
// pipeline.  One 1920x1080 image per call: wavefront strips with the smooth +
// flow in their first launch from 1920x1080 to 480x270 (16 steps per launch
// at 480x270), 32x32/1024t at 240x135.  With three images in flight a
// launch's share of the GPU counts as much as its duration: at 480x270 the
// strips' workgroups of 8 wavefronts leave more of it to the other pipelines
// than 32x32/1024t's 510 of 16, and at 960x540 and 1920x1080 they replace
// single-wavefront register regions of 32 rows whose launches were
// latency-bound (one or two wavefronts per SIMD).  Batches of 16: 48-row registers at 1920x1080,
// 32-row registers at 960x540 and 480x270, 64x64/1024t at 240x135.
int akz_fed_variant_for(int w, int h, int nz) {
  const size_t px = (size_t)w * h;
  if (nz > 1) {
    const size_t tot = px * (size_t)nz;
    if (tot >= 16000000) return 13;
    if (tot >= 1000000) return 16;
    if (tot >= 300000) return 6;
    return 5;
  }
  if (px >= 300000) return 18;
  if (px >= 100000) return 19;
  return 5;
}

void akz_launch_fed_multi(hipStream_t s, const float* Lin, const float* g, float* Lout, int pitch, int w, int h,
                          const float* half_tau, int n, akz_batch bt) {
  akz_launch_fed_variant(s, akz_fed_variant_for(w, h, bt.nz), Lin, g, Lout, pitch, w, h, half_tau, n, bt);
}

namespace {
// 64 consecutive floats per thread from x = 1.0f up; the count is n
__global__ __launch_bounds__(256) void k_rcp_check(unsigned long long* __restrict__ bad, uint32_t n) {
  const uint32_t start = (blockIdx.x * 256u + threadIdx.x) * 64u;
  uint32_t cnt = 0;
  for (uint32_t k = 0; k < 64u && start + k < n; k++) {
    const float x = __uint_as_float(0x3F800000u + start + k);
    if (__float_as_uint(akz_rcp_ge1(x)) != __float_as_uint(1.0f / x)) cnt++;
  }
  if (cnt) atomicAdd(bad, (unsigned long long)cnt);
}
}  // namespace

void akz_launch_rcp_check(hipStream_t s, unsigned long long* mismatches) {
  const uint32_t n = 0x7F800000u - 0x3F800000u + 1u;  // 1.0f .. +inf
  const uint32_t threads = (n + 63u) / 64u;
  hipLaunchKernelGGL(k_rcp_check, dim3((threads + 255u) / 256u), dim3(256), 0, s, mismatches, n);
}

void akz_launch_halfsample(hipStream_t s, const float* in, int iw, int ih, int ipitch, float* out, int ow, int oh,
                           int opitch, akz_batch bt) {
  if (iw == 2 * ow && ih == 2 * oh)
    hipLaunchKernelGGL(k_half2, grid64x4(ow, oh, bt.nz), dim3(256), 0, s, in, ipitch, out, ow, oh, opitch, bt.zs);
  else
    hipLaunchKernelGGL(k_half_general, grid64x4(ow, oh, bt.nz), dim3(256), 0, s, in, iw, ih, ipitch, out, ow, oh,
                       opitch, bt.zs);
}

void akz_launch_copy2d(hipStream_t s, const float* in, int ipitch, float* out, int opitch, int w, int h,
                       akz_batch bt) {
  hipLaunchKernelGGL(k_copy2d, grid64x4(w, h, bt.nz), dim3(256), 0, s, in, ipitch, out, opitch, w, h, bt.zs);
}
