// Learned tile proposal: a small convolutional network over 16x16-pixel tile
// maps of the Hessian-response levels, with BF16 MFMA forward and backward
// kernels (FP32 accumulation), Adam, the tile features, the labels, and the
// candidate mask the detector applies.  No counterpart in the reference.
//
// Network, per tile position of one level map (zero padding at map edges):
//   x0 (4)  = [f(max det of level l-1), f(max det of l), f(max det of l+1),
//              mean Lt of l], f(m) = log2(1 + max(m, 0) / thr) / 8
//   h1 (16) = relu(conv3x3(x0) + b1)
//   h2 (32) = relu(conv3x3(h1) + b2)
//   h3 (32) = relu(conv3x3(h2) + b3)
//   p       = sigmoid(w4 . h3 + b4)
// Scale neighbours in another octave are read at the matching tile: the
// maximum over the 2x2 finer tiles, or the coarser tile holding this one.
//
// Precision: weights, activations and back-propagated gradients are rounded
// to bfloat16 (round to nearest even) where they enter an MFMA; products
// accumulate in FP32; biases, the head, the loss and Adam run in FP32.
//
// MFMA: v_mfma_f32_16x16x32_bf16.  Lane l holds A[l%16][8(l/16)+e] and
// B[8(l/16)+e][l%16] (e = 0..7) and receives D[4(l/16)+e][l%16] (e = 0..3).
// Convolutions are implicit GEMMs: rows are 16 tile positions, K runs over
// (tap, input channel) with the tap major, columns are output channels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "akz_kernels.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int C0 = 4, C1 = 16, C2 = 32, C3 = 32;
constexpr int kpad32(int k) { return (k + 31) / 32 * 32; }
constexpr int kpad16(int k) { return (k + 15) / 16 * 16; }
// flat FP32 parameter layout
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + 9 * C0 * C1, OFF_W2 = OFF_B1 + C1, OFF_B2 = OFF_W2 + 9 * C1 * C2,
              OFF_W3 = OFF_B2 + C2, OFF_B3 = OFF_W3 + 9 * C2 * C3, OFF_W4 = OFF_B3 + C3, OFF_B4 = OFF_W4 + C3,
              NPARAM = OFF_B4 + 1;
static_assert(NPARAM == 14513, "parameter count");

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ f32x4 mfma(uint4 a, uint4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

// Tile position: image, level and tile coordinates of a global index.
struct Pos {
  int gbase;  // global index of the level map's first tile (image included)
  int ty, tx, tw, th;
  bool valid;
};

__device__ __forceinline__ Pos decode(const akz_prop_geom& g, int gp, int total) {
  Pos r;
  r.valid = gp < total;
  if (!r.valid) gp = 0;
  const int z = gp / g.P, p = gp - z * g.P;
  int l = 0;
  while (l + 1 < g.nlev && p >= g.base[l + 1]) l++;
  const int loc = p - g.base[l];
  r.tw = g.tw[l];
  r.th = g.th[l];
  r.ty = loc / r.tw;
  r.tx = loc - r.ty * r.tw;
  r.gbase = z * g.P + g.base[l];
  return r;
}

// global index of the tap-t neighbour (t = 3 (dy + 1) + dx + 1), -1 outside
__device__ __forceinline__ int nbr(const Pos& q, int t) {
  if (!q.valid || t > 8) return -1;
  const int y = q.ty + t / 3 - 1, x = q.tx + t % 3 - 1;
  if (y < 0 || y >= q.th || x < 0 || x >= q.tw) return -1;
  return q.gbase + y * q.tw + x;
}

// 8 consecutive K entries (k0 .. k0+7, k = tap * CIN + ci) of one position's
// im2col row
template <int CIN>
__device__ __forceinline__ uint4 im2col8(const uint16_t* __restrict__ X, const Pos& q, int k0) {
  if (CIN >= 8) {
    const int n = nbr(q, k0 / CIN);
    if (n < 0) return make_uint4(0, 0, 0, 0);
    return *reinterpret_cast<const uint4*>(X + (size_t)n * CIN + (k0 % CIN));
  } else {  // CIN == 4: two taps of four channels
    const int t = k0 / 4;
    const int n0 = nbr(q, t), n1 = nbr(q, t + 1);
    uint2 a = n0 < 0 ? make_uint2(0, 0) : *reinterpret_cast<const uint2*>(X + (size_t)n0 * 4);
    uint2 b = n1 < 0 ? make_uint2(0, 0) : *reinterpret_cast<const uint2*>(X + (size_t)n1 * 4);
    return make_uint4(a.x, a.y, b.x, b.y);
  }
}

// ---- features -------------------------------------------------------------
// Tile statistics per level: maximum of the determinant and mean of Lt over
// each 16x16 tile (pixels inside the level).  One wavefront covers four tiles
// of a tile row (64 columns x 16 rows); each lane loads the 16 rows of its
// column at once.  The sum of a column is ((S0 + S1) + (S2 + S3)) with Sw
// the rows w, w+4, w+8, w+12 summed in order, then a butterfly over the 16
// columns of a tile.  With `keys` the maxima were kept by the fused detector
// (order-preserving keys) and the determinant image is not read.
constexpr int TS_WAVES = 4;
__global__ __launch_bounds__(64 * TS_WAVES) void k_tile_stats(akz_prop_src src, float* __restrict__ tmax,
                                                              float* __restrict__ tmean,
                                                              const uint32_t* __restrict__ keys) {
  const int grp = blockIdx.x * TS_WAVES + akz_wave();
  if (grp >= src.tile_start[src.g.nlev]) return;
  int l = 0;
  while (l + 1 < src.g.nlev && grp >= src.tile_start[l + 1]) l++;
  const int t = grp - src.tile_start[l];
  const int ngx = (src.w[l] + 63) / 64;
  const int gx = t % ngx, ty = t / ngx;
  const int lane = threadIdx.x & 63;
  const int x = gx * 64 + lane;
  const int w = src.w[l], h = src.h[l], pitch = src.pitch[l];
  const float* D = akz_zoff(src.det[l], src.zs);
  const float* Lt = akz_zoff(src.Lt[l], src.zs);
  float m = -INFINITY, sum = 0.0f;
  if (x < w) {
    const int nr = akz_imin(AKZ_PROP_TILE, h - ty * AKZ_PROP_TILE);
    const size_t o0 = (size_t)(ty * AKZ_PROP_TILE) * pitch + x;
    float v[AKZ_PROP_TILE];
#pragma unroll
    for (int r = 0; r < AKZ_PROP_TILE; r++) v[r] = r < nr ? Lt[o0 + (size_t)r * pitch] : 0.0f;
    float sw[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      float a = 0.0f;
#pragma unroll
      for (int r = q; r < AKZ_PROP_TILE; r += 4)
        if (r < nr) a += v[r];
      sw[q] = a;
    }
    sum = (sw[0] + sw[1]) + (sw[2] + sw[3]);
    if (!keys) {
#pragma unroll
      for (int r = 0; r < AKZ_PROP_TILE; r++)
        if (r < nr) m = fmaxf(m, D[o0 + (size_t)r * pitch]);
    }
  }
  for (int off = 1; off < 16; off <<= 1) {
    m = fmaxf(m, __shfl_xor(m, off, 64));
    sum += __shfl_xor(sum, off, 64);
  }
  const int tx = gx * 4 + (lane >> 4);
  if ((lane & 15) == 0 && tx < src.g.tw[l]) {
    const int cw = akz_imin(AKZ_PROP_TILE, w - tx * AKZ_PROP_TILE);
    const int ch = akz_imin(AKZ_PROP_TILE, h - ty * AKZ_PROP_TILE);
    const size_t o = (size_t)blockIdx.z * src.g.P + src.g.base[l] + (size_t)ty * src.g.tw[l] + tx;
    tmax[o] = keys ? akz_det_unkey(keys[o]) : m;
    tmean[o] = sum / (float)(cw * ch);
  }
}

__device__ __forceinline__ float det_feature(float m, float thr) {
  return log2f(1.0f + fmaxf(m, 0.0f) / thr) * 0.125f;
}

// maximum statistic of level n over the area of tile (ty, tx) of level l
__device__ __forceinline__ float scale_nbr_max(const akz_prop_geom& g, const float* tm, int l, int n, int ty,
                                               int tx) {
  const int d = g.oct[l] - g.oct[n];
  if (d == 0) return tm[g.base[n] + ty * g.tw[n] + tx];
  if (d < 0) {  // coarser neighbour
    const int y = akz_imin(ty >> 1, g.th[n] - 1), x = akz_imin(tx >> 1, g.tw[n] - 1);
    return tm[g.base[n] + y * g.tw[n] + x];
  }
  float m = -INFINITY;  // finer neighbour: 2x2 tiles
  for (int dy = 0; dy < 2; dy++)
    for (int dx = 0; dx < 2; dx++) {
      const int y = 2 * ty + dy, x = 2 * tx + dx;
      if (y < g.th[n] && x < g.tw[n]) m = fmaxf(m, tm[g.base[n] + y * g.tw[n] + x]);
    }
  return m;
}

__global__ __launch_bounds__(256) void k_features(akz_prop_geom g, int total, float thr,
                                                  const float* __restrict__ tmax, const float* __restrict__ tmean,
                                                  uint16_t* __restrict__ x0) {
  const int gp = blockIdx.x * 256 + threadIdx.x;
  if (gp >= total) return;
  const int z = gp / g.P, p = gp - z * g.P;
  int l = 0;
  while (l + 1 < g.nlev && p >= g.base[l + 1]) l++;
  const int loc = p - g.base[l], ty = loc / g.tw[l], tx = loc - ty * g.tw[l];
  const float* tm = tmax + (size_t)z * g.P;
  const int lm = l > 0 ? l - 1 : l, lp = l + 1 < g.nlev ? l + 1 : l;
  uint16_t f[4];
  f[0] = f2bf(det_feature(scale_nbr_max(g, tm, l, lm, ty, tx), thr));
  f[1] = f2bf(det_feature(tm[p], thr));
  f[2] = f2bf(det_feature(scale_nbr_max(g, tm, l, lp, ty, tx), thr));
  f[3] = f2bf(tmean[gp]);
  *reinterpret_cast<uint2*>(x0 + (size_t)gp * 4) =
      make_uint2((uint32_t)f[0] | ((uint32_t)f[1] << 16), (uint32_t)f[2] | ((uint32_t)f[3] << 16));
}

// ---- convolutions ---------------------------------------------------------
// Forward: 16 positions per wavefront, all output channels; y = relu(acc + b).
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv_fwd(akz_prop_geom g, int total, const uint16_t* __restrict__ X,
                                                  const uint16_t* __restrict__ WT, const float* __restrict__ bias,
                                                  uint16_t* __restrict__ Y) {
  constexpr int KP = kpad32(9 * CIN), NT = COUT / 16;
  const int lane = threadIdx.x & 63, wv = akz_wave();
  const int p0 = (blockIdx.x * 4 + wv) * 16;
  if (p0 >= total) return;
  const int i = lane & 15, grp = lane >> 4;
  const Pos q = decode(g, p0 + i, total);
  f32x4 acc[NT];
#pragma unroll
  for (int n = 0; n < NT; n++) acc[n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int c = 0; c < KP / 32; c++) {
    const int k0 = c * 32 + grp * 8;
    const uint4 a = im2col8<CIN>(X, q, k0);
#pragma unroll
    for (int n = 0; n < NT; n++) {
      const uint4 b = *reinterpret_cast<const uint4*>(WT + (size_t)(n * 16 + i) * KP + k0);
      acc[n] = mfma(a, b, acc[n]);
    }
  }
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const int co = n * 16 + i;
    const float bv = bias[co];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int p = p0 + grp * 4 + e;
      if (p < total) Y[(size_t)p * COUT + co] = f2bf(fmaxf(acc[n][e] + bv, 0.0f));
    }
  }
}

// Backward through a convolution to its input (transposed convolution):
// dX[p][ci] = sum over taps t and output channels co of
// G[p - offset(t)][co] * W[t][ci][co], masked by the input's relu
// (dX = 0 where X == 0) and rounded to bf16.  COUT = 32: one K chunk per tap.
template <int CIN>
__global__ __launch_bounds__(256) void k_conv_bwd_data(akz_prop_geom g, int total, const uint16_t* __restrict__ G,
                                                       const uint16_t* __restrict__ WB,
                                                       const uint16_t* __restrict__ Xin, uint16_t* __restrict__ Gin) {
  constexpr int COUT = 32, NT = CIN / 16;
  const int lane = threadIdx.x & 63, wv = akz_wave();
  const int p0 = (blockIdx.x * 4 + wv) * 16;
  if (p0 >= total) return;
  const int i = lane & 15, grp = lane >> 4;
  const Pos q = decode(g, p0 + i, total);
  f32x4 acc[NT];
#pragma unroll
  for (int n = 0; n < NT; n++) acc[n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int t = 0; t < 9; t++) {
    const int src = nbr(q, 8 - t);  // the position whose tap-t neighbour is this one
    const uint4 a = src < 0 ? make_uint4(0, 0, 0, 0)
                            : *reinterpret_cast<const uint4*>(G + (size_t)src * COUT + grp * 8);
#pragma unroll
    for (int n = 0; n < NT; n++) {
      const uint4 b = *reinterpret_cast<const uint4*>(WB + (size_t)(n * 16 + i) * (9 * COUT) + t * COUT + grp * 8);
      acc[n] = mfma(a, b, acc[n]);
    }
  }
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const int ci = n * 16 + i;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int p = p0 + grp * 4 + e;
      if (p < total) {
        const float xv = bf2f(Xin[(size_t)p * CIN + ci]);
        Gin[(size_t)p * CIN + ci] = f2bf(xv > 0.0f ? acc[n][e] : 0.0f);
      }
    }
  }
}

// Weight and bias gradients of a convolution: dW[co][k] = sum over positions
// p of G[p][co] * im2col(X)[p][k], as MFMAs over 32-position chunks staged in
// LDS.  Each block reduces WG_CHUNK positions into per-block partials (summed
// in block order by k_grad_reduce, so the result does not depend on
// scheduling).
constexpr int WG_CHUNK = 1024;

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv_bwd_w(akz_prop_geom g, int total, const uint16_t* __restrict__ X,
                                                    const uint16_t* __restrict__ G, float* __restrict__ partW,
                                                    float* __restrict__ partB) {
  constexpr int K = 9 * CIN, KP = kpad16(K), MT = COUT / 16, NTK = KP / 16;
  constexpr int TPW = (NTK + 3) / 4;  // K tiles per wavefront
  constexpr int XS = KP + 8;          // LDS row stride (bf16), breaks bank alignment
  __shared__ uint16_t s_x[32 * XS];
  __shared__ uint16_t s_g[32 * COUT];
  const int lane = threadIdx.x & 63, wv = akz_wave();
  const int i = lane & 15, grp = lane >> 4;
  f32x4 acc[TPW][MT];
#pragma unroll
  for (int a = 0; a < TPW; a++)
#pragma unroll
    for (int m = 0; m < MT; m++) acc[a][m] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
  const int pbeg = blockIdx.x * WG_CHUNK;
  for (int pc = pbeg; pc < pbeg + WG_CHUNK && pc < total; pc += 32) {
    // stage im2col rows of 32 positions (tap-major, CIN channels per tap) and G
    for (int e = threadIdx.x; e < 32 * 9; e += 256) {
      const int r = e / 9, t = e - r * 9;
      const Pos q = decode(g, pc + r, total);
      const int n = nbr(q, t);
      for (int c = 0; c < CIN; c++) s_x[r * XS + t * CIN + c] = n < 0 ? (uint16_t)0 : X[(size_t)n * CIN + c];
    }
    if constexpr (KP > K) {
      for (int e = threadIdx.x; e < 32 * (KP - K); e += 256) {
        const int r = e / (KP - K);
        s_x[r * XS + K + (e - r * (KP - K))] = 0;
      }
    }
    for (int e = threadIdx.x; e < 32 * COUT; e += 256) {
      const int r = e / COUT;
      s_g[e] = (pc + r < total) ? G[(size_t)(pc + r) * COUT + (e - r * COUT)] : (uint16_t)0;
    }
    __syncthreads();
    if (threadIdx.x < COUT)
      for (int r = 0; r < 32; r++) bsum += bf2f(s_g[r * COUT + threadIdx.x]);
    // A[co][p] = G[p][co], B[p][k] = X[p][k]; lane rows/cols as in the header
    uint4 a[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) {
      uint16_t v[8];
#pragma unroll
      for (int e = 0; e < 8; e++) v[e] = s_g[(grp * 8 + e) * COUT + m * 16 + i];
      a[m] = make_uint4(v[0] | ((uint32_t)v[1] << 16), v[2] | ((uint32_t)v[3] << 16), v[4] | ((uint32_t)v[5] << 16),
                        v[6] | ((uint32_t)v[7] << 16));
    }
#pragma unroll
    for (int tt = 0; tt < TPW; tt++) {
      const int nt = wv + 4 * tt;
      if (nt < NTK) {
        uint16_t v[8];
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = s_x[(grp * 8 + e) * XS + nt * 16 + i];
        const uint4 b = make_uint4(v[0] | ((uint32_t)v[1] << 16), v[2] | ((uint32_t)v[3] << 16),
                                   v[4] | ((uint32_t)v[5] << 16), v[6] | ((uint32_t)v[7] << 16));
#pragma unroll
        for (int m = 0; m < MT; m++) acc[tt][m] = mfma(a[m], b, acc[tt][m]);
      }
    }
    __syncthreads();
  }
  float* pw = partW + (size_t)blockIdx.x * COUT * K;
#pragma unroll
  for (int tt = 0; tt < TPW; tt++) {
    const int nt = wv + 4 * tt;
    if (nt >= NTK) continue;
    const int k = nt * 16 + i;
    if (k >= K) continue;
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
      for (int e = 0; e < 4; e++) pw[(size_t)(m * 16 + grp * 4 + e) * K + k] = acc[tt][m][e];
  }
  if (threadIdx.x < COUT) partB[(size_t)blockIdx.x * COUT + threadIdx.x] = bsum;
}

// Sums per-block partials in block order into the flat gradient:
// W[tap][ci][co] (k = tap * CIN + ci) and b[co].
__global__ __launch_bounds__(256) void k_grad_reduce(const float* __restrict__ partW, const float* __restrict__ partB,
                                                     int nblk, int K, int COUT, float* __restrict__ gW,
                                                     float* __restrict__ gB) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < K * COUT) {
    const int co = e / K, k = e - co * K;
    float s = 0.0f;
    for (int b = 0; b < nblk; b++) s += partW[((size_t)b * COUT + co) * K + k];
    gW[(size_t)k * COUT + co] = s;
  } else if (e < K * COUT + COUT) {
    const int co = e - K * COUT;
    float s = 0.0f;
    for (int b = 0; b < nblk; b++) s += partB[(size_t)b * COUT + co];
    gB[co] = s;
  }
}

// ---- head, loss, optimiser --------------------------------------------------
// Head: z = b4 + w4 . h3 (channel order), p = sigmoid(z).  Training adds the
// weighted binary cross-entropy  L = mean(wpos * y * softplus(-z) +
// (1 - y) * softplus(z)),  its gradient to h3 through relu (bf16, G3) and
// per-block partials of dL/dw4, dL/db4 and L.
constexpr int HEAD_T = 256;

template <bool TRAIN>
__global__ __launch_bounds__(HEAD_T) void k_head(int total, const uint16_t* __restrict__ H3,
                                                 const float* __restrict__ params, const uint8_t* __restrict__ labels,
                                                 float wpos, float inv_n, float* __restrict__ prob,
                                                 uint16_t* __restrict__ G3, float* __restrict__ part) {
  __shared__ float s_red[HEAD_T / 64][C3 + 2];
  const float* w4 = params + OFF_W4;
  const int gp = blockIdx.x * HEAD_T + threadIdx.x;
  const bool ok = gp < total;
  float h[C3];
  float z = params[OFF_B4];
  if (ok) {
    const uint4* src = reinterpret_cast<const uint4*>(H3 + (size_t)gp * C3);
#pragma unroll
    for (int v = 0; v < C3 / 8; v++) {
      const uint4 u = src[v];
      const uint32_t ws[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        h[v * 8 + 2 * k] = bf2f((uint16_t)(ws[k] & 0xFFFFu));
        h[v * 8 + 2 * k + 1] = bf2f((uint16_t)(ws[k] >> 16));
      }
    }
#pragma unroll
    for (int c = 0; c < C3; c++) z += w4[c] * h[c];
  } else {
#pragma unroll
    for (int c = 0; c < C3; c++) h[c] = 0.0f;
  }
  const float pz = 1.0f / (1.0f + expf(-z));
  if (ok && prob) prob[gp] = pz;
  if (!TRAIN) return;
  float dz = 0.0f, loss = 0.0f;
  if (ok) {
    const float y = labels[gp] ? 1.0f : 0.0f;
    const float sp_pos = fmaxf(-z, 0.0f) + log1pf(expf(-fabsf(z)));  // softplus(-z)
    const float sp_neg = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));   // softplus(z)
    loss = (wpos * y * sp_pos + (1.0f - y) * sp_neg) * inv_n;
    dz = (wpos * y * (pz - 1.0f) + (1.0f - y) * pz) * inv_n;
    uint16_t gq[C3];
#pragma unroll
    for (int c = 0; c < C3; c++) gq[c] = f2bf(h[c] > 0.0f ? dz * w4[c] : 0.0f);
    uint4* dst = reinterpret_cast<uint4*>(G3 + (size_t)gp * C3);
#pragma unroll
    for (int v = 0; v < C3 / 8; v++)
      dst[v] = make_uint4(gq[v * 8] | ((uint32_t)gq[v * 8 + 1] << 16), gq[v * 8 + 2] | ((uint32_t)gq[v * 8 + 3] << 16),
                          gq[v * 8 + 4] | ((uint32_t)gq[v * 8 + 5] << 16),
                          gq[v * 8 + 6] | ((uint32_t)gq[v * 8 + 7] << 16));
  }
  // block partials: dw4[c] = sum dz * h[c], db4 = sum dz, loss
  const int lane = threadIdx.x & 63, wv = akz_wave();
  for (int c = 0; c < C3 + 2; c++) {
    float v = c < C3 ? dz * h[c] : (c == C3 ? dz : loss);
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) s_red[wv][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < C3 + 2) {
    float v = 0.0f;
    for (int k = 0; k < HEAD_T / 64; k++) v += s_red[k][threadIdx.x];
    part[(size_t)blockIdx.x * (C3 + 2) + threadIdx.x] = v;
  }
}

__global__ __launch_bounds__(64) void k_head_reduce(const float* __restrict__ part, int nblk,
                                                    float* __restrict__ grad, float* __restrict__ loss) {
  const int c = threadIdx.x;
  if (c >= C3 + 2) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; b++) s += part[(size_t)b * (C3 + 2) + c];
  if (c < C3) grad[OFF_W4 + c] = s;
  else if (c == C3) grad[OFF_B4] = s;
  else *loss = s;
}

// Adam (Kingma & Ba, 2015) with bias correction.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                              const float* __restrict__ g, int n, float lr, float b1, float b2,
                                              float eps, float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 weight layouts: forward WT[co][k] (k = tap * CIN + ci, zero padded),
// backward-data WB[ci][tap * COUT + co]
template <int CIN, int COUT>
__device__ __forceinline__ void pack_layer(const float* W, uint16_t* wt, uint16_t* wb, int e) {
  constexpr int K = 9 * CIN, KP = kpad32(K);
  if (e < COUT * KP) {
    const int co = e / KP, k = e - co * KP;
    wt[e] = k < K ? f2bf(W[k * COUT + co]) : (uint16_t)0;
  }
  if (wb && e < CIN * 9 * COUT) {
    const int ci = e / (9 * COUT), r = e - ci * 9 * COUT, t = r / COUT, co = r - t * COUT;
    wb[e] = f2bf(W[(t * CIN + ci) * COUT + co]);
  }
}

__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ params, uint16_t* wt1, uint16_t* wt2,
                                              uint16_t* wt3, uint16_t* wb2, uint16_t* wb3) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  pack_layer<C0, C1>(params + OFF_W1, wt1, nullptr, e);
  pack_layer<C1, C2>(params + OFF_W2, wt2, wb2, e);
  pack_layer<C2, C3>(params + OFF_W3, wt3, wb3, e);
}

// ---- pipeline hooks ---------------------------------------------------------
// detector tile-maximum keys to the key of -infinity (a kernel, not a memset:
// the pipeline is captured in graphs)
__global__ __launch_bounds__(256) void k_fill_keys(uint32_t* __restrict__ key, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) key[i] = akz_det_key(-INFINITY);
}

__global__ __launch_bounds__(256) void k_threshold(int total, const float* __restrict__ prob, float thr,
                                                   uint8_t* __restrict__ mask) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < total) mask[i] = prob[i] >= thr ? 1 : 0;
}

// Keeps the candidate slots of a row segment whose tile is masked in, in
// order (segment = 64 pixels of one row of one level).
__global__ __launch_bounds__(256) void k_mask_segments(akz_det_params p, akz_det_bufs b, akz_prop_geom g,
                                                       const uint8_t* __restrict__ mask) {
  const int seg = blockIdx.x * 256 + threadIdx.x;
  if (seg >= p.seg_total) return;
  int32_t* cnt = akz_zoff(b.seg_count, b.zs);
  const int c = cnt[seg];
  if (c == 0) return;
  int l = 0;
  while (l + 1 < p.nlev && seg >= p.lev[l + 1].seg_base) l++;
  const int ls = seg - p.lev[l].seg_base;
  const int y = ls / p.lev[l].nseg_x, sx = ls - y * p.lev[l].nseg_x;
  const uint8_t* m = mask + (size_t)blockIdx.z * g.P + g.base[l] + (size_t)(y / AKZ_PROP_TILE) * g.tw[l];
  uint8_t* slots = akz_zoff(b.seg_slots, b.zs) + (size_t)seg * 32;
  int k = 0;
  for (int r = 0; r < c; r++) {
    const int s = slots[r];
    if (m[(sx * 64 + s) / AKZ_PROP_TILE]) slots[k++] = (uint8_t)s;
  }
  cnt[seg] = k;
}

__global__ __launch_bounds__(256) void k_labels(akz_det_bufs b, akz_prop_geom g, uint8_t* __restrict__ labels) {
  const int n = akz_zoff(b.status, b.zs)[AKZ_ST_KEPT];
  const akz_kp* cand = akz_zoff(b.cand, b.zs);
  const int32_t* kidx = akz_zoff(b.kidx, b.zs);
  uint8_t* lab = labels + (size_t)blockIdx.z * g.P;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const akz_kp k = cand[kidx[i]];
    const int l = k.class_id;
    const float ratio = (float)(1 << g.oct[l]);
    const int x = (int)(k.x / ratio), y = (int)(k.y / ratio);
    lab[g.base[l] + (y / AKZ_PROP_TILE) * g.tw[l] + x / AKZ_PROP_TILE] = 1;
  }
}

inline int blocks_for(int n, int per) { return (n + per - 1) / per; }

}  // namespace

struct akz_prop_weights {
  float* params = nullptr;  // flat FP32, NPARAM
  uint16_t *wt1 = nullptr, *wt2 = nullptr, *wt3 = nullptr, *wb2 = nullptr, *wb3 = nullptr;
};

akz_prop_geom akz_prop_make_geom(int nlev, const int* w, const int* h, const int* oct) {
  akz_prop_geom g{};
  g.nlev = nlev;
  int P = 0;
  for (int l = 0; l < nlev; l++) {
    g.tw[l] = (w[l] + AKZ_PROP_TILE - 1) / AKZ_PROP_TILE;
    g.th[l] = (h[l] + AKZ_PROP_TILE - 1) / AKZ_PROP_TILE;
    g.oct[l] = oct[l];
    g.base[l] = P;
    P += g.tw[l] * g.th[l];
  }
  g.base[nlev] = P;
  g.P = P;
  return g;
}

void akz_prop_launch_features(hipStream_t s, const akz_prop_src& src, const akz_prop_act& act, int nz,
                              bool det_keys) {
  hipLaunchKernelGGL(k_tile_stats, dim3(blocks_for(src.tile_start[src.g.nlev], TS_WAVES), 1, nz),
                     dim3(64 * TS_WAVES), 0, s, src, act.tmax, act.tmean,
                     det_keys ? (const uint32_t*)act.tkey : (const uint32_t*)nullptr);
  const int total = src.g.P * nz;
  hipLaunchKernelGGL(k_features, dim3(blocks_for(total, 256)), dim3(256), 0, s, src.g, total, src.thr,
                     (const float*)act.tmax, (const float*)act.tmean, act.x0);
}

static void forward(hipStream_t s, const akz_prop_weights* w, const akz_prop_geom& g, int total, const uint16_t* x0,
                    uint16_t* h1, uint16_t* h2, uint16_t* h3) {
  const int nb = blocks_for(total, 64);
  hipLaunchKernelGGL((k_conv_fwd<C0, C1>), dim3(nb), dim3(256), 0, s, g, total, x0, w->wt1, w->params + OFF_B1, h1);
  hipLaunchKernelGGL((k_conv_fwd<C1, C2>), dim3(nb), dim3(256), 0, s, g, total, h1, w->wt2, w->params + OFF_B2, h2);
  hipLaunchKernelGGL((k_conv_fwd<C2, C3>), dim3(nb), dim3(256), 0, s, g, total, h2, w->wt3, w->params + OFF_B3, h3);
}

void akz_prop_launch_infer(hipStream_t s, const akz_prop_weights* w, const akz_prop_geom& g, const akz_prop_act& act,
                           int nz, float thr) {
  const int total = g.P * nz;
  forward(s, w, g, total, act.x0, act.h1, act.h2, act.h3);
  hipLaunchKernelGGL(k_head<false>, dim3(blocks_for(total, HEAD_T)), dim3(HEAD_T), 0, s, total,
                     (const uint16_t*)act.h3, (const float*)w->params, (const uint8_t*)nullptr, 0.0f, 0.0f, act.prob,
                     (uint16_t*)nullptr, (float*)nullptr);
  hipLaunchKernelGGL(k_threshold, dim3(blocks_for(total, 256)), dim3(256), 0, s, total, (const float*)act.prob, thr,
                     act.mask);
}

void akz_prop_launch_clear_keys(hipStream_t s, const akz_prop_geom& g, const akz_prop_act& act, int nz) {
  const int total = g.P * nz;
  hipLaunchKernelGGL(k_fill_keys, dim3(blocks_for(total, 256)), dim3(256), 0, s, act.tkey, total);
}

void akz_prop_launch_mask_segments(hipStream_t s, const akz_det_params& p, const akz_det_bufs& b,
                                   const akz_prop_geom& g, const uint8_t* mask, int nz) {
  hipLaunchKernelGGL(k_mask_segments, dim3(blocks_for(p.seg_total, 256), 1, nz), dim3(256), 0, s, p, b, g, mask);
}

void akz_prop_launch_labels(hipStream_t s, const akz_det_bufs& b, const akz_prop_geom& g, uint8_t* labels, int nz) {
  hipMemsetAsync(labels, 0, (size_t)g.P * nz, s);
  hipLaunchKernelGGL(k_labels, dim3(64, 1, nz), dim3(256), 0, s, b, g, labels);
}

// ---- network object and its C-ABI -------------------------------------------
struct akz_prop_net {
  akz_prop_weights w;
  float *m = nullptr, *v = nullptr, *grad = nullptr, *loss = nullptr;
  int step = 0;
  hipStream_t stream = nullptr;
  // training activations, grown on demand
  size_t cap_pos = 0;
  uint16_t *x0 = nullptr, *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *g1 = nullptr, *g2 = nullptr, *g3 = nullptr;
  uint8_t* lab = nullptr;
  float* prob = nullptr;
  size_t cap_part = 0;
  float *partW = nullptr, *partB = nullptr, *partH = nullptr;
  float* h_loss = nullptr;
};

namespace {

template <class T>
void release(T*& p) {
  if (p) hipFree(p);
  p = nullptr;
}

akz_status grow_act(akz_prop_net* n, size_t npos) {
  if (npos <= n->cap_pos) return AKZ_OK;
  AKZ_HIP_TRY(hipStreamSynchronize(n->stream));
  release(n->x0); release(n->h1); release(n->h2); release(n->h3);
  release(n->g1); release(n->g2); release(n->g3); release(n->lab); release(n->prob);
  release(n->partW); release(n->partB); release(n->partH);
  n->cap_pos = 0;
  AKZ_HIP_TRY(hipMalloc(&n->x0, npos * C0 * 2));
  AKZ_HIP_TRY(hipMalloc(&n->h1, npos * C1 * 2));
  AKZ_HIP_TRY(hipMalloc(&n->h2, npos * C2 * 2));
  AKZ_HIP_TRY(hipMalloc(&n->h3, npos * C3 * 2));
  AKZ_HIP_TRY(hipMalloc(&n->g1, npos * C1 * 2));
  AKZ_HIP_TRY(hipMalloc(&n->g2, npos * C2 * 2));
  AKZ_HIP_TRY(hipMalloc(&n->g3, npos * C3 * 2));
  AKZ_HIP_TRY(hipMalloc(&n->lab, npos));
  AKZ_HIP_TRY(hipMalloc(&n->prob, npos * 4));
  const size_t nblk = (npos + WG_CHUNK - 1) / WG_CHUNK + 1, nhead = (npos + HEAD_T - 1) / HEAD_T + 1;
  AKZ_HIP_TRY(hipMalloc(&n->partW, nblk * 9 * C2 * C3 * 4));
  AKZ_HIP_TRY(hipMalloc(&n->partB, nblk * C3 * 4));
  AKZ_HIP_TRY(hipMalloc(&n->partH, nhead * (C3 + 2) * 4));
  n->cap_pos = npos;
  return AKZ_OK;
}

void pack(akz_prop_net* n) {
  hipLaunchKernelGGL(k_pack, dim3(blocks_for(C3 * kpad32(9 * C2), 256)), dim3(256), 0, n->stream, n->w.params,
                     n->w.wt1, n->w.wt2, n->w.wt3, n->w.wb2, n->w.wb3);
}

// Gradients of the batch loss into n->grad (and the loss into n->loss).
void enqueue_gradients(akz_prop_net* n, const akz_prop_geom& g, int total, float wpos) {
  hipStream_t s = n->stream;
  forward(s, &n->w, g, total, n->x0, n->h1, n->h2, n->h3);
  const int nh = blocks_for(total, HEAD_T);
  hipLaunchKernelGGL(k_head<true>, dim3(nh), dim3(HEAD_T), 0, s, total, (const uint16_t*)n->h3,
                     (const float*)n->w.params, (const uint8_t*)n->lab, wpos, 1.0f / (float)total, n->prob, n->g3,
                     n->partH);
  hipLaunchKernelGGL(k_head_reduce, dim3(1), dim3(64), 0, s, (const float*)n->partH, nh, n->grad, n->loss);
  const int nw = blocks_for(total, WG_CHUNK), nd = blocks_for(total, 64);
  // layer 3
  hipLaunchKernelGGL((k_conv_bwd_w<C2, C3>), dim3(nw), dim3(256), 0, s, g, total, (const uint16_t*)n->h2,
                     (const uint16_t*)n->g3, n->partW, n->partB);
  hipLaunchKernelGGL(k_grad_reduce, dim3(blocks_for(9 * C2 * C3 + C3, 256)), dim3(256), 0, s,
                     (const float*)n->partW, (const float*)n->partB, nw, 9 * C2, C3, n->grad + OFF_W3,
                     n->grad + OFF_B3);
  hipLaunchKernelGGL(k_conv_bwd_data<C2>, dim3(nd), dim3(256), 0, s, g, total, (const uint16_t*)n->g3,
                     (const uint16_t*)n->w.wb3, (const uint16_t*)n->h2, n->g2);
  // layer 2
  hipLaunchKernelGGL((k_conv_bwd_w<C1, C2>), dim3(nw), dim3(256), 0, s, g, total, (const uint16_t*)n->h1,
                     (const uint16_t*)n->g2, n->partW, n->partB);
  hipLaunchKernelGGL(k_grad_reduce, dim3(blocks_for(9 * C1 * C2 + C2, 256)), dim3(256), 0, s,
                     (const float*)n->partW, (const float*)n->partB, nw, 9 * C1, C2, n->grad + OFF_W2,
                     n->grad + OFF_B2);
  hipLaunchKernelGGL(k_conv_bwd_data<C1>, dim3(nd), dim3(256), 0, s, g, total, (const uint16_t*)n->g2,
                     (const uint16_t*)n->w.wb2, (const uint16_t*)n->h1, n->g1);
  // layer 1
  hipLaunchKernelGGL((k_conv_bwd_w<C0, C1>), dim3(nw), dim3(256), 0, s, g, total, (const uint16_t*)n->x0,
                     (const uint16_t*)n->g1, n->partW, n->partB);
  hipLaunchKernelGGL(k_grad_reduce, dim3(blocks_for(9 * C0 * C1 + C1, 256)), dim3(256), 0, s,
                     (const float*)n->partW, (const float*)n->partB, nw, 9 * C0, C1, n->grad + OFF_W1,
                     n->grad + OFF_B1);
}

// copies the features / labels of images idx[0..nimg) of a dataset into the
// network's batch arrays
akz_status gather(akz_prop_net* n, int P, const uint16_t* feat, const uint8_t* labels, const int* idx, int nimg) {
  for (int z = 0; z < nimg; z++) {
    const size_t src = (size_t)(idx ? idx[z] : z) * P;
    AKZ_HIP_TRY(hipMemcpyAsync(n->x0 + (size_t)z * P * C0, feat + src * C0, (size_t)P * C0 * 2,
                               hipMemcpyDeviceToDevice, n->stream));
    if (labels)
      AKZ_HIP_TRY(hipMemcpyAsync(n->lab + (size_t)z * P, labels + src, (size_t)P, hipMemcpyDeviceToDevice,
                                 n->stream));
  }
  return AKZ_OK;
}

}  // namespace

const akz_prop_weights* akz_prop_net_weights(const akz_prop_net* n) { return &n->w; }

extern "C" {

akz_status akz_prop_num_params(int* n) {
  if (!n) return AKZ_ERR_ARG;
  *n = NPARAM;
  return AKZ_OK;
}

akz_status akz_prop_create(uint64_t seed, akz_prop_net** out) {
  if (!out) return AKZ_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { (void)hipGetLastError(); return AKZ_ERR_NO_DEVICE; }
  akz_prop_net* n = new (std::nothrow) akz_prop_net();
  if (!n) return AKZ_ERR_OOM;
  auto fail = [&](akz_status st) { akz_prop_destroy(n); return st; };
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) return fail(AKZ_ERR_HIP);
  if (hipMalloc(&n->w.params, NPARAM * 4) != hipSuccess || hipMalloc(&n->m, NPARAM * 4) != hipSuccess ||
      hipMalloc(&n->v, NPARAM * 4) != hipSuccess || hipMalloc(&n->grad, NPARAM * 4) != hipSuccess ||
      hipMalloc(&n->loss, 4) != hipSuccess || hipMalloc(&n->w.wt1, C1 * kpad32(9 * C0) * 2) != hipSuccess ||
      hipMalloc(&n->w.wt2, C2 * kpad32(9 * C1) * 2) != hipSuccess ||
      hipMalloc(&n->w.wt3, C3 * kpad32(9 * C2) * 2) != hipSuccess ||
      hipMalloc(&n->w.wb2, C1 * 9 * C2 * 2) != hipSuccess || hipMalloc(&n->w.wb3, C2 * 9 * C3 * 2) != hipSuccess ||
      hipHostMalloc((void**)&n->h_loss, 4, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return fail(AKZ_ERR_OOM);
  }
  // He-normal initialisation (fan-in 9 * CIN; head fan-in 32), zero biases,
  // from a seeded SplitMix64 / Box-Muller stream
  std::vector<float> p(NPARAM, 0.0f);
  uint64_t st = seed;
  auto next = [&]() {
    uint64_t z = (st += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  };
  auto normal = [&]() {
    double u1 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0), u2 = (next() >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  auto fill = [&](int off, int cnt, int fan_in) {
    const double sd = std::sqrt(2.0 / fan_in);
    for (int i = 0; i < cnt; i++) p[off + i] = (float)(sd * normal());
  };
  fill(OFF_W1, 9 * C0 * C1, 9 * C0);
  fill(OFF_W2, 9 * C1 * C2, 9 * C1);
  fill(OFF_W3, 9 * C2 * C3, 9 * C2);
  fill(OFF_W4, C3, C3);
  akz_status s2 = akz_prop_set_params(n, p.data());
  if (s2 != AKZ_OK) return fail(s2);
  *out = n;
  return AKZ_OK;
}

akz_status akz_prop_destroy(akz_prop_net* n) {
  if (!n) return AKZ_ERR_ARG;
  if (n->stream) hipStreamSynchronize(n->stream);
  void* ptrs[] = {n->w.params, n->w.wt1, n->w.wt2, n->w.wt3, n->w.wb2, n->w.wb3, n->m, n->v, n->grad, n->loss,
                  n->x0, n->h1, n->h2, n->h3, n->g1, n->g2, n->g3, n->lab, n->prob, n->partW, n->partB, n->partH};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (n->h_loss) hipHostFree(n->h_loss);
  if (n->stream) hipStreamDestroy(n->stream);
  delete n;
  return AKZ_OK;
}

akz_status akz_prop_set_params(akz_prop_net* n, const float* host_params) {
  if (!n || !host_params) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipMemcpyAsync(n->w.params, host_params, NPARAM * 4, hipMemcpyHostToDevice, n->stream));
  AKZ_HIP_TRY(hipMemsetAsync(n->m, 0, NPARAM * 4, n->stream));
  AKZ_HIP_TRY(hipMemsetAsync(n->v, 0, NPARAM * 4, n->stream));
  n->step = 0;
  pack(n);
  AKZ_HIP_TRY(hipGetLastError());
  AKZ_HIP_TRY(hipStreamSynchronize(n->stream));
  return AKZ_OK;
}

akz_status akz_prop_get_params(akz_prop_net* n, float* host_params) {
  if (!n || !host_params) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipStreamSynchronize(n->stream));
  AKZ_HIP_TRY(hipMemcpy(host_params, n->w.params, NPARAM * 4, hipMemcpyDeviceToHost));
  return AKZ_OK;
}

}  // extern "C"

// Entry points that need a context's geometry live beside the context
// (akz_capi.hip); these take the geometry directly.
akz_status akz_prop_forward_geom(akz_prop_net* n, const akz_prop_geom& g, const uint16_t* d_feat, const int* idx,
                                 int nimg, float* h_prob) {
  const size_t npos = (size_t)g.P * nimg;
  akz_status st = grow_act(n, npos);
  if (st != AKZ_OK) return st;
  st = gather(n, g.P, d_feat, nullptr, idx, nimg);
  if (st != AKZ_OK) return st;
  const int total = (int)npos;
  forward(n->stream, &n->w, g, total, n->x0, n->h1, n->h2, n->h3);
  hipLaunchKernelGGL(k_head<false>, dim3(blocks_for(total, HEAD_T)), dim3(HEAD_T), 0, n->stream, total,
                     (const uint16_t*)n->h3, (const float*)n->w.params, (const uint8_t*)nullptr, 0.0f, 0.0f, n->prob,
                     (uint16_t*)nullptr, (float*)nullptr);
  AKZ_HIP_TRY(hipGetLastError());
  if (h_prob) AKZ_HIP_TRY(hipMemcpyAsync(h_prob, n->prob, npos * 4, hipMemcpyDeviceToHost, n->stream));
  AKZ_HIP_TRY(hipStreamSynchronize(n->stream));
  return AKZ_OK;
}

akz_status akz_prop_grad_geom(akz_prop_net* n, const akz_prop_geom& g, const uint16_t* d_feat,
                              const uint8_t* d_labels, const int* idx, int nimg, float wpos, float lr,
                              float* h_grad, float* h_loss) {
  const size_t npos = (size_t)g.P * nimg;
  akz_status st = grow_act(n, npos);
  if (st != AKZ_OK) return st;
  st = gather(n, g.P, d_feat, d_labels, idx, nimg);
  if (st != AKZ_OK) return st;
  enqueue_gradients(n, g, (int)npos, wpos);
  if (lr > 0.0f) {
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - (float)std::pow(b1, n->step), c2 = 1.0f - (float)std::pow(b2, n->step);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(NPARAM, 256)), dim3(256), 0, n->stream, n->w.params, n->m, n->v,
                       (const float*)n->grad, NPARAM, lr, b1, b2, 1e-8f, c1, c2);
    pack(n);
  }
  AKZ_HIP_TRY(hipGetLastError());
  if (h_grad) AKZ_HIP_TRY(hipMemcpyAsync(h_grad, n->grad, NPARAM * 4, hipMemcpyDeviceToHost, n->stream));
  if (h_loss) AKZ_HIP_TRY(hipMemcpyAsync(n->h_loss, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
  AKZ_HIP_TRY(hipStreamSynchronize(n->stream));
  if (h_loss) *h_loss = *n->h_loss;
  return AKZ_OK;
}
