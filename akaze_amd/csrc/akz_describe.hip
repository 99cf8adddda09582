// Orientation and MLDB descriptor kernels: one wavefront per keypoint.
//
// Both kernels are persistent-style (grid-stride over the keypoint count read
// from device memory), so they are launched with a fixed grid and need no host
// knowledge of how many keypoints the detector kept.
#include <algorithm>

#include "akz_kernels.h"

namespace {

// The samples of FindOrientation's radius-6 disc (cuda_akaze.cu:2114-2167),
// offsets (i, j) with i^2 + j^2 < 36, in the order of their slot tx = 16 (j +
// 6) + (i + 6) in the reference's 16-wide sample grid: packed as (i + 6) |
// (j + 6) << 4 | (i^2 + j^2) << 8.
constexpr int OR_NS = 109;
struct OrientPts {
  int16_t v[OR_NS];
};
constexpr OrientPts make_orient_pts() {
  OrientPts t{};
  int c = 0;
  for (int tx = 0; tx < 208; tx++) {
    const int i = (tx & 15) - 6, j = (tx >> 4) - 6, r2 = i * i + j * j;
    if (r2 < 36) t.v[c++] = (int16_t)((i + 6) | ((j + 6) << 4) | (r2 << 8));
  }
  return t;
}
static_assert(make_orient_pts().v[OR_NS - 1] != 0, "109 disc samples");
__constant__ OrientPts d_orpts = make_orient_pts();

// exp(-r2 / 12.5) for r2 = 0..35 (cuda_akaze.cu:2133), rounded once on the host
struct OrientWeights { float w[36]; };

constexpr int OR_WAVES = 4;

// Dominant orientation (FindOrientation, cuda_akaze.cu:2114-2167).  The 109
// samples of the radius-6 disc go to the lanes in the reference's sample
// order (two per lane); their bin and weighted derivatives go to LDS, then
// lane b (b < 42) sums bin b in sample order, so the float sums do not depend
// on scheduling.  The strongest window is found by a wavefront maximum: the
// first window reaching it, as the reference's sequential scan keeps.
// Keypoint kp of image z (kps: that image's keypoints), one wavefront per
// keypoint: every wavefront of the block calls it once per round (active or
// not), so the block's barriers are reached by all of them.  The wavefronts
// of one block may work on different images (p is the wavefront's).
__device__ __forceinline__ void orient_one(const akz_desc_params& p, akz_kp* __restrict__ kps, int kp, bool active,
                                           const OrientWeights& wt, int z) {
  __shared__ int8_t s_bin[OR_WAVES][OR_NS + 3];
  __shared__ float s_dx[OR_WAVES][OR_NS + 3], s_dy[OR_WAVES][OR_NS + 3];
  __shared__ float s_rx[OR_WAVES][42], s_ry[OR_WAVES][42];
  __shared__ float s_ex[OR_WAVES][42], s_ey[OR_WAVES][42];
  const int lane = threadIdx.x & 63, wv = akz_wave();
  {
    if (active) {
      akz_kp k = kps[kp];
      const int lev = k.class_id;
      const float2* lxy = akz_zoff(p.Lxy[lev], p.zs, z);
      const int pitch = p.pitch[lev], w = p.w[lev], h = p.h[lev];
      const int octave = k.octave;
      const int step = (int)(0.5f * k.size + 0.5f) >> octave;
      const int x = (int)(k.x + 0.5f) >> octave;
      const int y = (int)(k.y + 0.5f) >> octave;
      // the lane's 2 samples: both gathers are issued before either is used
      float gxv[2], gyv[2], gwv[2];
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const int sn = lane + 64 * q;
        gwv[q] = gxv[q] = gyv[q] = 0.0f;
        if (sn < OR_NS) {
          const int pt = d_orpts.v[sn];
          const int i = (pt & 15) - 6, j = ((pt >> 4) & 15) - 6;
          gwv[q] = wt.w[pt >> 8];
          const int xs = akz_iclamp(x + step * i, 0, w - 1), ys = akz_iclamp(y + step * j, 0, h - 1);
          const float2 d = lxy[(size_t)ys * pitch + xs];
          gxv[q] = d.x;
          gyv[q] = d.y;
        }
      }
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const int sn = lane + 64 * q;
        if (sn >= OR_NS) break;
        const float dx = gwv[q] * gxv[q];
        const float dy = gwv[q] * gyv[q];
        const float angle = akz_atan2f(dy, dx);
        s_bin[wv][sn] = (int8_t)akz_imax(akz_imin((int)(angle * (21 / 3.14159265358979323846)) + 21, 41), 0);
        s_dx[wv][sn] = dx;
        s_dy[wv][sn] = dy;
      }
    }
    __syncthreads();
    // lane b gets the members of bin b as two 64-bit masks (samples 0..63
    // and 64..108, one ballot per bin and half), then adds them in sample
    // order: the bins' few members instead of a scan of all 109 samples
    uint64_t m0 = 0ull, m1 = 0ull;
    {
      const int b0 = active ? (int)s_bin[wv][lane] : -1;
      const int b1 = active && lane + 64 < OR_NS ? (int)s_bin[wv][lane + 64] : -1;
      for (int b = 0; b < 42; b++) {
        const uint64_t a0 = __ballot(b0 == b), a1 = __ballot(b1 == b);
        if (lane == b) {
          m0 = a0;
          m1 = a1;
        }
      }
    }
    if (active && lane < 42) {
      float rx = 0.0f, ry = 0.0f;
      while (m0) {
        const int sn = __ffsll((unsigned long long)m0) - 1;
        m0 &= m0 - 1ull;
        rx += s_dx[wv][sn];
        ry += s_dy[wv][sn];
      }
      while (m1) {
        const int sn = 64 + __ffsll((unsigned long long)m1) - 1;
        m1 &= m1 - 1ull;
        rx += s_dx[wv][sn];
        ry += s_dy[wv][sn];
      }
      s_rx[wv][lane] = rx;
      s_ry[wv][lane] = ry;
    }
    __syncthreads();
    float r = -1.0f;
    if (active && lane < 42) {
      float ex = s_rx[wv][lane], ey = s_ry[wv][lane];
      for (int q = lane + 1; q < lane + 7; q++) {
        ex += s_rx[wv][q < 42 ? q : q - 42];
        ey += s_ry[wv][q < 42 ? q : q - 42];
      }
      s_ex[wv][lane] = ex;
      s_ey[wv][lane] = ey;
      r = ex * ex + ey * ey;
    }
    // the reference keeps the first window whose r exceeds the running
    // maximum (from 0): the lowest window with the largest r, or window 0
    // when no r is positive
    float m = r;
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    const uint64_t at = __ballot(r == m && m > 0.0f);
    const int maxk = at ? __ffsll((unsigned long long)at) - 1 : 0;
    __syncthreads();
    if (active && lane == 0) {
      float angle = akz_atan2f(s_ey[wv][maxk], s_ex[wv][maxk]);
      kps[kp].angle = (angle < 0.0f ? angle + 2.0f * 3.14159265358979323846 : angle);
    }
    __syncthreads();
  }
}

// The keypoints of image z (grid stride over blocks bx of nb); every thread of
// the block calls it with the same arguments.
__device__ __forceinline__ void orient_image(const akz_desc_params& p, akz_kp* __restrict__ kps,
                                             const int32_t* nkps_p, const OrientWeights& wt, int z, int bx, int nb) {
  const int wv = akz_wave();
  kps = akz_zoff(kps, p.zs, z);
  const int n = *akz_zoff(nkps_p, p.zs, z);
  // the loop bound is uniform over the block so every barrier is reached by
  // all waves.  Consecutive keypoints (overlapping windows) stay on one XCD
  // and its L2: the blocks of XCD bx % 8 take the XCD's eighth of the
  // keypoints, OR_WAVES at a time with the stride of its blocks.
  const int nx = (nb & 7) == 0 ? 8 : 1, xcd = bx % nx;
  const int lo = (int)((long long)n * xcd / nx), hi = (int)((long long)n * (xcd + 1) / nx);
  for (int base = lo + (bx / nx) * OR_WAVES; base < hi; base += (nb / nx) * OR_WAVES) {
    const int kp = base + wv;
    orient_one(p, kps, kp, kp < hi, wt, z);
  }
}

__global__ __launch_bounds__(64 * OR_WAVES) void k_orientation(akz_desc_params p, akz_kp* __restrict__ kps,
                                                               const int32_t* nkps_p, OrientWeights wt) {
  orient_image(p, kps, nkps_p, wt, blockIdx.z, blockIdx.x, gridDim.x);
}

// Persistent form over a ragged batch: the keypoints of all images of the
// work list form one flat list (image i's after the counts of images 0..i-1,
// summed per block into LDS), walked with the grid stride.
__device__ __forceinline__ int list_counts(const akz_desc_worklist& wl, int32_t* s_start) {
  if (threadIdx.x < 64) {
    const int i = threadIdx.x;
    int v = 0;
    if (i < wl.nimg) v = *akz_zoff(wl.nkps[wl.cls[i]], wl.P[wl.cls[i]].zs, wl.z[i]);
    int inc = v;
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(inc, d, 64);
      if (i >= d) inc += o;
    }
    s_start[i] = inc - v;
    if (i == 63) s_start[64] = inc;
  }
  __syncthreads();
  return s_start[64];
}

__device__ __forceinline__ int list_image(const int32_t* s_start, int nimg, int item) {
  int a = 0, e = nimg;
  while (e - a > 1) {
    const int m = (a + e) >> 1;
    if (s_start[m] <= item) a = m; else e = m;
  }
  return a;
}

__global__ __launch_bounds__(64 * OR_WAVES) void k_orientation_list(akz_desc_worklist wl, OrientWeights wt) {
  __shared__ int32_t s_start[65];
  const int total = list_counts(wl, s_start);
  const int wv = akz_wave();
  // one wavefront per keypoint: wavefront wv of a block takes item base + wv
  for (int base = blockIdx.x * OR_WAVES; base < total; base += gridDim.x * OR_WAVES) {
    const int item = base + wv;
    const bool active = item < total;
    int i = 0, kp = 0;
    if (active) {
      i = __builtin_amdgcn_readfirstlane(list_image(s_start, wl.nimg, item));
      kp = item - s_start[i];
    }
    const int c = wl.cls[i];
    orient_one(wl.P[c], akz_zoff(wl.kps[c], wl.P[c].zs, wl.z[i]), kp, active, wt, wl.z[i]);
  }
}

// MLDB descriptor (ExtractDescriptors + BuildDescriptor, cuda_akaze.cu:
// 1420-1533, 1786-1807).  Lane l accumulates the samples i = l, l+64, ... of
// the rotated sampling grid into its own partial sums of the 87 values (29
// cells x 3 channels, LDS, value-major); the 64 partials of each value are
// then combined as the reference does: p[t] + p[t+32], then a pairwise tree
// over 32.  A value's 64 partials sit at a stride of 65 floats: the lanes
// writing them hit consecutive words, the lanes reducing them (one value
// each) hit different banks.
constexpr int MLDB_WAVES = 1;
constexpr int MLDB_VS = 65;

// Cells of sample i in the 2x2, 3x3 and 4x4 grids (-1: outside that grid).
struct mldb_cells { int c2, c3, c4; };
__device__ inline mldb_cells mldb_cell(int i, int size2, int size3, int size4, int winsize) {
  const int y = i / winsize, x = i - winsize * y;
  const int mm = x > y ? x : y;
  mldb_cells c{-1, -1, -1};
  if (mm < 2 * size2) c.c2 = (y < size2 ? 0 : 1) * 2 + (x < size2 ? 0 : 1);
  if (mm < 3 * size3) {
    int x3 = (x < size3 ? 0 : (x < 2 * size3 ? 1 : 2));
    int y3 = (y < size3 ? 0 : (y < 2 * size3 ? 1 : 2));
    c.c3 = 4 + y3 * 3 + x3;
  }
  if (mm < 4 * size4) {
    int x4 = (x < 2 * size4 ? (x < size4 ? 0 : 1) : (x < 3 * size4 ? 2 : 3));
    int y4 = (y < 2 * size4 ? (y < size4 ? 0 : 1) : (y < 3 * size4 ? 2 : 3));
    c.c4 = 13 + y4 * 4 + x4;
  }
  return c;
}

// The descriptors of image z (grid stride over blocks bx of nb).
// Items bx, bx + nb, ... < n: item_of(it) gives the item's image parameters,
// keypoints and 64-byte descriptor rows (of its own image) and keypoint
// index.  Pattern size and comparison list are p0's (the same for every
// image of a call).
struct mldb_item {
  const akz_desc_params* p;
  const akz_kp* kps;
  uint8_t* desc;
  int kp, z;
};
template <class ItemOf>
__device__ __forceinline__ void mldb_run(const akz_desc_params& p0, int n, int bx, int nb, ItemOf item_of) {
  // 29 cells x 3 channels, plus cell 29: where a sample outside a grid's
  // cells adds (never read)
  __shared__ float acc[90 * MLDB_VS];
  __shared__ float vals[88];
  const int lane = threadIdx.x & 63;
  const int size2 = p0.patsize;
  const int size3 = (int)ceilf(2.0f * p0.patsize / 3.0f);
  const int size4 = (int)ceilf(0.5f * p0.patsize);
  const int winsize = akz_imax(3 * size3, 4 * size4);
  constexpr int MS = 7;  // samples per lane (m < 7 covers grids up to 21x21)
  // comparison pairs of the lane's output byte (the variant's list, entries
  // past its last bit are (0, 0) and compare false; lanes past byte 60 read
  // the (0, 0) entry 487): loaded once, the same for every keypoint
  int ia[8], ib[8];
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const int c = lane < 61 ? lane * 8 + b : AKZ_MLDB_TABLE - 1;
    ia[b] = p0.cmp[c];
    ib[b] = p0.cmp[AKZ_MLDB_TABLE + c];
  }
  // first accumulator row (3 x cell) of the lane's samples in each grid: the
  // same for every keypoint
  int rb[3][MS];
#pragma unroll
  for (int m = 0; m < MS; m++) {
    const int i = lane + 64 * m;
    mldb_cells cc{-1, -1, -1};
    if (i < winsize * winsize) cc = mldb_cell(i, size2, size3, size4, winsize);
    rb[0][m] = 3 * (cc.c2 < 0 ? 29 : cc.c2);
    rb[1][m] = 3 * (cc.c3 < 0 ? 29 : cc.c3);
    rb[2][m] = 3 * (cc.c4 < 0 ? 29 : cc.c4);
  }
  // consecutive keypoints (sorted by level and position, so their sampling
  // windows overlap) stay on one XCD and its L2: the blocks of XCD bx % 8
  // take the XCD's eighth of the items with the stride of its blocks
  const int nx = (nb & 7) == 0 ? 8 : 1, xcd = bx % nx;
  const int lo = (int)((long long)n * xcd / nx), hi = (int)((long long)n * (xcd + 1) / nx);
  for (int it = lo + bx / nx; it < hi; it += nb / nx) {
    const mldb_item itm = item_of(it);
    const akz_desc_params& p = *itm.p;
    const int kp = itm.kp, z = itm.z;
    uint8_t* __restrict__ desc = itm.desc;
    const akz_kp k = itm.kps[kp];
    const int lev = k.class_id;
    const float* imd = akz_zoff(p.Lt[lev], p.zs, z);
    const float2* lxy = akz_zoff(p.Lxy[lev], p.zs, z);
    const int pitch = p.pitch[lev], w = p.w[lev], h = p.h[lev];
    const float iratio = 1.0f / (1 << k.octave);
    const int scale = (int)(0.5f * k.size * iratio + 0.5f);
    const float xf = k.x * iratio;
    const float yf = k.y * iratio;
    float co, si;
    akz_sincosf(k.angle, &si, &co);
    for (int v = 0; v < 87; v++) acc[v * MLDB_VS + lane] = 0.0f;
    // all gathers are issued first, then accumulated in increasing i
    float ch[3][MS];
#pragma unroll
    for (int m = 0; m < MS; m++) {
      const int i = lane + 64 * m;
      ch[0][m] = ch[1][m] = ch[2][m] = 0.0f;
      if (i < winsize * winsize) {
        const int y = i / winsize, x = i - winsize * y;
        const int l = x - size2, kk = y - size2;
        int xp = (int)(xf + scale * (kk * co - l * si) + 0.5f);
        int yp = (int)(yf + scale * (kk * si + l * co) + 0.5f);
        xp = akz_iclamp(xp, 0, w - 1);
        yp = akz_iclamp(yp, 0, h - 1);
        const size_t pos = (size_t)yp * pitch + xp;
        const float2 d = lxy[pos];
        ch[0][m] = imd[pos];
        ch[1][m] = d.x;
        ch[2][m] = d.y;
      }
    }
    // channels: L, rx = -dx sin + dy cos, ry = dx cos + dy sin; with two
    // channels L and the gradient magnitude sqrt(dx^2 + dy^2)
    if (p.nch == 2) {
#pragma unroll
      for (int m = 0; m < MS; m++) {
        const float dx = ch[1][m], dy = ch[2][m];
        ch[1][m] = sqrtf(dx * dx + dy * dy);
        ch[2][m] = 0.0f;
      }
    } else {
#pragma unroll
      for (int m = 0; m < MS; m++) {
        const float dx = ch[1][m], dy = ch[2][m];
        ch[1][m] = -dx * si + dy * co;
        ch[2][m] = dx * co + dy * si;
      }
    }
    // samples in increasing i.  The 2x2, 3x3 and 4x4 cells of one sample
    // are different accumulators (disjoint cell ranges), so the nine of a
    // sample are read together, then written: one LDS round trip per sample
    // instead of three
#pragma unroll
    for (int m = 0; m < MS; m++) {
      float r[3][3];
#pragma unroll
      for (int g = 0; g < 3; g++)
#pragma unroll
        for (int c = 0; c < 3; c++) r[g][c] = acc[(rb[g][m] + c) * MLDB_VS + lane];
#pragma unroll
      for (int g = 0; g < 3; g++)
#pragma unroll
        for (int c = 0; c < 3; c++) acc[(rb[g][m] + c) * MLDB_VS + lane] = r[g][c] + ch[c][m];
    }
    __syncthreads();
    for (int v = lane; v < 87; v += 64) {
      float a[32];
#pragma unroll
      for (int t = 0; t < 32; t++) a[t] = acc[v * MLDB_VS + t] + acc[v * MLDB_VS + t + 32];
#pragma unroll
      for (int s = 1; s < 32; s <<= 1)
#pragma unroll
        for (int t = 0; t + s < 32; t += 2 * s) a[t] = a[t] + a[t + s];
      vals[v] = a[0];
    }
    __syncthreads();
    uint32_t byte = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) byte |= (vals[ia[b]] > vals[ib[b]] ? 1u : 0u) << b;
    if (lane >= 61) byte = 0;
    desc[(size_t)kp * AKZ_DESC_STRIDE + lane] = (uint8_t)byte;
    __syncthreads();
  }
}

__global__ __launch_bounds__(64 * MLDB_WAVES) void k_mldb(akz_desc_params p, const akz_kp* __restrict__ kps,
                                                          const int32_t* nkps_p, uint8_t* __restrict__ desc) {
  const int z = blockIdx.z;
  const akz_kp* kz = akz_zoff(kps, p.zs, z);
  uint8_t* dz = akz_zoff(desc, p.zs, z);
  const int n = *akz_zoff(nkps_p, p.zs, z);
  mldb_run(p, n, blockIdx.x, gridDim.x, [&](int it) { return mldb_item{&p, kz, dz, it, z}; });
}

// Persistent form over a ragged batch: one flat list of the keypoints of all
// images (see k_orientation_list).
__global__ __launch_bounds__(64 * MLDB_WAVES) void k_mldb_list(akz_desc_worklist wl) {
  __shared__ int32_t s_start[65];
  const int total = list_counts(wl, s_start);
  mldb_run(wl.P[wl.cls[0]], total, blockIdx.x, gridDim.x, [&](int it) {
    const int i = __builtin_amdgcn_readfirstlane(list_image(s_start, wl.nimg, it));
    const int c = __builtin_amdgcn_readfirstlane(wl.cls[i]);
    const akz_desc_params* p = &wl.P[c];
    return mldb_item{p, akz_zoff(wl.kps[c], p->zs, wl.z[i]), akz_zoff(wl.desc[c], p->zs, wl.z[i]),
                     it - s_start[i], wl.z[i]};
  });
}

OrientWeights g_weights;
bool g_weights_ready = false;

}  // namespace

void akz_describe_init_tables() {
  if (g_weights_ready) return;
  for (int r2 = 0; r2 < 36; r2++)
    g_weights.w[r2] = (float)exp((double)((float)(-r2) / (2.5f * 2.5f * 2.0f)));
  g_weights_ready = true;
}

// Blocks per image shrink as the batch grows (the kernels stride over the
// keypoints), so a batch launch stays near 4096-8192 resident blocks.
static int per_image_blocks(int want, int nz, int floor_blocks) {
  int b = want / (nz > 0 ? nz : 1);
  if (b < floor_blocks) b = floor_blocks;
  return b;
}

void akz_launch_orientation(hipStream_t s, const akz_desc_params& p, akz_kp* kps, const int32_t* nkps, int cap,
                            int nz) {
  int blocks = (cap + OR_WAVES - 1) / OR_WAVES;
  blocks = std::min(std::max(blocks, 1), per_image_blocks(2048, nz, 256));
  hipLaunchKernelGGL(k_orientation, dim3(blocks, 1, nz), dim3(64 * OR_WAVES), 0, s, p, kps, nkps, g_weights);
}

void akz_launch_describe_list(hipStream_t s, const akz_desc_worklist& wl, bool orientation, int or_blocks,
                              int mldb_blocks) {
  if (orientation)
    hipLaunchKernelGGL(k_orientation_list, dim3(or_blocks), dim3(64 * OR_WAVES), 0, s, wl, g_weights);
  hipLaunchKernelGGL(k_mldb_list, dim3(mldb_blocks), dim3(64 * MLDB_WAVES), 0, s, wl);
}

void akz_launch_mldb(hipStream_t s, const akz_desc_params& p, const akz_kp* kps, const int32_t* nkps, uint8_t* desc64,
                     int cap, int nz) {
  int blocks = std::min(std::max(cap, 1), per_image_blocks(4096, nz, 512));
  hipLaunchKernelGGL(k_mldb, dim3(blocks, 1, nz), dim3(64 * MLDB_WAVES), 0, s, p, kps, nkps, desc64);
}
