// Detector kernels: Hessian response, in-level extrema with subpixel fit,
// ordered compaction, neighbour search and wave-synchronous suppression.
//
// All levels are handled by one launch per stage: blocks map to (level, tile)
// through the prefix table in akz_det_params, and blockIdx.z selects the
// image of a batch (its arrays sit z * zs bytes after image 0's).  Candidate order is
// (level, y, x), the order the reference establishes with its per-level
// bitonic sort (cuda_akaze.cu:882-939), produced here without sorting: each
// 64-pixel row segment is one wavefront whose ballot gives in-segment ranks,
// and an exclusive scan over the segments in raster order gives offsets.
#include "akz_kernels.h"

namespace {

__device__ __forceinline__ int find_level(const akz_det_params& p, int b) {
  int l = 0;
  while (l + 1 < p.nlev && b >= p.tile_start[l + 1]) l++;
  return l;
}

// Derivate (cuda_akaze.cu:651-671): first derivatives at integer step s,
// reflect-101 at offset s.  Element (y, x) of the source is f[ST * (y * pitch
// + x)]: ST = 2 reads one component of the interleaved Lx / Ly.
template <int ST>
__device__ __forceinline__ void deriv_xy(const float* f, int pitch, int w, int h, int x, int y, int s, float fac1,
                                         float fac2, float* dx, float* dy) {
  int xl = (x < s ? s - x : x - s);
  int xh = (x >= w - s ? 2 * w - x - s - 2 : x + s);
  int yl = (y < s ? s - y : y - s);
  int yh = (y >= h - s ? 2 * h - y - s - 2 : y + s);
  auto at = [&](int yy, int xx) { return f[ST * ((size_t)yy * pitch + xx)]; };
  float ul = at(yl, xl), ur = at(yl, xh);
  float ll = at(yh, xl), lr = at(yh, xh);
  if (dx) {
    float cl = at(y, xl), cr = at(y, xh);
    *dx = fac1 * (ur + lr - ul - ll) + fac2 * (cr - cl);
  }
  if (dy) {
    float uc = at(yl, x), lc = at(yh, x);
    *dy = fac1 * (lr + ll - ur - ul) + fac2 * (lc - uc);
  }
}

__global__ __launch_bounds__(256) void k_deriv(akz_det_params p) {
  const int l = find_level(p, blockIdx.x);
  const akz_level_desc& L = p.lev[l];
  int t = blockIdx.x - p.tile_start[l];
  int tiles_x = (L.w + 63) / 64;
  int x = (t % tiles_x) * 64 + (threadIdx.x & 63);
  int y0 = (t / tiles_x) * 16;
  if (x >= L.w) return;
  for (int y = y0 + akz_wave(); y < y0 + 16 && y < L.h; y += 4) {
    float dx, dy;
    deriv_xy<1>(akz_zoff(p.Lsm[l], p.zs), L.pitch, L.w, L.h, x, y, L.deriv_step, p.fac1, p.fac2, &dx, &dy);
    akz_zoff(p.Lxy[l], p.zs)[(size_t)y * L.pitch + x] = make_float2(dx, dy);
  }
}

// HessianDeterminant (cuda_akaze.cu:673-701): det = Lxx*Lyy - Lxy^2 with the
// same stencil applied to Lx and Ly.
__global__ __launch_bounds__(256) void k_hess(akz_det_params p) {
  const int l = find_level(p, blockIdx.x);
  const akz_level_desc& L = p.lev[l];
  int t = blockIdx.x - p.tile_start[l];
  int tiles_x = (L.w + 63) / 64;
  int x = (t % tiles_x) * 64 + (threadIdx.x & 63);
  int y0 = (t / tiles_x) * 16;
  if (x >= L.w) return;
  for (int y = y0 + akz_wave(); y < y0 + 16 && y < L.h; y += 4) {
    float lxx, lyx, lyy;
    const float* lxy = reinterpret_cast<const float*>(akz_zoff(p.Lxy[l], p.zs));
    deriv_xy<2>(lxy, L.pitch, L.w, L.h, x, y, L.deriv_step, p.fac1, p.fac2, &lxx, &lyx);
    deriv_xy<2>(lxy + 1, L.pitch, L.w, L.h, x, y, L.deriv_step, p.fac1, p.fac2, nullptr, &lyy);
    akz_zoff(p.det[l], p.zs)[(size_t)y * L.pitch + x] = lxx * lyy - lyx * lyx;
  }
}

// FindExtrema predicate (cuda_akaze.cu:734-743): inside the border, above the
// threshold and strictly greater than the 8 in-level neighbours.
__device__ __forceinline__ bool is_extremum(const float* imd, int pitch, int w, int h, int x, int y, float border,
                                            float thr) {
  if (x >= w || y >= h) return false;
  int left_x = (int)(x - border + 0.5f) - 1;
  int right_x = (int)(x + border + 0.5f) + 1;
  int up_y = (int)(y - border + 0.5f) - 1;
  int down_y = (int)(y + border + 0.5f) + 1;
  if (left_x < 0 || right_x >= w || up_y < 0 || down_y >= h) return false;
  size_t p = (size_t)y * pitch + x;
  float v = imd[p];
  return v > thr && v > imd[p - pitch - 1] && v > imd[p + pitch + 1] && v > imd[p + pitch - 1] &&
         v > imd[p - pitch + 1] && v > imd[p - 1] && v > imd[p + 1] && v > imd[p + pitch] && v > imd[p - pitch];
}

// One wavefront per 64-pixel row segment: the ballot of the predicate gives
// the segment's count and each candidate's rank within it.
__global__ __launch_bounds__(256) void k_extrema(akz_det_params p, int32_t* __restrict__ seg_count,
                                                 uint8_t* __restrict__ seg_slots) {
  seg_count = akz_zoff(seg_count, p.zs);
  seg_slots = akz_zoff(seg_slots, p.zs);
  const int l = find_level(p, blockIdx.x);
  const akz_level_desc& L = p.lev[l];
  int t = blockIdx.x - p.tile_start[l];
  int tx = t % L.nseg_x;
  int lane = threadIdx.x & 63;
  int x = tx * 64 + lane;
  int y0 = (t / L.nseg_x) * 16;
  for (int y = y0 + akz_wave(); y < y0 + 16 && y < L.h; y += 4) {
    bool pred = is_extremum(akz_zoff(p.det[l], p.zs), L.pitch, L.w, L.h, x, y, L.border, p.thr);
    uint64_t mask = __ballot(pred);
    int seg = L.seg_base + y * L.nseg_x + tx;
    if (lane == 0) seg_count[seg] = __popcll(mask);
    if (pred) {
      int rank = __popcll(mask & ((1ull << lane) - 1ull));
      seg_slots[(size_t)seg * 32 + rank] = (uint8_t)lane;
    }
  }
}

// Derivate + HessianDeterminant + FindExtrema (cuda_akaze.cu:651-774) fused
// per 64x30 tile of every level in one launch.  Lsmooth is loaded with a
// (2s+1)-pixel halo, Lx and Ly are computed on an (s+1)-pixel halo and the
// determinant on a one-pixel ring, all in LDS (a wavefront's lanes walk one
// row, the few columns past 64 are walked down).  The determinant reuses the
// Lsmooth buffer, dead once the first derivatives exist; with 30 rows and row
// strides of 83 (Lsmooth, determinant) and 75 floats (Lx, Ly) a block holds
// 40320 bytes of LDS (with the proposal's tile maxima), so four blocks share a
// CU (32 rows and one 97-float stride took 52 KB: three blocks, 6% slower).  Tiles whose halo stays inside
// the image load without clamps or bounds tests.  Every stencil addresses its
// reflect-101 neighbours through image coordinates, which stay inside the
// loaded regions.  Lx, Ly and det are written for the tile's own pixels and
// each 64-pixel row is one wavefront whose ballot gives the candidate count
// and ranks, as in k_extrema.
constexpr int DF_SMAX = 4, DF_TH = 30;
constexpr int DF_H1 = 2 * DF_SMAX + 1, DF_H2 = DF_SMAX + 1;
// LDS row strides, at least the loaded width and odd: the halo columns
// beyond the first 64 are walked down their columns, and consecutive rows of
// one column then fall into different banks
constexpr int DF_RL = 83, DF_RX = 75;
static_assert(DF_RL >= 64 + 2 * DF_H1 && DF_RX >= 64 + 2 * DF_H2 && (DF_RL & 1) && (DF_RX & 1), "odd row strides");
constexpr int DF_R1 = DF_TH + 2 * DF_H1, DF_R2 = DF_TH + 2 * DF_H2, DF_RD = DF_TH + 2;
static_assert(DF_RD <= DF_R1, "the determinant fits the Lsmooth buffer");

// Derivative and determinant phases of a tile whose stencils read plain
// LDS positions (interior tiles, and border tiles whose LDS grids hold the
// reflect-101 images of the positions outside the image): with the step S
// known at compile time, position (r, c) of each grid reads rows r, r+S,
// r+2S and columns c, c+S, c+2S of the grid below it, at constant LDS
// offsets.  Same operands in the same order as the general path.  `mid`
// runs between the two phases, after the first barrier.
template <int S, class Mid>
__device__ __forceinline__ void detect_interior(const float* s_l, float* s_x, float* s_y, float* s_d, float fac1,
                                                float fac2, int wv, int lane, Mid mid) {
  constexpr int NW = 8, H2 = S + 1, R2n = DF_TH + 2 * H2, E2 = 2 * H2, R = DF_RL, X = DF_RX;
  auto deriv = [&](int r, int c) {
    const float* b = s_l + r * R + c;
    const float ul = b[0], ur = b[2 * S], ll = b[2 * S * R], lr = b[2 * S * R + 2 * S];
    const float cl = b[S * R], cr = b[S * R + 2 * S];
    const float uc = b[S], lc = b[2 * S * R + S];
    s_x[r * X + c] = fac1 * (ur + lr - ul - ll) + fac2 * (cr - cl);
    s_y[r * X + c] = fac1 * (lr + ll - ur - ul) + fac2 * (lc - uc);
  };
  // rows wv, wv + NW, ...: the first R2n / NW rows of every wavefront
  // exist, unrolled without tests; then at most one more
#pragma unroll
  for (int k = 0; k < R2n / NW; k++) deriv(wv + k * NW, lane);
  if (R2n % NW && wv + (R2n / NW) * NW < R2n) deriv(wv + (R2n / NW) * NW, lane);
  for (int k = threadIdx.x; k < R2n * E2; k += 64 * NW) deriv(k % R2n, 64 + k / R2n);
  __syncthreads();
  mid();
  auto det = [&](int r, int c) {
    const float* bx = s_x + r * X + c;
    const float* by = s_y + r * X + c;
    float ul = bx[0], ur = bx[2 * S], ll = bx[2 * S * X], lr = bx[2 * S * X + 2 * S];
    const float cl = bx[S * X], cr = bx[S * X + 2 * S];
    const float lxx = fac1 * (ur + lr - ul - ll) + fac2 * (cr - cl);
    float uc = bx[S], lc = bx[2 * S * X + S];
    const float lyx = fac1 * (lr + ll - ur - ul) + fac2 * (lc - uc);
    ul = by[0]; ur = by[2 * S]; ll = by[2 * S * X]; lr = by[2 * S * X + 2 * S];
    uc = by[S]; lc = by[2 * S * X + S];
    const float lyy = fac1 * (lr + ll - ur - ul) + fac2 * (lc - uc);
    s_d[r * R + c] = lxx * lyy - lyx * lyx;
  };
#pragma unroll
  for (int k = 0; k < DF_RD / NW; k++) det(wv + k * NW, lane);
  if (DF_RD % NW && wv + (DF_RD / NW) * NW < DF_RD) det(wv + (DF_RD / NW) * NW, lane);
  for (int k = threadIdx.x; k < DF_RD * 2; k += 64 * NW) det(k % DF_RD, 64 + k / DF_RD);
}

// One 64 x DF_TH tile `bx` of image `z` (tiles numbered level-major through
// p.tile_start); the block's 512 threads all call it, with the same bx and z.
__device__ __forceinline__ void detect_tile(const akz_det_params& p, int bx, int z, int32_t* __restrict__ seg_count,
                                            uint8_t* __restrict__ seg_slots) {
  seg_count = akz_zoff(seg_count, p.zs, z);
  seg_slots = akz_zoff(seg_slots, p.zs, z);
  __shared__ float s_l[DF_R1 * DF_RL];
  __shared__ float s_x[DF_R2 * DF_RX];
  __shared__ float s_y[DF_R2 * DF_RX];
  float* const s_d = s_l;  // DF_RD <= DF_R1 rows
  const int l = find_level(p, bx);
  const akz_level_desc& L = p.lev[l];
  const int t = bx - p.tile_start[l];
  const int tx = t % L.nseg_x;
  const int x0 = tx * 64, y0 = (t / L.nseg_x) * DF_TH;
  const int w = L.w, h = L.h, pitch = L.pitch, s = L.deriv_step;
  const int H1 = 2 * s + 1, H2 = s + 1;
  const int W1 = 64 + 2 * H1, W2 = 64 + 2 * H2;
  const int lane = threadIdx.x & 63, wv = akz_wave();
  constexpr int NW = 8;
  const float* lsm = akz_zoff(p.Lsm[l], p.zs, z);
  const bool interior = x0 >= H1 && x0 + 64 + H1 <= w && y0 >= H1 && y0 + DF_TH + H1 <= h;
  // A border tile takes the interior path when its grids can hold the
  // reflect-101 images of their positions outside the image: Lsmooth is
  // loaded from reflected coordinates, and every Lx / Ly position outside
  // the image is then a copy of the reflected one, which lies in the tile's
  // own derivative grid when the tile keeps at least 33 columns (16 rows)
  // inside the image at its right (bottom) border.  One reflection must
  // bring every loaded coordinate into the image: c in [-(n - 1), 2n - 2]
  // (small levels, narrower than the tile plus its halo, may not allow it).
  const bool reflected = !interior && s >= 2 && s <= 4 && H1 <= w - 1 && H1 <= h - 1 &&
                         x0 + 63 + H1 <= 2 * w - 2 && y0 + DF_TH - 1 + H1 <= 2 * h - 2 &&
                         (x0 + 64 <= w || w - x0 >= 33) && (y0 + DF_TH <= h || h - y0 >= 16);
  // Lsmooth rows [y0-H1, y0+TH+H1), columns [x0-H1, x0+64+H1): every load of
  // a thread is issued before any is written to LDS
  if (reflected) {
    constexpr int ROWS = (DF_TH + 2 * DF_H1 + NW - 1) / NW;
    const int R1n = DF_TH + 2 * H1;
    float v[ROWS][2];
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
      const int r = wv + k * NW;
      if (r < R1n) {
        const float* row = lsm + (size_t)akz_reflect101(y0 - H1 + r, h) * pitch;
        v[k][0] = row[akz_reflect101(x0 - H1 + lane, w)];
        if (lane < W1 - 64) v[k][1] = row[akz_reflect101(x0 - H1 + 64 + lane, w)];
      }
    }
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
      const int r = wv + k * NW;
      if (r < R1n) {
        s_l[r * DF_RL + lane] = v[k][0];
        if (lane < W1 - 64) s_l[r * DF_RL + 64 + lane] = v[k][1];
      }
    }
  } else if (interior) {
    constexpr int ROWS = (DF_TH + 2 * DF_H1 + NW - 1) / NW;
    const int R1n = DF_TH + 2 * H1;
    const float* base = lsm + (size_t)(y0 - H1) * pitch + (x0 - H1);
    float v[ROWS][2];
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
      const int r = wv + k * NW;
      if (r < R1n) {
        v[k][0] = base[r * pitch + lane];
        if (lane < W1 - 64) v[k][1] = base[r * pitch + 64 + lane];
      }
    }
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
      const int r = wv + k * NW;
      if (r < R1n) {
        s_l[r * DF_RL + lane] = v[k][0];
        if (lane < W1 - 64) s_l[r * DF_RL + 64 + lane] = v[k][1];
      }
    }
  } else {
    constexpr int ROWS = (DF_TH + 2 * DF_H1 + NW - 1) / NW, COLS = 2;
    float v[ROWS][COLS];
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
      const int r = wv + k * NW, gy = y0 - H1 + r;
#pragma unroll
      for (int q = 0; q < COLS; q++) {
        const int c = lane + q * 64, gx = x0 - H1 + c;
        if (r < DF_TH + 2 * H1 && c < W1 && gy >= 0 && gy < h && gx >= 0 && gx < w)
          v[k][q] = lsm[(size_t)gy * pitch + gx];
      }
    }
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
      const int r = wv + k * NW, gy = y0 - H1 + r;
#pragma unroll
      for (int q = 0; q < COLS; q++) {
        const int c = lane + q * 64, gx = x0 - H1 + c;
        if (r < DF_TH + 2 * H1 && c < W1 && gy >= 0 && gy < h && gx >= 0 && gx < w) s_l[r * DF_RL + c] = v[k][q];
      }
    }
  }
  __syncthreads();
  const float fac1 = p.fac1, fac2 = p.fac2;
  if ((interior || reflected) && s >= 2 && s <= 4) {
    // reflected tiles: the Lx / Ly positions outside the image become copies
    // of their reflect-101 images (always inside the image and the grid)
    // before the determinant reads them
    auto mid = [&] {
      if (!reflected) return;
      const int R2n = DF_TH + 2 * H2, W2 = 64 + 2 * H2;
      for (int k = threadIdx.x; k < R2n * W2; k += 64 * NW) {
        const int r = k / W2, c = k - r * W2;
        const int y = y0 - H2 + r, x = x0 - H2 + c;
        if (y >= 0 && y < h && x >= 0 && x < w) continue;
        const int src = (akz_reflect101(y, h) - (y0 - H2)) * DF_RX + akz_reflect101(x, w) - (x0 - H2);
        s_x[r * DF_RX + c] = s_x[src];
        s_y[r * DF_RX + c] = s_y[src];
      }
      __syncthreads();
    };
    if (s == 2) detect_interior<2>(s_l, s_x, s_y, s_d, fac1, fac2, wv, lane, mid);
    else if (s == 3) detect_interior<3>(s_l, s_x, s_y, s_d, fac1, fac2, wv, lane, mid);
    else detect_interior<4>(s_l, s_x, s_y, s_d, fac1, fac2, wv, lane, mid);
  } else {
    // first derivatives on rows [y0-H2, y0+TH+H2), columns [x0-H2, x0+64+H2):
    // 64 columns per row-pass, then the 2*H2 leftover columns as one flat list
    auto deriv_at = [&](int r, int c) {
      const int y = y0 - H2 + r, x = x0 - H2 + c;
      if (x < 0 || x >= w || y < 0 || y >= h) return;
      const int yl = ((y < s ? s - y : y - s) - (y0 - H1)) * DF_RL;
      const int yh = ((y >= h - s ? 2 * h - y - s - 2 : y + s) - (y0 - H1)) * DF_RL;
      const int yc = (y - (y0 - H1)) * DF_RL;
      const int xl = (x < s ? s - x : x - s) - (x0 - H1);
      const int xh = (x >= w - s ? 2 * w - x - s - 2 : x + s) - (x0 - H1);
      const int xc = x - (x0 - H1);
      float ul = s_l[yl + xl], ur = s_l[yl + xh], ll = s_l[yh + xl], lr = s_l[yh + xh];
      float cl = s_l[yc + xl], cr = s_l[yc + xh];
      float uc = s_l[yl + xc], lc = s_l[yh + xc];
      s_x[r * DF_RX + c] = fac1 * (ur + lr - ul - ll) + fac2 * (cr - cl);
      s_y[r * DF_RX + c] = fac1 * (lr + ll - ur - ul) + fac2 * (lc - uc);
    };
    const int R2n = DF_TH + 2 * H2, E2 = 2 * H2;
    for (int r = wv; r < R2n; r += NW) deriv_at(r, lane);
    for (int k = threadIdx.x; k < R2n * E2; k += 64 * NW) deriv_at(k % R2n, 64 + k / R2n);
    __syncthreads();
    // determinant on rows [y0-1, y0+TH+1), columns [x0-1, x0+65)
    auto det_at = [&](int r, int c) {
      const int y = y0 - 1 + r, x = x0 - 1 + c;
      if (x < 0 || x >= w || y < 0 || y >= h) return;
      const int yl = ((y < s ? s - y : y - s) - (y0 - H2)) * DF_RX;
      const int yh = ((y >= h - s ? 2 * h - y - s - 2 : y + s) - (y0 - H2)) * DF_RX;
      const int yc = (y - (y0 - H2)) * DF_RX;
      const int xl = (x < s ? s - x : x - s) - (x0 - H2);
      const int xh = (x >= w - s ? 2 * w - x - s - 2 : x + s) - (x0 - H2);
      const int xc = x - (x0 - H2);
      float ul = s_x[yl + xl], ur = s_x[yl + xh], ll = s_x[yh + xl], lr = s_x[yh + xh];
      float cl = s_x[yc + xl], cr = s_x[yc + xh];
      float lxx = fac1 * (ur + lr - ul - ll) + fac2 * (cr - cl);
      float uc = s_x[yl + xc], lc = s_x[yh + xc];
      float lyx = fac1 * (lr + ll - ur - ul) + fac2 * (lc - uc);
      ul = s_y[yl + xl]; ur = s_y[yl + xh]; ll = s_y[yh + xl]; lr = s_y[yh + xh];
      uc = s_y[yl + xc]; lc = s_y[yh + xc];
      float lyy = fac1 * (lr + ll - ur - ul) + fac2 * (lc - uc);
      s_d[r * DF_RL + c] = lxx * lyy - lyx * lyx;
    };
    for (int r = wv; r < DF_RD; r += NW) det_at(r, lane);
    for (int k = threadIdx.x; k < DF_RD * 2; k += 64 * NW) det_at(k % DF_RD, 64 + k / DF_RD);
  }
  __syncthreads();
  const int x = x0 + lane;
  float2* Lxy = akz_zoff(p.Lxy[l], p.zs, z);
  float* D = akz_zoff(p.det[l], p.zs, z);
  // the detector border test splits into a column part and a row part,
  // each an interval precomputed on the host
  const bool col_in = x >= L.xlo && x < L.xhi;
  for (int r = wv; r < DF_TH; r += NW) {
    const int y = y0 + r;
    if (y >= h) break;
    bool pred = false;
    const int id = (r + 1) * DF_RL + lane + 1;
    const size_t o = (size_t)y * pitch + x;
    const bool row_in = y >= L.ylo && y < L.yhi;
    float v = 0.0f;
    if (x < w) {
      const int i2 = (r + H2) * DF_RX + lane + H2;
      Lxy[o] = make_float2(s_x[i2], s_y[i2]);
      v = s_d[id];
      if (p.det_full) D[o] = v;
    }
    // most rows of 64 pixels hold no response above the threshold (about 86%
    // at 1080p): the neighbour test runs only for a row that does
    const bool above = x < w && col_in && row_in && v > p.thr;
    if (__ballot(above) != 0ull) {
      // the eight neighbours are inside the tile's determinant rows for every
      // lane: read them all, combine without branches
      const int R = DF_RL;
      const float n0 = s_d[id - R - 1], n1 = s_d[id + R + 1], n2 = s_d[id + R - 1], n3 = s_d[id - R + 1];
      const float n4 = s_d[id - 1], n5 = s_d[id + 1], n6 = s_d[id + R], n7 = s_d[id - R];
      pred = above & (v > n0) & (v > n1) & (v > n2) & (v > n3) & (v > n4) & (v > n5) & (v > n6) & (v > n7);
    }
    // without the full determinant image, the subpixel fit of k_scan_scatter
    // still finds the 3x3 neighbourhood of every candidate (all inside the
    // image: candidates keep the detector border)
    if (pred && !p.det_full) {
#pragma unroll
      for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) D[o + (ptrdiff_t)dy * pitch + dx] = s_d[id + dy * DF_RL + dx];
    }
    uint64_t mask = __ballot(pred);
    int seg = L.seg_base + y * L.nseg_x + tx;
    if (lane == 0) seg_count[seg] = __popcll(mask);
    if (pred) seg_slots[(size_t)seg * 32 + __popcll(mask & ((1ull << lane) - 1ull))] = (uint8_t)lane;
  }
  if (p.tile_key) {
    // Maxima of the determinant over the learned proposal's 16x16 tiles (the
    // statistic k_tile_stats takes from a full determinant image): the tile's
    // 64 columns are four proposal tiles, its rows touch at most three
    // proposal tile rows; per wavefront, per lane group of 16, then over the
    // wavefronts, and into the tile's key with atomicMax (a proposal tile
    // row can span two detector tiles).
    static_assert(DF_TH <= 2 * AKZ_PROP_TILE + 1, "a tile's rows touch three proposal tile rows at most");
    __shared__ float s_tm[NW][3][4];
    const int pr0 = y0 / AKZ_PROP_TILE;
    float m0 = -INFINITY, m1 = -INFINITY, m2 = -INFINITY;
    for (int r = wv; r < DF_TH; r += NW) {
      const int y = y0 + r;
      if (y >= h) break;
      if (x < w) {
        const float v = s_d[(r + 1) * DF_RL + lane + 1];
        const int k = y / AKZ_PROP_TILE - pr0;
        if (k == 0) m0 = fmaxf(m0, v);
        else if (k == 1) m1 = fmaxf(m1, v);
        else m2 = fmaxf(m2, v);
      }
    }
    for (int off = 1; off < 16; off <<= 1) {
      m0 = fmaxf(m0, __shfl_xor(m0, off, 64));
      m1 = fmaxf(m1, __shfl_xor(m1, off, 64));
      m2 = fmaxf(m2, __shfl_xor(m2, off, 64));
    }
    if ((lane & 15) == 0) {
      s_tm[wv][0][lane >> 4] = m0;
      s_tm[wv][1][lane >> 4] = m1;
      s_tm[wv][2][lane >> 4] = m2;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
      const int k = threadIdx.x >> 2, c = threadIdx.x & 3;
      float mm = -INFINITY;
      for (int q = 0; q < NW; q++) mm = fmaxf(mm, s_tm[q][k][c]);
      const int pty = pr0 + k, ptx = x0 / AKZ_PROP_TILE + c;
      if (mm > -INFINITY && pty < p.tile_h[l] && ptx < p.tile_w[l])
        atomicMax(p.tile_key + (size_t)z * p.tiles_per_image + p.tile_base[l] + pty * p.tile_w[l] + ptx,
                  akz_det_key(mm));
    }
  }
}

__global__ __launch_bounds__(512) void k_detect_fused(akz_det_params p, int32_t* __restrict__ seg_count,
                                                      uint8_t* __restrict__ seg_slots) {
  detect_tile(p, blockIdx.x, blockIdx.z, seg_count, seg_slots);
}

// Persistent form over a ragged batch: the tiles of every image of every
// size class form one work list (image i owns items [start[i], start[i+1]),
// its level-major tiles in order); each block walks the list with the grid
// stride and runs the tile with its class's parameters.
// (eight wavefronts per SIMD, as the grid form reaches: four 40 KB blocks per CU)
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) void k_detect_fused_list(
    akz_det_worklist wl) {
  const int total = wl.start[wl.nimg];
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    int a = 0, e = wl.nimg;  // last image with start <= item
    while (e - a > 1) {
      const int m = (a + e) >> 1;
      if (wl.start[m] <= item) a = m; else e = m;
    }
    // the item is the block's: keep its decode in scalar registers
    const int ia = __builtin_amdgcn_readfirstlane(a);
    const int c = __builtin_amdgcn_readfirstlane(wl.cls[ia]);
    const int bx = __builtin_amdgcn_readfirstlane(item - wl.start[ia]);
    const int z = __builtin_amdgcn_readfirstlane(wl.z[ia]);
    detect_tile(wl.P[c], bx, z, wl.seg_count[c], wl.seg_slots[c]);
    __syncthreads();  // the next tile's loads overwrite this one's LDS
  }
}

// Scan blocks of k_seg_sums: 1024 threads, each owning four consecutive
// counts read as one int4.
constexpr int SCAN_T = 1024, SCAN_PER = 4, SCAN_BLK = SCAN_T * SCAN_PER;

__device__ __forceinline__ int4 load4(const int32_t* a, int i, int n) {
  if (i + 3 < n) return *reinterpret_cast<const int4*>(a + i);
  int4 v = make_int4(0, 0, 0, 0);
  if (i < n) v.x = a[i];
  if (i + 1 < n) v.y = a[i + 1];
  if (i + 2 < n) v.z = a[i + 2];
  return v;
}

// block-wide exclusive scan of one value per thread; returns the block total
__device__ __forceinline__ int block_exclusive_scan(int v, int* excl, int32_t* s_w) {
  const int lane = threadIdx.x & 63, wv = akz_wave();
  int inc = v;
  for (int d = 1; d < 64; d <<= 1) {
    int o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_w[wv] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) { int tv = s_w[i]; s_w[i] = run; run += tv; }
    s_w[16] = run;
  }
  __syncthreads();
  *excl = s_w[wv] + inc - v;
  int tot = s_w[16];
  __syncthreads();
  return tot;
}

// Candidate r of row segment `seg` (its rank r in seg_slots) in keypoint
// layout at cand[idx] (FindExtrema, cuda_akaze.cu:744-772): the subpixel fit
// on the determinant's 3x3 neighbourhood.
__device__ __forceinline__ void scatter_one(const akz_det_params& p, const akz_det_bufs& b, int seg, int r, int idx) {
  int l = 0;
  while (l + 1 < p.nlev && seg >= p.lev[l + 1].seg_base) l++;
  const akz_level_desc& L = p.lev[l];
  const int ls = seg - L.seg_base;
  const int y = ls / L.nseg_x, tx = ls - y * L.nseg_x;
  const uint8_t* slots = akz_zoff(b.seg_slots, b.zs);
  akz_kp* cand = akz_zoff(b.cand, b.zs);
  float2* cand_xy = akz_zoff(b.cand_xy, b.zs);
  const float* imd = akz_zoff(p.det[l], p.zs);
  const size_t pitch = L.pitch;
  {
    const int x = tx * 64 + slots[(size_t)seg * 32 + r];
    const size_t q = (size_t)y * pitch + x;
    float v = imd[q];
    float dx = 0.5f * (imd[q + 1] - imd[q - 1]);
    float dy = 0.5f * (imd[q + pitch] - imd[q - pitch]);
    float dxx = imd[q + 1] + imd[q - 1] - 2.0f * v;
    float dyy = imd[q + pitch] + imd[q - pitch] - 2.0f * v;
    float dxy = 0.25f * (imd[q + pitch + 1] + imd[q - pitch - 1] - imd[q + pitch - 1] - imd[q - pitch + 1]);
    float det = dxx * dyy - dxy * dxy;
    float idet = (det != 0.0f ? 1.0f / det : 0.0f);
    float dst0 = idet * (dxy * dy - dyy * dx);
    float dst1 = idet * (dxy * dx - dxx * dy);
    bool weak = !(dst0 >= -1.0f && dst0 <= 1.0f && dst1 >= -1.0f && dst1 <= 1.0f);
    akz_kp k;
    k.response = v;
    k.size = (weak ? -1 : 1) * 2.0 * L.det_size;
    float octsub = (dst0 < 0 ? -1 : 1) * (L.octave + fabsf(dst0));
    k.octave = akz_f2bits(weak ? (float)L.octave : octsub);
    k.class_id = l;
    int ratio = (1 << L.octave);
    k.x = ratio * (x);
    k.y = ratio * (y);
    k.angle = dst1;
    cand[idx] = k;
    cand_xy[idx] = make_float2(k.x, k.y);
  }
}

// Sums of the row segments' candidate counts over runs of 256 segments
// (one per wave: 64 lanes x 4 counts), the carries of k_scan_scatter's blocks.
__global__ __launch_bounds__(SCAN_T) void k_seg_sums(const int32_t* __restrict__ cnt, int n,
                                                      int32_t* __restrict__ wsum, size_t zs) {
  cnt = akz_zoff(cnt, zs);
  wsum = akz_zoff(wsum, zs);
  const int i = blockIdx.x * SCAN_BLK + threadIdx.x * SCAN_PER;
  const int4 v = load4(cnt, i, n);
  int t = v.x + v.y + v.z + v.w;
  for (int d = 32; d >= 1; d >>= 1) t += __shfl_xor(t, d, 64);
  if ((threadIdx.x & 63) == 0) wsum[blockIdx.x * (SCAN_T / 64) + akz_wave()] = t;
}

// The scan of the row segments' candidate counts fused with the scatter of
// the candidates, one segment per thread: a segment's offset is the sum of
// the 256-segment runs before its block (k_seg_sums) plus the block's
// exclusive scan, and its candidates are written there, so the candidate
// order is (level, y, x).  The candidates of a wavefront's 64 segments are
// contiguous; its lanes take them in turn (candidate j of the wavefront to
// lane j % 64, its segment from an LDS table), so a segment with many
// candidates does not hold one lane for all of them.  Candidates at or past
// the capacity are dropped.  The thread of a level's first segment writes
// the level's start; the thread of index seg_total (the grid covers it)
// writes the total, the candidate count (cut at the capacity), and clears
// the overflow flags of this detection run.
constexpr int SS_T = 256;
__global__ __launch_bounds__(SS_T) void k_scan_scatter(akz_det_params p, akz_det_bufs b) {
  __shared__ int32_t s_w[17];
  __shared__ uint8_t s_tab[SS_T / 64][64 * 32];  // the segment (lane) of each candidate of a wavefront
  const int32_t* wsum = akz_zoff(b.seg_wsum, b.zs);
  const int n = p.seg_total;
  const int seg = blockIdx.x * SS_T + threadIdx.x;
  const int c = seg < n ? akz_zoff(b.seg_count, b.zs)[seg] : 0;
  int part = 0;
  for (int j = threadIdx.x; j < (int)blockIdx.x; j += SS_T) part += wsum[j];
  int ex;
  const int carry = block_exclusive_scan(part, &ex, s_w);
  block_exclusive_scan(c, &ex, s_w);
  const int o = carry + ex;
  int32_t* ls = akz_zoff(b.level_start, b.zs);
  for (int l = 0; l < p.nlev; l++)
    if (seg == p.lev[l].seg_base) ls[l] = o;
  if (seg == n) {
    int32_t* st = akz_zoff(b.status, b.zs);
    ls[p.nlev] = o;
    st[AKZ_ST_TOTAL] = o;
    st[AKZ_ST_FLAGS] = 0;
    *akz_zoff(b.ncand, b.zs) = akz_imin(o, b.cap);
    *reinterpret_cast<unsigned long long*>(akz_zoff(b.nb_pool, b.zs)) = 0ull;  // k_nb_pool's allocator
  }
  // the wavefront's candidates: [o0, o0 + T)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int o0 = __shfl(o, 0, 64);
  const int T = __shfl(o + c, 63, 64) - o0;
  uint8_t* tab = s_tab[wv];
  for (int r = 0; r < c; r++) tab[o - o0 + r] = (uint8_t)lane;
  // the table's writes precede its reads in the wavefront's LDS order
  __builtin_amdgcn_wave_barrier();
  for (int j0 = 0; j0 < T; j0 += 64) {
    const int j = j0 + lane;
    int sl = 0;
    if (j < T) sl = tab[j];
    const int os = __shfl(o, sl, 64);  // the segment's offset (all lanes take part)
    if (j < T && o0 + j < b.cap) scatter_one(p, b, seg - lane + sl, o0 + j - os, o0 + j);
  }
}

// Neighbour list of FindNeighbors (cuda_akaze.cu:1005-1070) for candidate c
// at (kx, ky) of size `size`: earlier candidates of its level (indices
// start .. c-1, walked downward) and the candidates of the previous level
// (prev_lo .. start-1, when has_prev) within half the size.  A weak
// candidate (negative size) gets none, as the reference's early break
// gives.  The previous level is entered at the last candidate whose y is
// within reach (binary search) instead of being walked from its end.  xy(i)
// is candidate i's position; emit(k, i) takes the k-th neighbour i in list
// order.  Returns the count.
template <class XY, class Emit>
__device__ __forceinline__ int nb_search(int c, float kx, float ky, float size, int start, int prev_lo,
                                         bool has_prev, XY xy, Emit emit) {
  const float half = size * .5f;
  const double r2 = (double)(size * size) * 0.25;
  int cnt = 0;
  // walk i0, i0 - 1, ... down to lo, stopping at the first candidate more
  // than half above; four positions per round of loads, examined in order
  auto scan = [&](int i0, int lo) {
    for (int i = i0; i >= lo; i -= 4) {
      float2 q[4];
#pragma unroll
      for (int u = 0; u < 4; u++) q[u] = (i - u >= lo) ? xy(i - u) : make_float2(0.0f, 0.0f);
      bool stop = false;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (stop) break;
        if (i - u < lo || ky - q[u].y > half) { stop = true; break; }
        float dist = (kx - q[u].x) * (kx - q[u].x) + (ky - q[u].y) * (ky - q[u].y);
        if ((double)dist < r2) emit(cnt++, i - u);
      }
      if (stop) break;
    }
  };
  scan(c - 1, start);
  if (has_prev) {
    // first index in [prev_lo, start) with q.y - ky > half (monotone in the
    // sorted order; indices below a are false, from e on true): a four-way
    // search, three pivots loaded together
    int a = prev_lo, e = start;
    while (e - a >= 4) {
      const int d = (e - a) >> 2;
      const int m1 = a + d, m2 = a + 2 * d, m3 = a + 3 * d;
      const bool p1 = xy(m1).y - ky > half, p2 = xy(m2).y - ky > half, p3 = xy(m3).y - ky > half;
      if (p1) e = m1;
      else if (p2) { a = m1 + 1; e = m2; }
      else if (p3) { a = m2 + 1; e = m3; }
      else a = m3 + 1;
    }
    while (a < e) {
      const int m = (a + e) >> 1;
      if (xy(m).y - ky > half) e = m; else a = m + 1;
    }
    scan(a - 1, prev_lo);
  }
  return cnt;
}

constexpr int NMS_THREADS = 1024;
constexpr int NMS_LDS_N = 8192;     // candidates whose suppression runs in LDS
constexpr int NMS_PER = NMS_LDS_N / NMS_THREADS;  // candidates (or list entries) per thread
constexpr int NMS_FEXT = 4096;      // LDS pool of the neighbours past the first, for lists of three or more
constexpr int NMS_CNT_MAX = 63;     // longest list the packed form holds
constexpr int NMS_TAIL = 64;        // pending candidates from which one wavefront runs the waves
static_assert(NMS_TAIL % 64 == 0, "the tail keeps NMS_TAIL / 64 candidates per lane");
// A neighbour list packed in 32 bits: the first neighbour (13 bits), the
// second neighbour or, for three or more, the offset of neighbours 1.. in the
// pool (13 bits), and the count (6 bits).
static_assert(NMS_LDS_N <= 8192 && NMS_FEXT <= 8192, "13-bit fields");
__device__ __forceinline__ uint32_t fwd_pack(int q0, int f2, int cnt) {
  return (uint32_t)q0 | ((uint32_t)f2 << 13) | ((uint32_t)cnt << 26);
}
constexpr int CODE_WAIT = -1, CODE_ADD = -2;
constexpr int NMS_BITMAP_WORDS = 2048;  // bitmap ranking covers up to 65536 candidates

// Neighbour lists of every candidate, for the suppression kernel: each
// candidate's list is written by one thread, in list order, into space taken
// from a pool (nb_off = start, nb_count = length); a wavefront takes the
// space of its 64 lists with one atomic add, so no scan over the counts is
// needed, and where the lists sit in the pool does not matter.  A block
// takes 256 consecutive candidates; every position their searches can read
// lies between the start of the previous level of the first one and the
// last one, and that window is staged in LDS from the compact cand_xy array
// when it fits (NB_WIN positions; else the searches read cand_xy).  Entries
// past nb_cap are dropped and flag an overflow; the suppression reads only
// entries written in this run.
constexpr int NB_KEEP = 8;
constexpr int NB_WIN = 8192;
__global__ __launch_bounds__(256) void k_nb_lists(akz_det_bufs b) {
  __shared__ float2 s_win[NB_WIN];
  const int n = *akz_zoff(b.ncand, b.zs);
  const akz_kp* __restrict__ cand = akz_zoff(b.cand, b.zs);
  const float2* __restrict__ cxy = akz_zoff(b.cand_xy, b.zs);
  const int32_t* __restrict__ level_start = akz_zoff(b.level_start, b.zs);
  int32_t* nb_count = akz_zoff(b.nb_count, b.zs);
  int32_t* nb_off = akz_zoff(b.nb_off, b.zs);
  int32_t* nb_list = akz_zoff(b.nb_list, b.zs);
  unsigned long long* pool = reinterpret_cast<unsigned long long*>(akz_zoff(b.nb_pool, b.zs));
  const int lane = threadIdx.x & 63;
  // the loop runs while the block has candidates (uniform)
  for (int cb = blockIdx.x * 256; cb < n; cb += gridDim.x * 256) {
    const int ce = akz_imin(cb + 256, n);
    const int lv0 = cand[cb].class_id;
    const int win_lo = level_start[lv0 > 0 ? lv0 - 1 : 0];
    const bool staged = ce - win_lo <= NB_WIN;
    if (staged)
      for (int i = win_lo + (int)threadIdx.x; i < ce; i += 256) s_win[i - win_lo] = cxy[i];
    __syncthreads();
    auto xy = [&](int i) { return staged ? s_win[i - win_lo] : cxy[i]; };
    const int c = cb + (int)threadIdx.x;
    int cnt = 0, lvl = 0, start = 0, lo = 0;
    float2 p = make_float2(0.0f, 0.0f);
    float size = 0.0f;
    // the first NB_KEEP neighbours found stay in registers; longer lists are
    // searched again for the write
    int r[NB_KEEP];
#pragma unroll
    for (int u = 0; u < NB_KEEP; u++) r[u] = 0;
    if (c < n) {
      p = cxy[c];
      size = cand[c].size;
      lvl = cand[c].class_id;
      start = level_start[lvl];
      lo = lvl > 0 ? level_start[lvl - 1] : 0;
      cnt = nb_search(c, p.x, p.y, size, start, lo, lvl > 0, xy, [&](int k, int i) {
#pragma unroll
        for (int u = 0; u < NB_KEEP; u++) r[u] = (k == u) ? i : r[u];
      });
    }
    // the wavefront's space: inclusive prefix of the counts, one atomic add
    int inc = cnt;
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(inc, d, 64);
      if (lane >= d) inc += o;
    }
    const int total = __shfl(inc, 63, 64);
    unsigned long long wbase = 0;
    if (lane == 63 && total > 0) wbase = atomicAdd(pool, (unsigned long long)total);
    wbase = __shfl(wbase, 63, 64);
    if (c < n) {
      const unsigned long long base = wbase + (unsigned long long)(inc - cnt);
      const long long room_ll = (long long)b.nb_cap - (long long)base;
      const int room = room_ll <= 0 ? 0 : (room_ll < cnt ? (int)room_ll : cnt);
      int32_t* dst = nb_list + base;
      if (cnt <= NB_KEEP) {
#pragma unroll
        for (int u = 0; u < NB_KEEP; u++)
          if (u < room) dst[u] = r[u];
      } else if (room > 0) {
        nb_search(c, p.x, p.y, size, start, lo, lvl > 0, xy, [&](int j, int i) {
          if (j < room) dst[j] = i;
        });
      }
      nb_off[c] = room > 0 ? (int)base : 0;
      nb_count[c] = room;
      // the first two entries and min(count, 3) in one 8-byte word
      akz_zoff(b.nb_pack, b.zs)[c] = make_uint2((uint32_t)(room > 0 ? r[0] : 0) | ((uint32_t)akz_imin(room, 3) << 30),
                                                (uint32_t)(room > 1 ? r[1] : 0));
      if (room < cnt) atomicOr(akz_zoff(b.status, b.zs) + AKZ_ST_FLAGS, 1);
    }
    __syncthreads();  // the next block of candidates overwrites the window
  }
}

// block-wide exclusive scan for the 1024-thread suppression kernel
__device__ __forceinline__ int nms_block_scan(int v, int* excl, int32_t* s_w) {
  const int lane = threadIdx.x & 63, wv = akz_wave();
  int inc = v;
  for (int d = 1; d < 64; d <<= 1) {
    int o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_w[wv] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int i = 0; i < NMS_THREADS / 64; i++) { int tv = s_w[i]; s_w[i] = run; run += tv; }
    s_w[16] = run;
  }
  __syncthreads();
  *excl = s_w[wv] + inc - v;
  int tot = s_w[16];
  __syncthreads();
  return tot;
}

// Emitted keypoint of sortFiltered_kernel (cuda_akaze.cu:1271-1286): the
// subpixel offsets are restored from the float bits kept in `octave`.
__device__ __forceinline__ akz_kp emit(const akz_kp& ok) {
  float octf = akz_bits2f(ok.octave);
  float octsub = fabsf(octf);
  int octave = (int)octsub;
  float subp = (octf < 0 ? -1 : 1) * (octsub - octave);
  float ratio = 1 << octave;
  akz_kp t;
  t.y = ratio * ((int)(0.5f + ok.y / ratio) + ok.angle);
  t.x = ratio * ((int)(0.5f + ok.x / ratio) + subp);
  t.class_id = ok.class_id;
  t.octave = octave;
  t.response = ok.response;
  t.size = ok.size;
  t.angle = 0.0f;
  return t;
}

// Output of sortFiltered_kernel (cuda_akaze.cu:1227-1316): per 2048-slot
// chunk, the kept member values whose point is strong (strong_at: size >= 0),
// ascending.  The
// member values of a chunk are distinct candidate indices, so a bitmap over
// [0, n) and a prefix popcount of its words give every value its rank;
// candidate counts beyond the bitmap use a bitonic sort.  Returns the number
// of keypoints (thread 0's value).
template <class MemAt, class StrongAt>
__device__ int nms_output(const akz_kp* __restrict__ cand, int n, MemAt mem_at, StrongAt strong_at,
                          akz_kp* __restrict__ out, int32_t* __restrict__ kidx, int cap, int32_t* s_sort,
                          uint32_t* s_bits, int32_t* s_pre, int32_t* s_w, int32_t* s_run) {
  const int t = threadIdx.x;
  if (t == 0) *s_run = 0;
  __syncthreads();
  const int nwords = (n + 31) >> 5;
  for (int base = 0; base < n; base += 2048) {
    if (nwords <= NMS_BITMAP_WORDS) {
      for (int i = t; i < nwords; i += NMS_THREADS) s_bits[i] = 0u;
      __syncthreads();
      for (int e = t; e < 2048; e += NMS_THREADS) {
        const int i = base + e;
        const int v = (i < n) ? mem_at(i) : -1;
        if (v >= 0 && strong_at(v)) atomicOr(&s_bits[v >> 5], 1u << (v & 31));
      }
      __syncthreads();
      // exclusive prefix of word popcounts, two words per thread
      const int c0 = (2 * t < nwords) ? __popc(s_bits[2 * t]) : 0;
      const int c1 = (2 * t + 1 < nwords) ? __popc(s_bits[2 * t + 1]) : 0;
      int ex;
      const int tot = nms_block_scan(c0 + c1, &ex, s_w);
      if (2 * t < nwords) s_pre[2 * t] = ex;
      if (2 * t + 1 < nwords) s_pre[2 * t + 1] = ex + c0;
      __syncthreads();
      const int run = *s_run;
      for (int e = t; e < 2048; e += NMS_THREADS) {
        const int i = base + e;
        const int v = (i < n) ? mem_at(i) : -1;
        if (v >= 0 && strong_at(v)) {
          const int rank = s_pre[v >> 5] + __popc(s_bits[v >> 5] & ((1u << (v & 31)) - 1u));
          if (run + rank < cap) {
            out[run + rank] = emit(cand[v]);
            kidx[run + rank] = v;
          }
        }
      }
      __syncthreads();
      if (t == 0) *s_run = run + tot;
      __syncthreads();
    } else {
      for (int e = t; e < 2048; e += NMS_THREADS) {
        const int i = base + e;
        const int v = (i < n) ? mem_at(i) : -1;
        const bool ok = v >= 0 && strong_at(v);
        s_sort[e] = ok ? v : INT32_MAX;
      }
      __syncthreads();
      for (int k = 2; k <= 2048; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          const int a = ((t & ~(j - 1)) << 1) | (t & (j - 1));
          const int bb = a | j;
          const int va = s_sort[a], vb = s_sort[bb];
          const bool asc = (a & k) == 0;
          if ((va > vb) == asc) { s_sort[a] = vb; s_sort[bb] = va; }
          __syncthreads();
        }
      }
      const int cnt = __syncthreads_count((s_sort[t] != INT32_MAX) ? 1 : 0) +
                      __syncthreads_count((s_sort[t + 1024] != INT32_MAX) ? 1 : 0);
      const int run = *s_run;
      for (int e = t; e < 2048; e += NMS_THREADS) {
        const int v = s_sort[e];
        if (v != INT32_MAX && run + e < cap) {
          out[run + e] = emit(cand[v]);
          kidx[run + e] = v;
        }
      }
      __syncthreads();
      if (t == 0) *s_run = run + cnt;
      __syncthreads();
    }
  }
  return *s_run;
}

// The same output in one pass when every chunk's bitmap fits side by side
// (chunks x words <= NMS_BITMAP_WORDS): chunk c's bitmap follows chunk c-1's,
// so one prefix scan over all of them ranks every value at its position
// behind the earlier chunks' keypoints.  Returns the number of keypoints.
template <class MemAt, class StrongAt>
__device__ int nms_output_flat(const akz_kp* __restrict__ cand, int n, MemAt mem_at, StrongAt strong_at,
                               akz_kp* __restrict__ out, int32_t* __restrict__ kidx, int cap, uint32_t* s_bits,
                               int32_t* s_pre, int32_t* s_w) {
  const int t = threadIdx.x;
  const int nwords = (n + 31) >> 5, total = ((n + 2047) >> 11) * nwords;
  for (int i = t; i < total; i += NMS_THREADS) s_bits[i] = 0u;
  __syncthreads();
  for (int i = t; i < n; i += NMS_THREADS) {
    const int v = mem_at(i);
    if (v >= 0 && strong_at(v)) atomicOr(&s_bits[(i >> 11) * nwords + (v >> 5)], 1u << (v & 31));
  }
  __syncthreads();
  const int c0 = (2 * t < total) ? __popc(s_bits[2 * t]) : 0;
  const int c1 = (2 * t + 1 < total) ? __popc(s_bits[2 * t + 1]) : 0;
  int ex;
  const int kept = nms_block_scan(c0 + c1, &ex, s_w);
  if (2 * t < total) s_pre[2 * t] = ex;
  if (2 * t + 1 < total) s_pre[2 * t + 1] = ex + c0;
  __syncthreads();
  for (int i = t; i < n; i += NMS_THREADS) {
    const int v = mem_at(i);
    if (v >= 0 && strong_at(v)) {
      const int wi = (i >> 11) * nwords + (v >> 5);
      const int rank = s_pre[wi] + __popc(s_bits[wi] & ((1u << (v & 31)) - 1u));
      if (rank < cap) {
        out[rank] = emit(cand[v]);
        kidx[rank] = v;
      }
    }
  }
  return kept;
}

// Wave loop of the suppression with its state in global memory (more than
// NMS_LDS_N candidates, or neighbour lists too long for LDS), lists at
// (nb_off, nb_count): each wave visits a compacted list of the candidates
// still pending.  Returns the number of waves.
__device__ int nms_waves_global(const akz_det_bufs& b, const akz_kp* __restrict__ cand, int n, int32_t* s_nact) {
  const int t = threadIdx.x;
  const int32_t* __restrict__ nb_off = akz_zoff(b.nb_off, b.zs);
  const int32_t* __restrict__ nb_cnt = akz_zoff(b.nb_count, b.zs);
  const int32_t* __restrict__ nb_list = akz_zoff(b.nb_list, b.zs);
  int32_t* mem = akz_zoff(b.mem, b.zs);
  int32_t* minnb = akz_zoff(b.minnb, b.zs);
  int32_t* g_code = akz_zoff(b.code, b.zs);
  int32_t* act[2] = {akz_zoff(b.act, b.zs), akz_zoff(b.act, b.zs) + b.cap};
  for (int i = t; i < n; i += NMS_THREADS) {
    mem[i] = -1;
    minnb[i] = INT32_MAX;
    act[0][i] = i;
  }
  if (t == 0) { s_nact[0] = n; s_nact[1] = 0; }
  __syncthreads();
  int cur = 0, wave = 0;
  while (wave < b.max_waves) {
    const int na = s_nact[cur];
    if (na == 0) break;
    const int32_t* A = act[cur];
    for (int j = t; j < na; j += NMS_THREADS) {
      const int c = A[j];
      const int e0 = nb_off[c], e1 = e0 + nb_cnt[c];
      bool ready = true, anykept = false;
      int best = -1, bestv = INT32_MAX;
      for (int k = e0; k < e1; k++) {
        const int q = nb_list[k];
        const int m = mem[q];
        if (m == -1) { ready = false; break; }
        if (m >= 0) {
          anykept = true;
          if (m < bestv) { bestv = m; best = q; }
        }
      }
      int code = CODE_WAIT;
      if (ready) {
        if (!anykept) code = CODE_ADD;
        else { code = best; atomicMin(&minnb[best], c); }
      }
      g_code[c] = code;
    }
    __syncthreads();
    for (int j = t; j < na; j += NMS_THREADS) {
      const int c = A[j];
      const int code = g_code[c];
      if (code == CODE_ADD) {
        mem[c] = c;
      } else if (code >= 0 && minnb[code] == c) {
        if (cand[c].response > cand[code].response) {
          mem[c] = code;
          mem[code] = -2;
        } else {
          mem[c] = -2;
        }
      }
    }
    __syncthreads();
    const int nxt = cur ^ 1;
    for (int j = t; j < na; j += NMS_THREADS) {
      const int c = A[j];
      const int code = g_code[c];
      if (code >= 0) minnb[code] = INT32_MAX;
      if (mem[c] == -1) {
        int pos = atomicAdd(&s_nact[nxt], 1);
        act[nxt][pos] = c;
      }
    }
    __syncthreads();
    if (t == 0) s_nact[cur] = 0;
    cur = nxt;
    wave++;
    __syncthreads();
  }
  return wave;
}

// One wave's decision for a pending candidate with packed list f whose first
// two neighbours have member values m0 and m1 (m1 is unused for lists of one
// or of three and more): CODE_WAIT while a neighbour is pending, CODE_ADD
// when no neighbour is kept, else the kept neighbour with the smallest
// member value, first in list order on ties.
__device__ __forceinline__ int nms_classify(uint32_t f, int m0, int m1, const int16_t* s_mem,
                                            const int16_t* s_fext) {
  const int cnt = (int)(f >> 26);
  const int q0 = (int)(f & 0x1fffu), f2 = (int)((f >> 13) & 0x1fffu);
  bool ready = true;
  int best = -1, bestv = INT32_MAX;
  if (cnt >= 1) {
    if (m0 == -1) ready = false;
    else if (m0 >= 0) { bestv = m0; best = q0; }
  }
  if (cnt == 2) {
    if (m1 == -1) ready = false;
    else if (m1 >= 0 && m1 < bestv) { bestv = m1; best = f2; }
  } else if (cnt > 2) {
    for (int i = 0; i < cnt - 1; i++) {
      const int q = s_fext[f2 + i];
      const int m = s_mem[q];
      if (m == -1) ready = false;
      else if (m >= 0 && m < bestv) { bestv = m; best = q; }
    }
  }
  if (!ready) return CODE_WAIT;
  return best < 0 ? CODE_ADD : best;
}

// The second phase of a wave for candidate c with decision `code`: an
// addition, or for a contender that won its target's arbitration a
// replacement (response rc above the target's rq) or a drop.  Returns
// whether c is processed.
__device__ __forceinline__ bool nms_resolve(int c, int code, bool won, float rc, float rq, int16_t* s_mem) {
  if (code == CODE_ADD) {
    s_mem[c] = (int16_t)c;
    return true;
  }
  if (code >= 0 && won) {
    if (rc > rq) {
      s_mem[c] = (int16_t)code;
      s_mem[code] = -2;
    } else {
      s_mem[c] = -2;
    }
    return true;
  }
  return false;
}

constexpr int NMS_MID_N = 32768;    // candidates whose member values and pending list fit LDS as 16-bit values
constexpr int NMS_MID_CH = 8;       // list entries per round of staged loads
constexpr int NMS_MID_E = NMS_MID_N / NMS_THREADS;  // list entries per thread at most
static_assert(NMS_MID_N <= 32768 && NMS_MID_E % NMS_MID_CH == 0, "16-bit indices, whole rounds");

// The waves for NMS_LDS_N < n <= NMS_MID_N candidates: member values and the
// compacted list of pending candidates in LDS (16-bit), the neighbour lists
// of k_nb_lists read from global memory, arbitration by a 64-bit atomic
// minimum in global memory whose upper half carries the wave (as the LDS
// path's keys: no reset of the slots).  Phase 1 leaves each list entry's
// candidate index and code in one word of global memory (by list position,
// read back by the same thread in phase 2), so the registers stay free for
// the LDS path's waves; each round of NMS_MID_CH entries issues its loads
// together.
// Wave 0 is part of the set-up; the last waves (at most NMS_TAIL pending) run
// on one wavefront without barriers.  Returns the number of waves.
__device__ int nms_waves_mid(const akz_det_bufs& b, const akz_kp* __restrict__ cand, int n, int16_t* s_mem,
                             int16_t* s_list, int32_t* s_nact) {
  const int t = threadIdx.x, lane = t & 63;
  const int32_t* __restrict__ nb_off = akz_zoff(b.nb_off, b.zs);
  const int32_t* __restrict__ nb_count = akz_zoff(b.nb_count, b.zs);
  const int32_t* __restrict__ nb_list = akz_zoff(b.nb_list, b.zs);
  const uint2* __restrict__ nb_pack = akz_zoff(b.nb_pack, b.zs);
  unsigned long long* arb = akz_zoff(b.arb, b.zs);
  int32_t* g_pk = akz_zoff(b.code, b.zs);  // entry e: candidate | (code + 2) << 16, -1: none
  const bool wave0 = b.max_waves > 0 && n > 0;
  if (t == 0) s_nact[0] = 0;
  __syncthreads();
  for (int base = 0; base < n; base += NMS_THREADS) {
    const int c = base + t;
    bool pending = false;
    if (c < n) {
      const bool add = wave0 && nb_count[c] == 0;
      s_mem[c] = add ? (int16_t)c : (int16_t)-1;
      arb[c] = ~0ull;
      pending = !add;
    }
    const uint64_t bm = __ballot(pending);
    if (bm != 0ull) {
      const int leader = __ffsll((unsigned long long)bm) - 1;
      int o = 0;
      if (lane == leader) o = atomicAdd(&s_nact[0], __popcll(bm));
      o = __shfl(o, leader, 64);
      if (pending) s_list[o + __popcll(bm & ((1ull << lane) - 1ull))] = (int16_t)c;
    }
  }
  __syncthreads();
  int wave = wave0 ? 1 : 0;
  int cur = 0;
  int P = s_nact[0];
  // classification of candidate c (neighbour count cnt at nb_list[o]) whose
  // first two neighbours q0, q1 have member values m0, m1
  auto classify = [&](int cnt, int o, int q0, int q1, int m0, int m1) {
    bool ready = true;
    int best = -1, bestv = INT32_MAX;
    if (cnt >= 1) {
      if (m0 == -1) ready = false;
      else if (m0 >= 0) { bestv = m0; best = q0; }
    }
    if (cnt >= 2) {
      if (m1 == -1) ready = false;
      else if (m1 >= 0 && m1 < bestv) { bestv = m1; best = q1; }
    }
    for (int k = 2; k < cnt; k++) {
      const int q = nb_list[o + k];
      const int m = s_mem[q];
      if (m == -1) ready = false;
      else if (m >= 0 && m < bestv) { bestv = m; best = q; }
    }
    if (!ready) return CODE_WAIT;
    return best < 0 ? CODE_ADD : best;
  };
  while (wave < b.max_waves && P > NMS_TAIL) {
    const int nxt = cur ^ 1;
    if (t == 0) s_nact[nxt] = 0;
    const unsigned long long tag = (unsigned long long)(0xffffffffu - (unsigned)wave) << 32;
    const int ne = __builtin_amdgcn_readfirstlane(((t & ~63) < P) ? (P + NMS_THREADS - 1) / NMS_THREADS : 0);
#pragma unroll 1
    for (int r = 0; r < NMS_MID_E / NMS_MID_CH; r++) {
      if (r * NMS_MID_CH >= ne) break;
      int c[NMS_MID_CH], o[NMS_MID_CH], cn[NMS_MID_CH], q0[NMS_MID_CH], q1[NMS_MID_CH];
#pragma unroll
      for (int k = 0; k < NMS_MID_CH; k++) {
        const int e = t + (r * NMS_MID_CH + k) * NMS_THREADS;
        c[k] = e < P ? (int)s_list[e] : -1;
      }
      uint2 pw[NMS_MID_CH];
#pragma unroll
      for (int k = 0; k < NMS_MID_CH; k++) pw[k] = c[k] >= 0 ? nb_pack[c[k]] : make_uint2(0u, 0u);
#pragma unroll
      for (int k = 0; k < NMS_MID_CH; k++) {
        q0[k] = (int)(pw[k].x & 0x3fffffffu);
        q1[k] = (int)pw[k].y;
        cn[k] = (int)(pw[k].x >> 30);
        o[k] = 0;
        // lists of three or more: the count and start (rare)
        if (cn[k] == 3) { o[k] = nb_off[c[k]]; cn[k] = nb_count[c[k]]; }
      }
#pragma unroll
      for (int k = 0; k < NMS_MID_CH; k++) {
        int code = CODE_WAIT;
        if (c[k] >= 0) {
          code = classify(cn[k], o[k], q0[k], q1[k], s_mem[q0[k]], s_mem[q1[k]]);
          if (code >= 0) atomicMin(&arb[code], tag | (unsigned long long)c[k]);
        }
        const int e = t + (r * NMS_MID_CH + k) * NMS_THREADS;
        if (e < P) g_pk[e] = c[k] < 0 ? -1 : (c[k] | ((code + 2) << 16));
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int r = 0; r < NMS_MID_E / NMS_MID_CH; r++) {
      if (r * NMS_MID_CH >= ne) break;
      unsigned long long mn[NMS_MID_CH];
      float rc[NMS_MID_CH], rq[NMS_MID_CH];
      int pv[NMS_MID_CH];
#pragma unroll
      for (int k = 0; k < NMS_MID_CH; k++) {
        const int e = t + (r * NMS_MID_CH + k) * NMS_THREADS;
        pv[k] = e < P ? g_pk[e] : -1;
      }
#pragma unroll
      for (int k = 0; k < NMS_MID_CH; k++) {
        const int v = pv[k];
        const int c = v & 0xffff, code = (v >> 16) - 2;
        mn[k] = 0ull;
        rc[k] = rq[k] = 0.0f;
        if (v >= 0 && code >= 0) {
          mn[k] = __hip_atomic_load(&arb[code], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          rc[k] = cand[c].response;
          rq[k] = cand[code].response;
        }
      }
#pragma unroll
      for (int k = 0; k < NMS_MID_CH; k++) {
        const int v = pv[k];
        const int c = v & 0xffff, code = (v >> 16) - 2;
        bool pending = false;
        if (v >= 0) {
          const bool won = mn[k] == (tag | (unsigned long long)c);
          pending = !nms_resolve(c, code, won, rc[k], rq[k], s_mem);
        }
        const uint64_t bm = __ballot(pending);
        if (bm != 0ull) {
          const int leader = __ffsll((unsigned long long)bm) - 1;
          int o = 0;
          if (lane == leader) o = atomicAdd(&s_nact[nxt], __popcll(bm));
          o = __shfl(o, leader, 64);
          if (pending) s_list[o + __popcll(bm & ((1ull << lane) - 1ull))] = (int16_t)c;
        }
      }
    }
    __syncthreads();
    P = s_nact[nxt];
    cur = nxt;
    wave++;
  }
  // the last waves: wavefront 0, one pending candidate per lane
  if (t < 64 && wave < b.max_waves && P > 0) {
    static_assert(NMS_TAIL == 64, "one candidate per lane");
    int c = lane < P ? (int)s_list[lane] : -1;
    const uint2 pw = c >= 0 ? nb_pack[c] : make_uint2(0u, 0u);
    int o = 0, cn = (int)(pw.x >> 30);
    const int q0 = (int)(pw.x & 0x3fffffffu), q1 = (int)pw.y;
    if (cn == 3) { o = nb_off[c]; cn = nb_count[c]; }
    const float rc = c >= 0 ? cand[c].response : 0.0f;
    while (wave < b.max_waves && __ballot(c >= 0) != 0ull) {
      const unsigned long long tag = (unsigned long long)(0xffffffffu - (unsigned)wave) << 32;
      int code = CODE_WAIT;
      if (c >= 0) {
        code = classify(cn, o, q0, q1, s_mem[q0], s_mem[q1]);
        if (code >= 0) atomicMin(&arb[code], tag | (unsigned long long)c);
      }
      // every lane's atomic is done before the arbitration slots are read
      __threadfence_block();
      if (c >= 0 && code != CODE_WAIT) {
        unsigned long long mn = 0ull;
        float rq = 0.0f;
        if (code >= 0) {
          mn = __hip_atomic_load(&arb[code], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          rq = cand[code].response;
        }
        if (nms_resolve(c, code, mn == (tag | (unsigned long long)c), rc, rq, s_mem)) c = -1;
      }
      __builtin_amdgcn_wave_barrier();
      wave++;
    }
  }
  __syncthreads();
  return wave;
}

// Wave-synchronous greedy suppression of FilterExtrema_kernel
// (cuda_akaze.cu:1072-1224) and its output, one workgroup per image, on the
// neighbour lists of k_nb_lists.  Per wave: a candidate whose neighbours are
// all processed is "ready"; a ready candidate with no kept neighbour is kept;
// otherwise it contends for its kept neighbour with the smallest member
// value, the lowest contending index wins (atomicMin here, a racy min in the
// reference), and the winner replaces that neighbour if its response is
// higher, else is dropped.  No step of a wave depends on the order in which
// its candidates are visited.
//
// Up to NMS_LDS_N candidates everything is in LDS, the neighbour lists
// packed per candidate in 32 bits (the first neighbour, the second or the
// offset of the rest in a small pool, the count: almost every candidate has
// at most two).  Wave 0 is part of the set-up: nothing is processed before
// it, so exactly the candidates without neighbours are ready, and all of
// them are kept.  The candidates still pending are kept in a compacted list,
// so a wave costs work only for them.  While the list is longer than a
// wavefront, a wave is: list entries, packed lists and the neighbours'
// member values loaded in three rounds, classification and arbitration, a
// barrier, the arbitration results and responses loaded in one round,
// additions and replacements, the candidates left pending appended to the
// next list (one LDS atomic per wavefront), a barrier.  From then on one
// wavefront runs the remaining waves with one candidate per lane and no
// barriers: a wavefront's LDS instructions complete in order, so each step
// sees the previous one.  Beyond NMS_LDS_N candidates, or when a list does
// not fit the packed form or the pool, the state is in global memory.
__global__ __launch_bounds__(NMS_THREADS) void k_suppress(akz_det_bufs b) {
  // One LDS buffer in two layouts.  Up to NMS_LDS_N candidates: arbitration
  // slots, responses, packed neighbour lists, 16-bit member values, the
  // pending list (then the output's scratch), the pool of long lists and a
  // bitmap of strong candidates.  Up to NMS_MID_N: 16-bit member values and
  // the pending list (then the output's scratch).  Beyond: the output's
  // scratch.
  constexpr int O_MINNB = 0, O_RESP = O_MINNB + 4 * NMS_LDS_N, O_FWD = O_RESP + 4 * NMS_LDS_N;
  constexpr int O_MEM = O_FWD + 4 * NMS_LDS_N, O_LIST = O_MEM + 2 * NMS_LDS_N, O_FEXT = O_LIST + 2 * NMS_LDS_N;
  constexpr int O_STRONG = O_FEXT + 2 * NMS_FEXT, SMALL_BYTES = O_STRONG + NMS_LDS_N / 8;
  constexpr int MID_BYTES = 4 * NMS_MID_N;
  __shared__ __align__(16) unsigned char s_raw[SMALL_BYTES > MID_BYTES ? SMALL_BYTES : MID_BYTES];
  int32_t* s_minnb = reinterpret_cast<int32_t*>(s_raw + O_MINNB);
  float* s_resp = reinterpret_cast<float*>(s_raw + O_RESP);
  uint32_t* s_fwd = reinterpret_cast<uint32_t*>(s_raw + O_FWD);
  int16_t* s_mem = reinterpret_cast<int16_t*>(s_raw + O_MEM);
  int16_t* s_list = reinterpret_cast<int16_t*>(s_raw + O_LIST);
  int16_t* s_fext = reinterpret_cast<int16_t*>(s_raw + O_FEXT);
  uint32_t* s_strong = reinterpret_cast<uint32_t*>(s_raw + O_STRONG);
  __shared__ int32_t s_w[17];
  __shared__ int32_t s_nact[2];  // list lengths, by the parity of the wave
  __shared__ int32_t s_run;
  __shared__ int32_t s_ftop;     // pool entries taken
  __shared__ int32_t s_spill;    // a list does not fit the packed form or the pool
  const akz_kp* __restrict__ cand = akz_zoff(b.cand, b.zs);
  akz_kp* __restrict__ out = akz_zoff(b.kps, b.zs);
  int32_t* __restrict__ kidx = akz_zoff(b.kidx, b.zs);
  int32_t* st = akz_zoff(b.status, b.zs);
  const int cap = b.cap;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int n = *akz_zoff(b.ncand, b.zs);
  int wave = 0;
  int kept = 0;
  bool global_lists = n > NMS_LDS_N;
  if (t == 0) { s_ftop = 0; s_spill = 0; }
  if (!global_lists) {
    // the lists of k_nb_lists packed into LDS: counts and starts, then the
    // first two entries (one round of loads each); entries 1.. of lists of
    // three or more into the pool at space taken with one LDS atomic
    const int32_t* nb_off = akz_zoff(b.nb_off, b.zs);
    const int32_t* nb_count = akz_zoff(b.nb_count, b.zs);
    const int32_t* nb_list = akz_zoff(b.nb_list, b.zs);
    const uint2* nb_pack = akz_zoff(b.nb_pack, b.zs);
    int lo_[NMS_PER], lc_[NMS_PER], q0_[NMS_PER], q1_[NMS_PER];
    float rv[NMS_PER];
    bool sg[NMS_PER];
#pragma unroll
    for (int j = 0; j < NMS_PER; j++) {
      const int c = t + j * NMS_THREADS;
      uint2 pw = make_uint2(0u, 0u);
      rv[j] = 0.0f;
      sg[j] = false;
      if (c < n) {
        pw = nb_pack[c];
        rv[j] = cand[c].response;
        sg[j] = !(cand[c].size < 0);
      }
      q0_[j] = (int)(pw.x & 0x3fffffffu);
      q1_[j] = (int)pw.y;
      lc_[j] = (int)(pw.x >> 30);
      lo_[j] = 0;
    }
    // lists of three or more: the count and start (rare)
#pragma unroll
    for (int j = 0; j < NMS_PER; j++) {
      const int c = t + j * NMS_THREADS;
      if (lc_[j] == 3) { lo_[j] = nb_off[c]; lc_[j] = nb_count[c]; }
    }
    const bool wave0 = b.max_waves > 0 && n > 0;
    if (t == 0) s_nact[0] = 0;
    __syncthreads();
    // Wave 0 is part of the set-up: the candidates without neighbours are
    // kept, the others form the list of pending candidates.
#pragma unroll
    for (int j = 0; j < NMS_PER; j++) {
      if (j * NMS_THREADS >= n) break;
      const int c = t + j * NMS_THREADS;
      bool pending = false;
      if (c < n) {
        const int cnt = lc_[j];
        uint32_t f2 = (uint32_t)q1_[j];
        if (cnt >= 3) {
          int o = NMS_FEXT;
          if (cnt <= NMS_CNT_MAX) o = atomicAdd(&s_ftop, cnt - 1);
          if (o + cnt - 1 <= b.nms_pool) {
            for (int k = 1; k < cnt; k++) s_fext[o + k - 1] = (int16_t)nb_list[lo_[j] + k];
            f2 = (uint32_t)o;
          } else {
            s_spill = 1;
          }
        }
        s_fwd[c] = (uint32_t)q0_[j] | (f2 << 13) | ((uint32_t)akz_imin(cnt, NMS_CNT_MAX) << 26);
        const bool add = wave0 && cnt == 0;
        s_mem[c] = add ? (int16_t)c : (int16_t)-1;
        s_minnb[c] = INT32_MAX;
        s_resp[c] = rv[j];
        pending = !add;
      }
      // candidates c - lane .. c - lane + 63 are this wavefront's slot j
      const uint64_t sb = __ballot(sg[j]);
      if (lane == 0) {
        const int w0 = (c - lane) >> 5;
        s_strong[w0] = (uint32_t)sb;
        s_strong[w0 + 1] = (uint32_t)(sb >> 32);
      }
      const uint64_t bm = __ballot(pending);
      if (bm != 0ull) {
        const int leader = __ffsll((unsigned long long)bm) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&s_nact[0], __popcll(bm));
        base = __shfl(base, leader, 64);
        if (pending) s_list[base + __popcll(bm & ((1ull << lane) - 1ull))] = (int16_t)c;
      }
    }
    __syncthreads();
    // a list that does not fit the packed form or the pool: the whole
    // suppression on the lists in global memory
    if (s_spill) global_lists = true;
    if (!global_lists) {
      if (wave0) wave = 1;
      int cur = 0;
      int P = s_nact[0];
      // Arbitration key of a wave: every neighbour of a candidate has a
      // lower index, so the lowest pending candidate is always ready and
      // wins; a wave with work left therefore has wave < n <= NMS_LDS_N, and
      // keys of later waves are smaller than every key left from earlier
      // ones, which makes a reset of the slots unnecessary.
      while (wave < b.max_waves && P > NMS_TAIL) {
        const int nxt = cur ^ 1;
        // the other length was last read before the previous wave's barriers
        if (t == 0) s_nact[nxt] = 0;
        const int tag = (NMS_LDS_N - 1 - wave) << 13;
        // list entries per thread (uniform); wavefronts past the list skip
        // both phases
        const int ne = __builtin_amdgcn_readfirstlane(((t & ~63) < P) ? (P + NMS_THREADS - 1) / NMS_THREADS : 0);
        int cc[NMS_PER], cd[NMS_PER];
        uint32_t fw[NMS_PER];
        int m0[NMS_PER], m1[NMS_PER];
        // phase 1: entries, their packed lists, their neighbours' member
        // values (the second slot also for longer lists: an in-range pool
        // offset whose value is not used), one round each (the loops end at
        // the uniform ne: scalar branches skip the rest)
#pragma unroll
        for (int k = 0; k < NMS_PER; k++) {
          if (k >= ne) break;
          cc[k] = -1;
          if (t + k * NMS_THREADS < P) cc[k] = s_list[t + k * NMS_THREADS];
        }
#pragma unroll
        for (int k = 0; k < NMS_PER; k++) {
          if (k >= ne) break;
          fw[k] = cc[k] >= 0 ? s_fwd[cc[k]] : 0u;
        }
#pragma unroll
        for (int k = 0; k < NMS_PER; k++) {
          if (k >= ne) break;
          m0[k] = m1[k] = 0;
          if (cc[k] >= 0) {
            m0[k] = s_mem[fw[k] & 0x1fffu];
            m1[k] = s_mem[(fw[k] >> 13) & 0x1fffu];
          }
        }
        // classification in list order (smallest member value, first on
        // ties) and arbitration
#pragma unroll
        for (int k = 0; k < NMS_PER; k++) {
          if (k >= ne) break;
          cd[k] = CODE_WAIT;
          if (cc[k] < 0) continue;
          const int c = cc[k];
          const int code = nms_classify(fw[k], m0[k], m1[k], s_mem, s_fext);
          if (code >= 0) atomicMin(&s_minnb[code], tag | c);
          cd[k] = code;
        }
        __syncthreads();
        // phase 2: arbitration results and responses in one round, then
        // additions and replacements (targets are kept, never pending)
        int mn[NMS_PER];
        float rc[NMS_PER], rq[NMS_PER];
#pragma unroll
        for (int k = 0; k < NMS_PER; k++) {
          if (k >= ne) break;
          mn[k] = -1;
          rc[k] = rq[k] = 0.0f;
          if (cd[k] >= 0) {
            mn[k] = s_minnb[cd[k]];
            rc[k] = s_resp[cc[k]];
            rq[k] = s_resp[cd[k]];
          }
        }
#pragma unroll
        for (int k = 0; k < NMS_PER; k++) {
          if (k >= ne) break;
          const int c = cc[k];
          const bool pending = c >= 0 && !nms_resolve(c, cd[k], mn[k] == (tag | c), rc[k], rq[k], s_mem);
          // the candidates left pending go to the next list: one LDS atomic
          // per wavefront, lanes at their rank among the wavefront's
          const uint64_t bm = __ballot(pending);
          if (bm != 0ull) {
            const int leader = __ffsll((unsigned long long)bm) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(&s_nact[nxt], __popcll(bm));
            base = __shfl(base, leader, 64);
            if (pending) s_list[base + __popcll(bm & ((1ull << lane) - 1ull))] = (int16_t)c;
          }
        }
        __syncthreads();
        P = s_nact[nxt];
        cur = nxt;
        wave++;
      }
      // the last waves: wavefront 0, up to NMS_TAIL / 64 pending candidates
      // per lane, no barriers
      if (t < 64 && wave < b.max_waves && P > 0) {
        constexpr int TE = NMS_TAIL / 64;
        int c[TE];
        uint32_t f[TE];
#pragma unroll
        for (int u = 0; u < TE; u++) c[u] = lane + 64 * u < P ? (int)s_list[lane + 64 * u] : -1;
#pragma unroll
        for (int u = 0; u < TE; u++) f[u] = c[u] >= 0 ? s_fwd[c[u]] : 0u;
        auto any_left = [&]() {
          bool a = false;
#pragma unroll
          for (int u = 0; u < TE; u++) a |= c[u] >= 0;
          return __ballot(a) != 0ull;
        };
        while (wave < b.max_waves && any_left()) {
          const int tag = (NMS_LDS_N - 1 - wave) << 13;
          int m0[TE], m1[TE], code[TE];
#pragma unroll
          for (int u = 0; u < TE; u++) {
            m0[u] = m1[u] = 0;
            if (c[u] >= 0) { m0[u] = s_mem[f[u] & 0x1fffu]; m1[u] = s_mem[(f[u] >> 13) & 0x1fffu]; }
          }
#pragma unroll
          for (int u = 0; u < TE; u++) {
            code[u] = CODE_WAIT;
            if (c[u] >= 0) {
              code[u] = nms_classify(f[u], m0[u], m1[u], s_mem, s_fext);
              if (code[u] >= 0) atomicMin(&s_minnb[code[u]], tag | c[u]);
            }
          }
          // the atomics of every lane are done before these reads (in-order
          // LDS instructions of one wavefront)
          __builtin_amdgcn_wave_barrier();
          int mn[TE];
          float rc[TE], rq[TE];
#pragma unroll
          for (int u = 0; u < TE; u++) {
            mn[u] = -1;
            rc[u] = rq[u] = 0.0f;
            if (c[u] >= 0 && code[u] >= 0) { mn[u] = s_minnb[code[u]]; rc[u] = s_resp[c[u]]; rq[u] = s_resp[code[u]]; }
          }
#pragma unroll
          for (int u = 0; u < TE; u++)
            if (c[u] >= 0 && nms_resolve(c[u], code[u], mn[u] == (tag | c[u]), rc[u], rq[u], s_mem)) c[u] = -1;
          __builtin_amdgcn_wave_barrier();
          wave++;
        }
      }
      __syncthreads();
      // n <= NMS_LDS_N: at most 4 chunks of 256 words, side by side
      static_assert(((NMS_LDS_N + 2047) / 2048) * (NMS_LDS_N / 32) <= NMS_BITMAP_WORDS, "chunk bitmaps fit");
      static_assert(2 * NMS_BITMAP_WORDS * 4 <= NMS_LDS_N * 2, "output scratch fits the list");
      int32_t* scr = reinterpret_cast<int32_t*>(s_list);
      kept = nms_output_flat(
          cand, n, [&](int i) { return (int)s_mem[i]; },
          [&](int v) { return ((s_strong[v >> 5] >> (v & 31)) & 1u) != 0; }, out, kidx, cap,
          reinterpret_cast<uint32_t*>(scr), scr + NMS_BITMAP_WORDS, s_w);
    }
  }
  if (global_lists && n <= NMS_MID_N) {
    int16_t* m16 = reinterpret_cast<int16_t*>(s_raw);
    int16_t* l16 = m16 + NMS_MID_N;
    wave = nms_waves_mid(b, cand, n, m16, l16, s_nact);
    int32_t* scr = reinterpret_cast<int32_t*>(l16);
    static_assert(2048 + 2 * NMS_BITMAP_WORDS <= NMS_MID_N / 2, "output scratch fits the list");
    kept = nms_output(
        cand, n, [&](int i) { return (int)m16[i]; }, [&](int v) { return !(cand[v].size < 0); }, out, kidx, cap,
        scr, reinterpret_cast<uint32_t*>(scr) + 2048, scr + 2048 + NMS_BITMAP_WORDS, s_w, &s_run);
  } else if (global_lists) {
    const int32_t* mem = akz_zoff(b.mem, b.zs);
    int32_t* scr = reinterpret_cast<int32_t*>(s_raw);
    static_assert(4 * (2048 + 2 * NMS_BITMAP_WORDS) <= SMALL_BYTES, "output scratch fits");
    wave = nms_waves_global(b, cand, n, s_nact);
    kept = nms_output(
        cand, n, [&](int i) { return mem[i]; }, [&](int v) { return !(cand[v].size < 0); }, out, kidx, cap, scr,
        reinterpret_cast<uint32_t*>(scr) + 2048, scr + 2048 + NMS_BITMAP_WORDS, s_w, &s_run);
  }
  __syncthreads();
  if (t == 0) {
    const int flags = st[AKZ_ST_FLAGS];
    st[AKZ_ST_WAVES] = wave;
    st[AKZ_ST_KEPT] = akz_imin(kept, cap);
    // the whole block (the total comes from the candidate kernel) into the
    // contiguous per-batch array the host reads with one copy
    int4 blk = make_int4(akz_imin(kept, cap), flags, st[AKZ_ST_TOTAL], wave);
    reinterpret_cast<int4*>(b.status_all)[blockIdx.z] = blk;
  }
}

constexpr int NB_LIST_BLOCKS = 64;  // k_nb_lists' grid: 16384 threads walk the candidates

}  // namespace

int akz_suppression_pool_max() { return NMS_FEXT; }
int akz_scan_blocks(int len) { return (len + 1 + SCAN_BLK - 1) / SCAN_BLK; }
int akz_seg_runs(int seg_total) { return seg_total / SS_T + 1; }

void akz_launch_hessian(hipStream_t s, const akz_det_params& p, int nz) {
  const int nb = p.tile_start[p.nlev];
  hipLaunchKernelGGL(k_deriv, dim3(nb, 1, nz), dim3(256), 0, s, p);
  hipLaunchKernelGGL(k_hess, dim3(nb, 1, nz), dim3(256), 0, s, p);
}

int akz_detect_max_step() { return DF_SMAX; }

int akz_detect_tile_rows() { return DF_TH; }

void akz_launch_detect_fused(hipStream_t s, const akz_det_params& p32, const akz_det_bufs& b, int nz) {
  hipLaunchKernelGGL(k_detect_fused, dim3(p32.tile_start[p32.nlev], 1, nz), dim3(512), 0, s, p32, b.seg_count,
                     b.seg_slots);
}

void akz_launch_detect_fused_list(hipStream_t s, const akz_det_worklist& wl, int blocks) {
  hipLaunchKernelGGL(k_detect_fused_list, dim3(blocks), dim3(512), 0, s, wl);
}

void akz_launch_extrema(hipStream_t s, const akz_det_params& p, const akz_det_bufs& b, int nz) {
  hipLaunchKernelGGL(k_extrema, dim3(p.tile_start[p.nlev], 1, nz), dim3(256), 0, s, p, b.seg_count, b.seg_slots);
}

void akz_launch_candidates(hipStream_t s, const akz_det_params& p, const akz_det_bufs& b, int nz) {
  hipLaunchKernelGGL(k_seg_sums, dim3(akz_scan_blocks(p.seg_total), 1, nz), dim3(SCAN_T), 0, s, b.seg_count,
                     p.seg_total, b.seg_wsum, b.zs);
  hipLaunchKernelGGL(k_scan_scatter, dim3(akz_seg_runs(p.seg_total), 1, nz), dim3(SS_T), 0, s, p, b);
}

void akz_launch_suppression(hipStream_t s, const akz_det_bufs& b, int nz) {
  hipLaunchKernelGGL(k_nb_lists, dim3(NB_LIST_BLOCKS, 1, nz), dim3(256), 0, s, b);
  hipLaunchKernelGGL(k_suppress, dim3(1, 1, nz), dim3(NMS_THREADS), 0, s, b);
}
