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

// Exclusive scans over per-image count arrays in two launches: block totals
// (k_xscan_reduce), then each block adds the totals of the blocks before it
// and scans its 4096 counts (k_xscan_down).  Counts at indices >= nvalid
// count as zero, where nvalid is *n_ptr when given and `len` otherwise;
// offsets are written for indices 0..nvalid, so off[nvalid] is the total.
// Each thread owns four consecutive counts, read as one int4.
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

__global__ __launch_bounds__(SCAN_T) void k_xscan_reduce(const int32_t* __restrict__ cnt, int len, const int32_t* n_ptr,
                                                         int32_t* __restrict__ bsum, size_t zs) {
  __shared__ int32_t s_w[17];
  cnt = akz_zoff(cnt, zs);
  bsum = akz_zoff(bsum, zs);
  const int nvalid = n_ptr ? *akz_zoff(n_ptr, zs) : len;
  // the grid covers the capacity; blocks past the device-side count sum to 0
  if ((int)blockIdx.x * SCAN_BLK > nvalid) {
    if (threadIdx.x == 0) bsum[blockIdx.x] = 0;
    return;
  }
  const int i = blockIdx.x * SCAN_BLK + threadIdx.x * SCAN_PER;
  int4 v = load4(cnt, i, nvalid);
  int ex;
  int tot = block_exclusive_scan(v.x + v.y + v.z + v.w, &ex, s_w);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_T) void k_xscan_down(const int32_t* __restrict__ cnt, int len, const int32_t* n_ptr,
                                                       const int32_t* __restrict__ bsum, int32_t* __restrict__ off,
                                                       size_t zs) {
  __shared__ int32_t s_w[17];
  cnt = akz_zoff(cnt, zs);
  bsum = akz_zoff(bsum, zs);
  off = akz_zoff(off, zs);
  const int nvalid = n_ptr ? *akz_zoff(n_ptr, zs) : len;
  if ((int)blockIdx.x * SCAN_BLK > nvalid) return;  // no offset to write
  const int i = blockIdx.x * SCAN_BLK + threadIdx.x * SCAN_PER;
  int4 v = load4(cnt, i, nvalid);
  // the block's carry: the totals of the blocks before it (a few hundred at
  // most), summed here instead of in a separate launch
  int part = 0;
  for (int j = threadIdx.x; j < (int)blockIdx.x; j += SCAN_T) part += bsum[j];
  int ex;
  const int carry = block_exclusive_scan(part, &ex, s_w);
  block_exclusive_scan(v.x + v.y + v.z + v.w, &ex, s_w);
  const int o0 = carry + ex;
  const int o[4] = {o0, o0 + v.x, o0 + v.x + v.y, o0 + v.x + v.y + v.z};
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (i + k <= nvalid) off[i + k] = o[k];
}

// The c candidates of row segment `seg` (ranks in seg_slots) in keypoint
// layout at cand[base ...] (FindExtrema, cuda_akaze.cu:744-772): the subpixel
// fit on the determinant's 3x3 neighbourhood; slots at or past the capacity
// are dropped.
__device__ __forceinline__ void scatter_segment(const akz_det_params& p, const akz_det_bufs& b, int seg, int c,
                                                int base) {
  int l = 0;
  while (l + 1 < p.nlev && seg >= p.lev[l + 1].seg_base) l++;
  const akz_level_desc& L = p.lev[l];
  const int ls = seg - L.seg_base;
  const int y = ls / L.nseg_x, tx = ls - y * L.nseg_x;
  const uint8_t* slots = akz_zoff(b.seg_slots, b.zs);
  akz_kp* cand = akz_zoff(b.cand, b.zs);
  const float* imd = akz_zoff(p.det[l], p.zs);
  const size_t pitch = L.pitch;
  for (int r = 0; r < c; r++) {
    const int idx = base + r;
    if (idx >= b.cap) return;
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
// order is (level, y, x).  The thread of a level's first segment writes the
// level's start; the thread of index seg_total (the grid covers it) writes
// the total, the candidate count (cut at the capacity), and clears the
// overflow flags of this detection run.
constexpr int SS_T = 256;
__global__ __launch_bounds__(SS_T) void k_scan_scatter(akz_det_params p, akz_det_bufs b) {
  __shared__ int32_t s_w[17];
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
  }
  if (c > 0) scatter_segment(p, b, seg, c, o);
}

// Neighbour lists of FindNeighbors (cuda_akaze.cu:1005-1070): earlier
// candidates of the same level and all candidates of the previous level
// within half the candidate's size.  Weak candidates (negative size) get
// none, as the reference's early break gives.  The previous level is entered
// at the last candidate whose y is within reach (binary search) instead of
// being walked from its end.  FILL = false counts, FILL = true writes the
// CSR list at the scanned offsets.  Entries past nb_cap are dropped and flag
// an overflow; the suppression reads the lists cut at nb_cap, so every entry
// it reads was written by this run.
template <bool FILL>
__global__ __launch_bounds__(256) void k_neighbors(akz_det_bufs b) {
  const akz_kp* __restrict__ cand = akz_zoff(b.cand, b.zs);
  const int32_t* __restrict__ level_start = akz_zoff(b.level_start, b.zs);
  int32_t* nb_count = akz_zoff(b.nb_count, b.zs);
  const int32_t* nb_off = akz_zoff(b.nb_off, b.zs);
  int32_t* nb_list = akz_zoff(b.nb_list, b.zs);
  const int n = *akz_zoff(b.ncand, b.zs);
  for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256) {
    const akz_kp k = cand[c];
    const int lvl = k.class_id;
    const int start = level_start[lvl];
    const float size = k.size;
    const float half = size * .5f;
    const double r2 = (double)(size * size) * 0.25;
    int cnt = 0;
    const int o = FILL ? nb_off[c] : 0;
    for (int i = c - 1; i >= start; i--) {
      const akz_kp q = cand[i];
      if (k.y - q.y > half) break;
      float dist = (k.x - q.x) * (k.x - q.x) + (k.y - q.y) * (k.y - q.y);
      if ((double)dist < r2) {
        if (FILL && o + cnt < b.nb_cap) nb_list[o + cnt] = i;
        cnt++;
      }
    }
    if (lvl > 0) {
      const int lo = level_start[lvl - 1], hi = start;
      // first index in [lo, hi) with q.y - k.y > half (monotone in sorted order)
      int a = lo, e = hi;
      while (a < e) {
        int m = (a + e) >> 1;
        if (cand[m].y - k.y > half) e = m; else a = m + 1;
      }
      for (int i = a - 1; i >= lo; i--) {
        const akz_kp q = cand[i];
        if (k.y - q.y > half) break;
        float dist = (k.x - q.x) * (k.x - q.x) + (k.y - q.y) * (k.y - q.y);
        if ((double)dist < r2) {
          if (FILL && o + cnt < b.nb_cap) nb_list[o + cnt] = i;
          cnt++;
        }
      }
    }
    if (!FILL) nb_count[c] = cnt;
    else if (o + cnt > b.nb_cap) atomicOr(akz_zoff(b.status, b.zs) + AKZ_ST_FLAGS, 1);
  }
}

constexpr int NMS_THREADS = 1024;
constexpr int NMS_LDS_N = 8192;     // candidates whose suppression state fits LDS
constexpr int NMS_LDS_NB = 16384;   // neighbour entries that fit LDS beside it
constexpr int CODE_WAIT = -1, CODE_ADD = -2;
constexpr int NMS_BITMAP_WORDS = 2048;  // bitmap ranking covers up to 65536 candidates

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

// Wave-synchronous greedy suppression of FilterExtrema_kernel
// (cuda_akaze.cu:1072-1224), one workgroup per image.  Per wave: a candidate
// whose neighbours are all processed is "ready"; a ready candidate with no
// kept neighbour is kept; otherwise it contends for its kept neighbour with
// the smallest member value, the lowest contending index wins (atomicMin
// here, a racy min in the reference), and the winner replaces that neighbour
// if its response is higher, else is dropped.  No step of a wave depends on
// the order in which its candidates are visited.
//
// Up to NMS_LDS_N candidates with at most NMS_LDS_NB neighbour entries, the
// whole state lives in LDS (16-bit member values, codes and offsets, the CSR
// lists and the responses) and every wave visits all candidates, skipping
// the processed ones: a wave costs three barriers and LDS reads only.  Beyond
// that the state is in global memory and each wave visits a compacted list
// of the unprocessed candidates.
__global__ __launch_bounds__(NMS_THREADS) void k_nms(akz_det_bufs b) {
  __shared__ int16_t s_mem[NMS_LDS_N];
  __shared__ int32_t s_minnb[NMS_LDS_N];
  __shared__ uint16_t s_off[NMS_LDS_N + 2];
  __shared__ float s_resp[NMS_LDS_N];
  __shared__ uint8_t s_strong[NMS_LDS_N];  // 1: size >= 0 (emitted when kept)
  __shared__ __align__(16) int16_t s_nb[NMS_LDS_NB];  // the output phase reuses it (24 KB)
  __shared__ int32_t s_w[17];
  __shared__ int32_t s_nact[2];
  __shared__ int32_t s_left[2];  // work left after a wave, by the parity of the wave
  __shared__ int32_t s_run;
  int32_t* s_sort = reinterpret_cast<int32_t*>(s_nb);
  uint32_t* s_bits = reinterpret_cast<uint32_t*>(s_nb) + 2048;
  int32_t* s_pre = reinterpret_cast<int32_t*>(s_nb) + 2048 + NMS_BITMAP_WORDS;
  const akz_kp* __restrict__ cand = akz_zoff(b.cand, b.zs);
  const int32_t* __restrict__ nb_off = akz_zoff(b.nb_off, b.zs);
  const int32_t* __restrict__ nb_list = akz_zoff(b.nb_list, b.zs);
  akz_kp* __restrict__ out = akz_zoff(b.kps, b.zs);
  int32_t* __restrict__ kidx = akz_zoff(b.kidx, b.zs);
  int32_t* st = akz_zoff(b.status, b.zs);
  const int cap = b.cap;
  const int t = threadIdx.x;
  const int n = *akz_zoff(b.ncand, b.zs);
  const int ntot = n > 0 ? akz_imin(nb_off[n], b.nb_cap) : 0;
  const bool lds = n <= NMS_LDS_N && ntot <= NMS_LDS_NB;
  int wave = 0;
  int kept;
  if (lds) {
    // thread t owns candidates t + j * NMS_THREADS; their pending bits, the
    // readiness-scan resume points, list ends and codes stay in registers
    constexpr int PER = NMS_LDS_N / NMS_THREADS;
    int rs[PER], e1[PER], cd[PER];
    uint32_t pend = 0;
    // every global load of a thread is issued before its LDS stores
    {
      float rv[PER], sz[PER];
      int ov[PER + 1];
#pragma unroll
      for (int j = 0; j < PER; j++) {
        const int i = t + j * NMS_THREADS;
        if (i < n) { rv[j] = cand[i].response; sz[j] = cand[i].size; }
      }
#pragma unroll
      for (int j = 0; j <= PER; j++) {
        const int i = t + j * NMS_THREADS;
        if (i <= n) ov[j] = nb_off[i];
      }
      constexpr int PER4 = NMS_LDS_NB / 4 / NMS_THREADS;
      int4 lv[PER4];
      const int n4 = (ntot + 3) >> 2;
#pragma unroll
      for (int j = 0; j < PER4; j++) {
        const int k4 = t + j * NMS_THREADS;
        if (4 * k4 + 3 < ntot) {
          lv[j] = reinterpret_cast<const int4*>(nb_list)[k4];
        } else if (k4 < n4) {
          lv[j].x = nb_list[4 * k4];
          lv[j].y = 4 * k4 + 1 < ntot ? nb_list[4 * k4 + 1] : 0;
          lv[j].z = 4 * k4 + 2 < ntot ? nb_list[4 * k4 + 2] : 0;
          lv[j].w = 0;
        }
      }
#pragma unroll
      for (int j = 0; j < PER; j++) {
        const int i = t + j * NMS_THREADS;
        if (i < n) {
          s_mem[i] = -1;
          s_minnb[i] = INT32_MAX;
          s_resp[i] = rv[j];
          s_strong[i] = !(sz[j] < 0);
        }
      }
#pragma unroll
      for (int j = 0; j <= PER; j++) {
        const int i = t + j * NMS_THREADS;
        if (i <= n) s_off[i] = (uint16_t)akz_imin(ov[j], b.nb_cap);
        if (j < PER) {
          if (i < n) {
            rs[j] = akz_imin(ov[j], b.nb_cap);
            pend |= 1u << j;
          } else {
            rs[j] = 0;
          }
          cd[j] = CODE_WAIT;
        }
      }
#pragma unroll
      for (int j = 0; j < PER4; j++) {
        const int k4 = t + j * NMS_THREADS;
        if (k4 < n4) {
          // 4 * k4 + 3 < 4 * n4 <= NMS_LDS_NB: inside s_nb
          s_nb[4 * k4] = (int16_t)lv[j].x;
          s_nb[4 * k4 + 1] = (int16_t)lv[j].y;
          s_nb[4 * k4 + 2] = (int16_t)lv[j].z;
          s_nb[4 * k4 + 3] = (int16_t)lv[j].w;
        }
      }
    }
    __syncthreads();
    // list ends: the next candidate's start
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const int i = t + j * NMS_THREADS;
      e1[j] = i < n ? (int)s_off[i + 1] : 0;
    }
    bool any = n > 0;
    while (wave < b.max_waves && any) {
      // the flag of this wave's parity was last read before the previous
      // wave's first barrier; thread 0 clears it, wavefronts with work left
      // set it after this wave's second phase
      if (t == 0) s_left[wave & 1] = 0;
      // Arbitration key of this wave: every neighbour of a candidate has a
      // lower index, so the lowest pending candidate is always ready and
      // wins; a wave with work left therefore has wave < n <= NMS_LDS_N and
      // keys of later waves are smaller than every key left from earlier
      // ones, which makes a reset of the slots unnecessary.
      const int tag = (NMS_LDS_N - 1 - wave) << 13;
      // phase 1: classify ready candidates and arbitrate contention
#pragma unroll
      for (int j = 0; j < PER; j++) {
        if (!(pend & (1u << j))) continue;
        const int c = t + j * NMS_THREADS;
        // a processed neighbour never becomes pending again: resume the
        // readiness scan at the first neighbour that was pending
        // (two neighbours per step: both loads in flight together)
        int k = rs[j];
        const int e = e1[j];
        for (; k + 1 < e; k += 2) {
          const int ma = s_mem[s_nb[k]], mb = s_mem[s_nb[k + 1]];
          if (ma == -1) break;
          if (mb == -1) { k++; break; }
        }
        if (k == e - 1 && s_mem[s_nb[k]] != -1) k = e;
        rs[j] = k;
        int code = CODE_WAIT;
        if (k == e) {
          // the kept neighbour with the smallest member value, first in
          // list order on ties; four neighbours per step
          int best = -1, bestv = INT32_MAX;
          int q0 = s_off[c];
          for (; q0 + 3 < e; q0 += 4) {
            const int qa = s_nb[q0], qb = s_nb[q0 + 1], qc = s_nb[q0 + 2], qd = s_nb[q0 + 3];
            const int ma = s_mem[qa], mb = s_mem[qb], mc = s_mem[qc], md = s_mem[qd];
            if (ma >= 0 && ma < bestv) { bestv = ma; best = qa; }
            if (mb >= 0 && mb < bestv) { bestv = mb; best = qb; }
            if (mc >= 0 && mc < bestv) { bestv = mc; best = qc; }
            if (md >= 0 && md < bestv) { bestv = md; best = qd; }
          }
          for (; q0 < e; q0++) {
            const int q = s_nb[q0];
            const int m = s_mem[q];
            if (m >= 0 && m < bestv) { bestv = m; best = q; }
          }
          if (best < 0) code = CODE_ADD;
          else { code = best; atomicMin(&s_minnb[best], tag | c); }
        }
        cd[j] = code;
      }
      __syncthreads();
      // phase 2: additions and replacements (targets are kept, never pending)
#pragma unroll
      for (int j = 0; j < PER; j++) {
        if (!(pend & (1u << j))) continue;
        const int c = t + j * NMS_THREADS;
        const int code = cd[j];
        if (code == CODE_ADD) {
          s_mem[c] = (int16_t)c;
          pend &= ~(1u << j);
        } else if (code >= 0 && s_minnb[code] == (tag | c)) {
          if (s_resp[c] > s_resp[code]) {
            s_mem[c] = (int16_t)code;
            s_mem[code] = -2;
          } else {
            s_mem[c] = -2;
          }
          pend &= ~(1u << j);
        }
      }
      const bool wave_left = __ballot(pend != 0) != 0;
      if (wave_left && (t & 63) == 0) s_left[wave & 1] = 1;
      __syncthreads();
      any = s_left[wave & 1] != 0;
      wave++;
    }
    // n <= NMS_LDS_N: at most 4 chunks of 256 words, side by side in s_bits
    static_assert(((NMS_LDS_N + 2047) / 2048) * (NMS_LDS_N / 32) <= NMS_BITMAP_WORDS, "chunk bitmaps fit");
    kept = nms_output_flat(
        cand, n, [&](int i) { return (int)s_mem[i]; }, [&](int v) { return s_strong[v] != 0; }, out, kidx, cap,
        s_bits, s_pre, s_w);
  } else {
    int32_t* mem = akz_zoff(b.mem, b.zs);
    int32_t* minnb = akz_zoff(b.minnb, b.zs);
    int32_t* g_code = akz_zoff(b.code, b.zs);
    int32_t* act[2] = {akz_zoff(b.act, b.zs), akz_zoff(b.act, b.zs) + cap};
    for (int i = t; i < n; i += NMS_THREADS) {
      mem[i] = -1;
      minnb[i] = INT32_MAX;
      act[0][i] = i;
    }
    if (t == 0) { s_nact[0] = n; s_nact[1] = 0; }
    __syncthreads();
    int cur = 0;
    while (wave < b.max_waves) {
      const int na = s_nact[cur];
      if (na == 0) break;
      const int32_t* A = act[cur];
      for (int j = t; j < na; j += NMS_THREADS) {
        const int c = A[j];
        const int e0 = akz_imin(nb_off[c], b.nb_cap), e1 = akz_imin(nb_off[c + 1], b.nb_cap);
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
    kept = nms_output(
        cand, n, [&](int i) { return mem[i]; }, [&](int v) { return !(cand[v].size < 0); }, out, kidx, cap, s_sort,
        s_bits, s_pre, s_w, &s_run);
  }
  if (t == 0) {
    st[AKZ_ST_WAVES] = wave;
    st[AKZ_ST_KEPT] = akz_imin(kept, cap);
    // the whole block (flags and total come from earlier kernels) into the
    // contiguous per-batch array the host reads with one copy
    int4 blk = make_int4(akz_imin(kept, cap), st[AKZ_ST_FLAGS], st[AKZ_ST_TOTAL], wave);
    reinterpret_cast<int4*>(b.status_all)[blockIdx.z] = blk;
  }
}

inline int grid_stride_blocks(int cap) {
  int b = (cap + 255) / 256;
  return b < 2048 ? (b > 0 ? b : 1) : 2048;
}

void xscan(hipStream_t s, const int32_t* cnt, int len, const int32_t* n_ptr, int32_t* bsum, int32_t* off, int nz,
           size_t zs) {
  // grid covers indices 0..len (len + 1 offsets)
  const int nblk = (len + 1 + SCAN_BLK - 1) / SCAN_BLK;
  hipLaunchKernelGGL(k_xscan_reduce, dim3(nblk, 1, nz), dim3(SCAN_T), 0, s, cnt, len, n_ptr, bsum, zs);
  hipLaunchKernelGGL(k_xscan_down, dim3(nblk, 1, nz), dim3(SCAN_T), 0, s, cnt, len, n_ptr, bsum, off, zs);
}

}  // namespace

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
  const int g = grid_stride_blocks(b.cap);
  hipLaunchKernelGGL(k_neighbors<false>, dim3(g, 1, nz), dim3(256), 0, s, b);
  xscan(s, b.nb_count, b.cap, b.ncand, b.nb_bsum, b.nb_off, nz, b.zs);
  hipLaunchKernelGGL(k_neighbors<true>, dim3(g, 1, nz), dim3(256), 0, s, b);
  hipLaunchKernelGGL(k_nms, dim3(1, 1, nz), dim3(NMS_THREADS), 0, s, b);
}
