/*
 * akaze_oracle.c -- CPU restatement of the A-KAZE hot path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library, and only as the reference the results are compared with.  The
 * product (akaze_amd/) never links it.
 *
 * Each function restates the semantics of the reference CUDA fork of libAKAZE;
 * citations are file:line in that repository.  Operand order inside every
 * expression follows the reference, and the file is compiled with
 * -ffp-contract=off, so the result of every stage is a deterministic function
 * of its inputs (no FMA contraction, no reassociation).
 *
 * Decisions where the reference is racy or leaves values unset:
 *  - NMS arbitration (cuda_akaze.cu:1170, a non-atomic min) is "lowest
 *    candidate index wins".
 *  - Orientation bin sums (cuda_akaze.cu:2139-2140, float atomics) are summed
 *    in sample order tx = 0..207.
 *  - MLDB cell sums follow the reference's own fixed order: 64 strided
 *    partials then the pairwise tree of cuda_akaze.cu:1501-1522.
 *  - Matcher ties are lexicographic in (distance, trainIdx) (the reference's
 *    tie order depends on thread slots, cuda_akaze.cu:1873-1919).
 *  - atan2/sin/cos are the Cephes single-precision methods (S. L. Moshier,
 *    "Methods and Programs for Mathematical Functions", 1989) written out
 *    below, so the same expressions can be evaluated on any IEEE machine.
 *  - Orientation and MLDB sample coordinates are clamped into the level; for
 *    keypoints inside the detector border (AKAZE.cpp:353-363) the clamp never
 *    changes a coordinate.
 */
#define _DEFAULT_SOURCE /* random_r / initstate_r: glibc's rand() with private state */
#include "akaze_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline float bits2f(int32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline int32_t f2bits(float f) { int32_t b; memcpy(&b, &f, 4); return b; }
/* fRound, AKAZE.h:175-177 */
static inline int fround(float f) { return (int)(f + 0.5f); }

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

void orc_options_default(orc_options* o) {
  /* AKAZEConfig.h:86-111 */
  o->omax = 4;
  o->nsublevels = 4;
  o->soffset = 1.6f;
  o->derivative_factor = 1.5f;
  o->diffusivity = 1;
  o->dthreshold = 0.001f;
  o->min_dthreshold = 0.00001f;
  o->descriptor_pattern_size = 10;
  o->kcontrast_percentile = 0.7f;
  o->kcontrast_nbins = 300;
  o->nms_max_waves = 10000;
  o->descriptor = 5;
  o->descriptor_size = 0;
  o->descriptor_channels = 3;
}

/* ------------------------------------------------------------------------ */
/* FED schedule: fed.cpp:41-148                                              */
/* ------------------------------------------------------------------------ */
int orc_fed_is_prime_internal(int number) {
  /* fed.cpp:122-148 */
  if (number <= 1) return 0;
  if (number == 2 || number == 3 || number == 5 || number == 7) return 1;
  if ((number % 2) == 0 || (number % 3) == 0 || (number % 5) == 0 || (number % 7) == 0) return 0;
  int upper = (int)sqrt(number + 1.0);
  for (int d = 11; d <= upper; d += 2)
    if (number % d == 0) return 0;
  return 1;
}

int orc_fed_tau_internal(int n, float scale, float tau_max, int reordering, float* tau, int cap) {
  /* fed.cpp:63-119 */
  if (n <= 0) return 0;
  if (n > cap || n > ORC_MAX_TAU) return -1;
  float tauh[ORC_MAX_TAU];
  float c = 1.0f / (4.0f * (float)n + 2.0f);
  float d = scale * tau_max / 2.0f;
  for (int k = 0; k < n; ++k) {
    float h = (float)cos(ORC_PI * (2.0f * (float)k + 1.0f) * c);
    if (reordering)
      tauh[k] = d / (h * h);
    else
      tau[k] = d / (h * h);
  }
  if (reordering) {
    /* kappa-cycle permutation modulo the next prime above n (fed.cpp:95-116).
     * For n == 1 the reference reads tauh[-1]; the only step is used instead. */
    if (n == 1) {
      tau[0] = tauh[0];
      return 1;
    }
    int kappa = n / 2;
    int prime = n + 1;
    while (!orc_fed_is_prime_internal(prime)) prime++;
    for (int k = 0, l = 0; l < n; ++k, ++l) {
      int index;
      while ((index = ((k + 1) * kappa) % prime - 1) >= n) k++;
      tau[l] = tauh[index];
    }
  }
  return n;
}

int orc_fed_tau_by_cycle_time(float t, float tau_max, int reordering, float* tau, int cap) {
  /* fed.cpp:48-60 */
  int n = (int)(ceil(sqrt(3.0 * t / tau_max + 0.25f) - 0.5f - 1.0e-8f) + 0.5f);
  float scale = 3.0 * t / (tau_max * (float)(n * (n + 1)));
  return orc_fed_tau_internal(n, scale, tau_max, reordering, tau, cap);
}

int orc_fed_tau_by_process_time(float T, int M, float tau_max, int reordering, float* tau, int cap) {
  /* fed.cpp:41-45 */
  return orc_fed_tau_by_cycle_time(T / (float)M, tau_max, reordering, tau, cap);
}

/* ------------------------------------------------------------------------ */
/* Level geometry: AKAZE.cpp:146-197 (levels, FED), 322-364 (detector)       */
/* ------------------------------------------------------------------------ */
int orc_geometry(const orc_options* o, int W, int H, orc_geom* g) {
  memset(g, 0, sizeof(*g));
  int omax = o->omax;
  int nlev = 0;
  for (int i = 0; i <= omax - 1; i++) {
    float rfactor = (float)(1.0 / pow(2.0, (double)i));
    int lh = (int)(H * rfactor);
    int lw = (int)(W * rfactor);
    if ((lw < 80 || lh < 40) && i != 0) {
      omax = i;
      break;
    }
    for (int j = 0; j < o->nsublevels; j++) {
      if (nlev >= ORC_MAX_LEVELS) return -1;
      float esigma = o->soffset * powf(2.0f, (float)(j) / (float)(o->nsublevels) + i);
      g->w[nlev] = lw;
      g->h[nlev] = lh;
      g->octave[nlev] = i;
      g->sublevel[nlev] = j;
      g->esigma[nlev] = esigma;
      g->etime[nlev] = 0.5 * (esigma * esigma);
      nlev++;
    }
  }
  g->nlev = nlev;
  g->omax = omax;
  for (int i = 1; i < nlev; i++) {
    float ttime = g->etime[i] - g->etime[i - 1];
    float tmax = 0.25;
    int n = orc_fed_tau_by_process_time(ttime, 1, tmax, 1, g->tau[i], ORC_MAX_TAU);
    if (n < 0) return -1;
    g->nsteps[i] = n;
  }
  float smax = 10.0 * sqrtf(2.0f); /* MLDB/SURF, AKAZE.cpp:353-356 */
  for (int i = 0; i < nlev; i++) {
    float ratio = powf(2.0f, (float)g->octave[i]);
    g->deriv_step[i] = fround(g->esigma[i] * o->derivative_factor / ratio);
    float size = g->esigma[i] * o->derivative_factor;
    g->det_size[i] = size;
    g->border[i] = smax * fround(size / ratio);
  }
  return nlev;
}

/* ------------------------------------------------------------------------ */
/* Gaussian low-pass: cuda_akaze.cu:56-200                                   */
/* ------------------------------------------------------------------------ */
int orc_ksize_radius(int ksize) {
  /* cuda_akaze.cu:187-200 */
  if (ksize <= 5) return 2;
  if (ksize <= 7) return 3;
  if (ksize <= 9) return 4;
  return 5;
}

void orc_gauss_taps(double var, int R, float* k) {
  /* cuda_akaze.cu:175-185: taps rounded through float expf, normalised by a float sum */
  float sum = 0.0f;
  for (int j = -R; j <= R; j++) {
    k[j + R] = (float)expf((float)(-(double)j * j / 2.0 / var));
    sum += k[j + R];
  }
  for (int j = -R; j <= R; j++) k[j + R] /= sum;
}

void orc_lowpass(const float* in, float* out, int w, int h, double var, int ksize) {
  /* Row pass then column pass, replicate border (ConvRowGPU :56-80, ConvColGPU :85-128). */
  int R = orc_ksize_radius(ksize);
  float k[11];
  orc_gauss_taps(var, R, k);
  float* tmp = (float*)malloc(sizeof(float) * (size_t)w * h);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    const float* row = in + (size_t)y * w;
    for (int x = 0; x < w; x++) {
      float s = 0.0f;
      for (int i = 0; i <= 2 * R; i++) s += row[iclamp(x + i - R, 0, w - 1)] * k[i];
      tmp[(size_t)y * w + x] = s;
    }
  }
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      float s = 0.0f;
      for (int i = 0; i <= 2 * R; i++) s += tmp[(size_t)iclamp(y + i - R, 0, h - 1) * w + x] * k[i];
      out[(size_t)y * w + x] = s;
    }
  }
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Contrast factor: ContrastPercentile / MaxContrast / HistContrast,         */
/* cuda_akaze.cu:543-649                                                     */
/* ------------------------------------------------------------------------ */
float orc_contrast(const float* img, int w, int h, float perc, int nbins, int32_t* hist_out, float* max_out) {
  float* blur = (float*)malloc(sizeof(float) * (size_t)w * h);
  float* grad = (float*)calloc((size_t)w * h, sizeof(float));
  orc_lowpass(img, blur, w, h, 1.0f, 5);
  float maxv = 0.0f;
  /* gradient of the 3x3 window whose top-left is (x, y), stored at (x+1, y+1) (:567-573) */
#pragma omp parallel for schedule(static) reduction(max : maxv)
  for (int y = 0; y < h - 2; y++) {
    for (int x = 0; x < w - 2; x++) {
      const float* b = blur + (size_t)y * w + x;
      float dx = 3.0f * (b[0] - b[2] + b[2 * w] - b[2 * w + 2]) + 10.0f * (b[w] - b[w + 2]);
      float dy = 3.0f * (b[0] + b[2] - b[2 * w] - b[2 * w + 2]) + 10.0f * (b[1] - b[2 * w + 1]);
      float gv = sqrtf(dx * dx + dy * dy);
      grad[(size_t)(y + 1) * w + (x + 1)] = gv;
      if (gv > maxv) maxv = gv;
    }
  }
  if (nbins > 512) nbins = 512; /* :619-624 */
  int32_t hist[512];
  memset(hist, 0, sizeof(hist));
  float imaxval = 1.0f / maxv;
  for (int y = 1; y < h - 1; y++) {
    for (int x = 1; x < w - 1; x++) {
      float t = nbins * grad[(size_t)y * w + x] * imaxval; /* :594 */
      int idx = (t != t) ? 0 : (int)t;                     /* NaN (all-flat image) -> bin 0 */
      idx = imin(idx, nbins - 1);
      hist[idx]++;
    }
  }
  int npoints = (w - 2) * (h - 2);
  int nthreshold = (int)(npoints * perc);
  int k = 0, nelements = 0;
  for (k = 0; nelements < nthreshold && k < nbins; k++) nelements += hist[k];
  float contrast = (nelements < nthreshold ? 0.03f : maxv * ((float)k / nbins)); /* :637-642 */
  if (hist_out) memcpy(hist_out, hist, sizeof(int32_t) * nbins);
  if (max_out) *max_out = maxv;
  free(blur);
  free(grad);
  return contrast;
}

/* ------------------------------------------------------------------------ */
/* Conductance: Flow, cuda_akaze.cu:240-287 (unnormalised Scharr, reflect-101) */
/* ------------------------------------------------------------------------ */
static inline int reflect101(int v, int n) { return v < 0 ? -v : (v >= n ? 2 * n - 2 - v : v); }

void orc_flow(const float* L, float* g, int w, int h, int type, float kcontrast) {
  float invk = 1.0f / (kcontrast * kcontrast);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    int ym = reflect101(y - 1, h), yp = reflect101(y + 1, h);
    for (int x = 0; x < w; x++) {
      int xm = reflect101(x - 1, w), xp = reflect101(x + 1, w);
      float ul = L[(size_t)ym * w + xm], ur = L[(size_t)ym * w + xp];
      float ll = L[(size_t)yp * w + xm], lr = L[(size_t)yp * w + xp];
      float lx = 3.0f * (lr - ll + ur - ul) + 10.0f * (L[(size_t)y * w + xp] - L[(size_t)y * w + xm]);
      float ly = 3.0f * (lr + ll - ur - ul) + 10.0f * (L[(size_t)yp * w + x] - L[(size_t)ym * w + x]);
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
      g[(size_t)y * w + x] = r;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* One explicit diffusion step: NLDStep + NLDUpdate, cuda_akaze.cu:289-383.   */
/* Neighbours are clamped (zero flux across the border).                     */
/* ------------------------------------------------------------------------ */
void orc_fed_step(const float* L, const float* g, float* out, int w, int h, float tau) {
  float stepsize = 0.5 * tau; /* :359 */
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    int ym = imax(y - 1, 0), yp = imin(y + 1, h - 1);
    for (int x = 0; x < w; x++) {
      int xm = imax(x - 1, 0), xp = imin(x + 1, w - 1);
      size_t c = (size_t)y * w + x;
      float i0 = L[c], f0 = g[c];
      float xpos = (f0 + g[(size_t)y * w + xp]) * (L[(size_t)y * w + xp] - i0);
      float xneg = (f0 + g[(size_t)y * w + xm]) * (i0 - L[(size_t)y * w + xm]);
      float ypos = (f0 + g[(size_t)yp * w + x]) * (L[(size_t)yp * w + x] - i0);
      float yneg = (f0 + g[(size_t)ym * w + x]) * (i0 - L[(size_t)ym * w + x]);
      out[c] = i0 + stepsize * (xpos - xneg + ypos - yneg);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Downsampling: HalfSample2 (:432-440) for exact halving, HalfSample        */
/* (:385-430) otherwise, including the ptri[1] read of :413 (odd width and   */
/* odd height, on output rows y % 16 == 15).                                 */
/* ------------------------------------------------------------------------ */
static inline float hs_row(const float* in, int iw, int r, int x, int ow, int quirk) {
  const float* p = in + (size_t)r * iw + 2 * x;
  if (2 * ow == iw) return ow * (p[0] + p[1]);
  float f0 = ow - x, f2 = 1 + x;
  return f0 * p[0] + ow * p[1] + f2 * (quirk ? p[1] : p[2]);
}

void orc_halfsample(const float* in, int iw, int ih, float* out, int ow, int oh) {
  if (iw == 2 * ow && ih == 2 * oh) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < oh; y++)
      for (int x = 0; x < ow; x++) {
        const float* p = in + (size_t)(2 * y) * iw + 2 * x;
        out[(size_t)y * ow + x] = 0.25f * (p[0] + p[1] + p[iw] + p[iw + 1]);
      }
    return;
  }
  int oddw = (2 * ow != iw), oddh = (2 * oh != ih);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < oh; y++) {
    for (int x = 0; x < ow; x++) {
      float b0 = hs_row(in, iw, 2 * y, x, ow, 0);
      float b16 = hs_row(in, iw, 2 * y + 1, x, ow, 0);
      float r;
      if (!oddh) {
        r = oh * (b0 + b16) / (iw * ih);
      } else {
        float b32 = hs_row(in, iw, 2 * y + 2, x, ow, oddw && (y % 16 == 15));
        float f0 = oh - y, f2 = 1 + y;
        r = (f0 * b0 + oh * b16 + f2 * b32) / (iw * ih);
      }
      out[(size_t)y * ow + x] = r;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Hessian response: Derivate (:651-671) + HessianDeterminant (:673-701).    */
/* Reflect-101 at offset `step`.                                             */
/* ------------------------------------------------------------------------ */
static void deriv_pass(const float* f, float* dx_out, float* dy_out, int w, int h, int s, float fac1, float fac2) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    int yl = (y < s ? s - y : y - s);
    int yh = (y >= h - s ? 2 * h - y - s - 2 : y + s);
    for (int x = 0; x < w; x++) {
      int xl = (x < s ? s - x : x - s);
      int xh = (x >= w - s ? 2 * w - x - s - 2 : x + s);
      float ul = f[(size_t)yl * w + xl], ur = f[(size_t)yl * w + xh];
      float ll = f[(size_t)yh * w + xl], lr = f[(size_t)yh * w + xh];
      if (dx_out) {
        float cl = f[(size_t)y * w + xl], cr = f[(size_t)y * w + xh];
        dx_out[(size_t)y * w + x] = fac1 * (ur + lr - ul - ll) + fac2 * (cr - cl);
      }
      if (dy_out) {
        float uc = f[(size_t)yl * w + x], lc = f[(size_t)yh * w + x];
        dy_out[(size_t)y * w + x] = fac1 * (lr + ll - ur - ul) + fac2 * (lc - uc);
      }
    }
  }
}

void orc_hessian(const float* Lsm, float* Lx, float* Ly, float* det, int w, int h, int step) {
  float wgt = 10.0 / 3.0; /* :706-708 */
  float fac1 = 1.0 / (2.0 * (wgt + 2.0));
  float fac2 = wgt * fac1;
  deriv_pass(Lsm, Lx, Ly, w, h, step, fac1, fac2);
  size_t n = (size_t)w * h;
  float* lxx = (float*)malloc(sizeof(float) * n);
  float* lxy = (float*)malloc(sizeof(float) * n);
  float* lyy = (float*)malloc(sizeof(float) * n);
  deriv_pass(Lx, lxx, lxy, w, h, step, fac1, fac2);
  deriv_pass(Ly, NULL, lyy, w, h, step, fac1, fac2);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; i++) det[i] = lxx[i] * lyy[i] - lxy[i] * lxy[i];
  free(lxx);
  free(lxy);
  free(lyy);
}

/* ------------------------------------------------------------------------ */
/* In-level extrema + 2-D subpixel fit: FindExtrema, cuda_akaze.cu:727-774.  */
/* Candidates come out in raster (y, x) order, which is the per-level order  */
/* bitonicSort (:882-939) establishes in the reference.                      */
/* ------------------------------------------------------------------------ */
int orc_find_extrema(const float* imd, int w, int h, float border, float thr, int level, int octave,
                     float size, orc_kp* out, int cap) {
  int n = 0;
  for (int y = 0; y < h; y++) {
    int up_y = (int)(y - border + 0.5f) - 1;
    int down_y = (int)(y + border + 0.5f) + 1;
    if (up_y < 0 || down_y >= h) continue;
    for (int x = 0; x < w; x++) {
      int left_x = (int)(x - border + 0.5f) - 1;
      int right_x = (int)(x + border + 0.5f) + 1;
      if (left_x < 0 || right_x >= w) continue;
      size_t p = (size_t)y * w + x;
      size_t pitch = (size_t)w;
      float v = imd[p];
      if (v > thr && v > imd[p - pitch - 1] && v > imd[p + pitch + 1] && v > imd[p + pitch - 1] &&
          v > imd[p - pitch + 1] && v > imd[p - 1] && v > imd[p + 1] && v > imd[p + pitch] && v > imd[p - pitch]) {
        float dx = 0.5f * (imd[p + 1] - imd[p - 1]);
        float dy = 0.5f * (imd[p + pitch] - imd[p - pitch]);
        float dxx = imd[p + 1] + imd[p - 1] - 2.0f * v;
        float dyy = imd[p + pitch] + imd[p - pitch] - 2.0f * v;
        float dxy = 0.25f * (imd[p + pitch + 1] + imd[p - pitch - 1] - imd[p + pitch - 1] - imd[p - pitch + 1]);
        float det = dxx * dyy - dxy * dxy;
        float idet = (det != 0.0f ? 1.0f / det : 0.0f);
        float dst0 = idet * (dxy * dy - dyy * dx);
        float dst1 = idet * (dxy * dx - dxx * dy);
        int weak = 1;
        if (dst0 >= -1.0f && dst0 <= 1.0f && dst1 >= -1.0f && dst1 <= 1.0f) weak = 0;
        if (n < cap) {
          orc_kp* k = &out[n];
          k->response = v;
          k->size = (weak ? -1 : 1) * 2.0 * size;
          float octsub = (dst0 < 0 ? -1 : 1) * (octave + fabsf(dst0));
          k->octave = f2bits(weak ? (float)octave : octsub);
          k->class_id = level;
          int ratio = (1 << octave);
          k->x = ratio * (x);
          k->y = ratio * (y);
          k->angle = dst1;
        }
        n++;
      }
    }
  }
  return n;
}

/* ------------------------------------------------------------------------ */
/* Non-maximum suppression across the same and the previous level:           */
/* FindNeighbors (:1005-1070), FilterExtrema_kernel (:1072-1224),            */
/* sortFiltered_kernel (:1227-1316).                                         */
/* ------------------------------------------------------------------------ */
int orc_nms(const orc_kp* cand, int n, const int* level_start, int nlev, int max_waves, orc_kp* out, int* waves_out) {
  return orc_nms_ex(cand, n, level_start, nlev, max_waves, out, waves_out, NULL);
}

int orc_nms_ex(const orc_kp* cand, int n, const int* level_start, int nlev, int max_waves, orc_kp* out,
               int* waves_out, int* kidx_out) {
  /* neighbour lists (CSR) */
  int* nb_off = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  int nb_cap = n * 4 + 16, nb_n = 0;
  int* nb = (int*)malloc(sizeof(int) * (size_t)nb_cap);
  for (int c = 0; c < n; c++) {
    nb_off[c] = nb_n;
    const orc_kp* k = &cand[c];
    int lvl = k->class_id;
    int start = level_start[lvl];
    float size = k->size;
    for (int i = c - 1; i >= start; i--) {
      const orc_kp* q = &cand[i];
      if (k->y - q->y > size * .5f) break;
      float dist = (k->x - q->x) * (k->x - q->x) + (k->y - q->y) * (k->y - q->y);
      if ((double)dist < (double)(size * size) * 0.25) {
        if (nb_n == nb_cap) { nb_cap *= 2; nb = (int*)realloc(nb, sizeof(int) * (size_t)nb_cap); }
        nb[nb_n++] = i;
      }
    }
    if (lvl > 0) {
      int cmp = level_start[lvl - 1];
      for (int i = start - 1; i >= cmp; i--) {
        const orc_kp* q = &cand[i];
        if (q->y - k->y > size * .5f) continue;
        if (k->y - q->y > size * .5f) break;
        float dist = (k->x - q->x) * (k->x - q->x) + (k->y - q->y) * (k->y - q->y);
        if ((double)dist < (double)(size * size) * 0.25) {
          if (nb_n == nb_cap) { nb_cap *= 2; nb = (int*)realloc(nb, sizeof(int) * (size_t)nb_cap); }
          nb[nb_n++] = i;
        }
      }
    }
  }
  nb_off[n] = nb_n;
  (void)nlev;

  /* wave-synchronous greedy suppression; -1 unprocessed, -2 dropped, >=0 kept */
  int* mem = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  int* minnb = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  unsigned char* add = (unsigned char*)malloc((size_t)n + 1);
  for (int i = 0; i < n; i++) mem[i] = -1;
  int wave = 0;
  for (wave = 0; wave < max_waves; wave++) {
    int any = 0;
    for (int i = 0; i < n; i++) { minnb[i] = n + 1; add[i] = 1; }
    for (int i = 0; i < n; i++) {
      if (mem[i] != -1) continue;
      any = 1;
      int proc = 1;
      for (int e = nb_off[i]; e < nb_off[i + 1]; e++) {
        int m = mem[nb[e]];
        if (m == -1) { proc = 0; add[i] = 0; break; }
        if (m >= 0) add[i] = 0;
      }
      if (proc && !add[i]) {
        /* target: kept neighbour with the smallest member value (:1153-1170) */
        int best = -1, bestv = 0;
        for (int e = nb_off[i]; e < nb_off[i + 1]; e++) {
          int m = mem[nb[e]];
          if (m >= 0 && (best < 0 || m < bestv)) { best = nb[e]; bestv = m; }
        }
        if (i < minnb[best]) minnb[best] = i;
      }
    }
    for (int i = 0; i < n; i++)
      if (mem[i] == -1 && add[i]) mem[i] = i;
    for (int q = 0; q < n; q++) {
      int c = minnb[q];
      if (c == n + 1) continue;
      if (mem[c] == -1 && !add[c]) {
        if (cand[c].response > cand[q].response) {
          mem[c] = q;
          mem[q] = -2;
        } else {
          mem[c] = -2;
        }
      }
    }
    if (!any) break;
  }
  if (waves_out) *waves_out = wave;

  /* output: per 2048-chunk of candidate slots, ascending member values (:1243-1308) */
  int nout = 0;
  int vals[2048];
  for (int base = 0; base < n; base += 2048) {
    int cnt = 0;
    for (int i = base; i < base + 2048 && i < n; i++) {
      int v = mem[i];
      if (v >= 0 && !(cand[v].size < 0)) vals[cnt++] = v;
    }
    /* insertion sort: chunks are small and values distinct */
    for (int a = 1; a < cnt; a++) {
      int t = vals[a], b = a - 1;
      while (b >= 0 && vals[b] > t) { vals[b + 1] = vals[b]; b--; }
      vals[b + 1] = t;
    }
    for (int a = 0; a < cnt; a++) {
      const orc_kp* ok = &cand[vals[a]];
      float octf = bits2f(ok->octave);
      float octsub = fabsf(octf);
      int octave = (int)octsub;
      float subp = (octf < 0 ? -1 : 1) * (octsub - octave);
      float ratio = 1 << octave;
      if (kidx_out) kidx_out[nout] = vals[a];
      orc_kp* t = &out[nout++];
      t->y = ratio * ((int)(0.5f + ok->y / ratio) + ok->angle);
      t->x = ratio * ((int)(0.5f + ok->x / ratio) + subp);
      t->class_id = ok->class_id;
      t->octave = octave;
      t->response = ok->response;
      t->size = ok->size;
      t->angle = 0.0f;
    }
  }
  free(nb_off);
  free(nb);
  free(mem);
  free(minnb);
  free(add);
  return nout;
}

/* ------------------------------------------------------------------------ */
/* Cephes single-precision atan / atan2 / sin / cos (Moshier 1989).          */
/* ------------------------------------------------------------------------ */
static float cephes_atanf(float xx) {
  float x = xx, y, z;
  int sign = 0;
  if (x < 0.0f) { sign = 1; x = -x; }
  if (x > 2.414213562373095f) { /* tan(3pi/8) */
    y = 1.5707963267948966f;
    x = -(1.0f / x);
  } else if (x > 0.4142135623730950f) { /* tan(pi/8) */
    y = 0.7853981633974483f;
    x = (x - 1.0f) / (x + 1.0f);
  } else {
    y = 0.0f;
  }
  z = x * x;
  y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x + x);
  return sign ? -y : y;
}

float orc_atan2f(float y, float x) {
  /* quadrant handling as C99 atan2, including signed zeros */
  const float PIF = 3.14159265358979f, PIO2F = 1.5707963267948966f;
  if (x == 0.0f) {
    if (y > 0.0f) return PIO2F;
    if (y < 0.0f) return -PIO2F;
    /* y == +-0 */
    if (signbit(x)) return signbit(y) ? -PIF : PIF;
    return y;
  }
  if (y == 0.0f) {
    if (x > 0.0f) return y;
    return signbit(y) ? -PIF : PIF;
  }
  float z = cephes_atanf(y / x);
  if (x < 0.0f) z = (y < 0.0f) ? z - PIF : z + PIF;
  return z;
}

void orc_sincosf(float xx, float* s, float* c) {
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

float orc_orient_weight(int r2) {
  /* exp(-r2 / (2.5*2.5*2)) of cuda_akaze.cu:2133, rounded once from double */
  return (float)exp((double)((float)(-r2) / (2.5f * 2.5f * 2.0f)));
}

/* ------------------------------------------------------------------------ */
/* Dominant orientation: FindOrientation, cuda_akaze.cu:2114-2167            */
/* ------------------------------------------------------------------------ */
void orc_orientation(orc_kp* kp, const float* Lx, const float* Ly, int w, int h) {
  float resx[42], resy[42], re8x[42], re8y[42];
  for (int t = 0; t < 42; t++) resx[t] = resy[t] = 0.0f;
  int octave = kp->octave;
  int step = (int)(0.5f * kp->size + 0.5f) >> octave;
  int x = (int)(kp->x + 0.5f) >> octave;
  int y = (int)(kp->y + 0.5f) >> octave;
  for (int tx = 0; tx < 208; tx++) {
    int i = (tx & 15) - 6;
    int j = (tx / 16) - 6;
    int r2 = i * i + j * j;
    if (r2 < 36) {
      float gw = orc_orient_weight(r2);
      int xs = iclamp(x + step * i, 0, w - 1), ys = iclamp(y + step * j, 0, h - 1);
      float dx = gw * Lx[(size_t)ys * w + xs];
      float dy = gw * Ly[(size_t)ys * w + xs];
      float angle = orc_atan2f(dy, dx);
      int a = imax(imin((int)(angle * (21 / ORC_PI)) + 21, 41), 0);
      resx[a] += dx;
      resy[a] += dy;
    }
  }
  for (int t = 0; t < 42; t++) {
    re8x[t] = resx[t];
    re8y[t] = resy[t];
    for (int k = t + 1; k < t + 7; k++) {
      re8x[t] += resx[k < 42 ? k : k - 42];
      re8y[t] += resy[k < 42 ? k : k - 42];
    }
  }
  float maxr = 0.0f;
  int maxk = 0;
  for (int k = 0; k < 42; k++) {
    float r = re8x[k] * re8x[k] + re8y[k] * re8y[k];
    if (r > maxr) { maxr = r; maxk = k; }
  }
  float angle = orc_atan2f(re8y[maxk], re8x[maxk]);
  kp->angle = (angle < 0.0f ? angle + 2.0f * ORC_PI : angle);
}

/* ------------------------------------------------------------------------ */
/* MLDB descriptor: ExtractDescriptors (:1420-1533), BuildDescriptor          */
/* (:1786-1807), comparison tables of InitCompareIndices (:2034-2111).        */
/* ------------------------------------------------------------------------ */
void orc_compare_tables(int32_t* idx1, int32_t* idx2) {
  int c = 0;
  static const int lo[3] = {0, 4, 13}, hi[3] = {4, 13, 29};
  for (int g = 0; g < 3; g++)
    for (int ch = 0; ch < 3; ch++)
      for (int j = lo[g]; j < hi[g] - 1; j++)
        for (int i = j + 1; i < hi[g]; i++) {
          idx1[c] = 3 * j + ch;
          idx2[c] = 3 * i + ch;
          c++;
        }
  for (; c < 488; c++) idx1[c] = idx2[c] = 0;
}

void orc_mldb_ex(const orc_kp* kp, const float* Lt, const float* Lx, const float* Ly, int w, int h, int patsize,
                 int nch, const int32_t* idx1, const int32_t* idx2, int nbits, uint8_t* desc, float* vals_out) {
  int size2 = patsize;
  int size3 = (int)ceilf(2.0f * patsize / 3.0f);
  int size4 = (int)ceilf(0.5f * patsize);
  float iratio = 1.0f / (1 << kp->octave);
  int scale = (int)(0.5f * kp->size * iratio + 0.5f);
  float xf = kp->x * iratio;
  float yf = kp->y * iratio;
  float co, si;
  orc_sincosf(kp->angle, &si, &co);
  int winsize = imax(3 * size3, 4 * size4);
  float part[64][87];
  memset(part, 0, sizeof(part));
  for (int i = 0; i < winsize * winsize; i++) {
    int lane = i & 63;
    int y = i / winsize;
    int x = i - winsize * y;
    int m = imax(x, y);
    if (m >= winsize) continue;
    int l = x - size2;
    int k = y - size2;
    int xp = (int)(xf + scale * (k * co - l * si) + 0.5f);
    int yp = (int)(yf + scale * (k * si + l * co) + 0.5f);
    xp = iclamp(xp, 0, w - 1);
    yp = iclamp(yp, 0, h - 1);
    size_t pos = (size_t)yp * w + xp;
    float im = Lt[pos], dx = Lx[pos], dy = Ly[pos];
    float rx, ry;
    if (nch == 2) {
      /* intensity + gradient magnitude (the two-channel M-LDB of
       * AKAZEOptions::descriptor_channels, AKAZEConfig.h:131) */
      rx = sqrtf(dx * dx + dy * dy);
      ry = 0.0f;
    } else {
      rx = -dx * si + dy * co;
      ry = dx * co + dy * si;
    }
    if (m < 2 * size2) {
      int x2 = (x < size2 ? 0 : 1), y2 = (y < size2 ? 0 : 1);
      int cell = y2 * 2 + x2;
      part[lane][3 * cell] += im;
      part[lane][3 * cell + 1] += rx;
      part[lane][3 * cell + 2] += ry;
    }
    if (m < 3 * size3) {
      int x3 = (x < size3 ? 0 : (x < 2 * size3 ? 1 : 2));
      int y3 = (y < size3 ? 0 : (y < 2 * size3 ? 1 : 2));
      int cell = 4 + y3 * 3 + x3;
      part[lane][3 * cell] += im;
      part[lane][3 * cell + 1] += rx;
      part[lane][3 * cell + 2] += ry;
    }
    if (m < 4 * size4) {
      int x4 = (x < 2 * size4 ? (x < size4 ? 0 : 1) : (x < 3 * size4 ? 2 : 3));
      int y4 = (y < 2 * size4 ? (y < size4 ? 0 : 1) : (y < 3 * size4 ? 2 : 3));
      int cell = 13 + y4 * 4 + x4;
      part[lane][3 * cell] += im;
      part[lane][3 * cell + 1] += rx;
      part[lane][3 * cell + 2] += ry;
    }
  }
  float vals[87];
  for (int v = 0; v < 87; v++) {
    float a[32];
    for (int t = 0; t < 32; t++) a[t] = part[t][v] + part[t + 32][v];
    for (int s = 1; s < 32; s <<= 1)
      for (int t = 0; t + s < 32; t += 2 * s) a[t] = a[t] + a[t + s];
    vals[v] = a[0];
  }
  int nbytes = (nbits + 7) / 8;
  for (int b = 0; b < nbytes; b++) {
    uint8_t r = 0;
    for (int i = 0; i < 8 && b * 8 + i < nbits; i++)
      r |= (uint8_t)((vals[idx1[b * 8 + i]] > vals[idx2[b * 8 + i]] ? 1 : 0) << i);
    desc[b] = r;
  }
  if (vals_out) memcpy(vals_out, vals, sizeof(vals));
}

void orc_mldb(const orc_kp* kp, const float* Lt, const float* Lx, const float* Ly, int w, int h, int patsize,
              uint8_t* desc, float* vals_out) {
  int32_t i1[488], i2[488];
  orc_compare_tables(i1, i2);
  orc_mldb_ex(kp, Lt, Lx, Ly, w, h, patsize, 3, i1, i2, 486, desc, vals_out);
}

/* generateDescriptorSubsample, AKAZE.cpp:498-603: the quasi-random bit
 * selection of an nbits-long M-LDB descriptor.  The full comparison list
 * (every cell pair j < k of the 2x2, 3x3 and 4x4 grids, cells at pattern
 * offsets psz * (j % g) - P, psz * (j / g) - P, psz = ceil(2P / g)) is drawn
 * from with glibc's rand() after srand(1024) (here random_r on a private
 * TYPE_3 state, the generator rand() uses), the first 6 picks forced to the
 * 2x2 grid; a pick adds its two cells to the sample list unless present and
 * yields one comparison per channel.  The reference writes three comparison
 * rows per pick whatever the channel count (AKAZE.cpp:556-590), which for
 * fewer than three channels overwrites rows the next pick rewrites and runs
 * two rows past the matrix on the last pick; the rows it keeps are the
 * per-channel rows written here.  Returns the sample count, or -1 when nbits
 * exceeds the full descriptor. */
int orc_descriptor_subsample(int nbits, int patsize, int nch, int32_t* samples, int32_t* comps) {
  int ssz = 0;
  for (int i = 0; i < 3; i++) {
    int gz = (i + 2) * (i + 2);
    ssz += gz * (gz - 1) / 2;
  }
  if (nbits < 1 || nbits > ssz * nch) return -1;
  enum { FULL = 162 };
  int fullM[FULL][5];
  for (int i = 0, c = 0; i < 3; i++) {
    int gdiv = i + 2, gsz = gdiv * gdiv;
    int psz = (int)ceil(2. * patsize / (float)gdiv);
    for (int j = 0; j < gsz; j++)
      for (int k = j + 1; k < gsz; k++, c++) {
        fullM[c][0] = i;
        fullM[c][1] = psz * (j % gdiv) - patsize;
        fullM[c][2] = psz * (j / gdiv) - patsize;
        fullM[c][3] = psz * (k % gdiv) - patsize;
        fullM[c][4] = psz * (k / gdiv) - patsize;
      }
  }
  struct random_data rd;
  char state[128];
  memset(&rd, 0, sizeof rd);
  initstate_r(1024, state, sizeof state, &rd);
  int npicks = (int)ceil(nbits / (float)nch);
  int count = 0;
  for (int i = 0; i < npicks; i++) {
    int32_t r = 0;
    random_r(&rd, &r);
    size_t k = (size_t)r % (size_t)(FULL - i);
    if (i < 6) k = (size_t)i;
    for (int end = 0; end < 2; end++) {
      int sx = fullM[k][1 + 2 * end], sy = fullM[k][2 + 2 * end], found = -1;
      for (int j = 0; j < count; j++)
        if (samples[3 * j] == fullM[k][0] && samples[3 * j + 1] == sx && samples[3 * j + 2] == sy) {
          found = j;
          break;
        }
      if (found < 0) {
        samples[3 * count] = fullM[k][0];
        samples[3 * count + 1] = sx;
        samples[3 * count + 2] = sy;
        found = count++;
      }
      for (int ch = 0; ch < nch; ch++)
        if (i * nch + ch < nbits) comps[2 * (i * nch + ch) + end] = nch * found + ch;
    }
    memcpy(fullM[k], fullM[FULL - i - 1], sizeof fullM[k]);
  }
  return count;
}

/* The comparison list of the M-LDB variant, as indices into the 87 cell
 * values (3 * cell + channel, cells 0-3 of the 2x2 grid, 4-12 of the 3x3,
 * 13-28 of the 4x4): with descriptor_size 0, every cell pair j < i of each
 * grid for channels 0 .. nch-1 in turn (InitCompareIndices,
 * cuda_akaze.cu:2034-2111, for three channels); otherwise the comparisons of
 * orc_descriptor_subsample mapped from sample slots to cells.  Returns the bit
 * count, or -1. */
int orc_mldb_table(int descriptor_size, int patsize, int nch, int32_t* idx1, int32_t* idx2) {
  static const int lo[3] = {0, 4, 13}, hi[3] = {4, 13, 29};
  if (nch < 1 || nch > 3) return -1;
  int c = 0;
  if (descriptor_size == 0) {
    for (int g = 0; g < 3; g++)
      for (int ch = 0; ch < nch; ch++)
        for (int j = lo[g]; j < hi[g] - 1; j++)
          for (int i = j + 1; i < hi[g]; i++) {
            idx1[c] = 3 * j + ch;
            idx2[c] = 3 * i + ch;
            c++;
          }
    return c;
  }
  int32_t samples[3 * 29], comps[2 * 486];
  int ns = orc_descriptor_subsample(descriptor_size, patsize, nch, samples, comps);
  if (ns < 0) return -1;
  int cell_of[29];
  for (int s = 0; s < ns; s++) {
    int g = samples[3 * s], gdiv = g + 2;
    int psz = (int)ceil(2. * patsize / (float)gdiv);
    int cx = (samples[3 * s + 1] + patsize) / psz, cy = (samples[3 * s + 2] + patsize) / psz;
    cell_of[s] = lo[g] + cy * gdiv + cx;
  }
  for (int b = 0; b < descriptor_size; b++) {
    idx1[b] = 3 * cell_of[comps[2 * b] / nch] + comps[2 * b] % nch;
    idx2[b] = 3 * cell_of[comps[2 * b + 1] / nch] + comps[2 * b + 1] % nch;
  }
  return descriptor_size;
}

/* ------------------------------------------------------------------------ */
/* Brute-force Hamming 2-NN: MatchDescriptors, cuda_akaze.cu:1842-1967;       */
/* Matcher::bfmatch, AKAZE.cpp:36-79.  Rows are `stride` bytes, 61 used.      */
/* ------------------------------------------------------------------------ */
void orc_match2nn(const uint8_t* q, int nq, const uint8_t* t, int nt, int stride, orc_dmatch* out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int a = 0; a < nq; a++) {
    const uint8_t* qa = q + (size_t)a * stride;
    const int nb = stride < 61 ? stride : 61; /* shorter rows: the M-LDB subsets and channel variants */
    int b1 = 0, d1 = 512, b2 = 0, d2 = 512;
    for (int b = 0; b < nt; b++) {
      const uint8_t* tb = t + (size_t)b * stride;
      int d = 0;
      for (int k = 0; k < nb; k++) d += __builtin_popcount((unsigned)(qa[k] ^ tb[k]));
      if (d < d1) {
        d2 = d1; b2 = b1;
        d1 = d; b1 = b;
      } else if (d < d2) {
        d2 = d; b2 = b;
      }
    }
    out[2 * a].queryIdx = a; out[2 * a].trainIdx = b1; out[2 * a].imgIdx = 0; out[2 * a].distance = (float)d1;
    out[2 * a + 1].queryIdx = a; out[2 * a + 1].trainIdx = b2; out[2 * a + 1].imgIdx = 0; out[2 * a + 1].distance = (float)d2;
  }
}

/* ------------------------------------------------------------------------ */
/* Whole pipeline: Create_Nonlinear_Scale_Space (AKAZE.cpp:211-276),          */
/* Feature_Detection (:316-378), Compute_Descriptors (:405-449).              */
/* ------------------------------------------------------------------------ */
struct orc_ctx {
  orc_options o;
  orc_geom g;
  int W, H;
  float* lev[ORC_MAX_LEVELS][6];
  float kc[ORC_MAX_LEVELS];
  orc_kp* cand;
  int ncand, cand_cap;
  int lstart[ORC_MAX_LEVELS + 1];
  orc_kp* kps;
  int* kidx;
  int nkps;
  int waves;
  uint8_t* desc;
  int32_t cmp1[488], cmp2[488];
  int nbits;
};

orc_ctx* orc_create(const orc_options* o, int W, int H) {
  if (o->descriptor != 4 && o->descriptor != 5) return NULL;
  int nbits_probe[2][488];
  if (orc_mldb_table(o->descriptor_size, o->descriptor_pattern_size, o->descriptor_channels, nbits_probe[0],
                     nbits_probe[1]) < 0)
    return NULL;
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  c->o = *o;
  c->nbits = orc_mldb_table(o->descriptor_size, o->descriptor_pattern_size, o->descriptor_channels, c->cmp1, c->cmp2);
  c->W = W;
  c->H = H;
  if (orc_geometry(o, W, H, &c->g) <= 0) { free(c); return NULL; }
  for (int i = 0; i < c->g.nlev; i++)
    for (int k = 0; k < 6; k++) c->lev[i][k] = (float*)calloc((size_t)c->g.w[i] * c->g.h[i], sizeof(float));
  return c;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  for (int i = 0; i < c->g.nlev; i++)
    for (int k = 0; k < 6; k++) free(c->lev[i][k]);
  free(c->cand);
  free(c->kps);
  free(c->kidx);
  free(c->desc);
  free(c);
}

const orc_geom* orc_get_geom(const orc_ctx* c) { return &c->g; }
const float* orc_level(const orc_ctx* c, int lvl, int which) {
  if (lvl < 0 || lvl >= c->g.nlev || which < 0 || which > 5) return NULL;
  return c->lev[lvl][which];
}
float orc_get_kcontrast(const orc_ctx* c, int level) { return c->kc[level]; }

int orc_scale_space(orc_ctx* c, const float* img) {
  const orc_geom* g = &c->g;
  int w0 = g->w[0], h0 = g->h[0];
  float kc = orc_contrast(img, w0, h0, c->o.kcontrast_percentile, c->o.kcontrast_nbins, NULL, NULL);
  int ksize = 2 * ceil((c->o.soffset - 0.8) / 0.3) + 3; /* AKAZE.cpp:234-235 */
  orc_lowpass(img, c->lev[0][0], w0, h0, c->o.soffset * c->o.soffset, ksize);
  memcpy(c->lev[0][1], c->lev[0][0], sizeof(float) * (size_t)w0 * h0);
  c->kc[0] = kc;
  float* tmp = NULL;
  size_t tmpn = 0;
  for (int i = 1; i < g->nlev; i++) {
    int w = g->w[i], h = g->h[i];
    float* Lt = c->lev[i][0];
    if (g->octave[i] > g->octave[i - 1]) {
      orc_halfsample(c->lev[i - 1][0], g->w[i - 1], g->h[i - 1], Lt, w, h);
      kc = kc * 0.75;
    } else {
      memcpy(Lt, c->lev[i - 1][0], sizeof(float) * (size_t)w * h);
    }
    c->kc[i] = kc;
    orc_lowpass(Lt, c->lev[i][1], w, h, 1.0, 5);
    orc_flow(c->lev[i][1], c->lev[i][5], w, h, c->o.diffusivity, kc);
    if (tmpn < (size_t)w * h) {
      free(tmp);
      tmpn = (size_t)w * h;
      tmp = (float*)malloc(sizeof(float) * tmpn);
    }
    for (int j = 0; j < g->nsteps[i]; j++) {
      orc_fed_step(Lt, c->lev[i][5], tmp, w, h, g->tau[i][j]);
      memcpy(Lt, tmp, sizeof(float) * (size_t)w * h);
    }
  }
  free(tmp);
  return 0;
}

int orc_detect(orc_ctx* c) {
  const orc_geom* g = &c->g;
  float thr = c->o.dthreshold > c->o.min_dthreshold ? c->o.dthreshold : c->o.min_dthreshold;
  size_t cap = 0;
  for (int i = 0; i < g->nlev; i++) cap += (size_t)g->w[i] * g->h[i] / 4 + 4;
  free(c->cand);
  c->cand = (orc_kp*)malloc(sizeof(orc_kp) * cap);
  c->cand_cap = (int)cap;
  int n = 0;
  for (int i = 0; i < g->nlev; i++) {
    int w = g->w[i], h = g->h[i];
    orc_hessian(c->lev[i][1], c->lev[i][2], c->lev[i][3], c->lev[i][4], w, h, g->deriv_step[i]);
    c->lstart[i] = n;
    n += orc_find_extrema(c->lev[i][4], w, h, g->border[i], thr, i, g->octave[i], g->det_size[i], c->cand + n,
                          c->cand_cap - n);
  }
  c->lstart[g->nlev] = n;
  c->ncand = n;
  free(c->kps);
  c->kps = (orc_kp*)malloc(sizeof(orc_kp) * (size_t)(n + 1));
  free(c->kidx);
  c->kidx = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  c->nkps = orc_nms_ex(c->cand, n, c->lstart, g->nlev, c->o.nms_max_waves, c->kps, &c->waves, c->kidx);
  return c->nkps;
}

int orc_num_candidates(const orc_ctx* c) { return c->ncand; }
const orc_kp* orc_candidates(const orc_ctx* c) { return c->cand; }
const int* orc_level_start(const orc_ctx* c) { return c->lstart; }
int orc_nms_waves(const orc_ctx* c) { return c->waves; }
int orc_num_keypoints(const orc_ctx* c) { return c->nkps; }
const int* orc_kept_candidates(const orc_ctx* c) { return c->kidx; }
const orc_kp* orc_keypoints(const orc_ctx* c) { return c->kps; }
const uint8_t* orc_descriptors(const orc_ctx* c) { return c->desc; }

int orc_descriptor_bytes(const orc_ctx* c) { return (c->nbits + 7) / 8; }

/* Compute_Descriptors (AKAZE.cpp:405-449).  MLDB_UPRIGHT (descriptor 4) skips
 * the orientation: angle 0, the unrotated pattern.  Rows are
 * orc_descriptor_bytes() bytes. */
int orc_describe(orc_ctx* c) {
  int nb = orc_descriptor_bytes(c);
  free(c->desc);
  c->desc = (uint8_t*)calloc((size_t)c->nkps * nb + 1, 1);
  int n = c->nkps;
#pragma omp parallel for schedule(dynamic, 16)
  for (int p = 0; p < n; p++) {
    orc_kp* k = &c->kps[p];
    int l = k->class_id;
    int w = c->g.w[l], h = c->g.h[l];
    if (c->o.descriptor == 5)
      orc_orientation(k, c->lev[l][2], c->lev[l][3], w, h);
    else
      k->angle = 0.0f;
    orc_mldb_ex(k, c->lev[l][0], c->lev[l][2], c->lev[l][3], w, h, c->o.descriptor_pattern_size,
                c->o.descriptor_channels, c->cmp1, c->cmp2, c->nbits, c->desc + (size_t)p * nb, NULL);
  }
  return n;
}
