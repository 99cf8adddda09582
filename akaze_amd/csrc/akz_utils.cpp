// Host utilities of the reference's evaluation programs (src/lib/utils.cpp),
// without OpenCV: PGM input, the Mikolajczyk keypoint file, ground-truth
// homography files, NNDR match filtering, homography inlier counting, and a
// RANSAC homography estimator of this library's own (the reference calls
// cv::findHomography).  CPU only; no device calls.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/akaze_mi.h"

namespace {

// ---- PGM (netpbm P2 / P5) ------------------------------------------------
// Reads one whitespace-separated header token, skipping '#' comments.
bool pgm_token(FILE* f, char* buf, size_t cap) {
  int c = fgetc(f);
  for (;;) {
    while (c != EOF && (c == ' ' || c == '\t' || c == '\r' || c == '\n')) c = fgetc(f);
    if (c == '#') {
      while (c != EOF && c != '\n') c = fgetc(f);
      continue;
    }
    break;
  }
  if (c == EOF) return false;
  size_t n = 0;
  while (c != EOF && !(c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '#')) {
    if (n + 1 >= cap) return false;
    buf[n++] = (char)c;
    c = fgetc(f);
  }
  buf[n] = 0;
  // P5: exactly one whitespace byte separates maxval from the raster; it was
  // consumed above (c), unless it was a comment start
  if (c == '#') ungetc(c, f);
  return n > 0;
}

bool pgm_int(FILE* f, long* v) {
  char b[32];
  if (!pgm_token(f, b, sizeof b)) return false;
  char* e = nullptr;
  errno = 0;
  long x = strtol(b, &e, 10);
  if (errno || *e || x < 0) return false;
  *v = x;
  return true;
}

// ---- RANSAC homography ---------------------------------------------------
struct Rng {  // splitmix64
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  int below(int n) { return (int)(next() % (uint64_t)n); }
};

// Eigenvector of the smallest eigenvalue of a symmetric 9x9 matrix (cyclic
// Jacobi rotations).
void smallest_eigvec9(double A[9][9], double v_out[9]) {
  double V[9][9];
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0;
    for (int p = 0; p < 9; p++)
      for (int q = p + 1; q < 9; q++) off += A[p][q] * A[p][q];
    if (off < 1e-30) break;
    for (int p = 0; p < 9; p++)
      for (int q = p + 1; q < 9; q++) {
        if (std::fabs(A[p][q]) < 1e-300) continue;
        double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 9; k++) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 9; k++) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 9; k++) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  int best = 0;
  for (int i = 1; i < 9; i++)
    if (A[i][i] < A[best][best]) best = i;
  for (int k = 0; k < 9; k++) v_out[k] = V[k][best];
}

// Hartley normalisation: centroid to the origin, mean distance sqrt(2).
void normaliser(const double* x, const double* y, const int* idx, int n, double T[3]) {
  double cx = 0, cy = 0;
  for (int i = 0; i < n; i++) { cx += x[idx[i]]; cy += y[idx[i]]; }
  cx /= n; cy /= n;
  double d = 0;
  for (int i = 0; i < n; i++) d += std::hypot(x[idx[i]] - cx, y[idx[i]] - cy);
  d /= n;
  double s = d > 0 ? std::sqrt(2.0) / d : 1.0;
  T[0] = s; T[1] = -s * cx; T[2] = -s * cy;  // x' = s x - s cx
}

// Direct linear transform over the points idx[0..n): H maps (x1, y1) to (x2, y2).
bool dlt(const double* x1, const double* y1, const double* x2, const double* y2, const int* idx, int n,
         double H[9]) {
  double T1[3], T2[3];
  normaliser(x1, y1, idx, n, T1);
  normaliser(x2, y2, idx, n, T2);
  double A[9][9] = {};
  for (int i = 0; i < n; i++) {
    double u = T1[0] * x1[idx[i]] + T1[1], v = T1[0] * y1[idx[i]] + T1[2];
    double U = T2[0] * x2[idx[i]] + T2[1], V = T2[0] * y2[idx[i]] + T2[2];
    double r1[9] = {-u, -v, -1, 0, 0, 0, U * u, U * v, U};
    double r2[9] = {0, 0, 0, -u, -v, -1, V * u, V * v, V};
    for (int a = 0; a < 9; a++)
      for (int b = 0; b < 9; b++) A[a][b] += r1[a] * r1[b] + r2[a] * r2[b];
  }
  double h[9];
  smallest_eigvec9(A, h);
  // H = T2^-1 Hn T1
  double Hn[3][3] = {{h[0], h[1], h[2]}, {h[3], h[4], h[5]}, {h[6], h[7], h[8]}};
  double t1[3][3] = {{T1[0], 0, T1[1]}, {0, T1[0], T1[2]}, {0, 0, 1}};
  double t2i[3][3] = {{1 / T2[0], 0, -T2[1] / T2[0]}, {0, 1 / T2[0], -T2[2] / T2[0]}, {0, 0, 1}};
  double M[3][3], R[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      M[i][j] = 0;
      for (int k = 0; k < 3; k++) M[i][j] += Hn[i][k] * t1[k][j];
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      R[i][j] = 0;
      for (int k = 0; k < 3; k++) R[i][j] += t2i[i][k] * M[k][j];
    }
  double s = R[2][2];
  if (!std::isfinite(s) || std::fabs(s) < 1e-12) return false;
  for (int i = 0; i < 9; i++) H[i] = R[i / 3][i % 3] / s;
  for (int i = 0; i < 9; i++)
    if (!std::isfinite(H[i])) return false;
  return true;
}

double reproj_err2(const double H[9], double x, double y, double X, double Y) {
  double w = H[6] * x + H[7] * y + H[8];
  if (std::fabs(w) < 1e-15) return 1e300;
  double u = (H[0] * x + H[1] * y + H[2]) / w, v = (H[3] * x + H[4] * y + H[5]) / w;
  return (u - X) * (u - X) + (v - Y) * (v - Y);
}

// twice the signed area of the triangle (a, b, c)
double area2(const double* x, const double* y, int a, int b, int c) {
  return (x[b] - x[a]) * (y[c] - y[a]) - (y[b] - y[a]) * (x[c] - x[a]);
}

bool degenerate(const double* x, const double* y, const int* s) {
  for (int i = 0; i < 4; i++) {
    int a = s[(i + 1) % 4], b = s[(i + 2) % 4], c = s[(i + 3) % 4];
    if (std::fabs(area2(x, y, a, b, c)) < 1e-6) return true;
  }
  return false;
}

int score(const double H[9], const double* x1, const double* y1, const double* x2, const double* y2, int n,
          double thr2, uint8_t* mask, double* err_sum) {
  int cnt = 0;
  double es = 0;
  for (int i = 0; i < n; i++) {
    double e = reproj_err2(H, x1[i], y1[i], x2[i], y2[i]);
    bool in = e <= thr2;
    if (mask) mask[i] = in ? 1 : 0;
    if (in) { cnt++; es += e; }
  }
  if (err_sum) *err_sum = es;
  return cnt;
}

}  // namespace

extern "C" {

akz_status akz_read_pgm(const char* path, uint8_t* dst, size_t cap, int* width, int* height) {
  if (!path || !width || !height) return AKZ_ERR_ARG;
  FILE* f = fopen(path, "rb");
  if (!f) return AKZ_ERR_ARG;
  char magic[8];
  long w = 0, h = 0, maxval = 0;
  akz_status st = AKZ_OK;
  if (!pgm_token(f, magic, sizeof magic) || (strcmp(magic, "P5") && strcmp(magic, "P2")) || !pgm_int(f, &w) ||
      !pgm_int(f, &h) || !pgm_int(f, &maxval) || w < 1 || h < 1 || w > 65535 || h > 65535 || maxval < 1 ||
      maxval > 65535) {
    fclose(f);
    return AKZ_ERR_UNSUPPORTED;
  }
  *width = (int)w;
  *height = (int)h;
  size_t n = (size_t)w * (size_t)h;
  if (!dst) {
    fclose(f);
    return AKZ_OK;
  }
  if (cap < n) {
    fclose(f);
    return AKZ_ERR_ARG;
  }
  // 16-bit samples keep their high byte, as cv::imread(..., 0) converts them
  if (magic[1] == '5') {
    if (maxval < 256) {
      if (fread(dst, 1, n, f) != n) st = AKZ_ERR_ARG;
    } else {
      std::vector<uint8_t> buf(2 * n);
      if (fread(buf.data(), 1, 2 * n, f) != 2 * n) st = AKZ_ERR_ARG;
      for (size_t i = 0; i < n && st == AKZ_OK; i++) dst[i] = buf[2 * i];
    }
  } else {
    for (size_t i = 0; i < n; i++) {
      long v = 0;
      if (!pgm_int(f, &v) || v > maxval) { st = AKZ_ERR_ARG; break; }
      dst[i] = (uint8_t)(maxval < 256 ? v : (v >> 8));
    }
  }
  fclose(f);
  return st;
}

// save_keypoints, utils.cpp:97-151: "dsize\nN\n" then per keypoint
// "x y 1/r^2 0 1/r^2 [descriptor bytes]" with r = size, ostream's default
// six significant digits.
akz_status akz_save_keypoints(const char* path, const akz_keypoint* kps, int n, const uint8_t* desc, int desc_cols,
                              int save_desc) {
  if (!path || n < 0 || (n > 0 && !kps) || desc_cols < 0 || (save_desc && desc_cols > 0 && n > 0 && !desc))
    return AKZ_ERR_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return AKZ_ERR_ARG;
  int dsize = save_desc ? desc_cols : 1;
  fprintf(f, "%d\n%d\n", dsize, n);
  for (int i = 0; i < n; i++) {
    float sc = kps[i].size;
    sc *= sc;
    double inv = 1.0 / sc;
    fprintf(f, "%g %g %g %g %g", (double)kps[i].x, (double)kps[i].y, inv, 0.0, inv);
    if (save_desc)
      for (int j = 0; j < desc_cols; j++) fprintf(f, " %d", (int)desc[(size_t)i * desc_cols + j]);
    fputc('\n', f);
  }
  int bad = ferror(f);
  if (fclose(f) != 0 || bad) return AKZ_ERR_ARG;
  return AKZ_OK;
}

// read_homography, utils.cpp:340-380: three lines of three floats, divided by h33.
akz_status akz_read_homography(const char* path, double* H9) {
  if (!path || !H9) return AKZ_ERR_ARG;
  FILE* f = fopen(path, "r");
  if (!f) return AKZ_ERR_ARG;
  float h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  char line[256];
  for (int r = 0; r < 3; r++) {
    if (!fgets(line, sizeof line, f)) break;
    sscanf(line, "%f %f %f", &h[3 * r], &h[3 * r + 1], &h[3 * r + 2]);
  }
  fclose(f);
  float h33 = h[8];
  for (int i = 0; i < 9; i++) H9[i] = (double)(float)(h[i] / h33);
  return AKZ_OK;
}

// matches2points_nndr, utils.cpp:154-170: the pair (train[queryIdx].pt,
// query[trainIdx].pt) of every row whose best distance is below nndr times
// the second; pairs as x1 y1 x2 y2.
akz_status akz_matches2points_nndr(const akz_keypoint* train, int ntrain, const akz_keypoint* query, int nquery,
                                   const akz_dmatch* matches, int nrows, float nndr, float* pairs, int* npairs) {
  if (nrows < 0 || !npairs || (nrows > 0 && (!matches || !pairs))) return AKZ_ERR_ARG;
  int k = 0;
  for (int i = 0; i < nrows; i++) {
    const akz_dmatch& m0 = matches[2 * i];
    const akz_dmatch& m1 = matches[2 * i + 1];
    float dist1 = m0.distance, dist2 = m1.distance;
    if (dist1 < nndr * dist2) {
      if (m0.queryIdx < 0 || m0.queryIdx >= ntrain || m0.trainIdx < 0 || m0.trainIdx >= nquery) return AKZ_ERR_ARG;
      pairs[4 * k + 0] = train[m0.queryIdx].x;
      pairs[4 * k + 1] = train[m0.queryIdx].y;
      pairs[4 * k + 2] = query[m0.trainIdx].x;
      pairs[4 * k + 3] = query[m0.trainIdx].y;
      k++;
    }
  }
  *npairs = k;
  return AKZ_OK;
}

// compute_inliers_homography, utils.cpp:203-243, in its float arithmetic: a
// pair is an inlier when H maps its first point within min_error pixels of
// its second.
akz_status akz_inliers_homography(const float* pairs, int npairs, const double* H9, float min_error, uint8_t* mask,
                                  int* ninliers) {
  if (npairs < 0 || !H9 || !ninliers || (npairs > 0 && (!pairs || !mask))) return AKZ_ERR_ARG;
  float h[9];
  for (int i = 0; i < 9; i++) h[i] = (float)H9[i];
  int k = 0;
  for (int i = 0; i < npairs; i++) {
    float x1 = pairs[4 * i], y1 = pairs[4 * i + 1], x2 = pairs[4 * i + 2], y2 = pairs[4 * i + 3];
    float s = h[6] * x1 + h[7] * y1 + h[8];
    float x2m = (h[0] * x1 + h[1] * y1 + h[2]) / s;
    float y2m = (h[3] * x1 + h[4] * y1 + h[5]) / s;
    float dist = (float)std::sqrt(std::pow((double)(x2m - x2), 2) + std::pow((double)(y2m - y2), 2));
    mask[i] = dist <= min_error ? 1 : 0;
    k += mask[i];
  }
  *ninliers = k;
  return AKZ_OK;
}

// compute_inliers_ransac (utils.cpp:173-200) with this library's estimator in
// place of cv::findHomography: seeded 4-point samples (collinear triples
// skipped), normalised DLT, inliers at reprojection distance <= error, the
// sample count adapted to the inlier ratio for `confidence`, then one DLT
// refit on the best consensus set, kept if it holds at least as many
// inliers.  Fewer than 9 pairs give no inliers, as in the reference.
akz_status akz_inliers_ransac(const float* pairs, int npairs, float error, uint64_t seed, int max_iters,
                              double confidence, double* H9, uint8_t* mask, int* ninliers) {
  if (npairs < 0 || !ninliers || (npairs > 0 && (!pairs || !mask)) || !(error > 0) || max_iters < 1 ||
      !(confidence > 0 && confidence < 1))
    return AKZ_ERR_ARG;
  *ninliers = 0;
  for (int i = 0; i < npairs; i++) mask[i] = 0;
  if (H9)
    for (int i = 0; i < 9; i++) H9[i] = i % 4 == 0 ? 1.0 : 0.0;
  if (npairs <= 8) return AKZ_OK;
  std::vector<double> x1(npairs), y1(npairs), x2(npairs), y2(npairs);
  for (int i = 0; i < npairs; i++) {
    x1[i] = pairs[4 * i]; y1[i] = pairs[4 * i + 1]; x2[i] = pairs[4 * i + 2]; y2[i] = pairs[4 * i + 3];
  }
  const double thr2 = (double)error * (double)error;
  Rng rng{seed};
  double best[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  int best_n = 0;
  double best_e = 0;
  long iters = max_iters;
  for (long it = 0; it < iters; it++) {
    int s[4];
    for (int k = 0; k < 4; k++) {
      bool dup;
      do {
        s[k] = rng.below(npairs);
        dup = false;
        for (int j = 0; j < k; j++) dup |= s[j] == s[k];
      } while (dup);
    }
    if (degenerate(x1.data(), y1.data(), s) || degenerate(x2.data(), y2.data(), s)) continue;
    double H[9];
    if (!dlt(x1.data(), y1.data(), x2.data(), y2.data(), s, 4, H)) continue;
    double e = 0;
    int n = score(H, x1.data(), y1.data(), x2.data(), y2.data(), npairs, thr2, nullptr, &e);
    if (n > best_n || (n == best_n && n > 0 && e < best_e)) {
      best_n = n;
      best_e = e;
      memcpy(best, H, sizeof best);
      double w = (double)n / npairs;
      double p4 = w * w * w * w;
      if (p4 >= 1.0) {
        iters = it + 1;
      } else if (p4 > 0) {
        double need = std::log(1.0 - confidence) / std::log(1.0 - p4);
        if (need < (double)iters) iters = std::max((long)std::ceil(need), it + 1);
      }
    }
  }
  if (best_n < 4) return AKZ_OK;
  std::vector<uint8_t> m(npairs);
  score(best, x1.data(), y1.data(), x2.data(), y2.data(), npairs, thr2, m.data(), nullptr);
  std::vector<int> in;
  for (int i = 0; i < npairs; i++)
    if (m[i]) in.push_back(i);
  double R[9];
  if (dlt(x1.data(), y1.data(), x2.data(), y2.data(), in.data(), (int)in.size(), R)) {
    std::vector<uint8_t> m2(npairs);
    int n2 = score(R, x1.data(), y1.data(), x2.data(), y2.data(), npairs, thr2, m2.data(), nullptr);
    if (n2 >= best_n) {
      best_n = n2;
      m.swap(m2);
      memcpy(best, R, sizeof best);
    }
  }
  for (int i = 0; i < npairs; i++) mask[i] = m[i];
  *ninliers = best_n;
  if (H9) memcpy(H9, best, sizeof best);
  return AKZ_OK;
}

}  // extern "C"
