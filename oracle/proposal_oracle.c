/*
 * proposal_oracle.c -- CPU oracle of the learned tile proposal (test
 * infrastructure; the product is akaze_amd/csrc/akz_proposal.hip).  The
 * proposal has no reference counterpart: it is pinned by finite differences
 * of this file's own double-precision loss (tests/test_proposal_oracle.py),
 * by closed forms on hand-built inputs, and the device kernels are compared
 * with it in bf16-emulation mode.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "akaze_oracle.h"

#define TILE 16
enum { C0 = 4, C1 = 16, C2 = 32, C3 = 32 };
enum {
  OFF_W1 = 0,
  OFF_B1 = OFF_W1 + 9 * C0 * C1,
  OFF_W2 = OFF_B1 + C1,
  OFF_B2 = OFF_W2 + 9 * C1 * C2,
  OFF_W3 = OFF_B2 + C2,
  OFF_B3 = OFF_W3 + 9 * C2 * C3,
  OFF_W4 = OFF_B3 + C3,
  OFF_B4 = OFF_W4 + C3,
  NPARAM = OFF_B4 + 1
};

/* bfloat16, round to nearest even (finite inputs) */
static uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float bf2f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static double rb(double v, int emulate) { return emulate ? (double)bf2f(f2bf((float)v)) : v; }

void orc_prop_make_geom(const orc_ctx* c, orc_prop_geom* g) {
  const orc_geom* og = orc_get_geom(c);
  memset(g, 0, sizeof(*g));
  g->nlev = og->nlev;
  int P = 0;
  for (int l = 0; l < og->nlev; l++) {
    g->tw[l] = (og->w[l] + TILE - 1) / TILE;
    g->th[l] = (og->h[l] + TILE - 1) / TILE;
    g->oct[l] = og->octave[l];
    g->base[l] = P;
    P += g->tw[l] * g->th[l];
  }
  g->base[og->nlev] = P;
  g->P = P;
}

static float det_feature(float m, float thr) {
  return log2f(1.0f + (m > 0.0f ? m : 0.0f) / thr) * 0.125f;
}

/* maximum tile statistic of level n over the area of tile (ty, tx) of level l */
static float scale_nbr_max(const orc_prop_geom* g, const float* tm, int l, int n, int ty, int tx) {
  int d = g->oct[l] - g->oct[n];
  if (d == 0) return tm[g->base[n] + ty * g->tw[n] + tx];
  if (d < 0) {
    int y = ty >> 1, x = tx >> 1;
    if (y > g->th[n] - 1) y = g->th[n] - 1;
    if (x > g->tw[n] - 1) x = g->tw[n] - 1;
    return tm[g->base[n] + y * g->tw[n] + x];
  }
  float m = -INFINITY;
  for (int dy = 0; dy < 2; dy++)
    for (int dx = 0; dx < 2; dx++) {
      int y = 2 * ty + dy, x = 2 * tx + dx;
      if (y < g->th[n] && x < g->tw[n]) {
        float v = tm[g->base[n] + y * g->tw[n] + x];
        if (v > m) m = v;
      }
    }
  return m;
}

void orc_prop_features(const orc_ctx* c, float thr, uint16_t* x0) {
  orc_prop_geom g;
  orc_prop_make_geom(c, &g);
  const orc_geom* og = orc_get_geom(c);
  float* tmax = (float*)malloc(sizeof(float) * (size_t)g.P);
  float* tmean = (float*)malloc(sizeof(float) * (size_t)g.P);
  for (int l = 0; l < g.nlev; l++) {
    const float* D = orc_level(c, l, 4);
    const float* Lt = orc_level(c, l, 0);
    int w = og->w[l], h = og->h[l];
    for (int ty = 0; ty < g.th[l]; ty++)
      for (int tx = 0; tx < g.tw[l]; tx++) {
        float m = -INFINITY;
        double s = 0.0;
        int cnt = 0;
        for (int y = ty * TILE; y < ty * TILE + TILE && y < h; y++)
          for (int x = tx * TILE; x < tx * TILE + TILE && x < w; x++) {
            float v = D[(size_t)y * w + x];
            if (v > m) m = v;
            s += Lt[(size_t)y * w + x];
            cnt++;
          }
        tmax[g.base[l] + ty * g.tw[l] + tx] = m;
        tmean[g.base[l] + ty * g.tw[l] + tx] = (float)(s / cnt);
      }
  }
  for (int l = 0; l < g.nlev; l++) {
    int lm = l > 0 ? l - 1 : l, lp = l + 1 < g.nlev ? l + 1 : l;
    for (int ty = 0; ty < g.th[l]; ty++)
      for (int tx = 0; tx < g.tw[l]; tx++) {
        int p = g.base[l] + ty * g.tw[l] + tx;
        x0[4 * p + 0] = f2bf(det_feature(scale_nbr_max(&g, tmax, l, lm, ty, tx), thr));
        x0[4 * p + 1] = f2bf(det_feature(tmax[p], thr));
        x0[4 * p + 2] = f2bf(det_feature(scale_nbr_max(&g, tmax, l, lp, ty, tx), thr));
        x0[4 * p + 3] = f2bf(tmean[p]);
      }
  }
  free(tmax);
  free(tmean);
}

void orc_prop_labels(const orc_ctx* c, uint8_t* lab) {
  orc_prop_geom g;
  orc_prop_make_geom(c, &g);
  memset(lab, 0, (size_t)g.P);
  const orc_kp* cand = orc_candidates(c);
  const int* kidx = orc_kept_candidates(c);
  int n = orc_num_keypoints(c);
  for (int i = 0; i < n; i++) {
    const orc_kp* k = &cand[kidx[i]];
    int l = k->class_id;
    float ratio = (float)(1 << g.oct[l]);
    int x = (int)(k->x / ratio), y = (int)(k->y / ratio);
    lab[g.base[l] + (y / TILE) * g.tw[l] + x / TILE] = 1;
  }
}

/* global index of the tap-t neighbour of position gp, -1 outside its map */
static int nbr(const orc_prop_geom* g, int gp, int t) {
  int z = gp / g->P, p = gp - z * g->P, l = 0;
  while (l + 1 < g->nlev && p >= g->base[l + 1]) l++;
  int loc = p - g->base[l], ty = loc / g->tw[l], tx = loc - ty * g->tw[l];
  int y = ty + t / 3 - 1, x = tx + t % 3 - 1;
  if (y < 0 || y >= g->th[l] || x < 0 || x >= g->tw[l]) return -1;
  return z * g->P + g->base[l] + y * g->tw[l] + x;
}

/* y[p][co] = relu(b[co] + sum_t sum_ci W[t][ci][co] x[nbr(p,t)][ci]) */
static void conv_fwd(const orc_prop_geom* g, int T, const int* nb, int cin, int cout, const double* x,
                     const double* W, const float* b, double* y, int emulate) {
  for (int p = 0; p < T; p++)
    for (int co = 0; co < cout; co++) {
      double acc = 0.0;
      for (int t = 0; t < 9; t++) {
        int q = nb[p * 9 + t];
        if (q < 0) continue;
        for (int ci = 0; ci < cin; ci++) acc += W[(t * cin + ci) * cout + co] * x[(size_t)q * cin + ci];
      }
      double v = acc + b[co];
      y[(size_t)p * cout + co] = rb(v > 0.0 ? v : 0.0, emulate);
    }
  (void)g;
}

/* weight and bias gradients: dW[t][ci][co] = sum_p G[p][co] x[nbr(p,t)][ci] */
static void conv_bwd_w(int T, const int* nb, int cin, int cout, const double* x, const double* G, double* dW,
                       double* db) {
  for (int p = 0; p < T; p++)
    for (int co = 0; co < cout; co++) {
      double gv = G[(size_t)p * cout + co];
      if (gv == 0.0) continue;
      db[co] += gv;
      for (int t = 0; t < 9; t++) {
        int q = nb[p * 9 + t];
        if (q < 0) continue;
        for (int ci = 0; ci < cin; ci++) dW[(t * cin + ci) * cout + co] += gv * x[(size_t)q * cin + ci];
      }
    }
}

/* input gradient through the convolution, masked by the input's relu:
 * Gin[q][ci] = [x[q][ci] > 0] sum_{p, t: nbr(p,t) = q} sum_co G[p][co] W[t][ci][co] */
static void conv_bwd_data(int T, const int* nb, int cin, int cout, const double* x, const double* G,
                          const double* W, double* Gin, int emulate) {
  double* acc = (double*)calloc((size_t)T * cin, sizeof(double));
  for (int p = 0; p < T; p++)
    for (int t = 0; t < 9; t++) {
      int q = nb[p * 9 + t];
      if (q < 0) continue;
      for (int ci = 0; ci < cin; ci++) {
        double s = 0.0;
        for (int co = 0; co < cout; co++) s += G[(size_t)p * cout + co] * W[(t * cin + ci) * cout + co];
        acc[(size_t)q * cin + ci] += s;
      }
    }
  for (size_t i = 0; i < (size_t)T * cin; i++) Gin[i] = x[i] > 0.0 ? rb(acc[i], emulate) : 0.0;
  free(acc);
}

static double softplus(double z) { return (z > 0 ? z : 0.0) + log1p(exp(-fabs(z))); }

double orc_prop_loss_grad(const orc_prop_geom* g, int nimg, const float* params, const uint16_t* x0,
                          const uint8_t* lab, float wpos, int emulate, double* grad, float* prob) {
  const int T = g->P * nimg;
  int* nb = (int*)malloc(sizeof(int) * (size_t)T * 9);
  for (int p = 0; p < T; p++)
    for (int t = 0; t < 9; t++) nb[p * 9 + t] = nbr(g, p, t);
  double* W1 = (double*)malloc(sizeof(double) * 9 * C0 * C1);
  double* W2 = (double*)malloc(sizeof(double) * 9 * C1 * C2);
  double* W3 = (double*)malloc(sizeof(double) * 9 * C2 * C3);
  for (int i = 0; i < 9 * C0 * C1; i++) W1[i] = rb(params[OFF_W1 + i], emulate);
  for (int i = 0; i < 9 * C1 * C2; i++) W2[i] = rb(params[OFF_W2 + i], emulate);
  for (int i = 0; i < 9 * C2 * C3; i++) W3[i] = rb(params[OFF_W3 + i], emulate);
  double* a0 = (double*)malloc(sizeof(double) * (size_t)T * C0);
  double* a1 = (double*)malloc(sizeof(double) * (size_t)T * C1);
  double* a2 = (double*)malloc(sizeof(double) * (size_t)T * C2);
  double* a3 = (double*)malloc(sizeof(double) * (size_t)T * C3);
  for (size_t i = 0; i < (size_t)T * C0; i++) a0[i] = bf2f(x0[i]);
  conv_fwd(g, T, nb, C0, C1, a0, W1, params + OFF_B1, a1, emulate);
  conv_fwd(g, T, nb, C1, C2, a1, W2, params + OFF_B2, a2, emulate);
  conv_fwd(g, T, nb, C2, C3, a2, W3, params + OFF_B3, a3, emulate);
  double loss = 0.0;
  double* g3 = grad ? (double*)calloc((size_t)T * C3, sizeof(double)) : NULL;
  if (grad) memset(grad, 0, sizeof(double) * NPARAM);
  for (int p = 0; p < T; p++) {
    double z = params[OFF_B4];
    for (int c = 0; c < C3; c++) z += (double)params[OFF_W4 + c] * a3[(size_t)p * C3 + c];
    double pz = 1.0 / (1.0 + exp(-z));
    if (prob) prob[p] = (float)pz;
    if (!lab) continue;
    double y = lab[p] ? 1.0 : 0.0;
    loss += (wpos * y * softplus(-z) + (1.0 - y) * softplus(z)) / T;
    if (!grad) continue;
    double dz = (wpos * y * (pz - 1.0) + (1.0 - y) * pz) / T;
    grad[OFF_B4] += dz;
    for (int c = 0; c < C3; c++) {
      double h = a3[(size_t)p * C3 + c];
      grad[OFF_W4 + c] += dz * h;
      g3[(size_t)p * C3 + c] = h > 0.0 ? rb(dz * params[OFF_W4 + c], emulate) : 0.0;
    }
  }
  if (grad) {
    double* g2 = (double*)malloc(sizeof(double) * (size_t)T * C2);
    double* g1 = (double*)malloc(sizeof(double) * (size_t)T * C1);
    conv_bwd_w(T, nb, C2, C3, a2, g3, grad + OFF_W3, grad + OFF_B3);
    conv_bwd_data(T, nb, C2, C3, a2, g3, W3, g2, emulate);
    conv_bwd_w(T, nb, C1, C2, a1, g2, grad + OFF_W2, grad + OFF_B2);
    conv_bwd_data(T, nb, C1, C2, a1, g2, W2, g1, emulate);
    conv_bwd_w(T, nb, C0, C1, a0, g1, grad + OFF_W1, grad + OFF_B1);
    free(g2);
    free(g1);
  }
  free(g3);
  free(nb);
  free(W1);
  free(W2);
  free(W3);
  free(a0);
  free(a1);
  free(a2);
  free(a3);
  return loss;
}

void orc_prop_adam(float* p, float* m, float* v, const double* g, int n, float lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  for (int i = 0; i < n; i++) {
    float gi = (float)g[i];
    m[i] = b1 * m[i] + (1.0f - b1) * gi;
    v[i] = b2 * v[i] + (1.0f - b2) * gi * gi;
    p[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
  }
}
