/*
 * akaze_oracle.h -- CPU restatement of the A-KAZE hot path (test infrastructure).
 *
 * This library is the correctness oracle for the HIP implementation in akaze_amd/.
 * It restates, in plain C, the semantics of the reference CUDA fork of libAKAZE
 * (src/lib/cuda_akaze.cu, src/lib/AKAZE.cpp, src/lib/fed.cpp).  Every function
 * cites the reference lines it follows.  It is linked only by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py -- never by
 * the product path.
 *
 * Floating-point contract: compiled with -ffp-contract=off, every expression
 * keeps the reference's operand order, so results are reproducible bit-for-bit
 * by any IEEE-754 implementation that evaluates the same expressions.
 *
 * How each function is pinned is in tests/test_oracle_stages.py and
 * tests/test_proposal_oracle.py; the functions with no external pin are marked
 * "parity unpinned" below.
 */
#ifndef AKAZE_ORACLE_H
#define AKAZE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LEVELS 32
#define ORC_MAX_TAU 256

/* Options: the subset of AKAZEOptions (src/lib/AKAZEConfig.h:84-111) on the path. */
typedef struct {
  int omax;
  int nsublevels;
  float soffset;
  float derivative_factor;
  int diffusivity; /* 0 PM_G1, 1 PM_G2, 2 WEICKERT, 3 CHARBONNIER */
  float dthreshold;
  float min_dthreshold;
  int descriptor_pattern_size;
  float kcontrast_percentile;
  int kcontrast_nbins;
  int nms_max_waves; /* cuda_akaze.cu:1100 uses 10000 */
  int descriptor;          /* 5 MLDB, 4 MLDB_UPRIGHT (AKAZEConfig.h:41-57) */
  int descriptor_size;     /* 0 = full length, else the bits of the subsample */
  int descriptor_channels; /* 1 intensity, 2 + gradient magnitude, 3 + rotated derivatives */
} orc_options;

/* cv::KeyPoint wire layout (28 bytes).  For detector candidates `octave` holds
 * the float bits written at cuda_akaze.cu:763-764 and `angle` holds the y
 * subpixel offset (cuda_akaze.cu:769). */
typedef struct {
  float x, y, size, angle, response;
  int32_t octave;
  int32_t class_id;
} orc_kp;

/* cv::DMatch wire layout (16 bytes). */
typedef struct {
  int32_t queryIdx, trainIdx, imgIdx;
  float distance;
} orc_dmatch;

typedef struct {
  int nlev;
  int omax; /* after the <80x40 cut of AKAZE.cpp:157-160 */
  int w[ORC_MAX_LEVELS], h[ORC_MAX_LEVELS];
  int octave[ORC_MAX_LEVELS], sublevel[ORC_MAX_LEVELS];
  float esigma[ORC_MAX_LEVELS], etime[ORC_MAX_LEVELS];
  int nsteps[ORC_MAX_LEVELS]; /* FED steps from level i-1 to level i (i >= 1) */
  float tau[ORC_MAX_LEVELS][ORC_MAX_TAU];
  int deriv_step[ORC_MAX_LEVELS]; /* fRound(esigma*derivative_factor/2^o), AKAZE.cpp:328-330 */
  float border[ORC_MAX_LEVELS];   /* AKAZE.cpp:353-363 */
  float det_size[ORC_MAX_LEVELS]; /* esigma*derivative_factor, AKAZE.cpp:362 */
} orc_geom;

void orc_options_default(orc_options* o);

/* ---- FED schedule (fed.cpp:41-148) ---- */
int orc_fed_tau_by_process_time(float T, int M, float tau_max, int reordering, float* tau, int cap);
int orc_fed_tau_by_cycle_time(float t, float tau_max, int reordering, float* tau, int cap);
int orc_fed_tau_internal(int n, float scale, float tau_max, int reordering, float* tau, int cap);
int orc_fed_is_prime_internal(int number);

/* ---- level geometry (AKAZE.cpp:146-197, 322-364) ---- */
int orc_geometry(const orc_options* o, int W, int H, orc_geom* g);

/* ---- stage functions on dense (pitch == width) float images ---- */
int orc_ksize_radius(int ksize);
void orc_gauss_taps(double var, int R, float* taps);
void orc_lowpass(const float* in, float* out, int w, int h, double var, int ksize);
float orc_contrast(const float* img, int w, int h, float perc, int nbins, int32_t* hist_out, float* max_out);
/* Conductance of Lsm (cuda_akaze.cu:240-287).  type 1 (PM g2) and 3
 * (Charbonnier) are pinned (closed form of a ramp; bit-exact against the
 * device).  type 0 (PM g1) and 2 (Weickert): parity unpinned -- they call
 * expf / exp, whose last bit differs between glibc and the device library. */
void orc_flow(const float* Lsm, float* g, int w, int h, int type, float kcontrast);
void orc_fed_step(const float* L, const float* g, float* out, int w, int h, float tau);
void orc_halfsample(const float* in, int iw, int ih, float* out, int ow, int oh);
void orc_hessian(const float* Lsm, float* Lx, float* Ly, float* det, int w, int h, int step);
int orc_find_extrema(const float* det, int w, int h, float border, float thr, int level, int octave,
                     float size, orc_kp* out, int cap);
int orc_nms(const orc_kp* cand, int n, const int* level_start, int nlev, int max_waves, orc_kp* out,
            int* waves_out);
/* orc_nms, also returning the candidate index of every output keypoint. */
int orc_nms_ex(const orc_kp* cand, int n, const int* level_start, int nlev, int max_waves, orc_kp* out,
               int* waves_out, int* kidx_out);
void orc_orientation(orc_kp* kp, const float* Lx, const float* Ly, int w, int h);
void orc_mldb(const orc_kp* kp, const float* Lt, const float* Lx, const float* Ly, int w, int h,
              int patsize, uint8_t* desc61, float* vals87);
/* M-LDB with nch channels and a comparison list (orc_mldb_table) of nbits
 * entries into the 87 cell values; desc receives ceil(nbits / 8) bytes. */
void orc_mldb_ex(const orc_kp* kp, const float* Lt, const float* Lx, const float* Ly, int w, int h, int patsize,
                 int nch, const int32_t* idx1, const int32_t* idx2, int nbits, uint8_t* desc, float* vals87);
int orc_descriptor_subsample(int nbits, int patsize, int nch, int32_t* samples3, int32_t* comps2);
int orc_mldb_table(int descriptor_size, int patsize, int nch, int32_t* idx1, int32_t* idx2);
void orc_match2nn(const uint8_t* q, int nq, const uint8_t* t, int nt, int stride, orc_dmatch* out);
void orc_compare_tables(int32_t* idx1, int32_t* idx2); /* 488 entries each, cuda_akaze.cu:2034-2111 */

/* ---- math used where the reference calls device transcendentals ---- */
float orc_atan2f(float y, float x);
void orc_sincosf(float x, float* s, float* c);
float orc_orient_weight(int r2);

/* ---- whole pipeline (AKAZE.cpp:211-449) ---- */
typedef struct orc_ctx orc_ctx;
orc_ctx* orc_create(const orc_options* o, int W, int H);
void orc_destroy(orc_ctx* c);
const orc_geom* orc_get_geom(const orc_ctx* c);
int orc_scale_space(orc_ctx* c, const float* img);
float orc_get_kcontrast(const orc_ctx* c, int level);
/* which: 0 Lt, 1 Lsmooth, 2 Lx, 3 Ly, 4 Ldet, 5 Lflow */
const float* orc_level(const orc_ctx* c, int lvl, int which);
int orc_detect(orc_ctx* c);
int orc_num_candidates(const orc_ctx* c);
const orc_kp* orc_candidates(const orc_ctx* c);
const int* orc_level_start(const orc_ctx* c);
int orc_nms_waves(const orc_ctx* c);
int orc_num_keypoints(const orc_ctx* c);
const orc_kp* orc_keypoints(const orc_ctx* c);
/* candidate index of every kept keypoint (orc_candidates order) */
const int* orc_kept_candidates(const orc_ctx* c);

/* ---- Learned tile proposal (test oracle of akaze_amd/csrc/akz_proposal.hip;
 * no reference counterpart).  Tile maps of 16x16 pixels per level, numbered
 * level-major and row-major; features x0 (4 bf16 per tile), a 3x3 conv net
 * 4 -> 16 -> 32 -> 32 with relu and a 1x1 sigmoid head.  Parameters are the
 * flat FP32 vector documented in include/akaze_mi.h (14513 values). ---- */
typedef struct {
  int nlev, P;
  int tw[ORC_MAX_LEVELS], th[ORC_MAX_LEVELS], oct[ORC_MAX_LEVELS];
  int base[ORC_MAX_LEVELS + 1];
} orc_prop_geom;
void orc_prop_make_geom(const orc_ctx* c, orc_prop_geom* g);
/* features of the context's levels (after orc_detect): P x 4 bf16 */
void orc_prop_features(const orc_ctx* c, float thr, uint16_t* x0);
/* labels (after orc_detect): 1 where a kept keypoint's candidate pixel lies */
void orc_prop_labels(const orc_ctx* c, uint8_t* lab);
/* Weighted BCE loss over nimg images' tiles and, if grad is not null, its
 * gradient (flat parameter order); prob (optional) receives the tile
 * probabilities.  emulate_bf16 = 1 rounds weights, activations and
 * back-propagated gradients to bf16 where the device feeds MFMAs; 0 computes
 * everything in double (the reference for finite differences). */
double orc_prop_loss_grad(const orc_prop_geom* g, int nimg, const float* params, const uint16_t* x0,
                          const uint8_t* lab, float wpos, int emulate_bf16, double* grad, float* prob);
/* Adam (Kingma & Ba 2015), beta 0.9 / 0.999, eps 1e-8, bias-corrected; `step`
 * counts from 1. */
void orc_prop_adam(float* p, float* m, float* v, const double* g, int n, float lr, int step);
int orc_describe(orc_ctx* c);
int orc_descriptor_bytes(const orc_ctx* c);
const uint8_t* orc_descriptors(const orc_ctx* c);
int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
