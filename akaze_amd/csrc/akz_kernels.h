// Host-side launch wrappers of the HIP kernels (internal interface between
// the C-ABI layer and the kernel translation units).
#pragma once
#include "akz_common.h"

struct akz_taps { float k[11]; };

// ---- scale space (akz_scale_space.hip) ----
void akz_launch_gauss(hipStream_t s, const float* in, int in_pitch, float* out, int out_pitch, int w, int h, int R,
                      const akz_taps& taps);
void akz_launch_contrast_init(hipStream_t s, uint32_t* maxbits, int32_t* hist, uint32_t* ticket);
void akz_launch_contrast_grad(hipStream_t s, const float* blur, int pitch, float* grad, int w, int h,
                              uint32_t* maxbits);
// The last block of the histogram computes the contrast factor on the device
// and writes the per-level value (0.75 decay per octave) into kc[0..nlev).
void akz_launch_contrast_hist(hipStream_t s, const float* grad, int pitch, int w, int h, const uint32_t* maxbits,
                              int nbins, float perc, int32_t* hist, uint32_t* ticket, float* kc, int nlev,
                              const int* octaves_host);
void akz_launch_flow(hipStream_t s, const float* Lsm, int pitch, float* g, int w, int h, int type, const float* kc);
void akz_launch_fed_step(hipStream_t s, const float* Lin, const float* g, float* Lout, int pitch, int w, int h,
                         float half_tau);
void akz_launch_halfsample(hipStream_t s, const float* in, int iw, int ih, int ipitch, float* out, int ow, int oh,
                           int opitch);
void akz_launch_copy2d(hipStream_t s, const float* in, int ipitch, float* out, int opitch, int w, int h);

// ---- detection (akz_detect.hip) ----
struct akz_det_params {
  int nlev;
  akz_level_desc lev[AKZ_MAX_LEVELS];
  const float* Lsm[AKZ_MAX_LEVELS];
  float* Lx[AKZ_MAX_LEVELS];
  float* Ly[AKZ_MAX_LEVELS];
  float* det[AKZ_MAX_LEVELS];
  int tile_start[AKZ_MAX_LEVELS + 1];  // prefix of 64x16 tiles per level
  int seg_total;                       // row segments over all levels
  float thr;
  float fac1, fac2;
};
void akz_launch_hessian(hipStream_t s, const akz_det_params& p);
void akz_launch_extrema(hipStream_t s, const akz_det_params& p, int32_t* seg_count, uint8_t* seg_slots);
// scratch holds ceil(seg_total/4096) + 1 ints
void akz_launch_scan(hipStream_t s, const akz_det_params& p, const int32_t* seg_count, int32_t* seg_off,
                     int32_t* level_start, int32_t* ncand, int cap, int32_t* scratch);
void akz_launch_scatter(hipStream_t s, const akz_det_params& p, const int32_t* seg_count, const int32_t* seg_off,
                        const uint8_t* seg_slots, akz_kp* cand, int cap);
void akz_launch_neighbors(hipStream_t s, const akz_kp* cand, const int32_t* ncand, const int32_t* level_start,
                          int32_t* nb_count, int32_t* nb_list, int32_t* flags, int cap);
void akz_launch_nms(hipStream_t s, const akz_kp* cand, const int32_t* ncand, const int32_t* nb_count,
                    const int32_t* nb_list, int32_t* mem, int32_t* minnb, int32_t* target, int32_t* act,
                    akz_kp* out, int32_t* nout, int32_t* waves, int max_waves, int cap);

// ---- description (akz_describe.hip) ----
struct akz_desc_params {
  const float* Lt[AKZ_MAX_LEVELS];
  const float* Lx[AKZ_MAX_LEVELS];
  const float* Ly[AKZ_MAX_LEVELS];
  int w[AKZ_MAX_LEVELS], h[AKZ_MAX_LEVELS], pitch[AKZ_MAX_LEVELS];
  int patsize;
};
void akz_describe_init_tables();
void akz_launch_orientation(hipStream_t s, const akz_desc_params& p, akz_kp* kps, const int32_t* nkps, int cap);
void akz_launch_mldb(hipStream_t s, const akz_desc_params& p, const akz_kp* kps, const int32_t* nkps,
                     uint8_t* desc64, int cap);

// ---- matcher (akz_match.hip) ----
int akz_match_splits(int nq, int nt);
void akz_launch_match(hipStream_t s, const uint8_t* q64, int nq, const uint8_t* t64, int nt, uint32_t* partial,
                      int nsplit);
void akz_launch_match_merge(hipStream_t s, const uint32_t* partial, int nq, int nsplit, akz_dmatch* out);
