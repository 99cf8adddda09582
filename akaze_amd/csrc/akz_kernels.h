// Host-side launch wrappers of the HIP kernels (internal interface between
// the C-ABI layer and the kernel translation units).
#pragma once
#include "akz_common.h"

struct akz_taps { float k[11]; };

// ---- scale space (akz_scale_space.hip) ----
void akz_launch_gauss(hipStream_t s, const float* in, int in_pitch, float* out, int out_pitch, int w, int h, int R,
                      const akz_taps& taps);
// Contrast factor of ContrastPercentile on the device: gradient of `blur`
// (written to `grad`), its maximum, the histogram and the percentile, leaving
// the per-level factor in kc[0..nlev).  blockmax: 2048 entries; hist_part:
// 512 x 512 + 512 ints.
void akz_launch_contrast(hipStream_t s, const float* blur, int pitch, float* grad, int w, int h, int nbins,
                         float perc, uint32_t* maxbits, uint32_t* blockmax, int32_t* hist_part, float* kc, int nlev,
                         const int* octaves_host);
// FED kernel variant chosen for a level size
int akz_fed_variant_for(int w, int h);
// LowPass(sigma=1) + Flow of one level in one kernel; taps: 5 taps of sigma 1
void akz_launch_smooth_flow(hipStream_t s, const float* in, int pitch, float* lsm, float* g, int w, int h, int type,
                            const float* kc, const akz_taps& taps);
void akz_launch_fed_step(hipStream_t s, const float* Lin, const float* g, float* Lout, int pitch, int w, int h,
                         float half_tau);
// up to akz_fed_block_steps() FED steps in one launch (n <= that many)
int akz_fed_block_steps();
void akz_launch_fed_multi(hipStream_t s, const float* Lin, const float* g, float* Lout, int pitch, int w, int h,
                          const float* half_tau, int n);
// n FED steps with one kernel variant (0 single step, 1 single step LDS-tiled,
// 2..4 blocked); variants 0 and 1 need Lin != Lout and n == 1 for a
// ping-pong-free call, the blocked ones accept n <= akz_fed_block_steps()
void akz_launch_fed_variant(hipStream_t s, int variant, const float* Lin, const float* g, float* Lout, int pitch,
                            int w, int h, const float* half_tau, int n);
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
// Hessian response + extrema in one kernel (derivative steps up to
// akz_detect_max_step())
// akz_detect_tile_rows()); `p32.tile_start` must count 64 x akz_detect_tile_rows()
// tiles per level
int akz_detect_max_step();
int akz_detect_tile_rows();
void akz_launch_detect_fused(hipStream_t s, const akz_det_params& p32, int32_t* seg_count, uint8_t* seg_slots);
// scratch holds ceil(seg_total/4096) ints; total receives the candidate count
// before the capacity cut; flags is cleared for the detection run
void akz_launch_scan(hipStream_t s, const akz_det_params& p, const int32_t* seg_count, int32_t* seg_off,
                     int32_t* level_start, int32_t* ncand, int cap, int32_t* scratch, int32_t* total,
                     int32_t* flags);
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
