// Host-side launch wrappers of the HIP kernels (internal interface between
// the C-ABI layer and the kernel translation units).
#pragma once
#include "akz_common.h"

struct akz_taps { float k[11]; };
constexpr int AKZ_MAX_CLASSES_DEV = 16;  // size classes of a ragged batch (AKZ_MAX_CLASSES)

// ---- scale space (akz_scale_space.hip) ----
// Every launcher takes the batch (akz_batch): pointers name image 0's arrays,
// image z's sit z * bt.zs bytes further; the input image of
// akz_launch_contrast has its own image stride in_zs.
//
// Level 0 and the contrast factor of ContrastPercentile on the device: the
// soffset Gaussian of `img` into `lt` (radius R0, taps0) and the gradient of
// its sigma=1 blur (taps_c) into `grad` in one launch, with the tile maxima
// in `blockmax` (akz_contrast_front_blocks(w, h) entries); then the
// histogram and the percentile (bin totals included), leaving the per-level factor
// in kc[0..nlev).  hist_tot: 512 ints (the histogram); maxbits: one word.
int akz_contrast_front_blocks(int w, int h);
void akz_launch_contrast(hipStream_t s, const float* img, int in_pitch, size_t in_zs, float* lt, float* grad,
                         int pitch, int w, int h, int R0, const akz_taps& taps0, const akz_taps& taps_c, int nbins,
                         float perc, uint32_t* maxbits, uint32_t* blockmax, int32_t* hist_tot, float* kc, int nlev,
                         const int* octaves_host, akz_batch bt);
// FED kernel variant chosen for a level size and the images per launch
int akz_fed_variant_for(int w, int h, int nz);
// blocked FED variants are 2..AKZ_FED_VARIANT_MAX (akz_launch_fed_variant)
constexpr int AKZ_FED_VARIANT_MAX = 19;
// LowPass(sigma=1) + Flow of one level in one kernel; taps: 5 taps of sigma 1
void akz_launch_smooth_flow(hipStream_t s, const float* in, int pitch, float* lsm, float* g, int w, int h, int type,
                            const float* kc, const akz_taps& taps, akz_batch bt);
void akz_launch_fed_step(hipStream_t s, const float* Lin, const float* g, float* Lout, int pitch, int w, int h,
                         float half_tau, akz_batch bt);
// most FED steps any variant runs in one launch
int akz_fed_block_steps();
// FED steps per launch of a variant (its halo width; 1 for variants 0 and 1)
int akz_fed_variant_steps(int variant);
void akz_launch_fed_multi(hipStream_t s, const float* Lin, const float* g, float* Lout, int pitch, int w, int h,
                          const float* half_tau, int n, akz_batch bt);
// n FED steps with one kernel variant (0 single step, 1 single step LDS-tiled,
// 2..AKZ_FED_VARIANT_MAX blocked); variants 0 and 1 need Lin != Lout and
// n == 1 for a ping-pong-free call, the blocked ones accept
// n <= akz_fed_variant_steps(variant)
void akz_launch_fed_variant(hipStream_t s, int variant, const float* Lin, const float* g, float* Lout, int pitch,
                            int w, int h, const float* half_tau, int n, akz_batch bt);
// whether a level's first launch of n steps with this variant also builds
// the level's Lsmooth and conductance (k_fed_multi_sf, k_fed_reg_sf): one
// launch instead of two; every LDS variant does, the register variants for
// one image per launch (nz = 1) and otherwise when the region's halo is small
bool akz_fed_variant_fusable(int variant, int n, int nz);
// With `parent` non-null (only for variants where akz_fed_variant_halves
// holds; `in` is then unused), the input is HalfSample2 of `parent`, an
// image of exactly twice the level's width and height with row pitch ppitch.
bool akz_fed_variant_halves(int variant);
void akz_launch_fed_variant_sf(hipStream_t s, int variant, const float* in, const float* parent, int ppitch,
                               float* lsm, float* gout, float* Lout, int pitch, int w, int h, int type,
                               const float* kc, const akz_taps& taps, const float* half_tau, int n, akz_batch bt);
// every float x in [1, +inf]: akz_rcp_ge1(x) against 1.0f / x, differing
// results added to *mismatches
void akz_launch_rcp_check(hipStream_t s, unsigned long long* mismatches);
void akz_launch_halfsample(hipStream_t s, const float* in, int iw, int ih, int ipitch, float* out, int ow, int oh,
                           int opitch, akz_batch bt);
void akz_launch_copy2d(hipStream_t s, const float* in, int ipitch, float* out, int opitch, int w, int h,
                       akz_batch bt);

// ---- detection (akz_detect.hip) ----
struct akz_det_params {
  int nlev;
  akz_level_desc lev[AKZ_MAX_LEVELS];
  const float* Lsm[AKZ_MAX_LEVELS];
  float2* Lxy[AKZ_MAX_LEVELS];  // Lx and Ly interleaved: element (y, x) at y * pitch + x
  float* det[AKZ_MAX_LEVELS];
  int tile_start[AKZ_MAX_LEVELS + 1];  // prefix of 64x16 tiles per level
  int seg_total;                       // row segments over all levels
  float thr;
  float fac1, fac2;
  int det_full;                        // 1: write every determinant; 0: only candidates' 3x3
  size_t zs;                           // image stride of the batch in bytes
  // maxima of the determinant over the 16x16 tiles of the learned proposal,
  // kept by the fused detector when tile_key is set: order-preserving keys
  // (akz_det_key) raised with atomicMax, image z's tiles at z * tiles_per_image
  uint32_t* tile_key;
  int tile_base[AKZ_MAX_LEVELS], tile_w[AKZ_MAX_LEVELS], tile_h[AKZ_MAX_LEVELS];
  int tiles_per_image;
};
// Detector work arrays of image 0 (image z's sit z * zs bytes further).
struct akz_det_bufs {
  int32_t* seg_count;    // seg_total (+4) in-segment candidate counts
  uint8_t* seg_slots;    // 32 x-offsets per segment
  int32_t* seg_wsum;     // akz_scan_blocks(seg_total) * 16 sums of 256 counts
  int32_t* level_start;  // nlev + 1
  int32_t* ncand;        // candidates after the capacity cut
  akz_kp* cand;          // cap
  float2* cand_xy;       // cap: the candidates' positions (x, y), the neighbour search's input
  int32_t* nb_count;     // cap: list lengths (lists of k_nb_pool or of the suppression's global mode)
  int32_t* nb_off;       // cap: list starts in nb_list
  int32_t* nb_pool;      // 64-bit allocator of nb_list entries (k_nb_pool), 8-byte aligned
  int32_t* nb_list;      // nb_cap
  uint2* nb_pack;        // cap: first two list entries and min(count, 3) << 30 (x)
  int32_t *mem, *minnb, *code;  // cap each (suppression state beyond LDS)
  int32_t* act;          // 2 x cap
  unsigned long long* arb;  // min(cap, 32768) 64-bit arbitration slots of the suppression's middle path
  akz_kp* kps;           // cap output keypoints
  int32_t* kidx;         // cap: candidate index of each output keypoint
  int32_t* status;       // AKZ_ST_* block
  int32_t* status_all;   // contiguous copies of every image's block (4 ints per image, not z-strided)
  int cap, nb_cap, max_waves;
  int nms_pool;          // LDS pool entries the suppression may take for lists of three or more
  size_t zs;
};
int akz_scan_blocks(int len);
int akz_suppression_pool_max();  // the suppression's LDS pool of neighbour entries
int akz_seg_runs(int seg_total);  // blocks of k_scan_scatter: 256 segments each, index seg_total included
void akz_launch_hessian(hipStream_t s, const akz_det_params& p, int nz);
void akz_launch_extrema(hipStream_t s, const akz_det_params& p, const akz_det_bufs& b, int nz);
// Hessian response + extrema in one kernel (derivative steps up to
// akz_detect_max_step()); `p32.tile_start` must count 64 x
// akz_detect_tile_rows() tiles per level
int akz_detect_max_step();
int akz_detect_tile_rows();
void akz_launch_detect_fused(hipStream_t s, const akz_det_params& p32, const akz_det_bufs& b, int nz);
// Work list of a ragged batch for the persistent detector: image i is image
// z[i] of size class cls[i], whose tiles are items start[i] .. start[i+1]-1;
// P points to the classes' fused-detector parameters in device memory.
constexpr int AKZ_LIST_IMAGES = 64;
struct akz_det_worklist {
  int nimg;
  int cls[AKZ_LIST_IMAGES], z[AKZ_LIST_IMAGES], start[AKZ_LIST_IMAGES + 1];
  const akz_det_params* P;
  int32_t* seg_count[AKZ_MAX_CLASSES_DEV];
  uint8_t* seg_slots[AKZ_MAX_CLASSES_DEV];
};
void akz_launch_detect_fused_list(hipStream_t s, const akz_det_worklist& wl, int blocks);
// segment scan, level starts, status reset and candidate materialisation
void akz_launch_candidates(hipStream_t s, const akz_det_params& p, const akz_det_bufs& b, int nz);
// CSR neighbour lists and the wave-synchronous suppression with its output
void akz_launch_suppression(hipStream_t s, const akz_det_bufs& b, int nz);

// ---- learned tile proposal (akz_proposal.hip) ----
// Tile maps: every level is cut into 16x16-pixel tiles; the tiles of an
// image are numbered level-major, row-major (base[l] is level l's first
// tile, P the tiles per image); image z's tiles follow at z * P.
constexpr int AKZ_PROP_TILE = 16;
struct akz_prop_geom {
  int nlev, P;
  int tw[AKZ_MAX_LEVELS], th[AKZ_MAX_LEVELS], oct[AKZ_MAX_LEVELS];
  int base[AKZ_MAX_LEVELS + 1];
};
// What the proposal reads from a context's batch.
struct akz_prop_src {
  akz_prop_geom g;
  const float* det[AKZ_MAX_LEVELS];
  const float* Lt[AKZ_MAX_LEVELS];
  int w[AKZ_MAX_LEVELS], h[AKZ_MAX_LEVELS], pitch[AKZ_MAX_LEVELS];
  int tile_start[AKZ_MAX_LEVELS + 1];  // prefix of 64x16-pixel tile groups per level
  float thr;                           // detector threshold (feature scale)
  size_t zs;
};
akz_prop_geom akz_prop_make_geom(int nlev, const int* w, const int* h, const int* oct);
// per-image device work arrays of the proposal inference (all images of a batch)
struct akz_prop_act {
  uint16_t *x0, *h1, *h2, *h3;  // bf16 activations, P x {4, 16, 32, 32} per image
  float *tmax, *tmean;          // tile statistics, P per image
  uint32_t* tkey;               // P per image: determinant maxima from the fused detector (keys)
  float* prob;                  // P per image
  uint8_t* mask;                // P per image
};
struct akz_prop_weights;        // device weights of a network (opaque here)
// features of images 0..nz-1 into act.x0 (tile statistics in act.tmax / tmean);
// with det_keys the determinant maxima come from act.tkey (filled by the
// fused detector) instead of the full determinant images
void akz_prop_launch_features(hipStream_t s, const akz_prop_src& src, const akz_prop_act& act, int nz,
                              bool det_keys = false);
// act.tkey of nz images to the key of -infinity, before the detector raises them
void akz_prop_launch_clear_keys(hipStream_t s, const akz_prop_geom& g, const akz_prop_act& act, int nz);
// inference on nz images' features: probabilities and the tile mask (prob >= thr)
void akz_prop_launch_infer(hipStream_t s, const akz_prop_weights* w, const akz_prop_geom& g, const akz_prop_act& act,
                           int nz, float thr);
// drops detector candidates outside masked-in tiles (segment slots, in order)
void akz_prop_launch_mask_segments(hipStream_t s, const akz_det_params& p, const akz_det_bufs& b,
                                   const akz_prop_geom& g, const uint8_t* mask, int nz);
// labels: 1 for the tiles holding a kept keypoint of the last detection
void akz_prop_launch_labels(hipStream_t s, const akz_det_bufs& b, const akz_prop_geom& g, uint8_t* labels, int nz);

struct akz_prop_net;
const akz_prop_weights* akz_prop_net_weights(const akz_prop_net* n);
// forward pass on images idx[0..nimg) (all if idx is null) of a feature set
akz_status akz_prop_forward_geom(akz_prop_net* n, const akz_prop_geom& g, const uint16_t* d_feat, const int* idx,
                                 int nimg, float* h_prob);
// loss gradient (and an Adam step when lr > 0) on images idx[0..nimg)
akz_status akz_prop_grad_geom(akz_prop_net* n, const akz_prop_geom& g, const uint16_t* d_feat,
                              const uint8_t* d_labels, const int* idx, int nimg, float wpos, float lr,
                              float* h_grad, float* h_loss);

// ---- description (akz_describe.hip) ----
struct akz_desc_params {
  const float* Lt[AKZ_MAX_LEVELS];
  const float2* Lxy[AKZ_MAX_LEVELS];  // Lx and Ly interleaved (one gather reads both)
  int w[AKZ_MAX_LEVELS], h[AKZ_MAX_LEVELS], pitch[AKZ_MAX_LEVELS];
  int patsize;
  int nch;                 // M-LDB channels: 1 intensity, 2 + gradient magnitude, 3 + rotated derivatives
  const int16_t* cmp;      // AKZ_MLDB_TABLE first then AKZ_MLDB_TABLE second value slots of each bit
  size_t zs;
};
// M-LDB comparison list of a descriptor variant (akz_mldb_table.cpp): value
// slots (3 * cell + channel) of each bit's two cell sums, (0, 0) past the last
// bit.  Returns the bit count, -1 for an invalid variant.
constexpr int AKZ_MLDB_TABLE = 488;
int akz_mldb_table(int descriptor_size, int patsize, int nch, int16_t* a, int16_t* b);
void akz_describe_init_tables();
void akz_launch_orientation(hipStream_t s, const akz_desc_params& p, akz_kp* kps, const int32_t* nkps, int cap,
                            int nz);
void akz_launch_mldb(hipStream_t s, const akz_desc_params& p, const akz_kp* kps, const int32_t* nkps,
                     uint8_t* desc64, int cap, int nz);
// Work list of a ragged batch for the persistent orientation and MLDB
// kernels: image i is image z[i] of class cls[i]; P points to the classes'
// descriptor parameters in device memory, the per-class arrays are image 0's.
struct akz_desc_worklist {
  int nimg;
  int cls[AKZ_LIST_IMAGES], z[AKZ_LIST_IMAGES];
  const akz_desc_params* P;
  akz_kp* kps[AKZ_MAX_CLASSES_DEV];
  const int32_t* nkps[AKZ_MAX_CLASSES_DEV];
  uint8_t* desc[AKZ_MAX_CLASSES_DEV];
};
void akz_launch_describe_list(hipStream_t s, const akz_desc_worklist& wl, bool orientation, int or_blocks,
                              int mldb_blocks);

// ---- matcher (akz_match.hip) ----
int akz_match_splits(int nq, int nt);
void akz_launch_match(hipStream_t s, const uint8_t* q64, int nq, const uint8_t* t64, int nt, uint32_t* partial,
                      int nsplit);
void akz_launch_match_merge(hipStream_t s, const uint32_t* partial, int nq, int nsplit, akz_dmatch* out);
