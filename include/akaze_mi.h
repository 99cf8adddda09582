/*
 * akaze_mi.h -- C ABI of the MI355X A-KAZE library (libakaze_amd.so).
 *
 * Plain C types, plain pointers and sizes.  Every entry point returns an
 * akz_status; none of them exits the process (the reference's safeCall calls
 * exit(-1), src/lib/cudautils.h:15-21).  All state lives in a per-context
 * object with its own HIP stream (the reference keeps module-global device
 * symbols and one global copy stream, src/lib/cuda_akaze.cu:21-28).
 *
 * The functions replace the reference's kernel-wrapper layer
 * (src/lib/cuda_akaze.h:4-28) and carry what its public C++ classes need
 * (src/lib/AKAZE.h:29-143).  The C++ classes themselves are in
 * include/akaze/AKAZE.h.
 */
#ifndef AKAZE_MI_H
#define AKAZE_MI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  AKZ_OK = 0,
  AKZ_ERR_ARG = 1,        /* invalid argument */
  AKZ_ERR_HIP = 2,        /* HIP runtime error */
  AKZ_ERR_OOM = 3,        /* device or host allocation failed */
  AKZ_ERR_OVERFLOW = 4,   /* a capacity was exceeded: output truncated */
  AKZ_ERR_STATE = 5,      /* call order violated */
  AKZ_ERR_UNSUPPORTED = 6,/* option not implemented */
  AKZ_ERR_NO_DEVICE = 7   /* no HIP device visible */
} akz_status;

/* AKAZEOptions, src/lib/AKAZEConfig.h:84-143 (fields on the GPU path). */
typedef struct {
  int omax;                     /* octaves, default 4 */
  int nsublevels;               /* sublevels per octave, default 4 */
  int img_width, img_height;    /* image size fixed at creation (AKAZEConfig.h:113-114) */
  float soffset;                /* 1.6 */
  float derivative_factor;      /* 1.5; with soffset finite and > 0, and every
                                   level's derivative step below its width and
                                   height, else AKZ_ERR_ARG */
  float sderivatives;           /* 1.0 (unused by the path, kept for layout) */
  int diffusivity;              /* 0 PM_G1, 1 PM_G2 (default), 2 WEICKERT, 3 CHARBONNIER */
  float dthreshold;             /* 0.001 */
  float min_dthreshold;         /* 0.00001 */
  int descriptor;               /* 5 MLDB (the reference's GPU descriptor), 4 MLDB_UPRIGHT (angle 0, no
                                   orientation); 0-3 (SURF, M-SURF) give AKZ_ERR_UNSUPPORTED */
  int descriptor_size;          /* 0 = full length (162 bits per channel); 1..162*channels: the bits
                                   generateDescriptorSubsample picks (AKAZE.cpp:498-603) */
  int descriptor_channels;      /* 3 intensity + rotated derivatives; 2 intensity + gradient magnitude;
                                   1 intensity */
  int descriptor_pattern_size;  /* 10 */
  float kcontrast;              /* output of the contrast stage */
  float kcontrast_percentile;   /* 0.7; in (0, 1], else AKZ_ERR_ARG */
  int kcontrast_nbins;          /* 300; 1..512, else AKZ_ERR_ARG */
  int maxkeypoints;             /* 0 = capacity for every possible candidate */
  int nms_max_waves;            /* 10000, cuda_akaze.cu:1100 */
  int verbosity;
  int fed_steps_per_launch;     /* 0 = library choice; 1 = one FED step per kernel launch */
} akz_options;

/* cv::KeyPoint wire layout, 28 bytes (device arrays are memcpy'd into
 * std::vector<cv::KeyPoint> at src/lib/cuda_akaze.cu:1405-1411). */
typedef struct {
  float x, y, size, angle, response;
  int32_t octave, class_id;
} akz_keypoint;

/* cv::DMatch wire layout, 16 bytes (src/lib/AKAZE.cpp:62-77). */
typedef struct {
  int32_t queryIdx, trainIdx, imgIdx;
  float distance;
} akz_dmatch;

/* AKAZETiming, src/lib/AKAZEConfig.h:61-80 (milliseconds, GPU event time). */
typedef struct {
  double kcontrast, scale, derivatives, detector, extrema, subpixel, descriptor;
} akz_timing;

typedef struct akz_ctx akz_ctx;
typedef struct akz_matcher akz_matcher;

/* Defaults of AKAZEOptions::AKAZEOptions(), AKAZEConfig.h:86-111. */
akz_status akz_options_default(akz_options* opts);
const char* akz_status_string(akz_status s);
akz_status akz_device_count(int* n);

/* AKAZE::AKAZE + Allocate_Memory_Evolution (AKAZE.cpp:126-208; AllocBuffers,
 * cuda_akaze.cu:479-538).  Level geometry, FED schedule and one device slab. */
akz_status akz_create(const akz_options* opts, akz_ctx** out);
/* AKAZE::~AKAZE (AKAZE.cpp:140-143; FreeBuffers, cuda_akaze.cu:541). */
akz_status akz_destroy(akz_ctx* ctx);
/* Level count after the <80x40 cut (AKAZE.cpp:157-160) and each level's size. */
akz_status akz_num_levels(const akz_ctx* ctx, int* nlev);
akz_status akz_level_size(const akz_ctx* ctx, int lvl, int* w, int* h, int* octave, int* sublevel,
                          float* esigma);

/* AKAZE::Create_Nonlinear_Scale_Space (AKAZE.cpp:211-276): contrast factor,
 * Gaussian, conductance, FED cycles and downsampling.  `img` is a host
 * float32 image in [0,1], `pitch_bytes` apart per row (0 = dense). */
akz_status akz_build_scale_space(akz_ctx* ctx, const float* img, size_t pitch_bytes);
/* Same, for an image already in device memory.  Returns without waiting. */
akz_status akz_build_scale_space_device(akz_ctx* ctx, const float* d_img, size_t pitch_bytes);

/* AKAZE::Feature_Detection (AKAZE.cpp:316-378): Hessian response, in-level
 * extrema with subpixel fit, and the cross-level suppression of FilterExtrema
 * (cuda_akaze.cu:1318-1402).  Writes the number of kept keypoints. */
akz_status akz_detect(akz_ctx* ctx, int* n_out);
/* AKAZE::Compute_Descriptors (AKAZE.cpp:405-449): orientation
 * (FindOrientation, cuda_akaze.cu:2114-2188) and MLDB descriptors
 * (ExtractDescriptors + BuildDescriptor, :1420-1533, :1786-1837).  Copies at
 * most `cap` keypoints (28 B each) and descriptors (akz_descriptor_bytes
 * each: 61 for the default 486 bits) to the host. */
akz_status akz_describe(akz_ctx* ctx, akz_keypoint* kps, uint8_t* desc, int cap, int* n_out);
/* Bytes (and bits) of one descriptor row of the context's M-LDB variant:
 * ceil(bits / 8), bits = descriptor_size or 162 * descriptor_channels
 * (the reference's desc.cols, AKAZE.cpp:412-423).  Host readbacks use rows of
 * this many bytes; device rows stay 64 bytes, zero past the last bit. */
akz_status akz_descriptor_bytes(const akz_ctx* ctx, int* nbytes, int* nbits);

/* Asynchronous whole-image pipeline for throughput: scale space + detection +
 * description of a device-resident image, results left on the device.  No
 * host synchronisation; `akz_wait` waits for the context's stream. */
akz_status akz_process_device_async(akz_ctx* ctx, const float* d_img, size_t pitch_bytes);
akz_status akz_wait(akz_ctx* ctx);
/* Device pointers of the results of the last pipeline run (count, keypoints
 * and 64-byte-padded descriptor rows); valid until the next run. */
akz_status akz_results_device(akz_ctx* ctx, const int32_t** d_count, const akz_keypoint** d_kps,
                              const uint8_t** d_desc64);
/* Host copy of results after akz_process_device_async + akz_wait. */
akz_status akz_read_results(akz_ctx* ctx, akz_keypoint* kps, uint8_t* desc, int cap, int* n_out);
/* The context's own device input buffer (level-0 size, `pitch_bytes` per
 * row).  An image written there is processed by akz_process_device_async
 * without a device-to-device copy. */
akz_status akz_input_buffer(akz_ctx* ctx, float** d_ptr, size_t* pitch_bytes);
/* Capture the asynchronous pipeline into a HIP graph on the next call (1) or
 * run it as eager launches (0). */
akz_status akz_set_graph_mode(akz_ctx* ctx, int enabled);
/* The asynchronous pipeline (akz_process_*) writes only the determinant
 * pixels its subpixel fit reads (the 3x3 neighbourhood of every candidate);
 * keep = 1 makes it write the whole determinant image of every level, as
 * akz_detect always does, for akz_read_level(..., 4, ...) and
 * akz_prop_extract.  Both return AKZ_ERR_STATE after a run without it.  A
 * context with a proposal network always writes them. */
akz_status akz_set_keep_determinant(akz_ctx* ctx, int keep);
/* No counterpart in the reference: a test and tuning hook of the
 * suppression.  Up to 8192 candidates it keeps the neighbour lists in LDS,
 * the entries past the first of lists of three or more in a pool of at most
 * 4096; a list that does not fit sends the image's suppression to the
 * path that reads the lists from global memory (same result).  `entries` (0..4096, default 4096) caps
 * that pool; 0 sends every image with such a list down the global path. */
akz_status akz_set_suppression_pool(akz_ctx* ctx, int entries);

/* ---- Batches of same-size images.  The reference processes one image per
 * AKAZE object and call (AKAZE.cpp:211-449); these entry points have no
 * counterpart there and serve callers with many images of one size.  A batch
 * context holds `batch` images with identical device layouts; every stage
 * runs as one launch over all images of a call (grid z = image), and results
 * equal those of one single-image context per image. ---- */
#define AKZ_MAX_BATCH 64
akz_status akz_create_batch(const akz_options* opts, int batch, akz_ctx** out);
akz_status akz_batch_capacity(const akz_ctx* ctx, int* batch);
/* Image 0's input buffer; image z's starts `image_stride_bytes` * z later. */
akz_status akz_batch_input_buffer(akz_ctx* ctx, float** d_ptr, size_t* pitch_bytes, size_t* image_stride_bytes);
/* Images 0..n-1 (n <= batch) at d_imgs + z * image_stride_bytes (0 = dense
 * images), scale space + detection + description, asynchronous.  In graph
 * mode one graph per batch size is captured on first use. */
akz_status akz_process_batch_device_async(akz_ctx* ctx, const float* d_imgs, size_t pitch_bytes,
                                          size_t image_stride_bytes, int n);
/* Waits; per-image kept counts and statuses (AKZ_OK or AKZ_ERR_OVERFLOW). */
akz_status akz_read_batch_counts(akz_ctx* ctx, int n, int* counts, int* statuses);
/* Host copy of image z's results (as akz_read_results); waits. */
akz_status akz_read_batch_results(akz_ctx* ctx, int z, akz_keypoint* kps, uint8_t* desc, int cap, int* n_out);
akz_status akz_batch_results_device(akz_ctx* ctx, int z, const int32_t** d_count, const akz_keypoint** d_kps,
                                    const uint8_t** d_desc64);
akz_status akz_read_batch_level(akz_ctx* ctx, int z, int lvl, int which, float* dst);
/* Autotuner (no reference counterpart): times every blocked FED kernel
 * variant (LDS or register region shape, wavefront strips, threads per
 * block, 8 or 16 steps per launch) over each level's smooth + flow and
 * whole FED run, as the
 * pipeline runs them (the LDS variants build Lsmooth and the conductance in
 * their first launch), on the level's own buffers with batches of nz images,
 * and keeps the fastest per level size.  Overwrites the level contents; captured graphs are rebuilt on
 * the next call.  Variants are readable and settable per level (2..19). */
akz_status akz_autotune(akz_ctx* ctx, int nz, int reps);
akz_status akz_get_fed_variants(const akz_ctx* ctx, int* variants);
akz_status akz_set_fed_variants(akz_ctx* ctx, const int* variants);

/* Ragged batches (no reference counterpart): images of several sizes go to
 * one batch context per size class; each class runs as one batch call on
 * its own stream, concurrently with the others.  At most `capacity[k]`
 * images of class k per call. */
#define AKZ_MAX_CLASSES 16
typedef struct akz_ragged akz_ragged;
akz_status akz_ragged_create(const akz_options* base, int nclass, const int* widths, const int* heights,
                             const int* capacity, akz_ragged** out);
akz_status akz_ragged_destroy(akz_ragged* r);
/* The batch context of class `cls` (owned by r). */
akz_status akz_ragged_context(akz_ragged* r, int cls, akz_ctx** ctx);
akz_status akz_ragged_set_graph_mode(akz_ragged* r, int enabled);
/* Work lists (1): the scale space runs per class on concurrent streams, then
 * the detector, the orientation and the MLDB stages each run as ONE
 * persistent launch over all images of all classes, walking
 * (image, level, tile) and (image, keypoint) work lists; the candidate and
 * suppression stages run per class in between.  In graph mode the whole call
 * is one captured graph per class composition (images per class).
 * 0 (the default): one batch call per class (each class's own graph), the
 * classes on concurrent streams, which overlap one class's scale space with
 * another's detection.  Results are identical either way. */
akz_status akz_ragged_set_work_lists(akz_ragged* r, int enabled);
akz_status akz_ragged_autotune(akz_ragged* r, int reps);
/* Processes n device images (sizes widths[i] x heights[i], each one of the
 * classes; pitch 0 = dense) and waits; counts[i] receives image i's kept
 * keypoint count. */
akz_status akz_ragged_process(akz_ragged* r, int n, const float* const* d_imgs, const size_t* pitch_bytes,
                              const int* widths, const int* heights, int* counts);
akz_status akz_ragged_read(akz_ragged* r, int i, akz_keypoint* kps, uint8_t* desc, int cap, int* n_out);
akz_status akz_ragged_results_device(akz_ragged* r, int i, const int32_t** d_count, const akz_keypoint** d_kps,
                                     const uint8_t** d_desc64);

/* akz_profile_fed over a batch of n images (bytes count all n images). */
akz_status akz_profile_fed_batch(akz_ctx* ctx, const float* d_imgs, size_t pitch_bytes, size_t image_stride_bytes,
                                 int n, float* fed_ms, double* fed_bytes, int* fed_launches, double* min_bytes);

/* Debug readback of a level image (AKAZE::Save_Scale_Space reads Lt back,
 * AKAZE.cpp:453-463).  which: 0 Lt, 1 Lsmooth, 2 Lx, 3 Ly, 4 Ldet.
 * dst is dense w*h floats. */
akz_status akz_read_level(akz_ctx* ctx, int lvl, int which, float* dst);
/* Contrast factor used at level `lvl` (options_.kcontrast after the 0.75 per
 * octave decay, AKAZE.cpp:254-256). */
akz_status akz_read_kcontrast(akz_ctx* ctx, int lvl, float* k);
/* Detector candidates before suppression in (level, y, x) order, candidate
 * count per level, and suppression wave count: the state FilterExtrema
 * works on (cuda_akaze.cu:1318-1402). */
akz_status akz_read_candidates(akz_ctx* ctx, akz_keypoint* cand, int cap, int* n_out, int32_t* level_start,
                               int* waves);
/* Measurement: runs the whole pipeline on `d_img` with HIP events around each
 * level's FED launches and returns their summed time, the algorithmic bytes of
 * those launches (12 B per pixel and step, as if every step were a pass of
 * its own), the launch count, and the bytes the launches must move however
 * many steps each runs (L and g read once and L written once per launch;
 * min_bytes may be null).  Waits. */
akz_status akz_profile_fed(akz_ctx* ctx, const float* d_img, size_t pitch_bytes, float* fed_ms, double* fed_bytes,
                           int* fed_launches, double* min_bytes);
/* Measurement: a plain device copy of `nfloats` floats (a multiple of 4096),
 * `iters` times on the null stream; average ms per launch.  Used to calibrate
 * hardware byte counters against a kernel of known traffic. */
akz_status akz_copy_kernel_run(const float* d_src, float* d_dst, size_t nfloats, int iters, float* avg_ms);
/* AKAZE::Get_Computation_Times (AKAZE.h:139-141): stage times of the last
 * eager run.  AKZ_ERR_STATE when that run recorded none (a captured graph
 * replays without events). */
akz_status akz_get_timing(akz_ctx* ctx, akz_timing* t);

/* ---- Learned tile proposal.  No counterpart in the reference.  A small
 * convolutional network over 16x16-pixel tile maps of every Hessian-response
 * level predicts the tiles that can hold kept keypoints.  With a network set
 * on a context, detector candidates outside predicted tiles are dropped
 * before the suppression; the full scan (no network) stays the fallback.
 * Parameters: 14513 FP32 values in the flat order W1[9][4][16], b1[16],
 * W2[9][16][32], b2[32], W3[9][32][32], b3[32], w4[32], b4, with
 * W[tap][in][out] and tap = 3 (dy + 1) + (dx + 1).  Features are 4 bf16 per
 * tile, tiles numbered level-major, row-major per image. ---- */
typedef struct akz_prop_net akz_prop_net;
akz_status akz_prop_num_params(int* n);
/* He-normal weights from `seed`, zero biases, Adam state cleared. */
akz_status akz_prop_create(uint64_t seed, akz_prop_net** out);
akz_status akz_prop_destroy(akz_prop_net* net);
akz_status akz_prop_set_params(akz_prop_net* net, const float* params);
akz_status akz_prop_get_params(akz_prop_net* net, float* params);
/* Tiles per image of the context's level geometry. */
akz_status akz_prop_positions(const akz_ctx* ctx, int* per_image);
/* Features of images 0..n-1 of the context's last run into d_feat
 * (n x P x 4 bf16) and, if d_labels is not null, the labels (n x P bytes:
 * 1 where a kept keypoint's detector pixel lies in the tile).  Waits. */
akz_status akz_prop_extract(akz_ctx* ctx, int n, uint16_t* d_feat, uint8_t* d_labels);
/* Probabilities (host, nimg x P) for images idx[0..nimg) of a feature set
 * (idx null: images 0..nimg-1).  Waits. */
akz_status akz_prop_forward(akz_prop_net* net, const akz_ctx* ctx, const uint16_t* d_feat, const int* idx, int nimg,
                            float* probs);
/* Weighted binary cross-entropy (weight pos_weight on positive tiles, mean
 * over tiles) of images idx[0..nimg): gradient (host, optional), loss
 * (optional) and, when lr > 0, one Adam step (beta 0.9 / 0.999, eps 1e-8).
 * Waits. */
akz_status akz_prop_train_step(akz_prop_net* net, const akz_ctx* ctx, const uint16_t* d_feat,
                               const uint8_t* d_labels, const int* idx, int nimg, float pos_weight, float lr,
                               float* grad, float* loss);
/* Use `net` (null: full scan) with probability threshold `threshold` in the
 * context's detection.  Captured graphs are rebuilt on the next call. */
akz_status akz_set_proposal(akz_ctx* ctx, akz_prop_net* net, float threshold);
/* Tile probabilities of image z from the last detection with a network. */
akz_status akz_prop_read_probs(akz_ctx* ctx, int z, float* probs);
/* Synthetic images rendered on the device (seeds seed0 .. seed0+n-1), float
 * in [0,1], same image model as akz_synth_image with a hashed noise lattice.
 * Waits. */
akz_status akz_synth_device_batch(uint64_t seed0, int n, int w, int h, float* d_out, size_t pitch_bytes,
                                  size_t image_stride_bytes);
/* Device homography warp of a float [0,1] image (destination pixel centres
 * mapped through H^-1, bilinear, 0 outside) plus Gaussian noise of
 * `noise_std`, clipped and quantised to 1/255.  Waits. */
akz_status akz_synth_device_warp(const float* d_src, int w, int h, size_t src_pitch_bytes, const double* H9,
                                 float noise_std, uint64_t noise_seed, float* d_dst, size_t dst_pitch_bytes);

/* ---- Brute-force Hamming 2-NN matcher: Matcher (AKAZE.h:29-55, AKAZE.cpp:
 * 36-121) and MatchDescriptors (cuda_akaze.cu:1842-2031). ---- */
akz_status akz_matcher_create(akz_matcher** out);
akz_status akz_matcher_destroy(akz_matcher* m);
/* Matcher::bfmatch: q is nq rows, t is nt rows, `stride` bytes apart, of
 * which the first min(stride, 61) are the descriptor (61 for 486 bits; fewer
 * for the shorter M-LDB variants).  Writes 2*nq DMatches: best then second
 * best by (distance, trainIdx). */
akz_status akz_match2nn(akz_matcher* m, const uint8_t* q, int nq, const uint8_t* t, int nt, size_t stride,
                        akz_dmatch* out);
/* Device-resident variant: 64-byte rows (61 used, 3 zero), output in device
 * memory; asynchronous on the matcher's stream. */
akz_status akz_match2nn_device(akz_matcher* m, const uint8_t* d_q64, int nq, const uint8_t* d_t64, int nt,
                               akz_dmatch* d_out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the same 2-NN as
 * akz_match2nn_device on device-resident 64-byte rows, enqueued on `stream`
 * (a hipStream_t; null = the null stream), with no matcher object and no
 * allocation of any kind.  Its partial results live in a fixed 8 MiB array
 * per GPU, held until the process ends, so nq is at most
 * AKZ_NOALLOC_MAX_QUERIES (else AKZ_ERR_ARG); calls that overlap on one GPU
 * share that array and race.  Output equals akz_match2nn_device's. */
#define AKZ_NOALLOC_MAX_QUERIES 1048576
akz_status akz_match2nn_device_noalloc(const uint8_t* d_q64, int nq, const uint8_t* d_t64, int nt,
                                       akz_dmatch* d_out, void* stream);
akz_status akz_matcher_wait(akz_matcher* m);
/* Average duration in ms of the matcher kernels of the last call (HIP events). */
akz_status akz_matcher_last_ms(akz_matcher* m, float* ms);

/* ---- Evaluation and I/O utilities of the reference programs
 * (src/lib/utils.cpp), host only, no OpenCV.  C++ forms with the
 * reference's names are in include/akaze/utils.h. ---- */
/* cv::imread(path, 0) for netpbm greyscale files (P5 binary or P2 text;
 * 16-bit samples keep their high byte).  dst null: only the size. */
akz_status akz_read_pgm(const char* path, uint8_t* dst, size_t cap, int* width, int* height);
/* save_keypoints (utils.cpp:97-151): Mikolajczyk's text format, "dsize\nN\n"
 * then "x y 1/r^2 0 1/r^2 [desc bytes]" per keypoint, r = size; dsize is
 * desc_cols with save_desc, else 1.  desc rows are desc_cols bytes. */
akz_status akz_save_keypoints(const char* path, const akz_keypoint* kps, int n, const uint8_t* desc, int desc_cols,
                              int save_desc);
/* read_homography (utils.cpp:340-380): 3x3 floats divided by h33, row-major. */
akz_status akz_read_homography(const char* path, double* H9);
/* matches2points_nndr (utils.cpp:154-170): for each of the nrows 2-NN rows
 * (2 DMatches each) with best < nndr * second, the pair
 * (train[queryIdx].pt, query[trainIdx].pt) as 4 floats; pairs holds up to
 * 4 * nrows floats. */
akz_status akz_matches2points_nndr(const akz_keypoint* train, int ntrain, const akz_keypoint* query, int nquery,
                                   const akz_dmatch* matches, int nrows, float nndr, float* pairs, int* npairs);
/* compute_inliers_homography (utils.cpp:203-243): mask[i] = 1 where H maps
 * pair i's first point within min_error px of its second (float arithmetic
 * as the reference). */
akz_status akz_inliers_homography(const float* pairs, int npairs, const double* H9, float min_error, uint8_t* mask,
                                  int* ninliers);
/* compute_inliers_ransac (utils.cpp:173-200) with this library's own RANSAC
 * homography in place of cv::findHomography: seeded 4-point samples,
 * normalised DLT, reprojection distance <= error, adaptive sample count for
 * `confidence` (at most max_iters), a final DLT refit on the consensus set.
 * Writes H (optional), the inlier mask and count; npairs <= 8 gives none. */
akz_status akz_inliers_ransac(const float* pairs, int npairs, float error, uint64_t seed, int max_iters,
                              double confidence, double* H9, uint8_t* mask, int* ninliers);

/* ---- Device memory helpers (callers without their own HIP runtime) ---- */
akz_status akz_dev_alloc(size_t bytes, void** ptr);
akz_status akz_dev_free(void* ptr);
akz_status akz_memcpy_h2d(void* dst, const void* src, size_t bytes);
akz_status akz_memcpy_d2h(void* dst, const void* src, size_t bytes);
akz_status akz_memcpy_d2d(void* dst, const void* src, size_t bytes);
/* 2-D copy between any host or device memory (pitches and width in bytes). */
akz_status akz_memcpy_2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes,
                         size_t height);
akz_status akz_device_synchronize(void);

/* ---- Self-test (no counterpart in the reference) ----
 * The conductance's 1/x (x >= 1) uses the hardware reciprocal and one FMA
 * Newton step instead of the division sequence.  This compares that function
 * with IEEE 1.0f / x for every float x in [1, +inf] on the current device and
 * returns the number of differing results (0 expected). */
akz_status akz_selftest_reciprocal(uint64_t* mismatches);

/* ---- Seeded synthetic inputs (no reference counterpart: the benchmark
 * inputs).  uint8 image, homography about the image centre, warp. ---- */
akz_status akz_synth_image(uint64_t seed, int w, int h, uint8_t* out);
akz_status akz_synth_homography(uint64_t seed, int w, int h, double max_rot_deg, double min_scale,
                                double max_scale, double max_persp, double* H9);
akz_status akz_synth_warp(const uint8_t* src, int w, int h, const double* H9, double noise_std,
                          uint64_t noise_seed, uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif
