// C-ABI of the MI355X A-KAZE library: per-context state, device memory
// layout and the orchestration of the pipeline stages on the context's
// stream.  Host code; kernels live in the other translation units.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "akz_internal.h"
#include "akz_kernels.h"

akz_status akz_hip_fail(hipError_t e, const char* what, const char* file, int line) {
  if (getenv("AKZ_VERBOSE")) fprintf(stderr, "akaze_amd: %s failed (%s) at %s:%d\n", what, hipGetErrorString(e), file, line);
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? AKZ_ERR_OOM : (e == hipErrorNoDevice ? AKZ_ERR_NO_DEVICE : AKZ_ERR_HIP);
}

namespace {

constexpr size_t ALIGN = 256;
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
inline int fround(float f) { return (int)(f + 0.5f); }

// FED schedule, the published method of Grewenig et al. as fed.cpp:41-148
// computes it (number of steps, cosine step sizes, kappa-cycle reordering
// modulo the next prime above n).
bool fed_is_prime(int number) {
  if (number <= 1) return false;
  if (number == 2 || number == 3 || number == 5 || number == 7) return true;
  if ((number % 2) == 0 || (number % 3) == 0 || (number % 5) == 0 || (number % 7) == 0) return false;
  int upper = (int)std::sqrt(number + 1.0);
  for (int d = 11; d <= upper; d += 2)
    if (number % d == 0) return false;
  return true;
}

int fed_tau(float T, float tau_max, float* tau, int cap) {
  const float t = T / (float)1;
  int n = (int)(std::ceil(std::sqrt(3.0 * t / tau_max + 0.25f) - 0.5f - 1.0e-8f) + 0.5f);
  float scale = 3.0 * t / (tau_max * (float)(n * (n + 1)));
  if (n <= 0) return 0;
  if (n > cap) return -1;
  std::vector<float> tauh(n);
  float c = 1.0f / (4.0f * (float)n + 2.0f);
  float d = scale * tau_max / 2.0f;
  for (int k = 0; k < n; ++k) {
    float h = (float)std::cos(3.14159265358979323846 * (2.0f * (float)k + 1.0f) * c);
    tauh[k] = d / (h * h);
  }
  if (n == 1) { tau[0] = tauh[0]; return 1; }
  int kappa = n / 2, prime = n + 1;
  while (!fed_is_prime(prime)) prime++;
  for (int k = 0, l = 0; l < n; ++k, ++l) {
    int index;
    while ((index = ((k + 1) * kappa) % prime - 1) >= n) k++;
    tau[l] = tauh[index];
  }
  return n;
}

// Gaussian taps of LowPass (cuda_akaze.cu:175-185).
void gauss_taps(double var, int R, akz_taps* t) {
  float sum = 0.0f;
  for (int j = -R; j <= R; j++) {
    t->k[j + R] = (float)expf((float)(-(double)j * j / 2.0 / var));
    sum += t->k[j + R];
  }
  for (int j = -R; j <= R; j++) t->k[j + R] /= sum;
}
int ksize_radius(int ksize) { return ksize <= 5 ? 2 : (ksize <= 7 ? 3 : (ksize <= 9 ? 4 : 5)); }

struct Level {
  int w, h, pitch, octave, sublevel;
  float esigma, etime;
  int nsteps;
  float tau[AKZ_MAX_TAU];
  int deriv_step;
  float border, det_size;
  float *Lt, *Lsm, *det;
  float2* Lxy;  // Lx and Ly interleaved, element (y, x) at y * pitch + x
};

enum { EV_START, EV_CONTRAST, EV_SCALE, EV_DERIV, EV_EXTREMA, EV_NMS, EV_DESC, EV_COUNT };

}  // namespace

struct akz_ctx {
  akz_options opts;
  int device = 0;
  hipStream_t stream = nullptr;
  int nlev = 0;
  Level lev[AKZ_MAX_LEVELS];
  int octs[AKZ_MAX_LEVELS];
  void* slab = nullptr;
  size_t slab_bytes = 0;
  float* img_in = nullptr;  // input staging, pitch lev[0].pitch
  float* g = nullptr;       // conductance scratch
  float* fed_tmp = nullptr; // FED ping-pong scratch
  uint32_t* maxbits = nullptr;
  int32_t* hist = nullptr;
  uint32_t* blockmax = nullptr;
  float* kc = nullptr;
  akz_det_bufs db{};  // detector arrays of image 0
  akz_kp* cand = nullptr;
  int cap = 0;
  int32_t* level_start = nullptr;
  int32_t* ncand = nullptr;
  akz_kp* kps = nullptr;
  int32_t* status = nullptr;  // AKZ_ST_* block of image 0
  uint8_t* desc = nullptr;
  // batch: `batch` images of identical layout, image z at slab + z * zs
  int batch = 1;
  size_t zs = 0;
  int last_nz = 1;            // images in the last processed batch
  int status_synced = 0;      // images whose status blocks are in h_small and current
  bool keep_det = false;      // throughput path: write full determinant images
  bool det_full_last = true;  // the last detection wrote full determinant images
  int32_t* h_small = nullptr;  // pinned copy of the status blocks (4 ints per image)
  uint8_t* h_stage = nullptr;  // pinned staging for keypoints + 64-byte descriptor rows
  size_t h_stage_bytes = 0;
  akz_det_params det{};
  akz_det_params det32{};  // same, with the 64 x akz_detect_tile_rows() tiling of the fused detector
  akz_desc_params dp{};
  akz_taps taps0{}, taps1{}, taps_c{};
  int R0 = 4;
  hipEvent_t ev[EV_COUNT] = {};
  // events of the last eager run (bit e: ev[e] recorded); graph captures and
  // replays record none, so their stage times are not reported as current
  unsigned ev_mask = 0;
  // M-LDB variant: comparison list on the device, bits and bytes per descriptor
  int16_t* d_cmp = nullptr;
  int desc_bits = 486, desc_bytes = 61;
  // bumped whenever captured graphs of this context go stale (proposal,
  // FED variants, determinant setting): graphs built over several contexts
  // compare it
  unsigned config_gen = 0;
  int state = 0;  // 0 created, 1 scale space, 2 detected, 3 described
  bool graph_mode = false;
  bool capturing = false;
  // one captured graph per batch size (index n - 1), instantiated on first use
  std::vector<hipGraphExec_t> graph_exec;
  std::vector<hipGraph_t> graph;
  // optional per-level events around the FED launches (akz_profile_fed)
  // learned tile proposal
  akz_prop_src psrc{};
  akz_prop_act pact{};  // inference arrays of `batch` images, allocated on first use
  void* prop_mem = nullptr;
  akz_prop_net* prop_net = nullptr;
  float prop_thr = 0.5f;
  bool fed_profile = false;
  int fed_block = 1;  // FED steps per launch
  int fed_variant[AKZ_MAX_LEVELS] = {};  // blocked FED kernel variant per level (akz_autotune)
  hipEvent_t fed_ev[2 * AKZ_MAX_LEVELS] = {};
};

struct akz_ragged {
  int ncls = 0;
  akz_ctx* cls[AKZ_MAX_CLASSES] = {};
  std::vector<int> img_cls, img_slot;  // class and batch slot of each image of the last call
  // work-list mode: the detector, orientation and MLDB run as one persistent
  // launch each over every image of every class (ragged_enqueue_lists)
  bool lists = false;
  bool graph = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_mid = nullptr;
  hipEvent_t ev_join[AKZ_MAX_CLASSES] = {}, ev_join2[AKZ_MAX_CLASSES] = {};
  akz_det_params* d_detp = nullptr;    // the classes' fused-detector parameters
  akz_desc_params* d_descp = nullptr;  // and descriptor parameters, in device memory
  std::vector<akz_det_params> detp_host;
  bool descp_ready = false;
  unsigned gen = ~0u;  // sum of the class contexts' config_gen the graphs were built with
  // one captured graph per class composition (images per class)
  std::map<std::vector<int>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
};

struct akz_matcher {
  hipStream_t stream = nullptr;
  uint8_t* q = nullptr;
  uint8_t* t = nullptr;
  akz_dmatch* out = nullptr;
  uint32_t* partial = nullptr;
  size_t qcap = 0, tcap = 0, pcap = 0, ocap = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

akz_status compute_geometry(akz_ctx* c) {
  const akz_options& o = c->opts;
  // the derivative stencils reflect at offset s <= 4 and need levels wider than that
  if (o.img_width < 16 || o.img_height < 16 || o.omax < 1 || o.nsublevels < 1) return AKZ_ERR_ARG;
  if (o.img_width >= 32768 || o.img_height >= 32768) return AKZ_ERR_ARG;
  if (o.diffusivity < 0 || o.diffusivity > 3 || o.fed_steps_per_launch < 0) return AKZ_ERR_ARG;
  // M-LDB (5) and M-LDB upright (4); the reference's GPU path has only the
  // former (AKAZE.cpp:428-443).  SURF and M-SURF (0-3) are not provided.
  if (o.descriptor != 5 && o.descriptor != 4) return AKZ_ERR_UNSUPPORTED;
  if (o.descriptor_channels < 1 || o.descriptor_channels > 3) return AKZ_ERR_ARG;
  // the subsample picks from the 162 comparisons per channel (AKAZE.cpp:506-515)
  if (o.descriptor_size < 0 || o.descriptor_size > 162 * o.descriptor_channels) return AKZ_ERR_ARG;
  // the MLDB grid (max(3*ceil(2p/3), 4*ceil(p/2)))^2 must fit 7 x 64 samples per wavefront
  if (o.descriptor_pattern_size < 1 || o.descriptor_pattern_size > 10) return AKZ_ERR_UNSUPPORTED;
  // scales and thresholds: finite and positive (the level sigmas, FED times and
  // derivative steps are built from them)
  if (!(std::isfinite(o.soffset) && o.soffset > 0.0f)) return AKZ_ERR_ARG;
  if (!(std::isfinite(o.derivative_factor) && o.derivative_factor > 0.0f)) return AKZ_ERR_ARG;
  if (!std::isfinite(o.dthreshold) || !std::isfinite(o.min_dthreshold)) return AKZ_ERR_ARG;
  // ContrastPercentile (cuda_akaze.cu:619-641) caps the bins at 512 and takes a
  // fraction of the interior pixels; fewer than one bin or a fraction outside
  // (0, 1] has no percentile
  if (o.kcontrast_nbins < 1 || o.kcontrast_nbins > 512) return AKZ_ERR_ARG;
  if (!(std::isfinite(o.kcontrast_percentile) && o.kcontrast_percentile > 0.0f && o.kcontrast_percentile <= 1.0f))
    return AKZ_ERR_ARG;
  int nlev = 0;
  // level geometry and scales, AKAZE.cpp:151-184
  for (int i = 0; i <= o.omax - 1; i++) {
    float rfactor = (float)(1.0 / std::pow(2.0, (double)i));
    int lh = (int)(o.img_height * rfactor);
    int lw = (int)(o.img_width * rfactor);
    if ((lw < 80 || lh < 40) && i != 0) break;
    for (int j = 0; j < o.nsublevels; j++) {
      if (nlev >= AKZ_MAX_LEVELS) return AKZ_ERR_ARG;
      Level& L = c->lev[nlev];
      L.w = lw;
      L.h = lh;
      L.pitch = (int)align_up((size_t)lw, 64);
      L.octave = i;
      L.sublevel = j;
      L.esigma = o.soffset * powf(2.0f, (float)(j) / (float)(o.nsublevels) + i);
      L.etime = 0.5 * (L.esigma * L.esigma);
      L.nsteps = 0;
      nlev++;
    }
  }
  c->nlev = nlev;
  // FED steps between consecutive levels, AKAZE.cpp:187-197
  for (int i = 1; i < nlev; i++) {
    float ttime = c->lev[i].etime - c->lev[i - 1].etime;
    int n = fed_tau(ttime, 0.25f, c->lev[i].tau, AKZ_MAX_TAU);
    if (n < 0) return AKZ_ERR_ARG;
    c->lev[i].nsteps = n;
  }
  // detector parameters, AKAZE.cpp:322-364
  float smax = 10.0 * sqrtf(2.0f);
  for (int i = 0; i < nlev; i++) {
    Level& L = c->lev[i];
    float ratio = powf(2.0f, (float)L.octave);
    L.deriv_step = fround(L.esigma * o.derivative_factor / ratio);
    L.det_size = L.esigma * o.derivative_factor;
    L.border = smax * fround(L.det_size / ratio);
    c->octs[i] = L.octave;
    // the reflect-101 indices of Derivate (cuda_akaze.cu:651-671) stay inside a
    // level only for steps below its width and height
    if (L.deriv_step < 0 || L.deriv_step >= std::min(L.w, L.h)) return AKZ_ERR_ARG;
  }
  return AKZ_OK;
}

struct Carver {
  size_t off = 0;
  template <class T>
  size_t take(size_t count) {
    size_t o = off;
    off = align_up(off + sizeof(T) * count, ALIGN);
    return o;
  }
};

akz_status allocate(akz_ctx* c) {
  Carver cv;
  // per level: Lt, Lsmooth, the interleaved first derivatives, det
  size_t o_lev[AKZ_MAX_LEVELS][4];
  for (int i = 0; i < c->nlev; i++) {
    size_t n = (size_t)c->lev[i].pitch * c->lev[i].h;
    for (int k = 0; k < 4; k++) o_lev[i][k] = cv.take<float>(k == 2 ? 2 * n : n);
  }
  const Level& L0 = c->lev[0];
  size_t n0 = (size_t)L0.pitch * (L0.h + 1);
  size_t o_img = cv.take<float>(n0), o_g = cv.take<float>(n0), o_tmp = cv.take<float>(n0);
  size_t o_max = cv.take<uint32_t>(1), o_hist = cv.take<int32_t>(512);
  size_t o_bmax = cv.take<uint32_t>(akz_contrast_front_blocks(L0.w, L0.h));
  size_t o_kc = cv.take<float>(AKZ_MAX_LEVELS);
  // row segments of 64 pixels, level-major, raster order
  int seg_total = 0, tiles = 0;
  for (int i = 0; i < c->nlev; i++) {
    Level& L = c->lev[i];
    akz_level_desc& D = c->det.lev[i];
    D.w = L.w; D.h = L.h; D.pitch = L.pitch; D.octave = L.octave; D.sublevel = L.sublevel;
    D.deriv_step = L.deriv_step; D.border = L.border; D.det_size = L.det_size;
    akz_border_interval(L.border, L.w, &D.xlo, &D.xhi);
    akz_border_interval(L.border, L.h, &D.ylo, &D.yhi);
    D.nseg_x = (L.w + 63) / 64;
    D.seg_base = seg_total;
    seg_total += D.nseg_x * L.h;
    c->det.tile_start[i] = tiles;
    tiles += D.nseg_x * ((L.h + 15) / 16);
  }
  c->det.tile_start[c->nlev] = tiles;
  c->det.seg_total = seg_total;
  c->det.nlev = c->nlev;
  const int dr = akz_detect_tile_rows();
  for (int i = 0, t32 = 0; i <= c->nlev; i++) {
    c->det32.tile_start[i] = t32;
    if (i < c->nlev) t32 += c->det.lev[i].nseg_x * ((c->lev[i].h + dr - 1) / dr);
  }
  const int nblk_seg = akz_scan_blocks(seg_total);
  size_t o_segc = cv.take<int32_t>(seg_total + 4);
  size_t o_slots = cv.take<uint8_t>((size_t)seg_total * 32);
  size_t o_scan = cv.take<int32_t>((size_t)nblk_seg * 16 + 4);
  // candidate capacity: every strict 3x3 maximum (at most one per 2x2 block)
  size_t capz = 0;
  for (int i = 0; i < c->nlev; i++) capz += (size_t)((c->lev[i].w + 1) / 2) * ((c->lev[i].h + 1) / 2);
  if (c->opts.maxkeypoints > 0) capz = std::min(capz, (size_t)c->opts.maxkeypoints);
  if (capz > (size_t)(1u << 22) - 1) capz = (1u << 22) - 1;
  c->cap = (int)std::max<size_t>(capz, 1);
  const size_t nb_cap = (size_t)c->cap * AKZ_NB_PER_CAND;
  size_t o_cand = cv.take<akz_kp>(c->cap), o_ls = cv.take<int32_t>(AKZ_MAX_LEVELS + 1);
  size_t o_cxy = cv.take<float2>(c->cap);
  size_t o_nc = cv.take<int32_t>(4), o_status = cv.take<int32_t>(4);
  size_t o_nbc = cv.take<int32_t>(c->cap + 4), o_nbo = cv.take<int32_t>(c->cap + 4);
  size_t o_nbs = cv.take<int32_t>(4);  // 64-bit pool counter
  size_t o_nbl = cv.take<int32_t>(nb_cap);
  size_t o_nbp = cv.take<uint2>(c->cap);
  size_t o_mem = cv.take<int32_t>(c->cap), o_minnb = cv.take<int32_t>(c->cap), o_code = cv.take<int32_t>(c->cap);
  size_t o_act = cv.take<int32_t>((size_t)c->cap * 2);
  size_t o_arb = cv.take<unsigned long long>(std::min(c->cap, 32768));
  size_t o_kps = cv.take<akz_kp>(c->cap), o_kidx = cv.take<int32_t>(c->cap);
  size_t o_desc = cv.take<uint8_t>((size_t)c->cap * AKZ_DESC_STRIDE);
  c->zs = align_up(cv.off, ALIGN);
  // after the batch's sub-slabs: the contiguous status blocks of all images
  c->slab_bytes = c->zs * (size_t)c->batch + align_up(16 * (size_t)c->batch, ALIGN);
  AKZ_HIP_TRY(hipMalloc(&c->slab, c->slab_bytes));
  char* b = (char*)c->slab;
  for (int i = 0; i < c->nlev; i++) {
    Level& L = c->lev[i];
    L.Lt = (float*)(b + o_lev[i][0]);
    L.Lsm = (float*)(b + o_lev[i][1]);
    L.Lxy = (float2*)(b + o_lev[i][2]);
    L.det = (float*)(b + o_lev[i][3]);
  }
  c->img_in = (float*)(b + o_img);
  c->g = (float*)(b + o_g);
  c->fed_tmp = (float*)(b + o_tmp);
  c->maxbits = (uint32_t*)(b + o_max);
  c->hist = (int32_t*)(b + o_hist);
  c->blockmax = (uint32_t*)(b + o_bmax);
  c->kc = (float*)(b + o_kc);
  c->cand = (akz_kp*)(b + o_cand);
  c->level_start = (int32_t*)(b + o_ls);
  c->ncand = (int32_t*)(b + o_nc);
  c->status = (int32_t*)(b + o_status);
  c->kps = (akz_kp*)(b + o_kps);
  c->desc = (uint8_t*)(b + o_desc);
  akz_det_bufs& D = c->db;
  D.seg_count = (int32_t*)(b + o_segc);
  D.seg_slots = (uint8_t*)(b + o_slots);
  D.seg_wsum = (int32_t*)(b + o_scan);
  D.level_start = c->level_start;
  D.ncand = c->ncand;
  D.cand = c->cand;
  D.cand_xy = (float2*)(b + o_cxy);
  D.nb_count = (int32_t*)(b + o_nbc);
  D.nb_off = (int32_t*)(b + o_nbo);
  D.nb_pool = (int32_t*)(b + o_nbs);
  D.nb_list = (int32_t*)(b + o_nbl);
  D.nb_pack = (uint2*)(b + o_nbp);
  D.mem = (int32_t*)(b + o_mem);
  D.minnb = (int32_t*)(b + o_minnb);
  D.code = (int32_t*)(b + o_code);
  D.act = (int32_t*)(b + o_act);
  D.arb = (unsigned long long*)(b + o_arb);
  D.kps = c->kps;
  D.kidx = (int32_t*)(b + o_kidx);
  D.status = c->status;
  D.status_all = (int32_t*)(b + c->zs * (size_t)c->batch);
  D.cap = c->cap;
  D.nb_cap = (int)nb_cap;
  D.max_waves = c->opts.nms_max_waves;
  D.nms_pool = akz_suppression_pool_max();
  D.zs = c->zs;
  AKZ_HIP_TRY(hipMemsetAsync(c->slab, 0, c->slab_bytes, c->stream));
  AKZ_HIP_TRY(hipHostMalloc((void**)&c->h_small, 16 * (size_t)c->batch, hipHostMallocDefault));
  // kernel parameter blocks
  for (int i = 0; i < c->nlev; i++) {
    Level& L = c->lev[i];
    c->det.Lsm[i] = (i == 0) ? L.Lt : L.Lsm;  // level 0: Lsmooth is Lt (AKAZE.cpp:236)
    c->det.Lxy[i] = L.Lxy;
    c->det.det[i] = L.det;
    c->dp.Lt[i] = L.Lt;
    c->dp.Lxy[i] = L.Lxy;
    c->dp.w[i] = L.w;
    c->dp.h[i] = L.h;
    c->dp.pitch[i] = L.pitch;
  }
  // proposal source view: level geometry, determinant and Lt arrays
  {
    int w[AKZ_MAX_LEVELS], h[AKZ_MAX_LEVELS];
    for (int i = 0; i < c->nlev; i++) { w[i] = c->lev[i].w; h[i] = c->lev[i].h; }
    c->psrc.g = akz_prop_make_geom(c->nlev, w, h, c->octs);
    int ts = 0;
    for (int i = 0; i < c->nlev; i++) {
      c->psrc.det[i] = c->lev[i].det;
      c->psrc.Lt[i] = c->lev[i].Lt;
      c->psrc.w[i] = w[i];
      c->psrc.h[i] = h[i];
      c->psrc.pitch[i] = c->lev[i].pitch;
      c->psrc.tile_start[i] = ts;
      ts += ((w[i] + 63) / 64) * c->psrc.g.th[i];
    }
    c->psrc.tile_start[c->nlev] = ts;
    c->psrc.thr = std::max(c->opts.dthreshold, c->opts.min_dthreshold);
    c->psrc.zs = c->zs;
  }
  c->dp.patsize = c->opts.descriptor_pattern_size;
  c->dp.zs = c->zs;
  c->det.zs = c->zs;
  c->det.thr = std::max(c->opts.dthreshold, c->opts.min_dthreshold);
  float wgt = 10.0 / 3.0;  // HessianDeterminant, cuda_akaze.cu:706-708
  c->det.fac1 = 1.0 / (2.0 * (wgt + 2.0));
  c->det.fac2 = wgt * c->det.fac1;
  int t32[AKZ_MAX_LEVELS + 1];
  for (int i = 0; i <= c->nlev; i++) t32[i] = c->det32.tile_start[i];
  c->det32 = c->det;
  for (int i = 0; i <= c->nlev; i++) c->det32.tile_start[i] = t32[i];
  return AKZ_OK;
}

// FED steps per launch at level i: at most the variant's (its halo), capped
// by the option; a level needing several launches gets them balanced (21
// steps as 7 + 7 + 7, not 8 + 8 + 5): the same launches with smaller halos.
int level_block(const akz_ctx* c, int i) {
  if (c->fed_block == 1) return 1;
  const int cap = std::min(c->fed_block, akz_fed_variant_steps(c->fed_variant[i]));
  const int n = c->lev[i].nsteps;
  if (n <= cap) return cap;
  const int nl = (n + cap - 1) / cap;
  return (n + nl - 1) / nl;
}

// Whether level i's first FED launch forms the level's input, an exact
// HalfSample2 of the previous level, while loading it (fused launches of the
// LDS variants at an octave's first level).
bool level_half_in_fed(const akz_ctx* c, int i, int nz) {
  const Level& L = c->lev[i];
  const Level& P = c->lev[i - 1];
  const int bs = level_block(c, i), n = L.nsteps;
  const bool fuse = bs > 1 && n > 0 && akz_fed_variant_fusable(c->fed_variant[i], std::min(bs, n), nz);
  return fuse && L.octave > P.octave && P.w == 2 * L.w && P.h == 2 * L.h && akz_fed_variant_halves(c->fed_variant[i]);
}

// Create_Nonlinear_Scale_Space (AKAZE.cpp:211-276) on the context stream.
inline void record(akz_ctx* c, int e) {
  if (e == EV_START) c->ev_mask = 0;
  if (c->capturing) return;
  hipEventRecord(c->ev[e], c->stream);
  c->ev_mask |= 1u << e;
}

// Images 0..nz-1 of the batch: input z at img + z * in_zs bytes.
void enqueue_scale_space(akz_ctx* c, const float* img, int in_pitch, size_t in_zs, int nz) {
  hipStream_t s = c->stream;
  const akz_batch bt{nz, c->zs};
  Level& L0 = c->lev[0];
  record(c, EV_START);
  // level 0, the Gaussian of soffset (Lsmooth aliases Lt, AKAZE.cpp:234-236),
  // and the contrast factor of the input (ContrastPercentile,
  // cuda_akaze.cu:604-649); level-0 Ly holds the gradient until detection
  akz_launch_contrast(s, img, in_pitch, in_zs, L0.Lt, (float*)L0.Lxy, L0.pitch, L0.w, L0.h, c->R0, c->taps0, c->taps_c,
                      std::min(c->opts.kcontrast_nbins, 512), c->opts.kcontrast_percentile, c->maxbits,
                      c->blockmax, c->hist, c->kc, c->nlev, c->octs, bt);
  record(c, EV_CONTRAST);
  for (int i = 1; i < c->nlev; i++) {
    Level& L = c->lev[i];
    Level& P = c->lev[i - 1];
    const int n = L.nsteps;
    // FED launches alternate between Lt and the scratch so that the last one
    // lands in Lt; each launch runs up to `bs` steps (bs = 1: one per launch)
    const int bs = level_block(c, i);
    const int nl = (n + bs - 1) / bs;
    float* first_dst = (nl % 2 == 1) ? L.Lt : c->fed_tmp;
    // the level's first FED launch builds Lsmooth and g where that pays
    // (akz_fed_variant_fusable); g goes through HBM only when further
    // launches of the level read it
    const bool fuse = bs > 1 && n > 0 && akz_fed_variant_fusable(c->fed_variant[i], std::min(bs, n), nz);
    // an exact halving at an octave's first level is formed inside that
    // fused launch's load where the variant can (the LDS ones)
    const bool half_in_fed = level_half_in_fed(c, i, nz);
    const float* base;
    if (half_in_fed) {
      base = nullptr;
    } else if (L.octave > P.octave) {
      float* hs = (nl % 2 == 1) ? c->fed_tmp : L.Lt;  // must differ from first_dst
      if (n == 0) hs = L.Lt;
      akz_launch_halfsample(s, P.Lt, P.w, P.h, P.pitch, hs, L.w, L.h, L.pitch, bt);
      base = hs;
    } else {
      base = P.Lt;  // Copy (AKAZE.cpp:258) is folded into the first step's read
      if (n == 0) akz_launch_copy2d(s, P.Lt, P.pitch, L.Lt, L.pitch, L.w, L.h, bt);
    }
    if (!fuse)
      akz_launch_smooth_flow(s, base, L.pitch, L.Lsm, c->g, L.w, L.h, c->opts.diffusivity, c->kc + i, c->taps1, bt);
    const float* cur = base;
    float* dst = first_dst;
    if (c->fed_profile) hipEventRecord(c->fed_ev[2 * i], s);
    for (int j = 0; j < n; j += bs) {
      int k = std::min(bs, n - j);
      float ht[AKZ_MAX_TAU];
      for (int q = 0; q < k; q++) ht[q] = 0.5 * L.tau[j + q];
      if (bs == 1)
        akz_launch_fed_step(s, cur, c->g, dst, L.pitch, L.w, L.h, ht[0], bt);
      else if (fuse && j == 0)
        akz_launch_fed_variant_sf(s, c->fed_variant[i], cur, half_in_fed ? P.Lt : nullptr, P.pitch, L.Lsm,
                                  nl > 1 ? c->g : nullptr, dst, L.pitch, L.w, L.h, c->opts.diffusivity, c->kc + i,
                                  c->taps1, ht, k, bt);
      else
        akz_launch_fed_variant(s, c->fed_variant[i], cur, c->g, dst, L.pitch, L.w, L.h, ht, k, bt);
      cur = dst;
      dst = (dst == L.Lt) ? c->fed_tmp : L.Lt;
    }
    if (c->fed_profile) hipEventRecord(c->fed_ev[2 * i + 1], s);
  }
  record(c, EV_SCALE);
}

// Detection in three parts, so that a ragged batch can run the middle one as
// a single persistent launch over all its size classes: the parameters and
// the proposal keys (returns whether the fused detector applies), the
// detector kernels, and everything after them.
bool detect_params(akz_ctx* c, bool full_det) {
  int smax = 0;
  for (int i = 0; i < c->nlev; i++) smax = std::max(smax, c->lev[i].deriv_step);
  const bool fused = smax <= akz_detect_max_step();
  const bool prop_keys = c->prop_net != nullptr && fused;
  full_det = full_det || (c->prop_net != nullptr && !fused);
  c->det32.det_full = full_det ? 1 : 0;
  c->det32.tile_key = prop_keys ? c->pact.tkey : nullptr;
  if (prop_keys) {
    const akz_prop_geom& g = c->psrc.g;
    for (int i = 0; i < c->nlev; i++) {
      c->det32.tile_base[i] = g.base[i];
      c->det32.tile_w[i] = g.tw[i];
      c->det32.tile_h[i] = g.th[i];
    }
    c->det32.tiles_per_image = g.P;
  }
  c->det_full_last = full_det || !fused;  // the unfused kernels write all of it
  return fused;
}

bool detect_prepare(akz_ctx* c, int nz, bool full_det) {
  c->status_synced = 0;
  const bool fused = detect_params(c, full_det);
  if (c->det32.tile_key) akz_prop_launch_clear_keys(c->stream, c->psrc.g, c->pact, nz);
  return fused;
}

void detect_kernels(akz_ctx* c, int nz, bool fused) {
  hipStream_t s = c->stream;
  if (fused) {
    akz_launch_detect_fused(s, c->det32, c->db, nz);
    record(c, EV_DERIV);
  } else {
    akz_launch_hessian(s, c->det, nz);
    record(c, EV_DERIV);
    akz_launch_extrema(s, c->det, c->db, nz);
  }
}

void detect_finish(akz_ctx* c, int nz, bool fused) {
  hipStream_t s = c->stream;
  if (c->prop_net) {
    akz_prop_launch_features(s, c->psrc, c->pact, nz, c->prop_net != nullptr && fused);
    akz_prop_launch_infer(s, akz_prop_net_weights(c->prop_net), c->psrc.g, c->pact, nz, c->prop_thr);
    akz_prop_launch_mask_segments(s, c->det, c->db, c->psrc.g, c->pact.mask, nz);
  }
  akz_launch_candidates(s, c->det, c->db, nz);
  record(c, EV_EXTREMA);
  akz_launch_suppression(s, c->db, nz);
  record(c, EV_NMS);
}

// Feature_Detection (AKAZE.cpp:316-378) + FilterExtrema (cuda_akaze.cu:1318-1402).
// full_det: write every determinant pixel (parity readback, tile features);
// otherwise only the candidates' 3x3 neighbourhoods the subpixel fit reads.
// With a learned proposal the fused detector keeps the tile maxima of the
// determinant itself; the unfused kernels write whole determinant images,
// which the tile statistics then read.
void enqueue_detect(akz_ctx* c, int nz, bool full_det) {
  const bool fused = detect_prepare(c, nz, full_det);
  detect_kernels(c, nz, fused);
  detect_finish(c, nz, fused);
}

// Compute_Descriptors (AKAZE.cpp:405-449): orientation, then MLDB.
void enqueue_describe(akz_ctx* c, int nz) {
  hipStream_t s = c->stream;
  // M-LDB upright keeps the detector's angle 0 (no orientation)
  if (c->opts.descriptor == 5) akz_launch_orientation(s, c->dp, c->kps, c->status + AKZ_ST_KEPT, c->cap, nz);
  akz_launch_mldb(s, c->dp, c->kps, c->status + AKZ_ST_KEPT, c->desc, c->cap, nz);
  record(c, EV_DESC);
}

float ev_ms(akz_ctx* c, int a, int b) {
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) {
    (void)hipGetLastError();
    return 0.0f;
  }
  return ms;
}

// Waits for the stream and reads the status blocks of images 0..nz-1 (one
// copy): kept count, overflow flags, total candidates, waves.
akz_status sync_status(akz_ctx* c, int nz) {
  AKZ_HIP_TRY(hipMemcpyAsync(c->h_small, c->db.status_all, 16 * (size_t)nz, hipMemcpyDeviceToHost, c->stream));
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  c->status_synced = nz;
  return AKZ_OK;
}

// Result status of image z after sync_status.
akz_status image_status(const akz_ctx* c, int z, int* nk) {
  const int32_t* st = c->h_small + 4 * z;
  if (nk) *nk = st[AKZ_ST_KEPT];
  if (st[AKZ_ST_FLAGS] != 0 || st[AKZ_ST_TOTAL] > c->cap) return AKZ_ERR_OVERFLOW;
  return AKZ_OK;
}

// Inference arrays of the proposal for the context's batch.
akz_status ensure_prop_act(akz_ctx* c) {
  if (c->prop_mem) return AKZ_OK;
  const size_t P = (size_t)c->psrc.g.P * c->batch;
  Carver cv;
  size_t o_x0 = cv.take<uint16_t>(P * 4), o_h1 = cv.take<uint16_t>(P * 16), o_h2 = cv.take<uint16_t>(P * 32);
  size_t o_h3 = cv.take<uint16_t>(P * 32), o_tm = cv.take<float>(P), o_ts = cv.take<float>(P);
  size_t o_pr = cv.take<float>(P), o_mk = cv.take<uint8_t>(P), o_tk = cv.take<uint32_t>(P);
  AKZ_HIP_TRY(hipMalloc(&c->prop_mem, cv.off));
  char* b = (char*)c->prop_mem;
  c->pact.x0 = (uint16_t*)(b + o_x0);
  c->pact.h1 = (uint16_t*)(b + o_h1);
  c->pact.h2 = (uint16_t*)(b + o_h2);
  c->pact.h3 = (uint16_t*)(b + o_h3);
  c->pact.tmax = (float*)(b + o_tm);
  c->pact.tmean = (float*)(b + o_ts);
  c->pact.prob = (float*)(b + o_pr);
  c->pact.mask = (uint8_t*)(b + o_mk);
  c->pact.tkey = (uint32_t*)(b + o_tk);
  return AKZ_OK;
}

void drop_graphs(akz_ctx* c) {
  c->config_gen++;  // graphs that include this context's kernels are stale
  for (size_t i = 0; i < c->graph_exec.size(); i++) {
    if (c->graph_exec[i]) (void)hipGraphExecDestroy(c->graph_exec[i]);
    if (c->graph[i]) (void)hipGraphDestroy(c->graph[i]);
    c->graph_exec[i] = nullptr;
    c->graph[i] = nullptr;
  }
}

template <class T>
T* zptr(const akz_ctx* c, T* p, int z) {
  return reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(p) + (size_t)z * c->zs);
}

}  // namespace

extern "C" {

const char* akz_status_string(akz_status s) {
  switch (s) {
    case AKZ_OK: return "ok";
    case AKZ_ERR_ARG: return "invalid argument";
    case AKZ_ERR_HIP: return "HIP runtime error";
    case AKZ_ERR_OOM: return "out of memory";
    case AKZ_ERR_OVERFLOW: return "capacity exceeded, output truncated";
    case AKZ_ERR_STATE: return "call order violated";
    case AKZ_ERR_UNSUPPORTED: return "option not implemented";
    case AKZ_ERR_NO_DEVICE: return "no HIP device";
  }
  return "unknown status";
}

akz_status akz_options_default(akz_options* o) {
  if (!o) return AKZ_ERR_ARG;
  memset(o, 0, sizeof(*o));
  o->omax = 4;
  o->nsublevels = 4;
  o->soffset = 1.6f;
  o->derivative_factor = 1.5f;
  o->sderivatives = 1.0f;
  o->diffusivity = 1;
  o->dthreshold = 0.001f;
  o->min_dthreshold = 0.00001f;
  o->descriptor = 5;
  o->descriptor_size = 0;
  o->descriptor_channels = 3;
  o->descriptor_pattern_size = 10;
  o->kcontrast = 0.001f;
  o->kcontrast_percentile = 0.7f;
  o->kcontrast_nbins = 300;
  o->maxkeypoints = 0;
  o->nms_max_waves = 10000;
  return AKZ_OK;
}

akz_status akz_device_count(int* n) {
  if (!n) return AKZ_ERR_ARG;
  *n = 0;
  hipError_t e = hipGetDeviceCount(n);
  if (e != hipSuccess) { (void)hipGetLastError(); *n = 0; return AKZ_ERR_NO_DEVICE; }
  return AKZ_OK;
}

akz_status akz_create_batch(const akz_options* opts, int batch, akz_ctx** out) {
  if (!opts || !out || batch < 1 || batch > AKZ_MAX_BATCH) return AKZ_ERR_ARG;
  *out = nullptr;
  akz_ctx* c = new (std::nothrow) akz_ctx();
  if (!c) return AKZ_ERR_OOM;
  c->opts = *opts;
  c->batch = batch;
  akz_status st = compute_geometry(c);
  if (st != AKZ_OK) { delete c; return st; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    delete c;
    return AKZ_ERR_NO_DEVICE;
  }
  if (hipGetDevice(&c->device) != hipSuccess) { delete c; return AKZ_ERR_HIP; }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return AKZ_ERR_HIP; }
  for (int i = 0; i < EV_COUNT; i++) hipEventCreate(&c->ev[i]);
  int ksize = 2 * std::ceil((c->opts.soffset - 0.8) / 0.3) + 3;  // AKAZE.cpp:234-235
  c->R0 = ksize_radius(ksize);
  c->fed_block = c->opts.fed_steps_per_launch > 0 ? std::min(c->opts.fed_steps_per_launch, akz_fed_block_steps())
                                                 : akz_fed_block_steps();
  gauss_taps(c->opts.soffset * c->opts.soffset, c->R0, &c->taps0);
  gauss_taps(1.0, 2, &c->taps1);
  gauss_taps(1.0f, 2, &c->taps_c);
  akz_describe_init_tables();
  for (int i = 0; i < c->nlev; i++) c->fed_variant[i] = akz_fed_variant_for(c->lev[i].w, c->lev[i].h, batch);
  c->graph_exec.assign(batch, nullptr);
  c->graph.assign(batch, nullptr);
  st = allocate(c);
  if (st != AKZ_OK) { akz_destroy(c); return st; }
  {
    int16_t tab[2 * AKZ_MLDB_TABLE];
    int nb = akz_mldb_table(c->opts.descriptor_size, c->opts.descriptor_pattern_size, c->opts.descriptor_channels,
                            tab, tab + AKZ_MLDB_TABLE);
    if (nb < 1) { akz_destroy(c); return AKZ_ERR_ARG; }
    c->desc_bits = nb;
    c->desc_bytes = (nb + 7) / 8;
    if (hipMalloc((void**)&c->d_cmp, sizeof tab) != hipSuccess) { (void)hipGetLastError(); akz_destroy(c); return AKZ_ERR_OOM; }
    if (hipMemcpy(c->d_cmp, tab, sizeof tab, hipMemcpyHostToDevice) != hipSuccess) { akz_destroy(c); return AKZ_ERR_HIP; }
    c->dp.cmp = c->d_cmp;
    c->dp.nch = c->opts.descriptor_channels;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) { akz_destroy(c); return AKZ_ERR_HIP; }
  *out = c;
  return AKZ_OK;
}

akz_status akz_create(const akz_options* opts, akz_ctx** out) { return akz_create_batch(opts, 1, out); }

akz_status akz_destroy(akz_ctx* c) {
  if (!c) return AKZ_ERR_ARG;
  if (c->stream) hipStreamSynchronize(c->stream);
  for (size_t i = 0; i < c->graph_exec.size(); i++)
    if (c->graph_exec[i]) hipGraphExecDestroy(c->graph_exec[i]);
  for (size_t i = 0; i < c->graph.size(); i++)
    if (c->graph[i]) hipGraphDestroy(c->graph[i]);
  for (int i = 0; i < EV_COUNT; i++)
    if (c->ev[i]) hipEventDestroy(c->ev[i]);
  for (int i = 0; i < 2 * AKZ_MAX_LEVELS; i++)
    if (c->fed_ev[i]) hipEventDestroy(c->fed_ev[i]);
  if (c->slab) hipFree(c->slab);
  if (c->prop_mem) hipFree(c->prop_mem);
  if (c->d_cmp) hipFree(c->d_cmp);
  if (c->h_small) hipHostFree(c->h_small);
  if (c->h_stage) hipHostFree(c->h_stage);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return AKZ_OK;
}

akz_status akz_batch_capacity(const akz_ctx* c, int* batch) {
  if (!c || !batch) return AKZ_ERR_ARG;
  *batch = c->batch;
  return AKZ_OK;
}

akz_status akz_descriptor_bytes(const akz_ctx* c, int* nbytes, int* nbits) {
  if (!c || !nbytes) return AKZ_ERR_ARG;
  *nbytes = c->desc_bytes;
  if (nbits) *nbits = c->desc_bits;
  return AKZ_OK;
}

akz_status akz_num_levels(const akz_ctx* c, int* n) {
  if (!c || !n) return AKZ_ERR_ARG;
  *n = c->nlev;
  return AKZ_OK;
}

akz_status akz_level_size(const akz_ctx* c, int lvl, int* w, int* h, int* octave, int* sublevel, float* esigma) {
  if (!c || lvl < 0 || lvl >= c->nlev) return AKZ_ERR_ARG;
  const Level& L = c->lev[lvl];
  if (w) *w = L.w;
  if (h) *h = L.h;
  if (octave) *octave = L.octave;
  if (sublevel) *sublevel = L.sublevel;
  if (esigma) *esigma = L.esigma;
  return AKZ_OK;
}

akz_status akz_build_scale_space(akz_ctx* c, const float* img, size_t pitch_bytes) {
  if (!c || !img) return AKZ_ERR_ARG;
  const Level& L0 = c->lev[0];
  size_t sp = pitch_bytes ? pitch_bytes : sizeof(float) * L0.w;
  if (sp < sizeof(float) * L0.w) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipMemcpy2DAsync(c->img_in, sizeof(float) * L0.pitch, img, sp, sizeof(float) * L0.w, L0.h,
                               hipMemcpyHostToDevice, c->stream));
  enqueue_scale_space(c, c->img_in, L0.pitch, 0, 1);
  AKZ_HIP_TRY(hipGetLastError());
  c->state = 1;
  c->last_nz = 1;
  return AKZ_OK;
}

akz_status akz_build_scale_space_device(akz_ctx* c, const float* d_img, size_t pitch_bytes) {
  if (!c || !d_img) return AKZ_ERR_ARG;
  const Level& L0 = c->lev[0];
  size_t sp = pitch_bytes ? pitch_bytes : sizeof(float) * L0.w;
  if (sp < sizeof(float) * L0.w || sp % sizeof(float)) return AKZ_ERR_ARG;
  enqueue_scale_space(c, d_img, (int)(sp / sizeof(float)), 0, 1);
  AKZ_HIP_TRY(hipGetLastError());
  c->state = 1;
  c->last_nz = 1;
  return AKZ_OK;
}

akz_status akz_detect(akz_ctx* c, int* n_out) {
  if (!c) return AKZ_ERR_ARG;
  if (c->state < 1) return AKZ_ERR_STATE;
  enqueue_detect(c, 1, true);  // the reference API keeps every level image readable
  AKZ_HIP_TRY(hipGetLastError());
  akz_status st = sync_status(c, 1);
  if (st != AKZ_OK) return st;
  int nk = 0;
  st = image_status(c, 0, &nk);
  if (n_out) *n_out = nk;
  c->state = 2;
  return st;
}

akz_status akz_describe(akz_ctx* c, akz_keypoint* kps, uint8_t* desc, int cap, int* n_out) {
  if (!c || cap < 0) return AKZ_ERR_ARG;
  if (c->state < 2) return AKZ_ERR_STATE;
  enqueue_describe(c, 1);
  AKZ_HIP_TRY(hipGetLastError());
  c->state = 3;
  return akz_read_results(c, kps, desc, cap, n_out);
}

akz_status akz_read_batch_results(akz_ctx* c, int z, akz_keypoint* kps, uint8_t* desc, int cap, int* n_out) {
  if (!c || cap < 0 || z < 0 || z >= c->batch) return AKZ_ERR_ARG;
  if (z >= c->last_nz) return AKZ_ERR_STATE;  // image not part of the last run
  akz_status st = AKZ_OK;
  // the status block is current if it was read back after the last detection
  if (c->status_synced <= z) st = sync_status(c, z + 1);
  if (st != AKZ_OK) return st;
  int nk = 0;
  st = image_status(c, z, &nk);
  int m = std::min(nk, cap);
  if (m > 0 && (kps || desc)) {
    // one contiguous copy each into pinned staging, then repack the 64-byte
    // rows to the descriptor's bytes
    size_t kb = sizeof(akz_kp) * (size_t)m, db = (size_t)AKZ_DESC_STRIDE * m;
    if (c->h_stage_bytes < kb + db) {
      if (c->h_stage) hipHostFree(c->h_stage);
      c->h_stage = nullptr;
      c->h_stage_bytes = 0;
      size_t want = std::max(kb + db, (size_t)1 << 20);
      AKZ_HIP_TRY(hipHostMalloc((void**)&c->h_stage, want, hipHostMallocDefault));
      c->h_stage_bytes = want;
    }
    if (kps) AKZ_HIP_TRY(hipMemcpyAsync(c->h_stage, zptr(c, c->kps, z), kb, hipMemcpyDeviceToHost, c->stream));
    if (desc)
      AKZ_HIP_TRY(hipMemcpyAsync(c->h_stage + kb, zptr(c, c->desc, z), db, hipMemcpyDeviceToHost, c->stream));
    AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
    if (kps) memcpy(kps, c->h_stage, kb);
    if (desc)
      for (int i = 0; i < m; i++)
        memcpy(desc + (size_t)i * c->desc_bytes, c->h_stage + kb + (size_t)i * AKZ_DESC_STRIDE, c->desc_bytes);
  }
  if (n_out) *n_out = nk;
  return (nk > cap) ? AKZ_ERR_OVERFLOW : st;
}

akz_status akz_read_results(akz_ctx* c, akz_keypoint* kps, uint8_t* desc, int cap, int* n_out) {
  return akz_read_batch_results(c, 0, kps, desc, cap, n_out);
}

akz_status akz_read_batch_counts(akz_ctx* c, int n, int* counts, int* statuses) {
  if (!c || n < 1 || n > c->batch) return AKZ_ERR_ARG;
  if (n > c->last_nz) return AKZ_ERR_STATE;
  akz_status st = sync_status(c, n);
  if (st != AKZ_OK) return st;
  for (int z = 0; z < n; z++) {
    int nk = 0;
    akz_status s = image_status(c, z, &nk);
    if (counts) counts[z] = nk;
    if (statuses) statuses[z] = (int)s;
  }
  return AKZ_OK;
}

akz_status akz_process_batch_device_async(akz_ctx* c, const float* d_imgs, size_t pitch_bytes,
                                          size_t image_stride_bytes, int n) {
  if (!c || !d_imgs || n < 1 || n > c->batch) return AKZ_ERR_ARG;
  const Level& L0 = c->lev[0];
  size_t sp = pitch_bytes ? pitch_bytes : sizeof(float) * L0.w;
  if (sp < sizeof(float) * L0.w || sp % sizeof(float)) return AKZ_ERR_ARG;
  size_t is = image_stride_bytes ? image_stride_bytes : sp * L0.h;
  if (n > 1 && is < sp * L0.h) return AKZ_ERR_ARG;
  c->status_synced = 0;  // a graph replay rewrites the status blocks too
  if (!c->graph_mode) {
    enqueue_scale_space(c, d_imgs, (int)(sp / sizeof(float)), is, n);
    enqueue_detect(c, n, c->keep_det);
    enqueue_describe(c, n);
    AKZ_HIP_TRY(hipGetLastError());
  } else {
    // the captured graphs read the context's own input buffers
    if (d_imgs != c->img_in || (int)(sp / sizeof(float)) != L0.pitch || (n > 1 && is != c->zs)) {
      for (int z = 0; z < n; z++)
        AKZ_HIP_TRY(hipMemcpy2DAsync(zptr(c, c->img_in, z), sizeof(float) * L0.pitch,
                                     reinterpret_cast<const char*>(d_imgs) + (size_t)z * is, sp, sizeof(float) * L0.w,
                                     L0.h, hipMemcpyDeviceToDevice, c->stream));
    }
    hipGraphExec_t& ge = c->graph_exec[n - 1];
    if (!ge) {
      AKZ_HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      c->capturing = true;
      enqueue_scale_space(c, c->img_in, L0.pitch, c->zs, n);
      enqueue_detect(c, n, c->keep_det);
      enqueue_describe(c, n);
      c->capturing = false;
      hipError_t e = hipStreamEndCapture(c->stream, &c->graph[n - 1]);
      if (e != hipSuccess) return akz_hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
      AKZ_HIP_TRY(hipGraphInstantiate(&ge, c->graph[n - 1], nullptr, nullptr, 0));
    }
    AKZ_HIP_TRY(hipGraphLaunch(ge, c->stream));
  }
  c->state = 3;
  c->last_nz = n;
  return AKZ_OK;
}

akz_status akz_process_device_async(akz_ctx* c, const float* d_img, size_t pitch_bytes) {
  return akz_process_batch_device_async(c, d_img, pitch_bytes, 0, 1);
}

akz_status akz_batch_input_buffer(akz_ctx* c, float** d_ptr, size_t* pitch_bytes, size_t* image_stride_bytes) {
  if (!c || !d_ptr) return AKZ_ERR_ARG;
  *d_ptr = c->img_in;
  if (pitch_bytes) *pitch_bytes = sizeof(float) * c->lev[0].pitch;
  if (image_stride_bytes) *image_stride_bytes = c->zs;
  return AKZ_OK;
}

akz_status akz_input_buffer(akz_ctx* c, float** d_ptr, size_t* pitch_bytes) {
  return akz_batch_input_buffer(c, d_ptr, pitch_bytes, nullptr);
}

akz_status akz_set_keep_determinant(akz_ctx* c, int keep) {
  if (!c) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  c->keep_det = keep != 0;
  drop_graphs(c);
  return AKZ_OK;
}

akz_status akz_set_suppression_pool(akz_ctx* c, int entries) {
  if (!c || entries < 0 || entries > akz_suppression_pool_max()) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  c->db.nms_pool = entries;
  drop_graphs(c);
  return AKZ_OK;
}

akz_status akz_set_graph_mode(akz_ctx* c, int enabled) {
  if (!c) return AKZ_ERR_ARG;
  c->graph_mode = enabled != 0;
  return AKZ_OK;
}

akz_status akz_wait(akz_ctx* c) {
  if (!c) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  return AKZ_OK;
}

akz_status akz_batch_results_device(akz_ctx* c, int z, const int32_t** d_count, const akz_keypoint** d_kps,
                                    const uint8_t** d_desc64) {
  if (!c || z < 0 || z >= c->batch) return AKZ_ERR_ARG;
  if (d_count) *d_count = zptr(c, c->status, z) + AKZ_ST_KEPT;
  if (d_kps) *d_kps = reinterpret_cast<const akz_keypoint*>(zptr(c, c->kps, z));
  if (d_desc64) *d_desc64 = zptr(c, c->desc, z);
  return AKZ_OK;
}

akz_status akz_results_device(akz_ctx* c, const int32_t** d_count, const akz_keypoint** d_kps,
                              const uint8_t** d_desc64) {
  return akz_batch_results_device(c, 0, d_count, d_kps, d_desc64);
}

akz_status akz_read_level(akz_ctx* c, int lvl, int which, float* dst) {
  return akz_read_batch_level(c, 0, lvl, which, dst);
}

akz_status akz_read_batch_level(akz_ctx* c, int z, int lvl, int which, float* dst) {
  if (!c || !dst || lvl < 0 || lvl >= c->nlev || which < 0 || which > 4 || z < 0 || z >= c->batch)
    return AKZ_ERR_ARG;
  if (which == 4 && !c->det_full_last) return AKZ_ERR_STATE;  // see akz_set_keep_determinant
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  const Level& L = c->lev[lvl];
  if (which == 2 || which == 3) {  // one component of the interleaved derivatives
    std::vector<float> tmp((size_t)2 * L.w * L.h);
    AKZ_HIP_TRY(hipMemcpy2D(tmp.data(), sizeof(float2) * L.w, zptr(c, L.Lxy, z), sizeof(float2) * L.pitch,
                            sizeof(float2) * L.w, L.h, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < (size_t)L.w * L.h; i++) dst[i] = tmp[2 * i + (which - 2)];
    return AKZ_OK;
  }
  const float* src = which == 0 ? L.Lt : which == 1 ? (lvl == 0 ? L.Lt : L.Lsm) : L.det;
  AKZ_HIP_TRY(hipMemcpy2D(dst, sizeof(float) * L.w, zptr(c, src, z), sizeof(float) * L.pitch, sizeof(float) * L.w,
                          L.h, hipMemcpyDeviceToHost));
  return AKZ_OK;
}

akz_status akz_read_kcontrast(akz_ctx* c, int lvl, float* k) {
  if (!c || !k || lvl < 0 || lvl >= c->nlev) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  AKZ_HIP_TRY(hipMemcpy(k, c->kc + lvl, sizeof(float), hipMemcpyDeviceToHost));
  return AKZ_OK;
}

akz_status akz_read_candidates(akz_ctx* c, akz_keypoint* cand, int cap, int* n_out, int32_t* level_start,
                               int* waves) {
  if (!c || cap < 0) return AKZ_ERR_ARG;
  if (c->state < 2) return AKZ_ERR_STATE;
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  int n = 0, w = 0;
  AKZ_HIP_TRY(hipMemcpy(&n, c->ncand, 4, hipMemcpyDeviceToHost));
  AKZ_HIP_TRY(hipMemcpy(&w, c->status + AKZ_ST_WAVES, 4, hipMemcpyDeviceToHost));
  int m = std::min(n, cap);
  if (cand && m > 0) AKZ_HIP_TRY(hipMemcpy(cand, c->cand, sizeof(akz_kp) * m, hipMemcpyDeviceToHost));
  if (level_start) AKZ_HIP_TRY(hipMemcpy(level_start, c->level_start, 4 * (c->nlev + 1), hipMemcpyDeviceToHost));
  if (n_out) *n_out = n;
  if (waves) *waves = w;
  return n > cap ? AKZ_ERR_OVERFLOW : AKZ_OK;
}

akz_status akz_get_timing(akz_ctx* c, akz_timing* t) {
  if (!c || !t) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  auto have = [&](int a, int b) { return (c->ev_mask >> a & 1u) && (c->ev_mask >> b & 1u); };
  // no stage timed by the last run (a graph replay records no events)
  if (!have(EV_START, EV_SCALE)) return AKZ_ERR_STATE;
  akz_timing r{};
  r.kcontrast = have(EV_START, EV_CONTRAST) ? ev_ms(c, EV_START, EV_CONTRAST) : 0.0;
  r.scale = ev_ms(c, EV_START, EV_SCALE);
  if (c->state >= 2 && have(EV_SCALE, EV_NMS)) {
    r.derivatives = have(EV_SCALE, EV_DERIV) ? ev_ms(c, EV_SCALE, EV_DERIV) : 0.0;
    r.extrema = have(EV_DERIV, EV_NMS) ? ev_ms(c, EV_DERIV, EV_NMS) : 0.0;
    r.detector = ev_ms(c, EV_SCALE, EV_NMS);
    r.subpixel = 0.0;  // folded into the extrema kernels, as in the reference
  }
  if (c->state >= 3 && have(EV_NMS, EV_DESC)) r.descriptor = ev_ms(c, EV_NMS, EV_DESC);
  *t = r;
  return AKZ_OK;
}

akz_status akz_profile_fed_batch(akz_ctx* c, const float* d_imgs, size_t pitch_bytes, size_t image_stride_bytes,
                                 int n, float* fed_ms, double* fed_bytes, int* fed_launches, double* min_bytes) {
  if (!c || !d_imgs || n < 1 || n > c->batch) return AKZ_ERR_ARG;
  const Level& L0 = c->lev[0];
  size_t sp = pitch_bytes ? pitch_bytes : sizeof(float) * L0.w;
  if (sp < sizeof(float) * L0.w || sp % sizeof(float)) return AKZ_ERR_ARG;
  size_t is = image_stride_bytes ? image_stride_bytes : sp * L0.h;
  for (int i = 0; i < 2 * c->nlev; i++)
    if (!c->fed_ev[i]) AKZ_HIP_TRY(hipEventCreate(&c->fed_ev[i]));
  c->fed_profile = true;
  enqueue_scale_space(c, d_imgs, (int)(sp / sizeof(float)), is, n);
  c->fed_profile = false;
  enqueue_detect(c, n, c->keep_det);
  enqueue_describe(c, n);
  AKZ_HIP_TRY(hipGetLastError());
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  c->state = 3;
  c->last_nz = n;
  float total = 0.0f;
  double bytes = 0.0, least = 0.0;
  int launches = 0;
  for (int i = 1; i < c->nlev; i++) {
    float ms = 0.0f;
    AKZ_HIP_TRY(hipEventElapsedTime(&ms, c->fed_ev[2 * i], c->fed_ev[2 * i + 1]));
    total += ms;
    const double px = (double)c->lev[i].w * c->lev[i].h * n;
    const int bs = level_block(c, i), ns = c->lev[i].nsteps, nl = (ns + bs - 1) / bs;
    const bool fused = bs > 1 && ns > 0 && akz_fed_variant_fusable(c->fed_variant[i], std::min(bs, ns), n);
    // algorithmic traffic of one step: read L and g, write L (4 B each per
    // pixel); a level whose first launch also builds Lsmooth and g adds that
    // pass's (read L, write Lsmooth and g)
    bytes += 12.0 * px * ns;
    if (fused) bytes += 12.0 * px;
    // what the launches must move however many steps each runs: every launch
    // reads L and g and writes L once; a fused first launch reads its input
    // and writes Lsmooth, L and (when later launches read it) g
    least += 12.0 * px * nl;
    if (fused && nl > 1) least += 4.0 * px;
    // a launch that halves the previous level reads it (16 B per pixel of
    // the level) instead of the halved image
    if (level_half_in_fed(c, i, n)) least += 12.0 * px;
    launches += nl;
  }
  if (fed_ms) *fed_ms = total;
  if (fed_bytes) *fed_bytes = bytes;
  if (fed_launches) *fed_launches = launches;
  if (min_bytes) *min_bytes = least;
  return AKZ_OK;
}

akz_status akz_profile_fed(akz_ctx* c, const float* d_img, size_t pitch_bytes, float* fed_ms, double* fed_bytes,
                           int* fed_launches, double* min_bytes) {
  return akz_profile_fed_batch(c, d_img, pitch_bytes, 0, 1, fed_ms, fed_bytes, fed_launches, min_bytes);
}

akz_status akz_prop_positions(const akz_ctx* c, int* per_image) {
  if (!c || !per_image) return AKZ_ERR_ARG;
  *per_image = c->psrc.g.P;
  return AKZ_OK;
}

akz_status akz_prop_extract(akz_ctx* c, int n, uint16_t* d_feat, uint8_t* d_labels) {
  if (!c || !d_feat || n < 1 || n > c->batch) return AKZ_ERR_ARG;
  if (c->state < 2 || n > c->last_nz || !c->det_full_last) return AKZ_ERR_STATE;
  akz_status st = ensure_prop_act(c);
  if (st != AKZ_OK) return st;
  akz_prop_act act = c->pact;
  act.x0 = d_feat;
  akz_prop_launch_features(c->stream, c->psrc, act, n);
  if (d_labels) akz_prop_launch_labels(c->stream, c->db, c->psrc.g, d_labels, n);
  AKZ_HIP_TRY(hipGetLastError());
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  return AKZ_OK;
}

akz_status akz_prop_forward(akz_prop_net* net, const akz_ctx* c, const uint16_t* d_feat, const int* idx, int nimg,
                            float* probs) {
  if (!net || !c || !d_feat || nimg < 1) return AKZ_ERR_ARG;
  return akz_prop_forward_geom(net, c->psrc.g, d_feat, idx, nimg, probs);
}

akz_status akz_prop_train_step(akz_prop_net* net, const akz_ctx* c, const uint16_t* d_feat,
                               const uint8_t* d_labels, const int* idx, int nimg, float pos_weight, float lr,
                               float* grad, float* loss) {
  if (!net || !c || !d_feat || !d_labels || nimg < 1 || !(pos_weight > 0.0f) || lr < 0.0f) return AKZ_ERR_ARG;
  return akz_prop_grad_geom(net, c->psrc.g, d_feat, d_labels, idx, nimg, pos_weight, lr, grad, loss);
}

akz_status akz_set_proposal(akz_ctx* c, akz_prop_net* net, float threshold) {
  if (!c || !(threshold >= 0.0f && threshold <= 1.0f)) return AKZ_ERR_ARG;
  if (net) {
    akz_status st = ensure_prop_act(c);
    if (st != AKZ_OK) return st;
  }
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  c->prop_net = net;
  c->prop_thr = threshold;
  drop_graphs(c);
  return AKZ_OK;
}

akz_status akz_prop_read_probs(akz_ctx* c, int z, float* probs) {
  if (!c || !probs || z < 0 || z >= c->batch) return AKZ_ERR_ARG;
  if (!c->prop_net || z >= c->last_nz) return AKZ_ERR_STATE;
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  AKZ_HIP_TRY(hipMemcpy(probs, c->pact.prob + (size_t)z * c->psrc.g.P, sizeof(float) * c->psrc.g.P,
                        hipMemcpyDeviceToHost));
  return AKZ_OK;
}

// ---- autotuner ----
akz_status akz_autotune(akz_ctx* c, int nz, int reps) {
  if (!c || nz < 1 || nz > c->batch || reps < 1) return AKZ_ERR_ARG;
  static const int kVariants[] = {2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19};
  hipEvent_t a, b;
  AKZ_HIP_TRY(hipEventCreate(&a));
  AKZ_HIP_TRY(hipEventCreate(&b));
  const akz_batch bt{nz, c->zs};
  akz_status result = AKZ_OK;
  for (int i = 1; i < c->nlev && result == AKZ_OK; i++) {
    Level& L = c->lev[i];
    if (L.nsteps == 0 || c->fed_block < 2) continue;
    // same-size levels share the choice of the first of them
    bool done = false;
    for (int j = 1; j < i; j++)
      if (c->lev[j].w == L.w && c->lev[j].h == L.h && c->lev[j].nsteps > 0) {
        c->fed_variant[i] = c->fed_variant[j];
        done = true;
        break;
      }
    if (done) continue;
    float ht[AKZ_MAX_TAU];
    for (int q = 0; q < L.nsteps; q++) ht[q] = 0.5f * L.tau[q];
    float best = 1e30f;
    int bestv = c->fed_variant[i];
    for (int v : kVariants) {
      // the level's whole FED run with this variant (its steps per launch),
      // once to warm up, then `reps` timed times, ping-ponging Lt and the scratch
      const int bs = std::min(c->fed_block, akz_fed_variant_steps(v));
      const int nl = (L.nsteps + bs - 1) / bs;
      const bool fuse = bs > 1 && akz_fed_variant_fusable(v, std::min(bs, L.nsteps), nz);
      // as in the pipeline: smooth + flow, fused into the first launch of
      // the blocked variants
      auto run_level = [&](int parity) {
        int launch = parity;
        if (!fuse)
          akz_launch_smooth_flow(c->stream, (parity & 1) ? c->fed_tmp : L.Lt, L.pitch, L.Lsm, c->g, L.w, L.h,
                                 c->opts.diffusivity, c->kc + i, c->taps1, bt);
        for (int j = 0; j < L.nsteps; j += bs, launch++) {
          const bool even = (launch & 1) == 0;
          const float* src = even ? L.Lt : c->fed_tmp;
          float* dst = even ? c->fed_tmp : L.Lt;
          if (fuse && j == 0)
            akz_launch_fed_variant_sf(c->stream, v, src, nullptr, 0, L.Lsm, nl > 1 ? c->g : nullptr, dst, L.pitch,
                                      L.w, L.h, c->opts.diffusivity, c->kc + i, c->taps1, ht + j,
                                      std::min(bs, L.nsteps - j), bt);
          else
            akz_launch_fed_variant(c->stream, v, src, c->g, dst, L.pitch, L.w, L.h, ht + j,
                                   std::min(bs, L.nsteps - j), bt);
        }
      };
      run_level(0);
      if (hipEventRecord(a, c->stream) != hipSuccess) { result = AKZ_ERR_HIP; break; }
      for (int r = 0; r < reps; r++) run_level(r);
      if (hipEventRecord(b, c->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess) {
        result = AKZ_ERR_HIP;
        break;
      }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, a, b);
      if (ms < best) { best = ms; bestv = v; }
    }
    c->fed_variant[i] = bestv;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (result != AKZ_OK) { (void)hipGetLastError(); return result; }
  AKZ_HIP_TRY(hipGetLastError());
  drop_graphs(c);
  c->state = 0;  // level contents were overwritten by the timing launches
  return AKZ_OK;
}

akz_status akz_get_fed_variants(const akz_ctx* c, int* variants) {
  if (!c || !variants) return AKZ_ERR_ARG;
  for (int i = 0; i < c->nlev; i++) variants[i] = c->fed_variant[i];
  return AKZ_OK;
}

akz_status akz_set_fed_variants(akz_ctx* c, const int* variants) {
  if (!c || !variants) return AKZ_ERR_ARG;
  for (int i = 0; i < c->nlev; i++)
    if (variants[i] < 2 || variants[i] > AKZ_FED_VARIANT_MAX) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->nlev; i++) c->fed_variant[i] = variants[i];
  drop_graphs(c);
  return AKZ_OK;
}

// ---- ragged batches: one batch context (and stream) per size class ----
akz_status akz_ragged_create(const akz_options* base, int nclass, const int* widths, const int* heights,
                             const int* capacity, akz_ragged** out) {
  if (!base || !out || nclass < 1 || nclass > AKZ_MAX_CLASSES || !widths || !heights || !capacity)
    return AKZ_ERR_ARG;
  *out = nullptr;
  for (int k = 0; k < nclass; k++)
    for (int j = 0; j < k; j++)
      if (widths[j] == widths[k] && heights[j] == heights[k]) return AKZ_ERR_ARG;
  akz_ragged* r = new (std::nothrow) akz_ragged();
  if (!r) return AKZ_ERR_OOM;
  r->ncls = nclass;
  for (int k = 0; k < nclass; k++) {
    akz_options o = *base;
    o.img_width = widths[k];
    o.img_height = heights[k];
    akz_status st = akz_create_batch(&o, capacity[k], &r->cls[k]);
    if (st != AKZ_OK) { akz_ragged_destroy(r); return st; }
  }
  bool ok = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&r->ev_fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&r->ev_mid, hipEventDisableTiming) == hipSuccess &&
            hipMalloc((void**)&r->d_detp, sizeof(akz_det_params) * nclass) == hipSuccess &&
            hipMalloc((void**)&r->d_descp, sizeof(akz_desc_params) * nclass) == hipSuccess;
  for (int k = 0; k < nclass && ok; k++)
    ok = hipEventCreateWithFlags(&r->ev_join[k], hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&r->ev_join2[k], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    akz_ragged_destroy(r);
    return AKZ_ERR_HIP;
  }
  r->detp_host.resize(nclass);
  std::memset(r->detp_host.data(), 0xff, sizeof(akz_det_params) * nclass);  // uploaded on first use
  *out = r;
  return AKZ_OK;
}

namespace {

void ragged_drop_graphs(akz_ragged* r) {
  for (auto& g : r->graphs) {
    if (g.second.second) hipGraphExecDestroy(g.second.second);
    if (g.second.first) hipGraphDestroy(g.second.first);
  }
  r->graphs.clear();
}

// Work-list execution of a ragged call whose images are staged in slots
// 0..cnt[k]-1 of the class contexts: per class on its own stream, forked
// from r->stream, the contrast factor and scale space; then on r->stream one
// persistent detector launch walking the (image, level, tile) work list of
// all classes; per class again the candidates and the suppression; then one
// persistent orientation and one MLDB launch walking the (image, keypoint)
// lists.  Everything joins back into r->stream.
akz_status ragged_enqueue_lists(akz_ragged* r, const int* cnt) {
  hipStream_t s = r->stream;
  AKZ_HIP_TRY(hipEventRecord(r->ev_fork, s));
  akz_det_worklist dw{};
  akz_desc_worklist ew{};
  int tiles = 0;
  for (int k = 0; k < r->ncls; k++) {
    if (cnt[k] == 0) continue;
    akz_ctx* c = r->cls[k];
    AKZ_HIP_TRY(hipStreamWaitEvent(c->stream, r->ev_fork, 0));
    enqueue_scale_space(c, c->img_in, c->lev[0].pitch, c->zs, cnt[k]);
    detect_prepare(c, cnt[k], c->keep_det);
    AKZ_HIP_TRY(hipEventRecord(r->ev_join[k], c->stream));
    AKZ_HIP_TRY(hipStreamWaitEvent(s, r->ev_join[k], 0));
    dw.seg_count[k] = c->db.seg_count;
    dw.seg_slots[k] = c->db.seg_slots;
    ew.kps[k] = c->kps;
    ew.nkps[k] = c->status + AKZ_ST_KEPT;
    ew.desc[k] = c->desc;
    for (int z = 0; z < cnt[k]; z++) {
      dw.cls[dw.nimg] = k;
      dw.z[dw.nimg] = z;
      dw.start[dw.nimg] = tiles;
      tiles += c->det32.tile_start[c->nlev];
      dw.nimg++;
      ew.cls[ew.nimg] = k;
      ew.z[ew.nimg] = z;
      ew.nimg++;
    }
  }
  dw.start[dw.nimg] = tiles;
  dw.P = r->d_detp;
  ew.P = r->d_descp;
  if (dw.nimg > 0) akz_launch_detect_fused_list(s, dw, std::min(tiles, 1024));
  AKZ_HIP_TRY(hipEventRecord(r->ev_mid, s));
  bool orient = false;
  for (int k = 0; k < r->ncls; k++) {
    if (cnt[k] == 0) continue;
    akz_ctx* c = r->cls[k];
    AKZ_HIP_TRY(hipStreamWaitEvent(c->stream, r->ev_mid, 0));
    detect_finish(c, cnt[k], true);
    AKZ_HIP_TRY(hipEventRecord(r->ev_join2[k], c->stream));
    AKZ_HIP_TRY(hipStreamWaitEvent(s, r->ev_join2[k], 0));
    orient |= c->opts.descriptor == 5;
  }
  if (ew.nimg > 0) akz_launch_describe_list(s, ew, orient, 2048, 4096);
  return AKZ_OK;
}

// Whether every class can use the persistent detector (derivative steps
// within the fused kernel's limit) and the shared descriptor settings.
bool ragged_lists_apply(const akz_ragged* r) {
  for (int k = 0; k < r->ncls; k++) {
    const akz_ctx* c = r->cls[k];
    for (int i = 0; i < c->nlev; i++)
      if (c->lev[i].deriv_step > akz_detect_max_step()) return false;
  }
  return true;
}

}  // namespace

akz_status akz_ragged_destroy(akz_ragged* r) {
  if (!r) return AKZ_ERR_ARG;
  if (r->stream) hipStreamSynchronize(r->stream);
  ragged_drop_graphs(r);
  for (int k = 0; k < r->ncls; k++)
    if (r->cls[k]) akz_destroy(r->cls[k]);
  if (r->d_detp) hipFree(r->d_detp);
  if (r->d_descp) hipFree(r->d_descp);
  if (r->ev_fork) hipEventDestroy(r->ev_fork);
  if (r->ev_mid) hipEventDestroy(r->ev_mid);
  for (int k = 0; k < AKZ_MAX_CLASSES; k++) {
    if (r->ev_join[k]) hipEventDestroy(r->ev_join[k]);
    if (r->ev_join2[k]) hipEventDestroy(r->ev_join2[k]);
  }
  if (r->stream) hipStreamDestroy(r->stream);
  delete r;
  return AKZ_OK;
}

akz_status akz_ragged_set_work_lists(akz_ragged* r, int enabled) {
  if (!r) return AKZ_ERR_ARG;
  r->lists = enabled != 0;
  return AKZ_OK;
}

akz_status akz_ragged_context(akz_ragged* r, int cls, akz_ctx** ctx) {
  if (!r || !ctx || cls < 0 || cls >= r->ncls) return AKZ_ERR_ARG;
  *ctx = r->cls[cls];
  return AKZ_OK;
}

akz_status akz_ragged_process(akz_ragged* r, int n, const float* const* d_imgs, const size_t* pitch_bytes,
                              const int* widths, const int* heights, int* counts) {
  if (!r || n < 1 || !d_imgs || !widths || !heights) return AKZ_ERR_ARG;
  std::vector<int> cls(n), slot(n);
  int cnt[AKZ_MAX_CLASSES] = {};
  for (int i = 0; i < n; i++) {
    int k = 0;
    while (k < r->ncls && !(r->cls[k]->lev[0].w == widths[i] && r->cls[k]->lev[0].h == heights[i])) k++;
    if (k == r->ncls || !d_imgs[i]) return AKZ_ERR_ARG;
    const size_t sp = pitch_bytes && pitch_bytes[i] ? pitch_bytes[i] : sizeof(float) * widths[i];
    if (sp < sizeof(float) * widths[i]) return AKZ_ERR_ARG;
    cls[i] = k;
    slot[i] = cnt[k]++;
    if (cnt[k] > r->cls[k]->batch) return AKZ_ERR_ARG;
  }
  if (r->lists && ragged_lists_apply(r)) {
    // stage every image into its class context's input slot on r->stream,
    // then the work-list pipeline (captured once per class composition in
    // graph mode)
    for (int i = 0; i < n; i++) {
      akz_ctx* c = r->cls[cls[i]];
      const Level& L0 = c->lev[0];
      const size_t sp = pitch_bytes && pitch_bytes[i] ? pitch_bytes[i] : sizeof(float) * widths[i];
      AKZ_HIP_TRY(hipMemcpy2DAsync(zptr(c, c->img_in, slot[i]), sizeof(float) * L0.pitch, d_imgs[i], sp,
                                   sizeof(float) * L0.w, L0.h, hipMemcpyDeviceToDevice, r->stream));
    }
    if (!r->descp_ready) {
      for (int k = 0; k < r->ncls; k++)
        AKZ_HIP_TRY(hipMemcpy(r->d_descp + k, &r->cls[k]->dp, sizeof(akz_desc_params), hipMemcpyHostToDevice));
      r->descp_ready = true;
    }
    // the detector parameters as the class contexts hold them now (flags and
    // proposal keys change with their settings), in device memory before
    // anything reads them; graphs built with other settings are dropped
    bool changed = false;
    unsigned gen = 0;
    for (int k = 0; k < r->ncls; k++) {
      akz_ctx* c = r->cls[k];
      detect_params(c, c->keep_det);
      gen += c->config_gen;
      if (std::memcmp(&r->detp_host[k], &c->det32, sizeof(akz_det_params)) != 0) {
        r->detp_host[k] = c->det32;
        changed = true;
      }
    }
    if (changed) {
      AKZ_HIP_TRY(hipStreamSynchronize(r->stream));
      AKZ_HIP_TRY(hipMemcpy(r->d_detp, r->detp_host.data(), sizeof(akz_det_params) * r->ncls,
                            hipMemcpyHostToDevice));
    }
    if (changed || gen != r->gen) {
      ragged_drop_graphs(r);
      r->gen = gen;
    }
    std::vector<int> key(cnt, cnt + r->ncls);
    if (!r->graph) {
      akz_status st = ragged_enqueue_lists(r, cnt);
      if (st != AKZ_OK) return st;
    } else {
      auto it = r->graphs.find(key);
      if (it == r->graphs.end()) {
        AKZ_HIP_TRY(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < r->ncls; k++) r->cls[k]->capturing = true;
        akz_status st = ragged_enqueue_lists(r, cnt);
        for (int k = 0; k < r->ncls; k++) r->cls[k]->capturing = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(r->stream, &g);
        if (st != AKZ_OK) {
          if (g) hipGraphDestroy(g);
          return st;
        }
        if (e != hipSuccess) return akz_hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        hipGraphExec_t ge = nullptr;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
          hipGraphDestroy(g);
          return akz_hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        it = r->graphs.emplace(key, std::make_pair(g, ge)).first;
      }
      AKZ_HIP_TRY(hipGraphLaunch(it->second.second, r->stream));
    }
    AKZ_HIP_TRY(hipGetLastError());
    AKZ_HIP_TRY(hipStreamSynchronize(r->stream));
    for (int k = 0; k < r->ncls; k++) {
      if (cnt[k] == 0) continue;
      r->cls[k]->state = 3;
      r->cls[k]->last_nz = cnt[k];
      r->cls[k]->status_synced = 0;
    }
  } else {
    // stage every image into its class context's input slot, then one batch
    // call per class; the classes run concurrently on their own streams
    for (int i = 0; i < n; i++) {
      akz_ctx* c = r->cls[cls[i]];
      const Level& L0 = c->lev[0];
      const size_t sp = pitch_bytes && pitch_bytes[i] ? pitch_bytes[i] : sizeof(float) * widths[i];
      AKZ_HIP_TRY(hipMemcpy2DAsync(zptr(c, c->img_in, slot[i]), sizeof(float) * L0.pitch, d_imgs[i], sp,
                                   sizeof(float) * L0.w, L0.h, hipMemcpyDeviceToDevice, c->stream));
    }
    for (int k = 0; k < r->ncls; k++) {
      if (cnt[k] == 0) continue;
      akz_ctx* c = r->cls[k];
      akz_status st = akz_process_batch_device_async(c, c->img_in, sizeof(float) * c->lev[0].pitch, c->zs, cnt[k]);
      if (st != AKZ_OK) return st;
    }
  }
  r->img_cls = cls;
  r->img_slot = slot;
  for (int k = 0; k < r->ncls; k++) {
    if (cnt[k] == 0) continue;
    akz_status st = sync_status(r->cls[k], cnt[k]);
    if (st != AKZ_OK) return st;
  }
  if (counts)
    for (int i = 0; i < n; i++) image_status(r->cls[cls[i]], slot[i], &counts[i]);
  return AKZ_OK;
}

akz_status akz_ragged_read(akz_ragged* r, int i, akz_keypoint* kps, uint8_t* desc, int cap, int* n_out) {
  if (!r || i < 0 || i >= (int)r->img_cls.size()) return AKZ_ERR_ARG;
  return akz_read_batch_results(r->cls[r->img_cls[i]], r->img_slot[i], kps, desc, cap, n_out);
}

akz_status akz_ragged_results_device(akz_ragged* r, int i, const int32_t** d_count, const akz_keypoint** d_kps,
                                     const uint8_t** d_desc64) {
  if (!r || i < 0 || i >= (int)r->img_cls.size()) return AKZ_ERR_ARG;
  return akz_batch_results_device(r->cls[r->img_cls[i]], r->img_slot[i], d_count, d_kps, d_desc64);
}

akz_status akz_ragged_set_graph_mode(akz_ragged* r, int enabled) {
  if (!r) return AKZ_ERR_ARG;
  for (int k = 0; k < r->ncls; k++) akz_set_graph_mode(r->cls[k], enabled);
  r->graph = enabled != 0;
  return AKZ_OK;
}

akz_status akz_ragged_autotune(akz_ragged* r, int reps) {
  if (!r) return AKZ_ERR_ARG;
  for (int k = 0; k < r->ncls; k++) {
    akz_status st = akz_autotune(r->cls[k], r->cls[k]->batch, reps);
    if (st != AKZ_OK) return st;
  }
  return AKZ_OK;
}

akz_status akz_copy_kernel_run(const float* d_src, float* d_dst, size_t nfloats, int iters, float* avg_ms) {
  if (!d_src || !d_dst || iters < 1 || nfloats == 0) return AKZ_ERR_ARG;
  int w = 4096;
  int h = (int)((nfloats + w - 1) / w);
  if ((size_t)w * h != nfloats) return AKZ_ERR_ARG;
  hipEvent_t a, b;
  AKZ_HIP_TRY(hipEventCreate(&a));
  AKZ_HIP_TRY(hipEventCreate(&b));
  AKZ_HIP_TRY(hipEventRecord(a, nullptr));
  for (int i = 0; i < iters; i++) akz_launch_copy2d(nullptr, d_src, w, d_dst, w, w, h, akz_batch{1, 0});
  AKZ_HIP_TRY(hipEventRecord(b, nullptr));
  AKZ_HIP_TRY(hipEventSynchronize(b));
  float ms = 0.0f;
  AKZ_HIP_TRY(hipEventElapsedTime(&ms, a, b));
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (avg_ms) *avg_ms = ms / iters;
  return AKZ_OK;
}

// ---- matcher ----
akz_status akz_matcher_create(akz_matcher** out) {
  if (!out) return AKZ_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { (void)hipGetLastError(); return AKZ_ERR_NO_DEVICE; }
  akz_matcher* m = new (std::nothrow) akz_matcher();
  if (!m) return AKZ_ERR_OOM;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) { delete m; return AKZ_ERR_HIP; }
  hipEventCreate(&m->e0);
  hipEventCreate(&m->e1);
  *out = m;
  return AKZ_OK;
}

akz_status akz_matcher_destroy(akz_matcher* m) {
  if (!m) return AKZ_ERR_ARG;
  hipStreamSynchronize(m->stream);
  if (m->q) hipFree(m->q);
  if (m->t) hipFree(m->t);
  if (m->out) hipFree(m->out);
  if (m->partial) hipFree(m->partial);
  if (m->e0) hipEventDestroy(m->e0);
  if (m->e1) hipEventDestroy(m->e1);
  hipStreamDestroy(m->stream);
  delete m;
  return AKZ_OK;
}

static akz_status grow(void** p, size_t* cap, size_t need) {
  if (need <= *cap) return AKZ_OK;
  if (*p) hipFree(*p);
  *p = nullptr;
  *cap = 0;
  AKZ_HIP_TRY(hipMalloc(p, need));
  *cap = need;
  return AKZ_OK;
}

akz_status akz_match2nn_device(akz_matcher* m, const uint8_t* d_q64, int nq, const uint8_t* d_t64, int nt,
                               akz_dmatch* d_out) {
  if (!m || nq < 0 || nt < 0 || (nq > 0 && (!d_q64 || !d_out)) || (nt > 0 && !d_t64)) return AKZ_ERR_ARG;
  if (nt >= (1 << 22)) return AKZ_ERR_UNSUPPORTED;
  if (nq == 0) return AKZ_OK;
  int ns = akz_match_splits(nq, nt);
  akz_status st = grow((void**)&m->partial, &m->pcap, sizeof(uint32_t) * 2 * (size_t)nq * ns);
  if (st != AKZ_OK) return st;
  hipEventRecord(m->e0, m->stream);
  akz_launch_match(m->stream, d_q64, nq, d_t64, nt, m->partial, ns);
  akz_launch_match_merge(m->stream, m->partial, nq, ns, d_out);
  hipEventRecord(m->e1, m->stream);
  AKZ_HIP_TRY(hipGetLastError());
  return AKZ_OK;
}

akz_status akz_match2nn(akz_matcher* m, const uint8_t* q, int nq, const uint8_t* t, int nt, size_t stride,
                        akz_dmatch* out) {
  if (!m || nq < 0 || nt < 0 || stride < 1 || (nq > 0 && (!q || !out)) || (nt > 0 && !t)) return AKZ_ERR_ARG;
  const size_t row = stride < 61 ? stride : 61;  // descriptor bytes used per row
  if (nq == 0) return AKZ_OK;
  akz_status st;
  if ((st = grow((void**)&m->q, &m->qcap, (size_t)nq * 64)) != AKZ_OK) return st;
  if ((st = grow((void**)&m->t, &m->tcap, (size_t)std::max(nt, 1) * 64)) != AKZ_OK) return st;
  if ((st = grow((void**)&m->out, &m->ocap, sizeof(akz_dmatch) * 2 * (size_t)nq)) != AKZ_OK) return st;
  // rows into zero-padded 64-byte rows (AKAZE.cpp:39-60)
  AKZ_HIP_TRY(hipMemsetAsync(m->q, 0, (size_t)nq * 64, m->stream));
  AKZ_HIP_TRY(hipMemcpy2DAsync(m->q, 64, q, stride, row, nq, hipMemcpyHostToDevice, m->stream));
  if (nt > 0) {
    AKZ_HIP_TRY(hipMemsetAsync(m->t, 0, (size_t)nt * 64, m->stream));
    AKZ_HIP_TRY(hipMemcpy2DAsync(m->t, 64, t, stride, row, nt, hipMemcpyHostToDevice, m->stream));
  }
  st = akz_match2nn_device(m, m->q, nq, m->t, nt, m->out);
  if (st != AKZ_OK) return st;
  AKZ_HIP_TRY(hipMemcpyAsync(out, m->out, sizeof(akz_dmatch) * 2 * (size_t)nq, hipMemcpyDeviceToHost, m->stream));
  AKZ_HIP_TRY(hipStreamSynchronize(m->stream));
  return AKZ_OK;
}

akz_status akz_matcher_wait(akz_matcher* m) {
  if (!m) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipStreamSynchronize(m->stream));
  return AKZ_OK;
}

akz_status akz_matcher_last_ms(akz_matcher* m, float* ms) {
  if (!m || !ms) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipEventSynchronize(m->e1));
  AKZ_HIP_TRY(hipEventElapsedTime(ms, m->e0, m->e1));
  return AKZ_OK;
}

// ---- device memory helpers ----
akz_status akz_dev_alloc(size_t bytes, void** ptr) {
  if (!ptr) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
  return AKZ_OK;
}
akz_status akz_dev_free(void* ptr) {
  if (ptr) AKZ_HIP_TRY(hipFree(ptr));
  return AKZ_OK;
}
akz_status akz_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  AKZ_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return AKZ_OK;
}
akz_status akz_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  AKZ_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return AKZ_OK;
}
akz_status akz_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  if ((!dst || !src) && bytes) return AKZ_ERR_ARG;
  if (bytes) AKZ_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return AKZ_OK;
}

akz_status akz_memcpy_2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes,
                         size_t height) {
  if (!dst || !src || dpitch < width_bytes || spitch < width_bytes) return AKZ_ERR_ARG;
  AKZ_HIP_TRY(hipMemcpy2D(dst, dpitch, src, spitch, width_bytes, height, hipMemcpyDefault));
  return AKZ_OK;
}

akz_status akz_device_synchronize(void) {
  AKZ_HIP_TRY(hipDeviceSynchronize());
  return AKZ_OK;
}

akz_status akz_selftest_reciprocal(uint64_t* mismatches) {
  if (!mismatches) return AKZ_ERR_ARG;
  unsigned long long* d = nullptr;
  AKZ_HIP_TRY(hipMalloc(&d, sizeof(*d)));
  unsigned long long h = 0;
  hipError_t e = hipMemcpy(d, &h, sizeof(h), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    akz_launch_rcp_check(nullptr, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return AKZ_ERR_HIP;
  *mismatches = h;
  return AKZ_OK;
}

// ---- synthetic inputs ----
akz_status akz_synth_image(uint64_t seed, int w, int h, uint8_t* out) {
  if (!out || w <= 0 || h <= 0) return AKZ_ERR_ARG;
  akz_synth_render(seed, w, h, out);
  return AKZ_OK;
}
akz_status akz_synth_homography(uint64_t seed, int w, int h, double max_rot_deg, double min_scale, double max_scale,
                                double max_persp, double* H9) {
  if (!H9 || w <= 0 || h <= 0) return AKZ_ERR_ARG;
  akz_synth_homography_impl(seed, w, h, max_rot_deg, min_scale, max_scale, max_persp, H9);
  return AKZ_OK;
}
akz_status akz_synth_warp(const uint8_t* src, int w, int h, const double* H9, double noise_std, uint64_t noise_seed,
                          uint8_t* out) {
  if (!src || !out || !H9 || w <= 0 || h <= 0) return AKZ_ERR_ARG;
  return akz_synth_warp_u8(src, w, h, H9, noise_std, noise_seed, out) == 0 ? AKZ_OK : AKZ_ERR_ARG;
}

}  // extern "C"
