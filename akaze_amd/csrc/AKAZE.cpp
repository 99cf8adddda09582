// libAKAZE-compatible C++ classes over the C-ABI (host code).
// Follows the call structure of the reference's AKAZE.cpp:36-495.
#include "../../include/akaze/AKAZE.h"

#include <algorithm>
#include <iomanip>
#include <iostream>

namespace libAKAZEMI {

static void check(akz_status s, const char* where) {
  if (s != AKZ_OK) throw Error(s, where);
}

std::ostream& operator<<(std::ostream& os, const AKAZEOptions& o) {
  // the option dump of AKAZEConfig.h:145-171
  os << std::left;
  os << std::setw(33) << "omax" << " =  " << o.omax << "\n";
  os << std::setw(33) << "nsublevels" << " =  " << o.nsublevels << "\n";
  os << std::setw(33) << "soffset" << " =  " << o.soffset << "\n";
  os << std::setw(33) << "sderivatives" << " =  " << o.sderivatives << "\n";
  os << std::setw(33) << "diffusivity" << " =  " << o.diffusivity << "\n";
  os << std::setw(33) << "dthreshold" << " =  " << o.dthreshold << "\n";
  os << std::setw(33) << "descriptor" << " =  " << o.descriptor << "\n";
  os << std::setw(33) << "descriptor_channels" << " =  " << o.descriptor_channels << "\n";
  os << std::setw(33) << "descriptor_size" << " =  " << o.descriptor_size << "\n";
  os << std::setw(33) << "save_scale_space" << " =  " << o.save_scale_space << "\n";
  os << std::setw(33) << "verbosity" << " =  " << o.verbosity << "\n";
  return os;
}

static akz_options to_c(const AKAZEOptions& o) {
  akz_options c;
  akz_options_default(&c);
  c.omax = o.omax;
  c.nsublevels = o.nsublevels;
  c.img_width = o.img_width;
  c.img_height = o.img_height;
  c.soffset = o.soffset;
  c.derivative_factor = o.derivative_factor;
  c.sderivatives = o.sderivatives;
  c.diffusivity = (int)o.diffusivity;
  c.dthreshold = o.dthreshold;
  c.min_dthreshold = o.min_dthreshold;
  c.descriptor = (int)o.descriptor;
  c.descriptor_size = o.descriptor_size;
  c.descriptor_channels = o.descriptor_channels;
  c.descriptor_pattern_size = o.descriptor_pattern_size;
  c.kcontrast = o.kcontrast;
  c.kcontrast_percentile = o.kcontrast_percentile;
  c.kcontrast_nbins = (int)o.kcontrast_nbins;
  c.maxkeypoints = o.maxkeypoints;
  c.verbosity = o.verbosity;
  return c;
}

// ---- Matcher ----
Matcher::Matcher() { check(akz_matcher_create(&m_), "Matcher"); }
Matcher::~Matcher() {
  if (m_) akz_matcher_destroy(m_);
}

void Matcher::bfmatch(Mat& q, Mat& t, std::vector<std::vector<DMatch>>& dmatches) {
  if (q.type != CV_8U || t.type != CV_8U || (q.rows && t.rows && q.cols != t.cols) || (q.rows && q.cols < 1))
    throw Error(AKZ_ERR_ARG, "Matcher::bfmatch");
  const int cols = q.rows ? std::min(q.cols, 61) : 61;
  std::vector<akz_dmatch> out(2 * (size_t)q.rows);
  // rows of a Mat are `step` bytes apart; both sets must share one stride for the C call
  Mat tq = q, tt = t;
  if (q.step != t.step) {
    tq = Mat(q.rows, cols, CV_8U);
    for (int r = 0; r < q.rows; r++) std::memcpy(tq.ptr<uint8_t>(r), q.ptr<uint8_t>(r), cols);
    tt = Mat(t.rows, cols, CV_8U);
    for (int r = 0; r < t.rows; r++) std::memcpy(tt.ptr<uint8_t>(r), t.ptr<uint8_t>(r), cols);
  }
  check(akz_match2nn(m_, tq.data, tq.rows, tt.data, tt.rows, tq.rows ? tq.step : (size_t)cols, out.data()),
        "Matcher::bfmatch");
  dmatches.clear();
  dmatches.reserve(q.rows);
  for (int i = 0; i < q.rows; ++i) {
    std::vector<DMatch> two(2);
    std::memcpy(&two[0], &out[2 * (size_t)i], sizeof(DMatch));
    std::memcpy(&two[1], &out[2 * (size_t)i + 1], sizeof(DMatch));
    dmatches.push_back(two);
  }
}

Mat Matcher::bfmatch_(Mat q, Mat t) {
  std::vector<std::vector<DMatch>> v;
  bfmatch(q, t, v);
  Mat m((int)v.size(), 8, CV_32F);
  for (size_t i = 0; i < v.size(); ++i) {
    float* d = m.ptr<float>((int)i);
    d[0] = (float)v[i][0].queryIdx; d[1] = (float)v[i][0].trainIdx; d[2] = 0.f; d[3] = v[i][0].distance;
    d[4] = (float)v[i][1].queryIdx; d[5] = (float)v[i][1].trainIdx; d[6] = 0.f; d[7] = v[i][1].distance;
  }
  return m;
}

void MatchDescriptors(Mat& q, Mat& t, std::vector<std::vector<DMatch>>& dmatches) {
  Matcher m;
  std::vector<std::vector<DMatch>> out;
  m.bfmatch(q, t, out);
  for (auto& r : out) dmatches.push_back(r);  // the reference appends (cuda_akaze.cu:2018-2024)
}

// ---- AKAZE ----
AKAZE::AKAZE(const AKAZEOptions& options) : options_(options) {
  akz_options c = to_c(options_);
  check(akz_create(&c, &ctx_), "AKAZE::AKAZE");
}

AKAZE::~AKAZE() {
  if (ctx_) akz_destroy(ctx_);
}

int AKAZE::Num_Levels() const {
  int n = 0;
  akz_num_levels(ctx_, &n);
  return n;
}

int AKAZE::Create_Nonlinear_Scale_Space(const Mat& img) {
  if (img.type != CV_32F || img.rows != options_.img_height || img.cols != options_.img_width) {
    std::cerr << "Error generating the nonlinear scale space!!" << std::endl;
    return -1;
  }
  if (akz_build_scale_space(ctx_, reinterpret_cast<const float*>(img.data), img.step) != AKZ_OK) return -1;
  refresh_timing();
  return 0;
}

void AKAZE::Feature_Detection(std::vector<KeyPoint>& kpts) {
  int n = 0;
  check(akz_detect(ctx_, &n), "AKAZE::Feature_Detection");
  nump_ = n;
  kpts.resize(n);
  int got = 0;
  if (n > 0) {
    // keypoints without orientation: read the current detector output
    std::vector<akz_keypoint> tmp(n);
    check(akz_read_results(ctx_, tmp.data(), nullptr, n, &got), "AKAZE::Feature_Detection");
    std::memcpy(kpts.data(), tmp.data(), sizeof(KeyPoint) * (size_t)n);
  }
  refresh_timing();
}

Mat AKAZE::Feature_Detection_() {
  std::vector<KeyPoint> k;
  Feature_Detection(k);
  Mat m((int)k.size(), 7, CV_32F);
  for (size_t i = 0; i < k.size(); i++) {
    float* d = m.ptr<float>((int)i);  // mat2kpvec order, AKAZE.cpp:290-298
    d[0] = k[i].pt.x; d[1] = k[i].pt.y; d[2] = k[i].size; d[3] = k[i].angle;
    d[4] = k[i].response; d[5] = (float)k[i].octave; d[6] = (float)k[i].class_id;
  }
  return m;
}

void AKAZE::Compute_Descriptors(std::vector<KeyPoint>& kpts, Mat& desc) {
  int cap = nump_;
  kpts.resize(cap);
  int nbytes = 61;
  check(akz_descriptor_bytes(ctx_, &nbytes, nullptr), "AKAZE::Compute_Descriptors");
  desc = Mat(cap, nbytes, CV_8U);  // ceil(bits / 8) columns, AKAZE.cpp:412-423
  int n = 0;
  check(akz_describe(ctx_, reinterpret_cast<akz_keypoint*>(kpts.data()), desc.data, cap, &n),
        "AKAZE::Compute_Descriptors");
  kpts.resize(n);
  desc.rows = n;
  refresh_timing();
}

void AKAZE::refresh_timing() {
  akz_timing t;
  if (akz_get_timing(ctx_, &t) == AKZ_OK) {
    timing_.kcontrast = t.kcontrast;
    timing_.scale = t.scale;
    timing_.derivatives = t.derivatives;
    timing_.detector = t.detector;
    timing_.extrema = t.extrema;
    timing_.subpixel = t.subpixel;
    timing_.descriptor = t.descriptor;
  }
}

void AKAZE::Show_Computation_Times() const {
  // AKAZE.cpp:486-495
  std::cout << "(*) Time Scale Space: " << timing_.scale << std::endl;
  std::cout << "   - Time KContrast: " << timing_.kcontrast << std::endl;
  std::cout << "(*) Time Detector: " << timing_.detector << std::endl;
  std::cout << "   - Time Derivatives: " << timing_.derivatives << std::endl;
  std::cout << "   - Time Extrema: " << timing_.extrema << std::endl;
  std::cout << "   - Time Subpixel: " << timing_.subpixel << std::endl;
  std::cout << "(*) Time Descriptor: " << timing_.descriptor << std::endl;
  std::cout << std::endl;
}

Mat AKAZE::Read_Level(int level, int which) {
  int w = 0, h = 0;
  check(akz_level_size(ctx_, level, &w, &h, nullptr, nullptr, nullptr), "AKAZE::Read_Level");
  Mat m(h, w, CV_32F);
  check(akz_read_level(ctx_, level, which, reinterpret_cast<float*>(m.data)), "AKAZE::Read_Level");
  return m;
}

}  // namespace libAKAZEMI
