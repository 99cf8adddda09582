// libAKAZE-compatible C++ API over the MI355X C-ABI (include/akaze_mi.h).
//
// Mirrors libAKAZECU::AKAZE and libAKAZECU::Matcher (src/lib/AKAZE.h:29-143 of
// the reference) so that a caller changes only the namespace.  OpenCV is not
// required: KeyPoint, DMatch and Mat are plain types with the same memory
// layout as cv::KeyPoint (28 B), cv::DMatch (16 B) and the CV_8U / CV_32F
// cv::Mat cases the reference uses.  Errors throw libAKAZEMI::Error instead of
// exiting the process.
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../akaze_mi.h"

namespace libAKAZEMI {

// AKAZEConfig.h:41-57
enum DESCRIPTOR_TYPE { SURF_UPRIGHT = 0, SURF = 1, MSURF_UPRIGHT = 2, MSURF = 3, MLDB_UPRIGHT = 4, MLDB = 5 };
enum DIFFUSIVITY_TYPE { PM_G1 = 0, PM_G2 = 1, WEICKERT = 2, CHARBONNIER = 3 };

// AKAZETiming, AKAZEConfig.h:61-80
struct AKAZETiming {
  double kcontrast = 0, scale = 0, derivatives = 0, detector = 0, extrema = 0, subpixel = 0, descriptor = 0;
};

// AKAZEOptions, AKAZEConfig.h:84-143 (defaults of :86-111)
struct AKAZEOptions {
  int omin = 0;
  int omax = 4;
  int nsublevels = 4;
  int img_width = 0;
  int img_height = 0;
  float soffset = 1.6f;
  float derivative_factor = 1.5f;
  float sderivatives = 1.0f;
  DIFFUSIVITY_TYPE diffusivity = PM_G2;
  float dthreshold = 0.001f;
  float min_dthreshold = 0.00001f;
  DESCRIPTOR_TYPE descriptor = MLDB;
  int descriptor_size = 0;
  int descriptor_channels = 3;
  int descriptor_pattern_size = 10;
  float kcontrast = 0.001f;
  float kcontrast_percentile = 0.7f;
  size_t kcontrast_nbins = 300;
  bool save_scale_space = false;
  bool save_keypoints = false;
  bool verbosity = false;
  int ncudaimages = 4;
  int maxkeypoints = 0;  // 0: room for every candidate (the reference caps at 16*8192)
  void setWidth(int w) { img_width = w; }
  void setHeight(int h) { img_height = h; }
  friend std::ostream& operator<<(std::ostream& os, const AKAZEOptions& o);
};

struct Point2f {
  float x = 0, y = 0;
};

// cv::KeyPoint layout
struct KeyPoint {
  Point2f pt;
  float size = 0, angle = -1, response = 0;
  int octave = 0, class_id = -1;
  KeyPoint() = default;
  KeyPoint(float x, float y, float size_, float angle_ = -1, float response_ = 0, int octave_ = 0, int class_id_ = -1)
      : size(size_), angle(angle_), response(response_), octave(octave_), class_id(class_id_) {
    pt.x = x;
    pt.y = y;
  }
};
static_assert(sizeof(KeyPoint) == 28, "cv::KeyPoint layout");

// cv::DMatch layout
struct DMatch {
  int queryIdx = -1, trainIdx = -1, imgIdx = -1;
  float distance = 0;
};
static_assert(sizeof(DMatch) == 16, "cv::DMatch layout");

// Minimal owning matrix for the CV_8U (descriptors) and CV_32F (images,
// bfmatch_ output) cases.
enum { CV_8U = 0, CV_32F = 5 };
struct Mat {
  int rows = 0, cols = 0, type = CV_8U;
  size_t step = 0;  // bytes per row
  uint8_t* data = nullptr;
  Mat() = default;
  Mat(int r, int c, int t) { create(r, c, t); }
  Mat(int r, int c, int t, void* ext, size_t step_ = 0)
      : rows(r), cols(c), type(t), step(step_ ? step_ : (size_t)c * elem(t)), data((uint8_t*)ext) {}
  void create(int r, int c, int t) {
    rows = r; cols = c; type = t;
    step = (size_t)c * elem(t);
    own_ = std::make_shared<std::vector<uint8_t>>((size_t)r * step);
    data = own_->data();
  }
  static size_t elem(int t) { return t == CV_32F ? 4 : 1; }
  bool empty() const { return rows == 0 || cols == 0 || !data; }
  template <class T> T* ptr(int r) { return reinterpret_cast<T*>(data + (size_t)r * step); }
  template <class T> const T* ptr(int r) const { return reinterpret_cast<const T*>(data + (size_t)r * step); }

 private:
  std::shared_ptr<std::vector<uint8_t>> own_;
};

struct Error : std::runtime_error {
  akz_status status;
  Error(akz_status s, const std::string& where)
      : std::runtime_error(where + ": " + akz_status_string(s)), status(s) {}
};

// libAKAZECU::Matcher (AKAZE.h:29-55): brute-force Hamming 2-NN.
class Matcher {
 public:
  Matcher();
  ~Matcher();
  Matcher(const Matcher&) = delete;
  Matcher& operator=(const Matcher&) = delete;
  // two closest train rows per query row, ordered by (distance, trainIdx)
  void bfmatch(Mat& desc_query, Mat& desc_train, std::vector<std::vector<DMatch>>& dmatches);
  // n_query x 8 CV_32F: [q, t1, 0, d1, q, t2, 0, d2] per row (AKAZE.cpp:82-105)
  Mat bfmatch_(Mat desc_query, Mat desc_train);

 private:
  akz_matcher* m_ = nullptr;
};

// Free function of cuda_akaze.h:28.
void MatchDescriptors(Mat& desc_query, Mat& desc_train, std::vector<std::vector<DMatch>>& dmatches);

// libAKAZECU::AKAZE (AKAZE.h:58-143).
class AKAZE {
 public:
  explicit AKAZE(const AKAZEOptions& options);
  ~AKAZE();
  AKAZE(const AKAZE&) = delete;
  AKAZE& operator=(const AKAZE&) = delete;
  void Allocate_Memory_Evolution() {}  // done by the constructor
  // img: CV_32F, img_height x img_width, values in [0,1]; 0 on success, -1 otherwise
  int Create_Nonlinear_Scale_Space(const Mat& img);
  // fills kpts with the kept keypoints (angle not yet computed); the
  // reference leaves kpts empty here (AKAZE.cpp:372), a superset is harmless
  void Feature_Detection(std::vector<KeyPoint>& kpts);
  Mat Feature_Detection_();
  // orientation + 486-bit MLDB; kpts is overwritten with the detector's
  // keypoints, as the reference does (AKAZE.cpp:430-434)
  void Compute_Descriptors(std::vector<KeyPoint>& kpts, Mat& desc);
  void Show_Computation_Times() const;
  AKAZETiming Get_Computation_Times() const { return timing_; }
  // debug readback of a level (which: 0 Lt, 1 Lsmooth, 2 Lx, 3 Ly, 4 Ldet)
  Mat Read_Level(int level, int which);
  int Num_Levels() const;

 private:
  void refresh_timing();
  AKAZEOptions options_;
  akz_ctx* ctx_ = nullptr;
  AKAZETiming timing_;
  int nump_ = 0;
};

// fRound, AKAZE.h:175-177
inline int fRound(float flt) { return (int)(flt + 0.5f); }

}  // namespace libAKAZEMI
