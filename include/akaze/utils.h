// Evaluation and I/O helpers of the reference programs (src/lib/utils.h,
// src/lib/utils.cpp) over the C ABI, without OpenCV: the same names and
// argument meanings, with libAKAZEMI::Mat / KeyPoint / DMatch / Point2f in
// place of the cv:: types.  Host only.
#pragma once

#include <string>
#include <vector>

#include "AKAZE.h"

namespace libAKAZEMI {

// cv::imread(path, 0) for PGM files: CV_8U, empty on failure.
inline Mat imread_pgm(const std::string& path) {
  int w = 0, h = 0;
  if (akz_read_pgm(path.c_str(), nullptr, 0, &w, &h) != AKZ_OK) return Mat();
  Mat m(h, w, CV_8U);
  if (akz_read_pgm(path.c_str(), m.data, (size_t)w * h, &w, &h) != AKZ_OK) return Mat();
  return m;
}

// img.convertTo(out, CV_32F, 1.0 / 255.0, 0) of the programs (akaze_features.cpp:72)
inline Mat to_float(const Mat& img) {
  Mat out(img.rows, img.cols, CV_32F);
  for (int y = 0; y < img.rows; y++)
    for (int x = 0; x < img.cols; x++) out.ptr<float>(y)[x] = (float)(img.ptr<uint8_t>(y)[x] * (1.0 / 255.0));
  return out;
}

// utils.cpp:97-151; 0 on success, -1 when the file cannot be written
inline int save_keypoints(const std::string& outFile, const std::vector<KeyPoint>& kpts, const Mat& desc,
                          bool save_desc) {
  std::vector<uint8_t> rows((size_t)kpts.size() * (desc.empty() ? 0 : desc.cols));
  for (int i = 0; i < (int)kpts.size() && !desc.empty(); i++)
    std::memcpy(rows.data() + (size_t)i * desc.cols, desc.ptr<uint8_t>(i), desc.cols);
  akz_status st = akz_save_keypoints(outFile.c_str(), reinterpret_cast<const akz_keypoint*>(kpts.data()),
                                     (int)kpts.size(), rows.data(), desc.empty() ? 0 : desc.cols, save_desc ? 1 : 0);
  return st == AKZ_OK ? 0 : -1;
}

// utils.cpp:340-380; false when the file cannot be opened
inline bool read_homography(const std::string& hFile, Mat& H1toN) {
  double h[9];
  if (akz_read_homography(hFile.c_str(), h) != AKZ_OK) return false;
  H1toN.create(3, 3, CV_32F);
  for (int i = 0; i < 9; i++) H1toN.ptr<float>(i / 3)[i % 3] = (float)h[i];
  return true;
}

// utils.cpp:154-170
inline void matches2points_nndr(const std::vector<KeyPoint>& train, const std::vector<KeyPoint>& query,
                                const std::vector<std::vector<DMatch>>& matches, std::vector<Point2f>& pmatches,
                                float nndr) {
  std::vector<DMatch> flat;
  for (const auto& m : matches) {
    flat.push_back(m[0]);
    flat.push_back(m[1]);
  }
  std::vector<float> pairs(4 * matches.size() + 4);
  int n = 0;
  akz_status st = akz_matches2points_nndr(reinterpret_cast<const akz_keypoint*>(train.data()), (int)train.size(),
                                          reinterpret_cast<const akz_keypoint*>(query.data()), (int)query.size(),
                                          reinterpret_cast<const akz_dmatch*>(flat.data()), (int)matches.size(),
                                          nndr, pairs.data(), &n);
  if (st != AKZ_OK) throw Error(st, "matches2points_nndr");
  for (int i = 0; i < n; i++) {
    pmatches.push_back(Point2f{pairs[4 * i], pairs[4 * i + 1]});
    pmatches.push_back(Point2f{pairs[4 * i + 2], pairs[4 * i + 3]});
  }
}

namespace detail {
inline std::vector<float> flatten(const std::vector<Point2f>& m) {
  std::vector<float> p(2 * m.size());
  for (size_t i = 0; i < m.size(); i++) {
    p[2 * i] = m[i].x;
    p[2 * i + 1] = m[i].y;
  }
  return p;
}
inline void keep(const std::vector<Point2f>& m, const std::vector<uint8_t>& mask, std::vector<Point2f>& inliers) {
  for (size_t i = 0; i < mask.size(); i++)
    if (mask[i]) {
      inliers.push_back(m[2 * i]);
      inliers.push_back(m[2 * i + 1]);
    }
}
}  // namespace detail

// utils.cpp:203-243; H is 3x3 CV_32F
inline void compute_inliers_homography(const std::vector<Point2f>& matches, std::vector<Point2f>& inliers,
                                       const Mat& H, float min_error) {
  inliers.clear();
  double h[9];
  for (int i = 0; i < 9; i++) h[i] = H.ptr<float>(i / 3)[i % 3];
  std::vector<float> p = detail::flatten(matches);
  std::vector<uint8_t> mask(matches.size() / 2);
  int n = 0;
  akz_status st = akz_inliers_homography(p.data(), (int)mask.size(), h, min_error, mask.data(), &n);
  if (st != AKZ_OK) throw Error(st, "compute_inliers_homography");
  detail::keep(matches, mask, inliers);
}

// utils.cpp:173-200 (homography only: the fundamental-matrix branch needs
// cv::findFundamentalMat and is not provided; use_fund must be false)
inline void compute_inliers_ransac(const std::vector<Point2f>& matches, std::vector<Point2f>& inliers, float error,
                                   bool use_fund) {
  if (use_fund) throw Error(AKZ_ERR_UNSUPPORTED, "compute_inliers_ransac: fundamental matrix");
  std::vector<float> p = detail::flatten(matches);
  std::vector<uint8_t> mask(matches.size() / 2);
  int n = 0;
  akz_status st = akz_inliers_ransac(p.data(), (int)mask.size(), error, 0x5eed, 2000, 0.995, nullptr, mask.data(), &n);
  if (st != AKZ_OK) throw Error(st, "compute_inliers_ransac");
  detail::keep(matches, mask, inliers);
}

}  // namespace libAKAZEMI
