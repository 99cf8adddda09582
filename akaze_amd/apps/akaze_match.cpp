// akaze_match without the GUI (src/akaze_match.cpp:53-197): features of two
// PGM images, brute-force Hamming 2-NN on the GPU, NNDR 0.8, and inliers at
// 2.5 px under the ground-truth homography if one is given, else under a
// RANSAC homography; prints the reference's statistics lines.
//
// The reference matches image 1 against image 2 and then image 2 against
// image 1 into the same vector (akaze_match.cpp:134-135), so the second call
// is the one used: queries are image 2's descriptors.  Its point pairs are
// therefore (image-2 point, image-1 point), and the inlier test maps the
// image-2 point through H1toN (akaze_match.cpp:162-167).  The same order is
// kept here so that the statistics compare with the reference's.
#include <chrono>
#include <iostream>

#include "../../include/akaze/AKAZE.h"
#include "../../include/akaze/utils.h"
#include "cli_options.h"

using namespace libAKAZEMI;

static const float MIN_H_ERROR = 2.50f;  // akaze_match.cpp:35
static const float DRATIO = 0.80f;       // akaze_match.cpp:36

int main(int argc, char* argv[]) {
  if (argc < 3 || !strcmp(argv[1], "--help")) {
    akz_cli::help(1);
    return -1;
  }
  AKAZEOptions options;
  std::string img_path1 = argv[1], img_path2 = argv[2], homography_path;
  int first = 3;
  if (argc >= 4 && strncmp(argv[3], "--", 2) != 0) {
    homography_path = argv[3];
    first = 4;
  }
  if (akz_cli::parse(options, nullptr, first, argc, argv)) {
    akz_cli::help(1);
    return -1;
  }
  Mat img1 = imread_pgm(img_path1);
  if (img1.empty()) {
    std::cerr << "Error loading image 1: " << img_path1 << std::endl;
    return -1;
  }
  Mat img2 = imread_pgm(img_path2);
  if (img2.empty()) {
    std::cerr << "Error loading image 2: " << img_path2 << std::endl;
    return -1;
  }
  Mat HG;
  bool use_ransac = homography_path.empty() || !read_homography(homography_path, HG);
  Mat img1_32 = to_float(img1), img2_32 = to_float(img2);
  try {
    options.img_width = img1.cols;
    options.img_height = img1.rows;
    AKAZE evolution1(options);
    options.img_width = img2.cols;
    options.img_height = img2.rows;
    AKAZE evolution2(options);
    std::vector<KeyPoint> kpts1, kpts2;
    Mat desc1, desc2;
    auto t0 = std::chrono::steady_clock::now();
    evolution1.Create_Nonlinear_Scale_Space(img1_32);
    evolution1.Feature_Detection(kpts1);
    evolution1.Compute_Descriptors(kpts1, desc1);
    evolution2.Create_Nonlinear_Scale_Space(img2_32);
    evolution2.Feature_Detection(kpts2);
    evolution2.Compute_Descriptors(kpts2, desc2);
    auto t1 = std::chrono::steady_clock::now();
    Matcher cuda_matcher;
    std::vector<std::vector<DMatch>> dmatches;
    cuda_matcher.bfmatch(desc1, desc2, dmatches);
    cuda_matcher.bfmatch(desc2, desc1, dmatches);
    auto t2 = std::chrono::steady_clock::now();
    std::cout << "#matches: " << dmatches.size() << std::endl;
    std::cout << "#kptsq:   " << kpts1.size() << std::endl;
    std::cout << "#kptst:   " << kpts2.size() << std::endl;
    std::vector<Point2f> matches, inliers;
    matches2points_nndr(kpts2, kpts1, dmatches, matches, DRATIO);
    if (!use_ransac)
      compute_inliers_homography(matches, inliers, HG, MIN_H_ERROR);
    else
      compute_inliers_ransac(matches, inliers, MIN_H_ERROR, false);
    int nmatches = (int)matches.size() / 2, ninliers = (int)inliers.size() / 2;
    float ratio = 100.0f * ((float)ninliers / (float)nmatches);
    std::cout << "Number of Keypoints Image 1: " << kpts1.size() << std::endl;
    std::cout << "Number of Keypoints Image 2: " << kpts2.size() << std::endl;
    std::cout << "A-KAZE Features Extraction Time (ms): "
              << std::chrono::duration<double, std::milli>(t1 - t0).count() << std::endl;
    std::cout << "Matching Descriptors Time (ms): " << std::chrono::duration<double, std::milli>(t2 - t1).count()
              << std::endl;
    std::cout << "Number of Matches: " << nmatches << std::endl;
    std::cout << "Number of Inliers: " << ninliers << std::endl;
    std::cout << "Number of Outliers: " << nmatches - ninliers << std::endl;
    std::cout << "Inliers Ratio: " << ratio << std::endl << std::endl;
  } catch (const Error& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return -1;
  }
  return 0;
}
