// akaze_features without the GUI (src/akaze_features.cpp:44-187): reads a
// PGM image, builds the nonlinear scale space, detects keypoints, computes
// M-LDB descriptors on the GPU, prints the timings and the keypoint count,
// and saves keypoints with descriptors in Mikolajczyk's format.
#include <chrono>
#include <iostream>

#include "../../include/akaze/AKAZE.h"
#include "../../include/akaze/utils.h"
#include "cli_options.h"

using namespace libAKAZEMI;

int main(int argc, char* argv[]) {
  if (argc < 2 || !strcmp(argv[1], "--help")) {
    akz_cli::help(0);
    return -1;
  }
  AKAZEOptions options;
  std::string img_path = argv[1], kpts_path = "./keypoints.txt";
  if (akz_cli::parse(options, &kpts_path, 2, argc, argv)) {
    akz_cli::help(0);
    return -1;
  }
  if (options.verbosity) std::cout << "Check AKAZE options:\n" << options << std::endl;
  Mat img = imread_pgm(img_path);
  if (img.empty()) {
    std::cerr << "Error: cannot load image from file:" << std::endl << img_path << std::endl;
    return -1;
  }
  Mat img_32 = to_float(img);
  options.img_width = img.cols;
  options.img_height = img.rows;
  try {
    AKAZE evolution(options);
    std::vector<KeyPoint> kpts;
    auto t0 = std::chrono::steady_clock::now();
    evolution.Create_Nonlinear_Scale_Space(img_32);
    evolution.Feature_Detection(kpts);
    auto t1 = std::chrono::steady_clock::now();
    Mat desc;
    evolution.Compute_Descriptors(kpts, desc);
    auto t2 = std::chrono::steady_clock::now();
    double tdet = std::chrono::duration<double, std::milli>(t1 - t0).count();
    double tdesc = std::chrono::duration<double, std::milli>(t2 - t1).count();
    evolution.Show_Computation_Times();
    std::cout << "Number of points: " << kpts.size() << std::endl;
    std::cout << "Time Detector: " << tdet << " ms" << std::endl;
    std::cout << "Time Descriptor: " << tdesc << " ms" << std::endl;
    if (!kpts_path.empty() && save_keypoints(kpts_path, kpts, desc, true) != 0) {
      std::cerr << "Couldn't open file '" << kpts_path << "'!" << std::endl;
      return -1;
    }
  } catch (const Error& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return -1;
  }
  return 0;
}
