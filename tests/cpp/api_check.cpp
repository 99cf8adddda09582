// Drop-in check of the C++ API: the reference's usage pattern
// (src/akaze_features.cpp:80-116, src/akaze_match.cpp:106-138) with only the
// namespace changed.  Prints "ok <keypoints> <matches>" on success.
#include <akaze/AKAZE.h>

#include <cstdio>
#include <vector>

using namespace libAKAZEMI;

int main() {
  const int W = 640, H = 480;
  std::vector<uint8_t> u8(W * H);
  if (akz_synth_image(3, W, H, u8.data()) != AKZ_OK) return 2;
  Mat img(H, W, CV_32F);
  for (int i = 0; i < W * H; i++) reinterpret_cast<float*>(img.data)[i] = u8[i] / 255.0f;
  try {
    AKAZEOptions options;
    options.setWidth(W);
    options.setHeight(H);
    AKAZE evolution(options);
    if (evolution.Create_Nonlinear_Scale_Space(img) != 0) return 3;
    std::vector<KeyPoint> kpts;
    evolution.Feature_Detection(kpts);
    Mat desc;
    evolution.Compute_Descriptors(kpts, desc);
    if ((int)kpts.size() != desc.rows || desc.cols != 61) return 4;
    Matcher matcher;
    std::vector<std::vector<DMatch>> dmatches;
    matcher.bfmatch(desc, desc, dmatches);
    for (size_t i = 0; i < dmatches.size(); i++)
      if (dmatches[i].size() != 2 || dmatches[i][0].distance != 0.0f) return 5;
    evolution.Show_Computation_Times();
    std::printf("ok %zu %zu\n", kpts.size(), dmatches.size());
  } catch (const Error& e) {
    std::printf("error %s\n", e.what());
    return e.status == AKZ_ERR_NO_DEVICE ? 10 : 6;
  }
  return 0;
}
