// Command-line options shared by the programs: the flags of the reference's
// parse_input_options (src/akaze_features.cpp:189-330,
// src/akaze_match.cpp:200-330) and show_input_options_help
// (src/lib/utils.cpp:390-470), with the same clamping of out-of-range
// descriptor values.  --show_results is accepted and ignored: there is no GUI.
#pragma once

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "../../include/akaze/AKAZE.h"

namespace akz_cli {

inline void help(int example) {
  std::cout << "A-KAZE Features (MI355X)\nUsage: ";
  if (example == 0)
    std::cout << "akaze_features img.pgm [options]\n";
  else
    std::cout << "akaze_match img1.pgm img2.pgm [homography.txt] [options]\n";
  std::cout << "\n  homography.txt is optional; without it a homography is estimated by RANSAC.\n\n"
               "  --help                 Show the command line options\n"
               "  --verbose              Print the options\n"
               "  --soffset S            Base scale offset (sigma units)\n"
               "  --omax N               Maximum octave of the image evolution\n"
               "  --nsublevels N         Sublevels per octave\n"
               "  --diffusivity D        0 PM g1, 1 PM g2, 2 Weickert, 3 Charbonnier\n"
               "  --dthreshold T         Detector response threshold (0.001)\n"
               "  --sderivatives S       Derivative smoothing (kept for compatibility)\n"
               "  --descriptor D         4 M-LDB upright, 5 M-LDB\n"
               "  --descriptor_channels C  1 intensity, 2 + gradient magnitude, 3 + x and y gradients\n"
               "  --descriptor_size B    M-LDB bits, 0 = full length (486)\n"
               "  --save_scale_space B   Accepted for compatibility\n"
               "  --show_results B       Accepted for compatibility (no GUI)\n"
               "  --output FILE          Keypoint file (akaze_features; default ./keypoints.txt)\n";
}

// Parses argv[first..) into options; 0 on success, -1 after an error message.
inline int parse(libAKAZEMI::AKAZEOptions& o, std::string* kpts_path, int first, int argc, char* argv[]) {
  for (int i = first; i < argc; i++) {
    const char* a = argv[i];
    auto next = [&](const char*& v) {
      if (i + 1 >= argc) {
        std::cerr << "Error introducing input options!!" << std::endl;
        return false;
      }
      v = argv[++i];
      return true;
    };
    const char* v = nullptr;
    if (!strcmp(a, "--verbose")) {
      o.verbosity = true;
    } else if (!strcmp(a, "--soffset")) {
      if (!next(v)) return -1;
      o.soffset = (float)atof(v);
    } else if (!strcmp(a, "--omax")) {
      if (!next(v)) return -1;
      o.omax = (int)atof(v);
    } else if (!strcmp(a, "--dthreshold")) {
      if (!next(v)) return -1;
      o.dthreshold = (float)atof(v);
    } else if (!strcmp(a, "--sderivatives")) {
      if (!next(v)) return -1;
      o.sderivatives = (float)atof(v);
    } else if (!strcmp(a, "--nsublevels")) {
      if (!next(v)) return -1;
      o.nsublevels = atoi(v);
    } else if (!strcmp(a, "--diffusivity")) {
      if (!next(v)) return -1;
      o.diffusivity = libAKAZEMI::DIFFUSIVITY_TYPE(atoi(v));
    } else if (!strcmp(a, "--descriptor")) {
      if (!next(v)) return -1;
      int d = atoi(v);
      o.descriptor = (d < 0 || d > libAKAZEMI::MLDB) ? libAKAZEMI::MLDB : libAKAZEMI::DESCRIPTOR_TYPE(d);
    } else if (!strcmp(a, "--descriptor_channels")) {
      if (!next(v)) return -1;
      int c = atoi(v);
      o.descriptor_channels = (c <= 0 || c > 3) ? 3 : c;
    } else if (!strcmp(a, "--descriptor_size")) {
      if (!next(v)) return -1;
      int s = atoi(v);
      o.descriptor_size = s < 0 ? 0 : s;
    } else if (!strcmp(a, "--save_scale_space")) {
      if (!next(v)) return -1;
      o.save_scale_space = atoi(v) != 0;
    } else if (!strcmp(a, "--show_results")) {
      if (!next(v)) return -1;
    } else if (!strcmp(a, "--output") && kpts_path) {
      if (!next(v)) return -1;
      *kpts_path = v;
    } else if (!strcmp(a, "--help")) {
      return -1;
    }
  }
  return 0;
}

}  // namespace akz_cli
