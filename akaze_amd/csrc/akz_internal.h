// Internal declarations shared by the host-side translation units.
#pragma once
#include <cstddef>
#include <cstdint>

void akz_synth_render(uint64_t seed, int w, int h, uint8_t* out);
void akz_synth_homography_impl(uint64_t seed, int w, int h, double max_rot_deg, double min_scale, double max_scale,
                          double max_persp, double* H);
int akz_synth_warp_u8(const uint8_t* src, int w, int h, const double* H, double noise_std, uint64_t noise_seed,
                      uint8_t* out);
