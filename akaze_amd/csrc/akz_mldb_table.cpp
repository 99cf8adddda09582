// M-LDB comparison lists of the descriptor variants (host code).
//
// Every variant compares cell sums of the same 29 cells (2x2, 3x3 and 4x4
// grids over the sampling pattern) and three value slots per cell (3 * cell +
// channel); only the list of compared pairs differs:
//   full length (descriptor_size 0): every cell pair j < i of each grid, for
//     channels 0 .. nch-1 in turn (InitCompareIndices, cuda_akaze.cu:2034-2111,
//     with three channels: 486 bits);
//   subsample (descriptor_size > 0): the pairs generateDescriptorSubsample
//     (AKAZE.cpp:498-603) picks from the full list with glibc's rand() after
//     srand(1024), the first six picks forced to the 2x2 grid, one comparison
//     per channel and pick.  The generator state is private (random_r on a
//     TYPE_3 table, the one rand() uses), so callers' rand() sequences are
//     untouched.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "akz_kernels.h"

namespace {

constexpr int kLo[3] = {0, 4, 13}, kHi[3] = {4, 13, 29};
constexpr int kPairs = 162;  // 6 + 36 + 120

struct Pick {
  int grid, x1, y1, x2, y2;  // pattern offsets of the two cells
};

}  // namespace

int akz_mldb_table(int descriptor_size, int patsize, int nch, int16_t* a, int16_t* b) {
  if (nch < 1 || nch > 3 || patsize < 1 || descriptor_size < 0 || descriptor_size > kPairs * nch) return -1;
  int n = 0;
  if (descriptor_size == 0) {
    for (int g = 0; g < 3; g++)
      for (int ch = 0; ch < nch; ch++)
        for (int j = kLo[g]; j < kHi[g] - 1; j++)
          for (int i = j + 1; i < kHi[g]; i++) {
            a[n] = (int16_t)(3 * j + ch);
            b[n] = (int16_t)(3 * i + ch);
            n++;
          }
  } else {
    Pick pool[kPairs];
    int c = 0;
    for (int g = 0; g < 3; g++) {
      const int gdiv = g + 2, psz = (int)std::ceil(2.0 * patsize / (float)gdiv);
      for (int j = 0; j < gdiv * gdiv; j++)
        for (int k = j + 1; k < gdiv * gdiv; k++)
          pool[c++] = Pick{g, psz * (j % gdiv) - patsize, psz * (j / gdiv) - patsize, psz * (k % gdiv) - patsize,
                           psz * (k / gdiv) - patsize};
    }
    // distinct cells met so far (grid, x, y) -> slot, and each slot's cell
    int sg[29], sx[29], sy[29], scell[29], ns = 0;
    random_data rd;
    char state[128];
    std::memset(&rd, 0, sizeof rd);
    initstate_r(1024, state, sizeof state, &rd);
    const int npicks = (descriptor_size + nch - 1) / nch;
    for (int i = 0; i < npicks; i++) {
      int32_t r = 0;
      random_r(&rd, &r);
      int k = (int)((size_t)r % (size_t)(kPairs - i));
      if (i < 6) k = i;
      int slot[2];
      for (int end = 0; end < 2; end++) {
        const int x = end ? pool[k].x2 : pool[k].x1, y = end ? pool[k].y2 : pool[k].y1;
        int f = -1;
        for (int j = 0; j < ns; j++)
          if (sg[j] == pool[k].grid && sx[j] == x && sy[j] == y) { f = j; break; }
        if (f < 0) {
          const int gdiv = pool[k].grid + 2, psz = (int)std::ceil(2.0 * patsize / (float)gdiv);
          sg[ns] = pool[k].grid; sx[ns] = x; sy[ns] = y;
          scell[ns] = kLo[pool[k].grid] + ((y + patsize) / psz) * gdiv + (x + patsize) / psz;
          f = ns++;
        }
        slot[end] = f;
      }
      for (int ch = 0; ch < nch && n < descriptor_size; ch++) {
        a[n] = (int16_t)(3 * scell[slot[0]] + ch);
        b[n] = (int16_t)(3 * scell[slot[1]] + ch);
        n++;
      }
      pool[k] = pool[kPairs - i - 1];
    }
  }
  for (int i = n; i < AKZ_MLDB_TABLE; i++) a[i] = b[i] = 0;  // (0, 0) compares false
  return n;
}
