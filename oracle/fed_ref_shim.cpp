// C entry point over the reference's own FED schedule (src/lib/fed.cpp:41-45),
// compiled from the reference sources in place by oracle/Makefile into
// oracle/_ref/libfed_ref.so.  Test infrastructure only: used to pin the
// oracle's FED restatement and to regenerate tests/golden/fed_tau_4x4.txt.
#include <vector>
#include "fed.h"

extern "C" int fedref_tau_by_process_time(float T, int M, float tau_max, int reordering, float* tau, int cap) {
  std::vector<float> v;
  int n = fed_tau_by_process_time(T, M, tau_max, reordering != 0, v);
  if (n > cap) return -1;
  for (int i = 0; i < n; i++) tau[i] = v[i];
  return n;
}
