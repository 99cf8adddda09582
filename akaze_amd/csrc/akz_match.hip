// Brute-force Hamming 2-NN matcher (MatchDescriptors, cuda_akaze.cu:1842-1967).
//
// One query per lane, held in 16 VGPRs (64-byte rows: 61 descriptor bytes and
// 3 zero bytes).  The wavefront walks a slice of the train set; the train
// row index is uniform, so its 16 dwords arrive through scalar loads and each
// pair costs 16 x (v_xor + v_bcnt).  Best and second best are kept as packed
// keys (distance << 22 | trainIdx), which orders pairs lexicographically by
// (distance, trainIdx) and makes the update three integer min/max ops.  The
// train set is split across blockIdx.y so small query sets still fill the
// chip; k_match_merge folds the per-slice pairs.  Integer work only: no MFMA.
#include "akz_kernels.h"

namespace {

constexpr int MT = 256;                      // queries per block
constexpr uint32_t IDX_BITS = 22;            // trainIdx < 4M
constexpr uint32_t SENTINEL = 512u << IDX_BITS;  // (distance 512, trainIdx 0): the reference's initial pair

__device__ __forceinline__ void top2_insert(uint32_t key, uint32_t& b1, uint32_t& b2) {
  uint32_t mx = b1 > key ? b1 : key;
  b1 = b1 < key ? b1 : key;
  b2 = b2 < mx ? b2 : mx;
}

__global__ __launch_bounds__(MT) void k_match(const uint4* __restrict__ q, int nq, const uint4* __restrict__ t, int nt,
                                              int tchunk, uint32_t* __restrict__ partial) {
  const int qi = blockIdx.x * MT + threadIdx.x;
  const int j0 = blockIdx.y * tchunk;
  const int j1 = akz_imin(j0 + tchunk, nt);
  uint4 qd[4];
  if (qi < nq) {
#pragma unroll
    for (int k = 0; k < 4; k++) qd[k] = q[(size_t)qi * 4 + k];
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) qd[k] = make_uint4(0, 0, 0, 0);
  }
  uint32_t b1 = SENTINEL, b2 = SENTINEL;
  for (int j = j0; j < j1; j++) {
    const uint4* tj = t + (size_t)j * 4;
    uint32_t d = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      uint4 v = tj[k];
      d += __popc(qd[k].x ^ v.x);
      d += __popc(qd[k].y ^ v.y);
      d += __popc(qd[k].z ^ v.z);
      d += __popc(qd[k].w ^ v.w);
    }
    top2_insert((d << IDX_BITS) | (uint32_t)j, b1, b2);
  }
  if (qi < nq) {
    size_t o = ((size_t)blockIdx.y * nq + qi) * 2;
    partial[o] = b1;
    partial[o + 1] = b2;
  }
}

__global__ __launch_bounds__(256) void k_match_merge(const uint32_t* __restrict__ partial, int nq, int nsplit,
                                                     akz_dmatch* __restrict__ out) {
  const int qi = blockIdx.x * 256 + threadIdx.x;
  if (qi >= nq) return;
  uint32_t b1 = SENTINEL, b2 = SENTINEL;
  for (int s = 0; s < nsplit; s++) {
    size_t o = ((size_t)s * nq + qi) * 2;
    top2_insert(partial[o], b1, b2);
    top2_insert(partial[o + 1], b1, b2);
  }
  const uint32_t mask = (1u << IDX_BITS) - 1u;
  akz_dmatch m1, m2;
  m1.queryIdx = qi; m1.trainIdx = (int)(b1 & mask); m1.imgIdx = 0; m1.distance = (float)(b1 >> IDX_BITS);
  m2.queryIdx = qi; m2.trainIdx = (int)(b2 & mask); m2.imgIdx = 0; m2.distance = (float)(b2 >> IDX_BITS);
  out[2 * (size_t)qi] = m1;
  out[2 * (size_t)qi + 1] = m2;
}

}  // namespace

int akz_match_splits(int nq, int nt) {
  int qblocks = (nq + MT - 1) / MT;
  if (qblocks < 1) qblocks = 1;
  int want = (2048 + qblocks - 1) / qblocks;           // aim at ~2048 blocks
  int maxsplit = (nt + 63) / 64;                        // at least 64 train rows per slice
  if (maxsplit < 1) maxsplit = 1;
  return want < maxsplit ? want : maxsplit;
}

void akz_launch_match(hipStream_t s, const uint8_t* q64, int nq, const uint8_t* t64, int nt, uint32_t* partial,
                      int nsplit) {
  int tchunk = (nt + nsplit - 1) / nsplit;
  if (tchunk < 1) tchunk = 1;
  dim3 grid((nq + MT - 1) / MT, nsplit);
  hipLaunchKernelGGL(k_match, grid, dim3(MT), 0, s, reinterpret_cast<const uint4*>(q64), nq,
                     reinterpret_cast<const uint4*>(t64), nt, tchunk, partial);
}

void akz_launch_match_merge(hipStream_t s, const uint32_t* partial, int nq, int nsplit, akz_dmatch* out) {
  hipLaunchKernelGGL(k_match_merge, dim3((nq + 255) / 256), dim3(256), 0, s, partial, nq, nsplit, out);
}

// ---- Allocation-free 2-NN: the per-slice top-2 partials of k_match live in a
// file-scope device array instead of a matcher's grow-only buffer.  Each GPU
// holds its own copy of the array; its address on a device is taken once, with
// that device current, by hipGetSymbolAddress, and only that address is handed
// to kernels (the array's name is a host-side stand-in in host code).  Every
// (slice, query) pair the merge reads is written by k_match of the same call,
// so nothing needs clearing; calls that overlap on one GPU share the array and
// race.
#include <mutex>

namespace {

constexpr size_t NOALLOC_WORDS = (size_t)2 * AKZ_NOALLOC_MAX_QUERIES;  // 8 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

akz_status scratch_on_current_device(uint32_t** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return AKZ_ERR_NO_DEVICE;
  }
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return AKZ_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return AKZ_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return AKZ_OK;
}

}  // namespace

extern "C" akz_status akz_match2nn_device_noalloc(const uint8_t* d_q64, int nq, const uint8_t* d_t64, int nt,
                                                  akz_dmatch* d_out, void* stream) {
  if (nq < 0 || nt < 0 || (nq > 0 && (!d_q64 || !d_out)) || (nt > 0 && !d_t64)) return AKZ_ERR_ARG;
  if (nq > AKZ_NOALLOC_MAX_QUERIES) return AKZ_ERR_ARG;  // the partials would not fit the scratch array
  if (nt >= (1 << IDX_BITS)) return AKZ_ERR_UNSUPPORTED;
  if (nq == 0) return AKZ_OK;
  uint32_t* partial = nullptr;
  akz_status st = scratch_on_current_device(&partial);
  if (st != AKZ_OK) return st;
  // as many train slices as the scratch holds, at most the allocating path's
  // choice; the merged pairs do not depend on the split
  int ns = akz_match_splits(nq, nt);
  int fit = (int)(NOALLOC_WORDS / (2 * (size_t)nq));
  if (ns > fit) ns = fit;
  hipStream_t s = static_cast<hipStream_t>(stream);
  akz_launch_match(s, d_q64, nq, d_t64, nt, partial, ns);
  akz_launch_match_merge(s, partial, nq, ns, d_out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AKZ_OK : AKZ_ERR_HIP;
}
