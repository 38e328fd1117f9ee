// k_batch.hip -- persistent gridding kernels for a batch of plans.
//
// A batch runs many independent transforms (plans of different N and M, and
// several images sharing one plan's nodes) as one launch per stage.  Each
// (plan, image) pair is a slot with its own copy of the plan's kernel
// arguments in device memory (its grid, scratch and node arrays); a work list
// enumerates the (slot, work item) pairs of all slots, and a fixed grid of
// workgroups -- as many as fit on the device at once -- strides through it.
// Consecutive work entries (neighbouring tiles of one slot) are dealt to the
// workgroups of one XCD so that they share that XCD's L2.
#include "launch.hpp"

namespace nfftm {

// Rank of this workgroup such that the gridDim.x / 8 workgroups the
// dispatcher places on one XCD (round-robin over the 8 XCDs) hold
// consecutive ranks.  gridDim.x is a multiple of 8.
__device__ __forceinline__ int xcd_rank() {
  const int per = gridDim.x >> 3;
  return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

template <typename R, int D, int W, int NT>
__global__ __launch_bounds__(NT) void k_interp_batch(const KArgs* __restrict__ ka, const int2* __restrict__ work,
                                                     int nwork, const void* const* __restrict__ grids,
                                                     void* const* __restrict__ outs) {
  using C = typename cpx<R>::T;
  for (int w = xcd_rank(); w < nwork; w += gridDim.x) {
    const int2 e = work[w];
    interp_tile_item<R, D, W, NT>(ka[e.x], e.y, (const C*)grids[e.x], (C*)outs[e.x]);
    __syncthreads();  // the next item restages the LDS tile
  }
}

template <typename R, int D, int W, int NT>
__global__ __launch_bounds__(NT) void k_spread_batch(const KArgs* __restrict__ ka, const int2* __restrict__ work,
                                                     int nwork, const void* const* __restrict__ ins) {
  using C = typename cpx<R>::T;
  for (int w = xcd_rank(); w < nwork; w += gridDim.x) {
    const int2 e = work[w];
    const KArgs& a = ka[e.x];
    spread_tile_item<R, D, W, NT>(a, e.y, (const C*)ins[e.x], (C*)a.scratch);
    __syncthreads();
  }
}

// Fold of a batch: work entry {slot, first, end, 0} -- grid rows for D = 2,
// points of the single row for D = 1; the workgroup sums the covering
// scratch slots of every point.
template <typename R, int D>
__global__ __launch_bounds__(256) void k_fold_batch(const KArgs* __restrict__ ka, const int4* __restrict__ work,
                                                    int nwork, void* const* __restrict__ grids) {
  using C = typename cpx<R>::T;
  for (int w = xcd_rank(); w < nwork; w += gridDim.x) {
    const int4 e = work[w];
    const KArgs& a = ka[e.x];
    const C* scr = (const C*)a.scratch;
    C* grid = (C*)grids[e.x];
    const int ny = D >= 2 ? a.Nt[1] : 1;
    int t1[3] = {0, 0, 0}, p1[3] = {0, 0, 0}, t2[3] = {0, 0, 0}, p2[3] = {0, 0, 0};
    int n1 = 1, n2 = 1;
    if (D == 1) {
      for (int x = e.y + threadIdx.x; x < e.z; x += 256)
        grid[x] = fold_sum<R, D>(a, scr, x, n1, t1, p1, n2, t2, p2);
      continue;
    }
    for (int row = e.y; row < e.z; ++row) {
      if (D >= 2) n1 = cover_pairs(a, 1, row % ny, t1, p1);
      if (D == 3) n2 = cover_pairs(a, 2, row / ny, t2, p2);
      for (int x = threadIdx.x; x < a.Nt[0]; x += 256)
        grid[(long long)row * a.Nt[0] + x] = fold_sum<R, D>(a, scr, x, n1, t1, p1, n2, t2, p2);
    }
  }
}

template <typename R, int D, int W, int NT>
static hipError_t batch_go(bool spread, int nblk, size_t lds, hipStream_t s, const KArgs* ka, const int2* work,
                           int nwork, const void* const* in, void* const* out) {
  if (spread) {
    auto k = k_spread_batch<R, D, W, NT>;
    if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k, dim3(nblk), dim3(NT), lds, s, ka, work, nwork, in);
  } else {
    auto k = k_interp_batch<R, D, W, NT>;
    if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k, dim3(nblk), dim3(NT), lds, s, ka, work, nwork, in, out);
  }
  return hipGetLastError();
}

template <typename R, int D, int W>
static hipError_t batch_nt(bool spread, int NT, int nblk, size_t lds, hipStream_t s, const KArgs* ka,
                           const int2* work, int nwork, const void* const* in, void* const* out) {
  if (NT == 256) return batch_go<R, D, W, 256>(spread, nblk, lds, s, ka, work, nwork, in, out);
  if (NT == 512) return batch_go<R, D, W, 512>(spread, nblk, lds, s, ka, work, nwork, in, out);
  return hipErrorInvalidValue;
}

template <typename R, int D>
static hipError_t batch_w(bool spread, int W, int NT, int nblk, size_t lds, hipStream_t s, const KArgs* ka,
                          const int2* work, int nwork, const void* const* in, void* const* out) {
  switch (W) {
#define NFFTM_BW(WW) \
  case WW:           \
    return batch_nt<R, D, WW>(spread, NT, nblk, lds, s, ka, work, nwork, in, out);
    NFFTM_BW(2) NFFTM_BW(4) NFFTM_BW(6) NFFTM_BW(8) NFFTM_BW(10) NFFTM_BW(12) NFFTM_BW(14) NFFTM_BW(16)
#undef NFFTM_BW
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_tile_batch(int D, bool spread, bool dbl, int W, int NT, int nblk, size_t lds, hipStream_t s,
                             const KArgs* ka, const int2* work, int nwork, const void* const* in,
                             void* const* out) {
  if (nwork == 0) return hipSuccess;
  if (D == 1)
    return dbl ? batch_w<double, 1>(spread, W, NT, nblk, lds, s, ka, work, nwork, in, out)
               : batch_w<float, 1>(spread, W, NT, nblk, lds, s, ka, work, nwork, in, out);
  if (D == 2)
    return dbl ? batch_w<double, 2>(spread, W, NT, nblk, lds, s, ka, work, nwork, in, out)
               : batch_w<float, 2>(spread, W, NT, nblk, lds, s, ka, work, nwork, in, out);
  return hipErrorInvalidValue;
}

hipError_t launch_fold_batch(int D, bool dbl, int nblk, hipStream_t s, const KArgs* ka, const int4* work,
                             int nwork, void* const* grids) {
  if (nwork == 0) return hipSuccess;
  if (D == 1) {
    if (dbl) hipLaunchKernelGGL((k_fold_batch<double, 1>), dim3(nblk), dim3(256), 0, s, ka, work, nwork, grids);
    else hipLaunchKernelGGL((k_fold_batch<float, 1>), dim3(nblk), dim3(256), 0, s, ka, work, nwork, grids);
  } else if (D == 2) {
    if (dbl) hipLaunchKernelGGL((k_fold_batch<double, 2>), dim3(nblk), dim3(256), 0, s, ka, work, nwork, grids);
    else hipLaunchKernelGGL((k_fold_batch<float, 2>), dim3(nblk), dim3(256), 0, s, ka, work, nwork, grids);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Resident workgroups of the batch tile kernel (occupancy x CUs), a multiple of 8.
int tile_batch_blocks(int D, bool spread, bool dbl, int W, int NT, size_t lds) {
  int dev = 0, cus = 0, per = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const void* k = nullptr;
#define NFFTM_KPTR(RR, DD, WW, NN) \
  k = spread ? (const void*)k_spread_batch<RR, DD, WW, NN> : (const void*)k_interp_batch<RR, DD, WW, NN>;
#define NFFTM_KW(RR, DD, NN)                          \
  switch (W) {                                        \
    case 2: NFFTM_KPTR(RR, DD, 2, NN) break;          \
    case 4: NFFTM_KPTR(RR, DD, 4, NN) break;          \
    case 6: NFFTM_KPTR(RR, DD, 6, NN) break;          \
    case 8: NFFTM_KPTR(RR, DD, 8, NN) break;          \
    case 10: NFFTM_KPTR(RR, DD, 10, NN) break;        \
    case 12: NFFTM_KPTR(RR, DD, 12, NN) break;        \
    case 14: NFFTM_KPTR(RR, DD, 14, NN) break;        \
    case 16: NFFTM_KPTR(RR, DD, 16, NN) break;        \
  }
  if (D == 1 && dbl && NT == 256) { NFFTM_KW(double, 1, 256) }
  else if (D == 1 && dbl) { NFFTM_KW(double, 1, 512) }
  else if (D == 1 && NT == 256) { NFFTM_KW(float, 1, 256) }
  else if (D == 1) { NFFTM_KW(float, 1, 512) }
  else if (dbl && NT == 256) { NFFTM_KW(double, 2, 256) }
  else if (dbl) { NFFTM_KW(double, 2, 512) }
  else if (NT == 256) { NFFTM_KW(float, 2, 256) }
  else { NFFTM_KW(float, 2, 512) }
#undef NFFTM_KW
#undef NFFTM_KPTR
  if (k) {
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, NT, lds);
  }
  per = std::max(per, 1);
  cus = std::max(cus, 8);
  return ((per * cus + 7) / 8) * 8;
}

}  // namespace nfftm
