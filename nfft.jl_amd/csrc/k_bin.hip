// k_bin.hip -- node binning on the device for nodes! (src/implementation.jl:
// 108-141): the shift into [0, 1) of shiftNodes! (src/utils.jl:32-44) and the
// binning of src/precomputation.jl:459-520 (_precomputeBlocks), here at tile
// granularity:
//   1. k_bin_keys: tile index of every node (FP64 k * Ntilde, truncated, as
//      the reference's floor of the shifted node), a histogram of the tiles;
//   2. a stable radix sort of (tile, caller index) pairs over the key bits:
//      within a tile the nodes keep the caller's order, so the binning is
//      deterministic;
//   3. k_bin_scan: tile offsets and work-item offsets (chunks of at most CH
//      nodes of one tile), one workgroup;
//   4. k_bin_items: the work items and per-tile item ranges;
//   5. k_bin_ks: the shifted coordinates in sorted order.
#include <rocprim/device/device_radix_sort.hpp>

#include "launch.hpp"

namespace nfftm {

namespace {

constexpr int BT = 256;  // threads of the elementwise binning kernels

// shiftNodes! (src/utils.jl:32-44): k < 0 moves up by 1, and a result of
// exactly 1 (a negative k that rounds to 1 - k == 1) is pulled below 1 by
// eps(Float64).
__device__ __forceinline__ double shifted(double x) {
  if (x < 0.0) x += 1.0;
  if (x == 1.0) x -= 2.220446049250313e-16;  // eps(Float64)
  return x;
}

// Tile of node j (k in the caller's layout, D values per node).
__device__ __forceinline__ int node_tile(const BinArgs& a, const double* __restrict__ k, long long j) {
  long long t = 0, str = 1;
  for (int d = 0; d < a.D; ++d) {
    const double kscale = shifted(k[d + (long long)a.D * j]) * (double)a.Nt[d];
    long long c = (long long)kscale;
    if (c >= a.Nt[d]) c -= a.Nt[d];
    t += (c / a.T[d]) * str;
    str *= a.ntiles[d];
  }
  return (int)t;
}

__global__ __launch_bounds__(BT) void k_bin_keys(BinArgs a, const double* __restrict__ k, int* __restrict__ key,
                                                 int* __restrict__ idx, int* __restrict__ cnt) {
  const long long j = (long long)blockIdx.x * BT + threadIdx.x;
  if (j >= a.M) return;
  const int t = node_tile(a, k, j);
  key[j] = t;
  idx[j] = (int)j;
  atomicAdd(&cnt[t], 1);
}

// The same with the histogram of each workgroup's nodes in LDS first (at most
// HIST tiles): crowded tiles (thousands of nodes in one tile) then cost one
// global atomic per workgroup instead of one per node.
constexpr int HIST = 16384;
constexpr int HIST_BLOCKS = 512;
__global__ __launch_bounds__(BT) void k_bin_keys_lds(BinArgs a, const double* __restrict__ k, int* __restrict__ key,
                                                     int* __restrict__ idx, int* __restrict__ cnt) {
  __shared__ int hist[HIST];
  for (int i = threadIdx.x; i < a.ntile; i += BT) hist[i] = 0;
  __syncthreads();
  for (long long j = (long long)blockIdx.x * BT + threadIdx.x; j < a.M; j += (long long)gridDim.x * BT) {
    const int t = node_tile(a, k, j);
    key[j] = t;
    idx[j] = (int)j;
    atomicAdd(&hist[t], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < a.ntile; i += BT)
    if (hist[i]) atomicAdd(&cnt[i], hist[i]);
}

// One workgroup: off = exclusive scan of cnt (ntile + 1 entries), ioff =
// exclusive scan of ceil(cnt / CH) (work items per tile); stats[0] = number
// of work items, stats[1] = most items of one tile.
constexpr int ST = 1024;
__global__ __launch_bounds__(ST) void k_bin_scan(BinArgs a, const int* __restrict__ cnt, int* __restrict__ off,
                                                 int* __restrict__ ioff, int* __restrict__ stats) {
  __shared__ long long s_n[ST], s_i[ST];
  __shared__ int s_max[ST];
  const int n = a.ntile;
  const int per = (n + ST - 1) / ST;
  const int b = threadIdx.x * per, e = min(n, b + per);
  long long sn = 0, si = 0;
  int mx = 0;
  for (int t = b; t < e; ++t) {
    const int c = cnt[t];
    const int it = (c + a.CH - 1) / a.CH;
    sn += c;
    si += it;
    mx = max(mx, it);
  }
  s_n[threadIdx.x] = sn;
  s_i[threadIdx.x] = si;
  s_max[threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {  // ST partial sums: serial exclusive scan
    long long an = 0, ai = 0;
    int am = 0;
    for (int w = 0; w < ST; ++w) {
      const long long tn = s_n[w], ti = s_i[w];
      s_n[w] = an;
      s_i[w] = ai;
      an += tn;
      ai += ti;
      am = max(am, s_max[w]);
    }
    off[n] = (int)an;
    ioff[n] = (int)ai;
    stats[0] = (int)ai;
    stats[1] = am;
  }
  __syncthreads();
  long long an = s_n[threadIdx.x], ai = s_i[threadIdx.x];
  for (int t = b; t < e; ++t) {
    const int c = cnt[t];
    off[t] = (int)an;
    ioff[t] = (int)ai;
    an += c;
    ai += (c + a.CH - 1) / a.CH;
  }
}

// Work items of tile t: {t, begin, end, 0} for chunks of at most CH nodes;
// tstart[t] = first item of tile t (tstart[ntile] = number of items).
__global__ __launch_bounds__(BT) void k_bin_items(BinArgs a, const int* __restrict__ off, const int* __restrict__ ioff,
                                                  int4* __restrict__ items, int* __restrict__ tstart) {
  const long long t = (long long)blockIdx.x * BT + threadIdx.x;
  if (t > a.ntile) return;
  tstart[t] = ioff[t];
  if (t == a.ntile) return;
  const int b0 = off[t], e0 = off[t + 1];
  int q = ioff[t];
  for (int s = b0; s < e0; s += a.CH) items[q++] = int4{(int)t, s, min(e0, s + a.CH), 0};
}

// ks[d * M + s] = shifted k of node perm[s] (perm == nullptr: caller order).
__global__ __launch_bounds__(BT) void k_bin_ks(BinArgs a, const double* __restrict__ k, const int* __restrict__ perm,
                                               double* __restrict__ ks) {
  const long long s = (long long)blockIdx.x * BT + threadIdx.x;
  if (s >= a.M) return;
  const long long j = perm ? perm[s] : s;
  for (int d = 0; d < a.D; ++d) ks[(long long)d * a.M + s] = shifted(k[d + (long long)a.D * j]);
}

__global__ __launch_bounds__(BT) void k_iota(int* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * BT + threadIdx.x;
  if (i < n) p[i] = (int)i;
}

inline unsigned blocks(long long n) { return (unsigned)((n + BT - 1) / BT); }

}  // namespace

size_t bin_sort_temp_bytes(long long M, int key_bits) {
  size_t bytes = 0;
  if (rocprim::radix_sort_pairs(nullptr, bytes, (int*)nullptr, (int*)nullptr, (int*)nullptr, (int*)nullptr,
                                (size_t)M, 0, key_bits) != hipSuccess)
    return 0;
  return std::max<size_t>(bytes, 16);
}

hipError_t launch_bin_keys(const BinArgs& a, const double* k, int* key, int* idx, int* cnt, hipStream_t s) {
  hipError_t e = launch_fill0(cnt, sizeof(int) * ((size_t)a.ntile + 1), s);
  if (e != hipSuccess || a.M == 0) return e;
  if (a.ntile <= HIST) {
    const unsigned nb = (unsigned)std::min<long long>(HIST_BLOCKS, (a.M + BT - 1) / BT);
    hipLaunchKernelGGL(k_bin_keys_lds, dim3(nb), dim3(BT), 0, s, a, k, key, idx, cnt);
  } else {
    hipLaunchKernelGGL(k_bin_keys, dim3(blocks(a.M)), dim3(BT), 0, s, a, k, key, idx, cnt);
  }
  return hipGetLastError();
}

hipError_t launch_bin_sort(const BinArgs& a, void* temp, size_t temp_bytes, int* key_in, int* key_out, int* idx_in,
                           int* idx_out, hipStream_t s) {
  if (a.M == 0) return hipSuccess;
  return rocprim::radix_sort_pairs(temp, temp_bytes, key_in, key_out, idx_in, idx_out, (size_t)a.M, 0, a.key_bits,
                                   s);
}

hipError_t launch_bin_scan(const BinArgs& a, const int* cnt, int* off, int* ioff, int* stats, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(ST), 0, s, a, cnt, off, ioff, stats);
  return hipGetLastError();
}

hipError_t launch_bin_items(const BinArgs& a, const int* off, const int* ioff, int4* items, int* tstart,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_bin_items, dim3(blocks((long long)a.ntile + 1)), dim3(BT), 0, s, a, off, ioff, items, tstart);
  return hipGetLastError();
}

hipError_t launch_bin_ks(const BinArgs& a, const double* k, const int* perm, double* ks, hipStream_t s) {
  if (a.M == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bin_ks, dim3(blocks(a.M)), dim3(BT), 0, s, a, k, perm, ks);
  return hipGetLastError();
}

hipError_t launch_iota(int* p, long long n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3(blocks(n)), dim3(BT), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace nfftm
