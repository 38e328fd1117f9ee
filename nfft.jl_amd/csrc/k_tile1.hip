// Tile kernels for D = 1 (see kernels.hpp).
#include "launch.hpp"
#define TILE_D 1
#include "tile_launch.inc"

namespace nfftm {

// D = 1 with the fitted-polynomial window (precompute TENSOR or POLYNOMIAL):
// one thread per node, its 2m window values evaluated in the kernel from the
// node coordinate, as the reference's POLYNOMIAL mode evaluates them inside
// the convolution (src/precomputation.jl:152-165, :203-218;
// src/convolution.jl:220-222) -- the same polynomials TENSOR tabulates per
// node at plan time (src/precomputation.jl:565-585).  No node records: the
// kernels read 12 bytes per node (sorted coordinate and caller index) where a
// record holds the 2m window values, the cell and the index (64 bytes at
// 2m = 12 in FP32), and no lanes idle on stencils narrower than a power of
// two.
//   interpolation: the padded tile staged in LDS, each thread's 2m taps read
//     from it and summed in the data precision (k = 1..2m in order);
//   spreading: the padded tile accumulated in LDS as two FP64 planes (FP64
//     adds of the data-precision products, as the record-based kernel), then
//     written to the item's scratch slot for the fold.
constexpr int NT1 = 256;
constexpr int STAGE1 = 8;  // tile cells each thread loads before writing them

// The 2m window values of the node at shifted coordinate ksv and its wrapped
// cell.  FP64: the
// coefficient pointer is made opaque per node so that the compiler loads the
// (2m + 1) x 2m coefficients into scalar registers tap by tap inside the node
// loop instead of keeping all of them live across it (at 2m = 12 that is 156
// values, more than the scalar register file, and they would be spilled to
// vector lanes).  FP32 up to 2m = 12 keeps half of them (at most 78), which
// fit, live across the loop: a scalar load in the loop would wait (the wait
// counter is shared with LDS operations) for the previous node's LDS atomics
// as well.
template <typename R, int W>
__device__ __forceinline__ int node_taps1(const KArgs& a, double ksv, R* w) {
  const double kscale = scaled_node(ksv, a.Ntd[0]);
  const long long c = (long long)kscale;
  const double t = kscale - (double)(c - a.m + 1);
  using CR = const __attribute__((address_space(4))) R;  // scalar loads
  CR* P = (CR*)a.poly;
  const R x = (R)(t - (double)a.m + 1.0 - 0.5);
  if constexpr (sizeof(R) == 4 && W <= 12) {
    // the window is symmetric and the fit's sample points are too, so tap
    // 2m - 1 - l is tap l's polynomial at -x: the m (2m + 1) coefficients
    // of taps 0 .. m - 1 stay in scalar registers across the node loop
    constexpr int NC = W + 1;
#pragma unroll
    for (int l = 0; l < W / 2; ++l) {
      R s = P[l * NC + NC - 1], u = s;
#pragma unroll
      for (int i = NC - 2; i >= 0; --i) {
        s = fma(x, s, P[l * NC + i]);
        u = fma(-x, u, P[l * NC + i]);
      }
      w[l] = s;
      w[W - 1 - l] = u;
    }
  } else {
    asm volatile("" : "+s"(P));
    poly_taps<R, W, CR*>(P, x, w);
  }
  return (int)(c >= a.Nt[0] ? c - a.Nt[0] : c);
}

// Padded tile of D = 1 at t0 - m + 1 into LDS, STAGE1 loads in flight per
// thread (the padded range lies within one period).
template <typename C>
__device__ __forceinline__ void stage1(const C* __restrict__ grid, C* tile, int g0, int n, int Nx) {
  for (int i0 = 0; i0 < n; i0 += STAGE1 * NT1) {
    C v[STAGE1];
#pragma unroll
    for (int u = 0; u < STAGE1; ++u) {
      const int i = i0 + u * NT1 + (int)threadIdx.x;
      int g = g0 + (i < n ? i : 0);
      g = g < 0 ? g + Nx : (g >= Nx ? g - Nx : g);
      v[u] = grid[g];
    }
#pragma unroll
    for (int u = 0; u < STAGE1; ++u) {
      const int i = i0 + u * NT1 + (int)threadIdx.x;
      if (i < n) tile[i] = v[u];
    }
  }
}

template <typename R, int W>
__global__ __launch_bounds__(NT1) void k_interp1p(KArgs a, const typename cpx<R>::T* __restrict__ grid,
                                                  typename cpx<R>::T* __restrict__ fhat) {
  using C = typename cpx<R>::T;
  extern __shared__ __align__(16) char smem[];
  // consecutive items on consecutive workgroups, i.e. round robin over the
  // XCDs: items of clustered nodes are full and sparse regions' items are
  // not, and contiguous item ranges per XCD left some XCDs with several
  // times the nodes of others
  const int item = blockIdx.x;
  if (item >= a.nitems) return;
  const int4 it = a.items[item];
  C* tile = (C*)smem;
  const int t0 = it.x * a.T[0];
  // the first node's coordinate and index load during the staging, each
  // later node's during the one before
  int j = it.y + (int)threadIdx.x;
  double kn = 0;
  int pn = 0;
  if (j < it.z) {
    kn = a.ks[j];
    if (!a.fsort) pn = a.perm[j];
  }
  stage1(grid, tile, t0 - a.m + 1, a.padCells, a.Nt[0]);
  __syncthreads();
  const R osc = (R)a.oscale;
  for (; j < it.z; j += NT1) {
    const double kj = kn;
    const int pj = pn;
    if (j + NT1 < it.z) {
      kn = a.ks[j + NT1];
      if (!a.fsort) pn = a.perm[j + NT1];
    }
    R w[W];
    const C* b = tile + (node_taps1<R, W>(a, kj, w) - t0);
    R ar = 0, ai = 0;
#pragma unroll
    for (int l = 0; l < W; ++l) {
      const C v = b[l];
      ar = fma(w[l], v.x, ar);
      ai = fma(w[l], v.y, ai);
    }
    if (a.fsort) ((C*)a.fsort)[j] = C{ar * osc, ai * osc};
    else fhat[(long long)pj * a.fst] = C{ar * osc, ai * osc};
  }
}

template <typename R, int W>
__global__ __launch_bounds__(NT1) void k_spread1p(KArgs a, const typename cpx<R>::T* __restrict__ fhat) {
  using C = typename cpx<R>::T;
  extern __shared__ __align__(16) char smem[];
  const int item = blockIdx.x;  // round robin over the XCDs (k_interp1p)
  if (item >= a.nitems) return;
  const int4 it = a.items[item];
  double* acc = (double*)smem;  // re plane [0, te), im plane [te, 2 te)
  const int te = a.tileElems;
  const int t0 = it.x * a.T[0];
  // node coordinate and value one node ahead
  int j = it.y + (int)threadIdx.x;
  double kn = 0;
  C vn{0, 0};
  if (j < it.z) {
    kn = a.ks[j];
    vn = fhat[(long long)a.perm[j] * a.fst];
  }
  for (int i = threadIdx.x; i < 2 * te; i += NT1) acc[i] = 0;
  __syncthreads();
  const R osc = (R)a.oscale;
  for (; j < it.z; j += NT1) {
    const double kj = kn;
    const R fr = vn.x * osc, fi = vn.y * osc;
    if (j + NT1 < it.z) {
      kn = a.ks[j + NT1];
      vn = fhat[(long long)a.perm[j + NT1] * a.fst];
    }
    R w[W];
    const int loc = node_taps1<R, W>(a, kj, w) - t0;
#pragma unroll
    for (int l = 0; l < W; ++l) {
      __hip_atomic_fetch_add(&acc[loc + l], (double)(fr * w[l]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(&acc[te + loc + l], (double)(fi * w[l]), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __syncthreads();
  C* out = (C*)a.scratch + (size_t)item * a.padCells;
  for (int i = threadIdx.x; i < a.padCells; i += NT1) out[i] = C{(R)acc[i], (R)acc[te + i]};
}

template <typename R, int W>
static hipError_t poly1_go(bool spread, const KArgs& a, hipStream_t s, const void* in, void* out) {
  using C = typename cpx<R>::T;
  const int nblk = 8 * ((a.nitems + 7) / 8);
  if (spread) {
    const size_t lds = 2 * (size_t)a.tileElems * sizeof(double);
    auto k = k_spread1p<R, W>;
    if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k, dim3(nblk), dim3(NT1), lds, s, a, (const C*)in);
  } else {
    const size_t lds = (size_t)a.padCells * sizeof(C);
    auto k = k_interp1p<R, W>;
    if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k, dim3(nblk), dim3(NT1), lds, s, a, (const C*)in, (C*)out);
  }
  return hipGetLastError();
}

template <typename R>
static hipError_t poly1_w(bool spread, int W, const KArgs& a, hipStream_t s, const void* in, void* out) {
  switch (W) {
    case 2: return poly1_go<R, 2>(spread, a, s, in, out);
    case 4: return poly1_go<R, 4>(spread, a, s, in, out);
    case 6: return poly1_go<R, 6>(spread, a, s, in, out);
    case 8: return poly1_go<R, 8>(spread, a, s, in, out);
    case 10: return poly1_go<R, 10>(spread, a, s, in, out);
    case 12: return poly1_go<R, 12>(spread, a, s, in, out);
    case 14: return poly1_go<R, 14>(spread, a, s, in, out);
    case 16: return poly1_go<R, 16>(spread, a, s, in, out);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_tile_d1(bool spread, bool dbl, int W, int NT, const KArgs& a, size_t lds,
                          hipStream_t s, const void* in, void* out) {
  if (a.mode == WIN_POLY)
    return dbl ? poly1_w<double>(spread, W, a, s, in, out) : poly1_w<float>(spread, W, a, s, in, out);
  return dbl ? tile_w<double>(spread, W, NT, a, lds, s, in, out)
             : tile_w<float>(spread, W, NT, a, lds, s, in, out);
}
}  // namespace nfftm
