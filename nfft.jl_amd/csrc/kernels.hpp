// kernels.hpp -- device code of the gfx950 NFFT gridding path.
//
// Spreading (convolve_transpose!, src/convolution.jl:115-149) and
// interpolation (convolve!, src/convolution.jl:20-45) over nodes that the host
// has bin-sorted into tiles of the oversampled grid.  One workgroup owns one
// work item = (tile, contiguous range of that tile's nodes).  The tile plus
// its stencil halo ("padded tile", extent T + 2m - 1 per dimension) lives in
// LDS:
//   * spreading accumulates every node's (2m)^D footprint into the padded tile
//     with LDS atomics, then writes the padded tile with plain coalesced
//     stores to a per-item scratch slot; a fold kernel then sums the (at most
//     2^D, or 3^D for narrow tiles) padded tiles covering each grid point and
//     writes every grid point once.  FP32 accumulates packed 32+32-bit
//     fixed-point (re, im) pairs with one 64-bit integer LDS add per tap (on
//     gfx950 the FP32 LDS float add costs ~40x an integer add); the scale is a
//     power of two per work item from the strict bound sum_j |f_j| max(w)^D,
//     so sums are exact integers, independent of order.  FP64 accumulates
//     with the FP64 LDS add.
//   * interpolation stages the padded tile from HBM once, then gathers every
//     node's footprint from LDS.
// Per node and dimension the FP64 index reduction of the reference
// (src/precomputation.jl:524-554) gives the cell and the window argument; the
// 2m window values come from the LINEAR table (src/precomputation.jl:179-201),
// the polynomial of TENSOR/POLYNOMIAL (src/precomputation.jl:215-222,
// 556-591) or the exact Kaiser-Bessel window of FULL
// (src/windowFunctions.jl:21-34).
//
// Lane layout (64-wide wavefronts):
//   D = 3: one node per wavefront step, lane = (lx, ly) of the footprint, each
//          lane walks lz.  With row stride Sx = 8 (mod 16) the 64 lanes of a
//          W = 8 footprint hit 64 distinct LDS banks.
//   D = 2: Wp = pow2ceil(2m) lanes per node (lx), 64/Wp nodes per step, each
//          lane walks ly.
//   D = 1: Wp lanes per node, one tap per lane.
// Interpolation reduces per-lane partial sums with DPP row operations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace nfftm {

constexpr int MAXD = 8;
enum { WIN_LINEAR = 0, WIN_POLY = 1, WIN_EXACT = 2 };

template <typename R>
struct cpx;
template <>
struct cpx<float> {
  using T = float2;
};
template <>
struct cpx<double> {
  using T = double2;
};

// Division of non-negative 32-bit integers (< 2^31) by a fixed divisor via a
// multiply-high (Granlund & Montgomery 1994, fig. 4.1).
struct FastDiv {
  uint32_t d, mul, shift;
};

__host__ inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  f.shift = l;
  f.mul = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
  return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t x, const FastDiv& f) {
  return (__umulhi(x, f.mul) + x) >> f.shift;
}

struct KArgs {
  int D, m, W, mode;
  int M;                 // number of nodes
  long long fst;         // element stride of the node values (directional plans: the product of the
                         // sizes of the axes before the transformed ones; otherwise 1)
  int Nt[MAXD];          // oversampled grid size per dim
  double Ntd[MAXD];
  long long gstr[MAXD];  // grid strides (dim 1 fastest)
  // tiling (D <= 3)
  int T[3];        // tile extent per dim
  int ntiles[3];   // tiles per dim
  int Pd[3];       // padded extent T + W - 1
  int Sx, Sxy;     // LDS strides of the padded tile (elements)
  int tileElems;   // elements of one LDS plane
  int padCells;    // Pd[0]*Pd[1]*Pd[2]
  FastDiv divP0, divP1, divNtx;
  // window
  const void* lut;  // R[LUTSize + 2]
  long long scale;  // LUTSize / m, a power of two
  const void* poly; // R[(2m+1) * 2m], column l holds coefficients of tap l
  const double* polyd;  // the same coefficients in FP64 (node gradient)
  double b;         // Kaiser-Bessel shape pi*(2 - 1/sigma)
  int win;          // window id (NFFTM_KAISER_BESSEL .. NFFTM_COSH_TYPE) of the exact (FULL) mode
  double wscale;    // power-of-two window normalisation
  double oscale;    // output scale: 1 inside a transform; wscale^-D for the
                    // standalone convolve!/convolve_transpose! entry points
  // nodes
  const double* ks;   // ks[d * M + j], shifted into [0, 1)
  const int* perm;    // sorted position -> caller's node index
  void* fsort;        // interpolation results in sorted order (D = 1 polynomial
                      // kernels; null: stored at the caller's positions)
  const int4* items;  // {tile, begin, end, 0}
  int nitems;
  const char* rec;    // node records in sorted order (NodeRec layout), tiled plans
  int recBytes;
  int blocked3;       // 3-D scratch slots stored as 27 halo/interior blocks
  int tvar;           // 3-D tile geometry variant (compile-time tile edge)
  int fix2;           // FP32 2-D spreading accumulates packed fixed point (one
                      // 64-bit LDS atomic per tap) instead of two FP64 planes
  // spreading
  double wmaxD;        // bound of any product of D window values
  void* scratch;       // nitems padded tiles (dense, px fastest)
  const int* tstart;   // items of tile t: [tstart[t], tstart[t+1])
  int ntile_all;       // product of ntiles
  int maxcand;         // covering (tile, offset) pairs per dim (fold)
};

// k * Ntilde rounded to FP64 before any further use, as the reference
// computes kscale (src/precomputation.jl:538): kept out of a fused multiply-add
// with the following subtraction, which at Ntilde ~ 2e7 would move the window
// argument by ~1e-9 grid units.
__device__ __forceinline__ double scaled_node(double ksv, double Ntd) {
#pragma clang fp contract(off)
  return ksv * Ntd;
}

// Window values of one node along dimension d.  Returns the wrapped cell
// floor(k * Ntilde) (in [0, Ntilde)); t = kscale - off with
// off = cell - m + 1 exactly as src/precomputation.jl:537-545.
// I0 by its positive power series, to the precision of R.
template <typename R>
__device__ __forceinline__ R i0_dev(R x) {
  const R y = (R)0.25 * x * x;
  R t = 1, s = 1;
  for (int k = 1; k < 200 && t > (sizeof(R) == 4 ? (R)1e-8 : (R)1e-17) * s; ++k) {
    t = t * y / (R)(k * k);
    s += t;
  }
  return s;
}

// Exact window psi(u), u in grid units, of window a.win (src/windowFunctions.jl:
// 21-134): Kaiser-Bessel, Gaussian, cardinal B-spline of order W = 2m
// (Cox-de Boor), reversed Kaiser-Bessel, cosh type.
template <typename R, int W>
__device__ __forceinline__ R window_exact(const KArgs& a, R u) {
  const R m = (R)a.m, b = (R)a.b, au = u < 0 ? -u : u;
  switch (a.win) {
    case 1: {  // gauss, b_g = m / pi
      const R bg = m / (R)M_PI;
      return au < m ? exp(-u * u / bg) / sqrt((R)M_PI * bg) : (R)0;
    }
    case 2: {  // spline: M_W(u + m)
      if (!(au < m)) return 0;
      const R x = u + m;
      R v[W];
#pragma unroll
      for (int s = 0; s < W; ++s) v[s] = (x - (R)s >= 0 && x - (R)s < 1) ? (R)1 : (R)0;
#pragma unroll
      for (int k = 2; k <= W; ++k) {
#pragma unroll
        for (int s = 0; s + k <= W; ++s) v[s] = ((x - (R)s) * v[s] + ((R)k - (x - (R)s)) * v[s + 1]) / (R)(k - 1);
      }
      return v[0];
    }
    case 3: {  // kaiser_bessel_rev
      if (!(au < m)) return 0;
      const R r = u / m;
      return (R)0.5 / m * i0_dev<R>(m * b * sqrt((R)1 - r * r));
    }
    case 4: {  // cosh_type, beta = m b
      if (!(au < m)) return 0;
      const R r = u / m, al = sqrt((R)1 - r * r), beta = m * b;
      return (cosh(beta * al) - (R)1) / ((cosh(beta) - (R)1) * al);
    }
    default: {  // kaiser_bessel
      if (au < m) {
        const R arg = sqrt(m * m - u * u);
        return sinh(b * arg) / (arg * (R)M_PI);
      }
      return au > m ? (R)0 : b / (R)M_PI;
    }
  }
}

// The 2m fitted window polynomials at x = t - m + 1/2 (evalpoly per tap,
// src/precomputation.jl:203-218): column l of P holds the 2m + 1 coefficients
// of tap l, lowest degree first; Horner in the data precision.
template <typename R, int W, typename PT = const R*>
__device__ __forceinline__ void poly_taps(PT P, R x, R* w) {
  constexpr int NC = W + 1;
#pragma unroll
  for (int l = 0; l < W; ++l) {
    R s = P[l * NC + NC - 1];
#pragma unroll
    for (int i = NC - 2; i >= 0; --i) s = fma(x, s, P[l * NC + i]);
    w[l] = s;
  }
}

template <typename R, int W>
__device__ __forceinline__ int node_weights(const KArgs& a, int d, double ksv, R* w) {
  const double kscale = scaled_node(ksv, a.Ntd[d]);
  const long long c = (long long)kscale;  // unsafe_trunc, kscale >= 0
  const long long off = c - a.m + 1;
  const double t = kscale - (double)off;  // in [m-1, m), exact
  if (a.mode == WIN_LINEAR) {
    const R* lut = (const R*)a.lut;
    const double idx = t * (double)a.scale;  // (kscale - off) * (LUTSize / m)
    const long long ii = (long long)floor(idx);
    const R alpha = (R)(idx - (double)ii);
#pragma unroll
    for (int l = 0; l < W; ++l) {
      long long v = ii - (long long)l * a.scale;
      long long i1 = v < 0 ? -v : v;
      long long i2 = (v + 1) < 0 ? -(v + 1) : (v + 1);
      R y1 = lut[i1], y2 = lut[i2];
      w[l] = y1 + alpha * (y2 - y1);
    }
  } else if (a.mode == WIN_POLY) {
    poly_taps<R, W>((const R*)a.poly, (R)(t - (double)a.m + 1.0 - 0.5), w);
  } else {
#pragma unroll
    for (int l = 0; l < W; ++l) w[l] = window_exact<R, W>(a, (R)(t - (double)l)) * (R)a.wscale;
  }
  return (int)(c >= a.Nt[d] ? c - a.Nt[d] : c);
}

__device__ __forceinline__ long long wrapidx(long long x, long long n) {
  while (x < 0) x += n;
  while (x >= n) x -= n;
  return x;
}

// Workgroup -> work item, XCD-aware: the dispatcher deals workgroups to the 8
// XCDs round-robin, so consecutive items (neighbouring tiles sharing halo
// lines) are given to the same XCD and its L2.
__device__ __forceinline__ int xcd_item(int nitems) {
  const int per = (nitems + 7) >> 3;
  return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// ---------------------------------------------------------------------------
// DPP reductions (all 64 lanes must be active)
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  int2 u = *reinterpret_cast<int2*>(&v);
  u.x = __builtin_amdgcn_update_dpp(0, u.x, CTRL, 0xF, 0xF, false);
  u.y = __builtin_amdgcn_update_dpp(0, u.y, CTRL, 0xF, 0xF, false);
  return *reinterpret_cast<double*>(&u);
}
// DPP move restricted to the rows of ROWMASK (other lanes read 0).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_rows(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xF, false));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_rows(double v) {
  int2 u = *reinterpret_cast<int2*>(&v);
  u.x = __builtin_amdgcn_update_dpp(0, u.x, CTRL, ROWMASK, 0xF, false);
  u.y = __builtin_amdgcn_update_dpp(0, u.y, CTRL, ROWMASK, 0xF, false);
  return *reinterpret_cast<double*>(&u);
}
__device__ __forceinline__ float readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane(double v, int l) {
  int2 u = *reinterpret_cast<int2*>(&v);
  u.x = __builtin_amdgcn_readlane(u.x, l);
  u.y = __builtin_amdgcn_readlane(u.y, l);
  return *reinterpret_cast<double*>(&u);
}

// Sum over aligned groups of G lanes (G in 1,2,4,8,16,64); every lane of the
// group ends with the group sum.
template <int G, typename R>
__device__ __forceinline__ R group_sum(R v) {
  if (G >= 2) v += dpp<0xB1>(v);   // quad_perm(1,0,3,2)
  if (G >= 4) v += dpp<0x4E>(v);   // quad_perm(2,3,0,1)
  if (G >= 8) v += dpp<0x141>(v);  // row_half_mirror
  if (G >= 16) v += dpp<0x140>(v); // row_mirror
  if (G == 64) v = readlane(v, 0) + readlane(v, 16) + readlane(v, 32) + readlane(v, 48);
  return v;
}

template <int W>
struct LanesPerNode {
  static constexpr int value = W <= 2 ? 2 : W <= 4 ? 4 : W <= 8 ? 8 : 16;
};

template <typename R>
__device__ __forceinline__ void lds_add(R* p, R v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename R>
__device__ __forceinline__ void gbl_add(R* p, R v) {
  unsafeAtomicAdd(p, v);
}

// ---------------------------------------------------------------------------
// Node records and the LDS layout of the tile kernels
// ---------------------------------------------------------------------------
// Node record, one per node in tile order, built once per node set: the
// window weights of every dimension (D x W values of R), the node's cell
// relative to its tile origin (D <= 2: packed in 32 / D bits per dimension;
// D = 3: the LDS element index x + Sx y + Sxy z of that cell in the padded
// tile), and the caller's index of the node.  16-byte aligned for block
// copies.
template <typename R, int D, int W>
struct NodeRec {
  static constexpr int WBYTES = D * W * (int)sizeof(R);
  static constexpr int BYTES = (WBYTES + 8 + 15) & ~15;
  static constexpr int LOCBITS = 32 / D;
};
__host__ inline int node_rec_bytes(int D, size_t sizeofR, int W) {
  return (int)((D * W * sizeofR + 8 + 15) & ~(size_t)15);
}

// LDS layout of the D <= 2 tile kernels: tile, node records of one batch,
// node values of one batch (spreading).
template <typename R, int D, int W>
struct TileLayout {
  // the spreading tile of D <= 2 accumulates in FP64 in both precisions
  static constexpr size_t ACC = D <= 2 ? sizeof(double) : sizeof(R);
  static __host__ __device__ size_t rec_off(int tileElems) { return (2 * (size_t)tileElems * ACC + 15) & ~(size_t)15; }
  static __host__ __device__ size_t fv_off(int tileElems, int NB) {
    return rec_off(tileElems) + (size_t)NB * NodeRec<R, D, W>::BYTES;
  }
  static __host__ __device__ size_t bytes(int tileElems, int NB) {
    return fv_off(tileElems, NB) + (size_t)NB * 2 * sizeof(R);
  }
};

// The same layout size for a runtime stencil width (host side).
__host__ inline size_t tile_layout_bytes(int D, size_t sizeofR, int W, int tileElems, int NB) {
  const size_t acc = D <= 2 ? sizeof(double) : sizeofR;
  const size_t o = (2 * (size_t)tileElems * acc + 15) & ~(size_t)15;
  return o + (size_t)NB * node_rec_bytes(D, sizeofR, W) + (size_t)NB * 2 * sizeofR;
}

// Nodes of one 3-D work item at most (items are chunks of one tile's nodes).
constexpr int ITEM3 = 1024;

// Nodes per phase-A batch.
template <int D, typename R>
struct Batch {
  static constexpr int NB = D == 3 ? (sizeof(R) == 4 ? 256 : 128) : 256;
};

__device__ __forceinline__ void tile_origin(const KArgs& a, int D, int tile, int* t0) {
  int r = tile;
  for (int d = 0; d < D; ++d) {
    int td = r % a.ntiles[d];
    r /= a.ntiles[d];
    t0[d] = td * a.T[d];
  }
}

// Node records of all nodes of one work item (one workgroup per item, one
// thread per node): weights from node_weights, the cell relative to the
// tile origin, the caller's index.
template <typename R, int D, int W>
__global__ __launch_bounds__(256) void k_node_records(KArgs a, char* __restrict__ rec) {
  using N = NodeRec<R, D, W>;
  const int4 it = a.items[blockIdx.x];
  int t0[3] = {0, 0, 0};
  {
    int r = it.x;
    for (int d = 0; d < D; ++d) {
      const int td = r % a.ntiles[d];
      r /= a.ntiles[d];
      t0[d] = td * a.T[d];
    }
  }
  for (int j = it.y + threadIdx.x; j < it.z; j += 256) {
    char* out = rec + (size_t)j * N::BYTES;
    R* w = (R*)out;
    unsigned pk = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      R wd[W];
      const int cw = node_weights<R, W>(a, d, a.ks[(size_t)d * a.M + j], wd);
#pragma unroll
      for (int l = 0; l < W; ++l) w[d * W + l] = wd[l];
      if (D == 3) pk += (unsigned)(cw - t0[d]) * (unsigned)(d == 0 ? 1 : d == 1 ? a.Sx : a.Sxy);
      else pk |= (unsigned)(cw - t0[d]) << (N::LOCBITS * d);
    }
    ((int*)(out + N::WBYTES))[0] = (int)pk;
    ((int*)(out + N::WBYTES))[1] = a.perm[j];
  }
}

// Phase A: copy the records of nodes [b0, b0 + nb) into LDS (16-byte
// blocks); for spreading also gather their values fhat[pj] * scale.
template <typename R, int D, int W, int NT, bool SPREAD>
__device__ __forceinline__ void phase_a(const KArgs& a, const typename cpx<R>::T* __restrict__ fhat, int b0, int nb,
                                        char* lrec, typename cpx<R>::T* fv, R scale) {
  using N = NodeRec<R, D, W>;
  const uint4* src = (const uint4*)(a.rec + (size_t)b0 * N::BYTES);
  uint4* dst = (uint4*)lrec;
  const int n16 = nb * (N::BYTES / 16);
  for (int i = threadIdx.x; i < n16; i += NT) dst[i] = src[i];
  if (SPREAD) {
    for (int n = threadIdx.x; n < nb; n += NT) {
      const int jj = ((const int*)(a.rec + (size_t)(b0 + n) * N::BYTES + N::WBYTES))[1];
      const typename cpx<R>::T v = fhat[(long long)jj * a.fst];
      fv[n] = typename cpx<R>::T{v.x * scale, v.y * scale};
    }
  }
}

// Fields of a record in LDS.
template <typename R, int D, int W>
__device__ __forceinline__ const R* rec_w(const char* lrec, int n) {
  return (const R*)(lrec + n * NodeRec<R, D, W>::BYTES);
}
template <typename R, int D, int W>
__device__ __forceinline__ unsigned rec_loc(const char* lrec, int n) {
  return ((const unsigned*)(lrec + n * NodeRec<R, D, W>::BYTES + NodeRec<R, D, W>::WBYTES))[0];
}
template <typename R, int D, int W>
__device__ __forceinline__ int rec_pj(const char* lrec, int n) {
  return ((const int*)(lrec + n * NodeRec<R, D, W>::BYTES + NodeRec<R, D, W>::WBYTES))[1];
}

// ---------------------------------------------------------------------------
// Spreading
// ---------------------------------------------------------------------------
// Packed fixed-point term: the 64-bit integer im * 2^32 + re (re, im signed
// 32-bit).  Sums of such terms decode exactly while |sum re|, |sum im| < 2^31.
__device__ __forceinline__ unsigned long long pack_term(float vr, float vi) {
  const int re = __float2int_rn(vr), im = __float2int_rn(vi);
  return ((unsigned long long)(unsigned)(im + (re >> 31)) << 32) | (unsigned)re;
}

template <typename R>
struct SpreadCell;
template <>
struct SpreadCell<float> {  // one 64-bit integer per cell
  static __device__ __forceinline__ void add(char* tile, int te, int idx, float vr, float vi) {
    __hip_atomic_fetch_add(&((unsigned long long*)tile)[idx], pack_term(vr, vi), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  static __device__ __forceinline__ float2 get(const char* tile, int te, int idx, float inv) {
    const long long s = (long long)((const unsigned long long*)tile)[idx];
    const int re = (int)(unsigned)s;
    const int im = (int)((s - (long long)re) >> 32);
    return float2{(float)re * inv, (float)im * inv};
  }
};
template <>
struct SpreadCell<double> {  // two FP64 planes
  static __device__ __forceinline__ void add(char* tile, int te, int idx, double vr, double vi) {
    double* t = (double*)tile;
    __hip_atomic_fetch_add(&t[idx], vr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&t[te + idx], vi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  static __device__ __forceinline__ double2 get(const char* tile, int te, int idx, double) {
    const double* t = (const double*)tile;
    return double2{t[idx], t[te + idx]};
  }
};

// One work item (<= CH nodes of one tile, D <= 2) of spreading: the padded
// tile is accumulated in LDS -- FP64 adds of the data-precision products, one
// rounding per term -- and written to the item's slot of `scratch`.
template <typename R, int D, int W, int NT>
__device__ __forceinline__ void spread_tile_item(const KArgs& a, int item,
                                                 const typename cpx<R>::T* __restrict__ fhat,
                                                 typename cpx<R>::T* __restrict__ scratch) {
  static_assert(D <= 2, "3-D spreading lives in k_tile3.hip");
  using C = typename cpx<R>::T;
  using L = TileLayout<R, D, W>;
  using A = double;
  constexpr int NB = Batch<D, R>::NB;
  constexpr int NW = NT / 64;
  constexpr unsigned LM = NodeRec<R, D, W>::LOCBITS >= 32 ? 0xFFFFFFFFu : (1u << NodeRec<R, D, W>::LOCBITS) - 1u;
  extern __shared__ __align__(16) char smem[];
  const int4 it = a.items[item];
  char* tile = smem;
  char* lrec = smem + L::rec_off(a.tileElems);
  C* fv = (C*)(smem + L::fv_off(a.tileElems, NB));
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;

  for (int i = threadIdx.x; i < 2 * a.tileElems; i += NT) ((A*)tile)[i] = 0;

  // FP32 2-D with fix2: packed 32+32-bit fixed point in the first tileElems
  // 64-bit words, scale s = 2^(30 - e) from a strict bound of the item's
  // sums, B = sum_j |f_j|_inf * wmaxD < 2^e (as the 3-D FP32 kernel); one
  // ds_add_u64 per tap instead of two ds_add_f64
  const bool fix = sizeof(R) == 4 && D == 2 && a.fix2;
  R fscale = 1, inv = 1;
  if (fix) {
    R part = 0;
    for (int j = it.y + (int)threadIdx.x; j < it.z; j += NT) {
      const int jj = ((const int*)(a.rec + (size_t)j * NodeRec<R, D, W>::BYTES + NodeRec<R, D, W>::WBYTES))[1];
      const C v = fhat[(long long)jj * a.fst];
      part += fmax(fabs(v.x), fabs(v.y));
    }
    part = group_sum<64>(part);
    R* red = (R*)lrec;
    if (lane == 0) red[wave] = part;
    __syncthreads();
    double B = 0;
    for (int w = 0; w < NW; ++w) B += (double)red[w];
    B *= a.oscale * a.wmaxD * 1.0000001;
    int e = 0;
    if (B > 0) (void)frexp(B, &e);
    // both powers stay normal floats; a non-finite bound (Inf or NaN data)
    // spreads zeros and turns the slot into NaN at the flush
    e = max(e, -96);
    fscale = (R)ldexp(1.0, 30 - e);
    inv = (R)ldexp(1.0, e - 30);
    if (!(B <= 1.0e300)) {
      fscale = 0;
      inv = (R)NAN;
    }
  }

  constexpr int G = LanesPerNode<W>::value;
  constexpr int NPW = 64 / G;
  const int slot = lane / G, lx = lane & (G - 1);
  for (int b0 = it.y; b0 < it.z; b0 += NB) {
    const int nb = min(NB, it.z - b0);
    __syncthreads();
    phase_a<R, D, W, NT, true>(a, fhat, b0, nb, lrec, fv, (R)a.oscale * fscale);
    __syncthreads();
    for (int n0 = wave * NPW; n0 < nb; n0 += NW * NPW) {
      const int n = n0 + slot;
      if (n < nb && lx < W) {
        const C f = fv[n];
        const R* w = rec_w<R, D, W>(lrec, n);
        const unsigned pk = rec_loc<R, D, W>(lrec, n);
        const R w1 = w[lx];
        const R vr = f.x * w1, vi = f.y * w1;
        if constexpr (D == 2) {
          const int base = ((int)(pk & LM) + lx) + a.Sx * (int)(pk >> 16);
          if (fix) {
#pragma unroll
            for (int ly = 0; ly < W; ++ly) {
              const R w2 = w[W + ly];
              SpreadCell<float>::add(tile, a.tileElems, base + a.Sx * ly, (float)(vr * w2), (float)(vi * w2));
            }
          } else {
#pragma unroll
            for (int ly = 0; ly < W; ++ly) {
              const R w2 = w[W + ly];
              SpreadCell<A>::add(tile, a.tileElems, base + a.Sx * ly, (A)(vr * w2), (A)(vi * w2));
            }
          }
        } else {
          SpreadCell<A>::add(tile, a.tileElems, (int)pk + lx, (A)vr, (A)vi);
        }
      }
    }
  }
  __syncthreads();

  // padded tile -> this item's scratch slot (dense, px fastest), plain stores
  C* out = scratch + (size_t)item * a.padCells;
  for (int i = threadIdx.x; i < a.padCells; i += NT) {
    int px = i, py = 0;
    if (D >= 2) {
      const int q = (int)fdiv((uint32_t)i, a.divP0);
      px = i - q * a.Pd[0];
      py = q;
    }
    if (fix) {
      const float2 v = SpreadCell<float>::get(tile, a.tileElems, px + a.Sx * py, (float)inv);
      out[i] = C{(R)v.x, (R)v.y};
    } else {
      const auto v = SpreadCell<A>::get(tile, a.tileElems, px + a.Sx * py, 1.0);
      out[i] = C{(R)v.x, (R)v.y};
    }
  }
}

template <typename R, int D, int W, int NT>
__global__ __launch_bounds__(NT) void k_spread_tile(KArgs a, const typename cpx<R>::T* __restrict__ fhat,
                                                    typename cpx<R>::T* __restrict__ unused) {
  // items round robin over the XCDs: where nodes cluster, the items of one
  // region are much heavier than the rest, and contiguous item ranges per
  // XCD (xcd_item) leave the XCDs unevenly loaded
  const int item = blockIdx.x;
  if (item >= a.nitems) return;
  spread_tile_item<R, D, W, NT>(a, item, fhat, (typename cpx<R>::T*)a.scratch);
}

// Fold: grid point x sums, over the padded tiles covering it, the scratch
// slots of every work item of those tiles.  Tile t's padded range is
// [tT - m + 1, tT + T + m - 1] modulo Ntilde; with that range no longer than
// Ntilde and 2T >= 2m - 1 (checked by the host) a point has at most 3
// covering (tile, offset) pairs per dimension.
// Covering pairs in the order of the images y = x - Nt, x, x + Nt and
// ascending tile: tile 0 (y = x - Nt, its left halo wrapped), tiles t - 1,
// t, t + 1 of x's own tile t (y = x), the last two tiles (y = x + Nt, a
// right halo wrapped).  With T >= m no other tile can cover x, so one
// division per point suffices.
__device__ __forceinline__ int cover_pairs(const KArgs& a, int d, int x, int* tl, int* pl) {
  const int T = a.T[d], P = a.Pd[d], Nt = a.Nt[d], nt = a.ntiles[d], m1 = a.m - 1;
  const int t = x / T;
  const int cu[6] = {0, t - 1, t, t + 1, nt - 2, nt - 1};
  const int cy[6] = {x - Nt, x, x, x, x + Nt, x + Nt};
  int n = 0;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int u = cu[q];
    const int p = cy[q] - u * T + m1;
    if (u >= 0 && u < nt && p >= 0 && p < P && n < 3) {
      tl[n] = u;
      pl[n] = p;
      ++n;
    }
  }
  return n;
}

// Sum at grid point x (dimension 1) of the scratch slots covering it, given
// the covering (tile, offset) pairs of dimensions 2 and 3.
template <typename R, int D>
__device__ __forceinline__ typename cpx<R>::T fold_sum(const KArgs& a, const typename cpx<R>::T* __restrict__ scr,
                                                       int x, int n1, const int* t1, const int* p1, int n2,
                                                       const int* t2, const int* p2) {
  using C = typename cpx<R>::T;
  int t0[3] = {0, 0, 0}, p0[3] = {0, 0, 0};
  const int n0 = cover_pairs(a, 0, x, t0, p0);
  R sr = 0, si = 0;
#pragma unroll
  for (int c2 = 0; c2 < 3; ++c2) {
    if (c2 >= n2) break;
#pragma unroll
    for (int c1 = 0; c1 < 3; ++c1) {
      if (c1 >= n1) break;
#pragma unroll
      for (int c0 = 0; c0 < 3; ++c0) {
        if (c0 >= n0) break;
        const int t = t0[c0] + a.ntiles[0] * (t1[c1] + a.ntiles[1] * t2[c2]);
        const int off = p0[c0] + a.Pd[0] * (p1[c1] + a.Pd[1] * p2[c2]);
        for (int itm = a.tstart[t]; itm < a.tstart[t + 1]; ++itm) {
          const C v = scr[(size_t)itm * a.padCells + off];
          sr += v.x;
          si += v.y;
        }
      }
    }
  }
  return C{sr, si};
}

template <typename R, int D>
__global__ __launch_bounds__(256) void k_fold(KArgs a, typename cpx<R>::T* __restrict__ grid) {
  using C = typename cpx<R>::T;
  const C* scr = (const C*)a.scratch;
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int ny = D >= 2 ? a.Nt[1] : 1;
  const int nrows = (D >= 2 ? a.Nt[1] : 1) * (D == 3 ? a.Nt[2] : 1);
  int t1[3] = {0, 0, 0}, p1[3] = {0, 0, 0}, t2[3] = {0, 0, 0}, p2[3] = {0, 0, 0};
  int n1 = 1, n2 = 1;
  for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
    const int y = row % ny, z = row / ny;
    if (D >= 2) n1 = cover_pairs(a, 1, y, t1, p1);
    if (D == 3) n2 = cover_pairs(a, 2, z, t2, p2);
    if (x >= a.Nt[0]) continue;
    grid[(long long)row * a.Nt[0] + x] = fold_sum<R, D>(a, scr, x, n1, t1, p1, n2, t2, p2);
  }
}

// ---------------------------------------------------------------------------
// Interpolation
// ---------------------------------------------------------------------------
// One work item of interpolation (D <= 2): stage the padded tile, gather
// every node with G lanes per node.
template <typename R, int D, int W, int NT>
__device__ __forceinline__ void interp_tile_item(const KArgs& a, int item,
                                                 const typename cpx<R>::T* __restrict__ grid,
                                                 typename cpx<R>::T* __restrict__ fhat) {
  static_assert(D <= 2, "3-D interpolation lives in k_tile3.hip");
  using C = typename cpx<R>::T;
  using L = TileLayout<R, D, W>;
  constexpr int NB = Batch<D, R>::NB;
  constexpr int NW = NT / 64;
  constexpr unsigned LM = NodeRec<R, D, W>::LOCBITS >= 32 ? 0xFFFFFFFFu : (1u << NodeRec<R, D, W>::LOCBITS) - 1u;
  extern __shared__ __align__(16) char smem[];
  const int4 it = a.items[item];
  C* tile = (C*)smem;
  char* lrec = smem + L::rec_off(a.tileElems);
  int t0[3] = {0, 0, 0};
  tile_origin(a, D, it.x, t0);

  // stage the padded tile
  for (int i = threadIdx.x; i < a.padCells; i += NT) {
    int px = i, py = 0;
    if (D >= 2) {
      const int q = (int)fdiv((uint32_t)i, a.divP0);
      px = i - q * a.Pd[0];
      py = q;
    }
    long long gi = wrapidx((long long)t0[0] - a.m + 1 + px, a.Nt[0]);
    if (D >= 2) gi += wrapidx((long long)t0[1] - a.m + 1 + py, a.Nt[1]) * a.gstr[1];
    tile[px + a.Sx * py] = grid[gi];
  }

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int G = LanesPerNode<W>::value;
  constexpr int NPW = 64 / G;
  const int slot = lane / G, lx = lane & (G - 1);
  for (int b0 = it.y; b0 < it.z; b0 += NB) {
    const int nb = min(NB, it.z - b0);
    __syncthreads();
    phase_a<R, D, W, NT, false>(a, nullptr, b0, nb, lrec, nullptr, (R)1);
    __syncthreads();
    for (int n0 = wave * NPW; n0 < nb; n0 += NW * NPW) {
      const int n = n0 + slot;
      R ar = 0, ai = 0;
      if (n < nb && lx < W) {
        const R* w = rec_w<R, D, W>(lrec, n);
        const unsigned pk = rec_loc<R, D, W>(lrec, n);
        if constexpr (D == 2) {
          const int base = ((int)(pk & LM) + lx) + a.Sx * (int)(pk >> 16);
#pragma unroll
          for (int ly = 0; ly < W; ++ly) {
            const C v = tile[base + a.Sx * ly];
            ar = fma(w[W + ly], v.x, ar);
            ai = fma(w[W + ly], v.y, ai);
          }
          ar *= w[lx];
          ai *= w[lx];
        } else {
          const C v = tile[(int)pk + lx];
          ar = w[lx] * v.x;
          ai = w[lx] * v.y;
        }
      }
      ar = group_sum<G>(ar);
      ai = group_sum<G>(ai);
      if (lx == 0 && n < nb)
        fhat[(long long)rec_pj<R, D, W>(lrec, n) * a.fst] = C{ar * (R)a.oscale, ai * (R)a.oscale};
    }
  }
}

template <typename R, int D, int W, int NT>
__global__ __launch_bounds__(NT) void k_interp_tile(KArgs a, const typename cpx<R>::T* __restrict__ grid,
                                                    typename cpx<R>::T* __restrict__ fhat) {
  const int item = blockIdx.x;  // round robin over the XCDs (k_spread_tile)
  if (item >= a.nitems) return;
  interp_tile_item<R, D, W, NT>(a, item, grid, fhat);
}

}  // namespace nfftm
