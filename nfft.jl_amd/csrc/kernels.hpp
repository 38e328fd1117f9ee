// kernels.hpp -- device code of the gfx950 NFFT gridding path.
//
// Spreading (convolve_transpose!, src/convolution.jl:115-149) and
// interpolation (convolve!, src/convolution.jl:20-45) over nodes that the host
// has bin-sorted into tiles of the oversampled grid.  One workgroup owns one
// work item = (tile, contiguous range of that tile's nodes).  The tile plus
// its stencil halo ("padded tile", extent T + 2m - 1 per dimension) lives in
// LDS:
//   * spreading accumulates every node's (2m)^D footprint into the padded tile
//     with LDS float atomics, then flushes it to the HBM grid with one
//     coalesced pass of global atomics (halos overlap neighbouring tiles);
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
  double b;         // Kaiser-Bessel shape pi*(2 - 1/sigma)
  double wscale;    // power-of-two window normalisation
  // nodes
  const double* ks;   // ks[d * M + j], shifted into [0, 1)
  const int* perm;    // sorted position -> caller's node index
  const int4* items;  // {tile, begin, end, 0}
  int nitems;
};

// Window values of one node along dimension d.  Returns the wrapped cell
// floor(k * Ntilde) (in [0, Ntilde)); t = kscale - off with
// off = cell - m + 1 exactly as src/precomputation.jl:537-545.
template <typename R, int W>
__device__ __forceinline__ int node_weights(const KArgs& a, int d, double ksv, R* w) {
  const double kscale = ksv * a.Ntd[d];
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
    const R* P = (const R*)a.poly;
    const int nc = 2 * a.m + 1;
    const R x = (R)(t - (double)a.m + 1.0 - 0.5);
#pragma unroll
    for (int l = 0; l < W; ++l) {
      const R* cf = P + l * nc;
      R s = cf[nc - 1];
      for (int i = nc - 2; i >= 0; --i) s = fma(x, s, cf[i]);
      w[l] = s;
    }
  } else {
    const R m = (R)a.m, b = (R)a.b;
#pragma unroll
    for (int l = 0; l < W; ++l) {
      R u = (R)(t - (double)l);
      R au = u < 0 ? -u : u;
      R y;
      if (au < m) {
        R arg = sqrt(m * m - u * u);
        y = sinh(b * arg) / (arg * (R)M_PI);
      } else if (au > m) {
        y = 0;
      } else {
        y = b / (R)M_PI;
      }
      w[l] = y * (R)a.wscale;
    }
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
// LDS layout of the tile kernels (dynamic shared memory):
//   [tile: 2 * tileElems R]  [wts: NB * D * W R]  [fv: NB complex]
//   [loc: NB * D int] [pj: NB int]
// ---------------------------------------------------------------------------
template <typename R, int D, int W>
struct TileLayout {
  static __host__ __device__ size_t wts_off(int tileElems) { return 2 * (size_t)tileElems * sizeof(R); }
  static __host__ __device__ size_t fv_off(int tileElems, int NB) {
    size_t o = wts_off(tileElems) + (size_t)NB * D * W * sizeof(R);
    return (o + 15) & ~(size_t)15;
  }
  static __host__ __device__ size_t loc_off(int tileElems, int NB) {
    return fv_off(tileElems, NB) + (size_t)NB * 2 * sizeof(R);
  }
  static __host__ __device__ size_t pj_off(int tileElems, int NB) {
    return loc_off(tileElems, NB) + (size_t)NB * D * sizeof(int);
  }
  static __host__ __device__ size_t bytes(int tileElems, int NB) {
    return pj_off(tileElems, NB) + (size_t)NB * sizeof(int);
  }
};

// The same layout size for a runtime stencil width (host side).
__host__ inline size_t tile_layout_bytes(int D, size_t sizeofR, int W, int tileElems, int NB) {
  size_t o = 2 * (size_t)tileElems * sizeofR + (size_t)NB * D * W * sizeofR;
  o = (o + 15) & ~(size_t)15;
  return o + (size_t)NB * 2 * sizeofR + (size_t)NB * D * sizeof(int) + (size_t)NB * sizeof(int);
}

// Nodes per phase-A batch.
template <int D>
struct Batch {
  static constexpr int NB = D == 3 ? 64 : D == 2 ? 128 : 256;
};

__device__ __forceinline__ void tile_origin(const KArgs& a, int D, int tile, int* t0) {
  int r = tile;
  for (int d = 0; d < D; ++d) {
    int td = r % a.ntiles[d];
    r /= a.ntiles[d];
    t0[d] = td * a.T[d];
  }
}

// Phase A: weights, local offsets, values and output indices of a batch of
// nodes [b0, b0 + nb) into LDS.  One thread per (node, dim).
template <typename R, int D, int W, int NT, bool SPREAD>
__device__ __forceinline__ void phase_a(const KArgs& a, const typename cpx<R>::T* __restrict__ fhat,
                                        int b0, int nb, const int* t0, R* wts,
                                        typename cpx<R>::T* fv, int* loc, int* pj) {
  for (int q = threadIdx.x; q < nb * D; q += NT) {
    const int n = q / D, d = q - n * D;
    const int j = b0 + n;
    const double ksv = a.ks[(size_t)d * a.M + j];
    R w[W];
    const int cw = node_weights<R, W>(a, d, ksv, w);
#pragma unroll
    for (int l = 0; l < W; ++l) wts[(n * D + d) * W + l] = w[l];
    loc[n * D + d] = cw - t0[d];
    if (d == 0) {
      const int jj = a.perm[j];
      pj[n] = jj;
      if (SPREAD) fv[n] = fhat[jj];
    }
  }
}

// ---------------------------------------------------------------------------
// Spreading
// ---------------------------------------------------------------------------
template <typename R, int D, int W, int NT>
__global__ __launch_bounds__(NT) void k_spread_tile(KArgs a, const typename cpx<R>::T* __restrict__ fhat,
                                                    typename cpx<R>::T* __restrict__ grid) {
  using C = typename cpx<R>::T;
  using L = TileLayout<R, D, W>;
  constexpr int NB = Batch<D>::NB;
  constexpr int NW = NT / 64;
  extern __shared__ __align__(16) char smem[];
  const int item = xcd_item(a.nitems);
  if (item >= a.nitems) return;
  const int4 it = a.items[item];
  R* tre = (R*)smem;
  R* tim = tre + a.tileElems;
  R* wts = (R*)(smem + L::wts_off(a.tileElems));
  C* fv = (C*)(smem + L::fv_off(a.tileElems, NB));
  int* loc = (int*)(smem + L::loc_off(a.tileElems, NB));
  int* pj = (int*)(smem + L::pj_off(a.tileElems, NB));
  int t0[3] = {0, 0, 0};
  tile_origin(a, D, it.x, t0);

  for (int i = threadIdx.x; i < 2 * a.tileElems; i += NT) tre[i] = 0;

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int b0 = it.y; b0 < it.z; b0 += NB) {
    const int nb = min(NB, it.z - b0);
    __syncthreads();
    phase_a<R, D, W, NT, true>(a, fhat, b0, nb, t0, wts, fv, loc, pj);
    __syncthreads();
    if constexpr (D == 3) {
      for (int n = wave; n < nb; n += NW) {
        const C f = fv[n];
        const R* w = wts + n * 3 * W;
        R w3[W];
#pragma unroll
        for (int l = 0; l < W; ++l) w3[l] = w[2 * W + l];
        const int bx = loc[n * 3], by = loc[n * 3 + 1], bz = loc[n * 3 + 2];
        for (int t = lane; t < W * W; t += 64) {
          const int ly = t / W, lx = t - ly * W;
          const R wxy = w[lx] * w[W + ly];
          const R vr = f.x * wxy, vi = f.y * wxy;
          const int base = (bx + lx) + a.Sx * (by + ly) + a.Sxy * bz;
#pragma unroll
          for (int lz = 0; lz < W; ++lz) {
            lds_add(&tre[base + a.Sxy * lz], vr * w3[lz]);
            lds_add(&tim[base + a.Sxy * lz], vi * w3[lz]);
          }
        }
      }
    } else {
      constexpr int G = LanesPerNode<W>::value;
      constexpr int NPW = 64 / G;
      const int slot = lane / G, lx = lane & (G - 1);
      for (int n0 = wave * NPW; n0 < nb; n0 += NW * NPW) {
        const int n = n0 + slot;
        if (n < nb && lx < W) {
          const C f = fv[n];
          const R* w = wts + n * D * W;
          const R w1 = w[lx];
          const R vr = f.x * w1, vi = f.y * w1;
          if constexpr (D == 2) {
            const int base = (loc[n * 2] + lx) + a.Sx * loc[n * 2 + 1];
#pragma unroll
            for (int ly = 0; ly < W; ++ly) {
              const R w2 = w[W + ly];
              lds_add(&tre[base + a.Sx * ly], vr * w2);
              lds_add(&tim[base + a.Sx * ly], vi * w2);
            }
          } else {
            const int base = loc[n] + lx;
            lds_add(&tre[base], vr);
            lds_add(&tim[base], vi);
          }
        }
      }
    }
  }
  __syncthreads();

  // flush the padded tile: global index (t0 - m + 1 + p) mod Ntilde per dim
  R* g = (R*)grid;
  for (int i = threadIdx.x; i < a.padCells; i += NT) {
    int px = i, py = 0, pz = 0;
    if (D >= 2) {
      const int q = (int)fdiv((uint32_t)i, a.divP0);
      px = i - q * a.Pd[0];
      py = q;
      if (D == 3) {
        const int q2 = (int)fdiv((uint32_t)q, a.divP1);
        py = q - q2 * a.Pd[1];
        pz = q2;
      }
    }
    const int li = px + a.Sx * py + a.Sxy * pz;
    const R vr = tre[li], vi = tim[li];
    if (vr != (R)0 || vi != (R)0) {
      long long gi = wrapidx((long long)t0[0] - a.m + 1 + px, a.Nt[0]);
      if (D >= 2) gi += wrapidx((long long)t0[1] - a.m + 1 + py, a.Nt[1]) * a.gstr[1];
      if (D == 3) gi += wrapidx((long long)t0[2] - a.m + 1 + pz, a.Nt[2]) * a.gstr[2];
      gbl_add(&g[2 * gi], vr);
      gbl_add(&g[2 * gi + 1], vi);
    }
  }
}

// ---------------------------------------------------------------------------
// Interpolation
// ---------------------------------------------------------------------------
template <typename R, int D, int W, int NT>
__global__ __launch_bounds__(NT) void k_interp_tile(KArgs a, const typename cpx<R>::T* __restrict__ grid,
                                                    typename cpx<R>::T* __restrict__ fhat) {
  using C = typename cpx<R>::T;
  using L = TileLayout<R, D, W>;
  constexpr int NB = Batch<D>::NB;
  constexpr int NW = NT / 64;
  extern __shared__ __align__(16) char smem[];
  const int item = xcd_item(a.nitems);
  if (item >= a.nitems) return;
  const int4 it = a.items[item];
  C* tile = (C*)smem;
  R* wts = (R*)(smem + L::wts_off(a.tileElems));
  int* loc = (int*)(smem + L::loc_off(a.tileElems, NB));
  int* pj = (int*)(smem + L::pj_off(a.tileElems, NB));
  int t0[3] = {0, 0, 0};
  tile_origin(a, D, it.x, t0);

  // stage the padded tile
  for (int i = threadIdx.x; i < a.padCells; i += NT) {
    int px = i, py = 0, pz = 0;
    if (D >= 2) {
      const int q = (int)fdiv((uint32_t)i, a.divP0);
      px = i - q * a.Pd[0];
      py = q;
      if (D == 3) {
        const int q2 = (int)fdiv((uint32_t)q, a.divP1);
        py = q - q2 * a.Pd[1];
        pz = q2;
      }
    }
    long long gi = wrapidx((long long)t0[0] - a.m + 1 + px, a.Nt[0]);
    if (D >= 2) gi += wrapidx((long long)t0[1] - a.m + 1 + py, a.Nt[1]) * a.gstr[1];
    if (D == 3) gi += wrapidx((long long)t0[2] - a.m + 1 + pz, a.Nt[2]) * a.gstr[2];
    tile[px + a.Sx * py + a.Sxy * pz] = grid[gi];
  }

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int b0 = it.y; b0 < it.z; b0 += NB) {
    const int nb = min(NB, it.z - b0);
    __syncthreads();
    phase_a<R, D, W, NT, false>(a, nullptr, b0, nb, t0, wts, nullptr, loc, pj);
    __syncthreads();
    if constexpr (D == 3) {
      for (int n = wave; n < nb; n += NW) {
        const R* w = wts + n * 3 * W;
        R w3[W];
#pragma unroll
        for (int l = 0; l < W; ++l) w3[l] = w[2 * W + l];
        const int bx = loc[n * 3], by = loc[n * 3 + 1], bz = loc[n * 3 + 2];
        R ar = 0, ai = 0;
        for (int t = lane; t < W * W; t += 64) {
          const int ly = t / W, lx = t - ly * W;
          const int base = (bx + lx) + a.Sx * (by + ly) + a.Sxy * bz;
          R sr = 0, si = 0;
#pragma unroll
          for (int lz = 0; lz < W; ++lz) {
            const C v = tile[base + a.Sxy * lz];
            sr = fma(w3[lz], v.x, sr);
            si = fma(w3[lz], v.y, si);
          }
          const R wxy = w[lx] * w[W + ly];
          ar = fma(wxy, sr, ar);
          ai = fma(wxy, si, ai);
        }
        ar = group_sum<64>(ar);
        ai = group_sum<64>(ai);
        if (lane == 0) fhat[pj[n]] = C{ar, ai};
      }
    } else {
      constexpr int G = LanesPerNode<W>::value;
      constexpr int NPW = 64 / G;
      const int slot = lane / G, lx = lane & (G - 1);
      for (int n0 = wave * NPW; n0 < nb; n0 += NW * NPW) {
        const int n = n0 + slot;
        R ar = 0, ai = 0;
        if (n < nb && lx < W) {
          const R* w = wts + n * D * W;
          if constexpr (D == 2) {
            const int base = (loc[n * 2] + lx) + a.Sx * loc[n * 2 + 1];
#pragma unroll
            for (int ly = 0; ly < W; ++ly) {
              const C v = tile[base + a.Sx * ly];
              ar = fma(w[W + ly], v.x, ar);
              ai = fma(w[W + ly], v.y, ai);
            }
            ar *= w[lx];
            ai *= w[lx];
          } else {
            const C v = tile[loc[n] + lx];
            ar = w[lx] * v.x;
            ai = w[lx] * v.y;
          }
        }
        ar = group_sum<G>(ar);
        ai = group_sum<G>(ai);
        if (lx == 0 && n < nb) fhat[pj[n]] = C{ar, ai};
      }
    }
  }
}

}  // namespace nfftm
