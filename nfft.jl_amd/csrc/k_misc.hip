// k_misc.hip -- deconvolution (apodisation + zero padding + fftshift
// placement), grid clearing, and the node-parallel any-D gridding kernels.
#include "launch.hpp"

namespace nfftm {

// Forward deconvolution fused with the grid clear: every grid point of the
// oversampled grid is written once, either with f_n * prod_d 1/psi_hat_d(n_d)
// (image index n inside the fftshifted corners) or with zero.  This is
// fill!(tmpVec, 0) (src/implementation.jl:159) followed by deconvolve!
// (src/deconvolution.jl:22-45): image index i_d < Na_d goes to grid index
// i_d - Na_d + Nt_d, the rest to i_d - Na_d.  The factor product is FP64 in
// both precisions.  One thread per pair of grid points along dimension 1
// (Ntilde is even), grid-stride, divisions by multiply-high.
__device__ __forceinline__ int img_index(int g, int N, int Nt, int Na) {
  if (g < N - Na) return g + Na;
  if (g >= Nt - Na) return g - (Nt - Na);
  return -1;
}

// Grid points 2i and 2i + 1 (i < Q / 2).
template <typename R>
__device__ __forceinline__ void deconv_pair(const DArgs& a, long long i, const typename cpx<R>::T* __restrict__ f,
                                            typename cpx<R>::T* __restrict__ g) {
  using C = typename cpx<R>::T;
  const int Nt0 = a.Nt[0], N0 = a.N[0], Na0 = a.Na[0];
  const unsigned half = (unsigned)(Nt0 >> 1);
  {
    const unsigned row = fdiv((unsigned)(i >> 0), a.div_half);  // i < 2^31 (checked on the host)
    const int x = 2 * (int)(i - (long long)row * half);
    bool inside = true;
    long long fo = 0;
    double facr = a.oscale;
    unsigned r = row;
    for (int d = 1; d < a.D; ++d) {
      const unsigned q = fdiv(r, a.div_nt[d]);
      const int gd = (int)(r - q * (unsigned)a.Nt[d]);
      r = q;
      const int id = img_index(gd, a.N[d], a.Nt[d], a.Na[d]);
      if (id < 0) {
        inside = false;
        break;
      }
      fo += (long long)id * a.nstr[d];
      facr *= a.fac[d][id];
    }
    C o0, o1;
    o0.x = o0.y = o1.x = o1.y = 0;
    if (inside) {
      const int i0 = img_index(x, N0, Nt0, Na0), i1 = img_index(x + 1, N0, Nt0, Na0);
      if (i0 >= 0) {
        const C v = f[fo + i0 * a.nstr[0]];
        const double s = a.fac[0][i0] * facr;
        o0.x = (R)((double)v.x * s);
        o0.y = (R)((double)v.y * s);
      }
      if (i1 >= 0) {
        const C v = f[fo + i1 * a.nstr[0]];
        const double s = a.fac[0][i1] * facr;
        o1.x = (R)((double)v.x * s);
        o1.y = (R)((double)v.y * s);
      }
    }
    C* dst = g + (long long)row * Nt0 + x;
    dst[0] = o0;
    dst[1] = o1;
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_deconv(DArgs a, const typename cpx<R>::T* __restrict__ f,
                                                typename cpx<R>::T* __restrict__ g) {
  const long long npairs = a.Q >> 1;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < npairs; i += (long long)gridDim.x * 256)
    deconv_pair<R>(a, i, f, g);
}

// deconvolve_transpose! (src/deconvolution.jl:69-92): f_n = g[(n mod Nt)] *
// prod_d 1/psi_hat_d(n_d).  One image row: the outer-dimension offset and
// factor product, then the threads of the workgroup along dimension 1.
template <typename R>
__device__ __forceinline__ void deconv_t_row(const DArgs& a, int row, const typename cpx<R>::T* __restrict__ g,
                                             typename cpx<R>::T* __restrict__ f, int i0, int step) {
  using C = typename cpx<R>::T;
  const int N0 = a.N[0], Na0 = a.Na[0];
  long long go = 0, fo = 0;
  double facr = 1.0;
  int r = row;
  for (int d = 1; d < a.D; ++d) {
    const int id = r % a.N[d];
    r /= a.N[d];
    const int gd = id < a.Na[d] ? id - a.Na[d] + a.Nt[d] : id - a.Na[d];
    go += (long long)gd * a.gstr[d];
    fo += (long long)id * a.nstr[d];
    facr *= a.fac[d][id];
  }
  for (int i = i0; i < N0; i += step) {
    const int gx = i < Na0 ? i - Na0 + a.Nt[0] : i - Na0;
    const C v = g[go + gx];
    const double s = a.fac[0][i] * facr * a.oscale;
    C out;
    out.x = (R)((double)v.x * s);
    out.y = (R)((double)v.y * s);
    f[fo + (long long)i * a.nstr[0]] = out;
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_deconv_t(DArgs a, const typename cpx<R>::T* __restrict__ g,
                                                  typename cpx<R>::T* __restrict__ f) {
  const int i0 = blockIdx.x * 256 + threadIdx.x;
  const int nrows = (int)(a.P / a.N[0]);
  for (int row = blockIdx.y; row < nrows; row += gridDim.y)
    if (i0 < a.N[0]) deconv_t_row<R>(a, row, g, f, i0, a.N[0]);
}

// Batch forms: work entries {slot, begin, end, 0} over grid-point pairs
// (deconvolve) or image rows (transpose), strided over a resident grid.
template <typename R>
__global__ __launch_bounds__(256) void k_deconv_batch(const DArgs* __restrict__ da, const int4* __restrict__ work,
                                                      int nwork, const void* const* __restrict__ in,
                                                      void* const* __restrict__ out) {
  using C = typename cpx<R>::T;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int4 e = work[w];
    const DArgs& a = da[e.x];
    for (long long i = (long long)e.y + threadIdx.x; i < e.z; i += 256)
      deconv_pair<R>(a, i, (const C*)in[e.x], (C*)out[e.x]);
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_deconv_t_batch(const DArgs* __restrict__ da, const int4* __restrict__ work,
                                                        int nwork, const void* const* __restrict__ in,
                                                        void* const* __restrict__ out) {
  using C = typename cpx<R>::T;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int4 e = work[w];
    const DArgs& a = da[e.x];
    for (int row = e.y; row < e.z; ++row)
      deconv_t_row<R>(a, row, (const C*)in[e.x], (C*)out[e.x], threadIdx.x, 256);
  }
}

__global__ __launch_bounds__(256) void k_fill0(uint4* __restrict__ p, long long n16) {
  const uint4 z = {0u, 0u, 0u, 0u};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n16; i += (long long)gridDim.x * 256)
    p[i] = z;
}

__global__ void k_fill0_tail(unsigned char* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Node-parallel gridding for any D (the non-blocked path of
// src/convolution.jl:56-99 and :152-200): one thread per node, footprint
// walked with an odometer, spreading by global atomics.
template <typename R, int W>
__global__ __launch_bounds__(256) void k_spread_generic(KArgs a, const typename cpx<R>::T* __restrict__ fhat,
                                                        typename cpx<R>::T* __restrict__ grid) {
  using C = typename cpx<R>::T;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= a.M) return;
  R w[MAXD][W];
  long long base[MAXD];
  for (int d = 0; d < a.D; ++d) {
    const int cw = node_weights<R, W>(a, d, a.ks[(size_t)d * a.M + j], w[d]);
    base[d] = (long long)cw - a.m + 1;
  }
  const C fj = fhat[(long long)j * a.fst];
  const C f = C{fj.x * (R)a.oscale, fj.y * (R)a.oscale};
  R* g = (R*)grid;
  int l[MAXD];
  for (int d = 0; d < MAXD; ++d) l[d] = 0;
  while (true) {
    R pw = 1;
    long long gi = 0;
    for (int d = a.D - 1; d >= 0; --d) {
      pw = pw * w[d][l[d]];
      gi += wrapidx(base[d] + l[d], a.Nt[d]) * a.gstr[d];
    }
    gbl_add(&g[2 * gi], f.x * pw);
    gbl_add(&g[2 * gi + 1], f.y * pw);
    int d = 0;
    for (; d < a.D; ++d) {
      if (++l[d] < W) break;
      l[d] = 0;
    }
    if (d == a.D) break;
  }
}

template <typename R, int W>
__global__ __launch_bounds__(256) void k_interp_generic(KArgs a, const typename cpx<R>::T* __restrict__ grid,
                                                        typename cpx<R>::T* __restrict__ fhat) {
  using C = typename cpx<R>::T;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= a.M) return;
  R w[MAXD][W];
  long long base[MAXD];
  for (int d = 0; d < a.D; ++d) {
    const int cw = node_weights<R, W>(a, d, a.ks[(size_t)d * a.M + j], w[d]);
    base[d] = (long long)cw - a.m + 1;
  }
  int l[MAXD];
  for (int d = 0; d < MAXD; ++d) l[d] = 0;
  R sr = 0, si = 0;
  while (true) {
    R pw = 1;
    long long gi = 0;
    for (int d = a.D - 1; d >= 0; --d) {
      pw = pw * w[d][l[d]];
      gi += wrapidx(base[d] + l[d], a.Nt[d]) * a.gstr[d];
    }
    const C v = grid[gi];
    sr = fma(pw, v.x, sr);
    si = fma(pw, v.y, si);
    int d = 0;
    for (; d < a.D; ++d) {
      if (++l[d] < W) break;
      l[d] = 0;
    }
    if (d == a.D) break;
  }
  fhat[(long long)j * a.fst] = C{sr * (R)a.oscale, si * (R)a.oscale};
}

template <typename R>
static hipError_t generic_w(bool spread, int W, const KArgs& a, hipStream_t s, const void* in, void* out) {
  using C = typename cpx<R>::T;
  const dim3 grid((a.M + 255) / 256), blk(256);
#define NFFTM_GEN(WW)                                                                        \
  case WW:                                                                                   \
    if (spread)                                                                              \
      hipLaunchKernelGGL((k_spread_generic<R, WW>), grid, blk, 0, s, a, (const C*)in, (C*)out); \
    else                                                                                     \
      hipLaunchKernelGGL((k_interp_generic<R, WW>), grid, blk, 0, s, a, (const C*)in, (C*)out); \
    break;
  switch (W) {
    NFFTM_GEN(2) NFFTM_GEN(4) NFFTM_GEN(6) NFFTM_GEN(8) NFFTM_GEN(10) NFFTM_GEN(12) NFFTM_GEN(14)
    NFFTM_GEN(16)
    default:
      return hipErrorInvalidValue;
  }
#undef NFFTM_GEN
  return hipGetLastError();
}

// Sparse-matrix convolution: the reference's precompute = FULL form, whose
// GPU backend holds B as a device sparse matrix (src/precomputation.jl:81-120,
// ext/NFFTGPUArraysExt/implementation.jl:3-136): row j holds the K = W^D
// window weights of node j (the plan's precompute mode, normalised as the
// generic kernels) and their grid indices, rows of fixed length (CSR without
// a row-pointer array).  Forward: fhat_j = sum_l B_jl g_l, a wavefront per
// row.  Adjoint: g_l += B_jl fhat_j by global atomics, a wavefront per row.
template <typename R, int W>
__global__ __launch_bounds__(256) void k_sparse_build(KArgs a, int* __restrict__ cols, R* __restrict__ vals) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= a.M) return;
  R w[MAXD][W];
  long long base[MAXD];
  for (int d = 0; d < a.D; ++d) {
    const int cw = node_weights<R, W>(a, d, a.ks[(size_t)d * a.M + j], w[d]);
    base[d] = (long long)cw - a.m + 1;
  }
  int l[MAXD];
  for (int d = 0; d < MAXD; ++d) l[d] = 0;
  size_t o = (size_t)j;
  for (int d = 0; d < a.D; ++d) o *= W;
  while (true) {
    R pw = 1;
    long long gi = 0;
    for (int d = a.D - 1; d >= 0; --d) {
      pw = pw * w[d][l[d]];
      gi += wrapidx(base[d] + l[d], a.Nt[d]) * a.gstr[d];
    }
    cols[o] = (int)gi;
    vals[o] = pw;
    ++o;
    int d = 0;
    for (; d < a.D; ++d) {
      if (++l[d] < W) break;
      l[d] = 0;
    }
    if (d == a.D) break;
  }
}

template <typename R>
__device__ __forceinline__ R wave_sum64(R v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename R>
__global__ __launch_bounds__(256) void k_sparse_interp(KArgs a, int K, const int* __restrict__ cols,
                                                       const R* __restrict__ vals,
                                                       const typename cpx<R>::T* __restrict__ grid,
                                                       typename cpx<R>::T* __restrict__ fhat) {
  using C = typename cpx<R>::T;
  const int j = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= a.M) return;  // whole wavefronts
  const size_t o = (size_t)j * K;
  R sr = 0, si = 0;
  for (int l = lane; l < K; l += 64) {
    const R v = vals[o + l];
    const C g = grid[cols[o + l]];
    sr = fma(v, g.x, sr);
    si = fma(v, g.y, si);
  }
  sr = wave_sum64(sr);
  si = wave_sum64(si);
  if (lane == 0) fhat[(long long)j * a.fst] = C{sr * (R)a.oscale, si * (R)a.oscale};
}

template <typename R>
__global__ __launch_bounds__(256) void k_sparse_spread(KArgs a, int K, const int* __restrict__ cols,
                                                       const R* __restrict__ vals,
                                                       const typename cpx<R>::T* __restrict__ fhat,
                                                       typename cpx<R>::T* __restrict__ grid) {
  using C = typename cpx<R>::T;
  const int j = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= a.M) return;
  const size_t o = (size_t)j * K;
  const C f = fhat[(long long)j * a.fst];
  const R fr = f.x * (R)a.oscale, fi = f.y * (R)a.oscale;
  for (int l = lane; l < K; l += 64) {
    const R v = vals[o + l];
    R* g = (R*)(grid + cols[o + l]);
    atomicAdd(g, v * fr);
    atomicAdd(g + 1, v * fi);
  }
}

template <typename R>
static hipError_t sparse_build_w(int W, const KArgs& a, hipStream_t s, int* cols, void* vals) {
  const dim3 grid((a.M + 255) / 256), blk(256);
#define NFFTM_SB(WW)                                                                       \
  case WW:                                                                                 \
    hipLaunchKernelGGL((k_sparse_build<R, WW>), grid, blk, 0, s, a, cols, (R*)vals);      \
    break;
  switch (W) {
    NFFTM_SB(2) NFFTM_SB(4) NFFTM_SB(6) NFFTM_SB(8) NFFTM_SB(10) NFFTM_SB(12) NFFTM_SB(14) NFFTM_SB(16)
    default:
      return hipErrorInvalidValue;
  }
#undef NFFTM_SB
  return hipGetLastError();
}

hipError_t launch_sparse_build(bool dbl, int W, const KArgs& a, hipStream_t s, int* cols, void* vals) {
  if (a.M == 0) return hipSuccess;
  return dbl ? sparse_build_w<double>(W, a, s, cols, vals) : sparse_build_w<float>(W, a, s, cols, vals);
}

hipError_t launch_sparse(bool spread, bool dbl, const KArgs& a, int K, const int* cols, const void* vals,
                         hipStream_t s, const void* in, void* out) {
  if (a.M == 0) return hipSuccess;
  const dim3 grid((unsigned)((a.M + 3) / 4)), blk(256);
  if (dbl) {
    using C = cpx<double>::T;
    if (spread)
      hipLaunchKernelGGL((k_sparse_spread<double>), grid, blk, 0, s, a, K, cols, (const double*)vals, (const C*)in, (C*)out);
    else
      hipLaunchKernelGGL((k_sparse_interp<double>), grid, blk, 0, s, a, K, cols, (const double*)vals, (const C*)in, (C*)out);
  } else {
    using C = cpx<float>::T;
    if (spread)
      hipLaunchKernelGGL((k_sparse_spread<float>), grid, blk, 0, s, a, K, cols, (const float*)vals, (const C*)in, (C*)out);
    else
      hipLaunchKernelGGL((k_sparse_interp<float>), grid, blk, 0, s, a, K, cols, (const float*)vals, (const C*)in, (C*)out);
  }
  return hipGetLastError();
}

hipError_t launch_generic(bool spread, bool dbl, int W, const KArgs& a, hipStream_t s, const void* in,
                          void* out) {
  if (a.M == 0) return hipSuccess;
  return dbl ? generic_w<double>(spread, W, a, s, in, out) : generic_w<float>(spread, W, a, s, in, out);
}

hipError_t launch_deconv(bool dbl, const DArgs& a, hipStream_t s, const void* f, void* g) {
  const long long blocks = std::min<long long>((a.Q / 2 + 255) / 256, 8192);
  const dim3 grid((unsigned)blocks), blk(256);
  if (dbl)
    hipLaunchKernelGGL(k_deconv<double>, grid, blk, 0, s, a, (const double2*)f, (double2*)g);
  else
    hipLaunchKernelGGL(k_deconv<float>, grid, blk, 0, s, a, (const float2*)f, (float2*)g);
  return hipGetLastError();
}

hipError_t launch_deconv_t(bool dbl, const DArgs& a, hipStream_t s, const void* g, void* f) {
  const long long rows = a.P / a.N[0];
  const dim3 grid((a.N[0] + 255) / 256, (unsigned)std::min<long long>(rows, 65535)), blk(256);
  if (dbl)
    hipLaunchKernelGGL(k_deconv_t<double>, grid, blk, 0, s, a, (const double2*)g, (double2*)f);
  else
    hipLaunchKernelGGL(k_deconv_t<float>, grid, blk, 0, s, a, (const float2*)g, (float2*)f);
  return hipGetLastError();
}

hipError_t launch_deconv_batch(bool dbl, int nblk, hipStream_t s, const DArgs* da, const int4* work, int nwork,
                               const void* const* in, void* const* out) {
  if (nwork == 0) return hipSuccess;
  if (dbl) hipLaunchKernelGGL(k_deconv_batch<double>, dim3(nblk), dim3(256), 0, s, da, work, nwork, in, out);
  else hipLaunchKernelGGL(k_deconv_batch<float>, dim3(nblk), dim3(256), 0, s, da, work, nwork, in, out);
  return hipGetLastError();
}

hipError_t launch_deconv_t_batch(bool dbl, int nblk, hipStream_t s, const DArgs* da, const int4* work, int nwork,
                                 const void* const* in, void* const* out) {
  if (nwork == 0) return hipSuccess;
  if (dbl) hipLaunchKernelGGL(k_deconv_t_batch<double>, dim3(nblk), dim3(256), 0, s, da, work, nwork, in, out);
  else hipLaunchKernelGGL(k_deconv_t_batch<float>, dim3(nblk), dim3(256), 0, s, da, work, nwork, in, out);
  return hipGetLastError();
}

template <typename R, int D>
static hipError_t records_w(int W, const KArgs& a, char* rec, hipStream_t s) {
  const dim3 g((unsigned)a.nitems), b(256);
  switch (W) {
#define NFFTM_REC(WW) \
  case WW:            \
    hipLaunchKernelGGL((k_node_records<R, D, WW>), g, b, 0, s, a, rec); \
    break;
    NFFTM_REC(2) NFFTM_REC(4) NFFTM_REC(6) NFFTM_REC(8) NFFTM_REC(10) NFFTM_REC(12) NFFTM_REC(14) NFFTM_REC(16)
#undef NFFTM_REC
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_node_records(int D, bool dbl, int W, const KArgs& a, void* rec, hipStream_t s) {
  if (a.nitems == 0) return hipSuccess;
  char* r = (char*)rec;
  if (D == 1) return dbl ? records_w<double, 1>(W, a, r, s) : records_w<float, 1>(W, a, r, s);
  if (D == 2) return dbl ? records_w<double, 2>(W, a, r, s) : records_w<float, 2>(W, a, r, s);
  if (D == 3) return dbl ? records_w<double, 3>(W, a, r, s) : records_w<float, 3>(W, a, r, s);
  return hipErrorInvalidValue;
}

hipError_t launch_fill0(void* p, size_t bytes, hipStream_t s) {
  const long long n16 = (long long)(bytes / 16);
  if (n16 > 0) {
    const long long blocks = std::min<long long>((n16 + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill0, dim3((unsigned)blocks), dim3(256), 0, s, (uint4*)p, n16);
  }
  const long long tail = (long long)(bytes - (size_t)n16 * 16);
  if (tail > 0)
    hipLaunchKernelGGL(k_fill0_tail, dim3(1), dim3(64), 0, s, (unsigned char*)p + n16 * 16, tail);
  return hipGetLastError();
}

// dst[i * ds] = src[i * ss], i < n: one slice of a directional plan's grid
// between the caller's strided array and the plan's contiguous grid.
template <typename C>
__global__ __launch_bounds__(256) void k_copy_strided(const C* __restrict__ src, long long ss, C* __restrict__ dst,
                                                      long long ds, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    dst[i * ds] = src[i * ss];
}

hipError_t launch_copy_strided(bool dbl, const void* src, long long ss, void* dst, long long ds, long long n,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 65536);
  if (dbl)
    hipLaunchKernelGGL(k_copy_strided<double2>, dim3(blocks), dim3(256), 0, s, (const double2*)src, ss,
                       (double2*)dst, ds, n);
  else
    hipLaunchKernelGGL(k_copy_strided<float2>, dim3(blocks), dim3(256), 0, s, (const float2*)src, ss,
                       (float2*)dst, ds, n);
  return hipGetLastError();
}

// Inverse of the binning permutation: iperm[perm[j]] = j.
__global__ __launch_bounds__(256) void k_inv_perm(const int* __restrict__ perm, int* __restrict__ iperm, long long n) {
  for (long long j = (long long)blockIdx.x * 256 + threadIdx.x; j < n; j += (long long)gridDim.x * 256)
    iperm[perm[j]] = (int)j;
}

hipError_t launch_inv_perm(const int* perm, int* iperm, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_inv_perm, dim3(blocks), dim3(256), 0, s, perm, iperm, n);
  return hipGetLastError();
}

// Node values between sorted and caller order: out[i * os] = in[idx[i] * is]
// (idx = iperm: results of the sorted interpolation back to the caller's
// order).  Coalesced stores, gathered loads: on node sets larger than the
// caches this is cheaper than the sorted kernel storing at the caller's
// positions (8-byte stores at random positions cost a partial-line write
// each; measured on 2^24 nodes: 337 us of stores against ~180 us here).
// Each XCD takes one contiguous range of outputs (the dispatcher deals
// workgroups to the 8 XCDs round robin; gridDim.x is a multiple of 8): the
// outputs of one range read, in every tile, one contiguous part of its run,
// one line at a time, so that each line is fetched into one L2 and used
// there by all its outputs, rather than by every XCD's L2.
// Four outputs per thread, their index loads and then their gathers issued
// together (one output per thread left the kernel waiting on two dependent
// loads per output).
template <typename C>
__global__ __launch_bounds__(256) void k_gather_nodes(const C* __restrict__ in, long long is,
                                                      const int* __restrict__ idx, C* __restrict__ out,
                                                      long long os, long long n) {
  constexpr int U = 4;
  const long long b = (long long)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  for (long long i0 = b * 256 * U + threadIdx.x; i0 < n; i0 += (long long)gridDim.x * 256 * U) {
    int j[U];
#pragma unroll
    for (int u = 0; u < U; ++u) j[u] = i0 + u * 256 < n ? idx[i0 + u * 256] : -1;
    C v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = in[(long long)(j[u] < 0 ? 0 : j[u]) * is];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j[u] >= 0) out[(i0 + u * 256) * os] = v[u];
  }
}

hipError_t launch_gather_nodes(bool dbl, const void* in, long long is, const int* idx, void* out, long long os,
                               long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)(8 * ((std::min<long long>((n + 1023) / 1024, 65536) + 7) / 8));
  if (dbl)
    hipLaunchKernelGGL(k_gather_nodes<double2>, dim3(blocks), dim3(256), 0, s, (const double2*)in, is, idx,
                       (double2*)out, os, n);
  else
    hipLaunchKernelGGL(k_gather_nodes<float2>, dim3(blocks), dim3(256), 0, s, (const float2*)in, is, idx,
                       (float2*)out, os, n);
  return hipGetLastError();
}

size_t tile_lds_bytes(int D, bool dbl, int W, int tileElems) {
  int NB;
  if (dbl) NB = D == 1 ? Batch<1, double>::NB : D == 2 ? Batch<2, double>::NB : Batch<3, double>::NB;
  else NB = D == 1 ? Batch<1, float>::NB : D == 2 ? Batch<2, float>::NB : Batch<3, float>::NB;
  return tile_layout_bytes(D, dbl ? sizeof(double) : sizeof(float), W, tileElems, NB);
}

hipError_t launch_fold(int D, bool dbl, int W, const KArgs& a, hipStream_t s, void* grid) {
  if (D == 3) return launch_fold_d3(dbl, W, a, s, grid);
  long long rows = 1;
  for (int d = 1; d < D; ++d) rows *= a.Nt[d];
  const dim3 g((a.Nt[0] + 255) / 256, (unsigned)std::min<long long>(rows, 65535)), blk(256);
#define NFFTM_FOLD(DD)                                                                   \
  if (D == DD) {                                                                         \
    if (dbl) hipLaunchKernelGGL((k_fold<double, DD>), g, blk, 0, s, a, (double2*)grid);  \
    else hipLaunchKernelGGL((k_fold<float, DD>), g, blk, 0, s, a, (float2*)grid);        \
  }
  NFFTM_FOLD(1) NFFTM_FOLD(2)
#undef NFFTM_FOLD
  return hipGetLastError();
}

hipError_t launch_tile(int D, bool spread, bool dbl, int W, int NT, const KArgs& a, size_t lds, hipStream_t s,
                       const void* in, void* out) {
  if (a.nitems == 0) return hipSuccess;
  if (D == 1) return launch_tile_d1(spread, dbl, W, NT, a, lds, s, in, out);
  if (D == 2) return launch_tile_d2(spread, dbl, W, NT, a, lds, s, in, out);
  return launch_tile_d3(spread, dbl, W, NT, a, lds, s, in, out);
}

}  // namespace nfftm
