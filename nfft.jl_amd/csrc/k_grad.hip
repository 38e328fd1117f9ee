// k_grad.hip -- node gradient of the forward transform.
//
// For fhat = A(k) f the forward path computes g = FFT(deconvolve(f)) and
// fhat_j = sum_l g[idx_l] prod_d w_d(l_d) (src/convolution.jl:56-99).  The
// derivative with respect to node coordinate k_{d,j} only moves the window
// of dimension d:
//   d fhat_j / d k_{d,j} = sum_l g[idx_l] dw_d(l_d) prod_{e != d} w_e(l_e),
//   dw_d(l) = Ntilde_d psi'(u_l),  u_l = k Ntilde_d - off - l,
// and the kernel returns kbar[d + D j] = Re(conj(ybar_j) * that).  Window
// values and derivatives are FP64 (analytic Kaiser-Bessel derivative; in the
// POLYNOMIAL and TENSOR modes the derivative of the fitted polynomial, so the
// gradient is that of the operator the plan applies); the gather accumulates
// in the data precision.
//
// D <= 3 (k_grad_lanes): a group of G lanes per node (G = 2m rounded up to a
// power of two), one lane per tap along dimension 1.  Lane q evaluates tap q
// of the window in every dimension; the outer taps reach the group by
// shuffles, each footprint row is read by the group as one run of 2m
// consecutive cells, and the lanes' partial sums meet in a butterfly.
// D > 3 (k_grad_nodes): one thread per node, the footprint reduced row by row
// over an odometer of the outer indices.
#include "launch.hpp"

namespace nfftm {

// Kaiser-Bessel window psi(u) (u in grid units, src/windowFunctions.jl:21-34)
// and its derivative -u (x cosh x - sinh x) / (pi s^3), s = sqrt(m^2 - u^2),
// x = b s; for x < 1/2 the series u b^3 / pi * sum_{n>=1} x^(2n-2) 2n/(2n+1)!
// avoids the cancellation.
__device__ __forceinline__ void kb_value_deriv(double u, double m, double b, double* v, double* dv) {
  const double au = fabs(u);
  if (au >= m) {
    *v = au > m ? 0.0 : b / M_PI;
    *dv = 0.0;
    return;
  }
  const double s = sqrt(m * m - u * u);
  const double x = b * s;
  const double sh = sinh(x), ch = cosh(x);
  *v = sh / (s * M_PI);
  double q;
  if (x < 0.5) {
    const double x2 = x * x;
    double term = 1.0 / 3.0, sum = term;
#pragma unroll
    for (int n = 1; n < 12; ++n) {
      term *= x2 / (2.0 * n * (2.0 * n + 3.0));
      sum += term;
    }
    q = b * b * b * sum;
  } else {
    q = (x * ch - sh) / (s * s * s);
  }
  *dv = -u * q / M_PI;
}

// Window value and node derivative of tap l (0 <= l < 2m) at window argument
// t = k Ntilde_d - off, with the plan's power-of-two window normalisation.
__device__ __forceinline__ void tap_weight(const KArgs& a, double Ntd, double t, int l, double* w, double* dw) {
  if (a.mode == WIN_POLY) {
    const int nc = 2 * a.m + 1;
    const double* cf = a.polyd + l * nc;
    const double x = t - (double)a.m + 0.5;
    double s = cf[nc - 1], ds = 0.0;
    for (int i = nc - 2; i >= 0; --i) {
      ds = fma(x, ds, s);
      s = fma(x, s, cf[i]);
    }
    *w = s;
    *dw = Ntd * ds;
  } else {
    double v, dv;
    kb_value_deriv(t - (double)l, (double)a.m, a.b, &v, &dv);
    *w = v * a.wscale;
    *dw = Ntd * dv * a.wscale;
  }
}

// Window argument of one node along dimension d; returns the wrapped cell.
__device__ __forceinline__ int grad_cell(const KArgs& a, int d, double ksv, double* t) {
  const double kscale = scaled_node(ksv, a.Ntd[d]);
  const long long c = (long long)kscale;
  *t = kscale - (double)(c - a.m + 1);
  return (int)(c >= a.Nt[d] ? c - a.Nt[d] : c);
}

// Window values and node derivatives of all 2m taps along dimension d.
template <int W>
__device__ __forceinline__ int grad_weights(const KArgs& a, int d, double ksv, double* w, double* dw) {
  double t;
  const int cw = grad_cell(a, d, ksv, &t);
#pragma unroll
  for (int l = 0; l < W; ++l) tap_weight(a, a.Ntd[d], t, l, &w[l], &dw[l]);
  return cw;
}

// Row of 2m grid points along dimension 1 at offset ofs: dot products with
// w_0 (s) and dw_0 (t).
template <typename R, int W, typename C>
__device__ __forceinline__ void grad_row(const C* __restrict__ grid, long long ofs, const int* x0, const R* w0,
                                         const R* dw0, R& sr, R& si, R& tr, R& ti) {
  sr = si = tr = ti = 0;
#pragma unroll
  for (int q = 0; q < W; ++q) {
    const C v = grid[ofs + x0[q]];
    sr = fma(w0[q], v.x, sr);
    si = fma(w0[q], v.y, si);
    tr = fma(dw0[q], v.x, tr);
    ti = fma(dw0[q], v.y, ti);
  }
}

// Lanes per node of k_grad_lanes: 2m rounded up to a power of two.
template <int W>
constexpr int grad_group() {
  return W <= 2 ? 2 : W <= 4 ? 4 : W <= 8 ? 8 : 16;
}

template <typename R, int W, int DD>
__global__ __launch_bounds__(256) void k_grad_lanes(KArgs a, const typename cpx<R>::T* __restrict__ grid,
                                                    const typename cpx<R>::T* __restrict__ ybar,
                                                    double* __restrict__ kbar) {
  using C = typename cpx<R>::T;
  constexpr int G = grad_group<W>();
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const int s = (int)(t / G), q = (int)(t % G);
  // the G lanes of a node are consecutive and aligned: they leave together
  if (s >= a.M) return;
  const bool tap = q < W;
  // the cotangent of the node, loaded ahead of the gather
  const int jj = a.perm[s];
  const C y = ybar[jj];
  // fitted-polynomial modes: this lane's tap has the same coefficients in
  // every dimension, loaded once (the arithmetic of tap_weight)
  const bool poly = a.mode == WIN_POLY;
  double cq[W + 1];
  if (poly) {
    const double* cf = a.polyd + (tap ? q : 0) * (W + 1);
#pragma unroll
    for (int i = 0; i <= W; ++i) cq[i] = cf[i];
  }
  R w[DD], dw[DD];
  long long base[DD];
#pragma unroll
  for (int d = 0; d < DD; ++d) {
    double td, wd = 0.0, dwd = 0.0;
    const int cw = grad_cell(a, d, a.ks[(size_t)d * a.M + s], &td);
    if (tap) {
      if (poly) {
        const double x = td - (double)a.m + 0.5;
        double v = cq[W], dv = 0.0;
#pragma unroll
        for (int i = W - 1; i >= 0; --i) {
          dv = fma(x, dv, v);
          v = fma(x, v, cq[i]);
        }
        wd = v;
        dwd = a.Ntd[d] * dv;
      } else {
        tap_weight(a, a.Ntd[d], td, q, &wd, &dwd);
      }
    }
    w[d] = (R)wd;
    dw[d] = (R)dwd;
    base[d] = (long long)cw - a.m + 1;
  }
  // cell of this lane's tap along dimension 1 (lanes past 2m read tap 0 with
  // zero weights)
  const long long xi = wrapidx(base[0] + (tap ? q : 0), a.Nt[0]);
  R gr[DD], gi[DD];
#pragma unroll
  for (int d = 0; d < DD; ++d) gr[d] = gi[d] = 0;
  if constexpr (DD == 1) {
    const C v = grid[xi];
    gr[0] = dw[0] * v.x;
    gi[0] = dw[0] * v.y;
  } else if constexpr (DD == 2) {
    // row l_1: d/dk_0 weight dw_0 w_1, d/dk_1 weight w_0 dw_1
#pragma unroll
    for (int ly = 0; ly < W; ++ly) {
      const R w1 = __shfl(w[1], ly, G), dw1 = __shfl(dw[1], ly, G);
      const C v = grid[wrapidx(base[1] + ly, a.Nt[1]) * a.gstr[1] + xi];
      const R c0 = dw[0] * w1, c1 = w[0] * dw1;
      gr[0] = fma(c0, v.x, gr[0]);
      gi[0] = fma(c0, v.y, gi[0]);
      gr[1] = fma(c1, v.x, gr[1]);
      gi[1] = fma(c1, v.y, gi[1]);
    }
  } else {
    // plane l_2, row l_1: weights dw_0 w_1 w_2, w_0 dw_1 w_2, w_0 w_1 dw_2
    R w1s[W], dw1s[W];
#pragma unroll
    for (int ly = 0; ly < W; ++ly) {
      w1s[ly] = __shfl(w[1], ly, G);
      dw1s[ly] = __shfl(dw[1], ly, G);
    }
#pragma unroll 1
    for (int lz = 0; lz < W; ++lz) {
      const R w2 = __shfl(w[2], lz, G), dw2 = __shfl(dw[2], lz, G);
      const R p0 = dw[0] * w2, p1 = w[0] * w2, p2 = w[0] * dw2;
      const long long oz = wrapidx(base[2] + lz, a.Nt[2]) * a.gstr[2] + xi;
#pragma unroll
      for (int ly = 0; ly < W; ++ly) {
        const C v = grid[oz + wrapidx(base[1] + ly, a.Nt[1]) * a.gstr[1]];
        const R c0 = p0 * w1s[ly], c1 = p1 * dw1s[ly], c2 = p2 * w1s[ly];
        gr[0] = fma(c0, v.x, gr[0]);
        gi[0] = fma(c0, v.y, gi[0]);
        gr[1] = fma(c1, v.x, gr[1]);
        gi[1] = fma(c1, v.y, gi[1]);
        gr[2] = fma(c2, v.x, gr[2]);
        gi[2] = fma(c2, v.y, gi[2]);
      }
    }
  }
#pragma unroll
  for (int o = G / 2; o >= 1; o >>= 1)
#pragma unroll
    for (int d = 0; d < DD; ++d) {
      gr[d] += __shfl_xor(gr[d], o, G);
      gi[d] += __shfl_xor(gi[d], o, G);
    }
  if (q == 0) {
#pragma unroll
    for (int d = 0; d < DD; ++d)
      kbar[(size_t)d + (size_t)DD * jj] = (double)y.x * (double)gr[d] + (double)y.y * (double)gi[d];
  }
}

// Any D (used for D > 3): one thread per node.
template <typename R, int W>
__global__ __launch_bounds__(256) void k_grad_nodes(KArgs a, const typename cpx<R>::T* __restrict__ grid,
                                                    const typename cpx<R>::T* __restrict__ ybar,
                                                    double* __restrict__ kbar) {
  using C = typename cpx<R>::T;
  constexpr int DM = MAXD;
  const int D = a.D;
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= a.M) return;
  R w[DM][W], dw[DM][W];
  long long base[DM];
  for (int d = 0; d < D; ++d) {
    double wd[W], dwd[W];
    const int cw = grad_weights<W>(a, d, a.ks[(size_t)d * a.M + s], wd, dwd);
#pragma unroll
    for (int l = 0; l < W; ++l) {
      w[d][l] = (R)wd[l];
      dw[d][l] = (R)dwd[l];
    }
    base[d] = (long long)cw - a.m + 1;
  }
  int x0[W];
#pragma unroll
  for (int l = 0; l < W; ++l) x0[l] = (int)wrapidx(base[0] + l, a.Nt[0]);
  R gr[DM], gi[DM];
#pragma unroll
  for (int d = 0; d < DM; ++d) gr[d] = gi[d] = 0;
  {
    // odometer over the outer indices
    int l[DM];
    for (int d = 0; d < DM; ++d) l[d] = 0;
    while (true) {
      long long ofs = 0;
      for (int d = 1; d < D; ++d) ofs += wrapidx(base[d] + l[d], a.Nt[d]) * a.gstr[d];
      R sr, si, tr, ti;
      grad_row<R, W>(grid, ofs, x0, w[0], dw[0], sr, si, tr, ti);
      R wo = 1;
      for (int d = 1; d < D; ++d) wo *= w[d][l[d]];
      gr[0] = fma(tr, wo, gr[0]);
      gi[0] = fma(ti, wo, gi[0]);
      for (int d = 1; d < D; ++d) {
        R pw = dw[d][l[d]];
        for (int e = 1; e < D; ++e)
          if (e != d) pw *= w[e][l[e]];
        gr[d] = fma(sr, pw, gr[d]);
        gi[d] = fma(si, pw, gi[d]);
      }
      int d = 1;
      for (; d < D; ++d) {
        if (++l[d] < W) break;
        l[d] = 0;
      }
      if (d >= D) break;
    }
  }
  const int jj = a.perm[s];
  const C y = ybar[jj];
#pragma unroll
  for (int d = 0; d < DM; ++d)
    if (d < D) kbar[(size_t)d + (size_t)D * jj] = (double)y.x * (double)gr[d] + (double)y.y * (double)gi[d];
}

template <typename R, int W>
static hipError_t grad_d(const KArgs& a, hipStream_t s, const void* g, const void* ybar, double* kbar) {
  using C = typename cpx<R>::T;
  constexpr long long G = grad_group<W>();
  const dim3 lanes((unsigned)(((long long)a.M * G + 255) / 256)), blk(256);
  // one thread per node where a lane per tap would pass 2^31 threads
  switch ((long long)a.M * G < (1LL << 31) ? a.D : 0) {
    case 1:
      hipLaunchKernelGGL((k_grad_lanes<R, W, 1>), lanes, blk, 0, s, a, (const C*)g, (const C*)ybar, kbar);
      break;
    case 2:
      hipLaunchKernelGGL((k_grad_lanes<R, W, 2>), lanes, blk, 0, s, a, (const C*)g, (const C*)ybar, kbar);
      break;
    case 3:
      hipLaunchKernelGGL((k_grad_lanes<R, W, 3>), lanes, blk, 0, s, a, (const C*)g, (const C*)ybar, kbar);
      break;
    default:
      hipLaunchKernelGGL((k_grad_nodes<R, W>), dim3((a.M + 255) / 256), blk, 0, s, a, (const C*)g,
                         (const C*)ybar, kbar);
  }
  return hipGetLastError();
}

template <typename R>
static hipError_t grad_w(int W, const KArgs& a, hipStream_t s, const void* g, const void* ybar, double* kbar) {
  switch (W) {
    case 2: return grad_d<R, 2>(a, s, g, ybar, kbar);
    case 4: return grad_d<R, 4>(a, s, g, ybar, kbar);
    case 6: return grad_d<R, 6>(a, s, g, ybar, kbar);
    case 8: return grad_d<R, 8>(a, s, g, ybar, kbar);
    case 10: return grad_d<R, 10>(a, s, g, ybar, kbar);
    case 12: return grad_d<R, 12>(a, s, g, ybar, kbar);
    case 14: return grad_d<R, 14>(a, s, g, ybar, kbar);
    case 16: return grad_d<R, 16>(a, s, g, ybar, kbar);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_grad_nodes(bool dbl, int W, const KArgs& a, hipStream_t s, const void* g, const void* ybar,
                             double* kbar) {
  if (a.M == 0) return hipSuccess;
  return dbl ? grad_w<double>(W, a, s, g, ybar, kbar) : grad_w<float>(W, a, s, g, ybar, kbar);
}

}  // namespace nfftm
