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
// in the data precision.  One thread per node; the D-dimensional footprint is
// reduced row by row: for every outer index (l_1 .. l_{D-1}) the two dot
// products of the 2m-point row with w_0 and dw_0 are formed once and then
// weighted by the outer products.
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

// Window values and node derivatives of one node along dimension d, with the
// plan's power-of-two window normalisation; returns the wrapped cell.
template <int W>
__device__ __forceinline__ int grad_weights(const KArgs& a, int d, double ksv, double* w, double* dw) {
  const double Ntd = a.Ntd[d];
  const double kscale = scaled_node(ksv, Ntd);
  const long long c = (long long)kscale;
  const long long off = c - a.m + 1;
  const double t = kscale - (double)off;
  if (a.mode == WIN_POLY) {
    const double* P = a.polyd;
    const int nc = 2 * a.m + 1;
    const double x = t - (double)a.m + 0.5;
#pragma unroll
    for (int l = 0; l < W; ++l) {
      const double* cf = P + l * nc;
      double s = cf[nc - 1], ds = 0.0;
      for (int i = nc - 2; i >= 0; --i) {
        ds = fma(x, ds, s);
        s = fma(x, s, cf[i]);
      }
      w[l] = s;
      dw[l] = Ntd * ds;
    }
  } else {
#pragma unroll
    for (int l = 0; l < W; ++l) {
      double v, dv;
      kb_value_deriv(t - (double)l, (double)a.m, a.b, &v, &dv);
      w[l] = v * a.wscale;
      dw[l] = Ntd * dv * a.wscale;
    }
  }
  return (int)(c >= a.Nt[d] ? c - a.Nt[d] : c);
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

// v[i] for a runtime i of a register array (a select chain: no private
// memory for the array).
template <int W, typename R>
__device__ __forceinline__ R pick(const R* v, int i) {
  R r = v[0];
#pragma unroll
  for (int l = 1; l < W; ++l) r = i == l ? v[l] : r;
  return r;
}

template <typename R, int W, int DD>
__global__ __launch_bounds__(256) void k_grad_nodes(KArgs a, const typename cpx<R>::T* __restrict__ grid,
                                                    const typename cpx<R>::T* __restrict__ ybar,
                                                    double* __restrict__ kbar) {
  using C = typename cpx<R>::T;
  constexpr int DM = DD ? DD : MAXD;
  const int D = DD ? DD : a.D;
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= a.M) return;
  R w[DM][W], dw[DM][W];
  long long base[DM];
  // compile-time dimension loop for D <= 3 (register arrays), runtime beyond
#pragma unroll(DD ? DD : 1)
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
  if constexpr (DD == 1) {
    R sr, si;
    grad_row<R, W>(grid, 0, x0, w[0], dw[0], sr, si, gr[0], gi[0]);
  } else if constexpr (DD == 2) {
    // rows l_1: d/dk_0 takes (dw_0 . row) w_1, d/dk_1 takes (w_0 . row) dw_1
#pragma unroll 1
    for (int ly = 0; ly < W; ++ly) {
      R sr, si, tr, ti;
      grad_row<R, W>(grid, wrapidx(base[1] + ly, a.Nt[1]) * a.gstr[1], x0, w[0], dw[0], sr, si, tr, ti);
      const R w1 = pick<W>(w[1], ly), dw1 = pick<W>(dw[1], ly);
      gr[0] = fma(tr, w1, gr[0]);
      gi[0] = fma(ti, w1, gi[0]);
      gr[1] = fma(sr, dw1, gr[1]);
      gi[1] = fma(si, dw1, gi[1]);
    }
  } else if constexpr (DD == 3) {
    // planes l_2 of rows l_1: plane sums A = sum (dw_0 . row) w_1,
    // B = sum (w_0 . row) dw_1, E = sum (w_0 . row) w_1, then weighted by
    // w_2 (A, B) and dw_2 (E)
#pragma unroll 1
    for (int lz = 0; lz < W; ++lz) {
      const long long oz = wrapidx(base[2] + lz, a.Nt[2]) * a.gstr[2];
      R Ar = 0, Ai = 0, Br = 0, Bi = 0, Er = 0, Ei = 0;
#pragma unroll 1
      for (int ly = 0; ly < W; ++ly) {
        R sr, si, tr, ti;
        grad_row<R, W>(grid, oz + wrapidx(base[1] + ly, a.Nt[1]) * a.gstr[1], x0, w[0], dw[0], sr, si, tr, ti);
        const R w1 = pick<W>(w[1], ly), dw1 = pick<W>(dw[1], ly);
        Ar = fma(tr, w1, Ar);
        Ai = fma(ti, w1, Ai);
        Br = fma(sr, dw1, Br);
        Bi = fma(si, dw1, Bi);
        Er = fma(sr, w1, Er);
        Ei = fma(si, w1, Ei);
      }
      const R w2 = pick<W>(w[2], lz), dw2 = pick<W>(dw[2], lz);
      gr[0] = fma(Ar, w2, gr[0]);
      gi[0] = fma(Ai, w2, gi[0]);
      gr[1] = fma(Br, w2, gr[1]);
      gi[1] = fma(Bi, w2, gi[1]);
      gr[2] = fma(Er, dw2, gr[2]);
      gi[2] = fma(Ei, dw2, gi[2]);
    }
  } else {
    // any D: odometer over the outer indices
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
  const dim3 grid((a.M + 255) / 256), blk(256);
  switch (a.D) {
    case 1:
      hipLaunchKernelGGL((k_grad_nodes<R, W, 1>), grid, blk, 0, s, a, (const C*)g, (const C*)ybar, kbar);
      break;
    case 2:
      hipLaunchKernelGGL((k_grad_nodes<R, W, 2>), grid, blk, 0, s, a, (const C*)g, (const C*)ybar, kbar);
      break;
    case 3:
      hipLaunchKernelGGL((k_grad_nodes<R, W, 3>), grid, blk, 0, s, a, (const C*)g, (const C*)ybar, kbar);
      break;
    default:
      hipLaunchKernelGGL((k_grad_nodes<R, W, 0>), grid, blk, 0, s, a, (const C*)g, (const C*)ybar, kbar);
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
