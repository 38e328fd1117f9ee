// nfft_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A plain FP64 C++ restatement of NFFT.jl's CPU plan (src/*.jl) and of its
// direct NDFT oracle (src/direct.jl).  It is the checker that the HIP library
// is compared against and the CPU baseline that bench.py times.  Nothing in the
// shipped library links, calls or includes this file.
//
// Every function cites the reference file:line it restates.  Layouts follow
// Julia: arrays are column-major with dimension 1 fastest; node coordinates are
// a D x M column-major matrix (node j's coordinates contiguous); complex values
// are interleaved (re, im) doubles.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;
typedef int64_t i64;

namespace {

constexpr int MAXD = 8;

// --------------------------------------------------------------------------
// Window functions: src/windowFunctions.jl:4-19 (getWindow) and the pairs
// (psi, psi_hat) at :21-134.  Window ids: 0 kaiser_bessel, 1 gauss,
// 2 spline, 3 kaiser_bessel_rev, 4 cosh_type.
// --------------------------------------------------------------------------

// Modified Bessel function I0 by its power series sum_k (x^2/4)^k / (k!)^2.
// All terms are positive, so the sum is accurate to a few ulps.
double bessel_i0(double x) {
  double q = 0.25 * x * x, term = 1.0, s = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / ((double)k * (double)k);
    s += term;
    if (term < 1e-18 * s) break;
  }
  return s;
}

// src/windowFunctions.jl:21-34
double window_kb(double k, double Nt, int m, double sigma) {
  double m_by_N = m / Nt;
  double b = M_PI * (2.0 - 1.0 / sigma);
  double y;
  if (std::fabs(k) < m_by_N) {
    double arg = std::sqrt((double)(m * m) - (Nt * Nt) * (k * k));
    double arg_times_pi = arg * M_PI;
    y = std::sinh(b * arg) / arg_times_pi;
  } else if (std::fabs(k) > m_by_N) {
    y = 0.0;
  } else {
    y = b / M_PI;
  }
  return y;
}

// src/windowFunctions.jl:36-39
double window_kb_hat(double n, double Nt, int m, double sigma) {
  double b = M_PI * (2.0 - 1.0 / sigma);
  double t = 2.0 * M_PI * n / Nt;
  return bessel_i0(m * std::sqrt(b * b - t * t));
}

// Bessel function J0 by its power series sum_k (-x^2/4)^k / (k!)^2 in long
// double (64-bit mantissa): the alternating terms lose ~log10 of the largest
// term, < 5 digits for the arguments of the cosh-type transform (|x| <= 13).
double bessel_j0(double x) {
  long double q = -0.25L * (long double)x * (long double)x, term = 1.0L, s = 1.0L;
  for (int k = 1; k < 400; ++k) {
    term *= q / ((long double)k * (long double)k);
    s += term;
    if (k > std::fabs(x) && std::fabs((double)term) < 1e-24 * std::fabs((double)s)) break;
  }
  return (double)s;
}

// Cardinal B-spline of order n at x (src/windowFunctions.jl:65-77 cbspline):
// the same recursion, tabulated over the shifts it visits.
double cbspline(int n, double x) {
  std::vector<double> v(n);
  for (int s = 0; s < n; ++s) v[s] = (x - s >= 0.0 && x - s < 1.0) ? 1.0 : 0.0;
  for (int k = 2; k <= n; ++k)
    for (int s = 0; s + k <= n; ++s)
      v[s] = (x - s) / (k - 1) * v[s] + (k - (x - s)) / (k - 1) * v[s + 1];
  return v[0];
}

inline double jl_sinc(double x) { return x == 0.0 ? 1.0 : std::sin(M_PI * x) / (M_PI * x); }

// psi(k) of window `win` (k in units of the period, support |k| < m / Nt).
double window_psi(int win, double k, double Nt, int m, double sigma) {
  const double ak = std::fabs(k), mN = m / Nt;
  switch (win) {
    case 0:
      return window_kb(k, Nt, m, sigma);
    case 1: {  // :54-62
      const double b = m / M_PI;
      return ak < mN ? 1.0 / std::sqrt(M_PI * b) * std::exp(-(Nt * k) * (Nt * k) / b) : 0.0;
    }
    case 2:  // :79-86
      return ak < mN ? cbspline(2 * m, Nt * k + m) : 0.0;
    case 3: {  // :41-49
      const double b = M_PI * (2.0 - 1.0 / sigma);
      if (!(ak < mN)) return 0.0;
      const double r = Nt * k / m;
      return 0.5 / m * bessel_i0(m * b * std::sqrt(1.0 - r * r));
    }
    case 4: {  // :95-107
      const double beta = M_PI * m * (2.0 - 1.0 / sigma);
      if (!(ak < mN)) return 0.0;
      const double arg = Nt * k / m, alpha = std::sqrt(1.0 - arg * arg);
      return 1.0 / (std::cosh(beta) - 1.0) * (std::cosh(beta * alpha) - 1.0) / alpha;
    }
  }
  return std::nan("");
}

// psi_hat(n) of window `win`.
double window_hat(int win, double n, double Nt, int m, double sigma) {
  switch (win) {
    case 0:
      return window_kb_hat(n, Nt, m, sigma);
    case 1: {  // :64-67
      const double b = m / M_PI, t = M_PI * n / Nt;
      return std::exp(-t * t * b);
    }
    case 2:  // :88-90
      return std::pow(jl_sinc(n / Nt), 2 * m);
    case 3: {  // :51-56: real(sinc(sqrt(complex(x)) / pi))
      const double b = M_PI * (2.0 - 1.0 / sigma);
      const double t = 2.0 * M_PI * m * n / Nt, x = t * t - (m * b) * (m * b);
      if (x >= 0.0) {
        const double a = std::sqrt(x);
        return a == 0.0 ? 1.0 : std::sin(a) / a;
      }
      const double y = std::sqrt(-x);
      return std::sinh(y) / y;
    }
    case 4: {  // :109-127
      const double beta = M_PI * m * (2.0 - 1.0 / sigma), gam = beta / (2.0 * M_PI);
      const double zeta = M_PI / (std::cosh(beta) - 1.0) * m, arg = m * n / Nt;
      const double w = 2.0 * M_PI * arg;
      if (std::fabs(arg) < gam) return zeta * (bessel_i0(std::sqrt(beta * beta - w * w)) - bessel_j0(w));
      if (std::fabs(arg) > gam) return zeta * (bessel_j0(std::sqrt(w * w - beta * beta)) - bessel_j0(w));
      return zeta * (1.0 - bessel_j0(beta));
    }
  }
  return std::nan("");
}

// Derivative d psi / du of the Kaiser-Bessel window of src/windowFunctions.jl:
// 21-34 as a function of u = Nt * k (grid units).  With s = sqrt(m^2 - u^2),
// x = b s:  psi = sinh(x) / (pi s)  and
//   d psi / du = -u (x cosh x - sinh x) / (pi s^3)
//              = -u b^3 / pi * sum_{n>=1} x^(2n-2) 2n / (2n+1)!,
// the series used for x < 1/2 where the closed form cancels.  Zero for
// |u| >= m (the reference window is 0 beyond m; its value b/pi at |u| = m is
// a single point).  Not in the reference: the node gradient is an addition.
double window_kb_du(double u, int m, double sigma) {
  double b = M_PI * (2.0 - 1.0 / sigma);
  double au = std::fabs(u);
  if (au >= (double)m) return 0.0;
  double s = std::sqrt((double)(m * m) - u * u);
  double x = b * s, q;
  if (x < 0.5) {
    double x2 = x * x, term = 1.0 / 3.0, sum = term;
    for (int n = 1; n < 12; ++n) {
      term *= x2 / (2.0 * n * (2.0 * n + 3.0));
      sum += term;
    }
    q = b * b * b * sum;
  } else {
    q = (x * std::cosh(x) - std::sinh(x)) / (s * s * s);
  }
  return -u * q / M_PI;
}

// --------------------------------------------------------------------------
// Parameters: src/precomputation.jl:3-56 (initParams), :59-77 (_blockSize).
// --------------------------------------------------------------------------
i64 lut_size(int m) {
  static const int m2K[9] = {1, 3, 7, 9, 14, 17, 20, 23, 24};  // :15
  int K = m2K[std::min(m + 1, 9) - 1];                           // :16
  return ((i64)1 << K) * m;                                      // :17
}

i64 block_size(int D, const i64* Nt, int d) {
  if (D == 1) return std::min<i64>(1024, Nt[d]);  // :59-61
  if (D == 2) return std::min<i64>(64, Nt[d]);    // :63-65
  if (d < 3) return std::min<i64>(16, Nt[d]);     // :67-73
  return 1;                                       // :74-75
}

// --------------------------------------------------------------------------
// Tables
// --------------------------------------------------------------------------

// src/precomputation.jl:291-300 (precomputeLinInterp): K+2 samples of the
// positive half of the window, step m/K, evaluated with Ntilde = 1.
void lin_interp_table(int win, int m, double sigma, i64 K, double* out) {
  double step = (double)m / (double)K;
  for (i64 l = 0; l < K + 2; ++l) out[l] = window_psi(win, (double)l * step, 1.0, m, sigma);
}

// Least squares min ||A c - y||_2 by Householder QR (A is rows x cols,
// column-major, rows >= cols).  Used for src/precomputation.jl:317 (V' \ s).
void lstsq(int rows, int cols, std::vector<double> A, std::vector<double> y, double* c) {
  for (int j = 0; j < cols; ++j) {
    double nrm = 0;
    for (int i = j; i < rows; ++i) nrm += A[i + j * rows] * A[i + j * rows];
    nrm = std::sqrt(nrm);
    double alpha = A[j + j * rows] > 0 ? -nrm : nrm;
    std::vector<double> v(rows, 0.0);
    for (int i = j; i < rows; ++i) v[i] = A[i + j * rows];
    v[j] -= alpha;
    double vn = 0;
    for (int i = j; i < rows; ++i) vn += v[i] * v[i];
    if (vn == 0) continue;
    for (int jj = j; jj < cols; ++jj) {
      double s = 0;
      for (int i = j; i < rows; ++i) s += v[i] * A[i + jj * rows];
      s = 2 * s / vn;
      for (int i = j; i < rows; ++i) A[i + jj * rows] -= s * v[i];
    }
    double s = 0;
    for (int i = j; i < rows; ++i) s += v[i] * y[i];
    s = 2 * s / vn;
    for (int i = j; i < rows; ++i) y[i] -= s * v[i];
  }
  for (int j = cols - 1; j >= 0; --j) {
    double s = y[j];
    for (int jj = j + 1; jj < cols; ++jj) s -= A[j + jj * rows] * c[jj];
    c[j] = s / A[j + j * rows];
  }
}

// src/precomputation.jl:302-320 (precomputePolyInterp).  Output P is
// (2m+1) x 2m column-major: column l holds the monomial coefficients
// (degree 0..2m) of the fit of psi(m - (l - 1/2) + x), x in [-1/2, 1/2].
void poly_interp_table(int win, int m, double sigma, double* P) {
  int deg = 2 * m + 1, K = 2 * m, NS = 2 * deg;
  std::vector<double> k(NS);
  for (int s = 0; s < NS; ++s)  // range(-0.5, 0.5, length=NS), correctly rounded
    k[s] = (double)(2 * s - (NS - 1)) / (double)(2 * (NS - 1));
  // A = V' (NS x deg), V[r, s] = k_s^(r-1) built by repeated products (:309-312)
  std::vector<double> A(NS * deg);
  for (int s = 0; s < NS; ++s) {
    double v = 1.0;
    for (int r = 0; r < deg; ++r) {
      A[s + r * NS] = v;
      v = v * k[s];
    }
  }
  for (int l = 1; l <= K; ++l) {
    std::vector<double> y(NS);
    double c0 = -((double)l - 0.5) + m;
    for (int s = 0; s < NS; ++s) y[s] = window_psi(win, c0 + k[s], 1.0, m, sigma);
    lstsq(NS, deg, A, y, P + (l - 1) * deg);
  }
}

// src/precomputation.jl:345
i64 index_offset(i64 N) { return (N % 2 == 0) ? (-1 - N / 2) : (-1 - (N - 1) / 2); }

// Chebyshev interpolant of the first kind with n nodes on [a, b], the
// construction named at src/precomputation.jl:351
// (BasicInterpolators.ChebyshevInterpolator(kappa, 1, N, 30)).
struct Cheb {
  double a, b;
  std::vector<double> c;
  template <class F>
  Cheb(F f, double a_, double b_, int n) : a(a_), b(b_), c(n) {
    std::vector<double> fv(n);
    for (int k = 0; k < n; ++k) {
      double x = std::cos(M_PI * (k + 0.5) / n);
      fv[k] = f(0.5 * (a + b) + 0.5 * (b - a) * x);
    }
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int k = 0; k < n; ++k) s += fv[k] * std::cos(M_PI * j * (k + 0.5) / n);
      c[j] = 2.0 * s / n;
    }
    c[0] *= 0.5;
  }
  double operator()(double t) const {
    double x = (2.0 * t - a - b) / (b - a);
    double b1 = 0, b2 = 0;
    for (int j = (int)c.size() - 1; j >= 1; --j) {
      double t0 = 2 * x * b1 - b2 + c[j];
      b2 = b1;
      b1 = t0;
    }
    return x * b1 - b2 + c[0];
  }
};

// src/precomputation.jl:347-358 (precomputeWindowHatInvLUT), one dimension.
// cheb != 0: the reference's 30-node Chebyshev interpolant; cheb == 0: the
// exact 1/psi_hat(n).
void deconv_factors(int win, i64 N, i64 Nt, int m, double sigma, int cheb, double* out) {
  i64 off = index_offset(N);
  auto kappa = [&](double n) { return window_hat(win, n + (double)off, (double)Nt, m, sigma); };
  if (cheb && N > 1) {
    Cheb ch(kappa, 1.0, (double)N, 30);
    for (i64 j = 1; j <= N; ++j) out[j - 1] = 1.0 / ch((double)j);
  } else {
    for (i64 j = 1; j <= N; ++j) out[j - 1] = 1.0 / kappa((double)j);
  }
}

// Exact phase fraction of (k * n) mod 1 using an FMA two-product, so the
// NDFT phase error does not grow with |n|.
inline double frac_prod(double k, double n) {
  double p = k * n;
  double e = std::fma(k, n, -p);
  double r = p - std::nearbyint(p);
  return r + e;
}

}  // namespace

// ==========================================================================
// The plan: src/implementation.jl:16-43 (NFFTPlan), :73-106 (constructor).
// ==========================================================================
struct OPlan {
  int D, m, precompute, blocking, win = 0;
  i64 N[MAXD], Nt[MAXD], M, P, Q, LUTSize;
  double sigma;
  std::vector<double> k;  // D x M, unshifted (the plan's k)
  std::vector<double> lut, poly;
  std::vector<std::vector<double>> deconv;
  // blocking: src/precomputation.jl:459-554
  i64 bs[MAXD], nb[MAXD], nblocks;
  std::vector<std::vector<i64>> nodesInBlock;
  std::vector<std::vector<i64>> blockOffsets;           // per block, D offsets
  std::vector<std::vector<std::pair<i64, double>>> idxInBlock;  // per block, D x n
  std::vector<std::vector<double>> windowTensor;         // per block, 2m x D x n
};

namespace {

enum { FULL = 1, TENSOR = 2, LINEAR = 3, POLYNOMIAL = 4 };

// src/utils.jl:32-44 (shiftNodes!)
void shift_nodes(std::vector<double>& k) {
  for (auto& x : k) {
    if (x < 0.0) x += 1.0;
    if (x == 1.0) x -= std::numeric_limits<double>::epsilon();
  }
}

// Horner evaluation as Julia's evalpoly (muladd -> fma on FMA hardware).
inline double evalpoly(double x, const double* c, int n) {
  double s = c[n - 1];
  for (int i = n - 2; i >= 0; --i) s = std::fma(x, s, c[i]);
  return s;
}

// src/precomputation.jl:179-201 (shiftedWindowEntries, LUT lerp)
inline void lin_weights(const OPlan& p, double idx, i64 scale, double* w) {
  int Z = 2 * p.m;
  i64 idxInt = (i64)std::floor(idx);
  double alpha = idx - (double)idxInt;
  for (int l = 0; l < Z; ++l) {
    i64 i1 = std::llabs(idxInt - (i64)l * scale);
    i64 i2 = std::llabs(idxInt - (i64)l * scale + 1);
    w[l] = p.lut[i1] + alpha * (p.lut[i2] - p.lut[i1]);
  }
}

// src/precomputation.jl:215-222 (shiftedWindowEntries, polynomial)
inline void poly_weights(const OPlan& p, double x, double* w) {
  int Z = 2 * p.m, deg = 2 * p.m + 1;
  for (int l = 0; l < Z; ++l) w[l] = evalpoly(x, &p.poly[l * deg], deg);
}

void precompute_blocks(OPlan& p) {
  // src/precomputation.jl:459-554 and :556-591
  std::vector<double> xs = p.k;
  shift_nodes(xs);
  int D = p.D, m = p.m;
  p.nblocks = 1;
  for (int d = 0; d < D; ++d) {
    p.bs[d] = block_size(D, p.Nt, d);
    p.nb[d] = (p.Nt[d] + p.bs[d] - 1) / p.bs[d];
    p.nblocks *= p.nb[d];
  }
  p.nodesInBlock.assign(p.nblocks, {});
  auto block_of = [&](i64 j) {
    i64 lin = 0, str = 1;
    for (int d = 0; d < D; ++d) {
      i64 c = (i64)(xs[d + D * j] * (double)p.Nt[d]);  // unsafe_trunc (:495)
      i64 bi = c / p.bs[d];
      lin += bi * str;
      str *= p.nb[d];
    }
    return lin;
  };
  for (i64 j = 0; j < p.M; ++j) p.nodesInBlock[block_of(j)].push_back(j);  // :494-504
  p.blockOffsets.assign(p.nblocks, {});
  p.idxInBlock.assign(p.nblocks, {});
  p.windowTensor.assign(p.nblocks, {});
  i64 scale = p.LUTSize / m;
  for (i64 l = 0; l < p.nblocks; ++l) {
    if (p.nodesInBlock[l].empty()) continue;
    std::vector<i64> off(D);
    i64 r = l;
    for (int d = 0; d < D; ++d) {
      i64 ld = r % p.nb[d];
      r /= p.nb[d];
      off[d] = ld * p.bs[d] - m - 1;  // (l[d]-1)*blockSize[d]-padding[d]-1 (:515)
    }
    p.blockOffsets[l] = off;
    auto& ib = p.idxInBlock[l];
    i64 n = p.nodesInBlock[l].size();
    ib.resize(D * n);
    if (p.precompute == TENSOR) p.windowTensor[l].resize(2 * m * D * n);
    for (i64 i = 0; i < n; ++i) {
      i64 j = p.nodesInBlock[l][i];
      for (int d = 0; d < D; ++d) {
        double kscale = xs[d + D * j] * (double)p.Nt[d];            // :538
        i64 offn = (i64)kscale - m + 1;                              // :539
        i64 y = offn - off[d] - 1;                                   // :540
        double idx;
        if (p.precompute == LINEAR)
          idx = (kscale - (double)offn) * (double)scale;             // :543
        else
          idx = (kscale - (double)offn - (double)m + 1.0 - 0.5);     // :545
        ib[d + D * i] = {y, idx};
        if (p.precompute == TENSOR) {                                // :577-584
          double kk = (kscale - (double)offn - (double)m + 1.0 - 0.5);
          poly_weights(p, kk, &p.windowTensor[l][2 * m * (d + D * i)]);
        }
      }
    }
  }
}

// Weights and (0-based) grid indices of one node along dimension d, for the
// non-blocked and FULL paths: src/precomputation.jl:124-166.
void node_weights_nonblocked(const OPlan& p, i64 j, int d, i64* idx, double* w) {
  int m = p.m, Z = 2 * m;
  i64 Nt = p.Nt[d];
  double kscale = p.k[d + p.D * j] * (double)Nt;
  i64 off = (i64)std::floor(kscale) - m + 1;
  for (int l = 0; l < Z; ++l) idx[l] = (((i64)l + off + Nt) % Nt + Nt) % Nt;
  if (p.precompute == FULL) {  // :124-134
    for (int l = 0; l < Z; ++l)
      w[l] = window_psi(p.win, (kscale - (double)l - (double)off) / (double)Nt, (double)Nt, m, p.sigma);
  } else if (p.precompute == LINEAR) {  // :138-150
    double idxf = ((kscale - (double)off) * (double)p.LUTSize) / (double)m;
    lin_weights(p, idxf, p.LUTSize / m, w);
  } else {  // POLYNOMIAL / TENSOR without blocks (:154-166)
    double x = (kscale - (double)off - (double)m + 0.5);
    poly_weights(p, x, w);
  }
}

// Block-local weights: src/precomputation.jl:170-271
void node_weights_blocked(const OPlan& p, i64 l, i64 i, int d, i64* y, double* w) {
  auto pr = p.idxInBlock[l][d + p.D * i];
  *y = pr.first;
  if (p.precompute == LINEAR) {
    lin_weights(p, pr.second, p.LUTSize / p.m, w);
  } else if (p.precompute == TENSOR) {
    const double* t = &p.windowTensor[l][2 * p.m * (d + p.D * i)];
    for (int q = 0; q < 2 * p.m; ++q) w[q] = t[q];
  } else {
    poly_weights(p, pr.second, w);
  }
}

// Window values w and their node derivatives dw = d w / d k (grid index
// arithmetic as node_weights_nonblocked) for the node gradient.  w_l depends on
// k through u_l = k Nt - l - off with off locally constant, so
// dw_l = Nt psi'(u_l).  POLYNOMIAL / TENSOR: w_l = P_l(x), x = k Nt - off - m
// + 1/2 (src/precomputation.jl:215-222), dw_l = Nt P_l'(x) by Horner.  FULL and
// LINEAR: the exact window and its analytic derivative (the LINEAR table
// samples the exact window).
void node_weights_grad(const OPlan& p, i64 j, int d, i64* idx, double* w, double* dw) {
  int m = p.m, Z = 2 * m;
  i64 Nt = p.Nt[d];
  double kscale = p.k[d + p.D * j] * (double)Nt;
  i64 off = (i64)std::floor(kscale) - m + 1;
  for (int l = 0; l < Z; ++l) idx[l] = (((i64)l + off + Nt) % Nt + Nt) % Nt;
  if (p.precompute == TENSOR || p.precompute == POLYNOMIAL) {
    int deg = 2 * m + 1;
    double x = (kscale - (double)off - (double)m + 0.5);
    for (int l = 0; l < Z; ++l) {
      const double* c = &p.poly[l * deg];
      double s = c[deg - 1], ds = 0.0;
      for (int i = deg - 2; i >= 0; --i) {
        ds = std::fma(x, ds, s);
        s = std::fma(x, s, c[i]);
      }
      w[l] = s;
      dw[l] = (double)Nt * ds;
    }
  } else {
    for (int l = 0; l < Z; ++l) {
      double u = kscale - (double)l - (double)off;
      w[l] = window_kb(u / (double)Nt, (double)Nt, m, p.sigma);
      dw[l] = p.win == 0 ? (double)Nt * window_kb_du(u, m, p.sigma) : std::nan("");
    }
  }
}

// Row-major odometer over a D-dimensional box of extent Z (dim 0 fastest).
inline bool next_multi(int D, int* l, int Z) {
  for (int d = 0; d < D; ++d) {
    if (++l[d] < Z) return true;
    l[d] = 0;
  }
  return false;
}

// Per-dimension wrapped grid index of each padded-block position: block
// position l (0-based) maps to g index mod(l + 1 + off + Nt, Nt) (0-based),
// the index arithmetic of toBlock!/addBlock! (src/convolution.jl:248-299,
// :379-443), whose split loops only avoid the modulo.
inline void block_maps(const OPlan& p, const std::vector<i64>& off, const i64* bsz,
                       std::vector<i64>* maps) {
  for (int d = 0; d < p.D; ++d) {
    maps[d].resize(bsz[d]);
    i64 str = 1;
    for (int e = 0; e < d; ++e) str *= p.Nt[e];
    for (i64 l = 0; l < bsz[d]; ++l)
      maps[d][l] = ((((l + 1 + off[d] + p.Nt[d]) % p.Nt[d]) + p.Nt[d]) % p.Nt[d]) * str;
  }
}

// toBlock! (src/convolution.jl:248-299)
void to_block(const OPlan& p, const cd* g, const std::vector<i64>& off, std::vector<cd>& block,
              const i64* bsz) {
  int D = p.D;
  std::vector<i64> maps[MAXD];
  block_maps(p, off, bsz, maps);
  i64 n = 1;
  for (int d = 0; d < D; ++d) n *= bsz[d];
  block.resize(n);
  i64 rows = n / bsz[0];
  for (i64 r = 0; r < rows; ++r) {
    i64 q = r, base = 0;
    for (int d = 1; d < D; ++d) {
      base += maps[d][q % bsz[d]];
      q /= bsz[d];
    }
    cd* dst = &block[r * bsz[0]];
    for (i64 l1 = 0; l1 < bsz[0]; ++l1) dst[l1] = g[base + maps[0][l1]];
  }
}

// addBlock! (src/convolution.jl:379-443)
void add_block(const OPlan& p, cd* g, const std::vector<i64>& off, const std::vector<cd>& block,
               const i64* bsz) {
  int D = p.D;
  std::vector<i64> maps[MAXD];
  block_maps(p, off, bsz, maps);
  i64 n = block.size();
  i64 rows = n / bsz[0];
  for (i64 r = 0; r < rows; ++r) {
    i64 q = r, base = 0;
    for (int d = 1; d < D; ++d) {
      base += maps[d][q % bsz[d]];
      q /= bsz[d];
    }
    const cd* src = &block[r * bsz[0]];
    for (i64 l1 = 0; l1 < bsz[0]; ++l1) g[base + maps[0][l1]] += src[l1];
  }
}


// calcOneNode! (src/convolution.jl:315-344): nested partial sums
// fHat = sum_lD w_D * ( ... sum_l2 w_2 * ( sum_l1 w_1 * block ) ), dim 1
// innermost.  ZC is the compile-time stencil width (0 = runtime).
template <int ZC>
cd calc_one_node(int D, int Zr, const double* w, const i64* y, const cd* block, const i64* bsz) {
  const int Z = ZC ? ZC : Zr;
  if (D == 1) {
    cd s1(0, 0);
    for (int l1 = 0; l1 < Z; ++l1) s1 += w[l1] * block[y[0] + l1];
    return s1;
  }
  if (D == 2) {
    cd s2(0, 0);
    for (int l2 = 0; l2 < Z; ++l2) {
      const cd* row = block + (y[1] + l2) * bsz[0] + y[0];
      cd s1(0, 0);
      for (int l1 = 0; l1 < Z; ++l1) s1 += w[l1] * row[l1];
      s2 += w[Z + l2] * s1;
    }
    return s2;
  }
  if (D == 3) {
    cd s3(0, 0);
    for (int l3 = 0; l3 < Z; ++l3) {
      cd s2(0, 0);
      for (int l2 = 0; l2 < Z; ++l2) {
        const cd* row = block + ((y[2] + l3) * bsz[1] + (y[1] + l2)) * bsz[0] + y[0];
        cd s1(0, 0);
        for (int l1 = 0; l1 < Z; ++l1) s1 += w[l1] * row[l1];
        s2 += w[Z + l2] * s1;
      }
      s3 += w[2 * Z + l3] * s2;
    }
    return s3;
  }
  // D >= 4: same nesting, written recursively over the outer dimension
  struct Rec {
    int D, Z;
    const double* w;
    const i64* y;
    const cd* block;
    const i64* bsz;
    cd go(int d, i64 base) const {  // base: linear offset of dims > d
      if (d == 0) {
        cd s1(0, 0);
        for (int l1 = 0; l1 < Z; ++l1) s1 += w[l1] * block[base + y[0] + l1];
        return s1;
      }
      i64 str = 1;
      for (int e = 0; e < d; ++e) str *= bsz[e];
      cd s(0, 0);
      for (int l = 0; l < Z; ++l) s += w[d * Z + l] * go(d - 1, base + (y[d] + l) * str);
      return s;
    }
  } r{D, Z, w, y, block, bsz};
  return r.go(D - 1, 0);
}

template <int ZC>
void conv_blocked(const OPlan& p, const cd* g, cd* fhat, const i64* bsz) {
  int D = p.D, Z = 2 * p.m;
#pragma omp parallel
  {
    std::vector<cd> block;
    std::vector<double> w(D * Z);
#pragma omp for schedule(dynamic)
    for (i64 l = 0; l < p.nblocks; ++l) {  // @cthreads over blocks (:238)
      if (p.nodesInBlock[l].empty()) continue;
      to_block(p, g, p.blockOffsets[l], block, bsz);  // :240
      for (size_t i = 0; i < p.nodesInBlock[l].size(); ++i) {
        i64 y[MAXD];
        for (int d = 0; d < D; ++d) node_weights_blocked(p, l, i, d, &y[d], &w[d * Z]);
        fhat[p.nodesInBlock[l][i]] = calc_one_node<ZC>(D, Z, w.data(), y, block.data(), bsz);
      }
    }
  }
}

// fillOneNode! (src/convolution.jl:465-492): innerWin = w_1 .* fHat_j,
// prodWin_1 = w_D*...*w_2, block[...] += innerWin[l1] * prodWin_1.
template <int ZC>
void fill_one_node(int D, int Zr, const double* w, const i64* y, cd fh, cd* block, const i64* bsz) {
  const int Z = ZC ? ZC : Zr;
  cd inner[16];
  for (int l1 = 0; l1 < Z; ++l1) inner[l1] = w[l1] * fh;
  if (D == 1) {
    for (int l1 = 0; l1 < Z; ++l1) block[y[0] + l1] += inner[l1] * 1.0;
    return;
  }
  if (D == 2) {
    for (int l2 = 0; l2 < Z; ++l2) {
      double pw = 1.0 * w[Z + l2];
      cd* row = block + (y[1] + l2) * bsz[0] + y[0];
      for (int l1 = 0; l1 < Z; ++l1) row[l1] += inner[l1] * pw;
    }
    return;
  }
  if (D == 3) {
    for (int l3 = 0; l3 < Z; ++l3) {
      double p3 = 1.0 * w[2 * Z + l3];
      for (int l2 = 0; l2 < Z; ++l2) {
        double pw = p3 * w[Z + l2];
        cd* row = block + ((y[2] + l3) * bsz[1] + (y[1] + l2)) * bsz[0] + y[0];
        for (int l1 = 0; l1 < Z; ++l1) row[l1] += inner[l1] * pw;
      }
    }
    return;
  }
  int ll[MAXD] = {0};  // odometer over dims 2..D (0-based 1..D-1)
  while (true) {
    double pw = 1.0;
    for (int d = D - 1; d >= 1; --d) pw = pw * w[d * Z + ll[d]];
    i64 base = 0, str = 1;
    for (int e = 1; e < D; ++e) {
      str *= bsz[e - 1];
      base += (y[e] + ll[e]) * str;
    }
    for (int l1 = 0; l1 < Z; ++l1) block[base + y[0] + l1] += inner[l1] * pw;
    int d = 1;
    for (; d < D; ++d) {
      if (++ll[d] < Z) break;
      ll[d] = 0;
    }
    if (d == D) break;
  }
}

template <int ZC>
void spread_blocked(const OPlan& p, const cd* fhat, cd* g, const i64* bsz) {
  int D = p.D, Z = 2 * p.m;
  i64 bn = 1;
  for (int d = 0; d < D; ++d) bn *= bsz[d];
  std::mutex lk;  // ReentrantLock (:361)
#pragma omp parallel
  {
    std::vector<cd> block(bn);
    std::vector<double> w(D * Z);
#pragma omp for schedule(dynamic)
    for (i64 l = 0; l < p.nblocks; ++l) {
      if (p.nodesInBlock[l].empty()) continue;
      std::fill(block.begin(), block.end(), cd(0, 0));  // memset (:365)
      for (size_t i = 0; i < p.nodesInBlock[l].size(); ++i) {
        i64 y[MAXD];
        for (int d = 0; d < D; ++d) node_weights_blocked(p, l, i, d, &y[d], &w[d * Z]);
        fill_one_node<ZC>(D, Z, w.data(), y, fhat[p.nodesInBlock[l][i]], block.data(), bsz);
      }
      std::lock_guard<std::mutex> guard(lk);  // lock(lk) do addBlock! (:371-373)
      add_block(p, g, p.blockOffsets[l], block, bsz);
    }
  }
}

}  // namespace

extern "C" {

double orc_bessel_i0(double x) { return bessel_i0(x); }
double orc_window(double k, double Nt, int m, double sigma) { return window_kb(k, Nt, m, sigma); }
double orc_window_hat(double n, double Nt, int m, double sigma) { return window_kb_hat(n, Nt, m, sigma); }
double orc_window_w(int win, double k, double Nt, int m, double sigma) { return window_psi(win, k, Nt, m, sigma); }
double orc_window_hat_w(int win, double n, double Nt, int m, double sigma) {
  return window_hat(win, n, Nt, m, sigma);
}
double orc_bessel_j0(double x) { return bessel_j0(x); }
long long orc_lut_size(int m) { return lut_size(m); }
void orc_lut(int m, double sigma, long long K, double* out) { lin_interp_table(0, m, sigma, K, out); }
void orc_poly(int m, double sigma, double* out) { poly_interp_table(0, m, sigma, out); }
void orc_deconv_factors(long long N, long long Nt, int m, double sigma, int cheb, double* out) {
  deconv_factors(0, N, Nt, m, sigma, cheb, out);
}
int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// src/implementation.jl:73-106 with initParams (src/precomputation.jl:3-56).
// Returns NULL on invalid nodes (src/utils.jl:46-55).
OPlan* orc_plan_create(int D, const long long* N, long long M, const double* k, int m, double sigma,
                       int precompute, int blocking, int cheb_deconv, int window) {
  if (window < 0 || window > 4) return nullptr;
  if (D < 1 || D > MAXD) return nullptr;
  for (long long j = 0; j < (long long)D * M; ++j)
    if (!(std::fabs(k[j]) <= 0.5)) return nullptr;
  OPlan* p = new OPlan();
  p->D = D;
  p->m = m;
  p->precompute = precompute;
  p->blocking = blocking;
  p->win = window;
  p->M = M;
  p->P = 1;
  p->Q = 1;
  for (int d = 0; d < D; ++d) {
    p->N[d] = N[d];
    p->Nt[d] = ((i64)std::ceil(sigma * (double)N[d]) / 2) * 2;  // :25-27
    p->P *= N[d];
    p->Q *= p->Nt[d];
  }
  p->sigma = (double)p->Nt[0] / (double)p->N[0];  // :29
  p->LUTSize = lut_size(m);
  p->k.assign(k, k + (size_t)D * M);
  if (precompute == LINEAR) {
    p->lut.resize(p->LUTSize + 2);
    lin_interp_table(p->win, m, p->sigma, p->LUTSize, p->lut.data());
  }
  if (precompute == TENSOR || precompute == POLYNOMIAL) {
    p->poly.resize((2 * m + 1) * 2 * m);
    poly_interp_table(p->win, m, p->sigma, p->poly.data());
  }
  p->deconv.resize(D);
  for (int d = 0; d < D; ++d) {
    p->deconv[d].resize(p->N[d]);
    deconv_factors(p->win, p->N[d], p->Nt[d], m, p->sigma, cheb_deconv, p->deconv[d].data());
  }
  // calcBlocks: src/implementation.jl:182-187
  if (blocking && precompute != FULL) precompute_blocks(*p);
  return p;
}

void orc_plan_destroy(OPlan* p) { delete p; }

void orc_plan_info(OPlan* p, long long* Nt, long long* lutsize, double* sigma) {
  for (int d = 0; d < p->D; ++d) Nt[d] = p->Nt[d];
  *lutsize = p->LUTSize;
  *sigma = p->sigma;
}

// src/deconvolution.jl:2-45 (deconvolve!) after fill!(tmpVec, 0)
// (src/implementation.jl:159).  f: P complex, g: Q complex (overwritten).
void orc_deconvolve(OPlan* p, const double* f_, double* g_) {
  const cd* f = (const cd*)f_;
  cd* g = (cd*)g_;
  std::fill(g, g + p->Q, cd(0, 0));
  int D = p->D;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < p->P; ++i) {
    i64 r = i, gi = 0, str = 1;
    cd v = f[i];
    for (int d = 0; d < D; ++d) {
      i64 id = r % p->N[d];
      r /= p->N[d];
      i64 Na = p->N[d] / 2;
      i64 gd = (id < Na) ? id - Na + p->Nt[d] : id - Na;  // :30-42
      gi += gd * str;
      str *= p->Nt[d];
      v *= p->deconv[d][id];
    }
    g[gi] = v;
  }
}

// src/deconvolution.jl:49-92 (deconvolve_transpose!)
void orc_deconvolve_transpose(OPlan* p, const double* g_, double* f_) {
  const cd* g = (const cd*)g_;
  cd* f = (cd*)f_;
  int D = p->D;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < p->P; ++i) {
    i64 r = i, gi = 0, str = 1;
    i64 id[MAXD];
    for (int d = 0; d < D; ++d) {
      id[d] = r % p->N[d];
      r /= p->N[d];
      i64 Na = p->N[d] / 2;
      i64 gd = (id[d] < Na) ? id[d] - Na + p->Nt[d] : id[d] - Na;  // :77-89
      gi += gd * str;
      str *= p->Nt[d];
    }
    cd v = g[gi];
    for (int d = 0; d < D; ++d) v *= p->deconv[d][id[d]];
    f[i] = v;
  }
}

// Interpolation, src/convolution.jl:20-45 dispatching to the blocked
// (:229-344), non-blocked (:56-99) or sparse FULL (:102-110, matrix built
// as src/precomputation.jl:81-120) variant.
void orc_convolve(OPlan* p, const double* g_, double* fhat_) {
  const cd* g = (const cd*)g_;
  cd* fhat = (cd*)fhat_;
  int D = p->D, Z = 2 * p->m;
  if (p->blocking && p->precompute != FULL) {
    i64 bsz[MAXD];
    for (int d = 0; d < D; ++d) bsz[d] = p->bs[d] + 2 * p->m;  // blockSizePadded (:491)
    switch (Z) {
#define ORC_CASE(ZZ) \
  case ZZ:           \
    conv_blocked<ZZ>(*p, g, fhat, bsz); \
    break;
      ORC_CASE(2) ORC_CASE(4) ORC_CASE(6) ORC_CASE(8) ORC_CASE(10) ORC_CASE(12) ORC_CASE(14)
      ORC_CASE(16)
#undef ORC_CASE
      default:
        conv_blocked<0>(*p, g, fhat, bsz);
    }
    return;
  }
  // non-blocked and FULL: fHat += prodWin * g[idx] with prodWin = w_D*...*w_1
#pragma omp parallel
  {
    std::vector<i64> idx(D * Z);
    std::vector<double> w(D * Z);
#pragma omp for schedule(static)
    for (i64 j = 0; j < p->M; ++j) {
      for (int d = 0; d < D; ++d) node_weights_nonblocked(*p, j, d, &idx[d * Z], &w[d * Z]);
      cd s(0, 0);
      int l[MAXD] = {0};
      do {
        double pw = 1.0;
        for (int d = D - 1; d >= 0; --d) pw = pw * w[d * Z + l[d]];
        i64 gi = 0, str = 1;
        for (int d = 0; d < D; ++d) {
          gi += idx[d * Z + l[d]] * str;
          str *= p->Nt[d];
        }
        s += pw * g[gi];  // FULL: transpose(nzv[j]) * g[rv[j]] (src/utils.jl:150)
      } while (next_multi(D, l, Z));
      fhat[j] = s;
    }
  }
}

// Spreading, src/convolution.jl:115-149 -> blocked (:349-443, fillOneNode!
// :465-492, merge under a lock :371-373), non-blocked (:152-200, serial) or
// sparse FULL (:203-211).  g is overwritten.
void orc_convolve_transpose(OPlan* p, const double* fhat_, double* g_) {
  const cd* fhat = (const cd*)fhat_;
  cd* g = (cd*)g_;
  int D = p->D, Z = 2 * p->m;
  std::fill(g, g + p->Q, cd(0, 0));  // memset (:358)
  if (p->blocking && p->precompute != FULL) {
    i64 bsz[MAXD];
    for (int d = 0; d < D; ++d) bsz[d] = p->bs[d] + 2 * p->m;
    switch (Z) {
#define ORC_CASE(ZZ) \
  case ZZ:           \
    spread_blocked<ZZ>(*p, fhat, g, bsz); \
    break;
      ORC_CASE(2) ORC_CASE(4) ORC_CASE(6) ORC_CASE(8) ORC_CASE(10) ORC_CASE(12) ORC_CASE(14)
      ORC_CASE(16)
#undef ORC_CASE
      default:
        spread_blocked<0>(*p, fhat, g, bsz);
    }
    return;
  }
  // non-blocked: serial (:170-173); FULL: mul!(vec(g), B, fHat)
  std::vector<i64> idx(D * Z);
  std::vector<double> w(D * Z);
  for (i64 j = 0; j < p->M; ++j) {
    for (int d = 0; d < D; ++d) node_weights_nonblocked(*p, j, d, &idx[d * Z], &w[d * Z]);
    int l[MAXD] = {0};
    do {
      double pw = 1.0;
      for (int d = D - 1; d >= 0; --d) pw = pw * w[d * Z + l[d]];
      i64 gi = 0, str = 1;
      for (int d = 0; d < D; ++d) {
        gi += idx[d * Z + l[d]] * str;
        str *= p->Nt[d];
      }
      g[gi] += pw * fhat[j];
    } while (next_multi(D, l, Z));
  }
}

// Direct NDFT, src/direct.jl:99-123: fhat_j = sum_n f_n exp(-2 pi i n.k_j),
// n_d = i_d - floor(N_d/2).  Evaluated separably with exact phase reduction.
// sel (optional, nsel entries) restricts the output to those node indices;
// fhat then has nsel entries.
void orc_ndft(int D, const long long* N, long long M, const double* k, const double* f_,
              double* fhat_, const long long* sel, long long nsel) {
  const cd* f = (const cd*)f_;
  cd* fhat = (cd*)fhat_;
  i64 P = 1;
  for (int d = 0; d < D; ++d) P *= N[d];
  i64 nout = sel ? nsel : M;
#pragma omp parallel
  {
    std::vector<std::vector<cd>> ph(D);
    std::vector<cd> part;
#pragma omp for schedule(dynamic, 4)
    for (i64 q = 0; q < nout; ++q) {
      i64 j = sel ? sel[q] : q;
      for (int d = 0; d < D; ++d) {
        ph[d].resize(N[d]);
        i64 Na = N[d] / 2;
        for (i64 i = 0; i < N[d]; ++i) {
          double fr = frac_prod(k[d + (i64)D * j], (double)(i - Na));
          double a = -2.0 * M_PI * fr;
          ph[d][i] = cd(std::cos(a), std::sin(a));
        }
      }
      // contract dimension by dimension, innermost (dim 0) first
      part.assign(f, f + P);
      i64 len = P;
      for (int d = 0; d < D; ++d) {
        i64 n = N[d], outer = len / n;
        for (i64 o = 0; o < outer; ++o) {
          cd s(0, 0);
          for (i64 i = 0; i < n; ++i) s += part[o * n + i] * ph[d][i];
          part[o] = s;
        }
        len = outer;
      }
      fhat[q] = part[0];
    }
  }
}

// Direct adjoint NDFT, src/direct.jl:126-151: f_n = sum_j fhat_j exp(+2 pi i n.k_j).
// sel (optional) lists linear grid indices (0-based, Julia order); f then has
// nsel entries.
void orc_ndft_adjoint(int D, const long long* N, long long M, const double* k, const double* fhat_,
                      double* f_, const long long* sel, long long nsel) {
  const cd* fhat = (const cd*)fhat_;
  cd* f = (cd*)f_;
  i64 P = 1;
  for (int d = 0; d < D; ++d) P *= N[d];
  i64 nout = sel ? nsel : P;
#pragma omp parallel for schedule(dynamic, 16)
  for (i64 q = 0; q < nout; ++q) {
    i64 lin = sel ? sel[q] : q;
    double n[MAXD];
    i64 r = lin;
    for (int d = 0; d < D; ++d) {
      n[d] = (double)(r % N[d] - N[d] / 2);
      r /= N[d];
    }
    cd s(0, 0);
    for (i64 j = 0; j < M; ++j) {
      double fr = 0;
      for (int d = 0; d < D; ++d) fr += frac_prod(k[d + (i64)D * j], n[d]);
      fr -= std::nearbyint(fr);
      double a = 2.0 * M_PI * fr;
      s += fhat[j] * cd(std::cos(a), std::sin(a));
    }
    f[q] = s;
  }
}

// Analytic window derivative of the window sample itself (grid units):
// exposed so the test suite can check it against finite differences.
double orc_window_du(double u, int m, double sigma) { return window_kb_du(u, m, sigma); }

// Node gradient of fhat = A(k) f contracted with a cotangent ybar, given the
// deconvolved and transformed grid g = FFT(deconvolve(f)) of the forward path
// (src/implementation.jl:155-172):
//   kbar[d + D j] = Re(conj(ybar_j) * sum_l g[idx_l] d/dk_{d,j} prod_e w_e(l_e)).
// Differentiating the interpolation sum of src/convolution.jl:56-99 term by
// term; the reference's rrule returns NoTangent() for the nodes
// (AbstractNFFTs/ext/AbstractNFFTsChainRulesCoreExt.jl:14-22, 26-45).
void orc_grad_gather(OPlan* p, const double* g_, const double* ybar_, double* kbar) {
  const cd* g = (const cd*)g_;
  const cd* ybar = (const cd*)ybar_;
  int D = p->D, Z = 2 * p->m;
#pragma omp parallel
  {
    std::vector<i64> idx(D * Z);
    std::vector<double> w(D * Z), dw(D * Z);
#pragma omp for schedule(static)
    for (i64 j = 0; j < p->M; ++j) {
      for (int d = 0; d < D; ++d) node_weights_grad(*p, j, d, &idx[d * Z], &w[d * Z], &dw[d * Z]);
      cd S[MAXD];
      for (int d = 0; d < D; ++d) S[d] = cd(0, 0);
      int l[MAXD] = {0};
      do {
        i64 gi = 0, str = 1;
        for (int d = 0; d < D; ++d) {
          gi += idx[d * Z + l[d]] * str;
          str *= p->Nt[d];
        }
        const cd v = g[gi];
        for (int d = 0; d < D; ++d) {
          double pw = dw[d * Z + l[d]];
          for (int e = 0; e < D; ++e)
            if (e != d) pw *= w[e * Z + l[e]];
          S[d] += pw * v;
        }
      } while (next_multi(D, l, Z));
      for (int d = 0; d < D; ++d) kbar[d + (i64)D * j] = (std::conj(ybar[j]) * S[d]).real();
    }
  }
}

// Closed-form node gradient of the direct NDFT (src/direct.jl:99-123):
//   d fhat_j / d k_{d,j} = sum_n f_n (-2 pi i n_d) exp(-2 pi i n.k_j),
//   kbar[d + D q] = Re(conj(ybar_j) * that), j = sel[q] (or q).
void orc_ndft_grad(int D, const long long* N, long long M, const double* k, const double* f_,
                   const double* ybar_, double* kbar, const long long* sel, long long nsel) {
  const cd* f = (const cd*)f_;
  const cd* ybar = (const cd*)ybar_;
  i64 P = 1;
  for (int d = 0; d < D; ++d) P *= N[d];
  i64 nout = sel ? nsel : M;
#pragma omp parallel
  {
    std::vector<std::vector<cd>> ph(D);
    std::vector<cd> part;
#pragma omp for schedule(dynamic, 4)
    for (i64 q = 0; q < nout; ++q) {
      i64 j = sel ? sel[q] : q;
      for (int d = 0; d < D; ++d) {
        ph[d].resize(N[d]);
        i64 Na = N[d] / 2;
        for (i64 i = 0; i < N[d]; ++i) {
          double a = -2.0 * M_PI * frac_prod(k[d + (i64)D * j], (double)(i - Na));
          ph[d][i] = cd(std::cos(a), std::sin(a));
        }
      }
      for (int dd = 0; dd < D; ++dd) {  // derivative along dimension dd
        part.assign(f, f + P);
        i64 len = P;
        for (int d = 0; d < D; ++d) {
          i64 n = N[d], outer = len / n, Na = N[d] / 2;
          for (i64 o = 0; o < outer; ++o) {
            cd s(0, 0);
            for (i64 i = 0; i < n; ++i) {
              cd t = part[o * n + i] * ph[d][i];
              if (d == dd) t *= cd(0.0, -2.0 * M_PI * (double)(i - Na));
              s += t;
            }
            part[o] = s;
          }
          len = outer;
        }
        kbar[dd + (i64)D * q] = (std::conj(ybar[j]) * part[0]).real();
      }
    }
  }
}

// Direct adjoint NDFT, src/direct.jl:126-151, on a product set of grid
// points: along dimension d the points are idx[off_d + 0 .. cnt[d]-1]
// (0-based image indices i_d, frequency n_d = i_d - floor(N_d/2)), off_d =
// cnt[0] + ... + cnt[d-1].  f has prod(cnt) entries, dimension 0 fastest.
// Per node the phase factorises, exp(+2 pi i n.k_j) = prod_d exp(2 pi i
// n_d k_dj), so each node costs sum(cnt) sin/cos pairs (exact phase
// reduction as above) plus prod(cnt) complex products, instead of prod(cnt)
// D-term phases.  Nodes are split over threads; each thread sums into its
// own output, and the partial outputs are added at the end.
void orc_ndft_adjoint_box(int D, const long long* N, long long M, const double* k, const double* fhat_,
                          double* f_, const long long* idx, const long long* cnt) {
  const cd* fhat = (const cd*)fhat_;
  cd* f = (cd*)f_;
  i64 nout = 1, tot = 0;
  for (int d = 0; d < D; ++d) {
    nout *= cnt[d];
    tot += cnt[d];
  }
  std::fill(f, f + nout, cd(0, 0));
#pragma omp parallel
  {
    std::vector<cd> mine(nout, cd(0, 0)), ph(tot), part(nout);
#pragma omp for schedule(static)
    for (i64 j = 0; j < M; ++j) {
      i64 o = 0;
      for (int d = 0; d < D; ++d) {
        const double kd = k[d + (i64)D * j];
        for (i64 q = 0; q < cnt[d]; ++q) {
          const double a = 2.0 * M_PI * frac_prod(kd, (double)(idx[o + q] - N[d] / 2));
          ph[o + q] = cd(std::cos(a), std::sin(a));
        }
        o += cnt[d];
      }
      // outer product of the per-dimension phases, built from the last
      // dimension (slowest) to dimension 0 (fastest); walking r downwards
      // never overwrites an entry before it is read
      i64 len = 1;
      part[0] = fhat[j];
      o = tot;
      for (int d = D - 1; d >= 0; --d) {
        o -= cnt[d];
        for (i64 r = len - 1; r >= 0; --r) {
          const cd v = part[r];
          for (i64 q = cnt[d] - 1; q >= 0; --q) part[r * cnt[d] + q] = v * ph[o + q];
        }
        len *= cnt[d];
      }
      for (i64 t = 0; t < nout; ++t) mine[t] += part[t];
    }
#pragma omp critical
    {
      for (i64 t = 0; t < nout; ++t) f[t] += mine[t];
    }
  }
}

}  // extern "C"
