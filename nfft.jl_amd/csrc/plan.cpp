// plan.cpp -- host side of libnfftm: plan construction (parameters, window
// tables, deconvolution factors, FP64 node binning into tiles), the rocFFT
// stage, the transform orchestration, and the C ABI of include/nfftm.h.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nfftm.h"
#include "fft1.hpp"
#include "launch.hpp"

using namespace nfftm;
typedef int64_t i64;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      return fail(e_ == hipErrorOutOfMemory ? NFFTM_ERR_OOM : NFFTM_ERR_HIP, "%s: %s", #x, \
                  hipGetErrorString(e_));                                                 \
  } while (0)

// ---------------------------------------------------------------------------
// Kaiser-Bessel window and its Fourier transform (src/windowFunctions.jl:21-39)
// ---------------------------------------------------------------------------
double i0_series(double x) {
  // I0(x) = sum_k ((x/2)^2)^k / (k!)^2; positive terms, relative error ~ulp.
  const double y = 0.25 * x * x;
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 1000 && t > 1e-17 * s; ++k) {
    t = t * y / ((double)k * k);
    s += t;
  }
  return s;
}

// psi(u) with u in grid units (the reference's window(k, Ntilde, m, sigma) at
// k = u / Ntilde; the lookup table uses Ntilde = 1).
double kb_psi(double u, int m, double b) {
  const double au = std::fabs(u);
  if (au < m) {
    const double arg = std::sqrt((double)m * m - u * u);
    return std::sinh(b * arg) / (arg * M_PI);
  }
  if (au > m) return 0.0;
  return b / M_PI;
}

double kb_psi_hat(double n, double Nt, int m, double b) {
  const double t = 2.0 * M_PI * n / Nt;
  return i0_series(m * std::sqrt(b * b - t * t));
}

// ---------------------------------------------------------------------------
// The other windows of src/windowFunctions.jl:41-134 (ids as the C ABI's
// NFFTM_GAUSS .. NFFTM_COSH_TYPE), psi in grid units u = Ntilde k.
// ---------------------------------------------------------------------------
// J0(x) = (1/pi) int_0^pi cos(x sin t) dt by the trapezoidal rule on the
// periodic integrand, which converges geometrically: 48 + |x| nodes give full
// FP64 precision for the arguments of the cosh-type transform.
double j0_trapz(double x) {
  const int n = 48 + (int)std::ceil(std::fabs(x));
  double s = 0.5 * (1.0 + std::cos(0.0));  // t = 0 and t = pi ends: cos(0) each, halved
  for (int i = 1; i < n; ++i) s += std::cos(x * std::sin(M_PI * i / n));
  return s / n;
}

// Cardinal B-spline of order n at x (cbspline, src/windowFunctions.jl:65-77),
// Cox-de Boor over the shifts.
double bspline(int n, double x) {
  std::vector<double> v(n);
  for (int s = 0; s < n; ++s) v[s] = (x - s >= 0.0 && x - s < 1.0) ? 1.0 : 0.0;
  for (int k = 2; k <= n; ++k)
    for (int s = 0; s + k <= n; ++s) v[s] = ((x - s) * v[s] + (k - (x - s)) * v[s + 1]) / (k - 1);
  return v[0];
}

double win_psi(int win, double u, int m, double sigma) {
  const double au = std::fabs(u);
  switch (win) {
    case NFFTM_KAISER_BESSEL:
      return kb_psi(u, m, M_PI * (2.0 - 1.0 / sigma));
    case NFFTM_GAUSS: {
      const double b = m / M_PI;
      return au < m ? std::exp(-u * u / b) / std::sqrt(M_PI * b) : 0.0;
    }
    case NFFTM_SPLINE:
      return au < m ? bspline(2 * m, u + m) : 0.0;
    case NFFTM_KAISER_BESSEL_REV: {
      const double b = M_PI * (2.0 - 1.0 / sigma), r = u / m;
      return au < m ? 0.5 / m * i0_series(m * b * std::sqrt(1.0 - r * r)) : 0.0;
    }
    case NFFTM_COSH_TYPE: {
      const double beta = M_PI * m * (2.0 - 1.0 / sigma), r = u / m;
      if (!(au < m)) return 0.0;
      const double al = std::sqrt(1.0 - r * r);
      return (std::cosh(beta * al) - 1.0) / ((std::cosh(beta) - 1.0) * al);
    }
  }
  return 0.0;
}

double win_psi_hat(int win, double n, double Nt, int m, double sigma) {
  switch (win) {
    case NFFTM_KAISER_BESSEL:
      return kb_psi_hat(n, Nt, m, M_PI * (2.0 - 1.0 / sigma));
    case NFFTM_GAUSS: {
      const double t = M_PI * n / Nt;
      return std::exp(-t * t * (m / M_PI));
    }
    case NFFTM_SPLINE: {
      const double x = n / Nt;
      return std::pow(x == 0.0 ? 1.0 : std::sin(M_PI * x) / (M_PI * x), 2 * m);
    }
    case NFFTM_KAISER_BESSEL_REV: {
      // real(sinc(sqrt(complex(q)) / pi)): sin(a)/a for q >= 0, sinh(y)/y below
      const double b = M_PI * (2.0 - 1.0 / sigma), t = 2.0 * M_PI * m * n / Nt;
      const double q = t * t - (m * b) * (m * b);
      if (q >= 0.0) {
        const double a = std::sqrt(q);
        return a == 0.0 ? 1.0 : std::sin(a) / a;
      }
      const double y = std::sqrt(-q);
      return std::sinh(y) / y;
    }
    case NFFTM_COSH_TYPE: {
      const double beta = M_PI * m * (2.0 - 1.0 / sigma), gam = beta / (2.0 * M_PI);
      const double zeta = M_PI * m / (std::cosh(beta) - 1.0), arg = m * n / Nt, w = 2.0 * M_PI * arg;
      if (std::fabs(arg) < gam) return zeta * (i0_series(std::sqrt(beta * beta - w * w)) - j0_trapz(w));
      if (std::fabs(arg) > gam) return zeta * (j0_trapz(std::sqrt(w * w - beta * beta)) - j0_trapz(w));
      return zeta * (1.0 - j0_trapz(beta));
    }
  }
  return 1.0;
}

// Chebyshev interpolant with n first-kind nodes on [a, b]
// (BasicInterpolators.ChebyshevInterpolator, used at src/precomputation.jl:351):
// coefficients from the discrete cosine sums, evaluation by Clenshaw.
struct ChebInterp {
  double a, b;
  std::vector<double> c;
  template <class F>
  ChebInterp(F f, double a_, double b_, int n) : a(a_), b(b_), c(n, 0.0) {
    std::vector<double> v(n);
    for (int i = 0; i < n; ++i) v[i] = f(0.5 * (b + a) + 0.5 * (b - a) * std::cos(M_PI * (i + 0.5) / n));
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int i = 0; i < n; ++i) s += v[i] * std::cos(M_PI * j * (i + 0.5) / n);
      c[j] = (j == 0 ? 1.0 : 2.0) * s / n;
    }
  }
  double operator()(double t) const {
    const double x = (2.0 * t - a - b) / (b - a);
    double b1 = 0, b2 = 0;
    for (int j = (int)c.size() - 1; j >= 1; --j) {
      const double b0 = 2.0 * x * b1 - b2 + c[j];
      b2 = b1;
      b1 = b0;
    }
    return x * b1 - b2 + c[0];
  }
};

// LUTSize = 2^K * m with K from NFFT3's table (src/precomputation.jl:14-17).
i64 lut_size(int m) {
  static const int K[9] = {1, 3, 7, 9, 14, 17, 20, 23, 24};
  return ((i64)1 << K[std::min(m, 8)]) * m;
}

// Least-squares polynomial fit of each stencil column
// (src/precomputation.jl:302-320): sample points s_q = -1/2 + q/(2*deg-1),
// q = 0..2*deg-1, deg = 2m+1 monomials; column l (0-based) fits
// psi(m - (l + 1/2) + x).  Solved by Householder QR of the Vandermonde matrix.
std::vector<double> poly_fit(int win, int m, double sigma) {
  const int deg = 2 * m + 1, W = 2 * m, ns = 2 * deg;
  std::vector<double> x(ns);
  for (int q = 0; q < ns; ++q) x[q] = (double)(2 * q - (ns - 1)) / (double)(2 * (ns - 1));
  // V (ns x deg, column-major) and its QR factors, shared by all columns
  std::vector<double> V((size_t)ns * deg);
  for (int q = 0; q < ns; ++q) {
    double p = 1.0;
    for (int r = 0; r < deg; ++r) {
      V[q + (size_t)r * ns] = p;
      p *= x[q];
    }
  }
  std::vector<std::vector<double>> house(deg);
  std::vector<double> beta(deg, 0.0);
  for (int c = 0; c < deg; ++c) {
    double nrm2 = 0;
    for (int q = c; q < ns; ++q) nrm2 += V[q + (size_t)c * ns] * V[q + (size_t)c * ns];
    const double nrm = std::sqrt(nrm2);
    const double a0 = V[c + (size_t)c * ns];
    const double alpha = a0 >= 0 ? -nrm : nrm;
    std::vector<double> v(ns, 0.0);
    v[c] = a0 - alpha;
    for (int q = c + 1; q < ns; ++q) v[q] = V[q + (size_t)c * ns];
    double vv = 0;
    for (int q = c; q < ns; ++q) vv += v[q] * v[q];
    beta[c] = vv > 0 ? 2.0 / vv : 0.0;
    for (int cc = c; cc < deg; ++cc) {
      double s = 0;
      for (int q = c; q < ns; ++q) s += v[q] * V[q + (size_t)cc * ns];
      s *= beta[c];
      for (int q = c; q < ns; ++q) V[q + (size_t)cc * ns] -= s * v[q];
    }
    house[c] = std::move(v);
  }
  std::vector<double> P((size_t)deg * W);
  for (int l = 0; l < W; ++l) {
    std::vector<double> y(ns);
    const double c0 = (double)m - ((double)l + 0.5);
    for (int q = 0; q < ns; ++q) y[q] = win_psi(win, c0 + x[q], m, sigma);
    for (int c = 0; c < deg; ++c) {
      double s = 0;
      for (int q = c; q < ns; ++q) s += house[c][q] * y[q];
      s *= beta[c];
      for (int q = c; q < ns; ++q) y[q] -= s * house[c][q];
    }
    for (int r = deg - 1; r >= 0; --r) {
      double s = y[r];
      for (int cc = r + 1; cc < deg; ++cc) s -= V[r + (size_t)cc * ns] * P[cc + (size_t)l * deg];
      P[r + (size_t)l * deg] = s / V[r + (size_t)r * ns];
    }
  }
  return P;
}

// ---------------------------------------------------------------------------
// rocFFT stage: in place, unnormalised, all D dimensions, Julia order
// (dimension 1 = rocFFT lengths[0]).  D <= 3 is one plan; D > 3 is a 3-D plan
// batched over the outer dimensions followed by strided 1-D plans.
// ---------------------------------------------------------------------------
std::once_flag g_rocfft_once;

// Work enqueued on `next` after a change of stream must not overtake work
// still running on `cur` (both use the same plan-owned buffers): `next`
// waits for everything enqueued on `cur` so far.  Skipped when either stream
// is capturing a graph (a capture may not wait on outside events); callers
// that capture order their streams themselves.
int order_streams(hipStream_t cur, hipStream_t next, hipEvent_t* ev) {
  if (next == cur) return NFFTM_OK;
  hipStreamCaptureStatus a = hipStreamCaptureStatusNone, b = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(cur, &a) != hipSuccess || hipStreamIsCapturing(next, &b) != hipSuccess) {
    (void)hipGetLastError();
    return NFFTM_OK;
  }
  if (a != hipStreamCaptureStatusNone || b != hipStreamCaptureStatusNone) return NFFTM_OK;
  if (!*ev) HIPCHK(hipEventCreateWithFlags(ev, hipEventDisableTiming));
  HIPCHK(hipEventRecord(*ev, cur));
  HIPCHK(hipStreamWaitEvent(next, *ev, 0));
  return NFFTM_OK;
}

struct FFTPass {
  rocfft_plan plan = nullptr;
  i64 nexec = 1, exec_stride = 0;  // executions at pointer offsets (elements)
};

struct FFTStage {
  std::vector<FFTPass> fwd, bwd;
  rocfft_execution_info info = nullptr;
  void* work = nullptr;
  size_t worksize = 0;

  ~FFTStage() { release(); }
  void release() {
    for (auto* v : {&fwd, &bwd})
      for (auto& p : *v)
        if (p.plan) rocfft_plan_destroy(p.plan);
    fwd.clear();
    bwd.clear();
    if (info) rocfft_execution_info_destroy(info);
    info = nullptr;
    if (work) (void)hipFree(work);
    work = nullptr;
  }

  int make_pass(std::vector<FFTPass>& out, bool forward, bool dbl, int dim, const size_t* len,
                const size_t* stride, size_t dist, size_t batch, i64 nexec, i64 exec_stride) {
    rocfft_plan_description desc = nullptr;
    if (rocfft_plan_description_create(&desc) != rocfft_status_success)
      return fail(NFFTM_ERR_FFT, "rocfft_plan_description_create failed");
    rocfft_status st = rocfft_plan_description_set_data_layout(
        desc, rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved, nullptr,
        nullptr, dim, stride, dist, dim, stride, dist);
    FFTPass p;
    if (st == rocfft_status_success)
      st = rocfft_plan_create(&p.plan, rocfft_placement_inplace,
                              forward ? rocfft_transform_type_complex_forward
                                      : rocfft_transform_type_complex_inverse,
                              dbl ? rocfft_precision_double : rocfft_precision_single, dim, len, batch,
                              desc);
    rocfft_plan_description_destroy(desc);
    if (st != rocfft_status_success) return fail(NFFTM_ERR_FFT, "rocfft_plan_create failed (%d)", (int)st);
    size_t ws = 0;
    rocfft_plan_get_work_buffer_size(p.plan, &ws);
    worksize = std::max(worksize, ws);
    p.nexec = nexec;
    p.exec_stride = exec_stride;
    out.push_back(p);
    return NFFTM_OK;
  }

  // nbatch consecutive grids of prod(Nt) points (D <= 3 when nbatch > 1)
  int build(int D, const i64* Nt, bool dbl, size_t nbatch = 1) {
    std::call_once(g_rocfft_once, [] { rocfft_setup(); });
    for (int dir = 0; dir < 2; ++dir) {
      auto& v = dir == 0 ? fwd : bwd;
      const bool f = dir == 0;
      const int d3 = std::min(D, 3);
      size_t len[3], str[3];
      size_t s = 1;
      for (int d = 0; d < d3; ++d) {
        len[d] = (size_t)Nt[d];
        str[d] = s;
        s *= (size_t)Nt[d];
      }
      size_t outer = 1;
      for (int d = 3; d < D; ++d) outer *= (size_t)Nt[d];
      int rc = make_pass(v, f, dbl, d3, len, str, s, outer * nbatch, 1, 0);
      if (rc) return rc;
      size_t inner = s;
      for (int d = 3; d < D; ++d) {
        size_t after = 1;
        for (int e = d + 1; e < D; ++e) after *= (size_t)Nt[e];
        size_t l1[1] = {(size_t)Nt[d]}, s1[1] = {inner};
        rc = make_pass(v, f, dbl, 1, l1, s1, 1, inner, (i64)after, (i64)(inner * Nt[d]));
        if (rc) return rc;
        inner *= (size_t)Nt[d];
      }
    }
    if (rocfft_execution_info_create(&info) != rocfft_status_success)
      return fail(NFFTM_ERR_FFT, "rocfft_execution_info_create failed");
    if (worksize) {
      if (hipMalloc(&work, worksize) != hipSuccess) return fail(NFFTM_ERR_OOM, "rocFFT work buffer");
      rocfft_execution_info_set_work_buffer(info, work, worksize);
    }
    return NFFTM_OK;
  }

  int exec(bool forward, void* g, size_t elem, hipStream_t s) {
    rocfft_execution_info_set_stream(info, s);
    for (auto& p : forward ? fwd : bwd)
      for (i64 e = 0; e < p.nexec; ++e) {
        void* buf[1] = {(char*)g + (size_t)(e * p.exec_stride) * elem};
        if (rocfft_execute(p.plan, buf, nullptr, info) != rocfft_status_success)
          return fail(NFFTM_ERR_FFT, "rocfft_execute failed");
      }
    return NFFTM_OK;
  }
};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    reset();
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e == hipSuccess) n = count;
    return e;
  }
};

}  // namespace

// ===========================================================================
// The plan
// ===========================================================================
struct nfftm_plan_s {
  // parameters (NFFTParams, src/implementation.jl:3-14)
  int prec = NFFTM_F32, D = 1, m = 4, W = 8, window = 0, precompute = NFFTM_LINEAR, flags = 0;
  double sigma_in = 2.0, sigma = 2.0, b = 0.0, wscale = 1.0;
  i64 N[MAXD] = {0}, Nt[MAXD] = {0}, P = 1, Q = 1, M = 0, LUTSize = 0;
  bool dbl() const { return prec == NFFTM_F64; }
  // Directional plan (plan_nfft(k, N; dims), src/directional.jl): D, N, ...
  // describe the transform over the axes dims = dlo .. dlo + D - 1 of the
  // caller's Dfull-dimensional arrays (Julia order).  Every slice of those
  // arrays along the other axes is one transform: its elements are Bsl apart
  // (the product of the sizes before dims) and slices follow one another in
  // Bsl x Asl order.  Dfull = 0: an ordinary plan.
  int Dfull = 0, dlo = 0;
  i64 Nfull[MAXD] = {0}, Bsl = 1, Asl = 1;
  bool dir() const { return Dfull > 0; }
  // element stride of the caller's image and node values in the kernels'
  // arguments: Bsl while a directional transform runs, 1 otherwise (the
  // plan's own work vectors are contiguous)
  i64 io_stride = 1;
  // 3-D scratch slots in halo/interior block layout: whole tiles per
  // dimension and halos (m - 1, m cells) no wider than a tile
  bool blocked3() const {
    if (D != 3 || !tiled) return false;
    for (int d = 0; d < 3; ++d)
      if (Nt[d] % T[d] != 0 || T[d] < m) return false;
    return true;
  }
  // cells of one scratch slot (3-D: rounded to a 128-byte multiple)
  size_t slot_cells() const {
    const size_t pad = (size_t)Pd[0] * Pd[1] * Pd[2];
    return D == 3 ? (pad + 15) & ~(size_t)15 : pad;
  }
  size_t celem() const { return dbl() ? 16 : 8; }
  bool tiled = true;
  hipStream_t stream = nullptr;

  // tables (host FP64 masters; device copies in the data precision)
  std::vector<double> lut, poly;
  std::vector<std::vector<double>> deconv;
  DevBuf<char> d_lut, d_poly, d_grid;
  DevBuf<double> d_deconv, d_ks, d_polyd;
  DevBuf<int> d_perm;
  DevBuf<int4> d_items;
  DevBuf<int> d_tstart;
  DevBuf<char> d_scratch, d_rec;
  // Tiled 1-D polynomial-window plans: the inverse binning permutation and
  // the sorted-order interpolation results it maps back to the caller's
  // order.  (Measured on 3-D plans of 2^21 nodes, whose node values stay in
  // the caches, the extra gather launch cost more than the scattered stores
  // it replaces.)
  DevBuf<int> d_iperm;
  DevBuf<char> d_fsort;
  bool sorted_io() const {
    return tiled && D == 1 && precompute != NFFTM_LINEAR && precompute != NFFTM_FULL;
  }
  // node-side staging vector (M values of the plan's complex type) of the
  // typed stage operators, allocated at their first call
  DevBuf<char> d_tmpM;
  // NFFTM_FLAG_SPARSE: rows of W^D grid indices and window weights per node
  bool sparse = false;
  int sparseK = 0;
  DevBuf<int> d_bcols;
  DevBuf<char> d_bvals;
  int nitems = 0;
  bool single_items = false;  // tiled, nitems > 0, no tile with two work items
  int gen = 0;        // incremented whenever the node arrays are rebuilt
  double wmax = 1.0;  // bound of one (normalised) window value

  // tiling
  int T[3] = {1, 1, 1}, ntiles[3] = {1, 1, 1}, Pd[3] = {1, 1, 1}, Sx = 1, Sxy = 1, tileElems = 1, NT = 256;
  int tvar = 0;  // 3-D tile geometry variant
  size_t lds = 0;

  std::unique_ptr<FFTStage> fft;
  // pruned FFT passes (power-of-two grids): twiddle tables per dimension
  bool pfft = false;
  DevBuf<char> d_tw;
  size_t tw_off[3] = {0, 0, 0};
  // 1-D grids of n1 Lm^2 points: the three-pass FFT of k_fft1.hip, its
  // tables (n1-point, Lm-point, and the two of exp(2 pi i e / Nt)) in d_tw
  bool f1 = false;
  int f1n1 = 0, f1Lm = 0, f1sh = 0;
  size_t f1off[4] = {0, 0, 0, 0};
  DevBuf<double> d_fac1;
  double timing[7] = {0};
  // kernel-level profiling (event pairs recorded without host sync)
  bool profiling = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev;
  std::vector<int> prof_stage;
  size_t prof_used = 0;
  double prof_ms[8] = {0};
  int64_t prof_cnt[8] = {0};
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<double> k;  // caller's nodes (D x M), kept for copy()

  // Replayed transforms: the launch sequence of one direction for one
  // (input, output) pair, captured into a HIP graph at the pair's second call
  // (first calls run directly), dropped when the node arrays are rebuilt.
  struct GraphEntry {
    int dir;
    const void* in;
    void* out;
    int uses;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  hipStream_t cap = nullptr;  // capture stream
  hipEvent_t switch_ev = nullptr;  // orders work across nfftm_set_stream changes
  bool graphs_ok = true;      // false after a failed capture: direct launches only
  void drop_graphs() {
    for (auto& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }

  ~nfftm_plan_s() {
    drop_graphs();
    if (cap) (void)hipStreamDestroy(cap);
    if (switch_ev) (void)hipEventDestroy(switch_ev);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& pr : prof_ev) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  }

  // fold recorded event pairs into the per-stage sums (synchronises)
  int prof_drain() {
    if (prof_used == 0) return NFFTM_OK;
    hipError_t e = hipEventSynchronize(prof_ev[prof_used - 1].second);
    if (e != hipSuccess) return NFFTM_ERR_HIP;
    for (size_t i = 0; i < prof_used; ++i) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, prof_ev[i].first, prof_ev[i].second) != hipSuccess) return NFFTM_ERR_HIP;
      prof_ms[prof_stage[i]] += ms;
      prof_cnt[prof_stage[i]] += 1;
    }
    prof_used = 0;
    return NFFTM_OK;
  }

  // returns the event pair to record around a stage launch, or null
  std::pair<hipEvent_t, hipEvent_t>* prof_slot(int stage) {
    if (!profiling) return nullptr;
    if (prof_used == prof_ev.size()) {
      if (prof_ev.size() >= 4096) {
        if (prof_drain() != NFFTM_OK) return nullptr;
      } else {
        std::pair<hipEvent_t, hipEvent_t> pr;
        if (hipEventCreate(&pr.first) != hipSuccess || hipEventCreate(&pr.second) != hipSuccess) return nullptr;
        prof_ev.push_back(pr);
        prof_stage.push_back(0);
      }
    }
    prof_stage[prof_used] = stage;
    return &prof_ev[prof_used++];
  }

  KArgs kargs() const {
    KArgs a;
    std::memset(&a, 0, sizeof a);
    a.D = D;
    a.m = m;
    a.W = W;
    a.mode = precompute == NFFTM_LINEAR ? WIN_LINEAR : precompute == NFFTM_FULL ? WIN_EXACT : WIN_POLY;
    a.M = (int)M;
    i64 s = 1;
    for (int d = 0; d < D; ++d) {
      a.Nt[d] = (int)Nt[d];
      a.Ntd[d] = (double)Nt[d];
      a.gstr[d] = s;
      s *= Nt[d];
    }
    for (int d = 0; d < 3; ++d) {
      a.T[d] = T[d];
      a.ntiles[d] = ntiles[d];
      a.Pd[d] = Pd[d];
    }
    a.Sx = Sx;
    a.Sxy = Sxy;
    a.tileElems = tileElems;
    a.padCells = Pd[0] * Pd[1] * Pd[2];
    a.divP0 = make_fastdiv((uint32_t)Pd[0]);
    a.divP1 = make_fastdiv((uint32_t)Pd[1]);
    a.divNtx = make_fastdiv((uint32_t)std::max<i64>(1, Nt[0]));
    a.lut = d_lut.p;
    a.scale = LUTSize / m;
    a.poly = d_poly.p;
    a.polyd = d_polyd.p;
    a.b = b;
    a.win = window;
    a.wscale = wscale;
    a.oscale = 1.0;
    a.ks = d_ks.p;
    a.perm = d_perm.p;
    a.fsort = nullptr;
    a.items = d_items.p;
    a.nitems = nitems;
    a.rec = d_rec.p;
    a.blocked3 = blocked3() ? 1 : 0;
    a.tvar = tvar;
    // FP32 2-D spreading in packed fixed point where the deconvolution does
    // not amplify its rounding much (sigma near 2 or above; at small sigma the
    // FP64 planes keep the adjoint within 1e-5)
    a.fix2 = !dbl() && D == 2 && sigma >= 1.9 ? 1 : 0;
    a.recBytes = node_rec_bytes(D, dbl() ? sizeof(double) : sizeof(float), W);
    a.wmaxD = std::pow(wmax, (double)D);
    a.scratch = d_scratch.p;
    a.tstart = d_tstart.p;
    a.ntile_all = ntiles[0] * ntiles[1] * ntiles[2];
    a.fst = io_stride;
    return a;
  }

  FArgs fargs() const {
    FArgs a;
    std::memset(&a, 0, sizeof a);
    a.D = D;
    i64 gs = 1, off = 0;
    for (int d = 0; d < 3; ++d) {
      a.N[d] = d < D ? (int)N[d] : 1;
      a.L[d] = d < D ? (int)Nt[d] : 1;
      a.Na[d] = d < D ? (int)(N[d] / 2) : 0;
      a.gstr[d] = gs;
      if (d < D) {
        gs *= Nt[d];
        a.fac[d] = d_deconv.p + off;
        off += N[d];
        a.tw[d] = d_tw.p + tw_off[d];
      }
    }
    a.oscale = 1.0;
    a.ist = io_stride;
    return a;
  }

  F1Args f1args() const {
    F1Args a;
    std::memset(&a, 0, sizeof a);
    a.n1 = f1n1;
    a.P = f1n1 % 5 == 0 ? 5 : 1;
    a.M = f1n1 / a.P;
    a.Lm = f1Lm;
    a.N = (int)N[0];
    a.Na = (int)(N[0] / 2);
    a.Nt = (int)Nt[0];
    a.ist = io_stride;
    a.sh = f1sh;
    a.B = fft1_lines(f1n1, dbl());
    a.fac1 = d_fac1.p;
    a.tw1 = d_tw.p + f1off[0];
    a.twl = d_tw.p + f1off[1];
    a.thi = d_tw.p + f1off[2];
    a.tlo = d_tw.p + f1off[3];
    a.oscale = 1.0;
    return a;
  }

  DArgs dargs() const {
    DArgs a;
    std::memset(&a, 0, sizeof a);
    a.D = D;
    i64 ns = 1, gs = 1, off = 0;
    for (int d = 0; d < D; ++d) {
      a.N[d] = (int)N[d];
      a.Nt[d] = (int)Nt[d];
      a.Na[d] = (int)(N[d] / 2);
      a.nstr[d] = ns * io_stride;
      a.gstr[d] = gs;
      ns *= N[d];
      gs *= Nt[d];
      a.fac[d] = d_deconv.p + off;
      off += N[d];
    }
    a.P = P;
    a.Q = Q;
    a.oscale = 1.0;
    a.div_half = make_fastdiv((uint32_t)std::max<i64>(1, Nt[0] / 2));
    for (int d = 0; d < D; ++d) a.div_nt[d] = make_fastdiv((uint32_t)Nt[d]);
    return a;
  }
};

namespace {

// Largest number of padded tiles covering one grid point along one dimension
// (tile t covers [tT - m + 1, tT + T + m - 1] mod Ntilde): exact near the
// periodic wrap (where a partial last tile brings origins closer than T),
// analytic in the interior.
int max_cover(i64 T, i64 P, i64 Nt, int m) {
  const i64 nt = (Nt + T - 1) / T;
  int best = (int)std::min<i64>((T + 2 * m - 2) / T + 1, nt);
  auto count = [&](i64 x) {
    int c = 0;
    for (i64 t = 0; t < nt; ++t) {
      i64 p = ((x - t * T + m - 1) % Nt + Nt) % Nt;
      for (; p < P; p += Nt) ++c;
    }
    return c;
  };
  // away from the periodic wrap every point has the interior count above;
  // only points within 2P of either end can exceed it
  for (i64 x = 0; x < std::min(Nt, 2 * P); ++x) best = std::max(best, count(x));
  for (i64 x = std::max<i64>(0, Nt - 2 * P); x < Nt; ++x) best = std::max(best, count(x));
  return best;
}

// Tile geometry of D <= 2 for tile edge t (clipped to the grid): padded
// extents, LDS strides and bytes; the row stride Sx is the padded extent
// rounded to 8 (mod 16) when that costs one element.  False if it exceeds the
// LDS budget.
bool set_tile_shape(nfftm_plan_s& p, int t) {
  const int D = p.D, W = p.W;
  const size_t budget = 158 * 1024;
  int T[3] = {t, D >= 2 ? t : 1, 1}, Pd[3] = {1, 1, 1};
  for (int d = 0; d < D; ++d) {
    T[d] = (int)std::min<i64>(T[d], p.Nt[d]);
    Pd[d] = T[d] + W - 1;
  }
  int Sx = Pd[0];
  if (D >= 2 && Pd[0] % 16 == 7) Sx = Pd[0] + 1;
  const int Sxy = Sx * (D >= 2 ? Pd[1] : 1);
  const int te = Sxy;
  const size_t bytes = tile_lds_bytes(D, p.dbl(), W, te);
  for (int d = 0; d < 3; ++d) {
    p.T[d] = D > d ? T[d] : 1;
    p.Pd[d] = D > d ? Pd[d] : 1;
    p.ntiles[d] = D > d ? (int)((p.Nt[d] + T[d] - 1) / T[d]) : 1;
  }
  p.Sx = Sx;
  p.Sxy = Sxy;
  p.tileElems = te;
  p.lds = bytes;
  return bytes <= budget;
}

// Default tile shape (D = 3: the compile-time geometries of k_tile3.hip).
// Grids smaller than a padded tile fail tiles_ok and take the node-parallel
// path.
bool tiles_ok(const nfftm_plan_s& p);

// 3-D geometry of variant V (compile-time tile of k_tile3.hip); false if
// that variant has no tile fitting its LDS budget.
bool set_tile3(nfftm_plan_s& p, int V) {
  int T[2] = {0, 0}, P[2] = {0, 0}, SX = 0;
  size_t lds = 0;
  bool valid = false;
  tile3_geometry(p.dbl(), p.W, V, T, &lds, P, &SX, &valid);
  if (!valid) return false;
  p.tvar = V;
  for (int d = 0; d < 3; ++d) {
    p.T[d] = T[d == 2 ? 1 : 0];
    p.Pd[d] = P[d == 2 ? 1 : 0];
    p.ntiles[d] = (int)((p.Nt[d] + p.T[d] - 1) / p.T[d]);
  }
  p.Sx = SX;  // LDS strides of the tile: the 3-D node records hold cell = x + Sx y + Sxy z
  p.Sxy = SX * P[0];
  p.tileElems = P[0] * P[0] * P[1];
  p.lds = lds;
  return true;
}

void choose_tiles(nfftm_plan_s& p) {
  const int D = p.D;
  const bool dbl = p.dbl();
  if (D == 3) {
    // the flat two-per-CU geometry unless it costs more than 1.7x the halo
    // of the cubic one (measured on the 256^3 grid: faster at m = 4 in both
    // precisions)
    p.NT = 1024;
    set_tile3(p, 1);
    const double h1 = (double)p.Pd[0] * p.Pd[1] * p.Pd[2] / ((double)p.T[0] * p.T[1] * p.T[2]);
    set_tile3(p, 0);
    const double h0 = (double)p.Pd[0] * p.Pd[1] * p.Pd[2] / ((double)p.T[0] * p.T[1] * p.T[2]);
    if (h1 <= 1.7 * h0 && set_tile3(p, 1) && tiles_ok(p)) return;
    set_tile3(p, 0);
    return;
  }
  // D = 2: 16^2 tiles; D = 1: 2048 cells (FP64: 1024); 512 threads -- the
  // autotuner's choice on the 2-D spiral, the batch of uneven 2-D plans and
  // the edge-clustered 1-D workload (nfftm_plan_tune explores the rest)
  p.NT = 512;
  if (D == 2) {
    int t = 16;
    while (!set_tile_shape(p, t) && t > 4) t /= 2;
  } else {
    int t = dbl ? 1024 : 2048;
    while (t > 256 && (double)p.Nt[0] / t < 256) t /= 2;
    while (!set_tile_shape(p, t) && t > 64) t /= 2;
  }
}

// The tiled path needs the padded range within one period and at most 3
// covering tiles per dimension (the fold kernel's bound); tiny grids use the
// node-parallel kernels.
bool tiles_ok(const nfftm_plan_s& p) {
  if (p.lds > 160 * 1024) return false;
  for (int d = 0; d < p.D; ++d) {
    if (p.Pd[d] > p.Nt[d]) return false;
    if (max_cover(p.T[d], p.Pd[d], p.Nt[d], p.m) > 3) return false;
  }
  return true;
}

// Window normalisation by an exact power of two so the FP32 weights stay near
// 1 (the unnormalised psi(0) reaches 1e15 at m = 8); the deconvolution factors
// carry the inverse power, so the transform is unchanged.
double window_scale(int win, int m, double sigma) {
  const double p0 = win_psi(win, 0.0, m, sigma);
  return std::ldexp(1.0, -(int)std::lround(std::log2(p0)));
}

template <typename R>
void upload_table(DevBuf<char>& buf, const std::vector<double>& v, double scale, hipError_t& e) {
  std::vector<R> h(v.size());
  for (size_t i = 0; i < v.size(); ++i) h[i] = (R)(v[i] * scale);
  e = buf.alloc(h.size() * sizeof(R));
  if (e == hipSuccess && !h.empty()) e = hipMemcpy(buf.p, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice);
}

// checkNodes (src/utils.jl:46-55): |k| <= 1/2, NaN rejected.
int check_nodes(const double* k, i64 n) {
  for (i64 i = 0; i < n; ++i)
    if (!(std::fabs(k[i]) <= 0.5))
      return fail(NFFTM_ERR_ARG, "Nodes k need to be within the range [-1/2, 1/2] but k[%lld] = %g",
                  (long long)i, k[i]);
  return NFFTM_OK;
}

// Node-dependent state (nodes!, src/implementation.jl:108-141): validation
// on the host (checkNodes, src/utils.jl:46-55, before any device work), then
// on the device (k_bin.hip): shift into [0, 1) (src/utils.jl:32-44), bin by
// tile (src/precomputation.jl:459-520 at tile granularity) with a stable
// Node records (NodeRec layout) of the plan's sorted nodes.
int build_records(nfftm_plan_s& p) {
  const size_t rb = (size_t)node_rec_bytes(p.D, p.dbl() ? sizeof(double) : sizeof(float), p.W);
  HIPCHK(p.d_rec.alloc(std::max<size_t>(1, (size_t)p.M) * rb));
  HIPCHK(launch_node_records(p.D, p.dbl(), p.W, p.kargs(), p.d_rec.p, p.stream));
  return NFFTM_OK;
}

// sort, split each tile's nodes into work items, gather the shifted nodes in
// tile order.
int set_nodes(nfftm_plan_s& p, const double* k, i64 M) {
  const int D = p.D;
  if (M < 0 || M > 0x7fffffff) return fail(NFFTM_ERR_UNSUPPORTED, "M = %lld out of range", (long long)M);
  if (int rc = check_nodes(k, (i64)D * M)) return rc;
  p.M = M;
  p.gen++;
  p.drop_graphs();
  if (k != p.k.data()) p.k.assign(k, k + (size_t)D * M);
  hipStream_t st = p.stream;
  BinArgs a;
  std::memset(&a, 0, sizeof a);
  a.D = D;
  a.M = M;
  i64 ntile = 1;
  for (int d = 0; d < 3; ++d) {
    a.Nt[d] = d < D ? (int)p.Nt[d] : 1;
    a.T[d] = p.tiled && d < D ? p.T[d] : 1;
    a.ntiles[d] = p.tiled && d < D ? p.ntiles[d] : 1;
    ntile *= a.ntiles[d];
  }
  a.ntile = (int)ntile;
  // nodes per work item: 3-D, the fixed item of the tile kernels; D <= 2,
  // twice the mean nodes per tile (256 .. 2048 / 4096), so that a crowded tile
  // (the centre of a spiral or radial trajectory) spreads over several
  // workgroups instead of serialising the launch
  if (D == 3) {
    a.CH = ITEM3;
  } else {
    const i64 mean = (M + ntile - 1) / ntile, cap = D == 2 ? 2048 : 4096;
    i64 ch = 256;
    while (ch < 2 * mean && ch < cap) ch *= 2;
    a.CH = (int)ch;
  }
  a.key_bits = 1;
  while (((i64)1 << a.key_bits) < ntile) ++a.key_bits;
  DevBuf<double> kd;
  HIPCHK(kd.alloc((size_t)D * M));
  if (M) HIPCHK(hipMemcpyAsync(kd.p, k, (size_t)D * M * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(p.d_ks.alloc((size_t)D * M));
  HIPCHK(p.d_perm.alloc((size_t)M));
  int nitems = 0, most = 0;
  if (!p.tiled) {
    HIPCHK(launch_iota(p.d_perm.p, M, st));
    HIPCHK(launch_bin_ks(a, kd.p, nullptr, p.d_ks.p, st));
    HIPCHK(p.d_items.alloc(0));
    HIPCHK(hipStreamSynchronize(st));  // before kd is freed
  } else {
    DevBuf<int> key_in, key_out, idx_in, cnt, off, ioff, stats;
    HIPCHK(key_in.alloc((size_t)M));
    HIPCHK(key_out.alloc((size_t)M));
    HIPCHK(idx_in.alloc((size_t)M));
    HIPCHK(cnt.alloc((size_t)ntile + 1));
    HIPCHK(off.alloc((size_t)ntile + 1));
    HIPCHK(ioff.alloc((size_t)ntile + 1));
    HIPCHK(stats.alloc(2));
    HIPCHK(launch_bin_keys(a, kd.p, key_in.p, idx_in.p, cnt.p, st));
    DevBuf<char> tmp;
    if (M) {
      const size_t tb = bin_sort_temp_bytes(M, a.key_bits);
      if (!tb) return fail(NFFTM_ERR_HIP, "radix sort of %lld nodes: temporary size query failed", (long long)M);
      HIPCHK(tmp.alloc(tb));
      HIPCHK(launch_bin_sort(a, tmp.p, tb, key_in.p, key_out.p, idx_in.p, p.d_perm.p, st));
    }
    HIPCHK(launch_bin_scan(a, cnt.p, off.p, ioff.p, stats.p, st));
    int hs[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(hs, stats.p, sizeof hs, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    nitems = hs[0];
    most = hs[1];
    HIPCHK(p.d_items.alloc((size_t)nitems));
    HIPCHK(p.d_tstart.alloc((size_t)ntile + 1));
    HIPCHK(launch_bin_items(a, off.p, ioff.p, p.d_items.p, p.d_tstart.p, st));
    HIPCHK(launch_bin_ks(a, kd.p, p.d_perm.p, p.d_ks.p, st));
    HIPCHK(hipStreamSynchronize(st));  // before the temporaries are freed
  }
  p.nitems = nitems;
  p.single_items = p.tiled && nitems > 0 && most <= 1;
  if (p.sparse && M > 0) {
    i64 K = 1;
    for (int d = 0; d < D; ++d) K *= p.W;
    const double bytes = (double)M * (double)K * (4.0 + (p.dbl() ? 8.0 : 4.0));
    if (K > 0x7fffffff || bytes > 128.0 * (1ull << 30))
      return fail(NFFTM_ERR_UNSUPPORTED, "sparse matrix of %lld x %lld entries (%.1f GB) is too large", (long long)M,
                  (long long)K, bytes / 1e9);
    p.sparseK = (int)K;
    HIPCHK(p.d_bcols.alloc((size_t)M * K));
    HIPCHK(p.d_bvals.alloc((size_t)M * K * (p.dbl() ? 8 : 4)));
    HIPCHK(launch_sparse_build(p.dbl(), p.W, p.kargs(), p.stream, p.d_bcols.p, p.d_bvals.p));
  }
  if (p.tiled) {
    HIPCHK(p.d_scratch.alloc(std::max<size_t>(1, (size_t)nitems) * p.slot_cells() * p.celem()));
    // node records in tile order, built on the device (1-D polynomial-window
    // plans evaluate the window in their kernels; a batch that takes such a
    // plan builds its records then)
    if (p.sorted_io()) p.d_rec.reset();
    else if (int rc = build_records(p)) return rc;
    p.d_iperm.reset();
    p.d_fsort.reset();
    if (p.sorted_io() && M > 0) {
      HIPCHK(p.d_fsort.alloc((size_t)M * p.celem()));
      HIPCHK(p.d_iperm.alloc((size_t)M));
      HIPCHK(launch_inv_perm(p.d_perm.p, p.d_iperm.p, M, p.stream));
    }
  }
  return NFFTM_OK;
}

int build_tables(nfftm_plan_s& p) {
  const int m = p.m;
  p.b = M_PI * (2.0 - 1.0 / p.sigma);
  p.wscale = window_scale(p.window, m, p.sigma);
  // bound of one window value: psi(0) for the table and the exact window; the
  // fitted polynomials stay within a hair of psi (5 % slack)
  p.wmax = win_psi(p.window, 0.0, m, p.sigma) * p.wscale *
           (p.precompute == NFFTM_LINEAR || p.precompute == NFFTM_FULL ? 1.0 : 1.05);
  hipError_t e = hipSuccess;
  if (p.precompute == NFFTM_LINEAR) {
    // precomputeLinInterp (src/precomputation.jl:291-300): K+2 samples, step m/K
    const i64 K = p.LUTSize;
    p.lut.resize(K + 2);
    const double step = (double)m / (double)K;
    for (i64 l = 0; l < K + 2; ++l) p.lut[l] = win_psi(p.window, (double)l * step, m, p.sigma);
    if (p.dbl()) upload_table<double>(p.d_lut, p.lut, p.wscale, e);
    else upload_table<float>(p.d_lut, p.lut, p.wscale, e);
  } else if (p.precompute == NFFTM_TENSOR || p.precompute == NFFTM_POLYNOMIAL) {
    p.poly = poly_fit(p.window, m, p.sigma);
    if (p.dbl()) upload_table<double>(p.d_poly, p.poly, p.wscale, e);
    else upload_table<float>(p.d_poly, p.poly, p.wscale, e);
    if (e == hipSuccess) {
      std::vector<double> pd(p.poly.size());
      for (size_t i = 0; i < pd.size(); ++i) pd[i] = p.poly[i] * p.wscale;
      e = p.d_polyd.alloc(pd.size());
      if (e == hipSuccess)
        e = hipMemcpy(p.d_polyd.p, pd.data(), pd.size() * sizeof(double), hipMemcpyHostToDevice);
    }
  }
  HIPCHK(e);
  // deconvolution factors (src/precomputation.jl:345-358): 1/cheb(j),
  // j = 1..N_d, cheb the 30-node Chebyshev interpolant on [1, N_d] of
  // kappa(j) = psi_hat(j - 1 - floor(N_d/2)); FP64, times the inverse window
  // scale.  (At m = 6, sigma = 1.25 the interpolant differs from psi_hat by
  // ~1e-11, so it is built exactly as the reference builds it.)
  p.deconv.assign(p.D, {});
  std::vector<double> flat;
  for (int d = 0; d < p.D; ++d) {
    p.deconv[d].resize(p.N[d]);
    const i64 Na = p.N[d] / 2;
    const double Ntd = (double)p.Nt[d];
    auto kappa = [&](double j) { return win_psi_hat(p.window, j - 1.0 - (double)Na, Ntd, m, p.sigma); };
    if (p.N[d] > 1) {
      ChebInterp cheb(kappa, 1.0, (double)p.N[d], 30);
      for (i64 i = 0; i < p.N[d]; ++i) p.deconv[d][i] = 1.0 / cheb((double)(i + 1)) / p.wscale;
    } else {
      p.deconv[d][0] = 1.0 / kappa(1.0) / p.wscale;
    }
    flat.insert(flat.end(), p.deconv[d].begin(), p.deconv[d].end());
  }
  HIPCHK(p.d_deconv.alloc(flat.size()));
  HIPCHK(hipMemcpy(p.d_deconv.p, flat.data(), flat.size() * sizeof(double), hipMemcpyHostToDevice));
  return NFFTM_OK;
}

// standalone: called as convolve!/convolve_transpose! on a caller grid, so the
// window normalisation is undone to give the reference's unnormalised values.
struct ProfScope {
  std::pair<hipEvent_t, hipEvent_t>* s;
  hipStream_t st;
  ProfScope(nfftm_plan_s& p, int stage) : s(p.prof_slot(stage)), st(p.stream) {
    if (s) (void)hipEventRecord(s->first, st);
  }
  ~ProfScope() {
    if (s) (void)hipEventRecord(s->second, st);
  }
};

// The adjoint's fold runs inside its first FFT pass unless
// NFFTM_FLAG_SEPARATE_FOLD is set: 3-D plans on the pruned passes (see
// launch_fold_fft3) and tiled 1-D plans on the three-pass FFT (see
// launch_fft1_adj_fold).
bool fused_fold(const nfftm_plan_s& p) {
  if (p.flags & NFFTM_FLAG_SEPARATE_FOLD) return false;
  if (p.f1) return p.tiled && p.D == 1;
  return p.pfft && p.blocked3() && p.single_items &&
         launch_fold_fft3(p.dbl(), p.W, p.kargs(), p.fargs(), nullptr, nullptr, true) == hipSuccess;
}

int run_conv(nfftm_plan_s& p, bool spread, const void* in, void* out, bool standalone = false, bool fold = true) {
  KArgs a = p.kargs();
  if (standalone) a.oscale = std::pow(p.wscale, -(double)p.D);
  hipError_t e;
  if (spread && !p.tiled) {
    ProfScope ps(p, 3);
    HIPCHK(launch_fill0(out, (size_t)p.Q * p.celem(), p.stream));
  }
  {
    ProfScope ps(p, spread ? 4 : 2);
    if (p.sorted_io() && p.d_iperm.p && !spread) {
      // interpolation results in sorted order, then gathered into the
      // caller's order
      a.fsort = p.d_fsort.p;
      e = launch_tile(p.D, false, p.dbl(), p.W, p.NT, a, p.lds, p.stream, in, nullptr);
      if (e == hipSuccess) e = launch_gather_nodes(p.dbl(), p.d_fsort.p, 1, p.d_iperm.p, out, a.fst, p.M, p.stream);
    } else if (p.tiled)
      e = launch_tile(p.D, spread, p.dbl(), p.W, p.NT, a, p.lds, p.stream, in, spread ? nullptr : out);
    else if (p.sparse)
      e = launch_sparse(spread, p.dbl(), a, p.sparseK, p.d_bcols.p, p.d_bvals.p, p.stream, in, out);
    else
      e = launch_generic(spread, p.dbl(), p.W, a, p.stream, in, out);
  }
  HIPCHK(e);
  if (spread && p.tiled && fold) {
    ProfScope ps(p, 3);
    HIPCHK(launch_fold(p.D, p.dbl(), p.W, a, p.stream, out));
  }
  return NFFTM_OK;
}

int run_deconv(nfftm_plan_s& p, bool fwd, const void* in, void* out, bool standalone = false) {
  DArgs a = p.dargs();
  if (standalone) a.oscale = std::pow(p.wscale, (double)p.D);
  ProfScope ps(p, fwd ? 0 : 6);
  if (fwd) HIPCHK(launch_deconv(p.dbl(), a, p.stream, in, out));
  else HIPCHK(launch_deconv_t(p.dbl(), a, p.stream, in, out));
  return NFFTM_OK;
}

// convolve! / convolve_transpose! on caller buffers of any element type
// (src/convolution.jl:20-53, :115-149, which take RealOrComplex arrays of any
// precision; test/convolve.jl:43-146).  The sums run in the plan's complex
// type: inputs of another type are converted into the plan's grid (grid side)
// or its node-side staging vector, outputs are rounded from there.  A real
// input has a real result (the window is real), so a real output takes the
// real part exactly.  Checked as the reference checks: complex input with
// real output first (ArgumentError), then the sizes (DimensionMismatch).
int typed_conv(nfftm_plan_s& p, bool spread, const void* in, int tin, i64 nin, void* out, int tout, i64 nout) {
  if (p.dir())
    return fail(NFFTM_ERR_UNSUPPORTED, "typed stage operators of a directional plan (use nfftm_convolve%s)",
                spread ? "_transpose" : "");
  auto valid = [](int t) { return t >= NFFTM_R32 && t <= NFFTM_C16; };
  auto cpx = [](int t) { return t == NFFTM_C32 || t == NFFTM_C64 || t == NFFTM_C16; };
  if (!valid(tin) || !valid(tout)) return fail(NFFTM_ERR_ARG, "element type codes %d, %d not supported", tin, tout);
  if (cpx(tin) && !cpx(tout))
    return fail(NFFTM_ERR_ARG, spread ? "Complex input fHat requires Complex output g"
                                      : "Complex input g requires Complex output fHat");
  const i64 want_in = spread ? p.M : p.Q, want_out = spread ? p.Q : p.M;
  if (spread) {
    if (nout != p.Q) return fail(NFFTM_ERR_DIM, "length(g) = %lld != prod(Ntilde) = %lld", (long long)nout, (long long)p.Q);
    if (nin != p.M) return fail(NFFTM_ERR_DIM, "length(fHat) = %lld != J = %lld", (long long)nin, (long long)p.M);
  } else {
    if (nin != p.Q) return fail(NFFTM_ERR_DIM, "length(g) = %lld != prod(Ntilde) = %lld", (long long)nin, (long long)p.Q);
    if (nout != p.M) return fail(NFFTM_ERR_DIM, "length(fHat) = %lld != J = %lld", (long long)nout, (long long)p.M);
  }
  if ((want_in && !in) || (want_out && !out)) return fail(NFFTM_ERR_ARG, "NULL data pointer");
  if (!spread && p.M == 0) return NFFTM_OK;
  const int native = p.dbl() ? NFFTM_C64 : NFFTM_C32;
  auto node_side = [&]() -> void* {
    const size_t bytes = (size_t)std::max<i64>(1, p.M) * p.celem();
    if (p.d_tmpM.n < bytes && p.d_tmpM.alloc(bytes) != hipSuccess) return nullptr;
    return p.d_tmpM.p;
  };
  const void* src = in;
  void* dst = out;
  if (tin != native) {
    void* st = spread ? node_side() : (void*)p.d_grid.p;
    if (!st) return fail(NFFTM_ERR_OOM, "staging vector of %lld values", (long long)p.M);
    HIPCHK(launch_convert(tin, in, native, st, want_in, p.stream));
    src = st;
  }
  if (tout != native) {
    dst = spread ? (void*)p.d_grid.p : node_side();
    if (!dst) return fail(NFFTM_ERR_OOM, "staging vector of %lld values", (long long)p.M);
  }
  if (int rc = run_conv(p, spread, src, dst, true)) return rc;
  if (tout != native) HIPCHK(launch_convert(native, dst, tout, out, want_out, p.stream));
  return NFFTM_OK;
}

// Forward deconvolution and FFT into the grid: stage 0 (deconvolve!, or the
// pruned dimension-1 pass with the deconvolution fused) and stage 1 (the FFT,
// or the remaining pruned passes), with an optional timer mark in between.
template <typename Mark>
int forward_to_grid(nfftm_plan_s& p, const void* f, Mark&& mark) {
  if (p.pfft) {
    const FArgs a = p.fargs();
    {
      ProfScope ps(p, 0);
      HIPCHK(launch_pruned_fft_fwd(p.dbl(), a, p.stream, f, p.d_grid.p, 0));
    }
    if (int rc = mark()) return rc;
    ProfScope ps(p, 1);
    HIPCHK(launch_pruned_fft_fwd(p.dbl(), a, p.stream, f, p.d_grid.p, 1));
    return NFFTM_OK;
  }
  if (p.f1) {
    const F1Args a = p.f1args();
    {
      ProfScope ps(p, 0);
      HIPCHK(launch_fft1_fwd(p.dbl(), a, p.stream, f, p.d_grid.p, 0));
    }
    if (int rc = mark()) return rc;
    ProfScope ps(p, 1);
    HIPCHK(launch_fft1_fwd(p.dbl(), a, p.stream, f, p.d_grid.p, 1));
    return NFFTM_OK;
  }
  if (int rc = run_deconv(p, true, f, p.d_grid.p)) return rc;
  if (int rc = mark()) return rc;
  ProfScope ps(p, 1);
  return p.fft->exec(true, p.d_grid.p, p.celem(), p.stream);
}

// Adjoint inverse FFT of the grid and deconvolution into f: stage 5 and
// stage 6, mark in between.  fused: the scratch slots are not folded yet and
// the first pruned pass folds them (stage 3).
template <typename Mark>
int grid_to_image(nfftm_plan_s& p, void* f, Mark&& mark, bool fused = false) {
  if (p.pfft) {
    const FArgs a = p.fargs();
    if (fused) {
      ProfScope ps(p, 3);
      HIPCHK(launch_fold_fft3(p.dbl(), p.W, p.kargs(), a, p.stream, p.d_grid.p, false));
    }
    {
      ProfScope ps(p, 5);
      HIPCHK(launch_pruned_fft_adj(p.dbl(), a, p.stream, p.d_grid.p, f, fused ? 2 : 0));
    }
    if (int rc = mark()) return rc;
    ProfScope ps(p, 6);
    HIPCHK(launch_pruned_fft_adj(p.dbl(), a, p.stream, p.d_grid.p, f, 1));
    return NFFTM_OK;
  }
  if (p.f1) {
    const F1Args a = p.f1args();
    if (fused) {
      ProfScope ps(p, 3);
      HIPCHK(launch_fft1_adj_fold(p.dbl(), a, p.kargs(), p.stream, p.d_grid.p));
    }
    {
      ProfScope ps(p, 5);
      HIPCHK(launch_fft1_adj(p.dbl(), a, p.stream, p.d_grid.p, f, fused ? 2 : 0));
    }
    if (int rc = mark()) return rc;
    ProfScope ps(p, 6);
    HIPCHK(launch_fft1_adj(p.dbl(), a, p.stream, p.d_grid.p, f, 1));
    return NFFTM_OK;
  }
  {
    ProfScope ps(p, 5);
    if (int rc = p.fft->exec(false, p.d_grid.p, p.celem(), p.stream)) return rc;
  }
  if (int rc = mark()) return rc;
  return run_deconv(p, false, p.d_grid.p, f);
}

struct Timer {
  nfftm_plan_s& p;
  bool on;
  explicit Timer(nfftm_plan_s& pp) : p(pp), on(pp.flags & NFFTM_FLAG_TIMING) {}
  int mark(int i) {
    if (on) HIPCHK(hipEventRecord(p.ev[i], p.stream));
    return NFFTM_OK;
  }
  // add: accumulate (the slices of a directional transform) instead of set
  int finish(double* t1, double* t2, double* t3, bool add = false) {
    if (!on) return NFFTM_OK;
    HIPCHK(hipEventSynchronize(p.ev[3]));
    float a = 0, b = 0, c = 0;
    HIPCHK(hipEventElapsedTime(&a, p.ev[0], p.ev[1]));
    HIPCHK(hipEventElapsedTime(&b, p.ev[1], p.ev[2]));
    HIPCHK(hipEventElapsedTime(&c, p.ev[2], p.ev[3]));
    *t1 = (add ? *t1 : 0.0) + a * 1e-3;
    *t2 = (add ? *t2 : 0.0) + b * 1e-3;
    *t3 = (add ? *t3 : 0.0) + c * 1e-3;
    return NFFTM_OK;
  }
};

int create_impl(nfftm_plan_s& p, const double* k) {
  // initParams (src/precomputation.jl:3-56)
  p.W = 2 * p.m;
  p.LUTSize = lut_size(p.m);
  p.P = p.Q = 1;
  for (int d = 0; d < p.D; ++d) {
    p.Nt[d] = ((i64)std::ceil(p.sigma_in * (double)p.N[d]) / 2) * 2;
    p.P *= p.N[d];
    p.Q *= p.Nt[d];
  }
  p.sigma = (double)p.Nt[0] / (double)p.N[0];
  for (int d = 0; d < p.D; ++d)
    if (p.Nt[d] > 0x7fffffff)
      return fail(NFFTM_ERR_UNSUPPORTED, "oversampled size %lld exceeds 2^31-1", (long long)p.Nt[d]);
  if (p.Q / 2 > 0x7fffffffLL)
    return fail(NFFTM_ERR_UNSUPPORTED, "oversampled grid of %lld points exceeds 2^32", (long long)p.Q);
  p.sparse = (p.flags & NFFTM_FLAG_SPARSE) != 0;
  p.tiled = (p.flags & NFFTM_FLAG_BLOCKING) && p.D <= 3 && !p.sparse;
  if (p.tiled) {
    choose_tiles(p);
    p.tiled = tiles_ok(p);
  }
  if (int rc = build_tables(p)) return rc;
  if (int rc = set_nodes(p, k, p.M)) return rc;
  HIPCHK(p.d_grid.alloc((size_t)p.Q * p.celem()));
  p.fft.reset(new FFTStage());
  if (int rc = p.fft->build(p.D, p.Nt, p.dbl())) return rc;
  p.pfft = !(p.flags & NFFTM_FLAG_ROCFFT) && pruned_fft_supported(p.D, p.N, p.Nt, p.dbl());
  if (p.pfft) {
    // twiddle tables exp(-2 pi i k / Nt_d) as (cos, sin) of 2 pi k / Nt_d,
    // FP64 on the host, stored in the data precision
    const size_t es = p.dbl() ? 16 : 8;
    size_t bytes = 0;
    for (int d = 0; d < p.D; ++d) {
      p.tw_off[d] = bytes;
      bytes += (size_t)p.Nt[d] * es;
    }
    std::vector<char> h(bytes);
    for (int d = 0; d < p.D; ++d) {
      const i64 L = p.Nt[d];
      for (i64 k = 0; k < L; ++k) {
        const double th = 2.0 * M_PI * (double)k / (double)L;
        const double c = std::cos(th), sn = std::sin(th);
        char* e = h.data() + p.tw_off[d] + (size_t)k * es;
        if (p.dbl()) {
          ((double*)e)[0] = c;
          ((double*)e)[1] = sn;
        } else {
          ((float*)e)[0] = (float)c;
          ((float*)e)[1] = (float)sn;
        }
      }
    }
    HIPCHK(p.d_tw.alloc(bytes));
    HIPCHK(hipMemcpy(p.d_tw.p, h.data(), bytes, hipMemcpyHostToDevice));
  }
  p.f1 = !p.pfft && p.D == 1 && !(p.flags & NFFTM_FLAG_ROCFFT) &&
         fft1_factor(p.Nt[0], p.N[0], p.dbl(), &p.f1n1, &p.f1Lm);
  if (p.f1) {
    // (cos, sin)(2 pi j / n1) and (2 pi j / Lm) in the data precision,
    // (2 pi (j << sh) / Nt) and (2 pi j / Nt), 2^sh >= sqrt(Nt), in FP64
    const i64 Nt = p.Nt[0];
    int sh = 0;
    while ((1LL << (2 * sh)) < Nt) ++sh;
    p.f1sh = sh;
    const i64 len[4] = {p.f1n1, p.f1Lm, ((Nt - 1) >> sh) + 1, 1LL << sh};
    const i64 per[4] = {p.f1n1, p.f1Lm, Nt, Nt};
    size_t bytes = 0;
    for (int i = 0; i < 4; ++i) {
      p.f1off[i] = bytes;
      bytes += (size_t)len[i] * (p.dbl() || i >= 2 ? 16 : 8);
    }
    std::vector<char> h(bytes);
    for (int i = 0; i < 4; ++i)
      for (i64 j = 0; j < len[i]; ++j) {
        const i64 e = (i == 2 ? (j << sh) : j) % per[i];
        const double th = 2.0 * M_PI * (double)e / (double)per[i];
        const bool d64 = p.dbl() || i >= 2;
        char* q = h.data() + p.f1off[i] + (size_t)j * (d64 ? 16 : 8);
        if (d64) {
          ((double*)q)[0] = std::cos(th);
          ((double*)q)[1] = std::sin(th);
        } else {
          ((float*)q)[0] = (float)std::cos(th);
          ((float*)q)[1] = (float)std::sin(th);
        }
      }
    HIPCHK(p.d_tw.alloc(bytes));
    HIPCHK(hipMemcpy(p.d_tw.p, h.data(), bytes, hipMemcpyHostToDevice));
    // deconvolution factors in the per-workgroup order of pass 1
    const int B = fft1_lines(p.f1n1, p.dbl()), Lm = p.f1Lm;
    const i64 N0 = p.N[0], Na = N0 / 2;
    std::vector<double> f1((size_t)Nt, 0.0);
    for (i64 i = 0; i < N0; ++i) {
      const i64 g = i < Na ? i - Na + Nt : i - Na;
      const i64 j1 = g / ((i64)Lm * Lm), j2 = g / Lm % Lm, j3 = g % Lm;
      f1[(size_t)(((j2 * (Lm / B) + j3 / B) * p.f1n1 + j1) * B + j3 % B)] = p.deconv[0][(size_t)i];
    }
    HIPCHK(p.d_fac1.alloc(f1.size()));
    HIPCHK(hipMemcpy(p.d_fac1.p, f1.data(), f1.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  for (auto& e : p.ev) HIPCHK(hipEventCreate(&e));
  return NFFTM_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: tile edge and workgroup size of the gridding kernels, chosen by
// timing spreading (with its fold) plus interpolation on the plan's own nodes.
// ---------------------------------------------------------------------------
struct TuneChoice {
  int t = 0, NT = 0;
  double ms = 0;
};

int time_gridding(nfftm_plan_s& p, void* fh, double* ms) {
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = NFFTM_OK;
  float best = 1e30f;
  for (int rep = 0; rep < 4 && rc == NFFTM_OK; ++rep) {
    (void)hipEventRecord(e0, p.stream);
    rc = run_conv(p, true, fh, p.d_grid.p);
    if (rc == NFFTM_OK) rc = run_conv(p, false, p.d_grid.p, fh);
    (void)hipEventRecord(e1, p.stream);
    if (hipEventSynchronize(e1) != hipSuccess) rc = fail(NFFTM_ERR_HIP, "tuning run failed");
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (rep > 0) best = std::min(best, t);  // first run warms up
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = best;
  return rc;
}

// Candidate shapes: D = 2 tile edges 16..128, D = 1 edges 512..8192, with 256
// or 512 threads (maxNT caps it); D = 3: both compile-time geometries with 512
// or 1024 threads.  Plans of more than 2^20 nodes are timed on every k-th node
// (the same spatial distribution at a bounded binning cost); the nodes are
// binned once per geometry.  Leaves the fastest applied to all nodes.
int tune_plan(nfftm_plan_s& p, int maxNT, std::vector<TuneChoice>* log) {
  if (!p.tiled || p.M == 0) return NFFTM_OK;
  const i64 Mfull = p.M;
  const std::vector<double> kfull = p.k;
  const i64 stride = (Mfull + (1 << 20) - 1) >> 20;
  std::vector<double> ksub;
  if (stride > 1) {
    for (i64 j = 0; j < Mfull; j += stride) ksub.insert(ksub.end(), &kfull[(size_t)j * p.D], &kfull[(size_t)(j + 1) * p.D]);
  }
  const double* ktune = stride > 1 ? ksub.data() : kfull.data();
  const i64 Mtune = stride > 1 ? (i64)(ksub.size() / p.D) : Mfull;
  DevBuf<char> fh;
  HIPCHK(fh.alloc((size_t)Mtune * p.celem()));
  HIPCHK(launch_fill0(fh.p, (size_t)Mtune * p.celem(), p.stream));
  std::vector<std::pair<int, int>> cand;  // (D <= 2: tile edge, D = 3: variant), threads
  if (p.D == 3) {
    for (int v : {0, 1})
      for (int nt : {512, 1024})
        if (nt <= maxNT) cand.push_back({v, nt});
  } else {
    const std::vector<int> ts = p.D == 2 ? std::vector<int>{16, 32, 64, 128}
                                         : std::vector<int>{512, 1024, 2048, 4096, 8192};
    for (int t : ts)
      for (int nt : {256, 512})
        if (nt <= maxNT) cand.push_back({t, nt});
  }
  TuneChoice best;
  best.ms = 1e30;
  int binned = -1;
  // every exit, failed or not, leaves the plan binned on all of its nodes
  auto restore = [&](int rc) {
    if (rc != NFFTM_OK || best.NT == 0) {
      choose_tiles(p);
    } else {
      if (p.D == 3) set_tile3(p, best.t);
      else set_tile_shape(p, best.t);
      p.NT = best.NT;
    }
    const std::string err = g_err;
    const int rc2 = set_nodes(p, kfull.data(), Mfull);
    if (rc != NFFTM_OK) {
      g_err = err;  // report the tuning failure, not the restore
      return rc;
    }
    return rc2;
  };
  for (auto& c : cand) {
    const bool ok = p.D == 3 ? set_tile3(p, c.first) : set_tile_shape(p, c.first);
    if (!ok || !tiles_ok(p)) continue;
    if (c.first != binned) {
      if (int rc = set_nodes(p, ktune, Mtune)) return restore(rc);
      binned = c.first;
    }
    p.NT = c.second;
    double ms = 0;
    if (int rc = time_gridding(p, fh.p, &ms)) return restore(rc);
    if (log) log->push_back(TuneChoice{c.first, c.second, ms});
    if (ms < best.ms) best = TuneChoice{c.first, c.second, ms};
  }
  return restore(NFFTM_OK);
}

}  // namespace

// ---------------------------------------------------------------------------
// Batches: many plans (and several images per plan's nodes) per launch.
// ---------------------------------------------------------------------------
struct nfftm_batch_s {
  std::vector<nfftm_plan_s*> plans;
  std::vector<int> gens;  // plan generations the batch was built against
  int S = 0;  // slots = sum of images per plan
  bool dbl = false;
  size_t celem = 8;
  std::vector<int> slot_plan;
  std::vector<KArgs> ka;  // host copies (scratch of each slot)
  std::vector<DArgs> da;
  DevBuf<KArgs> d_ka;
  DevBuf<DArgs> d_da;
  DevBuf<char> grids, scratch;
  std::vector<void*> grid_ptr;
  DevBuf<void*> d_grid_ptr, d_tab;  // d_tab: S input then S output pointers of a call
  void** h_tab = nullptr;           // pinned staging of d_tab
  hipEvent_t tab_done = nullptr;
  hipStream_t last = nullptr;      // stream of the previous call
  bool called = false;
  hipEvent_t switch_ev = nullptr;  // orders calls on different streams
  struct TileGroup {
    int D = 0, W = 0, NT = 0, nblk = 0, nwork = 0;
    size_t lds = 0;
    DevBuf<int2> work;
  };
  std::vector<std::unique_ptr<TileGroup>> tg;
  struct RowWork {
    DevBuf<int4> work;
    int nwork = 0, nblk = 0;
  };
  RowWork dec, dect, fold[2];
  struct FFTGroup {
    std::unique_ptr<FFTStage> fft;
    void* base = nullptr;
  };
  std::vector<FFTGroup> fg;
  hipStream_t cap = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  bool use_graph = false;

  ~nfftm_batch_s() {
    for (int i = 0; i < 2; ++i) {
      if (gexec[i]) (void)hipGraphExecDestroy(gexec[i]);
      if (graph[i]) (void)hipGraphDestroy(graph[i]);
    }
    if (tab_done) (void)hipEventDestroy(tab_done);
    if (switch_ev) (void)hipEventDestroy(switch_ev);
    if (h_tab) (void)hipHostFree(h_tab);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace {

inline int blocks_for(int nwork, int cap) { return std::max(8, std::min(((nwork + 7) / 8) * 8, cap)); }

// The launches of one direction on stream s, reading the call's data
// pointers from d_tab (inputs at [0, S), outputs at [S, 2S)).
int batch_enqueue(nfftm_batch_s& b, bool forward, hipStream_t s) {
  void* const* tin = b.d_tab.p;
  void* const* tout = b.d_tab.p + b.S;
  void* const* grids = b.d_grid_ptr.p;
  if (forward) {
    HIPCHK(launch_deconv_batch(b.dbl, b.dec.nblk, s, b.d_da.p, b.dec.work.p, b.dec.nwork, (const void* const*)tin,
                               grids));
    for (auto& g : b.fg)
      if (int rc = g.fft->exec(true, g.base, b.celem, s)) return rc;
    for (auto& g : b.tg)
      HIPCHK(launch_tile_batch(g->D, false, b.dbl, g->W, g->NT, g->nblk, g->lds, s, b.d_ka.p, g->work.p, g->nwork,
                               (const void* const*)grids, tout));
  } else {
    for (auto& g : b.tg)
      HIPCHK(launch_tile_batch(g->D, true, b.dbl, g->W, g->NT, g->nblk, g->lds, s, b.d_ka.p, g->work.p, g->nwork,
                               (const void* const*)tin, nullptr));
    for (int d = 0; d < 2; ++d)
      HIPCHK(launch_fold_batch(d + 1, b.dbl, b.fold[d].nblk, s, b.d_ka.p, b.fold[d].work.p, b.fold[d].nwork, grids));
    for (auto& g : b.fg)
      if (int rc = g.fft->exec(false, g.base, b.celem, s)) return rc;
    HIPCHK(launch_deconv_t_batch(b.dbl, b.dect.nblk, s, b.d_da.p, b.dect.work.p, b.dect.nwork,
                                 (const void* const*)grids, tout));
  }
  return NFFTM_OK;
}

template <typename T>
int upload(DevBuf<T>& d, const std::vector<T>& h) {
  HIPCHK(d.alloc(std::max<size_t>(1, h.size())));
  if (!h.empty()) HIPCHK(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return NFFTM_OK;
}

int batch_build(nfftm_batch_s& b, const nfftm_plan* plans, const int* ntrans, int nplans, int flags) {
  b.dbl = plans[0]->dbl();
  b.celem = plans[0]->celem();
  for (int i = 0; i < nplans; ++i) {
    nfftm_plan_s& p = *plans[i];
    if (p.dbl() != b.dbl) return fail(NFFTM_ERR_ARG, "plans of a batch must share the precision");
    if (p.D > 2) return fail(NFFTM_ERR_UNSUPPORTED, "plan %d: batches take 1-D and 2-D plans", i);
    if (p.dir()) return fail(NFFTM_ERR_UNSUPPORTED, "plan %d: a directional plan batches its own slices", i);
    if (!p.tiled) return fail(NFFTM_ERR_UNSUPPORTED, "plan %d: grid too small for the tiled kernels", i);
    const int K = ntrans ? ntrans[i] : 1;
    if (K < 1) return fail(NFFTM_ERR_ARG, "ntrans[%d] = %d must be positive", i, K);
    b.plans.push_back(&p);
    for (int c = 0; c < K; ++c) b.slot_plan.push_back(i);
  }
  b.S = (int)b.slot_plan.size();
  if (flags & NFFTM_BATCH_AUTOTUNE)
    for (auto* p : b.plans)
      if (int rc = tune_plan(*p, 512, nullptr)) return rc;
  for (auto* p : b.plans) {
    // the batch kernels read node records, which 1-D polynomial-window plans
    // do not keep for themselves
    if (!p->d_rec.p)
      if (int rc = build_records(*p)) return rc;
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->NT != 256 && p->NT != 512) return fail(NFFTM_ERR_UNSUPPORTED, "batch kernels take 256 or 512 threads");
    b.gens.push_back(p->gen);
  }

  // grids grouped by size (one batched rocFFT per group), scratch per slot
  std::vector<int> order(b.S);
  for (int s = 0; s < b.S; ++s) order[s] = s;
  auto key = [&](int s) {
    const nfftm_plan_s& p = *b.plans[b.slot_plan[s]];
    return std::make_pair(p.D, std::vector<i64>(p.Nt, p.Nt + p.D));
  };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key(x) < key(y); });
  std::vector<size_t> goff(b.S), soff(b.S);
  size_t gtot = 0, stot = 0;
  for (int s : order) {
    goff[s] = gtot;
    gtot += (size_t)b.plans[b.slot_plan[s]]->Q * b.celem;
  }
  for (int s = 0; s < b.S; ++s) {
    const nfftm_plan_s& p = *b.plans[b.slot_plan[s]];
    soff[s] = stot;
    stot += (size_t)std::max(1, p.nitems) * p.Pd[0] * p.Pd[1] * p.Pd[2] * b.celem;
    stot = (stot + 255) & ~(size_t)255;
  }
  HIPCHK(b.grids.alloc(gtot));
  HIPCHK(b.scratch.alloc(stot));
  b.grid_ptr.resize(b.S);
  b.ka.resize(b.S);
  b.da.resize(b.S);
  for (int s = 0; s < b.S; ++s) {
    const nfftm_plan_s& p = *b.plans[b.slot_plan[s]];
    b.grid_ptr[s] = b.grids.p + goff[s];
    b.ka[s] = p.kargs();
    b.ka[s].scratch = b.scratch.p + soff[s];
    b.da[s] = p.dargs();
  }
  for (size_t i = 0; i < order.size();) {
    size_t j = i;
    while (j < order.size() && key(order[j]) == key(order[i])) ++j;
    const nfftm_plan_s& p = *b.plans[b.slot_plan[order[i]]];
    nfftm_batch_s::FFTGroup g;
    g.fft.reset(new FFTStage());
    if (int rc = g.fft->build(p.D, p.Nt, b.dbl, j - i)) return rc;
    g.base = b.grid_ptr[order[i]];
    b.fg.push_back(std::move(g));
    i = j;
  }

  // work lists: tile items per (D, W, NT) group; rows or pairs per slot
  std::vector<std::vector<int2>> tw;
  std::vector<int4> dec, dect, fold[2];
  int dev = 0, cus = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  for (int s = 0; s < b.S; ++s) {
    const nfftm_plan_s& p = *b.plans[b.slot_plan[s]];
    int gi = -1;
    for (size_t g = 0; g < b.tg.size(); ++g)
      if (b.tg[g]->D == p.D && b.tg[g]->W == p.W && b.tg[g]->NT == p.NT) gi = (int)g;
    if (gi < 0) {
      b.tg.emplace_back(new nfftm_batch_s::TileGroup());
      b.tg.back()->D = p.D;
      b.tg.back()->W = p.W;
      b.tg.back()->NT = p.NT;
      tw.emplace_back();
      gi = (int)b.tg.size() - 1;
    }
    b.tg[gi]->lds = std::max(b.tg[gi]->lds, p.lds);
    for (int it = 0; it < p.nitems; ++it) tw[gi].push_back(int2{s, it});
    const i64 npairs = p.Q / 2, CHP = 2048;
    for (i64 a0 = 0; a0 < npairs; a0 += CHP) dec.push_back(int4{s, (int)a0, (int)std::min(npairs, a0 + CHP), 0});
    const i64 rows = p.P / p.N[0], CR = std::max<i64>(1, 1024 / p.N[0]);
    for (i64 r0 = 0; r0 < rows; r0 += CR) dect.push_back(int4{s, (int)r0, (int)std::min(rows, r0 + CR), 0});
    if (p.D == 2) {
      const i64 CF = std::max<i64>(1, 1024 / p.Nt[0]);
      for (i64 r0 = 0; r0 < p.Nt[1]; r0 += CF)
        fold[1].push_back(int4{s, (int)r0, (int)std::min<i64>(p.Nt[1], r0 + CF), 0});
    } else {
      const i64 CF = 4096;
      for (i64 x0 = 0; x0 < p.Nt[0]; x0 += CF)
        fold[0].push_back(int4{s, (int)x0, (int)std::min<i64>(p.Nt[0], x0 + CF), 0});
    }
  }
  for (size_t g = 0; g < b.tg.size(); ++g) {
    auto& G = *b.tg[g];
    if (int rc = upload(G.work, tw[g])) return rc;
    G.nwork = (int)tw[g].size();
    G.nblk = std::min(tile_batch_blocks(G.D, false, b.dbl, G.W, G.NT, G.lds), ((G.nwork + 7) / 8) * 8);
    G.nblk = std::max(G.nblk, 8);
  }
  if (int rc = upload(b.dec.work, dec)) return rc;
  b.dec.nwork = (int)dec.size();
  b.dec.nblk = blocks_for(b.dec.nwork, 8 * cus);
  if (int rc = upload(b.dect.work, dect)) return rc;
  b.dect.nwork = (int)dect.size();
  b.dect.nblk = blocks_for(b.dect.nwork, 8 * cus);
  for (int d = 0; d < 2; ++d) {
    if (int rc = upload(b.fold[d].work, fold[d])) return rc;
    b.fold[d].nwork = (int)fold[d].size();
    b.fold[d].nblk = blocks_for(b.fold[d].nwork, 8 * cus);
  }
  if (int rc = upload(b.d_ka, b.ka)) return rc;
  if (int rc = upload(b.d_da, b.da)) return rc;
  if (int rc = upload(b.d_grid_ptr, b.grid_ptr)) return rc;
  HIPCHK(b.d_tab.alloc(2 * (size_t)b.S));
  HIPCHK(hipHostMalloc((void**)&b.h_tab, 2 * (size_t)b.S * sizeof(void*), hipHostMallocDefault));
  HIPCHK(hipEventCreate(&b.tab_done));
  HIPCHK(hipStreamCreate(&b.cap));

  // one eager pass of each direction on zero data (sets kernel attributes,
  // lets rocFFT finish its lazy set-up), then capture both directions
  {
    DevBuf<char> zi, zo;
    size_t mi = 0, mo = 0;
    for (int s = 0; s < b.S; ++s) {
      const nfftm_plan_s& p = *b.plans[b.slot_plan[s]];
      mi = std::max<size_t>(mi, (size_t)std::max<i64>(p.P, p.M));
      mo = mi;
    }
    HIPCHK(zi.alloc(std::max<size_t>(1, mi) * b.celem));
    HIPCHK(zo.alloc(std::max<size_t>(1, mo) * b.celem));
    HIPCHK(launch_fill0(zi.p, mi * b.celem, b.cap));
    for (int s = 0; s < b.S; ++s) {
      b.h_tab[s] = zi.p;
      b.h_tab[b.S + s] = zo.p;
    }
    HIPCHK(hipMemcpyAsync(b.d_tab.p, b.h_tab, 2 * (size_t)b.S * sizeof(void*), hipMemcpyHostToDevice, b.cap));
    for (int dir = 0; dir < 2; ++dir)
      if (int rc = batch_enqueue(b, dir == 0, b.cap)) return rc;
    HIPCHK(hipStreamSynchronize(b.cap));
  }
  b.use_graph = !(flags & NFFTM_BATCH_NO_GRAPH);
  for (int dir = 0; dir < 2 && b.use_graph; ++dir) {
    if (hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      b.use_graph = false;
      break;
    }
    const int rc = batch_enqueue(b, dir == 0, b.cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b.cap, &g);
    if (rc != NFFTM_OK || e != hipSuccess || !g ||
        hipGraphInstantiate(&b.gexec[dir], g, nullptr, nullptr, 0) != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      b.use_graph = false;
      (void)hipGetLastError();
      g_err.clear();
      break;
    }
    b.graph[dir] = g;
  }
  return NFFTM_OK;
}

int batch_run(nfftm_batch_s& b, bool forward, const void* const* in, void* const* out, hipStream_t s) {
  for (size_t i = 0; i < b.plans.size(); ++i)
    if (b.plans[i]->gen != b.gens[i])
      return fail(NFFTM_ERR_ARG, "plan %d of the batch changed its nodes; create the batch again", (int)i);
  for (int i = 0; i < b.S; ++i)
    if (!in[i] || !out[i]) return fail(NFFTM_ERR_ARG, "NULL data pointer for slot %d", i);
  if (b.called)
    if (int rc = order_streams(b.last, s, &b.switch_ev)) return rc;
  b.last = s;
  b.called = true;
  HIPCHK(hipEventSynchronize(b.tab_done));  // the previous call has read the staging table
  for (int i = 0; i < b.S; ++i) {
    b.h_tab[i] = const_cast<void*>(in[i]);
    b.h_tab[b.S + i] = out[i];
  }
  HIPCHK(hipMemcpyAsync(b.d_tab.p, b.h_tab, 2 * (size_t)b.S * sizeof(void*), hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(b.tab_done, s));
  if (b.use_graph) {
    HIPCHK(hipGraphLaunch(b.gexec[forward ? 0 : 1], s));
    return NFFTM_OK;
  }
  return batch_enqueue(b, forward, s);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* nfftm_last_error(void) { return g_err.c_str(); }
const char* nfftm_version(void) { return "nfftm 0.1 (gfx950)"; }

int nfftm_lds_issue_cycles(double cycles[4], double* clock_hz, int* cus) {
  g_err.clear();
  if (!cycles || !clock_hz || !cus) return fail(NFFTM_ERR_ARG, "NULL argument");
  HIPCHK(lds_issue_cycles(cycles, clock_hz, cus));
  return NFFTM_OK;
}

int nfftm_plan_create(nfftm_plan* out, int prec, int D, const int64_t* N, int64_t M, const double* k_host,
                      int m, double sigma, int window, int precompute, int flags, void* hip_stream) {
  g_err.clear();
  if (!out) return fail(NFFTM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (prec != NFFTM_F32 && prec != NFFTM_F64 && prec != NFFTM_MIXED)
    return fail(NFFTM_ERR_UNSUPPORTED, "precision %d not supported", prec);
  if (D < 1 || D > MAXD) return fail(NFFTM_ERR_UNSUPPORTED, "D = %d not supported (1..%d)", D, MAXD);
  if (!N) return fail(NFFTM_ERR_ARG, "N is NULL");
  for (int d = 0; d < D; ++d)
    if (N[d] < 1) return fail(NFFTM_ERR_ARG, "N[%d] = %lld must be positive", d, (long long)N[d]);
  if (M > 0 && !k_host) return fail(NFFTM_ERR_ARG, "k is NULL");
  if (window < NFFTM_KAISER_BESSEL || window > NFFTM_COSH_TYPE)
    return fail(NFFTM_ERR_UNSUPPORTED, "Window %d not yet implemented!", window);
  if (precompute < NFFTM_FULL || precompute > NFFTM_POLYNOMIAL)
    return fail(NFFTM_ERR_UNSUPPORTED, "precompute = %d not supported", precompute);
  if (m < 1 || m > 8) return fail(NFFTM_ERR_UNSUPPORTED, "m = %d not supported (1..8)", m);
  if (!(sigma >= 1.0) || !std::isfinite(sigma))
    return fail(NFFTM_ERR_ARG, "oversampling factor sigma = %g must be >= 1", sigma);
  std::unique_ptr<nfftm_plan_s> p(new nfftm_plan_s());
  p->prec = prec == NFFTM_F64 ? NFFTM_F64 : NFFTM_F32;
  p->D = D;
  p->m = m;
  p->window = window;
  p->precompute = precompute;
  p->flags = flags;
  p->sigma_in = sigma;
  p->M = M;
  p->stream = (hipStream_t)hip_stream;
  for (int d = 0; d < D; ++d) p->N[d] = N[d];
  // node validation happens before any device work (checkNodes in the ctor)
  if (int rc = check_nodes(k_host, (i64)D * M)) return rc;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(NFFTM_ERR_HIP, "no HIP device available");
  if (int rc = create_impl(*p, k_host)) return rc;
  *out = p.release();
  return NFFTM_OK;
}

int nfftm_plan_create_dims(nfftm_plan* out, int prec, int D, const int64_t* N, int dim_lo, int dim_hi, int64_t M,
                           const double* k_host, int m, double sigma, int window, int precompute, int flags,
                           void* hip_stream) {
  g_err.clear();
  if (!out) return fail(NFFTM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (D < 1 || D > MAXD) return fail(NFFTM_ERR_UNSUPPORTED, "D = %d not supported (1..%d)", D, MAXD);
  if (!N) return fail(NFFTM_ERR_ARG, "N is NULL");
  for (int d = 0; d < D; ++d)
    if (N[d] < 1) return fail(NFFTM_ERR_ARG, "N[%d] = %lld must be positive", d, (long long)N[d]);
  if (dim_lo < 0 || dim_hi < dim_lo || dim_hi >= D)
    return fail(NFFTM_ERR_ARG, "dims = %d:%d is not a range of the axes 0..%d", dim_lo, dim_hi, D - 1);
  nfftm_plan p = nullptr;
  if (int rc = nfftm_plan_create(&p, prec, dim_hi - dim_lo + 1, N + dim_lo, M, k_host, m, sigma, window,
                                 precompute, flags, hip_stream))
    return rc;
  if (dim_hi - dim_lo + 1 < D) {
    p->Dfull = D;
    p->dlo = dim_lo;
    for (int d = 0; d < D; ++d) {
      p->Nfull[d] = N[d];
      if (d < dim_lo) p->Bsl *= N[d];
      if (d > dim_hi) p->Asl *= N[d];
    }
  }
  *out = p;
  return NFFTM_OK;
}

int nfftm_nodes(nfftm_plan p, const double* k_host, int64_t M) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  if (M > 0 && !k_host) return fail(NFFTM_ERR_ARG, "k is NULL");
  if (int rc = check_nodes(k_host, (i64)p->D * M)) return rc;
  HIPCHK(hipStreamSynchronize(p->stream));
  return set_nodes(*p, k_host, M);
}

}  // extern "C"

namespace {

// fn(in slice, out slice) for every slice of a directional plan's caller
// arrays (in_len and out_len elements per slice), with the kernels' element
// stride set to the slice stride meanwhile: slice (b, a) starts at element
// b + a * Bsl * len (src/directional.jl: the axes before dims, then the
// transformed ones, then those after).  Ordinary plans: fn(in, out) once.
template <typename F>
int for_slices(nfftm_plan_s& p, const void* in, i64 in_len, void* out, i64 out_len, F&& fn) {
  if (!p.dir()) return fn(in, out);
  struct Reset {
    nfftm_plan_s& p;
    ~Reset() { p.io_stride = 1; }
  } reset{p};
  p.io_stride = p.Bsl;
  const size_t ce = p.celem();
  for (i64 a = 0; a < p.Asl; ++a)
    for (i64 b = 0; b < p.Bsl; ++b) {
      const void* is = in ? (const char*)in + (size_t)(b + a * p.Bsl * in_len) * ce : nullptr;
      void* os = out ? (char*)out + (size_t)(b + a * p.Bsl * out_len) * ce : nullptr;
      if (int rc = fn(is, os)) return rc;
    }
  return NFFTM_OK;
}

// Launch sequence of one direction on p.stream: forward = deconvolve, FFT,
// interpolate; adjoint = spread (+ fold), backward FFT, deconvolve_transpose;
// for every slice of a directional plan.
int enqueue_transform(nfftm_plan_s& p, bool fwd, const void* in, void* out) {
  auto none = [] { return (int)NFFTM_OK; };
  return for_slices(p, in, fwd ? p.P : p.M, out, fwd ? p.M : p.P, [&](const void* i, void* o) -> int {
    if (fwd) {
      if (int rc = forward_to_grid(p, i, none)) return rc;
      if (p.M) return run_conv(p, false, p.d_grid.p, o);
      return NFFTM_OK;
    }
    const bool fused = fused_fold(p);
    if (int rc = run_conv(p, true, i, p.d_grid.p, false, !fused)) return rc;
    return grid_to_image(p, o, none, fused);
  });
}

// Graph replay of enqueue_transform for a repeated (input, output) pair.
int graph_transform(nfftm_plan_s& p, bool fwd, const void* in, void* out) {
  const int dir = fwd ? 0 : 1;
  for (auto& g : p.graphs) {
    if (g.dir != dir || g.in != in || g.out != out) continue;
    ++g.uses;
    if (!g.exec && p.graphs_ok) {
      if (!p.cap) HIPCHK(hipStreamCreate(&p.cap));
      hipStream_t saved = p.stream;
      p.stream = p.cap;
      int rc = NFFTM_ERR_HIP;
      hipGraph_t graph = nullptr;
      if (hipStreamBeginCapture(p.cap, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        rc = enqueue_transform(p, fwd, in, out);
        if (hipStreamEndCapture(p.cap, &graph) != hipSuccess) rc = NFFTM_ERR_HIP;
      }
      p.stream = saved;
      if (rc == NFFTM_OK && graph && hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) != hipSuccess)
        g.exec = nullptr;
      if (graph) (void)hipGraphDestroy(graph);
      if (!g.exec) {  // capture not possible here: launch directly from now on
        p.graphs_ok = false;
        (void)hipGetLastError();
        g_err.clear();
      }
    }
    if (g.exec) {
      HIPCHK(hipGraphLaunch(g.exec, p.stream));
      return NFFTM_OK;
    }
    return enqueue_transform(p, fwd, in, out);
  }
  if (p.graphs.size() >= 8) {
    if (p.graphs.front().exec) (void)hipGraphExecDestroy(p.graphs.front().exec);
    p.graphs.erase(p.graphs.begin());
  }
  p.graphs.push_back(nfftm_plan_s::GraphEntry{dir, in, out, 1, nullptr});
  return enqueue_transform(p, fwd, in, out);
}

}  // namespace

extern "C" {

int nfftm_forward(nfftm_plan p, const void* f, void* fhat) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  if ((!f && p->P) || (!fhat && p->M)) return fail(NFFTM_ERR_ARG, "NULL data pointer");
  if ((p->flags & NFFTM_FLAG_GRAPHS) && !p->profiling && !(p->flags & NFFTM_FLAG_TIMING))
    return graph_transform(*p, true, f, fhat);
  const bool acc = p->dir();
  if (acc) p->timing[1] = p->timing[2] = p->timing[3] = 0;
  return for_slices(*p, f, p->P, fhat, p->M, [&](const void* fs, void* fhs) -> int {
    Timer t(*p);
    if (int rc = t.mark(0)) return rc;
    if (int rc = forward_to_grid(*p, fs, [&] { return t.mark(1); })) return rc;
    if (int rc = t.mark(2)) return rc;
    if (p->M) {
      if (int rc = run_conv(*p, false, p->d_grid.p, fhs)) return rc;
    }
    if (int rc = t.mark(3)) return rc;
    return t.finish(&p->timing[3], &p->timing[2], &p->timing[1], acc);
  });
}

int nfftm_adjoint(nfftm_plan p, const void* fhat, void* f) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  if ((!f && p->P) || (!fhat && p->M)) return fail(NFFTM_ERR_ARG, "NULL data pointer");
  if ((p->flags & NFFTM_FLAG_GRAPHS) && !p->profiling && !(p->flags & NFFTM_FLAG_TIMING))
    return graph_transform(*p, false, fhat, f);
  const bool acc = p->dir();
  if (acc) p->timing[4] = p->timing[5] = p->timing[6] = 0;
  return for_slices(*p, fhat, p->M, f, p->P, [&](const void* fhs, void* fs) -> int {
    Timer t(*p);
    if (int rc = t.mark(0)) return rc;
    const bool fused = fused_fold(*p);
    if (int rc = run_conv(*p, true, fhs, p->d_grid.p, false, !fused)) return rc;
    if (int rc = t.mark(1)) return rc;
    if (int rc = grid_to_image(*p, fs, [&] { return t.mark(2); }, fused)) return rc;
    if (int rc = t.mark(3)) return rc;
    return t.finish(&p->timing[4], &p->timing[5], &p->timing[6], acc);
  });
}

static int grad_check(nfftm_plan p, const void* in, i64 nin, const void* ybar, double* kbar) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  if (p->window != NFFTM_KAISER_BESSEL && (p->precompute == NFFTM_LINEAR || p->precompute == NFFTM_FULL))
    return fail(NFFTM_ERR_UNSUPPORTED,
                "node gradient of window %d needs precompute TENSOR or POLYNOMIAL (the analytic window "
                "derivative is built for Kaiser-Bessel)", p->window);
  if ((!in && nin) || (!ybar && p->M) || (!kbar && p->M)) return fail(NFFTM_ERR_ARG, "NULL data pointer");
  if (p->dir()) return fail(NFFTM_ERR_UNSUPPORTED, "node gradient of a directional plan");
  return NFFTM_OK;
}

int nfftm_grad_nodes(nfftm_plan p, const void* f, const void* ybar, double* kbar) {
  if (int rc = grad_check(p, f, p ? p->P : 0, ybar, kbar)) return rc;
  if (int rc = forward_to_grid(*p, f, [] { return (int)NFFTM_OK; })) return rc;
  if (p->M) {
    ProfScope ps(*p, 7);
    HIPCHK(launch_grad_nodes(p->dbl(), p->W, p->kargs(), p->stream, p->d_grid.p, ybar, kbar));
  }
  return NFFTM_OK;
}

int nfftm_grad_nodes_grid(nfftm_plan p, const void* g, const void* ybar, double* kbar) {
  if (int rc = grad_check(p, g, p ? p->Q : 0, ybar, kbar)) return rc;
  if (p->M) {
    ProfScope ps(*p, 7);
    HIPCHK(launch_grad_nodes(p->dbl(), p->W, p->kargs(), p->stream, g, ybar, kbar));
  }
  return NFFTM_OK;
}

// The stage operators of a directional plan (src/directional.jl:1-156) take
// the whole arrays: f and g along all axes, the grid oversampled along dims
// only.  Each slice runs through the plan's own grid, copied from or to the
// caller's strided grid slice.
static int grid_in(nfftm_plan_s& p, const void* gs) {
  HIPCHK(launch_copy_strided(p.dbl(), gs, p.Bsl, p.d_grid.p, 1, p.Q, p.stream));
  return NFFTM_OK;
}
static int grid_out(nfftm_plan_s& p, void* gs) {
  HIPCHK(launch_copy_strided(p.dbl(), p.d_grid.p, 1, gs, p.Bsl, p.Q, p.stream));
  return NFFTM_OK;
}

int nfftm_deconvolve(nfftm_plan p, const void* f, void* g) {
  g_err.clear();
  if (!p || !f || !g) return fail(NFFTM_ERR_ARG, "NULL argument");
  if (!p->dir()) return run_deconv(*p, true, f, g, true);
  return for_slices(*p, f, p->P, g, p->Q, [&](const void* fs, void* gs) -> int {
    if (int rc = run_deconv(*p, true, fs, p->d_grid.p, true)) return rc;
    return grid_out(*p, gs);
  });
}

int nfftm_deconvolve_transpose(nfftm_plan p, const void* g, void* f) {
  g_err.clear();
  if (!p || !f || !g) return fail(NFFTM_ERR_ARG, "NULL argument");
  if (!p->dir()) return run_deconv(*p, false, g, f, true);
  return for_slices(*p, g, p->Q, f, p->P, [&](const void* gs, void* fs) -> int {
    if (int rc = grid_in(*p, gs)) return rc;
    return run_deconv(*p, false, p->d_grid.p, fs, true);
  });
}

int nfftm_convolve(nfftm_plan p, const void* g, void* fhat) {
  g_err.clear();
  if (!p || !g || (!fhat && p->M)) return fail(NFFTM_ERR_ARG, "NULL argument");
  if (!p->M) return NFFTM_OK;
  if (!p->dir()) return run_conv(*p, false, g, fhat, true);
  return for_slices(*p, g, p->Q, fhat, p->M, [&](const void* gs, void* fhs) -> int {
    if (int rc = grid_in(*p, gs)) return rc;
    return run_conv(*p, false, p->d_grid.p, fhs, true);
  });
}

int nfftm_convolve_transpose(nfftm_plan p, const void* fhat, void* g) {
  g_err.clear();
  if (!p || !g || (!fhat && p->M)) return fail(NFFTM_ERR_ARG, "NULL argument");
  if (!p->dir()) return run_conv(*p, true, fhat, g, true);
  return for_slices(*p, fhat, p->M, g, p->Q, [&](const void* fhs, void* gs) -> int {
    if (int rc = run_conv(*p, true, fhs, p->d_grid.p, true)) return rc;
    return grid_out(*p, gs);
  });
}

int nfftm_convolve_typed(nfftm_plan p, const void* g, int g_type, int64_t g_len, void* fhat, int fhat_type,
                         int64_t fhat_len) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  return typed_conv(*p, false, g, g_type, g_len, fhat, fhat_type, fhat_len);
}

int nfftm_convolve_transpose_typed(nfftm_plan p, const void* fhat, int fhat_type, int64_t fhat_len, void* g,
                                   int g_type, int64_t g_len) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  return typed_conv(*p, true, fhat, fhat_type, fhat_len, g, g_type, g_len);
}

int nfftm_fft(nfftm_plan p, void* g, int direction) {
  g_err.clear();
  if (!p || !g) return fail(NFFTM_ERR_ARG, "NULL argument");
  if (direction != -1 && direction != 1) return fail(NFFTM_ERR_ARG, "direction must be -1 or +1");
  if (!p->dir()) return p->fft->exec(direction == -1, g, p->celem(), p->stream);
  return for_slices(*p, g, p->Q, g, p->Q, [&](const void* gs, void*) -> int {
    if (int rc = grid_in(*p, gs)) return rc;
    if (int rc = p->fft->exec(direction == -1, p->d_grid.p, p->celem(), p->stream)) return rc;
    return grid_out(*p, (void*)gs);
  });
}

int nfftm_sizes(nfftm_plan p, int64_t* Ntilde, int64_t* M, double* sigma) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  if (Ntilde && p->dir()) {
    // Ntilde over all axes: oversampled along dims only (initParams,
    // src/precomputation.jl:25-27)
    for (int d = 0; d < p->Dfull; ++d)
      Ntilde[d] = d >= p->dlo && d < p->dlo + p->D ? p->Nt[d - p->dlo] : p->Nfull[d];
  } else if (Ntilde) {
    for (int d = 0; d < p->D; ++d) Ntilde[d] = p->Nt[d];
  }
  if (M) *M = p->M;
  if (sigma) *sigma = p->sigma;
  return NFFTM_OK;
}

int nfftm_grid(nfftm_plan p, void** g) {
  g_err.clear();
  if (!p || !g) return fail(NFFTM_ERR_ARG, "NULL argument");
  *g = p->d_grid.p;
  return NFFTM_OK;
}

int nfftm_timing(nfftm_plan p, double t[7]) {
  g_err.clear();
  if (!p || !t) return fail(NFFTM_ERR_ARG, "NULL argument");
  for (int i = 0; i < 7; ++i) t[i] = p->timing[i];
  return NFFTM_OK;
}

int nfftm_set_timing(nfftm_plan p, int on) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  if (on) p->flags |= NFFTM_FLAG_TIMING;
  else p->flags &= ~NFFTM_FLAG_TIMING;
  return NFFTM_OK;
}

int nfftm_profile(nfftm_plan p, int on) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  p->profiling = on != 0;
  return NFFTM_OK;
}

int nfftm_profile_read(nfftm_plan p, double ms[8], int64_t count[8]) {
  g_err.clear();
  if (!p || !ms || !count) return fail(NFFTM_ERR_ARG, "NULL argument");
  if (p->prof_drain() != NFFTM_OK) return fail(NFFTM_ERR_HIP, "profile event read failed");
  for (int i = 0; i < 8; ++i) {
    ms[i] = p->prof_ms[i];
    count[i] = p->prof_cnt[i];
    p->prof_ms[i] = 0;
    p->prof_cnt[i] = 0;
  }
  return NFFTM_OK;
}

int nfftm_set_stream(nfftm_plan p, void* s) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  if (int rc = order_streams(p->stream, (hipStream_t)s, &p->switch_ev)) return rc;
  p->stream = (hipStream_t)s;
  return NFFTM_OK;
}

int nfftm_plan_copy(nfftm_plan src, nfftm_plan* dst) {
  g_err.clear();
  if (!src || !dst) return fail(NFFTM_ERR_ARG, "NULL argument");
  *dst = nullptr;
  std::unique_ptr<nfftm_plan_s> p(new nfftm_plan_s());
  p->prec = src->prec;
  p->D = src->D;
  p->m = src->m;
  p->window = src->window;
  p->precompute = src->precompute;
  p->flags = src->flags;
  p->sigma_in = src->sigma_in;
  p->M = src->M;
  p->stream = src->stream;
  for (int d = 0; d < src->D; ++d) p->N[d] = src->N[d];
  p->Dfull = src->Dfull;
  p->dlo = src->dlo;
  p->Bsl = src->Bsl;
  p->Asl = src->Asl;
  for (int d = 0; d < MAXD; ++d) p->Nfull[d] = src->Nfull[d];
  if (int rc = create_impl(*p, src->k.data())) return rc;
  std::memcpy(p->timing, src->timing, sizeof p->timing);
  *dst = p.release();
  return NFFTM_OK;
}

void nfftm_plan_destroy(nfftm_plan p) {
  if (!p) return;
  (void)hipStreamSynchronize(p->stream);
  delete p;
}

int nfftm_binning(nfftm_plan p, int* nitems, int32_t* perm_host, int32_t* items_host) {
  g_err.clear();
  if (!p || !nitems) return fail(NFFTM_ERR_ARG, "NULL argument");
  *nitems = p->nitems;
  HIPCHK(hipStreamSynchronize(p->stream));
  if (perm_host && p->M)
    HIPCHK(hipMemcpy(perm_host, p->d_perm.p, (size_t)p->M * sizeof(int), hipMemcpyDeviceToHost));
  if (items_host && p->nitems)
    HIPCHK(hipMemcpy(items_host, p->d_items.p, (size_t)p->nitems * sizeof(int4), hipMemcpyDeviceToHost));
  return NFFTM_OK;
}

int nfftm_plan_tune(nfftm_plan p) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  return tune_plan(*p, 1024, nullptr);
}

int nfftm_plan_set_tiles(nfftm_plan p, int tile, int threads) {
  g_err.clear();
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  if (!p->tiled) return fail(NFFTM_ERR_UNSUPPORTED, "the plan uses the node-parallel kernels");
  const bool ok_nt = p->D == 3 ? (threads == 512 || threads == 1024) : (threads == 256 || threads == 512 || threads == 1024);
  if (!ok_nt) return fail(NFFTM_ERR_UNSUPPORTED, "threads = %d not available for D = %d", threads, p->D);
  HIPCHK(hipStreamSynchronize(p->stream));
  const bool ok = p->D == 3 ? (tile == 0 || tile == 1) && set_tile3(*p, tile) : tile > 0 && set_tile_shape(*p, tile);
  if (!ok || !tiles_ok(*p)) {
    choose_tiles(*p);
    if (int rc = set_nodes(*p, p->k.data(), p->M)) return rc;
    return fail(NFFTM_ERR_UNSUPPORTED, "tile %d does not fit this plan", tile);
  }
  p->NT = threads;
  return set_nodes(*p, p->k.data(), p->M);
}

int nfftm_tile_shape(nfftm_plan p, int* tile, int* threads) {
  g_err.clear();
  if (!p || !tile || !threads) return fail(NFFTM_ERR_ARG, "NULL argument");
  for (int d = 0; d < 3; ++d) tile[d] = p->tiled && d < p->D ? p->T[d] : 0;
  *threads = p->tiled ? p->NT : 0;
  return NFFTM_OK;
}

int nfftm_batch_create(nfftm_batch* out, int nplans, const nfftm_plan* plans, const int* ntrans, int flags) {
  g_err.clear();
  if (!out) return fail(NFFTM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (nplans < 1 || !plans) return fail(NFFTM_ERR_ARG, "a batch needs at least one plan");
  for (int i = 0; i < nplans; ++i)
    if (!plans[i]) return fail(NFFTM_ERR_ARG, "plan %d is NULL", i);
  std::unique_ptr<nfftm_batch_s> b(new nfftm_batch_s());
  if (int rc = batch_build(*b, plans, ntrans, nplans, flags)) return rc;
  *out = b.release();
  return NFFTM_OK;
}

int nfftm_batch_info(nfftm_batch b, int* nslots, int* graph) {
  g_err.clear();
  if (!b) return fail(NFFTM_ERR_ARG, "batch is NULL");
  if (nslots) *nslots = b->S;
  if (graph) *graph = b->use_graph ? 1 : 0;
  return NFFTM_OK;
}

int nfftm_batch_forward(nfftm_batch b, const void* const* f_dev, void* const* fhat_dev, void* hip_stream) {
  g_err.clear();
  if (!b || !f_dev || !fhat_dev) return fail(NFFTM_ERR_ARG, "NULL argument");
  return batch_run(*b, true, f_dev, fhat_dev, (hipStream_t)hip_stream);
}

int nfftm_batch_adjoint(nfftm_batch b, const void* const* fhat_dev, void* const* f_dev, void* hip_stream) {
  g_err.clear();
  if (!b || !f_dev || !fhat_dev) return fail(NFFTM_ERR_ARG, "NULL argument");
  return batch_run(*b, false, fhat_dev, f_dev, (hipStream_t)hip_stream);
}

void nfftm_batch_destroy(nfftm_batch b) {
  if (!b) return;
  if (b->tab_done) (void)hipEventSynchronize(b->tab_done);
  (void)hipDeviceSynchronize();
  delete b;
}

// ---------------------------------------------------------------------------
// NFFTTools: sampling density compensation and the Toeplitz embedding.
// ---------------------------------------------------------------------------
}  // extern "C"

// The no-allocation workspace (nfftm_sdc_noalloc): a static device array,
// one copy per GPU, whose address on each device is taken once with
// hipGetSymbolAddress (that device current) and kept; host code hands only
// that address to kernels and copies.
constexpr size_t NOALLOC_BYTES = (size_t)64 << 20;
static __device__ __align__(256) char scratch[NOALLOC_BYTES];

namespace {

// Work arrays of sdc!: the weights as complex values (wc) and their
// interpolation (wt), M each; scalars sc[0] (first grid maximum), sc[1]
// (sum Re v), sc[2] (sum |v|^2) then 2048 partials; the error flag; and for
// the final scale an image (u, P) and a node vector (f, M).
struct SdcWork {
  void *wc, *wt, *u, *f;
  double* sc;
  int* bad;
};

int sdc_impl(nfftm_plan_s* p, int iters, void* weights, int scale, const SdcWork& w) {
  const bool dbl = p->dbl();
  const i64 M = p->M;
  hipStream_t s = p->stream;
  double* part = w.sc + 8;
  HIPCHK(launch_fill0(w.bad, sizeof(int), s));
  HIPCHK(launch_real_to_cpx(dbl, weights, w.wc, M, s));
  // sdc!(p, iters, weights, weights_tmp, workg), samplingDensity.jl:93-118
  for (int i = 0; i < iters; ++i) {
    if (int rc = run_conv(*p, true, w.wc, p->d_grid.p, true)) return rc;
    if (i == 0) HIPCHK(launch_reduce(dbl, RED_MAX_RE, p->d_grid.p, p->Q, part, w.sc, s));
    HIPCHK(launch_div_scalar(dbl, p->d_grid.p, p->Q, w.sc, s));
    if (int rc = run_conv(*p, false, p->d_grid.p, w.wt, true)) return rc;
    HIPCHK(launch_sdc_update(dbl, w.wc, w.wt, M, w.sc, w.bad, s));
  }
  if (scale) {
    // c = real(sum(v)) / sum(abs2, v), v = A' D(w) A 1 (samplingDensity.jl:141-154)
    HIPCHK(launch_fill_one(dbl, w.u, p->P, s));
    if (int rc = enqueue_transform(*p, true, w.u, w.f)) return rc;
    HIPCHK(launch_mul_real(dbl, w.f, w.wc, M, s));
    if (int rc = enqueue_transform(*p, false, w.f, w.u)) return rc;
    HIPCHK(launch_reduce(dbl, RED_SUM_RE, w.u, p->P, part, w.sc + 1, s));
    HIPCHK(launch_reduce(dbl, RED_SUM_ABS2, w.u, p->P, part, w.sc + 2, s));
    HIPCHK(launch_cpx_to_real(dbl, w.wc, weights, M, w.sc + 1, s));
  } else {
    HIPCHK(launch_cpx_to_real(dbl, w.wc, weights, M, nullptr, s));
  }
  int hbad = 0;
  HIPCHK(hipMemcpyAsync(&hbad, w.bad, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (hbad) return fail(NFFTM_ERR_UNSUPPORTED, "non-positive weights");
  return NFFTM_OK;
}

int sdc_check(nfftm_plan_s* p, int iters, void* weights) {
  if (!p) return fail(NFFTM_ERR_ARG, "plan is NULL");
  if (iters < 0) return fail(NFFTM_ERR_ARG, "iters = %d < 0", iters);
  if (!weights && p->M) return fail(NFFTM_ERR_ARG, "NULL weights");
  if (p->dir()) return fail(NFFTM_ERR_UNSUPPORTED, "sampling density of a directional plan");
  return NFFTM_OK;
}

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

int scratch_address(void** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(NFFTM_ERR_UNSUPPORTED, "device %d beyond %d", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(NFFTM_ERR_HIP, "no device address for the workspace");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return NFFTM_OK;
}

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int nfftm_sdc(nfftm_plan p, int iters, void* weights, int scale) {
  g_err.clear();
  if (int rc = sdc_check(p, iters, weights)) return rc;
  if (p->M == 0) return NFFTM_OK;
  const size_t ce = p->celem();
  DevBuf<char> wc, wt, u, f;
  DevBuf<double> sc;
  DevBuf<int> bad;
  HIPCHK(wc.alloc((size_t)p->M * ce));
  HIPCHK(wt.alloc((size_t)p->M * ce));
  HIPCHK(sc.alloc(8 + 2048));
  HIPCHK(bad.alloc(1));
  if (scale) {
    HIPCHK(u.alloc((size_t)p->P * ce));
    HIPCHK(f.alloc((size_t)p->M * ce));
  }
  return sdc_impl(p, iters, weights, scale, SdcWork{wc.p, wt.p, u.p, f.p, sc.p, bad.p});
}

int nfftm_sdc_noalloc(nfftm_plan p, int iters, void* weights, int scale) {
  g_err.clear();
  if (int rc = sdc_check(p, iters, weights)) return rc;
  if (p->M == 0) return NFFTM_OK;
  const size_t ce = p->celem(), vM = align256((size_t)p->M * ce);
  const size_t o_bad = align256((8 + 2048) * sizeof(double)), o_wc = o_bad + 256, o_wt = o_wc + vM,
               o_u = o_wt + vM, o_f = o_u + (scale ? align256((size_t)p->P * ce) : 0),
               need = o_f + (scale ? vM : 0);
  if (need > NOALLOC_BYTES)
    return fail(NFFTM_ERR_ARG, "sdc without allocation needs %zu bytes of workspace (M = %lld, %s), more than its %zu",
                need, (long long)p->M, scale ? "with the scale" : "no scale", NOALLOC_BYTES);
  void* ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  char* b = (char*)ws;
  return sdc_impl(p, iters, weights, scale,
                  SdcWork{b + o_wc, b + o_wt, scale ? b + o_u : nullptr, scale ? b + o_f : nullptr, (double*)b,
                          (int*)(b + o_bad)});
}

}  // extern "C"

struct nfftm_toeplitz_s {
  int prec = NFFTM_F32, D = 1;
  i64 N[3] = {1, 1, 1}, L[3] = {1, 1, 1}, P = 1, Q = 1;
  hipStream_t stream = nullptr;
  std::unique_ptr<FFTStage> fft;
  DevBuf<char> x;
  bool dbl() const { return prec == NFFTM_F64; }
  size_t celem() const { return dbl() ? 16 : 8; }
  ToepArgs targs() const {
    ToepArgs a;
    std::memset(&a, 0, sizeof a);
    a.D = D;
    for (int d = 0; d < 3; ++d) {
      a.N[d] = N[d];
      a.L[d] = L[d];
    }
    a.P = P;
    a.Q = Q;
    return a;
  }
};

extern "C" {

int nfftm_toeplitz_create(nfftm_toeplitz* out, int prec, int D, const int64_t* N, void* hip_stream) {
  g_err.clear();
  if (!out || !N) return fail(NFFTM_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (D < 1 || D > 3) return fail(NFFTM_ERR_UNSUPPORTED, "Toeplitz embedding for D = %d (1..3 supported)", D);
  if (prec != NFFTM_F32 && prec != NFFTM_F64) return fail(NFFTM_ERR_ARG, "precision %d", prec);
  std::unique_ptr<nfftm_toeplitz_s> t(new nfftm_toeplitz_s());
  t->prec = prec;
  t->D = D;
  t->stream = (hipStream_t)hip_stream;
  for (int d = 0; d < D; ++d) {
    if (N[d] < 1) return fail(NFFTM_ERR_ARG, "N[%d] = %lld", d, (long long)N[d]);
    t->N[d] = N[d];
    t->L[d] = 2 * N[d];
    t->P *= N[d];
    t->Q *= 2 * N[d];
  }
  HIPCHK(t->x.alloc((size_t)t->Q * t->celem()));
  t->fft.reset(new FFTStage());
  if (int rc = t->fft->build(D, t->L, t->dbl())) return rc;
  *out = t.release();
  return NFFTM_OK;
}

int nfftm_toeplitz_kernel(nfftm_toeplitz t, nfftm_plan p, void* lambda) {
  g_err.clear();
  if (!t || !p || !lambda) return fail(NFFTM_ERR_ARG, "NULL argument");
  if (p->dir()) return fail(NFFTM_ERR_UNSUPPORTED, "Toeplitz kernel of a directional plan");
  if (p->D != t->D || p->prec != t->prec) return fail(NFFTM_ERR_DIM, "plan and Toeplitz object differ in D or precision");
  for (int d = 0; d < t->D; ++d)
    if (p->N[d] != t->L[d])
      return fail(NFFTM_ERR_DIM, "plan size %lld along dimension %d is not 2N = %lld", (long long)p->N[d], d + 1,
                  (long long)t->L[d]);
  const bool dbl = t->dbl();
  hipStream_t s = p->stream;
  DevBuf<char> ones;
  HIPCHK(ones.alloc((size_t)std::max<i64>(1, p->M) * t->celem()));
  HIPCHK(launch_fill_one(dbl, ones.p, p->M, s));
  if (int rc = enqueue_transform(*p, false, ones.p, lambda)) return rc;  // eigMat = adjoint(p) * ones
  if (int rc = t->fft->exec(true, lambda, t->celem(), s)) return rc;
  HIPCHK(launch_shift_sign(dbl, lambda, t->targs(), s));  // the fftshift before the FFT
  HIPCHK(hipStreamSynchronize(s));
  return NFFTM_OK;
}

int nfftm_toeplitz_apply(nfftm_toeplitz t, const void* lambda, void* y) {
  g_err.clear();
  if (!t || !lambda || !y) return fail(NFFTM_ERR_ARG, "NULL argument");
  const bool dbl = t->dbl();
  const ToepArgs a = t->targs();
  hipStream_t s = t->stream;
  HIPCHK(launch_pad_corner(dbl, y, t->x.p, a, s));
  if (int rc = t->fft->exec(true, t->x.p, t->celem(), s)) return rc;
  HIPCHK(launch_cmul(dbl, t->x.p, lambda, t->Q, s));
  if (int rc = t->fft->exec(false, t->x.p, t->celem(), s)) return rc;
  HIPCHK(launch_crop_corner(dbl, t->x.p, y, a, 1.0 / (double)t->Q, s));
  return NFFTM_OK;
}

void nfftm_toeplitz_destroy(nfftm_toeplitz t) {
  if (!t) return;
  (void)hipStreamSynchronize(t->stream);
  delete t;
}

}  // extern "C"
