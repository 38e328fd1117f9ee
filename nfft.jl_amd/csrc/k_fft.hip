// k_fft.hip -- the FFT stage of a transform on power-of-two grids as pruned
// line passes, with the deconvolution fused into the dimension-1 pass.
//
// The forward transform only fills the image corners of the oversampled
// grid (N_d of the Nt_d positions per dimension, src/deconvolution.jl:22-45)
// and the adjoint only keeps those positions of the inverse FFT
// (src/deconvolution.jl:69-92).  A D-dimensional FFT is D passes of 1-D
// FFTs along the lines of one dimension, so each pass can skip the lines
// that are zero (forward) or discarded (adjoint):
//   forward  X: lines (y, z) in S_y x S_z, input read from f (deconvolution
//               fused), zero elsewhere;
//            Y: lines (x, z), z in S_z, input nonzero on S_y only;
//            Z: all lines (x, y), input nonzero on S_z only;
//   adjoint  Z: all lines, output kept on S_z;  Y: lines (x, z), z in S_z,
//               output kept on S_y;  X: lines (y, z) in S_y x S_z, output
//               kept on S_x and written to f with the transposed
//               deconvolution fused.
// S_d is the set of grid positions of the image indices: image index i < Na
// goes to i - Na + Nt, the rest to i - Na.  Values equal the full FFT of the
// zero-padded grid (FFTW's forward / unnormalised backward transform,
// src/implementation.jl:87-88) up to rounding.
//
// The line FFTs are those of fft_line.hpp.  Strided passes (Y, Z) take
// E / L lines that are consecutive along x, so that every row of the load
// and store is a contiguous run of E / L elements.
#include "fft_line.hpp"
#include "launch.hpp"

namespace nfftm {

namespace {

using namespace lf;

// Deconvolution factor product of image line q = iz N1 + iy, as
// deconvolve! forms it (oscale times the outer factors).
__device__ __forceinline__ double line_factor(const FArgs& a, int iy, int iz, bool with_oscale) {
  double f = with_oscale ? a.oscale : 1.0;
  if (a.D >= 2) f *= a.fac[1][iy];
  if (a.D >= 3) f *= a.fac[2][iz];
  return f;
}

// Forward, dimension 1: deconvolution of f into lines (y, z) in S_y x S_z,
// zero padding, FFT, whole lines to the grid.
template <typename R, int L>
__global__ __launch_bounds__((LineGeo<R, L>::NT)) void k_fft_x_fwd(FArgs a, const typename cpx<R>::T* __restrict__ f,
                                                   typename cpx<R>::T* __restrict__ grid) {
  using C = typename cpx<R>::T;
  using G = LineGeo<R, L>;
  __shared__ __align__(16) R sm[2 * (G::LPB * G::LS + L)];
  C* buf = (C*)sm;
  C* tw = buf + G::LPB * G::LS;
  load_tw<G::NT>(tw, a.tw[0], L);
  const int N0 = a.N[0], Na0 = a.Na[0], N1 = a.N[1];
  const int nlines = a.N[1] * a.N[2];
  const int l = threadIdx.x / G::TL, j0 = threadIdx.x % G::TL;
  const int q = blockIdx.x * G::LPB + l;
  C* line = buf + l * G::LS;
  C x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = C{0, 0};
  if (q < nlines) {
    const int iy = q % N1, iz = q / N1;
    const double fr = line_factor(a, iy, iz, true);
    const C* src = f + (size_t)q * N0 * a.ist;
    double fc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int ix = image_idx(j0 + k * G::TL, N0, L, Na0);
      fc[k] = 0;
      if (ix >= 0) {
        x[k] = src[(size_t)ix * a.ist];
        fc[k] = a.fac[0][ix];
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const double s = fc[k] * fr;
      x[k] = C{(R)((double)x[k].x * s), (R)((double)x[k].y * s)};
    }
  }
  line_fft<R, L, false>(x, line, j0, line, j0, tw);
  if (q < nlines) {
    const int iy = q % N1, iz = q / N1;
    const size_t o = (size_t)grid_pos(iz, a.N[2], a.L[2], a.Na[2]) * a.gstr[2] +
                     (size_t)grid_pos(iy, N1, a.L[1], a.Na[1]) * a.gstr[1];
#pragma unroll
    for (int k = 0; k < 16; ++k) grid[o + j0 + k * G::TL] = x[k];
  }
}

// Adjoint, dimension 1: lines (y, z) in S_y x S_z of the grid, inverse FFT,
// positions S_x to f with the transposed deconvolution.
template <typename R, int L>
__global__ __launch_bounds__((LineGeo<R, L>::NT)) void k_fft_x_adj(FArgs a, const typename cpx<R>::T* __restrict__ grid,
                                                   typename cpx<R>::T* __restrict__ f) {
  using C = typename cpx<R>::T;
  using G = LineGeo<R, L>;
  __shared__ __align__(16) R sm[2 * (G::LPB * G::LS + L)];
  C* buf = (C*)sm;
  C* tw = buf + G::LPB * G::LS;
  load_tw<G::NT>(tw, a.tw[0], L);
  const int N0 = a.N[0], Na0 = a.Na[0], N1 = a.N[1];
  const int nlines = a.N[1] * a.N[2];
  const int l = threadIdx.x / G::TL, j0 = threadIdx.x % G::TL;
  const int q = blockIdx.x * G::LPB + l;
  C* line = buf + l * G::LS;
  C x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = C{0, 0};
  if (q < nlines) {
    const int iy = q % N1, iz = q / N1;
    const size_t o = (size_t)grid_pos(iz, a.N[2], a.L[2], a.Na[2]) * a.gstr[2] +
                     (size_t)grid_pos(iy, N1, a.L[1], a.Na[1]) * a.gstr[1];
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = grid[o + j0 + k * G::TL];
  }
  line_fft<R, L, true>(x, line, j0, line, j0, tw);
  if (q < nlines) {
    const int iy = q % N1, iz = q / N1;
    const double fr = line_factor(a, iy, iz, false) * a.oscale;
    C* dst = f + (size_t)q * N0 * a.ist;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int ix = image_idx(j0 + k * G::TL, N0, L, Na0);
      if (ix >= 0) {
        const double s = a.fac[0][ix] * fr;
        dst[(size_t)ix * a.ist] = C{(R)((double)x[k].x * s), (R)((double)x[k].y * s)};
      }
    }
  }
}

// Strided pass along dimension d (1 or 2) over lines (x, c): x in
// [0, Nt_0) in chunks of LPB, c the position along the remaining dimension
// o (all of [0, Nt_o) if call, else S_o).  MODE 1: input nonzero on S_d
// only (forward); MODE 2: output kept on S_d only (adjoint).  Thread t
// takes line x0 + t mod LPB and points t / LPB + u L / 16, so that a
// wavefront's loads and stores are rows of LPB consecutive x.
template <typename R, int L, bool INV, int MODE>
__global__ __launch_bounds__((LineGeo<R, L>::NT)) void k_fft_col(FArgs a, typename cpx<R>::T* __restrict__ grid, int d, int call) {
  using C = typename cpx<R>::T;
  using G = LineGeo<R, L>;
  __shared__ __align__(16) R sm[2 * (G::LPB * G::LS + L)];
  C* buf = (C*)sm;
  C* tw = buf + G::LPB * G::LS;
  load_tw<G::NT>(tw, a.tw[d], L);
  const int o = 3 - d;  // d = 1: o = 2 (Nt = 1 for D = 2); d = 2: o = 1
  const int c = blockIdx.y;
  const int gc = call ? c : grid_pos(c, a.N[o], a.L[o], a.Na[o]);
  const int x0 = blockIdx.x * G::LPB;
  const size_t base = (size_t)gc * a.gstr[o] + x0;
  const long long st = a.gstr[d];
  const int Nd = a.N[d], Nad = a.Na[d];
  const int xl = threadIdx.x % G::LPB, j0 = threadIdx.x / G::LPB;
  const bool xin = x0 + xl < a.L[0];
  C v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int k = j0 + u * G::TL;
    v[u] = C{0, 0};
    if (xin && (MODE == 2 || image_idx(k, Nd, L, Nad) >= 0)) v[u] = grid[base + xl + (size_t)k * st];
  }
  line_fft<R, L, INV>(v, buf + xl * G::LS, j0, buf + (threadIdx.x / G::TL) * G::LS, threadIdx.x % G::TL, tw);
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int k = j0 + u * G::TL;
    if (xin && (MODE == 1 || image_idx(k, Nd, L, Nad) >= 0)) grid[base + xl + (size_t)k * st] = v[u];
  }
}

template <typename R, int L>
hipError_t x_pass(bool fwd, const FArgs& a, hipStream_t s, const void* in, void* out) {
  using C = typename cpx<R>::T;
  const int nlines = a.N[1] * a.N[2];
  using G = LineGeo<R, L>;
  const dim3 g((unsigned)((nlines + G::LPB - 1) / G::LPB)), b(G::NT);
  if (fwd) hipLaunchKernelGGL((k_fft_x_fwd<R, L>), g, b, 0, s, a, (const C*)in, (C*)out);
  else hipLaunchKernelGGL((k_fft_x_adj<R, L>), g, b, 0, s, a, (const C*)in, (C*)out);
  return hipGetLastError();
}

template <typename R, int L>
hipError_t col_pass(bool fwd, const FArgs& a, hipStream_t s, void* grid, int d) {
  using C = typename cpx<R>::T;
  const int o = 3 - d;
  // lines along d: (x, z in S_z) for d = 1, (x, y) for d = 2
  const int call = d == 2 ? 1 : 0;
  const int nc = call ? a.L[o] : a.N[o];
  using G = LineGeo<R, L>;
  const dim3 g((unsigned)((a.L[0] + G::LPB - 1) / G::LPB), (unsigned)nc), b(G::NT);
  if (fwd) hipLaunchKernelGGL((k_fft_col<R, L, false, 1>), g, b, 0, s, a, (C*)grid, d, call);
  else hipLaunchKernelGGL((k_fft_col<R, L, true, 2>), g, b, 0, s, a, (C*)grid, d, call);
  return hipGetLastError();
}

template <typename R>
hipError_t x_pass_l(bool fwd, const FArgs& a, hipStream_t s, const void* in, void* out) {
  switch (a.L[0]) {
#define NFFTM_XL(LL) \
  case LL:           \
    return x_pass<R, LL>(fwd, a, s, in, out);
    NFFTM_XL(16) NFFTM_XL(32) NFFTM_XL(64) NFFTM_XL(128) NFFTM_XL(256) NFFTM_XL(512) NFFTM_XL(1024)
#undef NFFTM_XL
    case 2048:
      if constexpr (sizeof(R) == 4) return x_pass<R, 2048>(fwd, a, s, in, out);
      return hipErrorInvalidValue;
    default:
      return hipErrorInvalidValue;
  }
}

template <typename R>
hipError_t col_pass_l(bool fwd, const FArgs& a, hipStream_t s, void* grid, int d) {
  switch (a.L[d]) {
#define NFFTM_CL(LL) \
  case LL:           \
    return col_pass<R, LL>(fwd, a, s, grid, d);
    NFFTM_CL(16) NFFTM_CL(32) NFFTM_CL(64) NFFTM_CL(128) NFFTM_CL(256) NFFTM_CL(512) NFFTM_CL(1024)
#undef NFFTM_CL
    case 2048:
      if constexpr (sizeof(R) == 4) return col_pass<R, 2048>(fwd, a, s, grid, d);
      return hipErrorInvalidValue;
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace

bool pruned_fft_supported(int D, const int64_t* N, const int64_t* Nt, bool dbl) {
  if (D < 1 || D > 3) return false;
  long long q = 1;
  for (int d = 0; d < D; ++d) {
    const int64_t n = Nt[d];
    if (n < 16 || n > (dbl ? 1024 : 2048) || (n & (n - 1)) || N[d] > n) return false;
    q *= n;
  }
  return q < (1LL << 31);
}

// Forward: part 0 = dimension-1 pass from f (deconvolution fused), part 1 =
// the passes along dimensions 2 and 3.
hipError_t launch_pruned_fft_fwd(bool dbl, const FArgs& a, hipStream_t s, const void* f, void* grid, int part) {
  if (part == 0) return dbl ? x_pass_l<double>(true, a, s, f, grid) : x_pass_l<float>(true, a, s, f, grid);
  for (int d = 1; d < a.D; ++d) {
    const hipError_t e = dbl ? col_pass_l<double>(true, a, s, grid, d) : col_pass_l<float>(true, a, s, grid, d);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Adjoint: part 0 = the passes along dimensions 3 and 2, part 1 = the
// dimension-1 pass into f (transposed deconvolution fused), part 2 = the
// passes of part 0 after the first (which launch_fold_fft3 made).
hipError_t launch_pruned_fft_adj(bool dbl, const FArgs& a, hipStream_t s, void* grid, void* f, int part) {
  if (part == 1) return dbl ? x_pass_l<double>(false, a, s, grid, f) : x_pass_l<float>(false, a, s, grid, f);
  for (int d = part == 2 ? a.D - 2 : a.D - 1; d >= 1; --d) {
    const hipError_t e = dbl ? col_pass_l<double>(false, a, s, grid, d) : col_pass_l<float>(false, a, s, grid, d);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace nfftm
