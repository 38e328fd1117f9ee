// fft_line.hpp -- in-LDS FFTs of power-of-two lines for the pruned FFT
// passes (k_fft.hip) and the fold fused into the first adjoint pass
// (k_tile3.hip).
//
// One workgroup transforms E / L lines of length L = 2^k (FP32: E = 4096
// points, 256 threads, 16 <= L <= 2048; FP64: E = 2048, 128 threads,
// 16 <= L <= 1024): a first radix-2^(k mod 4) pass (radix 16 when k is a
// multiple of 4), then radix-16 Stockham passes, L / 16 threads per line,
// one radix-16 butterfly (or 16 / R radix-R ones) per thread and pass,
// twiddles from a table of exp(-2 pi i j / L) in LDS.  A thread loads the 16
// points of its first-pass butterflies and stores the 16 outputs of its
// last-pass butterfly, so the first and last passes run on registers and
// only the passes between them exchange through LDS (for L = 256: one LDS
// write, one barrier, one read).
#pragma once
#include "kernels.hpp"

namespace nfftm {
namespace lf {

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
  return C{a.x + b.x, a.y + b.y};
}
template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
  return C{a.x - b.x, a.y - b.y};
}

// x * conj(t) (forward, t = (cos, sin) of +2 pi j / L) or x * t (inverse).
template <bool INV, typename C>
__device__ __forceinline__ C tmul(C x, C t) {
  if (INV) return C{x.x * t.x - x.y * t.y, x.y * t.x + x.x * t.y};
  return C{x.x * t.x + x.y * t.y, x.y * t.x - x.x * t.y};
}

// cos and sin of 2 pi k / 16, k < 8
__device__ constexpr double kC16[8] = {1.0,
                                       0.92387953251128675613,
                                       0.70710678118654752440,
                                       0.38268343236508977173,
                                       0.0,
                                       -0.38268343236508977173,
                                       -0.70710678118654752440,
                                       -0.92387953251128675613};
__device__ constexpr double kS16[8] = {0.0,
                                       0.38268343236508977173,
                                       0.70710678118654752440,
                                       0.92387953251128675613,
                                       1.0,
                                       0.92387953251128675613,
                                       0.70710678118654752440,
                                       0.38268343236508977173};

// x * w_N^k, w_N = exp(-+2 pi i / N), N <= 16, k < N / 2
template <typename R, int N, int K, bool INV, typename C>
__device__ __forceinline__ C wmul(C x) {
  constexpr int q = K * (16 / N);
  if constexpr (q == 0) {
    return x;
  } else if constexpr (q == 4) {  // -i (forward) or +i (inverse)
    return INV ? C{-x.y, x.x} : C{x.y, -x.x};
  } else {
    const R c = (R)kC16[q], s = (R)kS16[q];
    return tmul<INV>(x, C{c, s});
  }
}

template <typename R, int N, int K, bool INV, typename C>
__device__ __forceinline__ void dft_combine(C* v, const C* e, const C* o) {
  if constexpr (K < N / 2) {
    const C t = wmul<R, N, K, INV>(o[K]);
    v[K] = cadd(e[K], t);
    v[K + N / 2] = csub(e[K], t);
    dft_combine<R, N, K + 1, INV>(v, e, o);
  }
}

// In-register DFT of N <= 16 points in natural order (radix-2 recursion).
template <typename R, int N, bool INV, typename C>
__device__ __forceinline__ void dft(C* v) {
  if constexpr (N > 1) {
    C e[N / 2], o[N / 2];
#pragma unroll
    for (int k = 0; k < N / 2; ++k) {
      e[k] = v[2 * k];
      o[k] = v[2 * k + 1];
    }
    dft<R, N / 2, INV>(e);
    dft<R, N / 2, INV>(o);
    dft_combine<R, N, 0, INV>(v, e, o);
  }
}

// One Stockham pass of radix RAD over a line of L points in LDS, stage NS
// (the product of the radices of the previous passes): butterfly j reads
// line[j + r L / RAD], applies w_{NS RAD}^{(j mod NS) r}, transforms, and
// writes line[(j / NS) NS RAD + j mod NS + r NS].  Thread j0 of the line
// does butterflies j0 + b L / 16, b < 16 / RAD.
template <typename R, int L, int RAD, int NS, bool INV, typename C>
__device__ __forceinline__ void stockham(C* line, int j0, const C* tw) {
  constexpr int TL = L / 16, B = 16 / RAD;
  C v[B][RAD];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = j0 + b * TL;
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[b][r] = line[j + r * (L / RAD)];
    if constexpr (NS > 1) {
      const int jm = j % NS;
#pragma unroll
      for (int r = 1; r < RAD; ++r) v[b][r] = tmul<INV>(v[b][r], tw[jm * r * (L / (NS * RAD))]);
    }
    dft<R, RAD, INV>(v[b]);
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = j0 + b * TL;
    const int o = (j / NS) * NS * RAD + (j % NS);
#pragma unroll
    for (int r = 0; r < RAD; ++r) line[o + r * NS] = v[b][r];
  }
  __syncthreads();
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// Radix of the first pass: 2^(log2 L mod 4), or 16 when L is a power of 16.
template <int L>
constexpr int first_radix() {
  return (ilog2(L) % 4) ? (1 << (ilog2(L) % 4)) : 16;
}

// Radix-16 Stockham passes of stages NS, 16 NS, ... below L / 16, in LDS.
template <typename R, int L, int NS, bool INV, typename C>
__device__ __forceinline__ void mid_passes(C* line, int j0, const C* tw) {
  if constexpr (NS < L / 16) {
    stockham<R, L, 16, NS, INV>(line, j0, tw);
    mid_passes<R, L, NS * 16, INV>(line, j0, tw);
  }
}

// FFT of a line of L points.  On entry the thread holds x[u] = point
// j0 + u L / 16 of its line, on exit output point j0 + u L / 16: these are
// exactly the inputs of the first pass's butterflies j0 + b L / 16 (stage 1,
// radix R0, b < 16 / R0) and the outputs of the last pass's butterfly j0
// (stage L / 16, radix 16), so both passes run on registers and only the
// passes between them go through the LDS line.  All threads of the
// workgroup call this together; which line and j0 a thread takes may differ
// from the (line, j0) = (t / (L / 16), t mod (L / 16)) of the middle passes.
template <typename R, int L, bool INV, typename C>
__device__ __forceinline__ void line_fft(C (&x)[16], C* line, int j0, C* mline, int mj0, const C* tw) {
  constexpr int TL = L / 16;
  if constexpr (L == 16) {
    dft<R, 16, INV>(x);
  } else {
    constexpr int R0 = first_radix<L>(), B = 16 / R0;
#pragma unroll
    for (int b = 0; b < B; ++b) {
      C v[R0];
#pragma unroll
      for (int r = 0; r < R0; ++r) v[r] = x[b + r * B];  // point j0 + b TL + r L / R0
      dft<R, R0, INV>(v);
      const int o = (j0 + b * TL) * R0;
#pragma unroll
      for (int r = 0; r < R0; ++r) line[o + r] = v[r];
    }
    __syncthreads();
    mid_passes<R, L, R0, INV>(mline, mj0, tw);
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = line[j0 + r * TL];
#pragma unroll
    for (int r = 1; r < 16; ++r) x[r] = tmul<INV>(x[r], tw[j0 * r]);
    dft<R, 16, INV>(x);
  }
}

// Grid position of image index i (one dimension).
__device__ __forceinline__ int grid_pos(int i, int N, int Nt, int Na) { return i < Na ? i - Na + Nt : i - Na; }
// Image index of grid position g, or -1 if g is not an image position.
__device__ __forceinline__ int image_idx(int g, int N, int Nt, int Na) {
  if (g < N - Na) return g + Na;
  if (g >= Nt - Na) return g - (Nt - Na);
  return -1;
}

template <typename R, int L>
struct LineGeo {
  static constexpr int E = sizeof(R) == 4 ? 4096 : 2048;  // points per workgroup
  static constexpr int NT = E / 16;                        // threads per workgroup
  static constexpr int TL = L / 16;                        // threads per line
  static constexpr int LPB = E / L;                        // lines per workgroup
  static constexpr int LS = L + 1;                         // LDS line stride (elements)
  static_assert(L >= 16 && L <= E, "line length");
};

template <int NT, typename C>
__device__ __forceinline__ void load_tw(C* tw, const void* src, int L) {
  for (int i = threadIdx.x; i < L; i += NT) tw[i] = ((const C*)src)[i];
}

}  // namespace lf
}  // namespace nfftm
