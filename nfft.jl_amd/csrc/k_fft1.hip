// k_fft1.hip -- the FFT stage of 1-D transforms on large grids.
//
// The grid of Nt = n1 n2 n3 points (n2 = n3 = Lm; n1 = P M, P = 1 or 5, M a
// power of two) is transformed in three passes of line FFTs over the three
// digits of the Cooley-Tukey index map.  A grid position is read as digits
// (a, b, c): position a + n1 b + n1 n2 c.  With x at natural index
// j = j1 n2 n3 + j2 n3 + j3, X at natural index k = k1 + n1 k2 + n1 n2 k3
// and W = exp(-+2 pi i / Nt):
//   X[k] = sum_j3 W_n3^(j3 k3) W^(j3 (k1 + n1 k2))
//            sum_j2 W_n2^(j2 k2) W^(n3 j2 k1)
//              sum_j1 W_n1^(j1 k1) x[j].
// Forward (image to grid):
//   1. lines (j2, j3): x[j] read at stride n2 n3 (the deconvolution of f
//      fused: zero, and not loaded, outside the image positions), DFT over
//      j1, twiddle W^(n3 j2 k1), written contiguous at digits (k1, j2, j3);
//   2. lines (k1, j3) at stride n1: DFT over j2, twiddle W^(j3 (k1 + n1 k2)),
//      in place at (k1, k2, j3);
//   3. lines (k1, k2) at stride n1 n2: DFT over j3, in place: (k1, k2, k3) is
//      the natural position of X[k].
// Adjoint (grid to image), the transposed sequence (the DFT matrix is
// symmetric): pass 3 followed by the twiddles of pass 2, pass 2 followed by
// those of pass 1, then pass 1 read contiguous and written to the natural
// positions (only the image positions, the transposed deconvolution fused).
// The forward fills the whole grid and the adjoint reads all of it.
//
// Pass 1 reads the deconvolution factors from a copy laid out per workgroup
// (fft1_lines): slab (j2, j3 / B), then j1, then j3 mod B, so that each
// workgroup reads its factors as one contiguous block (read at the image
// positions of the strided side, as the data are, they cost the adjoint's
// last pass 70 of its 150 us on a 2^22 x 5 grid).
//
// The strided passes take LPB lines consecutive in a, so that each load and
// store of a wavefront is a run of LPB consecutive elements; pass 1 takes B
// lines consecutive in j3 for the same reason on its strided side.  Pass 1
// with P = 5 runs five M-point FFTs on the sub-lines j1 = 5 j' + r and
// combines them in radix-5 butterflies:
//   X[k' + q M] = sum_r W_5^(r q) W_n1^(r k') S_r[k'].
// Values equal the full FFT of the zero-padded grid (FFTW's forward /
// unnormalised backward transform, src/implementation.jl:87-88) up to
// rounding.  Twiddles of W come from two FP64 tables of (cos, sin) pairs:
// W^e = thi[e >> sh] tlo[e mod 2^sh].  A thread of a strided pass looks up
// W^(al j0 + be) and W^(al L / 16) once and steps through its 16 points by
// FP64 products (table gathers per point, scattered over the tables, cost
// the passes 25-55 us each on a 2^22 x 5 grid); pass 1 tabulates
// W^(n3 j2 k1), k1 < n1, per workgroup in LDS.
#include <cstring>

#include "fft1.hpp"
#include "fft_line.hpp"

namespace nfftm {

namespace {

using namespace lf;

// DFT of 5 points in natural order.
template <typename R, bool INV, typename C>
__device__ __forceinline__ void dft5(C* v) {
  const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;  // cos 2pi/5, cos 4pi/5
  const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;   // sin 2pi/5, sin 4pi/5
  const C a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  const C a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  const C t1 = C{v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
  const C t2 = C{v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
  const C u1 = C{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
  const C u2 = C{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
  // forward: X1 = t1 - i u1, X4 = t1 + i u1, X2 = t2 - i u2, X3 = t2 + i u2
  const C m1 = INV ? C{-u1.y, u1.x} : C{u1.y, -u1.x};
  const C m2 = INV ? C{-u2.y, u2.x} : C{u2.y, -u2.x};
  v[0] = C{v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y};
  v[1] = cadd(t1, m1);
  v[4] = csub(t1, m1);
  v[2] = cadd(t2, m2);
  v[3] = csub(t2, m2);
}

using D2 = cpx<double>::T;

__device__ __forceinline__ D2 cmul(D2 x, D2 y) { return D2{x.x * y.x - x.y * y.y, x.y * y.x + x.x * y.y}; }

// (cos, sin)(2 pi e / Nt), 0 <= e < Nt, FP64
__device__ __forceinline__ D2 big_tw(const F1Args& a, int e) {
  return cmul(((const D2*)a.thi)[e >> a.sh], ((const D2*)a.tlo)[e & ((1 << a.sh) - 1)]);
}

// Pass 1 geometry (B as fft1_lines): B lines of n1 = P M points per workgroup, P B sub-lines
// of M points (LDS stride M + 1), M / 16 threads per sub-line; LDS also
// holds the M-point and n1-point line twiddles and the forward's W^(n3 j2 k1).
template <typename R, int P, int M>
struct P1Geo {
  static constexpr int L = P * M;
  static constexpr int B = (sizeof(R) == 4 ? 16 : 8) / (L >= 1024 ? 4 : (L >= 512 ? 2 : 1));
  static constexpr int TLm = M / 16;
  static constexpr int NT = B * P * TLm;
  static constexpr int SS = M + 1;
  static constexpr int LDS_C = B * P * SS + M + (P > 1 ? L : 0) + L;  // complex elements
  static_assert(LDS_C * sizeof(typename cpx<R>::T) <= 65536, "pass-1 LDS");
};

template <typename R, int P, int M>
__device__ __forceinline__ void p1_tables(const F1Args& a, typename cpx<R>::T* twm, typename cpx<R>::T* twn) {
  using C = typename cpx<R>::T;
  using G = P1Geo<R, P, M>;
  for (int i = threadIdx.x; i < M; i += G::NT) twm[i] = ((const C*)a.tw1)[i * P];
  if constexpr (P > 1)
    for (int i = threadIdx.x; i < G::L; i += G::NT) twn[i] = ((const C*)a.tw1)[i];
}

// Forward pass 1: lines (j2, j3 = j3_0 + l), l < B.
template <typename R, int P, int M>
__global__ __launch_bounds__((P1Geo<R, P, M>::NT)) void k_fft1_first(F1Args a, const typename cpx<R>::T* __restrict__ f,
                                                                      typename cpx<R>::T* __restrict__ grid) {
  using C = typename cpx<R>::T;
  using G = P1Geo<R, P, M>;
  __shared__ __align__(16) C sm[G::LDS_C];
  C* buf = sm;
  C* twm = sm + G::B * P * G::SS;
  C* twn = twm + M;
  C* tb = twn + (P > 1 ? G::L : 0);
  const int Lm = a.Lm, n23 = a.Lm * a.Lm;
  const int j2 = blockIdx.y, j3_0 = blockIdx.x * G::B;
  const int t = threadIdx.x;
  // loads: line fastest (consecutive grid positions), then sub-line, then j0
  const int l = t % G::B, r = (t / G::B) % P, j0 = t / (G::B * P);
  const int g0 = j2 * Lm + j3_0 + l;
  const double* fb = a.fac1 + (size_t)(j2 * (Lm / G::B) + blockIdx.x) * a.n1 * G::B + l;
  // all 32 loads issued before any is used (unconditional: positions
  // outside the image load element 0 and are zeroed after)
  C x[16];
  double fr[16];
  unsigned in = 0;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int j1 = P * (j0 + u * G::TLm) + r;
    const int ix = image_idx(j1 * n23 + g0, a.N, a.Nt, a.Na);
    in |= ix >= 0 ? 1u << u : 0u;
    x[u] = f[(long long)(ix >= 0 ? ix : 0) * a.ist];
    fr[u] = fb[ix >= 0 ? j1 * G::B : 0];
  }
  // the tables after the data loads (a wait for a table load also waits for
  // the data loads issued before it)
  p1_tables<R, P, M>(a, twm, twn);
  for (int i = threadIdx.x; i < G::L; i += G::NT) {
    const D2 w = big_tw(a, Lm * j2 * i);
    tb[i] = C{(R)w.x, (R)w.y};
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const double s = fr[u] * a.oscale;
    x[u] = (in >> u) & 1 ? C{(R)((double)x[u].x * s), (R)((double)x[u].y * s)} : C{0, 0};
  }
  line_fft<R, M, false>(x, buf + (l * P + r) * G::SS, j0, buf + (t / G::TLm) * G::SS, t % G::TLm, twm);
  __syncthreads();
  {
    C* sl = buf + (l * P + r) * G::SS;
#pragma unroll
    for (int u = 0; u < 16; ++u) sl[j0 + u * G::TLm] = x[u];
  }
  __syncthreads();
  // combine and store: butterfly k' of line ll gives outputs k1 = k' + q M,
  // written contiguous (consecutive k' on consecutive threads)
  for (int idx = t; idx < G::B * M; idx += G::NT) {
    const int ll = idx / M, kp = idx % M;
    C v[P];
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = buf[(ll * P + q) * G::SS + kp];
    if constexpr (P == 5) {
#pragma unroll
      for (int q = 1; q < P; ++q) v[q] = tmul<false>(v[q], twn[q * kp]);
      dft5<R, false>(v);
    }
    C* dst = grid + (size_t)a.n1 * j2 + (size_t)a.n1 * Lm * (j3_0 + ll);
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int k1 = kp + q * M;
      dst[k1] = tmul<false>(v[q], tb[k1]);
    }
  }
}

// Adjoint pass 1 (last): lines (j2, j3 = j3_0 + l) read contiguous, image
// positions written with the transposed deconvolution.
template <typename R, int P, int M>
__global__ __launch_bounds__((P1Geo<R, P, M>::NT)) void k_fft1_last(F1Args a, const typename cpx<R>::T* __restrict__ grid,
                                                                     typename cpx<R>::T* __restrict__ f) {
  using C = typename cpx<R>::T;
  using G = P1Geo<R, P, M>;
  __shared__ __align__(16) C sm[G::LDS_C];
  C* buf = sm;
  C* twm = sm + G::B * P * G::SS;
  C* twn = twm + M;
  const int Lm = a.Lm, n23 = a.Lm * a.Lm;
  const int j2 = blockIdx.y, j3_0 = blockIdx.x * G::B;
  const int t = threadIdx.x;
  // loads: sub-line fastest, then j0 (consecutive points of a line), then line
  const int r = t % P, j0 = (t / P) % G::TLm, l = t / (P * G::TLm);
  const C* src = grid + (size_t)a.n1 * j2 + (size_t)a.n1 * Lm * (j3_0 + l);
  C x[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) x[u] = src[P * (j0 + u * G::TLm) + r];
  // factors of the next butterfly's outputs loaded one butterfly ahead (the
  // first with the data), so that their latency overlaps the work before
  // they are used (zero at the positions that are not written)
  const double* fb = a.fac1 + (size_t)(j2 * (Lm / G::B) + blockIdx.x) * a.n1 * G::B;
  double fcur[P];
#pragma unroll
  for (int q = 0; q < P; ++q) fcur[q] = fb[((t / G::B) + q * M) * G::B + t % G::B];
  p1_tables<R, P, M>(a, twm, twn);
  line_fft<R, M, true>(x, buf + (l * P + r) * G::SS, j0, buf + (t / G::TLm) * G::SS, t % G::TLm, twm);
  __syncthreads();
  {
    C* sl = buf + (l * P + r) * G::SS;
#pragma unroll
    for (int u = 0; u < 16; ++u) sl[j0 + u * G::TLm] = x[u];
  }
  __syncthreads();
  // combine and store: consecutive lines on consecutive threads (consecutive
  // grid positions j1 n2 n3 + j2 n3 + j3)
  for (int idx = t; idx < G::B * M; idx += G::NT) {
    const int ll = idx % G::B, kp = idx / G::B;
    double fnext[P];
    const int nx = idx + G::NT < G::B * M ? idx + G::NT : idx;
#pragma unroll
    for (int q = 0; q < P; ++q) fnext[q] = fb[((nx / G::B) + q * M) * G::B + nx % G::B];
    C v[P];
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = buf[(ll * P + q) * G::SS + kp];
    if constexpr (P == 5) {
#pragma unroll
      for (int q = 1; q < P; ++q) v[q] = tmul<true>(v[q], twn[q * kp]);
      dft5<R, true>(v);
    }
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int g = (kp + q * M) * n23 + j2 * Lm + j3_0 + ll;
      const int ix = image_idx(g, a.N, a.Nt, a.Na);
      if (ix >= 0) {
        const double s = fcur[q] * a.oscale;
        f[(long long)ix * a.ist] = C{(R)((double)v[q].x * s), (R)((double)v[q].y * s)};
      }
      fcur[q] = fnext[q];
    }
  }
}

// Strided pass over lines of Lm points at element stride st, in place: the
// workgroup takes lines a0 + xl (xl < LPB) at offset y ystr, and multiplies
// output point k of line (a, y) by W^(al k + be) after the DFT:
//   TW 0: none (forward pass 3);
//   TW 1: forward pass 2, y = j3: al = n1 y, be = y a;
//   TW 2: adjoint pass 3, y = k2: al = a + n1 y, be = 0;
//   TW 3: adjoint pass 2, y = j3: al = n3 a, be = 0.
// Every exponent is below Nt.
template <typename R, int L, bool INV, int TW, bool FOLD>
__global__ __launch_bounds__((LineGeo<R, L>::NT)) void k_fft1_mid(F1Args a, KArgs k, typename cpx<R>::T* __restrict__ grid,
                                                                   int st, int ystr) {
  using C = typename cpx<R>::T;
  using G = LineGeo<R, L>;
  __shared__ __align__(16) R sm[2 * (G::LPB * G::LS + L)];
  C* buf = (C*)sm;
  C* tw = buf + G::LPB * G::LS;
  const int a0 = blockIdx.x * G::LPB, y = blockIdx.y;
  const int xl = threadIdx.x % G::LPB, j0 = threadIdx.x / G::LPB;
  C* base = grid + (size_t)y * ystr + a0 + xl;
  C v[16];
  if constexpr (FOLD) {
    // the fold of the 1-D spreading slots (k_fold) at the line's points.
    // Tile t's slot covers grid points [t T - m + 1, t T + T + m) (mod Nt):
    // a point g = t T + gm with m <= gm < T - m + 1, away from the ends
    // (covered by the last tile's slot through the wrap below
    // ntiles T - Nt + m, by tile 0's at and above Nt - m + 1), lies in its
    // own tile's slot only; in a tile of at most one work item its value is
    // that item's slot value.  Every other point takes every covering slot
    // (fold_sum).
    const C* scr = (const C*)k.scratch;
    const int T = k.T[0], m1 = k.m - 1, Ntg = k.Nt[0], glo = k.ntiles[0] * T - Ntg + k.m;
    const int g0 = y * ystr + a0 + xl;
    int s0[16], s1[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int t = (g0 + (j0 + u * G::TL) * st) / T;
      s0[u] = k.tstart[t];
      s1[u] = k.tstart[t + 1];
    }
    unsigned slow = 0;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int g = g0 + (j0 + u * G::TL) * st;
      const int gm = g - (g / T) * T;
      const bool fast = gm >= k.m && gm < T - m1 && g >= glo && g < Ntg - m1 && s1[u] - s0[u] <= 1;
      slow |= fast ? 0u : 1u << u;
      v[u] = fast && s1[u] > s0[u] ? scr[(size_t)s0[u] * k.padCells + gm + m1] : C{0, 0};
    }
    if (slow) {
      const int z[1] = {0};
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if ((slow >> u) & 1u) v[u] = fold_sum<R, 1>(k, scr, g0 + (j0 + u * G::TL) * st, 1, z, z, 1, z, z);
    }
  } else {
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = base[(size_t)(j0 + u * G::TL) * st];
  }
  // twiddle loads with the data, the line table after them
  D2 w{1, 0}, st16{1, 0};
  if constexpr (TW != 0) {
    const int aa = a0 + xl;
    const int al = TW == 1 ? a.n1 * y : (TW == 2 ? aa + a.n1 * y : a.Lm * aa);
    const int be = TW == 1 ? y * aa : 0;
    w = big_tw(a, al * j0 + be);
    st16 = big_tw(a, al * G::TL);
  }
  load_tw<G::NT>(tw, a.twl, L);
  line_fft<R, L, INV>(v, buf + xl * G::LS, j0, buf + (threadIdx.x / G::TL) * G::LS, threadIdx.x % G::TL, tw);
  if constexpr (TW != 0) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      v[u] = tmul<INV>(v[u], C{(R)w.x, (R)w.y});
      w = cmul(w, st16);
    }
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) base[(size_t)(j0 + u * G::TL) * st] = v[u];
}

template <typename R, int P, int M>
hipError_t go_p1(bool fwd, const F1Args& a, hipStream_t s, const void* in, void* out) {
  using C = typename cpx<R>::T;
  using G = P1Geo<R, P, M>;
  if (a.B != G::B) return hipErrorInvalidValue;
  const dim3 g((unsigned)(a.Lm / G::B), (unsigned)a.Lm), b(G::NT);
  if (fwd) hipLaunchKernelGGL((k_fft1_first<R, P, M>), g, b, 0, s, a, (const C*)in, (C*)out);
  else hipLaunchKernelGGL((k_fft1_last<R, P, M>), g, b, 0, s, a, (const C*)in, (C*)out);
  return hipGetLastError();
}

template <typename R>
hipError_t pass1(bool fwd, const F1Args& a, hipStream_t s, const void* in, void* out) {
  if (a.P == 5) {
    switch (a.M) {
      case 16: return go_p1<R, 5, 16>(fwd, a, s, in, out);
      case 32: return go_p1<R, 5, 32>(fwd, a, s, in, out);
      case 64: return go_p1<R, 5, 64>(fwd, a, s, in, out);
      case 128: return go_p1<R, 5, 128>(fwd, a, s, in, out);
      case 256:
        if constexpr (sizeof(R) == 4) return go_p1<R, 5, 256>(fwd, a, s, in, out);
        return hipErrorInvalidValue;
      default: return hipErrorInvalidValue;
    }
  }
  if (a.P != 1) return hipErrorInvalidValue;
  switch (a.M) {
    case 64: return go_p1<R, 1, 64>(fwd, a, s, in, out);
    case 128: return go_p1<R, 1, 128>(fwd, a, s, in, out);
    case 256: return go_p1<R, 1, 256>(fwd, a, s, in, out);
    case 512: return go_p1<R, 1, 512>(fwd, a, s, in, out);
    case 1024:
      if constexpr (sizeof(R) == 4) return go_p1<R, 1, 1024>(fwd, a, s, in, out);
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

template <typename R, int L, bool INV, int TW, bool FOLD = false>
hipError_t go_mid(const F1Args& a, const KArgs* k, hipStream_t s, void* grid, int st, int ystr) {
  using C = typename cpx<R>::T;
  using G = LineGeo<R, L>;
  if (a.n1 % G::LPB) return hipErrorInvalidValue;
  KArgs kk;
  if (k) kk = *k;
  else std::memset(&kk, 0, sizeof kk);
  const dim3 g((unsigned)(a.n1 / G::LPB), (unsigned)a.Lm), b(G::NT);
  hipLaunchKernelGGL((k_fft1_mid<R, L, INV, TW, FOLD>), g, b, 0, s, a, kk, (C*)grid, st, ystr);
  return hipGetLastError();
}

template <typename R, bool INV, int TW, bool FOLD = false>
hipError_t mid_l(const F1Args& a, const KArgs* k, hipStream_t s, void* grid, int st, int ystr) {
  switch (a.Lm) {
    case 64: return go_mid<R, 64, INV, TW, FOLD>(a, k, s, grid, st, ystr);
    case 128: return go_mid<R, 128, INV, TW, FOLD>(a, k, s, grid, st, ystr);
    case 256: return go_mid<R, 256, INV, TW, FOLD>(a, k, s, grid, st, ystr);
    default: return hipErrorInvalidValue;
  }
}

template <typename R>
hipError_t fwd_rest(const F1Args& a, hipStream_t s, void* grid) {
  const int n12 = a.n1 * a.Lm;
  // pass 2: lines (a, j3), points j2 at stride n1, line offset j3 n1 n2
  hipError_t e = mid_l<R, false, 1>(a, nullptr, s, grid, a.n1, n12);
  // pass 3: lines (a, k2), points j3 at stride n1 n2, line offset k2 n1
  if (e == hipSuccess) e = mid_l<R, false, 0>(a, nullptr, s, grid, n12, a.n1);
  return e;
}

// Adjoint passes 3 and 2 (part 0), pass 3 with the fold of the spreading
// slots (part 3), pass 2 alone (part 2).
template <typename R>
hipError_t adj_first(const F1Args& a, const KArgs* k, hipStream_t s, void* grid, int part) {
  const int n12 = a.n1 * a.Lm;
  hipError_t e = hipSuccess;
  if (part == 0) e = mid_l<R, true, 2>(a, nullptr, s, grid, n12, a.n1);
  if (part == 3) return mid_l<R, true, 2, true>(a, k, s, grid, n12, a.n1);
  if (e == hipSuccess) e = mid_l<R, true, 3>(a, nullptr, s, grid, a.n1, n12);
  return e;
}

}  // namespace

bool fft1_factor(int64_t Nt, int64_t N, bool dbl, int* n1o, int* Lmo) {
  if (N < 1 || N > Nt || Nt >= (1LL << 31)) return false;
  for (int Lm : {256, 128, 64}) {
    const int64_t q = (int64_t)Lm * Lm;
    if (Nt % q) continue;
    const int64_t n1 = Nt / q;
    const int P = n1 % 5 == 0 ? 5 : 1;
    const int64_t M = n1 / P;
    if (M & (M - 1)) continue;
    if (P == 1 && (M < 64 || M > 1024)) continue;
    if (P == 1 && dbl && M > 512) continue;
    if (P == 5 && (M < 16 || M > (dbl ? 128 : 256))) continue;
    if (n1 % ((dbl ? 2048 : 4096) / Lm)) continue;
    *n1o = (int)n1;
    *Lmo = Lm;
    return true;
  }
  return false;
}

hipError_t launch_fft1_fwd(bool dbl, const F1Args& a, hipStream_t s, const void* f, void* grid, int part) {
  if (part == 0) return dbl ? pass1<double>(true, a, s, f, grid) : pass1<float>(true, a, s, f, grid);
  return dbl ? fwd_rest<double>(a, s, grid) : fwd_rest<float>(a, s, grid);
}

hipError_t launch_fft1_adj(bool dbl, const F1Args& a, hipStream_t s, void* grid, void* f, int part) {
  if (part == 1) return dbl ? pass1<double>(false, a, s, grid, f) : pass1<float>(false, a, s, grid, f);
  if (part != 0 && part != 2) return hipErrorInvalidValue;
  return dbl ? adj_first<double>(a, nullptr, s, grid, part) : adj_first<float>(a, nullptr, s, grid, part);
}

hipError_t launch_fft1_adj_fold(bool dbl, const F1Args& a, const KArgs& k, hipStream_t s, void* grid) {
  if (k.D != 1) return hipErrorInvalidValue;
  return dbl ? adj_first<double>(a, &k, s, grid, 3) : adj_first<float>(a, &k, s, grid, 3);
}

}  // namespace nfftm
