// k_tools.hip -- elementwise and reduction kernels of the NFFTTools callers:
// sampling density compensation (NFFTTools/src/samplingDensity.jl:93-155) and
// the Toeplitz embedding of the normal operator A'A
// (NFFTTools/src/Toeplitz.jl:86-250).  Every scalar the iterations need (the
// first grid maximum, the final least-squares scale) stays in device memory,
// so an sdc run or a Toeplitz product enqueues without host round trips.
#include "launch.hpp"

namespace nfftm {

namespace {

constexpr int RB = 256;  // threads of the reduction and elementwise kernels

template <int OP>
__device__ __forceinline__ double red_init() {
  return OP == RED_MAX_RE ? -INFINITY : OP == RED_MIN_RE ? INFINITY : 0.0;
}
template <int OP>
__device__ __forceinline__ double red_op(double a, double b) {
  if (OP == RED_MAX_RE) return fmax(a, b);
  if (OP == RED_MIN_RE) return fmin(a, b);
  return a + b;
}
template <int OP, typename C>
__device__ __forceinline__ double red_val(C z) {
  if (OP == RED_SUM_ABS2) return (double)z.x * (double)z.x + (double)z.y * (double)z.y;
  return (double)z.x;
}

// Workgroup reduction of v (all RB threads); the result in thread 0.
template <int OP>
__device__ __forceinline__ double block_reduce(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = red_op<OP>(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < RB / 64; ++w) v = red_op<OP>(v, sh[w]);
  }
  return v;
}

// Partial reductions of x[0 .. n) into part[blockIdx.x]; the last kernel
// folds the partials into out[0] (deterministic order).
template <int OP, typename C>
__global__ __launch_bounds__(RB) void k_reduce(const C* __restrict__ x, long long n, double* __restrict__ part) {
  __shared__ double sh[RB / 64];
  double v = red_init<OP>();
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB)
    v = red_op<OP>(v, red_val<OP>(x[i]));
  v = block_reduce<OP>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

template <int OP>
__global__ __launch_bounds__(RB) void k_reduce_final(const double* __restrict__ part, int nb,
                                                     double* __restrict__ out) {
  __shared__ double sh[RB / 64];
  double v = red_init<OP>();
  for (int i = threadIdx.x; i < nb; i += RB) v = red_op<OP>(v, part[i]);
  v = block_reduce<OP>(v, sh);
  if (threadIdx.x == 0) out[0] = v;
}

// g /= s[0] (real scalar of the data precision, the maximum in FP64 is exact
// as a value of that precision).
template <typename C, typename R>
__global__ __launch_bounds__(RB) void k_div_scalar(C* __restrict__ g, long long n, const double* __restrict__ s) {
  const R d = (R)s[0];
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB) {
    const C v = g[i];
    g[i] = C{v.x / d, v.y / d};
  }
}

// One sdc update: t = Re(wt) / s; w /= t (the weights are the real parts of
// the complex vector wc); bad[blockIdx] = 1 if some t <= 0.
template <typename C, typename R>
__global__ __launch_bounds__(RB) void k_sdc_update(C* __restrict__ wc, const C* __restrict__ wt, long long n,
                                                   const double* __restrict__ s, int* __restrict__ bad) {
  const R d = (R)s[0];
  int b = 0;
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB) {
    const R t = wt[i].x / d;
    b |= !(t > (R)0);
    wc[i] = C{wc[i].x / t, (R)0};
  }
  if (b) bad[0] = 1;  // plain vector store of the same value from any lane
}

template <typename C, typename R>
__global__ __launch_bounds__(RB) void k_real_to_cpx(const R* __restrict__ w, C* __restrict__ wc, long long n) {
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB)
    wc[i] = C{w[i], (R)0};
}

// w = Re(wc) * (c = s[0] / s[1] if s, else 1)
template <typename C, typename R>
__global__ __launch_bounds__(RB) void k_cpx_to_real(const C* __restrict__ wc, R* __restrict__ w, long long n,
                                                    const double* __restrict__ s) {
  const R c = s ? (R)(s[0] / s[1]) : (R)1;
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB)
    w[i] = wc[i].x * c;
}

// Element-type conversion of a vector of n values (ELT_* codes): real to
// complex with a zero imaginary part, complex to real by the real part, and
// any precision change by one rounding (through double).
template <int T>
struct Elt;
template <>
struct Elt<ELT_R16> { using V = _Float16; static constexpr bool CPX = false; };
template <>
struct Elt<ELT_R32> { using V = float; static constexpr bool CPX = false; };
template <>
struct Elt<ELT_R64> { using V = double; static constexpr bool CPX = false; };
template <>
struct Elt<ELT_C16> { using V = _Float16; static constexpr bool CPX = true; };
template <>
struct Elt<ELT_C32> { using V = float; static constexpr bool CPX = true; };
template <>
struct Elt<ELT_C64> { using V = double; static constexpr bool CPX = true; };

template <int TI, int TO>
__global__ __launch_bounds__(RB) void k_convert(const typename Elt<TI>::V* __restrict__ in,
                                                typename Elt<TO>::V* __restrict__ out, long long n) {
  using VI = typename Elt<TI>::V;
  using VO = typename Elt<TO>::V;
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB) {
    const VI re = Elt<TI>::CPX ? in[2 * i] : in[i];
    const VI im = Elt<TI>::CPX ? in[2 * i + 1] : (VI)0;
    if constexpr (Elt<TO>::CPX) {
      out[2 * i] = (VO)(double)re;
      out[2 * i + 1] = (VO)(double)im;
    } else {
      out[i] = (VO)(double)re;
    }
  }
}

template <typename C, typename R>
__global__ __launch_bounds__(RB) void k_fill_one(C* __restrict__ u, long long n) {
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB)
    u[i] = C{(R)1, (R)0};
}

// f *= Re(wc) (the sdc weights)
template <typename C, typename R>
__global__ __launch_bounds__(RB) void k_mul_real(C* __restrict__ f, const C* __restrict__ wc, long long n) {
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB) {
    const R w = wc[i].x;
    const C v = f[i];
    f[i] = C{v.x * w, v.y * w};
  }
}

// Toeplitz: x (oversampled, sizes L) = y in the corner [0, N) per dimension,
// zero elsewhere.
template <typename C, typename R>
__global__ __launch_bounds__(RB) void k_pad_corner(const C* __restrict__ y, C* __restrict__ x, ToepArgs a) {
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < a.Q; i += (long long)gridDim.x * RB) {
    long long r = i, src = 0, st = 1;
    bool in = true;
    for (int d = 0; d < a.D; ++d) {
      const long long c = r % a.L[d];
      r /= a.L[d];
      in &= c < a.N[d];
      src += c * st;
      st *= a.N[d];
    }
    x[i] = in ? y[src] : C{(R)0, (R)0};
  }
}

// y = x[corner] * s (the inverse FFT's 1 / prod L)
template <typename C, typename R>
__global__ __launch_bounds__(RB) void k_crop_corner(const C* __restrict__ x, C* __restrict__ y, ToepArgs a, double s) {
  const R sc = (R)s;
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < a.P; i += (long long)gridDim.x * RB) {
    long long r = i, dst = 0, st = 1;
    for (int d = 0; d < a.D; ++d) {
      const long long c = r % a.N[d];
      r /= a.N[d];
      dst += c * st;
      st *= a.L[d];
    }
    const C v = x[dst];
    y[i] = C{v.x * sc, v.y * sc};
  }
}

// x *= lam (elementwise complex product)
template <typename C, typename R>
__global__ __launch_bounds__(RB) void k_cmul(C* __restrict__ x, const C* __restrict__ lam, long long n) {
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB) {
    const C a = x[i], b = lam[i];
    x[i] = C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
  }
}

// FFT(fftshift(e)) = FFT(e) * (-1)^(k_1 + ... + k_D) for even sizes: flip the
// sign of the points with an odd index sum.
template <typename C, typename R>
__global__ __launch_bounds__(RB) void k_shift_sign(C* __restrict__ x, ToepArgs a) {
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < a.Q; i += (long long)gridDim.x * RB) {
    long long r = i, s = 0;
    for (int d = 0; d < a.D; ++d) {
      s += r % a.L[d];
      r /= a.L[d];
    }
    if (s & 1) {
      const C v = x[i];
      x[i] = C{-v.x, -v.y};
    }
  }
}

int grid_blocks(long long n) {
  const long long b = (n + RB - 1) / RB;
  return (int)std::max<long long>(1, std::min<long long>(b, 2048));
}

}  // namespace

hipError_t launch_reduce(bool dbl, int op, const void* x, long long n, double* part, double* out, hipStream_t s) {
  const int nb = grid_blocks(n);  // <= 2048 partials
#define NFFTM_RED(OP)                                                                                          \
  case OP:                                                                                                     \
    if (dbl) hipLaunchKernelGGL((k_reduce<OP, double2>), dim3(nb), dim3(RB), 0, s, (const double2*)x, n, part); \
    else hipLaunchKernelGGL((k_reduce<OP, float2>), dim3(nb), dim3(RB), 0, s, (const float2*)x, n, part);     \
    hipLaunchKernelGGL((k_reduce_final<OP>), dim3(1), dim3(RB), 0, s, part, nb, out);                         \
    break;
  switch (op) {
    NFFTM_RED(RED_MAX_RE) NFFTM_RED(RED_MIN_RE) NFFTM_RED(RED_SUM_RE) NFFTM_RED(RED_SUM_ABS2)
    default:
      return hipErrorInvalidValue;
  }
#undef NFFTM_RED
  return hipGetLastError();
}

hipError_t launch_div_scalar(bool dbl, void* g, long long n, const double* sc, hipStream_t s) {
  if (dbl) hipLaunchKernelGGL((k_div_scalar<double2, double>), dim3(grid_blocks(n)), dim3(RB), 0, s, (double2*)g, n, sc);
  else hipLaunchKernelGGL((k_div_scalar<float2, float>), dim3(grid_blocks(n)), dim3(RB), 0, s, (float2*)g, n, sc);
  return hipGetLastError();
}

hipError_t launch_sdc_update(bool dbl, void* wc, const void* wt, long long n, const double* sc, int* bad,
                             hipStream_t s) {
  if (dbl)
    hipLaunchKernelGGL((k_sdc_update<double2, double>), dim3(grid_blocks(n)), dim3(RB), 0, s, (double2*)wc,
                       (const double2*)wt, n, sc, bad);
  else
    hipLaunchKernelGGL((k_sdc_update<float2, float>), dim3(grid_blocks(n)), dim3(RB), 0, s, (float2*)wc,
                       (const float2*)wt, n, sc, bad);
  return hipGetLastError();
}

hipError_t launch_real_to_cpx(bool dbl, const void* w, void* wc, long long n, hipStream_t s) {
  if (dbl)
    hipLaunchKernelGGL((k_real_to_cpx<double2, double>), dim3(grid_blocks(n)), dim3(RB), 0, s, (const double*)w,
                       (double2*)wc, n);
  else
    hipLaunchKernelGGL((k_real_to_cpx<float2, float>), dim3(grid_blocks(n)), dim3(RB), 0, s, (const float*)w,
                       (float2*)wc, n);
  return hipGetLastError();
}

hipError_t launch_cpx_to_real(bool dbl, const void* wc, void* w, long long n, const double* sc, hipStream_t s) {
  if (dbl)
    hipLaunchKernelGGL((k_cpx_to_real<double2, double>), dim3(grid_blocks(n)), dim3(RB), 0, s, (const double2*)wc,
                       (double*)w, n, sc);
  else
    hipLaunchKernelGGL((k_cpx_to_real<float2, float>), dim3(grid_blocks(n)), dim3(RB), 0, s, (const float2*)wc,
                       (float*)w, n, sc);
  return hipGetLastError();
}

hipError_t launch_fill_one(bool dbl, void* u, long long n, hipStream_t s) {
  if (dbl) hipLaunchKernelGGL((k_fill_one<double2, double>), dim3(grid_blocks(n)), dim3(RB), 0, s, (double2*)u, n);
  else hipLaunchKernelGGL((k_fill_one<float2, float>), dim3(grid_blocks(n)), dim3(RB), 0, s, (float2*)u, n);
  return hipGetLastError();
}

hipError_t launch_mul_real(bool dbl, void* f, const void* wc, long long n, hipStream_t s) {
  if (dbl)
    hipLaunchKernelGGL((k_mul_real<double2, double>), dim3(grid_blocks(n)), dim3(RB), 0, s, (double2*)f,
                       (const double2*)wc, n);
  else
    hipLaunchKernelGGL((k_mul_real<float2, float>), dim3(grid_blocks(n)), dim3(RB), 0, s, (float2*)f,
                       (const float2*)wc, n);
  return hipGetLastError();
}

hipError_t launch_pad_corner(bool dbl, const void* y, void* x, const ToepArgs& a, hipStream_t s) {
  const long long n = a.Q;
  if (dbl)
    hipLaunchKernelGGL((k_pad_corner<double2, double>), dim3(grid_blocks(n)), dim3(RB), 0, s, (const double2*)y,
                       (double2*)x, a);
  else
    hipLaunchKernelGGL((k_pad_corner<float2, float>), dim3(grid_blocks(n)), dim3(RB), 0, s, (const float2*)y,
                       (float2*)x, a);
  return hipGetLastError();
}

hipError_t launch_crop_corner(bool dbl, const void* x, void* y, const ToepArgs& a, double sc, hipStream_t s) {
  const long long n = a.P;
  if (dbl)
    hipLaunchKernelGGL((k_crop_corner<double2, double>), dim3(grid_blocks(n)), dim3(RB), 0, s, (const double2*)x,
                       (double2*)y, a, sc);
  else
    hipLaunchKernelGGL((k_crop_corner<float2, float>), dim3(grid_blocks(n)), dim3(RB), 0, s, (const float2*)x,
                       (float2*)y, a, sc);
  return hipGetLastError();
}

hipError_t launch_cmul(bool dbl, void* x, const void* lam, long long n, hipStream_t s) {
  if (dbl)
    hipLaunchKernelGGL((k_cmul<double2, double>), dim3(grid_blocks(n)), dim3(RB), 0, s, (double2*)x,
                       (const double2*)lam, n);
  else
    hipLaunchKernelGGL((k_cmul<float2, float>), dim3(grid_blocks(n)), dim3(RB), 0, s, (float2*)x,
                       (const float2*)lam, n);
  return hipGetLastError();
}

hipError_t launch_shift_sign(bool dbl, void* x, const ToepArgs& a, hipStream_t s) {
  const long long n = a.Q;
  if (dbl)
    hipLaunchKernelGGL((k_shift_sign<double2, double>), dim3(grid_blocks(n)), dim3(RB), 0, s, (double2*)x, a);
  else hipLaunchKernelGGL((k_shift_sign<float2, float>), dim3(grid_blocks(n)), dim3(RB), 0, s, (float2*)x, a);
  return hipGetLastError();
}

template <int TI>
static hipError_t convert_to(int tout, const void* in, void* out, long long n, hipStream_t s, unsigned nb) {
#define NFFTM_CV(TO)                                                                                  \
  case TO:                                                                                            \
    hipLaunchKernelGGL((k_convert<TI, TO>), dim3(nb), dim3(RB), 0, s, (const typename Elt<TI>::V*)in, \
                       (typename Elt<TO>::V*)out, n);                                                 \
    break;
  switch (tout) {
    NFFTM_CV(ELT_R32) NFFTM_CV(ELT_R64) NFFTM_CV(ELT_C32) NFFTM_CV(ELT_C64) NFFTM_CV(ELT_R16) NFFTM_CV(ELT_C16)
    default:
      return hipErrorInvalidValue;
  }
#undef NFFTM_CV
  return hipGetLastError();
}

hipError_t launch_convert(int tin, const void* in, int tout, void* out, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned nb = (unsigned)grid_blocks(n);
  switch (tin) {
    case ELT_R32: return convert_to<ELT_R32>(tout, in, out, n, s, nb);
    case ELT_R64: return convert_to<ELT_R64>(tout, in, out, n, s, nb);
    case ELT_C32: return convert_to<ELT_C32>(tout, in, out, n, s, nb);
    case ELT_C64: return convert_to<ELT_C64>(tout, in, out, n, s, nb);
    case ELT_R16: return convert_to<ELT_R16>(tout, in, out, n, s, nb);
    case ELT_C16: return convert_to<ELT_C16>(tout, in, out, n, s, nb);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace nfftm
