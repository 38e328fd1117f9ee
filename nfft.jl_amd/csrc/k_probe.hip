// k_probe.hip -- issue cost of the LDS instructions the gridding kernels are
// bound by, measured on the current device: the per-CU cycles of one
// wavefront instruction of ds_read_b64 (interpolation taps), ds_read_b128,
// ds_add_u64 (FP32 3-D spreading taps: packed fixed point) and ds_add_f64
// (FP64 spreading taps), with every lane on its own bank group.  The
// bench's secondary (LDS) roofline divides the taps of a kernel by these.
#include "launch.hpp"

namespace nfftm {

namespace {

constexpr int PNT = 1024;      // threads per workgroup, two workgroups per CU
constexpr int PITER = 1024;    // iterations of 8 instructions per thread
constexpr int PWORDS = 8192;   // 64 KiB of 8-byte words

template <int OP>
__global__ __launch_bounds__(PNT) void k_lds_probe(unsigned long long* __restrict__ out, int salt) {
  __shared__ __align__(16) unsigned long long lds[PWORDS];
  for (int i = threadIdx.x; i < PWORDS; i += PNT) lds[i] = (unsigned long long)(i * 3 + salt);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long acc0 = 0, acc1 = 0;
  double accd = 0;
  // 8-byte operations: lane * 8 B; 16-byte reads: lane * 16 B (aligned)
  int base = ((OP == 1 ? 2 * lane : lane) + wave * 128) & (PWORDS - 1);
  for (int it = 0; it < PITER; ++it) {
    int offs[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      int o = (base + u * 256) & (PWORDS - 1);
      asm volatile("" : "+v"(o));  // opaque address: one instruction per access, no merging
      offs[u] = o;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (OP == 0) {
        acc0 += lds[offs[u]];
      } else if (OP == 1) {
        const ulonglong2 v = *(const ulonglong2*)&lds[offs[u] & (PWORDS - 2)];
        acc0 += v.x;
        acc1 += v.y;
      } else if (OP == 2) {
        __hip_atomic_fetch_add(&lds[offs[u]], (unsigned long long)lane, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
      } else {
        __hip_atomic_fetch_add((double*)&lds[offs[u]], 1.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    base = (base + 2) & (PWORDS - 1);
  }
  __syncthreads();
  if (OP >= 2) {
    acc0 = lds[threadIdx.x];
    accd = ((const double*)lds)[threadIdx.x + PNT];
  }
  out[(size_t)blockIdx.x * PNT + threadIdx.x] = acc0 + acc1 + (unsigned long long)(accd != 0);
}

template <int OP>
hipError_t time_probe(unsigned long long* d, int blocks, hipEvent_t a, hipEvent_t b, float* ms) {
  hipLaunchKernelGGL(k_lds_probe<OP>, dim3(blocks), dim3(PNT), 0, 0, d, 1);
  hipError_t e = hipEventRecord(a, 0);
  for (int r = 0; r < 3 && e == hipSuccess; ++r) {
    hipLaunchKernelGGL(k_lds_probe<OP>, dim3(blocks), dim3(PNT), 0, 0, d, r);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(b, 0);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
  *ms /= 3;
  return e;
}

}  // namespace

hipError_t lds_issue_cycles(double cycles[4], double* clock_hz, int* cus) {
  int dev = 0, ncu = 0, khz = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, dev);
  if (e != hipSuccess) return e;
  const int blocks = 2 * ncu;
  unsigned long long* d = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  e = hipMalloc(&d, sizeof(unsigned long long) * (size_t)blocks * PNT);
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  float ms[4] = {0, 0, 0, 0};
  if (e == hipSuccess) e = time_probe<0>(d, blocks, a, b, &ms[0]);
  if (e == hipSuccess) e = time_probe<1>(d, blocks, a, b, &ms[1]);
  if (e == hipSuccess) e = time_probe<2>(d, blocks, a, b, &ms[2]);
  if (e == hipSuccess) e = time_probe<3>(d, blocks, a, b, &ms[3]);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (d) (void)hipFree(d);
  if (e != hipSuccess) return e;
  // wavefront instructions per CU: 2 workgroups x 16 waves x PITER x 8
  const double per_cu = 2.0 * (PNT / 64) * PITER * 8;
  for (int i = 0; i < 4; ++i) cycles[i] = ms[i] * 1e-3 * khz * 1e3 / per_cu;
  *clock_hz = khz * 1e3;
  *cus = ncu;
  return hipSuccess;
}

}  // namespace nfftm
