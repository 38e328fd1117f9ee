// fft1.hpp -- the FFT stage of 1-D transforms on large grids (k_fft1.hip):
// Nt = n1 Lm^2 points in three passes of line FFTs, the deconvolution fused
// into the first forward pass and the last adjoint pass.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nfftm {

struct KArgs;

struct F1Args {
  int n1, P, M, Lm;  // n1 = P M (P = 1 or 5, M a power of two), n2 = n3 = Lm
  int N, Na, Nt;     // image size, N / 2, grid size n1 Lm^2
  long long ist;     // element stride of the image
  int sh;            // W^e = thi[e >> sh] tlo[e mod 2^sh], W = exp(-+2 pi i / Nt)
  int B;              // lines per workgroup of pass 1 (fft1_lines)
  const double* fac1;  // deconvolution factors at grid position j1 n2 n3 + j2 n3 + j3 (image
                       // positions only) stored at ((j2 Lm / B + j3 / B) n1 + j1) B + j3 mod B
  const void* tw1;    // (cos, sin)(2 pi j / n1), j < n1
  const void* twl;    // (cos, sin)(2 pi j / Lm), j < Lm
  const void* thi;    // FP64 (cos, sin)(2 pi (j << sh) / Nt), j <= (Nt - 1) >> sh
  const void* tlo;    // FP64 (cos, sin)(2 pi j / Nt), j < 2^sh
  double oscale;
};

// n1 and Lm with Nt = n1 Lm^2 for which the passes are built (Lm in {256,
// 128, 64}; n1 = 2^k in [64, 1024] or 5 2^k in [80, 1280], FP64 at most
// 512 and 640; n1 a multiple of the lines per workgroup of the strided
// passes), else false.
bool fft1_factor(int64_t Nt, int64_t N, bool dbl, int* n1, int* Lm);
// Lines per workgroup of pass 1 (16 in FP32, 8 in FP64, halved for
// n1 >= 512 and quartered for n1 >= 1024).
inline int fft1_lines(int n1, bool dbl) { return (dbl ? 8 : 16) / (n1 >= 1024 ? 4 : (n1 >= 512 ? 2 : 1)); }
// Forward: part 0 = the first pass (f deconvolved into the grid), part 1 =
// the other two passes, in place.
hipError_t launch_fft1_fwd(bool dbl, const F1Args& a, hipStream_t s, const void* f, void* grid, int part);
// Adjoint: part 0 = the first two passes, in place, part 1 = the last pass
// into f with the transposed deconvolution, part 2 = the second pass alone
// (after launch_fft1_adj_fold).
hipError_t launch_fft1_adj(bool dbl, const F1Args& a, hipStream_t s, void* grid, void* f, int part);
// Adjoint first pass reading the 1-D spreading slots of k (the fold of
// k_fold fused): the grid needs no fold launch before it.
hipError_t launch_fft1_adj_fold(bool dbl, const F1Args& a, const KArgs& k, hipStream_t s, void* grid);

}  // namespace nfftm
