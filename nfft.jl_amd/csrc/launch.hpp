// launch.hpp -- host-side launchers of the device kernels (one translation
// unit per tile dimensionality so they compile in parallel).
#pragma once
#include <hip/hip_runtime.h>
#include "kernels.hpp"

#include <algorithm>

namespace nfftm {

// Arguments of the deconvolution kernels (src/deconvolution.jl).
struct DArgs {
  int D;
  int N[MAXD], Nt[MAXD], Na[MAXD];
  long long nstr[MAXD], gstr[MAXD];
  const double* fac[MAXD];  // 1/psi_hat per dim, FP64, N[d] entries
  long long P, Q;
  double oscale;  // 1 inside a transform; wscale^D for standalone deconvolve!
  FastDiv div_half;        // Nt[0] / 2
  FastDiv div_nt[MAXD];    // Nt[d]
};

// Arguments of the pruned FFT passes (k_fft.hip): per dimension the image
// size, the oversampled grid size (1 beyond D), Na = N / 2 of the image
// placement, the grid stride, the deconvolution factors (FP64) and the
// twiddle table exp(-2 pi i k / Nt) (cos, sin pairs of the data precision,
// k < Nt).
struct FArgs {
  int D;
  long long ist;  // element stride of the image (directional plans; otherwise 1)
  int N[3], L[3], Na[3];
  long long gstr[3];
  const double* fac[3];
  const void* tw[3];
  double oscale;
};

// Power-of-two grids within the line kernels' LDS (FP32: 16..2048 points
// per dimension, FP64: 16..1024), D <= 3, N_d <= Nt_d.
bool pruned_fft_supported(int D, const int64_t* N, const int64_t* Nt, bool dbl);
// Adjoint, 3-D block-layout slots with at most one work item per tile: the
// fold fused into the first pruned pass (dimension 3), grid positions S_z of
// every (x, y) line written.  hipErrorNotSupported (nothing launched) where
// the tile geometry does not match the pass; dry = only that check.
hipError_t launch_fold_fft3(bool dbl, int W, const KArgs& a, const FArgs& f, hipStream_t s, void* grid, bool dry);
hipError_t launch_pruned_fft_fwd(bool dbl, const FArgs& a, hipStream_t s, const void* f, void* grid, int part);
hipError_t launch_pruned_fft_adj(bool dbl, const FArgs& a, hipStream_t s, void* grid, void* f, int part);

// NFFTTools callers (k_tools.hip): reductions over complex vectors into a
// device scalar, the sdc elementwise steps, and the Toeplitz pad / product /
// crop of an image of size N in an oversampled array of size L (column-major,
// D <= 3).
enum { RED_MAX_RE = 0, RED_MIN_RE = 1, RED_SUM_RE = 2, RED_SUM_ABS2 = 3 };
struct ToepArgs {
  int D;
  long long N[3], L[3];
  long long P, Q;  // prod N, prod L
};
// part: at least 2048 doubles
hipError_t launch_reduce(bool dbl, int op, const void* x, long long n, double* part, double* out, hipStream_t s);
hipError_t launch_div_scalar(bool dbl, void* g, long long n, const double* sc, hipStream_t s);
hipError_t launch_sdc_update(bool dbl, void* wc, const void* wt, long long n, const double* sc, int* bad,
                             hipStream_t s);
hipError_t launch_real_to_cpx(bool dbl, const void* w, void* wc, long long n, hipStream_t s);
hipError_t launch_cpx_to_real(bool dbl, const void* wc, void* w, long long n, const double* sc, hipStream_t s);
hipError_t launch_fill_one(bool dbl, void* u, long long n, hipStream_t s);
// Element types of caller buffers (the values of the C ABI's NFFTM_R32 ...
// NFFTM_C16) and the conversion of n values between any two of them.
enum { ELT_R32 = 0, ELT_R64 = 1, ELT_C32 = 2, ELT_C64 = 3, ELT_R16 = 4, ELT_C16 = 5 };
hipError_t launch_convert(int tin, const void* in, int tout, void* out, long long n, hipStream_t s);
hipError_t launch_mul_real(bool dbl, void* f, const void* wc, long long n, hipStream_t s);
hipError_t launch_pad_corner(bool dbl, const void* y, void* x, const ToepArgs& a, hipStream_t s);
hipError_t launch_crop_corner(bool dbl, const void* x, void* y, const ToepArgs& a, double sc, hipStream_t s);
hipError_t launch_cmul(bool dbl, void* x, const void* lam, long long n, hipStream_t s);
hipError_t launch_shift_sign(bool dbl, void* x, const ToepArgs& a, hipStream_t s);

// iperm[perm[j]] = j, and out[i * os] = in[idx[i] * is] (node values between
// the caller's and the sorted order).
hipError_t launch_inv_perm(const int* perm, int* iperm, long long n, hipStream_t s);
hipError_t launch_gather_nodes(bool dbl, const void* in, long long is, const int* idx, void* out, long long os,
                               long long n, hipStream_t s);

// LDS bytes of a tile kernel.
size_t tile_lds_bytes(int D, bool dbl, int W, int tileElems);

// Tile kernels (D = 1, 2, 3).  grid = 8 * ceil(nitems / 8) workgroups of NT.
hipError_t launch_tile(int D, bool spread, bool dbl, int W, int NT, const KArgs& a, size_t lds,
                       hipStream_t s, const void* in, void* out);
hipError_t launch_tile_d1(bool spread, bool dbl, int W, int NT, const KArgs& a, size_t lds,
                          hipStream_t s, const void* in, void* out);
hipError_t launch_tile_d2(bool spread, bool dbl, int W, int NT, const KArgs& a, size_t lds,
                          hipStream_t s, const void* in, void* out);
hipError_t launch_tile_d3(bool spread, bool dbl, int W, int NT, const KArgs& a, size_t lds,
                          hipStream_t s, const void* in, void* out);

// Node-parallel kernels for any D (no tiles; global atomics for spreading).
hipError_t launch_generic(bool spread, bool dbl, int W, const KArgs& a, hipStream_t s,
                          const void* in, void* out);

// Node gradient kbar[d + D j] = Re(conj(ybar_j) d fhat_j / d k_{d,j}) from
// the transformed grid g of the forward path.
hipError_t launch_grad_nodes(bool dbl, int W, const KArgs& a, hipStream_t s, const void* g,
                             const void* ybar, double* kbar);

// Batch stages (k_batch.hip): persistent kernels striding over (slot, item)
// work lists; ka holds one KArgs per slot in device memory.
hipError_t launch_tile_batch(int D, bool spread, bool dbl, int W, int NT, int nblk, size_t lds, hipStream_t s,
                             const KArgs* ka, const int2* work, int nwork, const void* const* in,
                             void* const* out);
hipError_t launch_fold_batch(int D, bool dbl, int nblk, hipStream_t s, const KArgs* ka, const int4* work,
                             int nwork, void* const* grids);
int tile_batch_blocks(int D, bool spread, bool dbl, int W, int NT, size_t lds);
// work entries {slot, begin, end, 0}: element pairs of the grid (forward) or
// image rows (transpose)
hipError_t launch_deconv_batch(bool dbl, int nblk, hipStream_t s, const DArgs* da, const int4* work, int nwork,
                               const void* const* in, void* const* out);
hipError_t launch_deconv_t_batch(bool dbl, int nblk, hipStream_t s, const DArgs* da, const int4* work, int nwork,
                                 const void* const* in, void* const* out);

// Node binning on the device (k_bin.hip).  Tiles per dimension as the tile
// kernels use them; CH = most nodes of one work item; key_bits covers the
// tile indices (ntile <= 2^key_bits).
struct BinArgs {
  int D;
  long long M;
  int Nt[3], T[3], ntiles[3];
  int ntile, CH, key_bits;
};
size_t bin_sort_temp_bytes(long long M, int key_bits);
hipError_t launch_bin_keys(const BinArgs& a, const double* k, int* key, int* idx, int* cnt, hipStream_t s);
hipError_t launch_bin_sort(const BinArgs& a, void* temp, size_t temp_bytes, int* key_in, int* key_out, int* idx_in,
                           int* idx_out, hipStream_t s);
hipError_t launch_bin_scan(const BinArgs& a, const int* cnt, int* off, int* ioff, int* stats, hipStream_t s);
hipError_t launch_bin_items(const BinArgs& a, const int* off, const int* ioff, int4* items, int* tstart,
                            hipStream_t s);
hipError_t launch_bin_ks(const BinArgs& a, const double* k, const int* perm, double* ks, hipStream_t s);
hipError_t launch_iota(int* p, long long n, hipStream_t s);

// Node records (weights, local cell, caller index) of every work item.
hipError_t launch_node_records(int D, bool dbl, int W, const KArgs& a, void* rec, hipStream_t s);

hipError_t launch_deconv(bool dbl, const DArgs& a, hipStream_t s, const void* f, void* g);
hipError_t launch_deconv_t(bool dbl, const DArgs& a, hipStream_t s, const void* g, void* f);
hipError_t launch_fill0(void* p, size_t bytes, hipStream_t s);
// dst[i * ds] = src[i * ss] for n complex values of the precision
hipError_t launch_copy_strided(bool dbl, const void* src, long long ss, void* dst, long long ds, long long n,
                               hipStream_t s);

// Sum the per-item padded tiles in a.scratch into the grid (D <= 3).
hipError_t launch_fold(int D, bool dbl, int W, const KArgs& a, hipStream_t s, void* grid);
// Sparse-matrix convolution (NFFTM_FLAG_SPARSE): build the M x W^D rows of
// window weights and grid indices, then forward (interpolation) or adjoint
// (spreading by atomics into a zeroed grid) products with them.
hipError_t launch_sparse_build(bool dbl, int W, const KArgs& a, hipStream_t s, int* cols, void* vals);
hipError_t launch_sparse(bool spread, bool dbl, const KArgs& a, int K, const int* cols, const void* vals,
                         hipStream_t s, const void* in, void* out);
hipError_t launch_fold_d3(bool dbl, int W, const KArgs& a, hipStream_t s, void* grid);

// Per-CU cycles of one wavefront instruction of ds_read_b64, ds_read_b128,
// ds_add_u64 and ds_add_f64 on the current device (k_probe.hip), with the
// shader clock and the CU count they are counted against.
hipError_t lds_issue_cycles(double cycles[4], double* clock_hz, int* cus);

// Compile-time tile geometry of the 3-D kernels: tile edge T, padded edge P,
// dynamic LDS bytes.
// V = 0: largest cubic tile within 160 KiB of LDS; V = 1: flat tile within
// 80 KiB (two workgroups per CU).  T[0] / P[0]: x and y tile / padded edge,
// T[1] / P[1]: z.  valid is false when no tile of that variant fits.
void tile3_geometry(bool dbl, int W, int V, int* T, size_t* lds, int* P, int* SX, bool* valid);

}  // namespace nfftm
