/*
 * nfftm.h -- C ABI of libnfftm.so, an MI355X (gfx950) backend for the NFFT
 * plan of JuliaMath/NFFT.jl.
 *
 * The entry points are the operations of the AbstractNFFTs plan interface that
 * a Julia `ccall` (or Python ctypes) binding needs.  Each declaration names the
 * reference interface it replaces.  Conventions:
 *
 *  - Arrays follow Julia's column-major layout: an image of size
 *    N = (N_1, ..., N_D) is stored with dimension 1 fastest; the oversampled
 *    grid (size Ntilde) likewise.  Complex data are interleaved (re, im).
 *  - Node coordinates are a host D x M column-major double matrix (node j's
 *    D coordinates are contiguous), as `k::Matrix` in
 *    AbstractNFFTs/src/derived.jl:29-30.  Index arithmetic on nodes is FP64
 *    in every precision.
 *  - Data buffers (f, fHat, g) are DEVICE pointers owned by the caller, of
 *    element type complex<float> (NFFTM_F32) or complex<double> (NFFTM_F64).
 *  - Every call returns an nfftm_status; nfftm_last_error() gives the message
 *    of the last failure on the calling thread.
 *  - A plan is not re-entrant (it owns one grid buffer, like `p.tmpVec` in
 *    src/implementation.jl:16-43); use nfftm_plan_copy for concurrency.  All
 *    work is enqueued on the plan's HIP stream; the transform calls return
 *    without synchronising unless NFFTM_FLAG_TIMING is set.
 */
#ifndef NFFTM_H
#define NFFTM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nfftm_plan_s* nfftm_plan;

/* Status codes; the Julia/Python binding maps them to the exception the
 * reference raises for the same condition. */
typedef enum {
  NFFTM_OK = 0,
  NFFTM_ERR_ARG = 1,         /* ArgumentError   (src/utils.jl:46-55, src/convolution.jl:52) */
  NFFTM_ERR_DIM = 2,         /* DimensionMismatch (src/utils.jl:98-105, src/convolution.jl:26-29) */
  NFFTM_ERR_UNSUPPORTED = 3, /* error(...)  (src/windowFunctions.jl:16, src/precomputation.jl:401) */
  NFFTM_ERR_HIP = 4,         /* HIP runtime failure */
  NFFTM_ERR_FFT = 5,         /* rocFFT failure */
  NFFTM_ERR_OOM = 6          /* device allocation failed */
} nfftm_status;

/* Arithmetic of data, grid, spreading/interpolation and FFT.  Node index
 * reduction and the deconvolution factors stay FP64 in both modes. */
enum { NFFTM_F32 = 0, NFFTM_F64 = 1, NFFTM_MIXED = 2 /* alias of NFFTM_F32 */ };

/* Precompute flags, same values as AbstractNFFTs/src/misc.jl:13-18. */
enum { NFFTM_FULL = 1, NFFTM_TENSOR = 2, NFFTM_LINEAR = 3, NFFTM_POLYNOMIAL = 4 };

/* Window ids (src/windowFunctions.jl:4-19, getWindow): :kaiser_bessel,
 * :gauss, :spline, :kaiser_bessel_rev, :cosh_type (pairs psi, psi_hat at
 * :21-134).  The node gradient with precompute LINEAR or FULL is built for
 * Kaiser-Bessel only (the analytic window derivative); TENSOR and POLYNOMIAL
 * differentiate the fitted polynomial of any window. */
enum {
  NFFTM_KAISER_BESSEL = 0,
  NFFTM_GAUSS = 1,
  NFFTM_SPLINE = 2,
  NFFTM_KAISER_BESSEL_REV = 3,
  NFFTM_COSH_TYPE = 4
};

/* Plan flags. */
enum {
  NFFTM_FLAG_BLOCKING = 1, /* tiled kernels (NFFTParams.blocking, default true) */
  NFFTM_FLAG_TIMING = 2,   /* record per-stage device times (TimingStats) */
  NFFTM_FLAG_GRAPHS = 4,   /* replay repeated (input, output) transforms from HIP
                              graphs captured at the pair's second call */
  NFFTM_FLAG_ROCFFT = 8,   /* FFT stage as rocFFT over the whole oversampled grid
                              with separate deconvolution passes.  Default on
                              power-of-two grids (FP32: 16..2048 points per
                              dimension, FP64: 16..1024, D <= 3): pruned line
                              FFTs that skip the zero padding (forward) and
                              the discarded positions (adjoint), with the
                              deconvolution fused; on 1-D grids of n1 Lm^2
                              points (Lm = 64, 128, 256; n1 = 2^k or 5 2^k):
                              three passes of line FFTs, deconvolution
                              fused.  Results agree to rounding;
                              the grid (nfftm_grid) then holds the full FFT
                              after a forward transform and a partial one
                              after an adjoint. */
  NFFTM_FLAG_SEPARATE_FOLD = 16, /* keep the 3-D adjoint's fold and its first
                              pruned FFT pass as two kernels.  By default they
                              are one kernel where every tile holds at most
                              one work item and the tile shape matches the
                              pass (e.g. 256-point z lines): bit-identical
                              results, the folded grid is never written.
                              Also for tiled 1-D plans on the three-pass
                              FFT: their fold runs inside its first adjoint
                              pass. */
  NFFTM_FLAG_SPARSE = 32    /* convolution as a device sparse matrix, the
                              reference GPU backend's form of precompute =
                              FULL (ext/NFFTGPUArraysExt): at nodes! the plan
                              stores the W^D window weights (of its precompute
                              mode) and grid indices of every node, 4 + s_r
                              bytes each; forward = gather product, adjoint =
                              atomic scatter.  A comparison baseline: slower
                              than the default tiled kernels and
                              M W^D (4 + s_r) bytes of device memory. */
};

/* Plan creation: replaces NFFTPlan(k, N; m, sigma, window, precompute,
 * blocking) -- src/implementation.jl:73-106 with initParams
 * src/precomputation.jl:3-56 -- as reached from
 * plan_nfft(::Backend, ::Type{<:AbstractGPUArray}, k, N) at
 * ext/NFFTGPUArraysExt/implementation.jl:21-59.  m and sigma are the
 * resolved accuracy parameters (AbstractNFFTs/src/misc.jl:66-81 is applied
 * by the caller).  hip_stream may be NULL (default stream). */
int nfftm_plan_create(nfftm_plan* out, int prec, int D, const int64_t* N, int64_t M,
                      const double* k_host, int m, double sigma, int window, int precompute,
                      int flags, void* hip_stream);

/* Directional plan: NFFTPlan(k, N; dims = dim_lo+1:dim_hi+1, ...) --
 * src/implementation.jl:73-106 with initParams' dims handling
 * (src/precomputation.jl:3-56) and the stage operators of
 * src/directional.jl:1-156 -- which the reference GPU extension refuses
 * (ext/NFFTGPUArraysExt/implementation.jl:35-37).  N has D entries (Julia
 * order), dims = dim_lo .. dim_hi (0-based) a range of them, and k_host is
 * (dim_hi - dim_lo + 1) x M.  The transform runs over N[dims] for every
 * slice along the other axes, in place in the caller's arrays (no
 * permutation): f is prod(N) complex values, fHat is
 * (N[0..dim_lo-1], M, N[dim_hi+1..D-1]) (NOut), and the stage operators take
 * the grid (N[0..dim_lo-1], Ntilde[dims], N[dim_hi+1..]).  nfftm_sizes
 * returns Ntilde over all D axes (N[d] off dims).  Typed stage operators,
 * sdc, the node gradient, Toeplitz kernels and batches return
 * NFFTM_ERR_UNSUPPORTED for such a plan.  dims covering all axes gives an
 * ordinary plan. */
int nfftm_plan_create_dims(nfftm_plan* out, int prec, int D, const int64_t* N, int dim_lo, int dim_hi,
                           int64_t M, const double* k_host, int m, double sigma, int window, int precompute,
                           int flags, void* hip_stream);

/* nodes!(p, k): src/implementation.jl:108-141 (AbstractNFFTs/src/interface.jl:211).
 * M may differ from the plan's current M. */
int nfftm_nodes(nfftm_plan p, const double* k_host, int64_t M);

/* mul!(fHat, p, f): src/implementation.jl:155-172 (interface.jl:170).
 * f: prod(N) complex, fhat: M complex. */
int nfftm_forward(nfftm_plan p, const void* f_dev, void* fhat_dev);

/* mul!(f, adjoint(p), fHat): src/implementation.jl:176-193 (interface.jl:178). */
int nfftm_adjoint(nfftm_plan p, const void* fhat_dev, void* f_dev);

/* Node gradient of the forward transform fHat = A(k) f contracted with a
 * cotangent ybar:  kbar[d + D*j] = Re(conj(ybar_j) * d fHat_j / d k_{d,j}).
 * Replaces the NoTangent() that the reference's rrule returns for the nodes
 * (AbstractNFFTs/ext/AbstractNFFTsChainRulesCoreExt.jl:14-22, 26-45); the
 * pullback of y = A(k) f is then (kbar, adjoint(p) * ybar).  Computed as the
 * forward path (deconvolve!, forward FFT) followed by a gather with the
 * analytic derivative of the Kaiser-Bessel window (of the fitted polynomial
 * in POLYNOMIAL/TENSOR mode), window arithmetic in FP64.
 * f: prod(N) complex, ybar: M complex, kbar: D*M doubles laid out as k
 * (column-major D x M, caller's node order); all device pointers.  Uses the
 * plan's grid. */
int nfftm_grad_nodes(nfftm_plan p, const void* f_dev, const void* ybar_dev, double* kbar_dev);

/* The same node gradient from the transformed grid g = FFT(deconvolve(f)) of
 * the forward transform, as the plan's grid (nfftm_grid) holds it after
 * nfftm_forward: only the window-derivative gather runs.  No counterpart in
 * the reference; for pullbacks that keep their forward pass's grid (an
 * rrule's pullback, a loss evaluated and differentiated in one step).
 * g: prod(Ntilde) complex (may be the plan's own grid); ybar, kbar as in
 * nfftm_grad_nodes. */
int nfftm_grad_nodes_grid(nfftm_plan p, const void* g_dev, const void* ybar_dev, double* kbar_dev);

/* deconvolve!(p, f, g): src/deconvolution.jl:2-45, including the
 * fill!(tmpVec, 0) of src/implementation.jl:159.  g: prod(Ntilde) complex. */
int nfftm_deconvolve(nfftm_plan p, const void* f_dev, void* g_dev);

/* deconvolve_transpose!(p, g, f): src/deconvolution.jl:49-92. */
int nfftm_deconvolve_transpose(nfftm_plan p, const void* g_dev, void* f_dev);

/* convolve!(p, g, fHat): src/convolution.jl:20-45 (interpolation). */
int nfftm_convolve(nfftm_plan p, const void* g_dev, void* fhat_dev);

/* convolve_transpose!(p, fHat, g): src/convolution.jl:115-140 (spreading);
 * g is overwritten. */
int nfftm_convolve_transpose(nfftm_plan p, const void* fhat_dev, void* g_dev);

/* Element types of caller buffers for the typed stage operators below. */
enum { NFFTM_R32 = 0, NFFTM_R64 = 1, NFFTM_C32 = 2, NFFTM_C64 = 3, NFFTM_R16 = 4, NFFTM_C16 = 5 };

/* convolve!(p, g, fHat) and convolve_transpose!(p, fHat, g) on buffers of any
 * real or complex element type (half, single or double), as the reference's
 * methods take RealOrComplex arrays of any precision: src/convolution.jl:20-53
 * and :115-149, exercised by test/convolve.jl:43-146, and called with real
 * weights and grids by NFFTTools.sdc! (NFFTTools/src/samplingDensity.jl:106,
 * 112).  The sums run in the plan's precision; real inputs give real results
 * (the window is real).  Checks, in the reference's order: complex input with
 * a real output -> NFFTM_ERR_ARG (src/convolution.jl:48-53, :144-149); a
 * buffer length (elements, not bytes) other than prod(Ntilde) for g or M for
 * fHat -> NFFTM_ERR_DIM (:26-29, :121-124).  A buffer of another type than
 * the plan's complex type is staged through the plan's grid or a plan-owned
 * vector of M values (allocated at the first such call). */
int nfftm_convolve_typed(nfftm_plan p, const void* g_dev, int g_type, int64_t g_len, void* fhat_dev,
                         int fhat_type, int64_t fhat_len);
int nfftm_convolve_transpose_typed(nfftm_plan p, const void* fhat_dev, int fhat_type, int64_t fhat_len,
                                   void* g_dev, int g_type, int64_t g_len);

/* Oversampled FFT stage on a caller grid, in place, unnormalised:
 * p.forwardFFT * g (sign -1) or p.backwardFFT * g (sign +1),
 * src/implementation.jl:87-88, 161, 182.  direction: -1 or +1. */
int nfftm_fft(nfftm_plan p, void* g_dev, int direction);

/* Plan geometry: Ntilde (D entries), M, and the effective sigma
 * (src/precomputation.jl:25-29). */
int nfftm_sizes(nfftm_plan p, int64_t* Ntilde, int64_t* M, double* sigma);

/* The plan's internal grid (p.tmpVec) as a device pointer. */
int nfftm_grid(nfftm_plan p, void** g_dev);

/* TimingStats fields (AbstractNFFTs/src/misc.jl:22-30), seconds:
 * {pre, conv, fft, deconv, conv_adjoint, fft_adjoint, deconv_adjoint}. */
int nfftm_timing(nfftm_plan p, double t[7]);

/* Switch per-stage timing on (1) or off (0) for later transform calls; the
 * per-call `timing=` keyword of mul! (src/implementation.jl:155-193). */
int nfftm_set_timing(nfftm_plan p, int on);

/* Kernel-level profiling without host synchronisation: while on, every
 * stage launch of the transforms is bracketed by HIP events recorded on the
 * plan's stream.  nfftm_profile_read synchronises, returns per stage the
 * summed milliseconds and launch counts, and clears the record.  Stages:
 * 0 deconvolve, 1 forward FFT, 2 interpolation, 3 grid clear, 4 spreading,
 * 5 backward FFT, 6 deconvolve_transpose.  No counterpart in the reference
 * (its TimingStats is host-timed); used by the benchmark. */
int nfftm_profile(nfftm_plan p, int on);
int nfftm_profile_read(nfftm_plan p, double ms[8], int64_t count[8]);

/* The node binning of the last nodes! (or plan creation): the number of
 * work items; with non-NULL buffers also perm_host[s] (M entries: caller
 * index of the node at sorted position s) and items_host (4 per work item:
 * tile index, first and one-past-last sorted position, 0).  Tiles are
 * numbered with dimension 1 fastest; a tile's nodes keep the caller's order;
 * each work item holds at most 1024 (D = 3), 2048 (D = 2) or 4096 (D = 1)
 * nodes of one tile.  Plans on the node-parallel path have no items and the
 * identity permutation.  The reference's counterpart is the block layout of
 * _precomputeBlocks (src/precomputation.jl:459-520); exposed for tests and
 * tools.  Synchronises the plan's stream. */
int nfftm_binning(nfftm_plan p, int* nitems, int32_t* perm_host, int32_t* items_host);

/* Autotuner: times the plan's gridding kernels (spreading with its fold,
 * plus interpolation) on its own nodes for each candidate tile edge (D = 1:
 * 512..8192, D = 2: 16..128) and workgroup size (256, 512; D = 3: 512, 1024)
 * and keeps the fastest.  Synchronises; results are unchanged (same
 * arithmetic, different tiling).  No counterpart in the reference, whose
 * block size is fixed per dimension (src/precomputation.jl:59-77). */
int nfftm_plan_tune(nfftm_plan p);

/* Set the tile shape explicitly: D <= 2: tile edge (clipped to the grid);
 * D = 3: geometry variant (0: cubic tile within 160 KiB of LDS, one
 * workgroup per CU; 1: flat tile within 80 KiB, two per CU); threads 256,
 * 512 or 1024 (D = 3: 512 or 1024).  NFFTM_ERR_UNSUPPORTED (default shape
 * restored) when it does not fit. */
int nfftm_plan_set_tiles(nfftm_plan p, int tile, int threads);

/* Current tile edge per dimension (0 beyond D, or when the plan uses the
 * node-parallel kernels) and workgroup size of the gridding kernels. */
int nfftm_tile_shape(nfftm_plan p, int tile[3], int* threads);

/* Batches: many transforms per launch.  A batch takes plans of one precision
 * and D <= 2 (different N, M and nodes) and, per plan, ntrans[i] >= 1 images
 * that share that plan's nodes (NULL: one each).  Slot s = (plan i, image c)
 * in order of i, then c.  Every stage of all slots runs as one persistent
 * launch striding over the union of their work items (tiles of all slots in
 * one spreading / interpolation launch, one deconvolution launch, one fold),
 * the FFTs as one batched rocFFT per distinct grid size, and each direction
 * is captured once in a HIP graph and replayed.  Data pointers are passed per
 * call (f_dev[s], fhat_dev[s]: device buffers of slot s).  The plans must
 * outlive the batch; nfftm_nodes on a member plan invalidates it (its calls
 * then fail with NFFTM_ERR_ARG).  The reference transforms one image per
 * mul! (src/implementation.jl:155-193); a batch computes exactly the same
 * values as ntrans[i] such calls per plan.
 * flags: NFFTM_BATCH_AUTOTUNE (run nfftm_plan_tune on every plan first),
 * NFFTM_BATCH_NO_GRAPH (launch directly instead of replaying graphs). */
typedef struct nfftm_batch_s* nfftm_batch;
enum { NFFTM_BATCH_AUTOTUNE = 1, NFFTM_BATCH_NO_GRAPH = 2 };
int nfftm_batch_create(nfftm_batch* out, int nplans, const nfftm_plan* plans, const int* ntrans, int flags);
int nfftm_batch_info(nfftm_batch b, int* nslots, int* graph);
int nfftm_batch_forward(nfftm_batch b, const void* const* f_dev, void* const* fhat_dev, void* hip_stream);
int nfftm_batch_adjoint(nfftm_batch b, const void* const* fhat_dev, void* const* f_dev, void* hip_stream);
void nfftm_batch_destroy(nfftm_batch b);

/* NFFTTools on the plan.
 *
 * Sampling density compensation (Pipe & Menon): replaces
 * sdc!(p, iters, weights, weights_tmp, workg[, workf, workv]) of
 * NFFTTools/src/samplingDensity.jl:93-155 (reached from sdc(p; iters) at
 * :59-86).  weights_dev holds M real values of the plan's precision (the
 * initial weights, ones in sdc(p)) and receives the result.  Each iteration
 * spreads the weights (convolve_transpose!), divides the grid by its maximum
 * of the first iteration, interpolates (convolve!) and divides the weights by
 * the result.  scale = 1 adds the final least-squares scale c with
 * A' D(w) A 1 ~ 1 (the 7-argument sdc!); scale = 0 stops before it (the
 * 5-argument sdc!).  A non-positive interpolated weight fails with
 * NFFTM_ERR_UNSUPPORTED ("non-positive weights"), as the reference throws.
 * Synchronises the plan's stream. */
int nfftm_sdc(nfftm_plan p, int iters, void* weights_dev, int scale);

/* nfftm_sdc with no device allocation: no counterpart in the reference.  For
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and
 * errors as nfftm_sdc, but its work arrays (2 M complex values, plus P + M
 * with scale = 1, and 16.6 KB of scalars) live in a fixed 64 MiB device
 * workspace held by each GPU the library runs on until the process ends.
 * Sizes beyond it fail with NFFTM_ERR_ARG before any device work (FP32:
 * M <= ~4.1e6 without the scale).  Every call on one GPU shares that GPU's
 * workspace: calls that overlap on one GPU (other host threads, other
 * streams) race.  The plan's own grid and tables were allocated when it was
 * created. */
int nfftm_sdc_noalloc(nfftm_plan p, int iters, void* weights_dev, int scale);

/* Toeplitz embedding of A'A for an image of size N (D <= 3): an object
 * owning the FFTs of size 2N and one work array of that size.
 * nfftm_toeplitz_kernel: lambda = FFT(fftshift(adjoint(p) * ones(M))) for a
 * plan p over 2N, as calculateToeplitzKernel!(f, p, tr, fftplan)
 * (NFFTTools/src/Toeplitz.jl:119-125) after nodes!(p, tr).
 * nfftm_toeplitz_apply: y <- the N-corner of IFFT(FFT(pad(y)) .* lambda), as
 * convolveToeplitzKernel!(y, lambda, ...) (NFFTTools/src/Toeplitz.jl:236-250).
 * lambda_dev and y_dev are device buffers of complex values of the
 * precision (column-major, sizes 2N and N).  Work is enqueued on the
 * object's stream (nfftm_toeplitz_kernel: on the plan's stream). */
typedef struct nfftm_toeplitz_s* nfftm_toeplitz;
int nfftm_toeplitz_create(nfftm_toeplitz* out, int prec, int D, const int64_t* N, void* hip_stream);
int nfftm_toeplitz_kernel(nfftm_toeplitz t, nfftm_plan p, void* lambda_dev);
int nfftm_toeplitz_apply(nfftm_toeplitz t, const void* lambda_dev, void* y_dev);
void nfftm_toeplitz_destroy(nfftm_toeplitz t);

/* Change the HIP stream the plan enqueues on. */
int nfftm_set_stream(nfftm_plan p, void* hip_stream);

/* copy(p): src/implementation.jl:45-66.  The copy has its own grid. */
int nfftm_plan_copy(nfftm_plan src, nfftm_plan* dst);

/* finalizer (Wrappers/FINUFFT.jl:52-56 pattern). */
void nfftm_plan_destroy(nfftm_plan p);

/* Measurement utility, no counterpart in the reference: the issue cost of
 * the LDS instructions that bound the gridding taps, on the current device
 * (a short microbenchmark, about 10 ms).  cycles[0..3]: shader-clock cycles
 * per CU of one wavefront instruction of ds_read_b64 (interpolation taps),
 * ds_read_b128, ds_add_u64 (FP32 3-D spreading taps) and ds_add_f64 (FP64
 * spreading taps), conflict-free; clock_hz and cus: the clock and the CU
 * count they are counted against.  Allocates and synchronises the device. */
int nfftm_lds_issue_cycles(double cycles[4], double* clock_hz, int* cus);

/* Message of the last failure on this thread ("" if none). */
const char* nfftm_last_error(void);

/* Library version string. */
const char* nfftm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* NFFTM_H */
