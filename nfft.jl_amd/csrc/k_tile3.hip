// k_tile3.hip -- 3-D gridding kernels with compile-time tile geometry.
//
// Same algorithm as the generic tile kernels of kernels.hpp (padded tile in
// LDS, nodes bin-sorted into tiles, one workgroup per work item), specialised
// for D = 3 so that the LDS strides are constants (immediate ds offsets, no
// integer divisions), with 1024-thread workgroups for latency hiding and two
// nodes in flight per wavefront.
#include "fft_line.hpp"
#include "launch.hpp"

namespace nfftm {

// 3-D tile geometry per precision, stencil width W = 2m and variant V: tile
// T x T x TZ, padded P x P x PZ (P = T + W - 1).  V = 0: cubic, the largest
// edge whose padded tile plus a batch of node records fits 160 KiB of LDS
// (one workgroup per CU).  V = 1: the largest T x T x TZ (T, TZ in 16, 8, 4)
// within 80 KiB, so that two workgroups share a CU and one's staging and
// flush overlap the other's taps, at the price of more halo per tile.
template <typename R, int W, int V>
struct Geo3 {
  static constexpr int RB = NodeRec<R, 3, W>::BYTES;
  static constexpr __host__ __device__ int pad_of(int t) { return t + W - 1; }
  static constexpr __host__ __device__ int sx_of(int t) { return pad_of(t) % 16 == 7 ? pad_of(t) + 1 : pad_of(t); }
  static constexpr __host__ __device__ size_t tile_bytes(int t, int tz) {
    return ((2 * (size_t)sx_of(t) * pad_of(t) * pad_of(tz) * sizeof(R)) + 15) & ~(size_t)15;
  }
  static constexpr __host__ __device__ size_t batch_bytes(int nb) { return (size_t)nb * RB + (size_t)nb * 2 * sizeof(R); }
  static constexpr size_t CAP = V == 0 ? 163840 : 81920;
  // V = 1 choice, encoded t * 32 + tz
  static constexpr __host__ __device__ int flat_tile() {
    int best = 4 * 32 + 4, vol = 0;
    const int ts[3] = {16, 8, 4};
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) {
        const int t = ts[i], tz = ts[j];
        if (tile_bytes(t, tz) + batch_bytes(64) <= CAP && t * t * tz > vol) {
          vol = t * t * tz;
          best = t * 32 + tz;
        }
      }
    return best;
  }
  static constexpr int TC = sizeof(R) == 4 ? (W <= 10 ? 16 : W <= 12 ? 12 : W <= 14 ? 10 : 8)
                                           : (W <= 8 ? 12 : W <= 12 ? 8 : W <= 14 ? 6 : 4);
  static constexpr int T = V == 0 ? TC : flat_tile() / 32;
  static constexpr int TZ = V == 0 ? TC : flat_tile() % 32;
  static constexpr int P = T + W - 1;
  static constexpr int PZ = TZ + W - 1;
  static constexpr int SX = sx_of(T);  // W = 8: 8x8 lanes on distinct banks
  static constexpr int SXY = SX * P;
  static constexpr int CELLS = SXY * PZ;                // elements of one LDS plane
  static constexpr int PAD = P * P * PZ;                // dense padded cells
  static constexpr int SLOT = (PAD + 15) & ~15;         // scratch slot, 128-byte multiple
  static constexpr size_t TILE_BYTES = (2 * (size_t)CELLS * sizeof(R) + 15) & ~(size_t)15;
  // LDS: node records of one batch, their values (spreading), then the tile,
  // so that record fields sit at small immediate offsets
  static constexpr int NB = TILE_BYTES + batch_bytes(256) <= CAP ? 256 : TILE_BYTES + batch_bytes(128) <= CAP ? 128 : 64;
  static constexpr size_t REC_OFF = 0;
  static constexpr size_t FV_OFF = (size_t)NB * RB;
  static constexpr size_t TILE_OFF = FV_OFF + (size_t)NB * 2 * sizeof(R);
  static constexpr size_t LDS = TILE_OFF + TILE_BYTES;
  static constexpr bool VALID = LDS <= CAP;
  static_assert(TILE_OFF % 16 == 0, "tile offset");

};

__device__ __forceinline__ int wrap1(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

// Block layout of a scratch slot (grids of whole tiles, at least 3 per
// dimension, m <= T): per dimension the padded range splits into the left
// halo (m - 1 cells, belonging to the previous tile), the interior (T cells)
// and the right halo (m cells, belonging to the next tile); the 27 products
// are stored as contiguous blocks, the interior block first, each ordered
// x fastest, then z, then y (so that one y row of a block, as the fold fused
// into the first adjoint FFT pass reads it, is one contiguous run).
template <int W, int T, int TZ>
struct Blocks3 {
  static constexpr int m = W / 2;
  static constexpr __host__ __device__ int ext(int i, int t = T) { return i == 0 ? m - 1 : i == 1 ? t : m; }
  static constexpr __host__ __device__ int org(int i, int t = T) { return i == 0 ? 0 : i == 1 ? m - 1 : m - 1 + t; }
  static constexpr __host__ __device__ int size(int b) { return ext(b % 3) * ext((b / 3) % 3) * ext(b / 9, TZ); }
  // position of cell (x, y, z) in block b
  static constexpr __host__ __device__ int at(int b, int x, int y, int z) {
    return x + ext(b % 3) * (z + ext(b / 9, TZ) * y);
  }
  static constexpr __host__ __device__ int off(int b) {
    if (b == 13) return 0;
    int o = size(13);
    for (int q = 0; q < b; ++q)
      if (q != 13) o += size(q);
    return o;
  }
};

// Compile-time loop: f(std::integral_constant<int, I>) for I in [I0, N).
template <int I0, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I0 < N) {
    f(std::integral_constant<int, I0>{});
    static_for<I0 + 1, N>(f);
  }
}

// Sum over the 64 lanes; the result is valid in lane 63.
template <typename R>
__device__ __forceinline__ R wave_total(R v) {
  v += dpp<0xB1>(v);   // quad_perm(1,0,3,2)
  v += dpp<0x4E>(v);   // quad_perm(2,3,0,1)
  v += dpp<0x141>(v);  // row_half_mirror
  v += dpp<0x140>(v);  // row_mirror: every lane holds its row's sum
  v += dpp_rows<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_rows<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3
  return v;
}

// Rows of the padded tile each thread loads before writing them to LDS
// (measured on the 256^3 grid: FP32 gains up to 11 rows in flight, FP64 loses
// beyond 2).
template <typename R>
constexpr int STAGE_RB = sizeof(R) == 4 ? 11 : 2;

// Copy the padded tile at grid cell (x0, y0, z0) (periodic) into LDS, one
// padded row per 32 lanes (P <= 32): the x offsets are per lane, the row
// offsets per half-wave; each thread issues the loads of RB rows before
// writing them.  32-bit offsets unless the grid holds 2^32 or more elements.
template <typename G, int NT, int RB, typename C>
__device__ __forceinline__ void stage3(const C* __restrict__ grid, C* tile, int x0, int y0, int z0, int Nx, int Ny,
                                       int Nz, long long sy, long long sz, bool off32) {
  constexpr int P = G::P, ROWS = G::P * G::PZ, RPP = NT / 32;
  static_assert(P <= 32, "one padded row per 32 lanes");
  const int px = threadIdx.x & 31, r0 = threadIdx.x >> 5;
  const bool lx = px < P;
  const int gx = wrap1(x0 + (lx ? px : 0), Nx);
  for (int q0 = 0; q0 < ROWS; q0 += RB * RPP) {
    // plain scalars (not an array of vector structs) so that they stay in
    // registers across the predicated loads
    decltype(grid[0].x) vr[RB], vi[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int r = q0 + b * RPP + r0;
      const bool ok = lx && r < ROWS;
      const int rr = ok ? r : 0;
      const int py = rr % P, pz = rr / P;
      const int gy = wrap1(y0 + py, Ny), gz = wrap1(z0 + pz, Nz);
      const size_t g = off32 ? (size_t)((unsigned)gz * (unsigned)sz + (unsigned)gy * (unsigned)sy + (unsigned)gx)
                             : (size_t)(gz * sz + gy * sy + gx);
      C v{0, 0};
      if (ok) v = grid[g];
      vr[b] = v.x;
      vi[b] = v.y;
    }
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int r = q0 + b * RPP + r0;
      if (lx && r < ROWS) tile[px + G::SX * (r % P) + G::SXY * (r / P)] = C{vr[b], vi[b]};
    }
  }
}

// Interpolation: the padded tile is staged once; each batch of node records
// is copied to LDS; 16 lanes per node, 4 nodes per wavefront, each lane
// walking ceil(W^2 / 16) (x, y) columns of the footprint along z, and a
// 16-lane DPP reduction per node.
// The records of one batch of nodes held in registers between their global
// load and their LDS store, so that the load of the next batch overlaps the
// taps of this one.
template <typename G, int NT>
struct RecStage {
  static constexpr int N16 = G::NB * G::RB / 16;
  static constexpr int PER = (N16 + NT - 1) / NT;
  static_assert(G::RB % 16 == 0, "records are 16-byte blocks");
  // scalar arrays (not an array of vector structs): kept in registers
  unsigned v0[PER], v1[PER], v2[PER], v3[PER];
  __device__ __forceinline__ void load(const KArgs& a, int b0, int nb) {
    const uint4* src = (const uint4*)(a.rec + (size_t)b0 * G::RB);
    const int n16 = nb * (G::RB / 16);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = (int)threadIdx.x + u * NT;
      uint4 t = make_uint4(0, 0, 0, 0);
      if (i < n16) t = src[i];
      v0[u] = t.x;
      v1[u] = t.y;
      v2[u] = t.z;
      v3[u] = t.w;
    }
  }
  __device__ __forceinline__ void store(char* lrec, int nb) const {
    uint4* dst = (uint4*)lrec;
    const int n16 = nb * (G::RB / 16);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = (int)threadIdx.x + u * NT;
      if (i < n16) dst[i] = make_uint4(v0[u], v1[u], v2[u], v3[u]);
    }
  }
};

// V = 1 geometries hold two workgroups per CU: 2 NT / 256 waves per SIMD
// (1024 threads: 8 waves, at most 64 VGPRs) where that needs no spills.
template <typename R, int W, int NT, int V>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(V == 1 && W <= (sizeof(R) == 4 ? 12 : 8) ? NT / 128 : 1)))
void k_interp3(KArgs a, const typename cpx<R>::T* __restrict__ grid,
                                                typename cpx<R>::T* __restrict__ fhat) {
  using C = typename cpx<R>::T;
  using G = Geo3<R, W, V>;
  constexpr int NW = NT / 64;
  constexpr int NCOL = (W * W + 15) / 16;
  extern __shared__ __align__(16) char smem[];
  const int item = xcd_item(a.nitems);
  if (item >= a.nitems) return;
  const int4 it = a.items[item];
  C* tile = (C*)(smem + G::TILE_OFF);
  char* lrec = smem + G::REC_OFF;
  int t0[3];
  {
    const int tx = it.x % a.ntiles[0], q = it.x / a.ntiles[0];
    t0[0] = tx * G::T;
    t0[1] = (q % a.ntiles[1]) * G::T;
    t0[2] = (q / a.ntiles[1]) * G::TZ;
  }
  const int Nx = a.Nt[0], Ny = a.Nt[1], Nz = a.Nt[2];
  const long long sy = a.gstr[1], sz = a.gstr[2];

  // the first batch's records are loaded while the padded tile is staged,
  // each later batch's during the taps of the one before
  RecStage<G, NT> rs;
  rs.load(a, it.y, min(G::NB, it.z - it.y));
  stage3<G, NT, STAGE_RB<R>>(grid, tile, t0[0] - a.m + 1, t0[1] - a.m + 1, t0[2] - a.m + 1, Nx, Ny, Nz, sy, sz,
                             (unsigned long long)a.gstr[2] * (unsigned long long)Nz <= 0xFFFFFFFFull);

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int quarter = lane >> 4, r16 = lane & 15;
  const R osc = (R)a.oscale;
  for (int b0 = it.y; b0 < it.z; b0 += G::NB) {
    const int nb = min(G::NB, it.z - b0);
    __syncthreads();
    rs.store(lrec, nb);
    __syncthreads();
    if (b0 + G::NB < it.z) rs.load(a, b0 + G::NB, min(G::NB, it.z - b0 - G::NB));
    for (int n0 = 4 * wave; n0 < nb; n0 += 4 * NW) {
      const int n = n0 + quarter;
      const bool live = n < nb;
      const int nn = live ? n : n0;
      const R* w = rec_w<R, 3, W>(lrec, nn);
      R w3[W];
#pragma unroll
      for (int l = 0; l < W; ++l) w3[l] = w[2 * W + l];
      const C* base0 = tile + rec_loc<R, 3, W>(lrec, nn);
      R ar = 0, ai = 0;
#pragma unroll
      for (int c = 0; c < NCOL; ++c) {
        const int t = r16 + 16 * c;
        if (t < W * W) {
          const int ly = t / W, lx = t - ly * W;
          const C* col = base0 + lx + G::SX * ly;
          R sr = 0, si = 0;
#pragma unroll
          for (int lz = 0; lz < W; ++lz) {
            const C v = col[G::SXY * lz];
            sr = fma(w3[lz], v.x, sr);
            si = fma(w3[lz], v.y, si);
          }
          const R wxy = w[lx] * w[W + ly];
          ar = fma(wxy, sr, ar);
          ai = fma(wxy, si, ai);
        }
      }
      ar = group_sum<16>(ar);
      ai = group_sum<16>(ai);
      if (live && r16 == 0) fhat[(long long)rec_pj<R, 3, W>(lrec, n) * a.fst] = C{ar * osc, ai * osc};
    }
  }
}

// Spreading: lanes over the (x, y) columns of a node's footprint, walking z,
// one LDS atomic per tap (FP32: packed 32+32-bit fixed point with a per-item
// scale from a strict bound; FP64: two FP64 planes); the padded tile goes to
// the item's scratch slot.
template <typename R, int W, int NT, int V>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(V == 1 && W <= (sizeof(R) == 4 ? 12 : 8) ? NT / 128 : 1)))
void k_spread3(KArgs a, const typename cpx<R>::T* __restrict__ fhat) {
  using C = typename cpx<R>::T;
  using G = Geo3<R, W, V>;
  constexpr int NW = NT / 64;
  extern __shared__ __align__(16) char smem[];
  const int item = xcd_item(a.nitems);
  if (item >= a.nitems) return;
  const int4 it = a.items[item];
  char* tile = smem + G::TILE_OFF;
  char* lrec = smem + G::REC_OFF;
  C* fv = (C*)(smem + G::FV_OFF);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;

  // the first batch's records; the item's node values, held in registers
  // (thread t holds nodes it.y + t + h NT)
  RecStage<G, NT> rs;
  rs.load(a, it.y, min(G::NB, it.z - it.y));
  constexpr int HPT = ITEM3 / NT;
  static_assert(ITEM3 % NT == 0, "item size");
  R hre[HPT], him[HPT];
#pragma unroll
  for (int h = 0; h < HPT; ++h) {
    const int j = it.y + (int)threadIdx.x + h * NT;
    C v{0, 0};
    if (j < it.z) v = fhat[(long long)a.perm[j] * a.fst];
    hre[h] = v.x;
    him[h] = v.y;
  }
  // clear the tile while the gathers are in flight
  for (int i = threadIdx.x; i < (int)(G::TILE_BYTES / 16); i += NT) ((uint4*)tile)[i] = make_uint4(0, 0, 0, 0);

  // FP32: fixed-point scale s = 2^(30 - e), B = sum_j |f_j|_inf * wmaxD < 2^e
  R fscale = 1, inv = 1;
  if (sizeof(R) == 4) {
    R part = 0;
#pragma unroll
    for (int h = 0; h < HPT; ++h) part += fmax(fabs(hre[h]), fabs(him[h]));
    part = wave_total(part);
    R* red = (R*)lrec;
    if (lane == 63) red[wave] = part;
    __syncthreads();
    double B = 0;
    for (int w = 0; w < NW; ++w) B += (double)red[w];
    B *= a.oscale * a.wmaxD * 1.0000001;
    int e = 0;
    if (B > 0) (void)frexp(B, &e);
    // both powers stay normal floats: below 2^-96 the quantum stops at 2^-126
    // (values under ~1e-38 round to 0); a non-finite bound (Inf or NaN data)
    // spreads zeros and turns the whole tile into NaN at the flush
    e = max(e, -96);
    fscale = (R)ldexp(1.0, 30 - e);
    inv = (R)ldexp(1.0, e - 30);
    if (!(B <= 1.0e300)) {
      fscale = 0;
      inv = (R)NAN;
    }
  }

  const R sc = fscale * (R)a.oscale;
  for (int b0 = it.y; b0 < it.z; b0 += G::NB) {
    const int nb = min(G::NB, it.z - b0);
    __syncthreads();
    rs.store(lrec, nb);
#pragma unroll
    for (int h = 0; h < HPT; ++h) {
      const int n = it.y + (int)threadIdx.x + h * NT - b0;
      if (n >= 0 && n < nb) fv[n] = C{hre[h] * sc, him[h] * sc};
    }
    __syncthreads();
    if (b0 + G::NB < it.z) rs.load(a, b0 + G::NB, min(G::NB, it.z - b0 - G::NB));
    // z weights and cell of the wavefront's node from the record in global
    // memory (uniform: scalar loads, one node ahead), x and y weights per
    // lane from the LDS copy
    using CR = const __attribute__((address_space(4))) R;
    using CU = const __attribute__((address_space(4))) unsigned;
    const char* gr = a.rec + (size_t)b0 * G::RB;
    R w3n[W];
    unsigned celln = 0;
    if (wave < nb) {
      CR* g = (CR*)(gr + (size_t)wave * G::RB);
#pragma unroll
      for (int l = 0; l < W; ++l) w3n[l] = g[2 * W + l];
      celln = ((CU*)(g + 3 * W))[0];
    }
    for (int n = wave; n < nb; n += NW) {
      const C f = fv[n];
      const R* w = rec_w<R, 3, W>(lrec, n);
      R w3[W];
#pragma unroll
      for (int l = 0; l < W; ++l) w3[l] = w3n[l];
      const int cell = (int)celln;
      if (n + NW < nb) {
        CR* g = (CR*)(gr + (size_t)(n + NW) * G::RB);
#pragma unroll
        for (int l = 0; l < W; ++l) w3n[l] = g[2 * W + l];
        celln = ((CU*)(g + 3 * W))[0];
      }
#pragma unroll
      for (int t0l = 0; t0l < W * W; t0l += 64) {
        const int t = t0l + lane;
        if (t < W * W) {
          const int ly = t / W, lx = t - ly * W;
          const R wxy = w[lx] * w[W + ly];
          const R vr = f.x * wxy, vi = f.y * wxy;
          const int base = cell + lx + G::SX * ly;
          if constexpr (sizeof(R) == 4) {
            unsigned long long* tl = (unsigned long long*)tile + base;
#pragma unroll
            for (int lz = 0; lz < W; ++lz) {
              int re, im;
              asm volatile("v_cvt_rpi_i32_f32 %0, %1" : "=v"(re) : "v"((float)(vr * w3[lz])));
              asm volatile("v_cvt_rpi_i32_f32 %0, %1" : "=v"(im) : "v"((float)(vi * w3[lz])));
              const unsigned long long term =
                  ((unsigned long long)(unsigned)(im + (re >> 31)) << 32) | (unsigned)re;
              __hip_atomic_fetch_add(tl + G::SXY * lz, term, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
          } else {
            double* tr = (double*)tile + base;
#pragma unroll
            for (int lz = 0; lz < W; ++lz) {
              __hip_atomic_fetch_add(tr + G::SXY * lz, (double)(vr * w3[lz]), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
              __hip_atomic_fetch_add(tr + G::CELLS + G::SXY * lz, (double)(vi * w3[lz]), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
            }
          }
        }
      }
    }
  }
  __syncthreads();

  // padded tile -> this item's scratch slot
  C* out = (C*)a.scratch + (size_t)item * G::SLOT;
  if (a.blocked3) {
    using B = Blocks3<W, G::T, G::TZ>;
    static_for<0, 27>([&](auto bc) {
      constexpr int b = decltype(bc)::value;
      constexpr int ex = B::ext(b % 3), ey = B::ext((b / 3) % 3), ez = B::ext(b / 9, G::TZ);
      constexpr int ox = B::org(b % 3), oy = B::org((b / 3) % 3), oz = B::org(b / 9, G::TZ);
      constexpr int off = B::off(b), n = ex * ey * ez;
      C* dst = out + off;
      if constexpr (n > 0) {
        for (int i = threadIdx.x; i < n; i += NT) {
          const int lx = i % ex, q = i / ex;
          const int lz = q % ez, ly = q / ez;
          dst[i] = SpreadCell<R>::get(tile, G::CELLS, (ox + lx) + G::SX * (oy + ly) + G::SXY * (oz + lz), inv);
        }
      }
    });
    return;
  }
  const int px = threadIdx.x & 31;
  for (int r = threadIdx.x >> 5; r < G::P * G::PZ; r += NT / 32) {
    if (px < G::P) {
      const int li = px + G::SX * (r % G::P) + G::SXY * (r / G::P);
      out[px + G::P * r] = SpreadCell<R>::get(tile, G::CELLS, li, inv);
    }
  }
}

// Fold for D = 3, one workgroup per output tile O (origin O_d, extent
// E_d <= T): thread (x, y) owns the column O + (x, y, 0..E_z-1) in registers.
// Tile t covers the arc [tT - m + 1, tT - m + 1 + P) of the periodic axis; its
// intersection with [O_d, O_d + E_d) gives up to two (t, lo, hi, off) pieces
// per candidate t in {o-2 .. o+2} (deduplicated), with padded offset
// p = x - off.  For every combination of pieces, every work item of that
// tile adds its scratch slot over the intersection box.
struct Piece {
  int t, lo, hi, off;  // cells [lo, hi) of the axis, padded offset x - off
};

template <int T, int P>
__device__ __forceinline__ int pieces(int o, int Nt, int nt, int m1, Piece* pc) {
  const int O = o * T, E = min(T, Nt - O);
  int n = 0;
  int seen[5];
  int ns = 0;
  for (int dt = -2; dt <= 2; ++dt) {
    int t = o + dt;
    t = ((t % nt) + nt) % nt;
    bool dup = false;
    for (int q = 0; q < ns; ++q) dup |= seen[q] == t;
    if (dup) continue;
    seen[ns++] = t;
    int start = t * T - m1;
    start = ((start % Nt) + Nt) % Nt;
    const int a0[2] = {start, 0};
    const int b0[2] = {min(start + P, Nt), start + P > Nt ? start + P - Nt : 0};
    const int offs[2] = {start, start - Nt};
    for (int k = 0; k < 2; ++k) {
      const int lo = max(a0[k], O), hi = min(b0[k], O + E);
      if (lo < hi && n < 10) pc[n++] = Piece{t, lo, hi, offs[k]};
    }
  }
  return n;
}

template <typename R, int W, int V>
__global__ __launch_bounds__(256) void k_fold3(KArgs a, typename cpx<R>::T* __restrict__ grid) {
  using C = typename cpx<R>::T;
  using G = Geo3<R, W, V>;
  constexpr int T = G::T, P = G::P, TZ = G::TZ, PZ = G::PZ;
  const C* scr = (const C*)a.scratch;
  const int m1 = a.m - 1;
  const int tile = blockIdx.x;
  const int ox = tile % a.ntiles[0], oq = tile / a.ntiles[0];
  const int oy = oq % a.ntiles[1], oz = oq / a.ntiles[1];
  const int Nx = a.Nt[0], Ny = a.Nt[1], Nz = a.Nt[2];
  __shared__ Piece px[10], py[10], pz[10];
  __shared__ int np[3];
  if (threadIdx.x == 0) {
    np[0] = pieces<T, P>(ox, Nx, a.ntiles[0], m1, px);
    np[1] = pieces<T, P>(oy, Ny, a.ntiles[1], m1, py);
    np[2] = pieces<TZ, PZ>(oz, Nz, a.ntiles[2], m1, pz);
  }
  __syncthreads();
  const int npx = np[0], npy = np[1], npz = np[2];
  const int lx = threadIdx.x % T, ly = threadIdx.x / T;
  const int x = ox * T + lx, y = oy * T + ly, z0 = oz * TZ;
  const bool mine = threadIdx.x < T * T && x < Nx && y < Ny;
  C acc[TZ];
#pragma unroll
  for (int k = 0; k < TZ; ++k) acc[k] = C{0, 0};
  for (int iz = 0; iz < npz; ++iz) {
    const Piece qz = pz[iz];
    for (int iy = 0; iy < npy; ++iy) {
      const Piece qy = py[iy];
      const bool iny = y >= qy.lo && y < qy.hi;
      for (int ix = 0; ix < npx; ++ix) {
        const Piece qx = px[ix];
        const bool in = mine && iny && x >= qx.lo && x < qx.hi;
        const int t = qx.t + a.ntiles[0] * (qy.t + a.ntiles[1] * qz.t);
        const int i0 = a.tstart[t], i1 = a.tstart[t + 1];
        for (int itm = i0; itm < i1; ++itm) {
          const C* s = scr + (size_t)itm * G::SLOT + (x - qx.off) + P * ((y - qy.off) + P * (-qz.off));
#pragma unroll
          for (int k = 0; k < TZ; ++k) {
            const int z = z0 + k;
            if (z >= qz.lo && z < qz.hi && in) {
              const C v = s[P * P * z];
              acc[k].x += v.x;
              acc[k].y += v.y;
            }
          }
        }
      }
    }
  }
  if (mine) {
    const int E = min(TZ, Nz - z0);
#pragma unroll
    for (int k = 0; k < TZ; ++k)
      if (k < E) grid[((long long)(z0 + k) * Ny + y) * Nx + x] = acc[k];
  }
}

// Fold for block-layout slots: output tile o gathers its own tile's interior
// block and, from each of its 26 neighbours o + d, the halo block facing it
// (d = -1: the neighbour's right halo over local [0, m); d = +1: its left halo
// over [T - m + 1, T)), for every work item of those tiles.  Thread (x, y)
// owns a column of T cells in registers.
template <typename R, int W, int V>
__global__ __launch_bounds__(256) void k_fold3b(KArgs a, typename cpx<R>::T* __restrict__ grid) {
  using C = typename cpx<R>::T;
  using G = Geo3<R, W, V>;
  using B = Blocks3<W, G::T, G::TZ>;
  constexpr int T = G::T, TZ = G::TZ, m = W / 2;
  const C* scr = (const C*)a.scratch;
  const int tile = blockIdx.x;
  const int ntx = a.ntiles[0], nty = a.ntiles[1], ntz = a.ntiles[2];
  const int ox = tile % ntx, oq = tile / ntx;
  const int oy = oq % nty, oz = oq / nty;
  const int lx = threadIdx.x % T, ly = threadIdx.x / T;
  const bool mine = threadIdx.x < T * T;
  C acc[TZ];
#pragma unroll
  for (int k = 0; k < TZ; ++k) acc[k] = C{0, 0};
  static_for<0, 27>([&](auto dc) {
    constexpr int d = decltype(dc)::value;
    constexpr int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
    constexpr int bx = 1 - dx, by = 1 - dy, bz = 1 - dz;
    constexpr int sx = dx == 1 ? T - m + 1 : 0, sy = dy == 1 ? T - m + 1 : 0, sz = dz == 1 ? TZ - m + 1 : 0;
    constexpr int zlo = dz == 1 ? TZ - m + 1 : 0, zhi = dz == -1 ? m : TZ;
    constexpr int ex = B::ext(bx);
    constexpr int off = B::off(bx + 3 * by + 9 * bz);
    const bool in = mine && (dy == 0 || (dy == -1 ? ly < m : ly >= T - m + 1)) &&
                    (dx == 0 || (dx == -1 ? lx < m : lx >= T - m + 1));
    const int tx = (ox + dx + ntx) % ntx, ty = (oy + dy + nty) % nty, tz = (oz + dz + ntz) % ntz;
    const int t = tx + ntx * (ty + nty * tz);
    const int i0 = a.tstart[t], i1 = a.tstart[t + 1];
    for (int itm = i0; itm < i1; ++itm) {
      const C* s = scr + (size_t)itm * G::SLOT + off + B::at(bx + 3 * by + 9 * bz, lx - sx, ly - sy, 0);
#pragma unroll
      for (int k = zlo; k < zhi; ++k) {
        if (in) {
          const C v = s[ex * (k - sz)];
          acc[k].x += v.x;
          acc[k].y += v.y;
        }
      }
    }
  });
  if (mine) {
    const int x = ox * T + lx, y = oy * T + ly, z0 = oz * TZ;
    const int Nx = a.Nt[0], Ny = a.Nt[1];
#pragma unroll
    for (int k = 0; k < TZ; ++k) grid[((long long)(z0 + k) * Ny + y) * Nx + x] = acc[k];
  }
}

// Adjoint with the fold fused into the first pruned FFT pass (lines along
// dimension 3, inverse FFT, output kept on S_z), for plans whose tiles hold
// at most one work item each.  Workgroup (ox, y) takes the T lines x of tile
// column ox at row y (E / L = T lines of L points); thread (lx, j0) sums its
// 16 points z = j0 + u L / 16 from the slots exactly as k_fold3b sums them
// (same neighbours in the same order, so the grid values are bit-identical),
// transforms them with the thread mapping of k_fft_col and stores the S_z
// positions.  A wavefront's lanes share the tile row of each point
// ((64 / T) | TZ), so tile and item lookups are scalar.
template <typename R, int W, int V, int L>
struct FoldFft3 {
  using G = Geo3<R, W, V>;
  using LG = lf::LineGeo<R, L>;
  static constexpr bool OK = G::VALID && LG::LPB == G::T && G::T <= 64 && 64 % G::T == 0 &&
                             LG::TL % G::TZ == 0 && G::TZ % (64 / G::T) == 0;
};

template <typename R, int W, int V, int L>
__global__ __launch_bounds__((lf::LineGeo<R, L>::NT)) void k_fold_fft3(KArgs a, FArgs f,
                                                                      typename cpx<R>::T* __restrict__ grid) {
  using C = typename cpx<R>::T;
  using G = Geo3<R, W, V>;
  using B = Blocks3<W, G::T, G::TZ>;
  using LG = lf::LineGeo<R, L>;
  constexpr int T = G::T, TZ = G::TZ, m = W / 2, TL = LG::TL;
  constexpr int NZ = L / TZ;  // tile rows along z
  __shared__ __align__(16) R sm[2 * (LG::LPB * LG::LS + L)];
  __shared__ int item[9][NZ];  // work item of tile (ox + dx, oy + dy, z row), -1 if none
  C* buf = (C*)sm;
  C* tw = buf + LG::LPB * LG::LS;
  lf::load_tw<LG::NT>(tw, f.tw[2], L);
  const C* scr = (const C*)a.scratch;
  const int ntx = a.ntiles[0], nty = a.ntiles[1];
  // workgroups of consecutive (ox, y) on one XCD, so that the halo lines
  // neighbouring workgroups share stay in that XCD's L2 (gridDim.x =
  // ntiles[0] L[1], a multiple of 8)
  const int wg = xcd_item((int)gridDim.x);
  const int ox = wg % ntx, y = wg / ntx, oy = y / T, ly = y % T;
  for (int i = threadIdx.x; i < 9 * NZ; i += LG::NT) {
    const int dxy = i / NZ, tz = i % NZ;
    const int tx = (ox + dxy % 3 - 1 + ntx) % ntx, ty = (oy + dxy / 3 - 1 + nty) % nty;
    const int t = tx + ntx * (ty + nty * tz);
    const int i0 = a.tstart[t];
    item[dxy][tz] = a.tstart[t + 1] > i0 ? i0 : -1;
  }
  __syncthreads();
  const int lx = threadIdx.x % T, j0 = threadIdx.x / T, lz = j0 % TZ;
  // point u lies in tile row q + u TL / TZ
  const int q = __builtin_amdgcn_readfirstlane(j0 / TZ);
  C v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) v[u] = C{0, 0};
  static_for<0, 27>([&](auto dc) {
    constexpr int d = decltype(dc)::value;
    constexpr int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
    constexpr int bx = 1 - dx, by = 1 - dy, bz = 1 - dz;
    constexpr int sx = dx == 1 ? T - m + 1 : 0, sy = dy == 1 ? T - m + 1 : 0, sz = dz == 1 ? TZ - m + 1 : 0;
    constexpr int zlo = dz == 1 ? TZ - m + 1 : 0, zhi = dz == -1 ? m : TZ;
    constexpr int off = B::off(bx + 3 * by + 9 * bz);
    if (dy == 0 || (dy == -1 ? ly < m : ly >= T - m + 1)) {
      const bool use = (dx == 0 || (dx == -1 ? lx < m : lx >= T - m + 1)) && lz >= zlo && lz < zhi;
      if (use) {
        const int* row = item[(dy + 1) * 3 + dx + 1];
        const unsigned li = off + B::at(bx + 3 * by + 9 * bz, lx - sx, ly - sy, lz - sz);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int it = row[(q + u * (TL / TZ) + dz + NZ) % NZ];
          const C w = (scr + (size_t)(it >= 0 ? it : 0) * G::SLOT)[li];
          v[u].x += it >= 0 ? w.x : (R)0;
          v[u].y += it >= 0 ? w.y : (R)0;
        }
      }
    }
  });
  lf::line_fft<R, L, true>(v, buf + lx * LG::LS, j0, buf + (threadIdx.x / TL) * LG::LS, threadIdx.x % TL, tw);
  const size_t base = (size_t)y * f.gstr[1] + (size_t)ox * T;
  const long long st = f.gstr[2];
  const int Nd = f.N[2], Nad = f.Na[2];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int k = j0 + u * TL;
    if (lf::image_idx(k, Nd, L, Nad) >= 0) grid[base + lx + (size_t)k * st] = v[u];
  }
}

template <typename R, int W, int V, int L>
static hipError_t fold_fft3(const KArgs& a, const FArgs& f, hipStream_t s, void* grid, bool dry) {
  if constexpr (!FoldFft3<R, W, V, L>::OK) {
    return hipErrorNotSupported;
  } else {
    if (dry) return hipSuccess;
    hipLaunchKernelGGL((k_fold_fft3<R, W, V, L>), dim3((unsigned)(a.ntiles[0] * f.L[1])),
                       dim3(lf::LineGeo<R, L>::NT), 0, s, a, f, (typename cpx<R>::T*)grid);
    return hipGetLastError();
  }
}

template <typename R, int W, int V>
static hipError_t fold_fft3_l(const KArgs& a, const FArgs& f, hipStream_t s, void* grid, bool dry) {
  switch (f.L[2]) {
    case 128: return fold_fft3<R, W, V, 128>(a, f, s, grid, dry);
    case 256: return fold_fft3<R, W, V, 256>(a, f, s, grid, dry);
    case 512: return fold_fft3<R, W, V, 512>(a, f, s, grid, dry);
    default: return hipErrorNotSupported;
  }
}

template <typename R, int W>
static hipError_t fold_fft3_v(const KArgs& a, const FArgs& f, hipStream_t s, void* grid, bool dry) {
  return a.tvar ? fold_fft3_l<R, W, 1>(a, f, s, grid, dry) : fold_fft3_l<R, W, 0>(a, f, s, grid, dry);
}

hipError_t launch_fold_fft3(bool dbl, int W, const KArgs& a, const FArgs& f, hipStream_t s, void* grid, bool dry) {
  if (!a.blocked3 || f.D != 3) return hipErrorNotSupported;
#define NFFTM_FF(WW) \
  case WW:           \
    return dbl ? fold_fft3_v<double, WW>(a, f, s, grid, dry) : fold_fft3_v<float, WW>(a, f, s, grid, dry);
  switch (W) {
    NFFTM_FF(2) NFFTM_FF(4) NFFTM_FF(6) NFFTM_FF(8) NFFTM_FF(10) NFFTM_FF(12) NFFTM_FF(14) NFFTM_FF(16)
    default:
      return hipErrorNotSupported;
  }
#undef NFFTM_FF
}

template <typename R, int W, int V>
static hipError_t go3(bool spread, int NT, const KArgs& a, hipStream_t s, const void* in, void* out) {
  using C = typename cpx<R>::T;
  using G = Geo3<R, W, V>;
  if constexpr (!G::VALID) {
    return hipErrorInvalidConfiguration;
  } else {
    const int nblk = 8 * ((a.nitems + 7) / 8);
    const int lds = (int)G::LDS;
    if (spread) {
      auto k = NT == 1024 ? k_spread3<R, W, 1024, V> : k_spread3<R, W, 512, V>;
      if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess)
        return hipErrorInvalidConfiguration;
      hipLaunchKernelGGL(k, dim3(nblk), dim3(NT), lds, s, a, (const C*)in);
      (void)out;
    } else {
      auto k = NT == 1024 ? k_interp3<R, W, 1024, V> : k_interp3<R, W, 512, V>;
      if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess)
        return hipErrorInvalidConfiguration;
      hipLaunchKernelGGL(k, dim3(nblk), dim3(NT), lds, s, a, (const C*)in, (C*)out);
    }
    return hipGetLastError();
  }
}

template <typename R, int W>
static hipError_t go3v(bool spread, int NT, const KArgs& a, hipStream_t s, const void* in, void* out) {
  return a.tvar ? go3<R, W, 1>(spread, NT, a, s, in, out) : go3<R, W, 0>(spread, NT, a, s, in, out);
}

template <typename R>
static hipError_t go3w(bool spread, int W, int NT, const KArgs& a, hipStream_t s, const void* in, void* out) {
  switch (W) {
    case 2: return go3v<R, 2>(spread, NT, a, s, in, out);
    case 4: return go3v<R, 4>(spread, NT, a, s, in, out);
    case 6: return go3v<R, 6>(spread, NT, a, s, in, out);
    case 8: return go3v<R, 8>(spread, NT, a, s, in, out);
    case 10: return go3v<R, 10>(spread, NT, a, s, in, out);
    case 12: return go3v<R, 12>(spread, NT, a, s, in, out);
    case 14: return go3v<R, 14>(spread, NT, a, s, in, out);
    case 16: return go3v<R, 16>(spread, NT, a, s, in, out);
    default: return hipErrorInvalidValue;
  }
}

template <typename R, int W, int V>
static hipError_t fold3(const KArgs& a, hipStream_t s, void* grid) {
  using C = typename cpx<R>::T;
  using G = Geo3<R, W, V>;
  if constexpr (!G::VALID) {
    return hipErrorInvalidConfiguration;
  } else {
    static_assert(G::T * G::T <= 256, "fold3 maps one thread per tile column");
    if (a.blocked3)
      hipLaunchKernelGGL((k_fold3b<R, W, V>), dim3((unsigned)a.ntile_all), dim3(256), 0, s, a, (C*)grid);
    else
      hipLaunchKernelGGL((k_fold3<R, W, V>), dim3((unsigned)a.ntile_all), dim3(256), 0, s, a, (C*)grid);
    return hipGetLastError();
  }
}

hipError_t launch_fold_d3(bool dbl, int W, const KArgs& a, hipStream_t s, void* grid) {
#define NFFTM_F3(WW)                                                                          \
  case WW:                                                                                    \
    if (a.tvar) return dbl ? fold3<double, WW, 1>(a, s, grid) : fold3<float, WW, 1>(a, s, grid); \
    return dbl ? fold3<double, WW, 0>(a, s, grid) : fold3<float, WW, 0>(a, s, grid);
  switch (W) {
    NFFTM_F3(2) NFFTM_F3(4) NFFTM_F3(6) NFFTM_F3(8) NFFTM_F3(10) NFFTM_F3(12) NFFTM_F3(14) NFFTM_F3(16)
    default:
      return hipErrorInvalidValue;
  }
#undef NFFTM_F3
}

// 3-D tile kernels (spreading writes per-item padded tiles to a.scratch; the
// fold is launched separately by launch_fold).
hipError_t launch_tile_d3(bool spread, bool dbl, int W, int NT, const KArgs& a, size_t, hipStream_t s,
                          const void* in, void* out) {
  return dbl ? go3w<double>(spread, W, NT, a, s, in, out) : go3w<float>(spread, W, NT, a, s, in, out);
}

template <typename R, int V>
static void geo3(int W, int* T, size_t* lds, int* P, int* SX, bool* valid) {
#define NFFTM_G(WW)                    \
  case WW:                             \
    T[0] = Geo3<R, WW, V>::T;          \
    T[1] = Geo3<R, WW, V>::TZ;         \
    P[0] = Geo3<R, WW, V>::P;          \
    P[1] = Geo3<R, WW, V>::PZ;         \
    *SX = Geo3<R, WW, V>::SX;          \
    *lds = Geo3<R, WW, V>::LDS;        \
    *valid = Geo3<R, WW, V>::VALID;    \
    break;
  switch (W) {
    NFFTM_G(2) NFFTM_G(4) NFFTM_G(6) NFFTM_G(8) NFFTM_G(10) NFFTM_G(12) NFFTM_G(14) NFFTM_G(16)
    default:
      T[0] = T[1] = 0;
      P[0] = P[1] = 0;
      *SX = 0;
      *lds = 0;
      *valid = false;
  }
#undef NFFTM_G
}

void tile3_geometry(bool dbl, int W, int V, int* T, size_t* lds, int* P, int* SX, bool* valid) {
  if (dbl) {
    if (V) geo3<double, 1>(W, T, lds, P, SX, valid);
    else geo3<double, 0>(W, T, lds, P, SX, valid);
  } else {
    if (V) geo3<float, 1>(W, T, lds, P, SX, valid);
    else geo3<float, 0>(W, T, lds, P, SX, valid);
  }
}

}  // namespace nfftm
