// Dense block kernels of the eigensolvers: Gram AᴴB, combine Cᵀ·S, row
// reductions, row-wise linear combinations and the counter-based random blocks.
//
// Reference: eigenfns/solver.py:35-57 (gram, combine), :95-126 (deflate,
// ortho_block), :196-205 and :228-238 (random rows, Rayleigh quotients and
// residual norms); interior.py:75-82.
// Gram products of FP32 inputs are exact in FP64 and summed in FP64, so the
// Gram carries no FP32 accumulation error (the compensated Gram the SVQB drop
// threshold and the normalised Rayleigh quotient rely on).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace efn {

// ---------------------------------------------------------------------------
// Gram partials: part[s][i][j] = Σ_{k in split s} conj(A[i][k])·B[j][k]
constexpr int kMaxProb = 64;  // problems per multi-problem launch (pointer lists in kernel args)
struct PairPtrs {
  const float2* a[kMaxProb];
  const float2* b[kMaxProb];
};
// Per-problem sizes of a ragged launch (ragged = 0: every problem has the
// launch's sizes).  A ragged Gram's output keeps the launch's (largest) shape
// per problem, zero outside the problem's own ma × mb.
struct ProbDims {
  short ma[kMaxProb];
  short mb[kMaxProb];
  int ragged;
};
// Output row i < split goes to out + i·n, row i ≥ split to out2 + (i − split)·n
// (split = m_out: one contiguous output block).
// Ragged launches (ragged = 1) give each problem its own m_in, m_out and split
// (C keeps the launch's largest m_in × m_out per problem).
struct CombinePtrs {
  const float2* s[kMaxProb];
  const float2* x[kMaxProb];
  float2* out[kMaxProb];
  float2* out2[kMaxProb];
  short m_in_p[kMaxProb];
  short m_out_p[kMaxProb];
  short split_p[kMaxProb];
  int split;
  int ragged;
};

__device__ __forceinline__ float2* combine_row(const CombinePtrs& cp, int prob, int i, long long n) {
  const int split = cp.ragged ? cp.split_p[prob] : cp.split;
  return i < split ? cp.out[prob] + (long long)i * n : cp.out2[prob] + (long long)(i - split) * n;
}

// Output tile of blockIdx.y: tiles_j > 0 — row-major over a tiles_i × tiles_j
// grid; tiles_j = −T — the upper block triangle ti ≤ tj of a T × T grid, row by
// row (for Hermitian products; the sum kernel mirrors the lower triangle).
__device__ __forceinline__ void tile_of(int by, int tiles_j, int& ti, int& tj) {
  if (tiles_j > 0) {
    ti = by / tiles_j;
    tj = by - ti * tiles_j;
    return;
  }
  const int T = -tiles_j;
  ti = 0;
  while (by >= T - ti) {
    by -= T - ti;
    ++ti;
  }
  tj = ti + by;
}

constexpr int kGT = 64;   // output tile edge
constexpr int kGK = 32;   // k-chunk staged in LDS
constexpr int kGPad = kGT + 1;

// grid: (splits, tiles_i·tiles_j, problems); part[(prob·splits + s)·ma·mb + i·mb + j]
__global__ __launch_bounds__(256) void gram_partial_kernel(PairPtrs pp, int ma, int mb, int tiles_j,
                                                           long long n, long long kper,
                                                           double2* __restrict__ part) {
  const float2* __restrict__ A = pp.a[blockIdx.z];
  const float2* __restrict__ B = pp.b[blockIdx.z];
  __shared__ float2 As[kGK * kGPad];
  __shared__ float2 Bs[kGK * kGPad];
  const int tid = threadIdx.x;
  const int ti = tid >> 4, tj = tid & 15;
  int ti_, tj_;
  tile_of(blockIdx.y, tiles_j, ti_, tj_);
  const int i0 = ti_ * kGT, j0 = tj_ * kGT;
  const long long kb = (long long)blockIdx.x * kper;
  const long long ke = kb + kper < n ? kb + kper : n;
  double ar[4][4], ai[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) ar[q][r] = ai[q][r] = 0.0;
  for (long long k0 = kb; k0 < ke; k0 += kGK) {
    // stage A[i0.., k0..] and B[j0.., k0..] transposed into LDS: As[k][i]
#pragma unroll
    for (int s = 0; s < (kGT * kGK) / 256; ++s) {
      const int e = tid + 256 * s;
      const int row = e / kGK, col = e - row * kGK;
      const long long k = k0 + col;
      float2 va = make_float2(0.f, 0.f), vb = make_float2(0.f, 0.f);
      if (k < ke) {
        if (i0 + row < ma) va = A[(long long)(i0 + row) * n + k];
        if (j0 + row < mb) vb = B[(long long)(j0 + row) * n + k];
      }
      As[col * kGPad + row] = va;
      Bs[col * kGPad + row] = vb;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kGK; ++k) {
      double axr[4], axi[4], bxr[4], bxi[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float2 a = As[k * kGPad + ti + 16 * q];
        const float2 b = Bs[k * kGPad + tj + 16 * q];
        axr[q] = a.x; axi[q] = a.y; bxr[q] = b.x; bxi[q] = b.y;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // conj(a)·b = (ar·br + ai·bi) + i(ar·bi − ai·br)
          ar[q][r] = fma(axr[q], bxr[r], fma(axi[q], bxi[r], ar[q][r]));
          ai[q][r] = fma(axr[q], bxi[r], fma(-axi[q], bxr[r], ai[q][r]));
        }
    }
    __syncthreads();
  }
  double2* out = part + ((long long)blockIdx.z * gridDim.x + blockIdx.x) * ma * mb;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + ti + 16 * q, j = j0 + tj + 16 * r;
      if (i < ma && j < mb) out[(long long)i * mb + j] = make_double2(ar[q][r], ai[q][r]);
    }
}

// ---------------------------------------------------------------------------
// Gram on the matrix cores.  Complex AᴴB as two real products over k' = 2k
// interleaved reals:
//   Re = Σ_k' A'[i][k']·B'[j][k'],  Im = Σ_k' A'[i][k']·B''[j][k'],
//   A' = (a_r, a_i), B' = (b_r, b_i), B'' = (b_i, −b_r) per complex entry.
// A wave owns a 32×32 output tile (2×2 MFMA tiles of 16×16) over a quarter of
// its block's k range.  Per step lane l reads one float4 (2 complex) of rows
// l%16 and l%16+16 at k' offset 4·(l/16); component q feeds MFMA q, whose
// operand k index is l/16 for A and B alike, so each step covers 16 k' once.
// MFMA 16x16x4 layouts: A[i][k] in lane i+16k, B[k][j] in lane j+16k; D[i][j]
// in lane j+16·(i/4), element i%4 for f32, and in lane j+16·(i%4), element
// i/4 for f64 (P::row maps lane group and element to the output row).
//   GramF64: v_mfma_f64_16x16x4f64 on the exact FP64 images of the FP32
//            inputs — products exact, sums in FP64 (same precision class as
//            the VALU kernel).  Three real products per complex entry
//            (Gauss): P1 = Σ a_r·b_r, P2 = Σ a_i·b_i, P3 = Σ (a_r − a_i)(b_r + b_i),
//            Re = P1 + P2, Im = (P3 − P1) + P2.  a_r − a_i and b_r + b_i of
//            two FP32 values are exact in FP64 unless their exponents differ
//            by more than 29, and the final cancellation costs a few FP64
//            ulps of Σ|a||b| — far below the FP32 data — for 3/4 of the
//            MFMA work of the four-product form.
//   GramF32: v_mfma_f32_16x16x4f32, FP32 accumulators added into FP64 every
//            kMFlush steps (≤ 4·kMFlush·4 products per FP32 partial sum).
// The four waves' tiles are summed through LDS in the fixed order
// (w0 + w2) + (w1 + w3) and each block writes one split partial, as the
// VALU kernel does, so the work-buffer layout is shared.
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
constexpr int kMWaveTile = 32;
constexpr int kMFlush = 4;
constexpr int kMStep = 8;  // complex entries per wave step
#ifndef EFN_GRAM_PREFETCH
#define EFN_GRAM_PREFETCH 2
#endif
constexpr int kMPrefetch = EFN_GRAM_PREFETCH;  // steps in flight per wave

#ifndef EFN_GRAM_GAUSS
#define EFN_GRAM_GAUSS 1
#endif
struct GramF64 {
  typedef double in_t;
  typedef doublex4 acc_t;
  static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int q) { return g + 4 * q; }
  static constexpr bool kFlush = false;
  static constexpr bool kGauss = EFN_GRAM_GAUSS;
};
struct GramF32 {
  typedef float in_t;
  typedef floatx4 acc_t;
  static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int q) { return 4 * g + q; }
  static constexpr bool kFlush = true;
  static constexpr bool kGauss = false;
};

__device__ __forceinline__ float4 load_c2(const float2* __restrict__ row, long long k, long long ke) {
  // complex entries k, k+1 of a row as a float4, zero at and beyond ke
  if (k + 1 < ke) return *reinterpret_cast<const float4*>(row + k);
  if (k < ke) {
    const float2 v = row[k];
    return make_float4(v.x, v.y, 0.f, 0.f);
  }
  return make_float4(0.f, 0.f, 0.f, 0.f);
}

// Block → (k split, output tile).  The tiles of one split are dispatched
// together, so the A and B rows they share are read from HBM about once;
// order 2 also keeps them on one XCD (workgroups go round-robin over the 8
// XCDs, each with its own L2), provided the block count is a multiple of 8.
#ifndef EFN_GRAM_ORDER
#define EFN_GRAM_ORDER 2
#endif
__device__ __forceinline__ void gram_block_of(int L, int ntile, int nsplit, int& split, int& tile) {
  if (EFN_GRAM_ORDER == 0) {  // split-major
    split = L % nsplit;
    tile = L / nsplit;
    return;
  }
  const int N = ntile * nsplit;
  int v = L;
  if (EFN_GRAM_ORDER == 2 && (N & 7) == 0) v = (L & 7) * (N >> 3) + (L >> 3);
  split = v / ntile;
  tile = v - split * ntile;
}

// grid: (splits·tiles, 1, problems) with 32×32 tiles; kper a multiple of
// 4·kMStep; rows 16-byte aligned (n even, aligned bases).
// One block's 32×32 tile over its k range.  NA × NB of the 2 × 2 sub-tiles
// of 16×16 hold rows < ma / columns < mb; SKIP10 drops sub-tile (1, 0) of a
// diagonal tile of a Hermitian product (strictly below the diagonal, never
// read).  Only the kept sub-tiles are loaded and multiplied.
template <class P, int NA, int NB, bool SKIP10>
__device__ __forceinline__ void gram_tile(const float2* __restrict__ A, const float2* __restrict__ B,
                                          int ma, int mb, int ostride, int i0, int j0, long long n, long long kb,
                                          long long ke, int wave, double2 (*red)[kMWaveTile * kMWaveTile],
                                          double2* __restrict__ out) {
  typedef typename P::in_t T;
  typedef typename P::acc_t Acc;
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
  auto keep = [](int a, int b) { return a < NA && b < NB && !(SKIP10 && a == 1 && b == 0); };
  const bool ra[2] = {i0 + r < ma, i0 + 16 + r < ma};
  const bool rb[2] = {j0 + r < mb, j0 + 16 + r < mb};
  const float2* ap[2] = {A + (long long)(ra[0] ? i0 + r : 0) * n, A + (long long)(ra[1] ? i0 + 16 + r : 0) * n};
  const float2* bp[2] = {B + (long long)(rb[0] ? j0 + r : 0) * n, B + (long long)(rb[1] ? j0 + 16 + r : 0) * n};
  // four-product form: cre = Re, cim = Im; Gauss form: cre = P1, cim = P2, c3 = P3
  Acc cre[2][2], cim[2][2], c3[2][2];
  double dre[2][2][4], dim_[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        cre[a][b][q] = 0;
        cim[a][b][q] = 0;
        c3[a][b][q] = 0;
        dre[a][b][q] = dim_[a][b][q] = 0.0;
      }
    }
  int steps = 0;
  const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
  // rows past ma / mb were loaded from a clamped (valid) row and count as zero
  auto step = [&](const float4 (&av_)[2], const float4 (&bv_)[2]) {
    T x[2][4], y[2][4], w[2][4];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const float4 av = ra[a] ? av_[a] : z4;
      x[a][0] = (T)av.x; x[a][1] = (T)av.y; x[a][2] = (T)av.z; x[a][3] = (T)av.w;
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const float4 bv = rb[b] ? bv_[b] : z4;
      y[b][0] = (T)bv.x; y[b][1] = (T)bv.y; y[b][2] = (T)bv.z; y[b][3] = (T)bv.w;
    }
    if constexpr (P::kGauss) {
      // entries h = 0, 1 of the float4: (re, im) at components 2h, 2h + 1
      T sx[2][2], sy[2][2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int a = 0; a < NA; ++a) sx[a][h] = x[a][2 * h] - x[a][2 * h + 1];
#pragma unroll
        for (int b = 0; b < NB; ++b) sy[b][h] = y[b][2 * h] + y[b][2 * h + 1];
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            if (!keep(a, b)) continue;
            cre[a][b] = P::mfma(x[a][2 * h], y[b][2 * h], cre[a][b]);
            cim[a][b] = P::mfma(x[a][2 * h + 1], y[b][2 * h + 1], cim[a][b]);
            c3[a][b] = P::mfma(sx[a][h], sy[b][h], c3[a][b]);
          }
    } else {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        w[b][0] = y[b][1]; w[b][1] = -y[b][0]; w[b][2] = y[b][3]; w[b][3] = -y[b][2];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            if (!keep(a, b)) continue;
            cre[a][b] = P::mfma(x[a][q], y[b][q], cre[a][b]);
            cim[a][b] = P::mfma(x[a][q], w[b][q], cim[a][b]);
          }
    }
    if (P::kFlush && ++steps == kMFlush) {
      steps = 0;
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            dre[a][b][q] += (double)cre[a][b][q];
            dim_[a][b][q] += (double)cim[a][b][q];
            cre[a][b][q] = 0;
            cim[a][b][q] = 0;
          }
    }
  };
  // Loads of one step (rows of kept sub-tiles only).  fast: the step lies
  // wholly inside [kb, ke) — unconditional 16-byte loads.  fetch: any step
  // (per-entry bounds, zero beyond ke, no access there).
  auto fast = [&](long long kw0, float4 (&av)[2], float4 (&bv)[2]) {
#pragma unroll
    for (int a = 0; a < NA; ++a) av[a] = *reinterpret_cast<const float4*>(ap[a] + kw0 + 2 * g);
#pragma unroll
    for (int b = 0; b < NB; ++b) bv[b] = *reinterpret_cast<const float4*>(bp[b] + kw0 + 2 * g);
  };
  auto fetch = [&](long long kw0, float4 (&av)[2], float4 (&bv)[2]) {
    if (kw0 + kMStep <= ke) {
      fast(kw0, av, bv);
    } else {
#pragma unroll
      for (int a = 0; a < NA; ++a) av[a] = load_c2(ap[a], kw0 + 2 * g, ke);
#pragma unroll
      for (int b = 0; b < NB; ++b) bv[b] = load_c2(bp[b], kw0 + 2 * g, ke);
    }
  };
  constexpr int D = kMPrefetch;
  float4 ra_buf[D][2], rb_buf[D][2];
  const long long span = ke > kb ? ke - kb : 0;
  const long long nst = (span + kMStep - 1) / kMStep;  // steps, the last maybe partial
  const long long nfull = span / kMStep;               // steps wholly in range
  long long t = 0;
  if (nfull >= 2 * D) {
    // steady state, branch-free from the prologue on: slot d is consumed,
    // then refilled with step t+d+D (all full).  Per-lane bounds branches
    // here made the compiler wait for every load in flight at each join, and
    // a refill before the consume landed in a temporary copied back after
    // the load returned.  The compiler is left to schedule across steps:
    // pinning each step in place (sched_barrier) measured slower.
#pragma unroll
    for (int d = 0; d < D; ++d) fast(kb + d * kMStep, ra_buf[d], rb_buf[d]);
    for (; t + 2 * D <= nfull; t += D) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        step(ra_buf[d], rb_buf[d]);
        fast(kb + (t + d + D) * kMStep, ra_buf[d], rb_buf[d]);
      }
    }
  } else {
#pragma unroll
    for (int d = 0; d < D; ++d) fetch(kb + d * kMStep, ra_buf[d], rb_buf[d]);
  }
  // the remaining (fewer than 2D + 1) steps
  for (; t < nst; t += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (t + d < nst) {  // wave-uniform
        const float4 av[2] = {ra_buf[d][0], ra_buf[d][1]};
        const float4 bv[2] = {rb_buf[d][0], rb_buf[d][1]};
        fetch(kb + (t + d + D) * kMStep, ra_buf[d], rb_buf[d]);
        step(av, bv);
      }
    }
  }
  // this wave's tile: element (P::row(g, q), r) of sub-tile (a, b)
  double2 v[2][2][4];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if constexpr (P::kGauss) {
          const double p1 = (double)cre[a][b][q], p2 = (double)cim[a][b][q];
          v[a][b][q] = make_double2(p1 + p2, ((double)c3[a][b][q] - p1) + p2);
        } else {
          v[a][b][q] = make_double2(dre[a][b][q] + (double)cre[a][b][q],
                                    dim_[a][b][q] + (double)cim[a][b][q]);
        }
      }
  // fixed-order reduction over the waves: (w0 + w2) + (w1 + w3)
  auto lds_at = [&](int a, int b, int q) { return (16 * a + P::row(g, q)) * kMWaveTile + 16 * b + r; };
  if (wave >= 2)
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (keep(a, b)) red[wave - 2][lds_at(a, b, q)] = v[a][b][q];
  __syncthreads();
  if (wave < 2)
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (!keep(a, b)) continue;
          const double2 o = red[wave][lds_at(a, b, q)];
          v[a][b][q].x += o.x;
          v[a][b][q].y += o.y;
        }
  __syncthreads();
  if (wave == 1)
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (keep(a, b)) red[0][lds_at(a, b, q)] = v[a][b][q];
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!keep(a, b)) continue;
        const int i = i0 + 16 * a + P::row(g, q), j = j0 + 16 * b + r;
        const double2 o = red[0][lds_at(a, b, q)];
        if (i < ma && j < mb) out[(long long)i * ostride + j] = make_double2(v[a][b][q].x + o.x, v[a][b][q].y + o.y);
      }
}

template <class P>
__global__ __launch_bounds__(256, 2) void gram_mfma_kernel(PairPtrs pp, ProbDims pd, int ma, int mb,
                                                        int tiles_j, int ntile, int nsplit, long long n,
                                                        long long kper, double2* __restrict__ part) {
  __shared__ double2 red[2][kMWaveTile * kMWaveTile];
  const float2* __restrict__ A = pp.a[blockIdx.z];
  const float2* __restrict__ B = pp.b[blockIdx.z];
  // wave index through readfirstlane: the compiler then keeps the k range in
  // scalar registers and the step checks are scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int split, tile, ti_, tj_;
  gram_block_of(blockIdx.x, ntile, nsplit, split, tile);
  tile_of(tile, tiles_j, ti_, tj_);
  const int i0 = ti_ * kMWaveTile, j0 = tj_ * kMWaveTile;
  // this problem's sizes (the partials keep the launch's ma × mb layout)
  const int ma_p = pd.ragged ? pd.ma[blockIdx.z] : ma;
  const int mb_p = pd.ragged ? pd.mb[blockIdx.z] : mb;
  if (i0 >= ma_p || j0 >= mb_p) return;  // block-uniform: a tile past this problem's edge
  const long long kb_blk = (long long)split * kper;
  const long long ke_blk = kb_blk + kper < n ? kb_blk + kper : n;
  const long long kw = kper / 4;
  const long long kb = kb_blk + wave * kw;
  const long long ke = kb + kw < ke_blk ? kb + kw : ke_blk;
  double2* out = part + ((long long)blockIdx.z * nsplit + split) * ma * mb;
  // sub-tile shape of this tile (block-uniform)
  const bool fa = i0 + 16 < ma_p, fb = j0 + 16 < mb_p;
  const bool diag = tiles_j < 0 && ti_ == tj_;
  if (fa && fb) {
    if (diag)
      gram_tile<P, 2, 2, true>(A, B, ma_p, mb_p, mb, i0, j0, n, kb, ke, wave, red, out);
    else
      gram_tile<P, 2, 2, false>(A, B, ma_p, mb_p, mb, i0, j0, n, kb, ke, wave, red, out);
  } else if (fa) {
    gram_tile<P, 2, 1, false>(A, B, ma_p, mb_p, mb, i0, j0, n, kb, ke, wave, red, out);
  } else if (fb) {
    gram_tile<P, 1, 2, false>(A, B, ma_p, mb_p, mb, i0, j0, n, kb, ke, wave, red, out);
  } else {
    gram_tile<P, 1, 1, false>(A, B, ma_p, mb_p, mb, i0, j0, n, kb, ke, wave, red, out);
  }
}

// zero fill of count double2 (a kernel rather than a memset node, so the
// clear is safe inside captured graphs)
__global__ void zero_double2_kernel(double2* __restrict__ out, long long count) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
       e += (long long)gridDim.x * blockDim.x)
    out[e] = make_double2(0.0, 0.0);
}

// out[prob][e] = Σ_s part[prob][s][e] in fixed order (grid.y = problem).
// herm: only tiles ti ≤ tj were computed; the result is made exactly
// Hermitian — entry (i, j) with i > j is the conjugate of (j, i) and the
// diagonal is real.
__global__ void sum_partials_kernel(const double2* __restrict__ part, int nsplit, long long count,
                                    int mb, int herm, ProbDims pd, double2* __restrict__ out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  part += (long long)blockIdx.y * nsplit * count;
  out += (long long)blockIdx.y * count;
  const int i = (int)(e / mb), j = (int)(e - (long long)i * mb);
  if (pd.ragged && (i >= pd.ma[blockIdx.y] || j >= pd.mb[blockIdx.y])) {  // past this problem's edge
    out[e] = make_double2(0.0, 0.0);
    return;
  }
  if (herm && i > j) return;  // written by the thread of (j, i)
  // four independent chains (split s into chain s % 4 while four remain),
  // then (c0 + c1) + (c2 + c3): a fixed order that depends on nsplit only
  double re[4] = {0.0, 0.0, 0.0, 0.0}, im[4] = {0.0, 0.0, 0.0, 0.0};
  int s = 0;
  for (; s + 4 <= nsplit; s += 4) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double2 p = part[(long long)(s + c) * count + e];
      re[c] += p.x;
      im[c] += p.y;
    }
  }
  for (; s < nsplit; ++s) {
    const double2 p = part[(long long)s * count + e];
    re[0] += p.x;
    im[0] += p.y;
  }
  const double vr = (re[0] + re[1]) + (re[2] + re[3]);
  const double vi = (im[0] + im[1]) + (im[2] + im[3]);
  if (!herm) {
    out[e] = make_double2(vr, vi);
  } else if (i == j) {
    out[e] = make_double2(vr, 0.0);
  } else {
    out[e] = make_double2(vr, vi);
    out[(long long)j * mb + i] = make_double2(vr, -vi);
  }
}

// ---------------------------------------------------------------------------
// combine: out[i][k] = alpha·X[i][k] + Σ_j C[j][i]·S[j][k]   (C is m_in × m_out, c64)
// grid: (columns/256, row tiles, problems); C is problems × m_in × m_out
template <int TI>
__global__ __launch_bounds__(256) void combine_kernel(const float2* __restrict__ Call, int m_in, int m_out,
                                                      CombinePtrs cp, long long n, float alpha) {
  // each thread reads all m_in rows of its column before it stores, so with
  // one row tile (gridDim.y = 1) the outputs may lie in S
  const float2* __restrict__ C = Call + (long long)blockIdx.z * m_in * m_out;
  const float2* S = cp.s[blockIdx.z];
  const float2* X = cp.x[blockIdx.z];
  const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
  const int i0 = blockIdx.y * TI;
  if (k >= n) return;
  float2 acc[TI];
#pragma unroll
  for (int i = 0; i < TI; ++i) acc[i] = make_float2(0.f, 0.f);
  for (int j = 0; j < m_in; ++j) {
    const float2 s = S[(long long)j * n + k];
    const float2* crow = C + (long long)j * m_out + i0;
#pragma unroll
    for (int i = 0; i < TI; ++i) {
      if (i0 + i < m_out) {
        const float2 c = crow[i];
        acc[i].x = fmaf(c.x, s.x, fmaf(-c.y, s.y, acc[i].x));
        acc[i].y = fmaf(c.x, s.y, fmaf(c.y, s.x, acc[i].y));
      }
    }
  }
#pragma unroll
  for (int i = 0; i < TI; ++i) {
    if (i0 + i < m_out) {
      float2 r = acc[i];
      if (X != nullptr) {
        const float2 x = X[(long long)(i0 + i) * n + k];
        r.x = fmaf(alpha, x.x, r.x);
        r.y = fmaf(alpha, x.y, r.y);
      }
      combine_row(cp, blockIdx.z, i0 + i, n)[k] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// combine on the matrix cores (v_mfma_f32_16x16x4f32): out = Cᵀ·S (+ alpha·X)
// as real products of the complex operands, FP32 accumulation as in the
// vector-ALU kernel:
//   Re += Cr·Sr − Ci·Si,  Im += Cr·Si + Ci·Sr      (A = Cᵀ 16 rows × 4 j, B = S 4 j × 16 columns)
// A wave owns 16·NA output rows × 64 columns (NA × 4 tiles of 16×16); a
// block's four waves take consecutive 64-column ranges.  Lane l = r + 16·g
// owns the four consecutive columns c0 + 4r .. c0 + 4r + 3, column c0 + 4r + b
// belonging to tile b: its B operand for rows j0 + g is two 16-byte loads,
// and D[i][r] of tile b (lane r + 16·(i/4), element i%4) puts the four
// columns of output row i in one lane, stored as two 16-byte runs.  The rows
// of S stream through a register ring of kCPrefetch steps of 4 rows.  Every
// output element is the same MFMA chain in j order whatever the column
// mapping.  Each wave reads all m_in rows of its columns before it stores,
// so with one row tile (gridDim.y = 1) the outputs may lie in S.  m_out ≥ 1.
// vec: n even and S, X, out 16-byte aligned (float4 access); else per element.
#ifndef EFN_COMBINE_PREFETCH
#define EFN_COMBINE_PREFETCH 2
#endif
constexpr int kCPrefetch = EFN_COMBINE_PREFETCH;

template <int NA>
__global__ __launch_bounds__(256, 2) void combine_mfma_kernel(const float2* __restrict__ Call, int m_in,
                                                              int m_out, CombinePtrs cp, long long n,
                                                              float alpha, int vec) {
  // C keeps the launch's m_in × m_out per problem (row stride m_out); a
  // ragged problem uses its own first m_in_p rows and m_out_p columns
  const float2* __restrict__ C = Call + (long long)blockIdx.z * m_in * m_out;
  const int ldc = m_out;
  if (cp.ragged) {
    m_in = cp.m_in_p[blockIdx.z];
    m_out = cp.m_out_p[blockIdx.z];
  }
  const float2* S = cp.s[blockIdx.z];
  const float2* X = cp.x[blockIdx.z];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, g = lane >> 4;
  const long long c0 = ((long long)blockIdx.x * 4 + wave) * 64;
  const long long cl = c0 + 4 * r;
  const bool wfull = vec && c0 + 64 <= n;  // wave-uniform: all 64 columns, 16-byte access
  const bool full = vec && cl + 3 < n;
  const int i0 = blockIdx.y * 16 * NA;
  if (i0 >= m_out) return;  // block-uniform: a row tile past this problem's outputs
  // C column of each A row; rows past m_out read column m_out − 1 and count as zero
  int ic[NA];
  bool iok[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const int i = i0 + 16 * a + r;
    iok[a] = i < m_out;
    ic[a] = iok[a] ? i : m_out - 1;
  }
  floatx4 re[NA][4], im[NA][4];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      re[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
      im[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
  // rows j0 + g of S (this lane's 4 columns) and of C.  fast: rows j0..j0+3
  // exist and the wave's columns are full — unconditional loads; fetch: any.
  auto fast = [&](int j0, float4 (&sv)[2], float2 (&cv)[NA]) {
    const int j = j0 + g;
    const float2* row = S + (long long)j * n;
    sv[0] = *reinterpret_cast<const float4*>(row + cl);
    sv[1] = *reinterpret_cast<const float4*>(row + cl + 2);
#pragma unroll
    for (int a = 0; a < NA; ++a) cv[a] = C[(long long)j * ldc + ic[a]];
  };
  auto fetch = [&](int j0, float4 (&sv)[2], float2 (&cv)[NA]) {
    const int j = j0 + g;
    sv[0] = sv[1] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int a = 0; a < NA; ++a) cv[a] = make_float2(0.f, 0.f);
    if (j >= m_in) return;
    const float2* row = S + (long long)j * n;
    if (full) {
      sv[0] = *reinterpret_cast<const float4*>(row + cl);
      sv[1] = *reinterpret_cast<const float4*>(row + cl + 2);
    } else {
      float2 e[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) e[b] = cl + b < n ? row[cl + b] : make_float2(0.f, 0.f);
      sv[0] = make_float4(e[0].x, e[0].y, e[1].x, e[1].y);
      sv[1] = make_float4(e[2].x, e[2].y, e[3].x, e[3].y);
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) cv[a] = C[(long long)j * ldc + ic[a]];
  };
  auto step = [&](const float4 (&sv)[2], const float2 (&cv_)[NA]) {
    const float sr[4] = {sv[0].x, sv[0].z, sv[1].x, sv[1].z};
    const float si[4] = {sv[0].y, sv[0].w, sv[1].y, sv[1].w};
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const float2 cv = iok[a] ? cv_[a] : make_float2(0.f, 0.f);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        re[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(cv.x, sr[b], re[a][b], 0, 0, 0);
        re[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(-cv.y, si[b], re[a][b], 0, 0, 0);
        im[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(cv.x, si[b], im[a][b], 0, 0, 0);
        im[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(cv.y, sr[b], im[a][b], 0, 0, 0);
      }
    }
  };
  constexpr int D = kCPrefetch;
  float4 sb[D][2];
  float2 cb[D][NA];
  const int nst = (m_in + 3) / 4, nfull = m_in / 4;
  int t = 0;
  if (wfull && nfull >= 2 * D) {
    // branch-free steady state from the prologue on, slot consumed then
    // refilled with a full step (as in the Gram kernel: D − 1 steps in flight)
#pragma unroll
    for (int d = 0; d < D; ++d) {  // in slot order, so slot 0 is the oldest
      fast(4 * d, sb[d], cb[d]);
      __builtin_amdgcn_sched_barrier(0);
    }
    for (; t + 2 * D <= nfull; t += D) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        step(sb[d], cb[d]);
        fast(4 * (t + d + D), sb[d], cb[d]);
        __builtin_amdgcn_sched_barrier(0);  // steps kept in place: measured faster here
      }
    }
  } else {
#pragma unroll
    for (int d = 0; d < D; ++d) fetch(4 * d, sb[d], cb[d]);
  }
  for (; t < nst; t += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (t + d < nst) {  // wave-uniform
        const float4 sv[2] = {sb[d][0], sb[d][1]};
        float2 cv[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) cv[a] = cb[d][a];
        fetch(4 * (t + d + D), sb[d], cb[d]);
        step(sv, cv);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + 16 * a + 4 * g + q;
      if (i >= m_out) continue;
      float2 v[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) v[b] = make_float2(re[a][b][q], im[a][b][q]);
      float2* orow = combine_row(cp, blockIdx.z, i, n);
      if (full) {
        if (X != nullptr) {
          const float2* xrow = X + (long long)i * n;
          const float4 x0 = *reinterpret_cast<const float4*>(xrow + cl);
          const float4 x1 = *reinterpret_cast<const float4*>(xrow + cl + 2);
          const float xs[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            v[b].x = fmaf(alpha, xs[2 * b], v[b].x);
            v[b].y = fmaf(alpha, xs[2 * b + 1], v[b].y);
          }
        }
        *reinterpret_cast<float4*>(orow + cl) = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
        *reinterpret_cast<float4*>(orow + cl + 2) = make_float4(v[2].x, v[2].y, v[3].x, v[3].y);
      } else {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const long long col = cl + b;
          if (col >= n) continue;
          float2 w = v[b];
          if (X != nullptr) {
            const float2 xv = X[(long long)i * n + col];
            w.x = fmaf(alpha, xv.x, w.x);
            w.y = fmaf(alpha, xv.y, w.y);
          }
          orow[col] = w;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// Row reductions in FP64 over splits: part[i][s] = Σ_k f(A[i][k], B[i][k])
// op 0: |A|²  (re only)   op 1: conj(A)·B (re, im)
__device__ __forceinline__ double2 wave_sum2(double2 v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  return v;
}

// Logical row i of a strided row set: physical row (i / nrow)·stride + i % nrow.
__device__ __forceinline__ long long phys_row(int i, int nrow, long long stride) {
  const int slot = i / nrow;
  return (long long)slot * stride + (i - slot * nrow);
}

// Row maps of the row-statistics kernels: without a selection, logical row i
// is phys_row(i, nrow, stride); with one, sel[i] = slot << 16 | r names row r
// of slot `slot` (physical slot·stride + r), and a pair outside r < nrow,
// slot < nslot is no row at all (-1: nothing read or written, zero sums).
struct RowSel {
  const int* sel;
  int nrow, nslot;
};
__device__ __forceinline__ bool sel_pair(const RowSel& rs, int i, int& slot, int& r) {
  if (!rs.sel) {
    slot = i / rs.nrow;
    r = i - slot * rs.nrow;
    return true;
  }
  const int v = rs.sel[i];
  slot = v >> 16;
  r = v & 0xFFFF;
  if (v >= 0 && r < rs.nrow && slot < rs.nslot) return true;
  slot = r = 0;
  return false;
}

template <int OP>
__global__ __launch_bounds__(256) void rowreduce_partial_kernel(const float2* __restrict__ A, long long sa,
                                                                const float2* __restrict__ B, long long sb,
                                                                int nrow, long long n, long long kper,
                                                                double2* __restrict__ part) {
  __shared__ double2 red[4];
  const int i = blockIdx.y, s = blockIdx.x, nsplit = gridDim.x;
  const long long kb = (long long)s * kper;
  const long long ke = kb + kper < n ? kb + kper : n;
  const float2* a = A + phys_row(i, nrow, sa) * n;
  const float2* b = B + phys_row(i, nrow, sb) * n;
  double re = 0.0, im = 0.0;
  for (long long k = kb + threadIdx.x; k < ke; k += 256) {
    const float2 x = a[k];
    if (OP == 0) {
      re += (double)x.x * x.x + (double)x.y * x.y;
    } else {
      const float2 y = b[k];
      re += (double)x.x * y.x + (double)x.y * y.y;
      im += (double)x.x * y.y - (double)x.y * y.x;
    }
  }
  double2 v = wave_sum2(make_double2(re, im));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 t = red[0];
    for (int w = 1; w < 4; ++w) { t.x += red[w].x; t.y += red[w].y; }
    part[(long long)i * nsplit + s] = t;
  }
}

__global__ void rowreduce_final_kernel(const double2* __restrict__ part, int nsplit, int m,
                                       double2* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double re = 0.0, im = 0.0;
  for (int s = 0; s < nsplit; ++s) {
    re += part[(long long)i * nsplit + s].x;
    im += part[(long long)i * nsplit + s].y;
  }
  out[i] = make_double2(re, im);
}

// Fused row statistics of a pair of strided row sets in one read of each:
// part[i][s] = (Re⟨A,B⟩, Im⟨A,B⟩, |A|², |B|²) over split s, FP64 sums of exact
// FP32 products (same split layout as rowreduce_partial_kernel).
__global__ __launch_bounds__(256) void rowstats_partial_kernel(const float2* __restrict__ A, long long sa,
                                                               const float2* __restrict__ B, long long sb,
                                                               RowSel rs, long long n, long long kper,
                                                               double2* __restrict__ part) {
  __shared__ double2 red[8];
  const int i = blockIdx.y, s = blockIdx.x, nsplit = gridDim.x;
  const long long kb = (long long)s * kper;
  int slot, r;
  const bool valid = sel_pair(rs, i, slot, r);
  const long long ke = !valid ? kb : (kb + kper < n ? kb + kper : n);
  const float2* a = A + ((long long)slot * sa + r) * n;
  const float2* b = B + ((long long)slot * sb + r) * n;
  double dre = 0.0, dim_ = 0.0, aa = 0.0, bb = 0.0;
  for (long long k = kb + threadIdx.x; k < ke; k += 256) {
    const float2 x = a[k], y = b[k];
    dre += (double)x.x * y.x + (double)x.y * y.y;
    dim_ += (double)x.x * y.y - (double)x.y * y.x;
    aa += (double)x.x * x.x + (double)x.y * x.y;
    bb += (double)y.x * y.x + (double)y.y * y.y;
  }
  const double2 v0 = wave_sum2(make_double2(dre, dim_)), v1 = wave_sum2(make_double2(aa, bb));
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = v0;
    red[4 + (threadIdx.x >> 6)] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 t0 = red[0], t1 = red[4];
    for (int w = 1; w < 4; ++w) {
      t0.x += red[w].x; t0.y += red[w].y;
      t1.x += red[4 + w].x; t1.y += red[4 + w].y;
    }
    part[((long long)i * nsplit + s) * 2] = t0;
    part[((long long)i * nsplit + s) * 2 + 1] = t1;
  }
}

// out[i] = the four sums of row i over the splits, fixed order
__global__ void rowstats_final_kernel(const double2* __restrict__ part, int nsplit, int m,
                                      double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < nsplit; ++s) {
    const double2 p0 = part[((long long)i * nsplit + s) * 2], p1 = part[((long long)i * nsplit + s) * 2 + 1];
    v[0] += p0.x;
    v[1] += p0.y;
    v[2] += p1.x;
    v[3] += p1.y;
  }
  for (int c = 0; c < 4; ++c) out[(long long)i * 4 + c] = v[c];
}

// out[i] = a[i]·X[i] + b[i]·Y[i] and per-split partials of |out[i]|² (FP64 of
// the stored FP32 values), split layout as rowreduce_partial_kernel.
__global__ __launch_bounds__(256) void row_axpby_norm_kernel(const float2* __restrict__ X, long long sx,
                                                             const double* __restrict__ a,
                                                             const float2* __restrict__ Y, long long sy,
                                                             const double* __restrict__ b, RowSel rs,
                                                             long long n, long long kper,
                                                             float2* __restrict__ out, long long so,
                                                             double2* __restrict__ part) {
  __shared__ double2 red[4];
  const int i = blockIdx.y, s = blockIdx.x, nsplit = gridDim.x;
  const long long kb = (long long)s * kper;
  int slot, r;
  const bool valid = sel_pair(rs, i, slot, r);
  const long long ke = !valid ? kb : (kb + kper < n ? kb + kper : n);
  const float ai = (float)a[i];
  const float bi = Y ? (float)b[i] : 0.f;
  const float2* xr = X + ((long long)slot * sx + r) * n;
  const float2* yr = Y ? Y + ((long long)slot * sy + r) * n : nullptr;
  float2* orow = out + ((long long)slot * so + r) * n;
  double acc = 0.0;
  for (long long k = kb + threadIdx.x; k < ke; k += 256) {
    const float2 x = xr[k];
    float2 r = make_float2(ai * x.x, ai * x.y);
    if (yr) {
      const float2 y = yr[k];
      r.x = fmaf(bi, y.x, r.x);
      r.y = fmaf(bi, y.y, r.y);
    }
    orow[k] = r;
    acc += (double)r.x * r.x + (double)r.y * r.y;
  }
  const double2 v = wave_sum2(make_double2(acc, 0.0));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0].x;
    for (int w = 1; w < 4; ++w) t += red[w].x;
    part[(long long)i * nsplit + s] = make_double2(t, 0.0);
  }
}

// ---------------------------------------------------------------------------
// out[i][k] = a[i]·X[i][k] + b[i]·Y[i][k] (a, b real per row, FP64 on device,
// applied in FP32).  Y may be null.
__global__ void row_axpby_kernel(const float2* __restrict__ X, long long sx, const double* __restrict__ a,
                                 const float2* __restrict__ Y, long long sy, const double* __restrict__ b,
                                 int nrow, long long n, float2* __restrict__ out, long long so) {
  const int i = blockIdx.y;
  const float ai = (float)a[i];
  const float bi = Y ? (float)b[i] : 0.f;
  const float2* xr = X + phys_row(i, nrow, sx) * n;
  const float2* yr = Y ? Y + phys_row(i, nrow, sy) * n : nullptr;
  float2* orow = out + phys_row(i, nrow, so) * n;
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
    const float2 x = xr[k];
    float2 r = make_float2(ai * x.x, ai * x.y);
    if (yr) {
      const float2 y = yr[k];
      r.x = fmaf(bi, y.x, r.x);
      r.y = fmaf(bi, y.y, r.y);
    }
    orow[k] = r;
  }
}

// out = a·X + b·Y with scalar coefficients over a flat range
__global__ void axpby_kernel(const float2* __restrict__ X, float a, const float2* __restrict__ Y,
                             float b, long long count, float2* __restrict__ out) {
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long long)gridDim.x * 256) {
    const float2 x = X[e];
    float2 r = make_float2(a * x.x, a * x.y);
    if (Y) {
      const float2 y = Y[e];
      r.x = fmaf(b, y.x, r.x);
      r.y = fmaf(b, y.y, r.y);
    }
    out[e] = r;
  }
}

// dst_p[e] = src_p[e] for e < count, problem p = blockIdx.y (NULL src = zero fill)
struct CopyPtrs {
  const float2* src[kMaxProb];
  float2* dst[kMaxProb];
};
__global__ void copy_multi_kernel(CopyPtrs cp, long long count) {
  const float2* __restrict__ src = cp.src[blockIdx.y];
  float2* __restrict__ dst = cp.dst[blockIdx.y];
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long long)gridDim.x * 256)
    dst[e] = src ? src[e] : make_float2(0.f, 0.f);
}

// ---------------------------------------------------------------------------
// Counter-based random blocks (splitmix64 stream + Box–Muller / Rademacher),
// multiplied by the Γ mask (zero at the singular voxel).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void random_block_kernel(int kind, uint64_t key, long long row0, long long G3,
                                    long long count, long long sing_voxel,
                                    float2* __restrict__ out) {
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long long)gridDim.x * 256) {
    const uint64_t ge = (uint64_t)row0 * (uint64_t)(2 * G3) + (uint64_t)e;
    const uint64_t h = splitmix64(key + ge * 0x9E3779B97F4A7C15ull);
    float2 z;
    if (kind == 0) {
      const double u1 = ((double)(h >> 40) + 0.5) * 0x1p-24;
      const double u2 = ((double)((h >> 16) & 0xFFFFFFull) + 0.5) * 0x1p-24;
      const double r = sqrt(-2.0 * log(u1));
      z = make_float2((float)(r * cos(2.0 * 3.14159265358979323846 * u2)),
                      (float)(r * sin(2.0 * 3.14159265358979323846 * u2)));
    } else {
      z = make_float2((h >> 63) ? -1.f : 1.f, 0.f);
    }
    const long long vox = e % G3;
    if (vox == sing_voxel) z = make_float2(0.f, 0.f);
    out[e] = z;
  }
}

}  // namespace efn
