// Cholesky-QR coefficients for the small dense FP64 orthonormalisations of
// LOBPCG (the SVQB steps of eigenfns/solver.py:113-126 and the P whitening of
// :258-274): for a Hermitian positive semidefinite Gram matrix G = Yᴴ·Y of n
// block rows (n ≤ 96, complex128, row-major) find C with Cᴴ·G·C = I on the
// numerically independent rows, so that the rows q_i = Σ_j C[j][i]·y_j are
// orthonormal.
//
// Pivoted Cholesky: at step j the largest remaining diagonal of the Schur
// complement is swapped to position j; the factorisation stops (rank = j)
// once that diagonal falls to drop·max(diag G) or below — the rows beyond are
// numerically dependent on the ones taken, and their output rows are zero,
// as SVQB's dropped directions are (the threshold is on pivots instead of
// eigenvalues, so the two rules agree on clearly (in)dependent rows, not at
// the margin).  With ΠᵀGΠ = L·Lᴴ on the first `rank` pivots,
// C[π(j)][i] = conj((L⁻¹)[i][j]) for j ≤ i < rank (C = Π·L⁻ᴴ), zero
// elsewhere.  All arithmetic in FP64.
//
// One workgroup per matrix: G is copied into LDS (rows padded to n + 1),
// factorised in place with the whole Schur complement kept Hermitian (row and
// column swaps stay plain), wave 0 finds each pivot with a butterfly
// reduction, and thread c then inverts column c of L by forward substitution,
// writing row π(c) of C.
#pragma once

#include <hip/hip_runtime.h>

#include "eig_kernels.hpp"

namespace efn {

constexpr int kCholMaxN = 96;
constexpr int kCholThreads = 256;

// grid: (nprob); block: kCholThreads.  Dynamic LDS: n·(n + 1) double2 +
// n ints + 2 ints.  G: nprob × n × n, C: nprob × n × n, rank: nprob.
__global__ __launch_bounds__(kCholThreads) void cholqr_coeffs_kernel(const double2* __restrict__ G_in, int n,
                                                                     double drop, double2* __restrict__ C_out,
                                                                     int* __restrict__ rank_out) {
  extern __shared__ double2 lds[];
  const int ld = n + 1;
  double2* A = lds;
  int* piv = reinterpret_cast<int*>(A + (long long)n * ld);
  int* ctl = piv + n;  // ctl[0]: pivot index of the step, ctl[1]: rank
  const int tid = threadIdx.x;
  const long long base = (long long)blockIdx.x * n * n;
  const double2* G = G_in + base;
  double2* C = C_out + base;

  for (int e = tid; e < n * n; e += kCholThreads) {
    const int i = e / n, j = e - i * n;
    A[i * ld + j] = G[e];
    C[e] = make_double2(0.0, 0.0);
  }
  for (int i = tid; i < n; i += kCholThreads) piv[i] = i;
  __syncthreads();
  // threshold from the largest diagonal of G (wave 0)
  double thresh = 0.0;
  if (tid < 64) {
    double m = 0.0;
    for (int i = tid; i < n; i += 64) {
      const double v = A[i * ld + i].x;
      m = (v > m || v != v) ? v : m;  // NaN propagates
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(m, off, 64);
      m = (o > m || o != o) ? o : m;
    }
    if (tid == 0) ctl[1] = (m > 0.0 && m <= 1.7976931348623157e308) ? n : 0;  // NaN, inf, 0 → rank 0
    thresh = drop * m;  // used by wave 0 only
  }
  __syncthreads();
  int rank = ctl[1];

  for (int j = 0; j < rank; ++j) {
    // pivot: the largest remaining diagonal (wave 0), ties to the lowest index
    if (tid < 64) {
      double best = -1.0;
      int bi = j;
      for (int i = j + tid; i < n; i += 64) {
        const double v = A[i * ld + i].x;
        if (v > best) {
          best = v;
          bi = i;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (tid == 0) {
        ctl[0] = bi;
        if (!(best > thresh)) ctl[1] = j;  // rank reached (also catches NaN pivots)
      }
    }
    __syncthreads();
    if (ctl[1] == j) {
      rank = j;
      break;
    }
    const int p = ctl[0];
    if (p != j) {  // uniform
      for (int c = tid; c < n; c += kCholThreads) {
        const double2 t = A[j * ld + c];
        A[j * ld + c] = A[p * ld + c];
        A[p * ld + c] = t;
      }
      __syncthreads();
      for (int r = tid; r < n; r += kCholThreads) {
        const double2 t = A[r * ld + j];
        A[r * ld + j] = A[r * ld + p];
        A[r * ld + p] = t;
      }
      if (tid == 0) {
        const int t = piv[j];
        piv[j] = piv[p];
        piv[p] = t;
      }
      __syncthreads();
    }
    const double ljj = sqrt(A[j * ld + j].x);
    const double inv = 1.0 / ljj;
    // column j below the diagonal, and row j right of it (kept Hermitian)
    for (int i = j + 1 + tid; i < n; i += kCholThreads) {
      const double2 v = A[i * ld + j];
      A[i * ld + j] = make_double2(v.x * inv, v.y * inv);
      A[j * ld + i] = make_double2(v.x * inv, -v.y * inv);
    }
    __syncthreads();
    if (tid == 0) A[j * ld + j] = make_double2(ljj, 0.0);
    // Schur complement: A[i][l] -= L[i][j]·conj(L[l][j]) for i, l > j
    const int w = n - j - 1;
    for (int e = tid; e < w * w; e += kCholThreads) {
      const int i = j + 1 + e / w, l = j + 1 + e % w;
      const double2 a = A[i * ld + j], b = A[l * ld + j];
      double2 v = A[i * ld + l];
      v.x -= a.x * b.x + a.y * b.y;
      v.y -= a.y * b.x - a.x * b.y;
      A[i * ld + l] = v;
    }
    __syncthreads();
  }
  if (tid == 0) rank_out[blockIdx.x] = rank;
  __syncthreads();  // the zero fill of C before the rows written below

  // column c of M = L⁻¹ (rank × rank, lower) by forward substitution; row
  // π(c) of C holds conj(M[i][c]) at column i ≥ c, and is read back as the
  // substitution proceeds (the thread's own stores)
  for (int c = tid; c < rank; c += kCholThreads) {
    double2* row = C + (long long)piv[c] * n;
    const double2 mcc = make_double2(1.0 / A[c * ld + c].x, 0.0);
    row[c] = mcc;  // conj of a real value
    for (int i = c + 1; i < rank; ++i) {
      double2 s = make_double2(0.0, 0.0);
      for (int l = c; l < i; ++l) {
        const double2 m = zconj(row[l]);  // M[l][c]
        const double2 a = A[i * ld + l];  // L[i][l]
        s.x += a.x * m.x - a.y * m.y;
        s.y += a.x * m.y + a.y * m.x;
      }
      const double r = -1.0 / A[i * ld + i].x;
      row[i] = make_double2(s.x * r, -s.y * r);  // conj(M[i][c])
    }
  }
}

}  // namespace efn
