// Batched Hermitian eigensolver for the small dense FP64 problems of LOBPCG
// (SVQB Gram matrices, Rayleigh–Ritz, n ≤ 96): cyclic two-sided Jacobi, one
// workgroup per matrix.
//
// A (n × n, complex128) lives in LDS with rows padded to n + 1 entries; the
// eigenvector accumulator V is the output buffer itself (row-major, V[i][j]
// = component i of eigenvector j).  A sweep is N − 1 rounds of the circle
// ordering over N = n rounded up to even (an index n is a bye); a round
// rotates its N/2 disjoint pairs in parallel: rotation parameters, then
// A ← A·U with V ← V·U (columns), then A ← Uᴴ·A (rows), three barriers a
// round; each update phase spreads its (pair, row) items over the whole
// workgroup and issues a thread's loads before its stores.
//
// The rotation for pair (p, q): with b = A[p][q] = |b|e^{iφ}, D = diag(1,
// e^{−iφ}) makes the 2 × 2 block real, [[a, |b|], [|b|, d]], which the real
// Jacobi rotation R = [[c, s], [−s, c]] (t = tan θ the smaller root of
// t² + 2τt − 1 = 0, τ = (d − a)/(2|b|)) diagonalises; U = D·R =
// [[c, s], [−s e^{−iφ}, c e^{−iφ}]], and the new diagonal is a − t|b|,
// d + t|b| with A[p][q] = 0.  A pair is skipped when
// |b| ≤ kJacobiTol·sqrt(|a·d|) (relative criterion: small eigenvalues keep
// their relative accuracy next to large ones); a sweep without a rotation
// ends the iteration (or max_sweeps).  Eigenvalues come out ascending with
// their vectors.
#pragma once

#include <hip/hip_runtime.h>

namespace efn {

constexpr int kEigMaxN = 96;
constexpr int kEigThreads = 256;
// (pair, row) work items per thread in one update phase: ⌈(N/2)·n / threads⌉
constexpr int kEigItems = (kEigMaxN / 2 * kEigMaxN + kEigThreads - 1) / kEigThreads;
constexpr int kEigChunk = 6;  // items whose loads are in flight together (register budget)
constexpr double kJacobiTol = 1e-14;  // ≈ n·ε: below it round-off would re-trigger rotations

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }

// Circle ordering: round r, slot k of N/2 → the pair (L[k], L[N−1−k]) with
// L[0] = 0 and L[j] = 1 + (r + j − 1) mod (N − 1).
__device__ __forceinline__ int circle_at(int r, int j, int N) { return j == 0 ? 0 : 1 + (r + j - 1) % (N - 1); }

// grid: (nprob); block: kEigThreads.  A_in: nprob × n × n, w: nprob × n,
// V: nprob × n × n.  Dynamic LDS: n·(n + 1) double2 + 3·(N/2) double2-sized
// rotation slots + n doubles.  redo (may be NULL): only the matrices with
// redo[p] != 0 are solved, the others left as they are.
__global__ __launch_bounds__(kEigThreads) void heev_jacobi_kernel(const double2* __restrict__ A_in, int n,
                                                                  int max_sweeps, double* __restrict__ w_out,
                                                                  double2* __restrict__ V_out,
                                                                  const int* __restrict__ redo) {
  if (redo && !redo[blockIdx.x]) return;  // uniform per block
  extern __shared__ double2 lds[];
  const int ld = n + 1;
  double2* A = lds;                          // n × ld
  double2* rot = A + (long long)n * ld;      // per pair: (c, a − t|b|), (s, d + t|b|), e^{−iφ}
  const int N = n + (n & 1);
  const int P = N / 2;
  double* wdiag = reinterpret_cast<double*>(rot + 3 * P);
  __shared__ int pq[2 * kEigMaxN / 2 + 2];
  __shared__ int rotated;
  const double2* Ap = A_in + (long long)blockIdx.x * n * n;
  double2* V = V_out + (long long)blockIdx.x * n * n;
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += kEigThreads) {
    const int i = e / n, j = e - i * n;
    A[i * ld + j] = Ap[e];
    V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  // exactly Hermitian input: mirror the upper triangle, real diagonal
  for (int e = tid; e < n * n; e += kEigThreads) {
    const int i = e / n, j = e - i * n;
    if (i > j) A[i * ld + j] = zconj(A[j * ld + i]);
    if (i == j) A[i * ld + i].y = 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int r = 0; r < N - 1; ++r) {
      // rotation parameters, one thread per pair
      if (tid < P) {
        const int p0 = circle_at(r, tid, N), q0 = circle_at(r, N - 1 - tid, N);
        const int p = p0 < q0 ? p0 : q0, q = p0 < q0 ? q0 : p0;
        pq[2 * tid] = p;
        pq[2 * tid + 1] = q;
        double c = 1.0, s = 0.0, anew = 0.0, dnew = 0.0;
        double2 ph = make_double2(1.0, 0.0);
        if (q < n) {
          const double a = A[p * ld + p].x, d = A[q * ld + q].x;
          const double2 b = A[p * ld + q];
          const double ab = hypot(b.x, b.y);
          if (ab > kJacobiTol * sqrt(fabs(a * d)) && ab > 0.0) {
            const double tau = (d - a) / (2.0 * ab);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            if (t != 0.0) {
              c = 1.0 / sqrt(1.0 + t * t);
              s = t * c;
              anew = a - t * ab;
              dnew = d + t * ab;
              ph = make_double2(b.x / ab, -b.y / ab);  // e^{−iφ} = conj(b)/|b|
              rotated = 1;
            } else {  // |b| negligible next to |d − a|: drop the coupling
              A[p * ld + q] = make_double2(0.0, 0.0);
              A[q * ld + p] = make_double2(0.0, 0.0);
            }
          }
        }
        rot[3 * tid] = make_double2(c, anew);
        rot[3 * tid + 1] = make_double2(s, dnew);
        rot[3 * tid + 2] = ph;
      }
      __syncthreads();
      // columns of A and V: A[i][p], A[i][q] ← (c·x − s·e^{−iφ}·y, s·x + c·e^{−iφ}·y),
      // the same for V.  Work item (pair k, row i), all of a thread's loads
      // issued before its stores (the items touch distinct entries).
      for (int it0 = 0; it0 < kEigItems; it0 += kEigChunk) {
        double2 ax[kEigChunk], ay[kEigChunk], vx[kEigChunk], vy[kEigChunk];
#pragma unroll
        for (int it = 0; it < kEigChunk; ++it) {
          const int e = tid + (it0 + it) * kEigThreads;
          if (e < P * n) {
            const int k = e / n, i = e - k * n;
            if (rot[3 * k + 1].x != 0.0) {
              const int p = pq[2 * k], q = pq[2 * k + 1];
              ax[it] = A[i * ld + p];
              ay[it] = A[i * ld + q];
              vx[it] = V[(long long)i * n + p];
              vy[it] = V[(long long)i * n + q];
            }
          }
        }
#pragma unroll
        for (int it = 0; it < kEigChunk; ++it) {
          const int e = tid + (it0 + it) * kEigThreads;
          if (e < P * n) {
            const int k = e / n, i = e - k * n;
            const double c = rot[3 * k].x, s = rot[3 * k + 1].x;
            if (s != 0.0) {
              const int p = pq[2 * k], q = pq[2 * k + 1];
              const double2 ph = rot[3 * k + 2];
              const double2 ea = zmul(ph, ay[it]), ev = zmul(ph, vy[it]);
              A[i * ld + p] = make_double2(c * ax[it].x - s * ea.x, c * ax[it].y - s * ea.y);
              A[i * ld + q] = make_double2(s * ax[it].x + c * ea.x, s * ax[it].y + c * ea.y);
              V[(long long)i * n + p] = make_double2(c * vx[it].x - s * ev.x, c * vx[it].y - s * ev.y);
              V[(long long)i * n + q] = make_double2(s * vx[it].x + c * ev.x, s * vx[it].y + c * ev.y);
            }
          }
        }
      }
      __syncthreads();
      // rows: A[p][j], A[q][j] ← (c·x − s·e^{iφ}·y, s·x + c·e^{iφ}·y); the 2 × 2
      // block itself is set exactly: diagonal a − t|b|, d + t|b|, zero coupling
      for (int it0 = 0; it0 < kEigItems; it0 += kEigChunk) {
        double2 ax[kEigChunk], ay[kEigChunk];
#pragma unroll
        for (int it = 0; it < kEigChunk; ++it) {
          const int e = tid + (it0 + it) * kEigThreads;
          if (e < P * n) {
            const int k = e / n, j = e - k * n;
            if (rot[3 * k + 1].x != 0.0) {
              const int p = pq[2 * k], q = pq[2 * k + 1];
              ax[it] = A[p * ld + j];
              ay[it] = A[q * ld + j];
            }
          }
        }
#pragma unroll
        for (int it = 0; it < kEigChunk; ++it) {
          const int e = tid + (it0 + it) * kEigThreads;
          if (e < P * n) {
            const int k = e / n, j = e - k * n;
            const double c = rot[3 * k].x, s = rot[3 * k + 1].x;
            if (s != 0.0) {
              const int p = pq[2 * k], q = pq[2 * k + 1];
              double2 np_, nq;
              if (j == p) {
                np_ = make_double2(rot[3 * k].y, 0.0);
                nq = make_double2(0.0, 0.0);
              } else if (j == q) {
                np_ = make_double2(0.0, 0.0);
                nq = make_double2(rot[3 * k + 1].y, 0.0);
              } else {
                const double2 ey = zmul(zconj(rot[3 * k + 2]), ay[it]);
                np_ = make_double2(c * ax[it].x - s * ey.x, c * ax[it].y - s * ey.y);
                nq = make_double2(s * ax[it].x + c * ey.x, s * ax[it].y + c * ey.y);
              }
              A[p * ld + j] = np_;
              A[q * ld + j] = nq;
            }
          }
        }
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();
    if (!any) break;
  }
  // ascending order: rank of each eigenvalue (ties by index)
  for (int i = tid; i < n; i += kEigThreads) wdiag[i] = A[i * ld + i].x;
  __syncthreads();
  // reuse A's LDS as the permuted eigenvector buffer (n × n)
  double2* Vs = A;
  for (int i = tid; i < n; i += kEigThreads) {
    const double wi = wdiag[i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double wj = wdiag[j];
      rank += (wj < wi) || (wj == wi && j < i);
    }
    w_out[(long long)blockIdx.x * n + rank] = wi;
    for (int row = 0; row < n; ++row) Vs[row * ld + rank] = V[(long long)row * n + i];
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += kEigThreads) {
    const int i = e / n, j = e - i * n;
    V[e] = Vs[i * ld + j];
  }
}

}  // namespace efn
