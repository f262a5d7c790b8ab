// Batched Hermitian eigensolver for the small dense FP64 problems of LOBPCG
// (Rayleigh–Ritz matrices and Gram matrices, n ≤ 96): Householder reduction
// to a real symmetric tridiagonal matrix, implicit QL iterations with
// Wilkinson shifts on it, and the back-transformation of the requested
// lowest eigenvectors; one workgroup per matrix, everything in LDS, nothing
// allocated (the kernel can be captured in a graph).
//
//   A = Q·T·Qᴴ,  Q = H(0)·H(1)···H(n−2),  H(k) = I − τ_k v_k v_kᴴ
//   (v_k zero above row k+1, one at row k+1; H(k)ᴴ maps A's column k below
//   the diagonal onto a real multiple of e_{k+1}, so T is real);
//   T = Z·diag(λ)·Zᵀ by QL sweeps whose plane rotations accumulate into Z;
//   the eigenvectors of A are Q·Z, applied only to the selected columns.
//
// The per-matrix body is written for a team of threads (tid, nthr) with
// barriers between phases, so that the same code runs serially on a host
// thread (nthr = 1, barriers empty) for the unit tests of its arithmetic.
// On the device the QL phase splits the team: wave 0 computes the next
// sweep's rotation chain (it needs only the diagonals; its lanes test 64
// couplings at once for a split) while threads 64… apply the previous sweep's
// rotations to the rows of Z, double-buffered.
#pragma once

#include <math.h>
#include <string.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define EFN_QL_HD __host__ __device__
#else
#define EFN_QL_HD
#endif

namespace efn {
namespace ql {

constexpr int kMaxN = 96;
constexpr int kThreads = 256;
constexpr int kIterPerEig = 30;  // QL iterations allowed per eigenvalue (all of them: 30·n)
constexpr int kBackPer = 8;      // partial sums per column in the back-transformation
constexpr int kBackCols = kThreads / kBackPer;  // columns per back-transformation batch

struct cd {
  double x, y;
};
EFN_QL_HD inline cd cmk(double a, double b) {
  cd r;
  r.x = a;
  r.y = b;
  return r;
}
EFN_QL_HD inline cd cmul(cd a, cd b) { return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
EFN_QL_HD inline cd cconj(cd a) { return cmk(a.x, -a.y); }
// conj(a)·b
EFN_QL_HD inline cd cdotc(cd a, cd b) { return cmk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }

// packed lower triangle: element (i, j), i ≥ j
EFN_QL_HD inline int pk(int i, int j) { return i * (i + 1) / 2 + j; }

struct Team {
  int tid, nthr;
  EFN_QL_HD void sync() const {
#ifdef __HIP_DEVICE_COMPILE__
    __syncthreads();
#endif
  }
  // threads that apply rotations to rows of Z while thread 0 runs the chain
  EFN_QL_HD int row0() const { return nthr > 64 ? tid - 64 : tid; }
  EFN_QL_HD int rowstep() const { return nthr > 64 ? nthr - 64 : nthr; }
};

// LDS layout for launch size n (every problem of the launch has n_p ≤ n)
struct Layout {
  int n, ldz, npk;
  EFN_QL_HD explicit Layout(int n_) : n(n_), ldz(n_ + 1), npk(n_ * (n_ + 1) / 2) {}
  // byte offsets, each a multiple of 16
  EFN_QL_HD long a_off() const { return 0; }                                   // cd[npk]
  // Z: n·ldz doubles, at least 6n (one thread's inverse-iteration scratch)
  EFN_QL_HD long zsz() const {
    const long a = (long)n * ldz, b = 6L * n;
    return ((a > b ? a : b) + 1) & ~1L;
  }
  EFN_QL_HD long z_off() const { return a_off() + 16L * npk; }                 // double[zsz]
  EFN_QL_HD long d_off() const { return z_off() + 8L * zsz(); }                 // double[n]
  EFN_QL_HD long e_off() const { return d_off() + 8L * ((n + 1) & ~1); }        // double[n]
  EFN_QL_HD long tau_off() const { return e_off() + 8L * ((n + 1) & ~1); }      // cd[n]
  EFN_QL_HD long u_off() const { return tau_off() + 16L * n; }                  // cd[n] (v)
  EFN_QL_HD long y_off() const { return u_off() + 16L * n; }                    // cd[n] (τAv, w)
  EFN_QL_HD long cs_off() const { return y_off() + 16L * n; }                   // double[2][2][n]
  EFN_QL_HD long red_off() const { return cs_off() + 8L * 4 * ((n + 1) & ~1); } // cd[kThreads]
  EFN_QL_HD long ctl_off() const { return red_off() + 16L * kThreads; }         // int[16 + n]
  EFN_QL_HD long bytes() const { return ctl_off() + 4L * ((16 + n + 3) & ~3); }
};

// control words
// (QL: 0-9; bisection + inverse iteration: 2, 3, and 4-11 as two pairs of
// 64-bit words, the maxima and the split mask; both: the permutation at 16)
enum { kCtlL = 0, kCtlIter = 1, kCtlFail = 2, kCtlClus = 3, kCtlCnt = 4, kCtlLo = 6, kCtlDone = 8, kCtlMax = 4,
       kCtlMask = 8, kCtlPerm = 16 };

// Block reduction of a complex partial (every thread passes one); the sum
// is returned to every thread.  One barrier: on the device each wave sums by
// lane shuffles and the wave sums meet in red[4·slot …]; a slot is written
// again only after a later barrier (consecutive reductions alternate slots).
// On the host (one thread) the value is returned.
EFN_QL_HD inline cd block_sum(const Team& t, cd* red, cd v, int slot) {
#ifdef __HIP_DEVICE_COMPILE__
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  cd* r = red + 4 * slot;
  if ((t.tid & 63) == 0) r[t.tid >> 6] = v;
  t.sync();
  cd s = r[0];
  for (int i = 1; i < (t.nthr >> 6); ++i) {
    s.x += r[i].x;
    s.y += r[i].y;
  }
  return s;
#else
  (void)t;
  (void)red;
  (void)slot;
  return v;
#endif
}

// bit patterns of doubles, on both sides
EFN_QL_HD inline long long __double_as_longlong_hd(double v) {
#ifdef __HIP_DEVICE_COMPILE__
  return __double_as_longlong(v);
#else
  long long r;
  memcpy(&r, &v, sizeof r);
  return r;
#endif
}
EFN_QL_HD inline double __longlong_as_double_hd(long long v) {
#ifdef __HIP_DEVICE_COMPILE__
  return __longlong_as_double(v);
#else
  double r;
  memcpy(&r, &v, sizeof r);
  return r;
#endif
}

// In a 128-bit mask (two words): the highest set bit below i (−1: none) and
// the lowest set bit at or above i (the mask has one there)
EFN_QL_HD inline int last_bit_below(const unsigned long long* m, int i) {
  for (int w = (i - 1) >> 6; w >= 0; --w) {
    const int hi = i - 64 * w;  // bits [0, hi) of word w lie below i
    const unsigned long long x = hi >= 64 ? m[w] : (m[w] & ((1ull << hi) - 1ull));
    if (x) return 64 * w + 63 - __builtin_clzll(x);
  }
  return -1;
}
EFN_QL_HD inline int first_bit_from(const unsigned long long* m, int i) {
  for (int w = i >> 6; w < 2; ++w) {
    const int lo = i - 64 * w;  // bits [lo, 64) of word w lie at or above i
    const unsigned long long x = lo <= 0 ? m[w] : (m[w] & ~((1ull << lo) - 1ull));
    if (x) return 64 * w + __builtin_ctzll(x);
  }
  return 127;
}

// 1/√t for t > 0: the hardware estimate and two Newton steps on the device,
// the library square root on the host
EFN_QL_HD inline double rsqrt_d(double t) {
#ifdef __HIP_DEVICE_COMPILE__
  double y = __builtin_amdgcn_rsq(t);
  for (int k = 0; k < 2; ++k) {
    const double e = fma(-t * y, y, 1.0);  // 1 − t·y²
    y = fma(0.5 * y, e, y);
  }
  return y;
#else
  return 1.0 / sqrt(t);
#endif
}

// A condition every lane of the wave computes alike, as a wave-uniform
// (scalar) branch condition on the device
EFN_QL_HD inline bool uniform(bool v) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_readfirstlane(v ? 1 : 0) != 0;
#else
  return v;
#endif
}

// First m ≥ l with a negligible coupling sub[m] (m = n − 1 at the latest).
// On the device the 64 lanes of the calling wave test 64 indices at a time.
EFN_QL_HD inline int ql_split(int n, int l, const double* __restrict__ d, const double* __restrict__ sub, int lane) {
#ifdef __HIP_DEVICE_COMPILE__
  for (int base = l;; base += 64) {
    const int mm = base + lane;
    bool neg = mm >= n - 1;
    if (!neg) {
      const double dd = fabs(d[mm]) + fabs(d[mm + 1]);
      neg = fabs(sub[mm]) + dd == dd;
    }
    const unsigned long long mask = __ballot(neg);
    if (mask) return base + __ffsll((long long)mask) - 1;
  }
#else
  (void)lane;
  int m = l;
  for (; m < n - 1; ++m) {
    const double dd = fabs(d[m]) + fabs(d[m + 1]);
    if (fabs(sub[m]) + dd == dd) break;
  }
  return m;
#endif
}

// One QL pass of the rotation chain on (d, sub) (sub[i] couples d[i] and
// d[i+1], sub[n−1] = 0): advances the deflation index until one implicit QL
// iteration has run, recording its rotations (c_i, s_i) for i = hi … lo in
// cs_c[i], cs_s[i] and their range in ctl, or until every eigenvalue has
// converged (done).  Run by one whole wave on the device (every lane the
// same values, lane 0 stores; LDS accesses of one wave stay in order) and by
// the one thread on the host.  Every d and sub a rotation reads is still the
// sweep's old value, so the next rotation's operands are loaded ahead.
EFN_QL_HD inline void ql_chain(int n, double* __restrict__ d, double* __restrict__ sub, double* __restrict__ cs_c,
                               double* __restrict__ cs_s, int* __restrict__ ctl, int buf, int lane) {
  const bool w0 = lane == 0;
  int l = ctl[kCtlL];
  int iter = ctl[kCtlIter];
  int cnt = 0, lo = 0, done = 0;
  while (true) {
    if (l >= n) {
      done = 1;
      break;
    }
    const int m = ql_split(n, l, d, sub, lane);
    if (m == l) {  // d[l] converged
      ++l;
      continue;
    }
    if (++iter > kIterPerEig * n) {
      if (w0) ctl[kCtlFail] = 1 + l;
      done = 1;
      break;
    }
    // Wilkinson shift from the leading 2 × 2 block at l
    const double dl = d[l];
    double g = (d[l + 1] - dl) / (2.0 * sub[l]);
    const double r0 = sqrt(g * g + 1.0);
    g = d[m] - dl + sub[l] / (g + (g >= 0.0 ? r0 : -r0));
    double s = 1.0, c = 1.0, p = 0.0;
    double di1 = d[m];                    // d[i+1], old
    double di = d[m - 1], ei = sub[m - 1];  // d[i], sub[i], old
    int i = m - 1;
    bool split = false;
    for (; i >= l; --i) {
      double di_n = 0.0, ei_n = 0.0;
      if (i > l) {
        di_n = d[i - 1];
        ei_n = sub[i - 1];
      }
      const double f = s * ei, b = c * ei;
      const double tt = f * f + g * g;
      if (uniform(tt == 0.0)) {  // the rotation vanished: split here and start again
        if (w0) {
          sub[i + 1] = 0.0;
          d[i + 1] = di1 - p;
          sub[m] = 0.0;
        }
        split = true;
        break;
      }
      const double y = rsqrt_d(tt);
      s = f * y;
      c = g * y;
      const double g2 = di1 - p;
      const double r2 = (di - g2) * s + 2.0 * c * b;
      p = s * r2;
      if (w0) {
        sub[i + 1] = tt * y;
        d[i + 1] = g2 + p;
        cs_c[i] = c;
        cs_s[i] = s;
      }
      g = c * r2 - b;
      di1 = di;
      di = di_n;
      ei = ei_n;
    }
    // rotations recorded for indices i+1 … m−1 (i = l − 1 when complete)
    cnt = m - 1 - i;
    lo = i + 1;
    if (!split && w0) {
      d[l] = dl - p;
      sub[l] = g;
      sub[m] = 0.0;
    }
    if (cnt > 0) break;
  }
  if (w0) {
    ctl[kCtlL] = l;
    ctl[kCtlIter] = iter;
    ctl[kCtlCnt + buf] = cnt;
    ctl[kCtlLo + buf] = lo;
    ctl[kCtlDone + buf] = done;
  }
}

// Z ← Z·(rotations of one QL iteration) on row r: for i = hi … lo the pair
// of columns (i, i+1) turns by (c_i, s_i); the new column i carries down.
EFN_QL_HD inline void ql_apply_row(double* __restrict__ zr, const double* __restrict__ cs_c,
                                   const double* __restrict__ cs_s, int lo, int cnt) {
  if (cnt <= 0) return;
  const int hi = lo + cnt - 1;
  double carry = zr[hi + 1];
  double zi = zr[hi], c = cs_c[hi], s = cs_s[hi];
  for (int i = hi; i >= lo; --i) {
    double zn = 0.0, cn = 0.0, sn = 0.0;
    if (i > lo) {
      zn = zr[i - 1];
      cn = cs_c[i - 1];
      sn = cs_s[i - 1];
    }
    zr[i + 1] = s * zi + c * carry;
    carry = c * zi - s * carry;
    zi = zn;
    c = cn;
    s = sn;
  }
  zr[lo] = carry;
}

// Σ_j A₂₂(i, j)·v_j over the columns j ≡ q (mod G) of row i of the trailing
// block at k+1 (size m), A₂₂ Hermitian from the packed lower triangle (A(i, j)
// = conj(A(j, i)) above the diagonal, the diagonal real); four entries loaded
// ahead of their products, alternate entries into two sums added at the end
EFN_QL_HD inline void hemv_part(const cd* __restrict__ a, const cd* __restrict__ u, int k, int m, int i, int q,
                                int G, double& sx, double& sy) {
  const int gi = k + 1 + i;
  const int row = pk(gi, k + 1);
  double ax[2] = {0.0, 0.0}, ay[2] = {0.0, 0.0};
  for (int j0 = q; j0 < m; j0 += 4 * G) {
    cd av[4], uv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = j0 + r * G;
      if (j < m) {
        av[r] = a[j <= i ? row + j : pk(k + 1 + j, gi)];
        uv[r] = u[j];
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = j0 + r * G;
      if (j < m) {
        const double re = av[r].x, im = j < i ? av[r].y : (j == i ? 0.0 : -av[r].y);
        ax[r & 1] += re * uv[r].x - im * uv[r].y;
        ay[r & 1] += re * uv[r].y + im * uv[r].x;
      }
    }
  }
  sx += ax[0] + ax[1];
  sy += ay[0] + ay[1];
}

// y = τ·A₂₂·v over the trailing block at k+1 (size m).  G threads a row —
// the largest power of two up to 16 with G·m ≤ kThreads (2m ≤ kThreads for
// m < kMaxN) — adjacent lanes of one wave, each summing the columns j ≡ q
// (mod G); the partials meet by butterfly lane shuffles ((S₀+S₁)+(S₂+S₃))+….
// The host sums the same partials in the same order.
EFN_QL_HD inline int hemv_lanes(int m) { return 16 * m <= kThreads ? 16 : 8 * m <= kThreads ? 8 : 4 * m <= kThreads ? 4 : 2; }
// Returns this thread's part of yᴴv (Σ conj(y_i)·v_i over its rows; on the
// device the q = 0 lane of each row).
EFN_QL_HD inline cd ql_hemv(const Team& t, const cd* __restrict__ a, const cd* __restrict__ u, cd* __restrict__ y,
                            int k, int m, cd tk) {
  const int G = hemv_lanes(m);
#ifdef __HIP_DEVICE_COMPILE__
  const int i = t.tid / G, q = t.tid % G;
  double sx = 0.0, sy = 0.0;
  if (i < m) hemv_part(a, u, k, m, i, q, G, sx, sy);
  for (int off = 1; off < G; off <<= 1) {
    sx += __shfl_xor(sx, off, 64);
    sy += __shfl_xor(sy, off, 64);
  }
  cd part = cmk(0.0, 0.0);
  if (i < m && q == 0) {
    const cd yi = cmul(tk, cmk(sx, sy));
    y[i] = yi;
    part = cdotc(yi, u[i]);
  }
  return part;
#else
  for (int i = t.tid; i < m; i += t.nthr) {
    double px[16], py[16];
    for (int q = 0; q < G; ++q) px[q] = py[q] = 0.0;
    for (int q = 0; q < G; ++q) hemv_part(a, u, k, m, i, q, G, px[q], py[q]);
    for (int off = 1; off < G; off <<= 1)
      for (int q = 0; q < G; q += 2 * off) {
        px[q] += px[q + off];
        py[q] += py[q + off];
      }
    y[i] = cmul(tk, cmk(px[0], py[0]));
  }
  cd part = cmk(0.0, 0.0);
  for (int i = t.tid; i < m; i += t.nthr) {
    const cd c = cdotc(y[i], u[i]);
    part.x += c.x;
    part.y += c.y;
  }
  return part;
#endif
}

// A₂₂ ← A₂₂ − v·wᴴ − w·vᴴ over the trailing lower triangle (size m), the
// diagonal kept real.  Rows p and m−1−p are dealt as one run of m+1 entries
// (row p's p+1, then row m−1−p's m−p), the runs evenly over the team with
// incremental indices (for odd m the middle row's run stops at its length).
// Returns this thread's part of the next step's column norm: Σ|A₂₂(i, 0)|²
// over its updated entries with i ≥ 2 (global rows ≥ k+3 of column k+1).
EFN_QL_HD inline cd ql_her2(const Team& t, cd* __restrict__ a, const cd* __restrict__ u, const cd* __restrict__ w,
                            int k, int m) {
  cd part = cmk(0.0, 0.0);
  const int L = m + 1, tot = ((m + 1) / 2) * L;
  int p = t.tid / L, r = t.tid % L;
  const int dp = t.nthr / L, dr = t.nthr % L;
  for (int q = t.tid; q < tot; q += t.nthr) {
    const bool first = r <= p;
    const int i = first ? p : m - 1 - p, j = first ? r : r - p - 1;
    if (first || i != p) {
      const cd vi = u[i], wi = w[i], vj = u[j], wj = w[j];
      cd x = a[pk(k + 1 + i, k + 1 + j)];
      // v_i·conj(w_j) + w_i·conj(v_j)
      x.x -= (vi.x * wj.x + vi.y * wj.y) + (wi.x * vj.x + wi.y * vj.y);
      x.y -= (vi.y * wj.x - vi.x * wj.y) + (wi.y * vj.x - wi.x * vj.y);
      if (i == j) x.y = 0.0;
      a[pk(k + 1 + i, k + 1 + j)] = x;
      if (j == 0 && i >= 2) part.x += x.x * x.x + x.y * x.y;
    }
    p += dp;
    r += dr;
    if (r >= L) {
      r -= L;
      ++p;
    }
  }
  return part;
}

// The LDS arrays of one matrix (offsets from Layout)
struct Work {
  cd* a;        // packed lower triangle; Householder vectors below the subdiagonal
  double* z;    // rotation accumulator / scratch / back-transformation staging
  double* d;    // tridiagonal: diagonal
  double* sub;  // tridiagonal: subdiagonal
  cd* tau;
  cd* u;
  cd* y;
  double* cs;
  cd* red;
  int* ctl;
  int ldz, csn;
};
EFN_QL_HD inline Work carve(int nl, char* lds) {
  const Layout L(nl);
  Work W;
  W.a = reinterpret_cast<cd*>(lds + L.a_off());
  W.z = reinterpret_cast<double*>(lds + L.z_off());
  W.d = reinterpret_cast<double*>(lds + L.d_off());
  W.sub = reinterpret_cast<double*>(lds + L.e_off());
  W.tau = reinterpret_cast<cd*>(lds + L.tau_off());
  W.u = reinterpret_cast<cd*>(lds + L.u_off());
  W.y = reinterpret_cast<cd*>(lds + L.y_off());
  W.cs = reinterpret_cast<double*>(lds + L.cs_off());
  W.red = reinterpret_cast<cd*>(lds + L.red_off());
  W.ctl = reinterpret_cast<int*>(lds + L.ctl_off());
  W.ldz = L.ldz;
  W.csn = (nl + 1) & ~1;
  return W;
}

// Load A's lower triangle (rows and columns in the order rowmap gives, if
// any; Z = I when z_identity) and reduce it to the real tridiagonal (d, sub)
// by Householder reflectors (LAPACK zhetd2's steps); τ_k in tau, v_k below
// the subdiagonal of column k of a.
EFN_QL_HD inline void reduce_tridiag(const Team& t, const cd* A, int lda, int n, const Work& W, bool z_identity,
                                     const int* rowmap = nullptr) {
  cd* a = W.a;
  double* z = W.z;
  double* d = W.d;
  double* sub = W.sub;
  cd* tau = W.tau;
  cd* u = W.u;
  cd* y = W.y;
  const int ldz = W.ldz;
  for (int i = t.tid; i < n; i += t.nthr) {
    if (rowmap) {  // row i of the permuted matrix is row rowmap[i] of A
      const int ri = rowmap[i];
      for (int j = 0; j <= i; ++j) {
        const int rj = rowmap[j];
        a[pk(i, j)] = ri >= rj ? A[(long)ri * lda + rj] : cconj(A[(long)rj * lda + ri]);
      }
    } else {
      for (int j = 0; j <= i; ++j) a[pk(i, j)] = A[(long)i * lda + j];
    }
    if (z_identity)
      for (int j = 0; j < n; ++j) z[i * ldz + j] = i == j ? 1.0 : 0.0;
  }
  if (t.tid == 0) {
    for (int c = 0; c < 16; ++c) W.ctl[c] = 0;
  }
  t.sync();
  // four barriers a step: the column norm's reduction (whose partials the
  // previous step's rank-2 update leaves, or a loop after a skipped update),
  // v written, yᴴv's reduction (fused with A₂₂·v), w written
  cd npart = cmk(0.0, 0.0);
  bool have = false;
  for (int k = 0; k + 1 < n; ++k) {
    const int m = n - k - 1;  // rows k+1 … n−1 of column k
    // ‖column below row k+1‖²
    if (!have) {
      npart = cmk(0.0, 0.0);
      for (int i = k + 2 + t.tid; i < n; i += t.nthr) {
        const cd x = a[pk(i, k)];
        npart.x += x.x * x.x + x.y * x.y;
      }
    }
    const double xn2 = block_sum(t, W.red, npart, 0).x;
    have = false;
    const cd alpha = a[pk(k + 1, k)];
    cd tk = cmk(0.0, 0.0), scale = cmk(0.0, 0.0);
    double beta = alpha.x;
    if (!(xn2 == 0.0 && alpha.y == 0.0)) {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      beta = alpha.x >= 0.0 ? -nrm : nrm;
      tk = cmk((beta - alpha.x) / beta, -alpha.y / beta);
      // 1 / (α − β)
      const cd den = cmk(alpha.x - beta, alpha.y);
      const double dn = den.x * den.x + den.y * den.y;
      scale = cmk(den.x / dn, -den.y / dn);
    }
    // v (v[0] = 1) into u and into column k below row k+1 (kept for the
    // back-transformation)
    for (int i = t.tid; i < m; i += t.nthr) {
      if (i == 0) {
        u[0] = cmk(1.0, 0.0);
      } else {
        const cd v = cmul(a[pk(k + 1 + i, k)], scale);
        u[i] = v;
        a[pk(k + 1 + i, k)] = v;
      }
    }
    if (t.tid == 0) {
      tau[k] = tk;
      sub[k] = beta;
      d[k] = a[pk(k, k)].x;
    }
    t.sync();
    if (tk.x != 0.0 || tk.y != 0.0) {
      // y = τ·A₂₂·v over the trailing block (rows and columns k+1 …), and yᴴv
      const cd yv = block_sum(t, W.red, ql_hemv(t, a, u, y, k, m, tk), 1);
      // w = y − ½·τ·(yᴴv)·v
      const cd al = cmul(cmk(-0.5 * tk.x, -0.5 * tk.y), yv);
      for (int i = t.tid; i < m; i += t.nthr) {
        const cd c = cmul(al, u[i]);
        y[i] = cmk(y[i].x + c.x, y[i].y + c.y);
      }
      t.sync();
      // A₂₂ ← A₂₂ − v·wᴴ − w·vᴴ (lower triangle, real diagonal); no barrier
      // after it: the next step's first reduction orders its writes
      npart = ql_her2(t, a, u, y, k, m);
      have = true;
    }
  }
  t.sync();
  if (t.tid == 0 && n > 0) {
    d[n - 1] = a[pk(n - 1, n - 1)].x;
    sub[n - 1] = 0.0;
  }
  t.sync();
}

// Reflector k's vector at global row g: zero above k+1, one at k+1, below
// it the entries kept in column k of a
EFN_QL_HD inline cd refl(const cd* a, int k, int g) {
  return g <= k ? cmk(0.0, 0.0) : (g == k + 1 ? cmk(1.0, 0.0) : a[pk(g, k)]);
}

// V ← Q·V for the first `low` columns of V (lda × lowest row-major; on entry
// the real eigenvectors of T in rows < n, zeros elsewhere; row i stored to
// row rowmap[i] when a map is given), in batches of up to kBackCols columns
// staged over Z.  The reflectors apply in pairs (k, k−1),
// k = n−2, n−4, …: with a = v_k, b = v_{k−1} and c = bᴴa (one per pair,
// computed first), α = τ_k·aᴴy, β = τ_{k−1}·(bᴴy − α·c) and y ← y − a·α − b·β
// is H(k−1)·H(k)·y from one pass over y.  A column is owned by kBackPer
// adjacent lanes of one wave: lane q sums the rows ≡ q (mod kBackPer) of the
// pair's range and updates the same rows, the partials meeting by butterfly
// shuffles ((p₀+p₁)+(p₂+p₃))+((p₄+p₅)+(p₆+p₇)) — so no workgroup barrier
// separates the pairs, and a column's result does not depend on the batch
// it is in (a subset of columns equals the full solve's).
EFN_QL_HD inline void back_transform(const Team& t, const Work& W, int n, int nl, int lda, int lowest, int low,
                                     cd* V, const int* rowmap = nullptr) {
  (void)lda;
  static_assert(kBackPer == 8 && kBackCols * kBackPer == kThreads, "one column per 8 lanes");
  const cd* a = W.a;
  const cd* tau = W.tau;
  cd* cpair = W.u;  // c of the pair (k, k−1) at cpair[k]
  for (int k = n - 2 - 2 * t.tid; k >= 1; k -= 2 * t.nthr) {
    cd c = cmk(0.0, 0.0);
    for (int g = k + 1; g < n; ++g) {
      const cd x = cdotc(refl(a, k - 1, g), refl(a, k, g));
      c.x += x.x;
      c.y += x.y;
    }
    cpair[k] = c;
  }
#ifdef __HIP_DEVICE_COMPILE__
  __threadfence_block();
#endif
  t.sync();
  cd* yb = reinterpret_cast<cd*>(W.z);
  int bmax = (nl * W.ldz) / (2 * (n > 0 ? n : 1));
  if (bmax > kBackCols) bmax = kBackCols;
  constexpr int kRows = (kMaxN + kBackPer - 1) / kBackPer;  // rows a lane owns at most
  for (int c0 = 0; c0 < low; c0 += bmax) {
    const int nb = low - c0 < bmax ? low - c0 : bmax;
    for (int idx = t.tid; idx < n * nb; idx += t.nthr) {
      const int i = idx / nb, c = idx % nb;
      yb[i * nb + c] = V[(long)i * lowest + c0 + c];
    }
    t.sync();
#ifdef __HIP_DEVICE_COMPILE__
    const int c = t.tid / kBackPer, q = t.tid % kBackPer;
    const bool act = c < nb;
    for (int k = n - 2; k >= 0; k -= 2) {
      const cd ta = tau[k], tb = k >= 1 ? tau[k - 1] : cmk(0.0, 0.0);
      if (ta.x == 0.0 && ta.y == 0.0 && tb.x == 0.0 && tb.y == 0.0) continue;  // uniform: tau is in LDS
      const int r0 = k >= 1 ? k : k + 1;  // first row either vector reaches
      cd va[kRows], vb[kRows], yy[kRows];
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const int g = r0 + q + r * kBackPer;
        if (act && g < n) {
          va[r] = refl(a, k, g);
          vb[r] = k >= 1 ? refl(a, k - 1, g) : cmk(0.0, 0.0);
          yy[r] = yb[g * nb + c];
        }
      }
      double sax = 0.0, say = 0.0, sbx = 0.0, sby = 0.0;
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        if (act && r0 + q + r * kBackPer < n) {
          const cd xa = cdotc(va[r], yy[r]), xb = cdotc(vb[r], yy[r]);
          sax += xa.x;
          say += xa.y;
          sbx += xb.x;
          sby += xb.y;
        }
      for (int off = 1; off < kBackPer; off <<= 1) {
        sax += __shfl_xor(sax, off, 64);
        say += __shfl_xor(say, off, 64);
        sbx += __shfl_xor(sbx, off, 64);
        sby += __shfl_xor(sby, off, 64);
      }
      const cd al = cmul(ta, cmk(sax, say));
      const cd cc = k >= 1 ? cpair[k] : cmk(0.0, 0.0);
      const cd ac = cmul(al, cc);
      const cd be = cmul(tb, cmk(sbx - ac.x, sby - ac.y));
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const int g = r0 + q + r * kBackPer;
        if (act && g < n) {
          const cd fa = cmul(va[r], al), fb = cmul(vb[r], be);
          yb[g * nb + c] = cmk(yy[r].x - fa.x - fb.x, yy[r].y - fa.y - fb.y);
        }
      }
    }
#else
    for (int c = 0; c < nb; ++c)
      for (int k = n - 2; k >= 0; k -= 2) {
        const cd ta = tau[k], tb = k >= 1 ? tau[k - 1] : cmk(0.0, 0.0);
        if (ta.x == 0.0 && ta.y == 0.0 && tb.x == 0.0 && tb.y == 0.0) continue;
        const int r0 = k >= 1 ? k : k + 1;
        double sx[2][kBackPer], sy[2][kBackPer];
        for (int q = 0; q < kBackPer; ++q) {
          sx[0][q] = sy[0][q] = sx[1][q] = sy[1][q] = 0.0;
          for (int g = r0 + q; g < n; g += kBackPer) {
            const cd y = yb[g * nb + c];
            const cd xa = cdotc(refl(a, k, g), y);
            const cd xb = cdotc(k >= 1 ? refl(a, k - 1, g) : cmk(0.0, 0.0), y);
            sx[0][q] += xa.x;
            sy[0][q] += xa.y;
            sx[1][q] += xb.x;
            sy[1][q] += xb.y;
          }
        }
        for (int v = 0; v < 2; ++v)
          for (int off = 1; off < kBackPer; off <<= 1)
            for (int q = 0; q < kBackPer; q += 2 * off) {
              sx[v][q] += sx[v][q + off];
              sy[v][q] += sy[v][q + off];
            }
        const cd al = cmul(ta, cmk(sx[0][0], sy[0][0]));
        const cd cc = k >= 1 ? cpair[k] : cmk(0.0, 0.0);
        const cd ac = cmul(al, cc);
        const cd be = cmul(tb, cmk(sx[1][0] - ac.x, sy[1][0] - ac.y));
        for (int g = r0; g < n; ++g) {
          const cd fa = cmul(refl(a, k, g), al);
          const cd fb = cmul(k >= 1 ? refl(a, k - 1, g) : cmk(0.0, 0.0), be);
          cd& x = yb[g * nb + c];
          x.x -= fa.x + fb.x;
          x.y -= fa.y + fb.y;
        }
      }
#endif
    t.sync();
    for (int idx = t.tid; idx < n * nb; idx += t.nthr) {
      const int i = idx / nb, cc = idx % nb;
      V[(long)(rowmap ? rowmap[i] : i) * lowest + c0 + cc] = yb[i * nb + cc];
    }
    t.sync();
  }
}

// The QL solver for one matrix.  A: n_p × n_p of a row-major array with
// leading dimension lda (only the lower triangle is read); w: `lowest` values
// (ascending; +inf past min(lowest, n_p)); V: lda × lowest row-major
// (V[i·lowest + j] = component i of eigenvector j; zero past n_p rows and
// min(lowest, n_p) columns); info: 0, or 1 + l if the QL iteration for
// eigenvalue l did not converge (then every w is NaN).  lds:
// Layout(nl).bytes() bytes, nl ≥ n_p.
EFN_QL_HD inline void heev_ql_body(const Team& t, const cd* A, int lda, int n, int nl, int lowest, double* w,
                                   cd* V, int* info, char* lds) {
  const Work W = carve(nl, lds);
  double* z = W.z;
  double* d = W.d;
  double* sub = W.sub;
  double* cs = W.cs;
  int* ctl = W.ctl;
  const int ldz = W.ldz;
  const int csn = W.csn;
  const int low = lowest < n ? lowest : n;
  const double inf = HUGE_VAL;

  reduce_tridiag(t, A, lda, n, W, true);

  // ---- implicit QL on (d, sub), rotations into Z -----------------------------
  // buffer b: c at cs[(2b)·csn + i], s at cs[(2b + 1)·csn + i]
  const bool chain = t.nthr > 64 ? t.tid < 64 : t.tid == 0;  // wave 0 on the device
  if (chain) ql_chain(n, d, sub, cs, cs + csn, ctl, 0, t.tid);
  t.sync();
  for (int b = 0;; b ^= 1) {
    const int cnt = ctl[kCtlCnt + b], lo = ctl[kCtlLo + b], fin = ctl[kCtlDone + b];
    if (chain && !fin) ql_chain(n, d, sub, cs + 2 * (b ^ 1) * csn, cs + (2 * (b ^ 1) + 1) * csn, ctl, b ^ 1, t.tid);
    for (int r = t.row0(); r >= 0 && r < n; r += t.rowstep())
      ql_apply_row(z + r * ldz, cs + 2 * b * csn, cs + (2 * b + 1) * csn, lo, cnt);
    t.sync();
    if (fin) break;
  }

  // ---- the lowest `low` eigenvalues: rank of each (ties by index) ------------
  int* perm = ctl + kCtlPerm;
  for (int j = t.tid; j < n; j += t.nthr) {
    int rank = 0;
    const double dj = d[j];
    for (int i = 0; i < n; ++i) rank += (d[i] < dj) || (d[i] == dj && i < j);
    if (rank < low) perm[rank] = j;
  }
  t.sync();
  // a problem whose QL iteration failed reports NaN eigenvalues (visible
  // without reading info back)
  const bool failed = ctl[kCtlFail] != 0;
  for (int r = t.tid; r < lowest; r += t.nthr) w[r] = failed ? NAN : (r < low ? d[perm[r]] : inf);
  if (t.tid == 0) *info = ctl[kCtlFail];

  // ---- V = Q·Z[:, perm]: Z's selected columns to V (real), then turned ------
  for (long idx = t.tid; idx < (long)lda * lowest; idx += t.nthr) {
    const int i = (int)(idx / lowest), r = (int)(idx % lowest);
    V[idx] = (i < n && r < low) ? cmk(z[i * ldz + perm[r]], 0.0) : cmk(0.0, 0.0);
  }
  back_transform(t, W, n, nl, lda, lowest, low, V);
}

// ============================================================================
// Selected eigenpairs by bisection and inverse iteration (the scheme of
// LAPACK's zheevx, i.e. dstebz + zstein, after the same reduction):
//   1. T splits where a coupling is negligible (e² ≤ ulp²·|d_i·d_{i+1}| +
//      safmin) into unreduced blocks;
//   2. every eigenvalue of every block by multisection on Sturm counts: slot
//      s (row s of T) holds the (s − b0)-th eigenvalue of the block [b0, b1)
//      containing row s, searched by bisect_lanes(n) adjacent threads at
//      kSturmPts points each per step; the lowest `low` over all blocks are
//      selected (ties by slot);
//   3. per selected eigenvalue, inverse iteration on T_b − σI (LU with partial
//      pivoting, random start, at most kInvIts solves, stopped kInvExtra
//      solves after the iterate's growth shows σ is an eigenvalue to working
//      precision), one thread per vector; then the vectors of each cluster
//      (consecutive eigenvalues of a block closer than 10⁻³·‖T_b‖₁) made
//      orthonormal by modified Gram–Schmidt, one thread per cluster, one
//      more solve from each, and Gram–Schmidt again;
//   4. the same back-transformation.
// Every quantity of eigenpair s depends on the matrix, s and n_p only — not on
// `lowest`, the launch size or the thread that computes it — so a lowest-k
// solve returns the leading columns of the full one bit for bit.

constexpr double kUlp = 2.220446049250313e-16;      // 2⁻⁵²
constexpr double kSafMin = 2.2250738585072014e-308;  // smallest normal
constexpr int kMaxBisect = 128;                      // multisection steps (≥ 3 bits each)
constexpr int kInvIts = 5;                           // inverse-iteration solves at most
constexpr int kInvExtra = 1;                         // solves after the growth test first passes

// threads that search one eigenvalue: a function of the problem size only
EFN_QL_HD inline int bisect_lanes(int n) {
  const int g = kThreads / (n > 0 ? n : 1);
  return g > 16 ? 16 : g;
}

// Inverse-iteration scratch: per thread and row four doubles (1/u0, u1, l, x)
// and one int (the interchange; U's second superdiagonal is then the
// coupling below, read from T).  It lives in a region after the QL
// arrays, sized for kInvSlots threads of the launch size where the LDS has
// room, else over Z.
constexpr int kInvRow = 36;
constexpr int kInvSlots = 32;
constexpr long kLdsMax = 163840 - 256;  // dynamic LDS beside __syncthreads_or's own 256 B
EFN_QL_HD inline long heevx_extra_bytes(int nl) {
  const Layout L(nl);
  const long want = (long)kInvRow * nl * kInvSlots, inz = 8L * L.zsz();
  if (want <= inz) return 0;
  const long room = (kLdsMax - L.bytes()) & ~15L;
  return want < room ? want : (room > inz ? room : 0);
}
EFN_QL_HD inline long heevx_lds_bytes(int nl) { return Layout(nl).bytes() + heevx_extra_bytes(nl); }

// threads with scratch for problems of n rows, and where it starts
EFN_QL_HD inline int invit_slots(int n, int nl) {
  const long xb = heevx_extra_bytes(nl);
  const long k = (xb > 0 ? xb : 8L * Layout(nl).zsz()) / ((long)kInvRow * (n > 0 ? n : 1));
  return k > 64 ? 64 : (int)k;
}
EFN_QL_HD inline double* invit_scratch(int nl, char* lds, const Work& W) {
  return heevx_extra_bytes(nl) > 0 ? reinterpret_cast<double*>(lds + Layout(nl).bytes()) : W.z;
}

constexpr int kChunk = 8;  // rows whose operands a sequential sweep loads at once

// 1/t: the hardware estimate and two Newton steps on the device
EFN_QL_HD inline double rcp_d(double t) {
#ifdef __HIP_DEVICE_COMPILE__
  double y = __builtin_amdgcn_rcp(t);
  for (int k = 0; k < 2; ++k) {
    const double e = fma(-t, y, 1.0);
    y = fma(y, e, y);
  }
  return y;
#else
  return 1.0 / t;
#endif
}

// Sturm counts of the block [b0, b1) at kSturmPts points xs[0..): the sign
// changes of the leading minors p_i = det(T_{b0..i} − x), p_i = (d_i − x)·p_{i−1}
// − e²_i·p_{i−2} (p_{b0−1} = 1, p_{b0−2} = 0; e²_i couples rows i−1 and i) —
// the number of eigenvalues below x, the negative pivots of T_b − x = LDLᵀ —
// on a copy of T scaled by an exact power of two (ds, ee, xs = x·2^−σ, so
// |entries| ≤ 1), the pair renormalised by p's exponent every kChunk rows.  A
// minor that is exactly zero counts with its sign bit (+0 as positive, the
// usual rule for zeros of a Sturm sequence).  Two dependent operations a row,
// no division, no branch; the points' chains interleave.
constexpr int kSturmPts = 4;
EFN_QL_HD inline void sturm_step(double dv, double ev, const double* xs, double* p1, double* p2, int* c) {
#pragma unroll
  for (int u = 0; u < kSturmPts; ++u) {
    const double pn = fma(-ev, p2[u], (dv - xs[u]) * p1[u]);
    c[u] += (pn < 0.0) != (p1[u] < 0.0);
    p2[u] = p1[u];
    p1[u] = pn;
  }
}
EFN_QL_HD inline void sturm_counts(const double* __restrict__ ds, const double* __restrict__ ee, int b0, int b1,
                                   const double* xs, int* cnt) {
  double p1[kSturmPts], p2[kSturmPts];
  int c[kSturmPts];
#pragma unroll
  for (int u = 0; u < kSturmPts; ++u) {
    p1[u] = 1.0;
    p2[u] = 0.0;
    c[u] = 0;
  }
  int i = b0;
  for (; i + kChunk <= b1; i += kChunk) {
    double dv[kChunk], ev[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      dv[j] = ds[i + j];
      ev[j] = ee[i + j];
    }
#pragma unroll
    for (int j = 0; j < kChunk; ++j) sturm_step(dv[j], ev[j], xs, p1, p2, c);
#pragma unroll
    for (int u = 0; u < kSturmPts; ++u) {
      int e;
      (void)frexp(p1[u], &e);
      p1[u] = ldexp(p1[u], -e);
      p2[u] = ldexp(p2[u], -e);
    }
  }
  for (; i < b1; ++i) sturm_step(ds[i], ee[i], xs, p1, p2, c);
#pragma unroll
  for (int u = 0; u < kSturmPts; ++u) cnt[u] = c[u];
}

// Gershgorin interval of the block, widened as LAPACK dstebz does, and the
// absolute tolerance 8·ulp·‖T_b‖ (midpoints within 4 ulp·‖T_b‖)
EFN_QL_HD inline void bisect_init(const double* d, const double* sub, int b0, int b1, double pivmin, double& lo,
                                  double& hi, double& atol) {
  double gl = d[b0], gu = d[b0];
  for (int r = b0; r < b1; ++r) {
    double rad = 0.0;
    if (r > b0) rad += fabs(sub[r - 1]);
    if (r + 1 < b1) rad += fabs(sub[r]);
    gl = fmin(gl, d[r] - rad);
    gu = fmax(gu, d[r] + rad);
  }
  const double tnorm = fmax(fabs(gl), fabs(gu));
  const double fud = 2.1 * kUlp * tnorm * (b1 - b0) + 4.2 * pivmin;
  lo = gl - fud;
  hi = gu + fud;
  atol = 8.0 * kUlp * tnorm;
}

// multisection point k of K inside (lo, hi)
EFN_QL_HD inline double bisect_point(double lo, double hi, int k, int K) {
  return lo + (hi - lo) * ((double)k / (double)(K + 1));
}

EFN_QL_HD inline bool bisect_done(double lo, double hi, double atol, double pivmin) {
  const double tol = fmax(fmax(atol, pivmin), 2.0 * kUlp * fmax(fabs(lo), fabs(hi)));
  return hi - lo <= tol;
}

// the largest of the points k0 … k0 + kSturmPts − 1 whose count is ≤ j (0:
// none); sc = 2^−σ, the scale of ds and ee
EFN_QL_HD inline int bisect_pts(const double* ds, const double* ee, int b0, int b1, double lo, double hi, int k0,
                                int K, int j, double sc) {
  double xs[kSturmPts];
  int cnt[kSturmPts];
#pragma unroll
  for (int u = 0; u < kSturmPts; ++u) xs[u] = bisect_point(lo, hi, k0 + u, K) * sc;
  sturm_counts(ds, ee, b0, b1, xs, cnt);
  int best = 0;
#pragma unroll
  for (int u = 0; u < kSturmPts; ++u)
    if (cnt[u] <= j) best = k0 + u;
  return best;
}

// pseudo-random start entry in [−1, 1) for row i of the j-th vector of a block
EFN_QL_HD inline double start_entry(int j, int i) {
  unsigned h = (unsigned)(j + 1) * 0x9E3779B1u + (unsigned)(i + 1) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return (double)h * 0x1.0p-31 - 1.0;
}

// Inverse iteration for the selected eigenvalues of one cluster (slots lead …
// last of one block) by one thread with scratch column c of Kc: the unit
// vector of slot s into vs[rank[s]·vr + g·vi] for the block's rows g (the
// rest stays as it was, zero), each made orthogonal to the cluster's earlier
// vectors there after every solve.  With refine, one solve from the vector
// already there instead (no stopping test).  Returns the number of vectors
// whose iteration did not pass the stopping test (they are kept).
EFN_QL_HD inline int invit_cluster(const double* __restrict__ d, const double* __restrict__ sub,
                                   const double* __restrict__ ev, const double* __restrict__ onen,
                                   const int* __restrict__ bs, const int* __restrict__ be,
                                   const int* __restrict__ rank, int lead, int last, double* __restrict__ sc,
                                   int nrow, int c, int Kc, double* vs, long vr, long vi, bool refine = false) {
  // scratch: array a ∈ {1/u0, u1, l, x} at sc[(a·nrow + i)·Kc + c], piv after
  // (with an interchange at row k, U's second superdiagonal there is T's
  // coupling sub[k+1], otherwise zero); the sweeps load kChunk rows'
  // operands at a time ahead of their use
  double* __restrict__ ui = sc;
  double* __restrict__ u1 = sc + (long)nrow * Kc;
  double* __restrict__ lm = sc + 2L * nrow * Kc;
  double* __restrict__ X = sc + 3L * nrow * Kc;
  int* __restrict__ pv = reinterpret_cast<int*>(sc + 4L * nrow * Kc);
#define EFN_S(i) ((long)(i) * Kc + c)
  int fails = 0;
  double sigm = 0.0;
  for (int s = lead; s <= last; ++s) {
    const int b0 = bs[s], nb = be[s] - b0, r = rank[s];
    if (nb == 1) {
      vs[r * vr + b0 * vi] = 1.0;
      continue;
    }
    // shift: the eigenvalue, kept 10·ulp·|σ| above the previous one of the cluster
    double sig = ev[s];
    if (s > lead) {
      const double pert = 10.0 * fabs(kUlp * sig);
      if (sig - sigm < pert) sig = sigm + pert;
    }
    sigm = sig;
    // LU of T_b − σI with partial pivoting: row k of U is (u0, u1, u2) at
    // columns k, k+1, k+2; (l_k, interchange) eliminates row k+1
    double umax = 0.0;
    {
      double a = d[b0] - sig, bb = sub[b0];
      for (int k0 = 0; k0 + 1 < nb; k0 += kChunk) {
        double dd[kChunk], ss[kChunk + 1];
#pragma unroll
        for (int j = 0; j <= kChunk; ++j) {
          const int k = k0 + j;
          if (j < kChunk && k + 1 < nb) dd[j] = d[b0 + k + 1];
          if (k + 1 < nb) ss[j] = sub[b0 + k];
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const int k = k0 + j;
          if (k + 1 < nb) {
            const double cc = ss[j], a1 = dd[j] - sig;
            const double b1n = k + 2 < nb ? ss[j + 1] : 0.0;
            // interchange when the row below has the larger leading entry
            const bool p = fabs(cc) > fabs(a);
            const double num = p ? a : cc, den = p ? cc : a;
            const double l = den != 0.0 ? num * rcp_d(den) : 0.0;
            const double u0 = den, v1 = p ? a1 : bb, v2 = p ? b1n : 0.0;
            const double na = p ? bb - l * a1 : a1 - l * bb;
            const double nbb = p ? -l * b1n : b1n;
            ui[EFN_S(k)] = u0;
            u1[EFN_S(k)] = v1;
            lm[EFN_S(k)] = l;
            pv[EFN_S(k)] = p;
            umax = fmax(umax, fmax(fabs(u0), fmax(fabs(v1), fabs(v2))));
            a = na;
            bb = nbb;
          }
        }
      }
      ui[EFN_S(nb - 1)] = a;
      u1[EFN_S(nb - 1)] = 0.0;
      pv[EFN_S(nb - 1)] = 0;
      umax = fmax(umax, fabs(a));
    }
    const double ulast = ui[EFN_S(nb - 1)];
    // pivots below ulp·max|U| raised to that size (LAPACK dlagts' perturbation)
    const double tol = umax > 0.0 ? kUlp * umax : kUlp;
#pragma unroll 4
    for (int k = 0; k < nb; ++k) {
      double u0 = ui[EFN_S(k)];
      if (fabs(u0) < tol) u0 = u0 < 0.0 ? -tol : tol;
      ui[EFN_S(k)] = rcp_d(u0);
    }
    double xm = 0.0;  // |x|∞ of the current iterate
#pragma unroll 4
    for (int i = 0; i < nb; ++i) {
      const double v = refine ? vs[r * vr + (b0 + i) * vi] : start_entry(s - b0, i);
      X[EFN_S(i)] = v;
      xm = fmax(xm, fabs(v));
    }
    const double onenrm = onen[b0], dtpcrt = sqrt(0.1 / nb);
    int chk = 0;
    bool conv = false;
    for (int its = 0; its < (refine ? 1 : kInvIts) && !conv; ++its) {
      // scale so that |x|∞ = nb·‖T_b‖₁·max(ulp, |u_last|)
      const double scl = nb * onenrm * fmax(kUlp, fabs(ulast)) / xm;
      // y ← L⁻¹·P·(scl·x)
      double cur = X[EFN_S(0)] * scl;
      for (int k0 = 0; k0 + 1 < nb; k0 += kChunk) {
        double nx[kChunk], ll[kChunk];
        int pp[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const int k = k0 + j;
          if (k + 1 < nb) {
            nx[j] = X[EFN_S(k + 1)] * scl;
            ll[j] = lm[EFN_S(k)];
            pp[j] = pv[EFN_S(k)];
          }
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const int k = k0 + j;
          if (k + 1 < nb) {
            if (pp[j]) {
              X[EFN_S(k)] = nx[j];
              cur = cur - ll[j] * nx[j];
            } else {
              X[EFN_S(k)] = cur;
              cur = nx[j] - ll[j] * cur;
            }
          }
        }
      }
      X[EFN_S(nb - 1)] = cur;
      // x ← U⁻¹·y, rescaled by 2⁻⁵⁰⁰ whenever an entry passes 2⁵⁰⁰
      double x1 = 0.0, x2 = 0.0, big = 0.0;
      for (int k1 = nb - 1; k1 >= 0; k1 -= kChunk) {
        double yy[kChunk], v1[kChunk], v2[kChunk], vi[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const int k = k1 - j;
          if (k >= 0) {
            yy[j] = X[EFN_S(k)];
            v1[j] = u1[EFN_S(k)];
            v2[j] = pv[EFN_S(k)] && k + 2 < nb ? sub[b0 + k + 1] : 0.0;
            vi[j] = ui[EFN_S(k)];
          }
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const int k = k1 - j;
          if (k >= 0) {
            double xk = (yy[j] - v1[j] * x1 - v2[j] * x2) * vi[j];
            if (fabs(xk) > 0x1.0p500) {
              for (int i = 0; i < nb; ++i) X[EFN_S(i)] *= 0x1.0p-500;
#pragma unroll
              for (int jj = 0; jj < kChunk; ++jj)
                if (jj > j) yy[jj] *= 0x1.0p-500;
              xk *= 0x1.0p-500;
              x1 *= 0x1.0p-500;
              x2 *= 0x1.0p-500;
              big *= 0x1.0p-500;
            }
            big = fmax(big, fabs(xk));
            X[EFN_S(k)] = xk;
            x2 = x1;
            x1 = xk;
          }
        }
      }
      // modified Gram–Schmidt against the cluster's earlier vectors (read
      // kChunk rows at a time)
      if (s > lead) {
        for (int q = lead; q < s; ++q) {
          const double* vq = vs + rank[q] * vr + b0 * vi;
          double dot = 0.0;
          for (int i0 = 0; i0 < nb; i0 += kChunk) {
            double vv[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j)
              if (i0 + j < nb) vv[j] = vq[(i0 + j) * vi];
#pragma unroll
            for (int j = 0; j < kChunk; ++j)
              if (i0 + j < nb) dot += X[EFN_S(i0 + j)] * vv[j];
          }
          for (int i0 = 0; i0 < nb; i0 += kChunk) {
            double vv[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j)
              if (i0 + j < nb) vv[j] = vq[(i0 + j) * vi];
#pragma unroll
            for (int j = 0; j < kChunk; ++j)
              if (i0 + j < nb) X[EFN_S(i0 + j)] -= dot * vv[j];
          }
        }
        big = 0.0;
        for (int i = 0; i < nb; ++i) big = fmax(big, fabs(X[EFN_S(i)]));
      }
      xm = big;
      if (refine) break;
      if (big < dtpcrt) continue;
      if (++chk < kInvExtra + 1) continue;
      conv = true;
    }
    if (!conv && !refine) ++fails;
    // unit 2-norm, the largest entry (first of equals) positive
    double ss = 0.0, top = -1.0, bigv = 0.0;
#pragma unroll 8
    for (int i = 0; i < nb; ++i) {
      const double v = X[EFN_S(i)];
      ss += v * v;
      if (fabs(v) > top) {
        top = fabs(v);
        bigv = v;
      }
    }
    double scl = 1.0 / sqrt(ss);
    if (bigv < 0.0) scl = -scl;
#pragma unroll 8
    for (int i = 0; i < nb; ++i) vs[r * vr + (b0 + i) * vi] = X[EFN_S(i)] * scl;
  }
#undef EFN_S
  return fails;
}

// Classical Gram–Schmidt, twice, over one cluster's unit vectors in LDS
// (vector of slot s at zf[rank[s]·n + g], block rows g): in slot order each
// vector loses its components along all the earlier ones, taken from one
// copy of it, and again (the second pass restores the orthogonality the
// first loses to cancellation when vectors of a tight cluster are nearly
// dependent), then is normalised with its largest entry (first of equals)
// positive.  On the device one wave runs a cluster without workgroup
// barriers: lane l forms the dot product with the cluster's l-th earlier
// vector (rows in order), the dots meet in a per-wave LDS row, each lane
// updates its own rows (earlier vectors in order), and the norm and the
// largest entry come from lane shuffles; dsc: ≥ 96 doubles of the wave's own.
EFN_QL_HD inline void wave_order() {
#ifdef __HIP_DEVICE_COMPILE__
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#endif
}
EFN_QL_HD inline void orth_cluster(double* __restrict__ zf, int n, const int* __restrict__ bs,
                                   const int* __restrict__ be, const int* __restrict__ rank, int lead, int last,
                                   double* __restrict__ dsc, int lane) {
  const int b0 = bs[lead], nb = be[lead] - b0;
#ifdef __HIP_DEVICE_COMPILE__
  constexpr int W = 64;
#else
  constexpr int W = 1;
  (void)lane;
  lane = 0;
#endif
  for (int s = lead + 1; s <= last; ++s) {
    double* xj = zf + (long)rank[s] * n + b0;
    const int c = s - lead;  // earlier vectors
    for (int pass = 0; pass < 2; ++pass) {
      for (int l = lane; l < c; l += W) {
        const double* xq = zf + (long)rank[lead + l] * n + b0;
        double dot = 0.0;
        for (int i = 0; i < nb; ++i) dot += xq[i] * xj[i];
        dsc[l] = dot;
      }
      wave_order();
      for (int i = lane; i < nb; i += W) {
        double acc = xj[i];
        for (int l = 0; l < c; ++l) acc -= dsc[l] * zf[(long)rank[lead + l] * n + b0 + i];
        xj[i] = acc;
      }
      wave_order();
    }
    // ‖x‖² and the largest entry (first of equals)
    double ss = 0.0, top = -1.0, bigv = 0.0;
    int at = 1 << 30;
    for (int i = lane; i < nb; i += W) {
      const double v = xj[i];
      ss += v * v;
      if (fabs(v) > top) {
        top = fabs(v);
        bigv = v;
        at = i;
      }
    }
#ifdef __HIP_DEVICE_COMPILE__
    for (int off = 1; off < W; off <<= 1) {
      ss += __shfl_xor(ss, off, 64);
      const double ot = __shfl_xor(top, off, 64), ov = __shfl_xor(bigv, off, 64);
      const int oa = __shfl_xor(at, off, 64);
      if (ot > top || (ot == top && oa < at)) {
        top = ot;
        bigv = ov;
        at = oa;
      }
    }
#endif
    double scl = 1.0 / sqrt(ss);
    if (bigv < 0.0) scl = -scl;
    for (int i = lane; i < nb; i += W) xj[i] *= scl;
    wave_order();
  }
}

// orth_cluster over every cluster with two members or more: one wave a
// cluster on the device (waves take clusters in turn), then a barrier
EFN_QL_HD inline void orth_all(const Team& t, double* zf, int n, const int* bs, const int* be, const int* rank,
                               const int* gr, int nc, double* dsc) {
#ifdef __HIP_DEVICE_COMPILE__
  const int wave = t.tid >> 6, nw = t.nthr >> 6, lane = t.tid & 63;
  for (int q = wave; q < nc; q += nw)
    if (gr[2 * q + 1] > gr[2 * q]) orth_cluster(zf, n, bs, be, rank, gr[2 * q], gr[2 * q + 1], dsc + 96 * wave, lane);
  __threadfence_block();
#else
  for (int q = 0; q < nc; ++q)
    if (gr[2 * q + 1] > gr[2 * q]) orth_cluster(zf, n, bs, be, rank, gr[2 * q], gr[2 * q + 1], dsc, 0);
#endif
  t.sync();
}

// The bisection + inverse-iteration solver for one matrix: arguments and
// outputs as heev_ql_body's (lds: heevx_lds_bytes(nl) bytes), except info: the number of eigenvectors whose
// inverse iteration did not pass its stopping test (their vectors are still
// returned; the eigenvalues are always the bisection's).
EFN_QL_HD inline void heevx_body(const Team& t, const cd* A, int lda, int n, int nl, int lowest, double* w, cd* V,
                                 int* info, char* lds) {
  const Work W = carve(nl, lds);
  const double* d = W.d;
  const double* sub = W.sub;
  // the scaled copy of T: ee[i] = (2^−σ·sub[i−1])², zero at i = 0 and at splits; ds = 2^−σ·d
  double* ee = reinterpret_cast<double*>(W.u);
  double* ds = ee + nl;
  int* bs = reinterpret_cast<int*>(W.y);
  int* be = bs + n;
  int* rank = bs + 2 * n;
  double* ev = W.cs;
  double* onen = W.cs + W.csn;
  int* ctl = W.ctl;
  int* perm = ctl + kCtlPerm;
  int* gr = reinterpret_cast<int*>(W.red);
  const int low = lowest < n ? lowest : n;
  const double inf = HUGE_VAL;

  // ---- rows with no off-diagonal entry are eigenvectors already: ordered
  // last (live rows first, each group in its order), the reduction leaves
  // them untouched and T splits there, so the live block's norm — not a
  // penalty on such rows — sets the cluster scale below.  V's rows are
  // mapped back at the last store.
  int* rowmap = reinterpret_cast<int*>(W.cs + 2 * W.csn);
  int* iso = reinterpret_cast<int*>(W.cs + 3 * W.csn);
  for (int i = t.tid; i < n; i += t.nthr) {
    bool z = true;  // every off-diagonal entry of row i (lower triangle: row i and column i), loads issued 8 at a time
    for (int j0 = 0; j0 < n; j0 += kChunk) {
      cd v[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        const int j = j0 + u;
        v[u] = j < n && j != i ? (j < i ? A[(long)i * lda + j] : A[(long)j * lda + i]) : cmk(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < kChunk; ++u) z = z && v[u].x == 0.0 && v[u].y == 0.0;
    }
    iso[i] = z;
  }
  t.sync();
  if (t.tid == 0) {
    int nlive = 0;
    for (int i = 0; i < n; ++i) nlive += !iso[i];
    for (int i = 0, p = 0, q = nlive; i < n; ++i) rowmap[iso[i] ? q++ : p++] = i;
  }
  t.sync();

  reduce_tridiag(t, A, lda, n, W, false, rowmap);

  // ---- split into unreduced blocks; pivmin; ‖T_b‖₁ per row; the copy of T
  // scaled by 2^−σ ≥ 1/max_i(|d_i| + |sub_{i−1}| + |sub_i|) for the counts.
  // Row sums and split flags per row; the two maxima by 64-bit atomic max on
  // the bit patterns (non-negative doubles order as their bits); each row
  // then finds its block and the block's norm by scanning its own block.
  int* split = bs + 3 * n;
  double* rows = W.cs + 3 * W.csn;
  unsigned long long* mx = reinterpret_cast<unsigned long long*>(ctl + kCtlMax);  // rmax, emax
  if (t.tid == 0) mx[0] = mx[1] = 0ull;
  t.sync();
  {
    unsigned long long rb = 0ull, eb = 0ull;
    for (int i = t.tid; i < n; i += t.nthr) {
      const double s = fabs(d[i]) + (i > 0 ? fabs(sub[i - 1]) : 0.0) + (i + 1 < n ? fabs(sub[i]) : 0.0);
      rows[i] = s;
      bool sp = i == n - 1;
      if (!sp) {
        const double e = sub[i] * sub[i];
        sp = e <= kUlp * kUlp * fabs(d[i] * d[i + 1]) + kSafMin;
        const unsigned long long b = (unsigned long long)__double_as_longlong_hd(e);
        eb = b > eb ? b : eb;
      }
      split[i] = sp;
      const unsigned long long b = (unsigned long long)__double_as_longlong_hd(s);
      rb = b > rb ? b : rb;
    }
#ifdef __HIP_DEVICE_COMPILE__
    if (rb) atomicMax(&mx[0], rb);
    if (eb) atomicMax(&mx[1], eb);
#else
    mx[0] = rb > mx[0] ? rb : mx[0];
    mx[1] = eb > mx[1] ? eb : mx[1];
#endif
  }
  t.sync();
  const double rmax = __longlong_as_double_hd((long long)mx[0]), emax = __longlong_as_double_hd((long long)mx[1]);
  int ex = 0;
  if (rmax > 0.0) (void)frexp(rmax, &ex);
  const double scl = ldexp(1.0, -ex);
  const double pivmin = kSafMin * fmax(1.0, emax);  // of the unscaled matrix
  // block bounds from the split flags as a 128-bit mask (n ≤ 96): a block
  // ends at each split row
  unsigned long long* msk = reinterpret_cast<unsigned long long*>(ctl + kCtlMask);
#ifdef __HIP_DEVICE_COMPILE__
  if (t.tid < 64) {
    const unsigned long long m0 = __ballot(t.tid < n && split[t.tid]);
    const unsigned long long m1 = __ballot(64 + t.tid < n && split[64 + t.tid]);
    if (t.tid == 0) {
      msk[0] = m0;
      msk[1] = m1;
    }
  }
#else
  if (t.tid == 0) {
    msk[0] = msk[1] = 0ull;
    for (int i = 0; i < n; ++i)
      if (split[i]) msk[i >> 6] |= 1ull << (i & 63);
  }
#endif
  t.sync();
  for (int i = t.tid; i < n; i += t.nthr) {
    ds[i] = d[i] * scl;
    const double es = i > 0 ? sub[i - 1] * scl : 0.0;
    ee[i] = i > 0 && !split[i - 1] ? es * es : 0.0;
    const int b0 = i == 0 ? 0 : last_bit_below(msk, i) + 1, b1 = first_bit_from(msk, i);
    bs[i] = b0;
    be[i] = b1 + 1;
    if (i == b0) {  // the block's ‖T_b‖₁, kept at its first row
      double on = 0.0;
      for (int r = b0; r <= b1; ++r) on = fmax(on, fabs(d[r]) + (r > b0 ? fabs(sub[r - 1]) : 0.0) +
                                                       (r < b1 ? fabs(sub[r]) : 0.0));
      onen[b0] = on;
    }
  }
  t.sync();

  // ---- every eigenvalue of every block by multisection ----------------------
  {
    const int G = bisect_lanes(n), K = G * kSturmPts;
#ifdef __HIP_DEVICE_COMPILE__
    const int s = t.tid / G, g = t.tid % G;
    const bool mine = s < n;
    int b0 = 0, b1 = 0, j = 0;
    double lo = 0.0, hi = 0.0, atol = 0.0;
    bool done = true;
    if (mine) {
      b0 = bs[s];
      b1 = be[s];
      j = s - b0;
      if (b1 - b0 == 1) {
        lo = hi = d[s];
      } else {
        bisect_init(d, sub, b0, b1, pivmin, lo, hi, atol);
        done = bisect_done(lo, hi, atol, pivmin);
      }
    }
    for (int it = 0; it < kMaxBisect; ++it) {
      if (!__syncthreads_or(!done)) break;
      if (!done) gr[t.tid] = bisect_pts(ds, ee, b0, b1, lo, hi, g * kSturmPts + 1, K, j, scl);
      t.sync();
      if (!done) {
        int a = 0;
        for (int q = 0; q < G; ++q) a = max(a, gr[s * G + q]);
        const double nlo = a == 0 ? lo : bisect_point(lo, hi, a, K);
        const double nhi = a == K ? hi : bisect_point(lo, hi, a + 1, K);
        lo = nlo;
        hi = nhi;
        done = bisect_done(lo, hi, atol, pivmin);
      }
    }
    if (mine && g == 0) ev[s] = 0.5 * (lo + hi);
#else
    for (int s = 0; s < n; ++s) {
      const int b0 = bs[s], b1 = be[s], j = s - b0;
      if (b1 - b0 == 1) {
        ev[s] = 0.5 * (d[s] + d[s]);
        continue;
      }
      double lo, hi, atol;
      bisect_init(d, sub, b0, b1, pivmin, lo, hi, atol);
      for (int it = 0; it < kMaxBisect && !bisect_done(lo, hi, atol, pivmin); ++it) {
        int a = 0;
        for (int g = 0; g < G; ++g) {
          const int v = bisect_pts(ds, ee, b0, b1, lo, hi, g * kSturmPts + 1, K, j, scl);
          a = v > a ? v : a;
        }
        const double nlo = a == 0 ? lo : bisect_point(lo, hi, a, K);
        const double nhi = a == K ? hi : bisect_point(lo, hi, a + 1, K);
        lo = nlo;
        hi = nhi;
      }
      ev[s] = 0.5 * (lo + hi);
    }
#endif
  }
  t.sync();
  // ascending within each block (a Sturm count need not be monotone in x)
  if (t.tid == 0) {
    double prev = 0.0;
    for (int s = 0; s < n; ++s) {
      double v = ev[s];
      if (s > bs[s] && v < prev) v = prev;
      ev[s] = v;
      prev = v;
    }
  }
  t.sync();

  // ---- the lowest `low`: rank of each slot (ties by slot) -------------------
  for (int s = t.tid; s < n; s += t.nthr) {
    int r = 0;
    const double v = ev[s];
    for (int i = 0; i < n; ++i) r += (ev[i] < v) || (ev[i] == v && i < s);
    rank[s] = r;
    if (r < low) perm[r] = s;
  }
  t.sync();
  for (int r = t.tid; r < lowest; r += t.nthr) w[r] = r < low ? ev[perm[r]] : inf;
  // clusters with a selected member, as (first slot, last selected slot) in
  // slots taken by an atomic counter (a cluster's vectors do not depend on
  // its place in the list); within a block the selected slots are a prefix
  // (ranks grow with the slot).  A cluster leads with a block's first slot or
  // a gap above 10⁻³·‖T_b‖₁ to the slot before.
  for (int s = t.tid; s < n; s += t.nthr) {
    const bool leader = s == bs[s] || ev[s] - ev[s - 1] > 1e-3 * onen[bs[s]];
    if (leader && rank[s] < low) {
      int last = s;
      while (last + 1 < be[s] && rank[last + 1] < low && !(ev[last + 1] - ev[last] > 1e-3 * onen[bs[s]])) ++last;
#ifdef __HIP_DEVICE_COMPILE__
      const int slot = atomicAdd(&ctl[kCtlClus], 1);
#else
      const int slot = ctl[kCtlClus]++;
#endif
      gr[2 * slot] = s;
      gr[2 * slot + 1] = last;
    }
  }
  // ---- eigenvectors of T: every selected vector by inverse iteration on its
  // own thread, then each cluster made orthonormal once by one thread, in Z
  // (column-major, zf[rank·n + row]).  The iterations' scratch sits beside Z
  // where the launch leaves LDS room, the vectors then go straight into Z;
  // otherwise the scratch is Z itself, the vectors go to V and are staged
  // into Z afterwards (the same arithmetic either way).  An independent
  // vector's components along its cluster's other vectors are of size
  // ulp·‖T_b‖/gap, so removing them moves its residual by ulp·‖T_b‖.
  {
    const int Kc = invit_slots(n, nl), nc = ctl[kCtlClus];
    double* sc = invit_scratch(nl, lds, W);
    const bool inz = heevx_extra_bytes(nl) > 0;
    double* zf = W.z;
    // per-wave rows of 96 doubles for the Gram–Schmidt dots, in red past the
    // cluster list (2·96 ints)
    double* dsc = reinterpret_cast<double*>(W.red) + 96;
    double* vd = reinterpret_cast<double*>(V);
    int fails = 0;
    if (inz)
      for (long idx = t.tid; idx < (long)low * n; idx += t.nthr) zf[idx] = 0.0;
    else
      for (long idx = t.tid; idx < (long)lda * lowest; idx += t.nthr) V[idx] = cmk(0.0, 0.0);
#ifdef __HIP_DEVICE_COMPILE__
    __threadfence_block();
#endif
    t.sync();
    if (t.tid < Kc)
      for (int r = t.tid; r < low; r += (t.nthr > 1 ? Kc : 1))
        fails += invit_cluster(d, sub, ev, onen, bs, be, rank, perm[r], perm[r], sc, n, t.nthr > 1 ? t.tid : 0, Kc,
                               inz ? zf : vd, inz ? n : 2, inz ? 1 : 2L * lowest);
#ifdef __HIP_DEVICE_COMPILE__
    __threadfence_block();
#endif
    t.sync();
    if (!inz) {
      for (long idx = t.tid; idx < (long)low * n; idx += t.nthr) {
        const int r = (int)(idx / n), i = (int)(idx % n);
        zf[idx] = V[(long)i * lowest + r].x;
      }
      t.sync();
    }
    orth_all(t, zf, n, bs, be, rank, gr, nc, dsc);
    // members of clusters: one more solve from the orthonormal vector (it
    // damps what the Gram–Schmidt amplified from the other vectors' errors
    // outside the cluster) and Gram–Schmidt again
    // (a slot is a member when a neighbour of its block lies within the
    // cluster gap, selected or not, so the set does not depend on `lowest`)
    int* inclus = bs + 3 * n;  // the split flags' place, free again
    bool any = false;
    for (int s = t.tid; s < n; s += t.nthr) {
      const double gap = 1e-3 * onen[bs[s]];
      const bool m = (s > bs[s] && !(ev[s] - ev[s - 1] > gap)) || (s + 1 < be[s] && !(ev[s + 1] - ev[s] > gap));
      inclus[s] = m;
      any = any || (m && rank[s] < low);
    }
#ifdef __HIP_DEVICE_COMPILE__
    any = __syncthreads_or(any);
#else
    t.sync();
#endif
    if (any) {
      if (!inz)
        for (long idx = t.tid; idx < (long)low * n; idx += t.nthr) {
          const int r = (int)(idx / n), i = (int)(idx % n);
          V[(long)i * lowest + r] = cmk(zf[idx], 0.0);
        }
#ifdef __HIP_DEVICE_COMPILE__
      __threadfence_block();
#endif
      t.sync();
      if (t.tid < Kc)
        for (int r = t.tid; r < low; r += (t.nthr > 1 ? Kc : 1))
          if (inclus[perm[r]])
            invit_cluster(d, sub, ev, onen, bs, be, rank, perm[r], perm[r], sc, n, t.nthr > 1 ? t.tid : 0, Kc,
                          inz ? zf : vd, inz ? n : 2, inz ? 1 : 2L * lowest, true);
#ifdef __HIP_DEVICE_COMPILE__
      __threadfence_block();
#endif
      t.sync();
      if (!inz) {
        for (long idx = t.tid; idx < (long)low * n; idx += t.nthr) {
          const int r = (int)(idx / n), i = (int)(idx % n);
          zf[idx] = V[(long)i * lowest + r].x;
        }
        t.sync();
      }
      orth_all(t, zf, n, bs, be, rank, gr, nc, dsc);
    }
    for (long idx = t.tid; idx < (long)lda * lowest; idx += t.nthr) {
      const int i = (int)(idx / lowest), r = (int)(idx % lowest);
      V[idx] = cmk(i < n && r < low ? zf[(long)r * n + i] : 0.0, 0.0);
    }
    if (fails) {
#ifdef __HIP_DEVICE_COMPILE__
      atomicAdd(&ctl[kCtlFail], fails);
#else
      ctl[kCtlFail] += fails;
#endif
    }
  }
  back_transform(t, W, n, nl, lda, lowest, low, V, rowmap);
  if (t.tid == 0) *info = ctl[kCtlFail];
}

// per-problem sizes of a ragged launch, passed by value (no device array)
constexpr int kMaxRagged = 64;
struct Dims {
  short n[kMaxRagged];
  int ragged;
};

#ifdef __HIPCC__
// grid: (nprob); block: kThreads; dynamic LDS: Layout(lda).bytes().
// A: nprob × lda × lda; problem p is its leading n_p × n_p block (n_p =
// dims.n[p] when dims.ragged, else lda); w: nprob × lowest; V: nprob × lda ×
// lowest; info: nprob.
__global__ __launch_bounds__(kThreads) void heev_ql_kernel(const double2* __restrict__ A, int lda, Dims dims,
                                                           int lowest, double* __restrict__ w,
                                                           double2* __restrict__ V, int* __restrict__ info) {
  extern __shared__ __align__(16) char ql_lds[];
  const int p = blockIdx.x;
  const int n = dims.ragged ? dims.n[p] : lda;
  const Team t{(int)threadIdx.x, (int)blockDim.x};
  heev_ql_body(t, reinterpret_cast<const cd*>(A) + (long)p * lda * lda, lda, n, lda, lowest,
               w + (long)p * lowest, reinterpret_cast<cd*>(V) + (long)p * lda * lowest, info + p, ql_lds);
}

// grid: (nprob); block: kThreads; dynamic LDS: heevx_lds_bytes(lda); arguments
// as heev_ql_kernel's (info: eigenvectors whose inverse iteration did not pass
// its stopping test)
__global__ __launch_bounds__(kThreads) void heevx_kernel(const double2* __restrict__ A, int lda, Dims dims,
                                                         int lowest, double* __restrict__ w,
                                                         double2* __restrict__ V, int* __restrict__ info) {
  extern __shared__ __align__(16) char ql_lds[];
  const int p = blockIdx.x;
  const int n = dims.ragged ? dims.n[p] : lda;
  const Team t{(int)threadIdx.x, (int)blockDim.x};
  heevx_body(t, reinterpret_cast<const cd*>(A) + (long)p * lda * lda, lda, n, lda, lowest, w + (long)p * lowest,
             reinterpret_cast<cd*>(V) + (long)p * lda * lowest, info + p, ql_lds);
}
#endif

}  // namespace ql
}  // namespace efn
