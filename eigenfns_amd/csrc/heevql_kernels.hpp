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
  EFN_QL_HD long z_off() const { return a_off() + 16L * npk; }                 // double[n·ldz]
  EFN_QL_HD long d_off() const { return z_off() + 8L * ((n * ldz + 1) & ~1); }  // double[n]
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
enum { kCtlL = 0, kCtlIter = 1, kCtlFail = 2, kCtlCnt = 4, kCtlLo = 6, kCtlDone = 8, kCtlPerm = 16 };

// Block reduction of a complex partial (every thread passes one); the sum
// is returned to every thread.  Two barriers.  On the device each wave sums
// by lane shuffles and the wave sums meet in red; on the host (one thread)
// the value is returned.
EFN_QL_HD inline cd block_sum(const Team& t, cd* red, cd v) {
#ifdef __HIP_DEVICE_COMPILE__
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  if ((t.tid & 63) == 0) red[t.tid >> 6] = v;
  t.sync();
  cd s = red[0];
  for (int i = 1; i < (t.nthr >> 6); ++i) {
    s.x += red[i].x;
    s.y += red[i].y;
  }
  t.sync();
  return s;
#else
  cd s = cmk(0.0, 0.0);
  red[0] = v;
  for (int i = 0; i < t.nthr; ++i) {
    s.x += red[i].x;
    s.y += red[i].y;
  }
  return s;
#endif
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

// y_i = τ·Σ_j A₂₂(i, j)·v_j for rows i of the trailing block at k+1 (size m),
// A₂₂ Hermitian from the packed lower triangle
EFN_QL_HD inline void ql_hemv_row(const cd* __restrict__ a, const cd* __restrict__ u, cd* __restrict__ y, int k,
                                  int m, int i, cd tk) {
  const int gi = k + 1 + i;
  double sx = 0.0, sy = 0.0;
  const cd* __restrict__ row = a + pk(gi, k + 1);
  for (int j = 0; j < i; ++j) {
    const cd aij = row[j], v = u[j];
    sx += aij.x * v.x - aij.y * v.y;
    sy += aij.x * v.y + aij.y * v.x;
  }
  {
    const double aii = a[pk(gi, gi)].x;
    sx += aii * u[i].x;
    sy += aii * u[i].y;
  }
  for (int j = i + 1; j < m; ++j) {
    const cd aji = a[pk(k + 1 + j, gi)], v = u[j];  // A(i, j) = conj(A(j, i))
    sx += aji.x * v.x + aji.y * v.y;
    sy += aji.x * v.y - aji.y * v.x;
  }
  y[i] = cmul(tk, cmk(sx, sy));
}

// (row, column) of entry q of a packed lower triangle
EFN_QL_HD inline void tri_ij(int q, int& i, int& j) {
  int r = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
  while (r * (r + 1) / 2 > q) --r;
  while ((r + 1) * (r + 2) / 2 <= q) ++r;
  i = r;
  j = q - r * (r + 1) / 2;
}

// A₂₂ ← A₂₂ − v·wᴴ − w·vᴴ over the trailing lower triangle (size m), its
// entries dealt evenly over the team; the diagonal kept real
EFN_QL_HD inline void ql_her2(const Team& t, cd* __restrict__ a, const cd* __restrict__ u, const cd* __restrict__ w,
                              int k, int m) {
  const int tot = m * (m + 1) / 2;
  for (int q = t.tid; q < tot; q += t.nthr) {
    int i, j;
    tri_ij(q, i, j);
    const cd vi = u[i], wi = w[i], vj = u[j], wj = w[j];
    cd x = a[pk(k + 1 + i, k + 1 + j)];
    // v_i·conj(w_j) + w_i·conj(v_j)
    x.x -= (vi.x * wj.x + vi.y * wj.y) + (wi.x * vj.x + wi.y * vj.y);
    x.y -= (vi.y * wj.x - vi.x * wj.y) + (wi.y * vj.x - wi.x * vj.y);
    if (i == j) x.y = 0.0;
    a[pk(k + 1 + i, k + 1 + j)] = x;
  }
}

// The solver for one matrix.  A: n_p × n_p of a row-major array with leading
// dimension lda (only the lower triangle is read); w: `lowest` values
// (ascending; +inf past min(lowest, n_p)); V: lda × lowest row-major
// (V[i·lowest + j] = component i of eigenvector j; zero past n_p rows and
// min(lowest, n_p) columns); info: 0, or 1 + l if the QL iteration for
// eigenvalue l did not converge (then every w is NaN).  lds:
// Layout(nl).bytes() bytes, nl ≥ n_p.
EFN_QL_HD inline void heev_ql_body(const Team& t, const cd* A, int lda, int n, int nl, int lowest, double* w,
                                   cd* V, int* info, char* lds) {
  const Layout L(nl);
  cd* a = reinterpret_cast<cd*>(lds + L.a_off());
  double* z = reinterpret_cast<double*>(lds + L.z_off());
  double* d = reinterpret_cast<double*>(lds + L.d_off());
  double* sub = reinterpret_cast<double*>(lds + L.e_off());
  cd* tau = reinterpret_cast<cd*>(lds + L.tau_off());
  cd* u = reinterpret_cast<cd*>(lds + L.u_off());
  cd* y = reinterpret_cast<cd*>(lds + L.y_off());
  double* cs = reinterpret_cast<double*>(lds + L.cs_off());
  cd* red = reinterpret_cast<cd*>(lds + L.red_off());
  int* ctl = reinterpret_cast<int*>(lds + L.ctl_off());
  const int ldz = L.ldz;
  const int csn = (nl + 1) & ~1;
  const int low = lowest < n ? lowest : n;
  const double inf = HUGE_VAL;

  // ---- load the lower triangle, Z = I ----------------------------------------
  for (int i = t.tid; i < n; i += t.nthr) {
    for (int j = 0; j <= i; ++j) a[pk(i, j)] = A[(long)i * lda + j];
    for (int j = 0; j < n; ++j) z[i * ldz + j] = i == j ? 1.0 : 0.0;
  }
  if (t.tid == 0) {
    for (int c = 0; c < 16; ++c) ctl[c] = 0;
  }
  t.sync();

  // ---- Householder reduction to real tridiagonal (d, sub) ---------------------
  for (int k = 0; k + 1 < n; ++k) {
    const int m = n - k - 1;  // rows k+1 … n−1 of column k
    // ‖column below row k+1‖²
    cd part = cmk(0.0, 0.0);
    for (int i = k + 2 + t.tid; i < n; i += t.nthr) {
      const cd x = a[pk(i, k)];
      part.x += x.x * x.x + x.y * x.y;
    }
    const double xn2 = block_sum(t, red, part).x;
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
      // y = τ·A₂₂·v over the trailing block (rows and columns k+1 …)
      for (int i = t.tid; i < m; i += t.nthr) ql_hemv_row(a, u, y, k, m, i, tk);
      t.sync();
      // w = y − ½·τ·(yᴴv)·v
      cd pd = cmk(0.0, 0.0);
      for (int i = t.tid; i < m; i += t.nthr) {
        const cd c = cdotc(y[i], u[i]);
        pd.x += c.x;
        pd.y += c.y;
      }
      const cd yv = block_sum(t, red, pd);
      const cd al = cmul(cmk(-0.5 * tk.x, -0.5 * tk.y), yv);
      for (int i = t.tid; i < m; i += t.nthr) {
        const cd c = cmul(al, u[i]);
        y[i] = cmk(y[i].x + c.x, y[i].y + c.y);
      }
      t.sync();
      // A₂₂ ← A₂₂ − v·wᴴ − w·vᴴ (lower triangle, real diagonal)
      ql_her2(t, a, u, y, k, m);
      t.sync();
    }
  }
  if (t.tid == 0 && n > 0) {
    d[n - 1] = a[pk(n - 1, n - 1)].x;
    sub[n - 1] = 0.0;
  }
  t.sync();

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

  // ---- V = Q·Z[:, perm], in batches of columns staged in the Z region --------
  // Z's selected columns go to V first (real); each batch is read back into
  // LDS (complex, over Z), turned by H(n−2) … H(0) and stored.
  for (long idx = t.tid; idx < (long)lda * lowest; idx += t.nthr) {
    const int i = (int)(idx / lowest), r = (int)(idx % lowest);
    V[idx] = (i < n && r < low) ? cmk(z[i * ldz + perm[r]], 0.0) : cmk(0.0, 0.0);
  }
#ifdef __HIP_DEVICE_COMPILE__
  __threadfence_block();
#endif
  t.sync();
  cd* yb = reinterpret_cast<cd*>(z);
  // complex columns of n rows that fit over Z, at most kBackCols (the
  // partial sums of a batch fill red)
  int bmax = (nl * ldz) / (2 * (n > 0 ? n : 1));
  if (bmax > kBackCols) bmax = kBackCols;
  for (int c0 = 0; c0 < low; c0 += bmax) {
    const int nb = low - c0 < bmax ? low - c0 : bmax;
    for (int idx = t.tid; idx < n * nb; idx += t.nthr) {
      const int i = idx / nb, c = idx % nb;
      yb[i * nb + c] = V[(long)i * lowest + c0 + c];
    }
    t.sync();
    for (int k = n - 2; k >= 0; --k) {
      const cd tk = tau[k];
      if (tk.x == 0.0 && tk.y == 0.0) continue;  // uniform: tau is in LDS
      const int m = n - k - 1;
      // s_c = v_kᴴ·y_c over rows k+1 …: kBackPer partial sums a column, in
      // an order that depends on n only (a column's result is the same in
      // any batch, so a subset of columns equals the full solve's)
      constexpr int per = kBackPer;
      for (int slot = t.tid; slot < nb * per; slot += t.nthr) {
        const int c = slot / per, q = slot % per;
        cd part = cmk(0.0, 0.0);
        for (int i = q; i < m; i += per) {
          const cd v = i == 0 ? cmk(1.0, 0.0) : a[pk(k + 1 + i, k)];
          const cd x = cdotc(v, yb[(k + 1 + i) * nb + c]);
          part.x += x.x;
          part.y += x.y;
        }
        red[slot] = part;
      }
      t.sync();
      for (int cc = t.tid; cc < nb; cc += t.nthr) {
        cd s = cmk(0.0, 0.0);
        for (int j = 0; j < per; ++j) {
          s.x += red[cc * per + j].x;
          s.y += red[cc * per + j].y;
        }
        u[cc] = cmul(tk, s);  // τ·(vᴴy)
      }
      t.sync();
      for (int idx = t.tid; idx < m * nb; idx += t.nthr) {
        const int i = idx / nb, cc = idx % nb;
        const cd v = i == 0 ? cmk(1.0, 0.0) : a[pk(k + 1 + i, k)];
        const cd f = cmul(v, u[cc]);
        cd& x = yb[(k + 1 + i) * nb + cc];
        x.x -= f.x;
        x.y -= f.y;
      }
      t.sync();
    }
    for (int idx = t.tid; idx < n * nb; idx += t.nthr) {
      const int i = idx / nb, cc = idx % nb;
      V[(long)i * lowest + c0 + cc] = yb[i * nb + cc];
    }
    t.sync();
  }
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
#endif

}  // namespace ql
}  // namespace efn
