// Device FFT building blocks for gfx950: compile-time in-register DFTs and a
// four-step pencil FFT staged in LDS.
//
// Conventions: forward transforms use exp(-2πi nk/N) (S = -1), inverse ones
// exp(+2πi nk/N) (S = +1) with no 1/N factor; callers fold normalisation into
// the real-space multiply.  Data are float2 (re, im).
#pragma once

#include <hip/hip_runtime.h>

namespace efn {

// ---------------------------------------------------------------------------
// constexpr trigonometry (Taylor series in double after exact reduction of the
// rational angle p/q turns); used only to build compile-time twiddle tables.
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ce_sin_series(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 16; ++i) {
    term *= -x * x / double((2 * i) * (2 * i + 1));
    sum += term;
  }
  return sum;
}
constexpr double ce_cos_series(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 16; ++i) {
    term *= -x * x / double((2 * i - 1) * (2 * i));
    sum += term;
  }
  return sum;
}
// cos/sin of 2π·p/q
constexpr double ce_cos_turn(long p, long q) {
  long r = ((p % q) + q) % q;
  if (2 * r > q) r -= q;  // angle in (-π, π]
  return ce_cos_series(2.0 * kPi * double(r) / double(q));
}
constexpr double ce_sin_turn(long p, long q) {
  long r = ((p % q) + q) % q;
  if (2 * r > q) r -= q;
  return ce_sin_series(2.0 * kPi * double(r) / double(q));
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by S·i (S = ±1)
template <int S>
__device__ __forceinline__ float2 mul_si(float2 a) {
  return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// Multiply by the compile-time root exp(S·2πi·e/R); exact shortcuts for the
// quarter turns, folded when e is a constant after unrolling.
template <int R, int S>
__device__ __forceinline__ float2 twiddle_c(float2 v, int e) {
  e = ((e % R) + R) % R;
  if (e == 0) return v;
  if (4 * e == R) return mul_si<S>(v);
  if (2 * e == R) return make_float2(-v.x, -v.y);
  if (4 * e == 3 * R) return mul_si<-S>(v);
  const float c = float(ce_cos_turn(e, R));
  const float s = float(S * ce_sin_turn(e, R));
  return cmul(v, make_float2(c, s));
}

// ---------------------------------------------------------------------------
// In-register DFT of size R (natural order in and out).
constexpr bool is_prime_small(int r) { return r == 2 || r == 3 || r == 5 || r == 7; }
constexpr int smallest_factor(int r) {
  for (int f = 2; f <= r; ++f)
    if (r % f == 0) return f;
  return r;
}
// radix-4 steps first where possible (cheaper butterflies)
constexpr int first_factor(int r) { return (r % 4 == 0 && r != 4) ? 4 : smallest_factor(r); }

template <int R, int S>
struct DFT {
  // composite R = A·B: n = B·a + b, k = ka + A·kb
  static constexpr int A = first_factor(R);
  static constexpr int B = R / A;
  __device__ __forceinline__ static void run(float2 (&v)[R]) {
    float2 y[B][A];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      float2 t[A];
#pragma unroll
      for (int a = 0; a < A; ++a) t[a] = v[B * a + b];
      DFT<A, S>::run(t);
#pragma unroll
      for (int ka = 0; ka < A; ++ka) y[b][ka] = twiddle_c<R, S>(t[ka], b * ka);
    }
#pragma unroll
    for (int ka = 0; ka < A; ++ka) {
      float2 t[B];
#pragma unroll
      for (int b = 0; b < B; ++b) t[b] = y[b][ka];
      DFT<B, S>::run(t);
#pragma unroll
      for (int kb = 0; kb < B; ++kb) v[ka + A * kb] = t[kb];
    }
  }
};

template <int S>
struct DFT<1, S> {
  __device__ __forceinline__ static void run(float2 (&)[1]) {}
};

template <int S>
struct DFT<2, S> {
  __device__ __forceinline__ static void run(float2 (&v)[2]) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int S>
struct DFT<3, S> {
  __device__ __forceinline__ static void run(float2 (&v)[3]) {
    constexpr float kH = 0.866025403784438646763723170752936183f;  // √3/2
    float2 s = cadd(v[1], v[2]);
    float2 d = csub(v[1], v[2]);
    float2 m = make_float2(fmaf(-0.5f, s.x, v[0].x), fmaf(-0.5f, s.y, v[0].y));
    float2 r = mul_si<S>(cscale(d, kH));
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <int S>
struct DFT<4, S> {
  __device__ __forceinline__ static void run(float2 (&v)[4]) {
    float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    float2 s13 = cadd(v[1], v[3]), d13 = mul_si<S>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};

// small odd primes: direct O(R²) with symmetric pairing
template <int R, int S>
struct DFTPrime {
  __device__ __forceinline__ static void run(float2 (&v)[R]) {
    float2 out[R];
    float2 sum = v[0];
#pragma unroll
    for (int n = 1; n < R; ++n) sum = cadd(sum, v[n]);
    out[0] = sum;
#pragma unroll
    for (int k = 1; k <= R / 2; ++k) {
      float2 re = v[0], im = make_float2(0.f, 0.f);
#pragma unroll
      for (int n = 1; n <= R / 2; ++n) {
        const float c = float(ce_cos_turn(long(n) * k, R));
        const float s = float(ce_sin_turn(long(n) * k, R));
        float2 p = cadd(v[n], v[R - n]);
        float2 q = csub(v[n], v[R - n]);
        re = make_float2(fmaf(c, p.x, re.x), fmaf(c, p.y, re.y));
        im = make_float2(fmaf(s, q.x, im.x), fmaf(s, q.y, im.y));
      }
      // X_k = re + S·i·im, X_{R-k} = re − S·i·im
      float2 r = mul_si<S>(im);
      out[k] = cadd(re, r);
      out[R - k] = csub(re, r);
    }
#pragma unroll
    for (int n = 0; n < R; ++n) v[n] = out[n];
  }
};
template <int S>
struct DFT<5, S> : DFTPrime<5, S> {};
template <int S>
struct DFT<7, S> : DFTPrime<7, S> {};

// ---------------------------------------------------------------------------
// Four-step plan for a pencil of length N = N1·N2 (N1 ≤ N2, both ≤ 16 and
// built from radices 2,3,4,5,7).
constexpr bool radix_ok(int r) {
  if (r < 1 || r > 16) return false;
  while (r % 2 == 0) r /= 2;
  while (r % 3 == 0) r /= 3;
  while (r % 5 == 0) r /= 5;
  while (r % 7 == 0) r /= 7;
  return r == 1;
}
constexpr int plan_n1(int N) {
  int best = 0;
  for (int a = 1; a * a <= N; ++a)
    if (N % a == 0 && radix_ok(a) && radix_ok(N / a)) best = a;
  return best;
}

// SWAP exchanges the two factors (N1 ≥ N2): step 3 then writes runs of N1
// consecutive outputs.
template <int N, bool SWAP = false>
struct Plan {
  static constexpr int N1 = SWAP ? N / plan_n1(N) : plan_n1(N);  // step-1 DFT size (stride-N2 inputs)
  static constexpr int N2 = N / N1;                              // step-3 DFT size
  static_assert(N1 > 0, "pencil length not supported by the four-step plan");
  // padded LDS position of element i: one pad slot per run of N2, pencil stride odd
  static constexpr int pos(int i) { return i + i / N2; }
  static constexpr int kStrideRaw = N1 * (N2 + 1);
  static constexpr int kStride = (kStrideRaw % 2 == 0) ? kStrideRaw + 1 : kStrideRaw;
};

// Four-step building blocks on one pencil (padded LDS layout):
//   step1: the task n2 owns inputs x[N2·n1 + n2] (n1 < N1) in registers; DFT_N1,
//          twiddle by ω_N^{n2·k1}, store A[n2][k1] at pos(N2·k1 + n2).
//   step3: the task k1 loads A[·][k1] (the contiguous run pos(N2·k1 + n2)) and
//          leaves X[k1 + N1·k2] in r[k2].
// twid[e] = exp(-2πi e/N) for e in [0, N) (conjugated for S = +1).
template <int N, int S, class P = Plan<N>>
__device__ __forceinline__ void step1_store(float2 (&v)[P::N1], int n2, float2* __restrict__ pencil,
                                            const float2* __restrict__ twid) {
  constexpr int N1 = P::N1, N2 = P::N2;
  DFT<N1, S>::run(v);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) {
    float2 w = twid[n2 * k1];  // n2·k1 < N
    if (S > 0) w.y = -w.y;
    v[k1] = cmul(v[k1], w);
  }
  float2* col = pencil + n2;  // pos(N2·k1 + n2) = (N2 + 1)·k1 + n2 for n2 < N2
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) col[(N2 + 1) * k1] = v[k1];
}

// step1_store for register-bound kernels: the twiddles are read from LDS
// four at a time, each group consumed before the next is fetched, so at
// most four are held in registers (the plain version lets the compiler fetch
// all N1 − 1 of them ahead of the DFT).
template <int N, int S, class P = Plan<N>>
__device__ __forceinline__ void step1_store_lean(float2 (&v)[P::N1], int n2, float2* __restrict__ pencil,
                                                 const float2* __restrict__ twid) {
  constexpr int N1 = P::N1, N2 = P::N2;
  DFT<N1, S>::run(v);
#pragma unroll
  for (int k0 = 1; k0 < N1; k0 += 4) {
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k1 = k0; k1 < k0 + 4 && k1 < N1; ++k1) {
      float2 w = twid[n2 * k1];
      if (S > 0) w.y = -w.y;
      v[k1] = cmul(v[k1], w);
    }
  }
  float2* col = pencil + n2;
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) col[(N2 + 1) * k1] = v[k1];
}

template <int N, int S, class P = Plan<N>>
__device__ __forceinline__ void step3_load(float2 (&r)[P::N2], int k1, const float2* __restrict__ pencil) {
  constexpr int N2 = P::N2;
  const float2* run = pencil + (N2 + 1) * k1;  // pos(N2·k1 + n2) = (N2 + 1)·k1 + n2
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) r[n2] = run[n2];
  DFT<N2, S>::run(r);
}

// Step 3 of the four-step FFT over NP pencils in LDS, results written back in
// natural order (read all, barrier, write all, barrier).
template <int N, int S, int NP, int NT>
__device__ __forceinline__ void lds_step3_natural(float2* __restrict__ buf) {
  using P = Plan<N>;
  constexpr int N1 = P::N1, N2 = P::N2;
  const int tid = threadIdx.x;
  constexpr int T3 = NP * N1;
  constexpr int TPT = (T3 + NT - 1) / NT;
  float2 r[TPT][N2];
#pragma unroll
  for (int j = 0; j < TPT; ++j) {
    const int t = tid + j * NT;
    if (t < T3) {
      const int p = t / N1, k1 = t - p * N1;
      step3_load<N, S>(r[j], k1, buf + p * P::kStride);
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < TPT; ++j) {
    const int t = tid + j * NT;
    if (t < T3) {
      const int p = t / N1, k1 = t - p * N1;
      float2* base = buf + p * P::kStride;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) base[P::pos(k1 + N1 * k2)] = r[j][k2];
    }
  }
  __syncthreads();
}

// Four-step FFT over NP pencils resident in LDS (padded layout, natural order
// in and out).  Must be called by all NT threads of the block; contains __syncthreads().
template <int N, int S, int NP, int NT>
__device__ __forceinline__ void lds_fft(float2* __restrict__ buf, const float2* __restrict__ twid) {
  using P = Plan<N>;
  constexpr int N1 = P::N1, N2 = P::N2;
  constexpr int T1 = NP * N2;
  for (int t = threadIdx.x; t < T1; t += NT) {
    const int p = t / N2, n2 = t - p * N2;
    float2* base = buf + p * P::kStride;
    float2 v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = base[(N2 + 1) * n1 + n2];  // pos(N2·n1 + n2)
    step1_store<N, S>(v, n2, base, twid);
  }
  __syncthreads();
  lds_step3_natural<N, S, NP, NT>(buf);
}

}  // namespace efn
