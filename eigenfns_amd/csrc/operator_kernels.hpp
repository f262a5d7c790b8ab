// Fused operator passes of Θ·H = ∇×(ε⁻¹∇×H) and of the transverse-projection
// preconditioner in the 2-component transverse plane-wave basis.
//
// Reference: eigenfns/operator.py:82-95 (_theta_block), :104-127
// (_precondition_fancy_block).  One application is five HBM passes over a
// chunk of vectors, each a batched pencil FFT staged in LDS:
//   zin   (m,2,G³) → W (m,3,G³): curl into the frame, Cartesian components,
//                                 inverse z-FFT                     [pointwise D1 + D2]
//   ypass W in place:            inverse y-FFT                     [D2]
//   xmid  W in place:            inverse x-FFT, ×ε⁻¹/G³ (or ε/G³),  [D2 + D3 + D4]
//                                 forward x-FFT
//   ypass W in place:            forward y-FFT                     [D4]
//   zout  W → (m,2,G³):          forward z-FFT, project on t̂₁,t̂₂,  [D4 + D5]
//                                 curl back, fused epilogue (store / Chebyshev
//                                 step / KPM step with moment partials)
// The frame (t̂₁, t̂₂, |k+G|) is rebuilt per voxel from three 1-D tables and a
// bit-packed FP64 reference-axis switch computed on the host.
#pragma once

#include "fft.hpp"

namespace efn {

enum OpKind { kTheta = 0, kPrecond = 1 };
enum Epi { kStore = 0, kCheb = 1, kKpm = 2, kDot = 3 };

struct FrameTables {
  const float* kx;          // (G) K_x = g + k0_x, FP32 cast of the FP64 value
  const float* ky;
  const float* kz;
  const uint32_t* swap;     // G³ bits, z fastest: reference axis x̂ instead of ẑ
  long long sing_voxel;     // linear index where k+G = 0 (frame zero), or -1
  float kn_floor;           // preconditioner floor on |k+G|
};

struct Frame {
  float kn, t1x, t1y, t1z, t2x, t2y, t2z;
};

// operator.py:53-71: t̂₁ = K×ref/|K×ref|, t̂₂ = K×t̂₁/|K×t̂₁|, zero at K = 0.
__device__ __forceinline__ Frame make_frame(const FrameTables& ft, int x, int y, int z, long long vox) {
  Frame f;
  const float Kx = ft.kx[x], Ky = ft.ky[y], Kz = ft.kz[z];
  const float kn2 = fmaf(Kx, Kx, fmaf(Ky, Ky, Kz * Kz));
  if (vox == ft.sing_voxel) {
    f.kn = 0.f;
    f.t1x = f.t1y = f.t1z = f.t2x = f.t2y = f.t2z = 0.f;
    return f;
  }
  f.kn = sqrtf(kn2);
  const bool sw = (ft.swap[vox >> 5] >> (vox & 31)) & 1u;
  float ax, ay, az;
  if (!sw) {  // ref = ẑ: K×ẑ = (Ky, −Kx, 0)
    ax = Ky; ay = -Kx; az = 0.f;
  } else {    // ref = x̂: K×x̂ = (0, Kz, −Ky)
    ax = 0.f; ay = Kz; az = -Ky;
  }
  const float r1 = 1.0f / sqrtf(fmaf(ax, ax, fmaf(ay, ay, az * az)));
  f.t1x = ax * r1; f.t1y = ay * r1; f.t1z = az * r1;
  const float bx = fmaf(Ky, f.t1z, -Kz * f.t1y);
  const float by = fmaf(Kz, f.t1x, -Kx * f.t1z);
  const float bz = fmaf(Kx, f.t1y, -Ky * f.t1x);
  const float r2 = 1.0f / sqrtf(fmaf(bx, bx, fmaf(by, by, bz * bz)));
  f.t2x = bx * r2; f.t2y = by * r2; f.t2z = bz * r2;
  return f;
}

// Pencils per block for the contiguous-axis passes: a power of two dividing G²
// so a block never spans two vectors.
__host__ __device__ constexpr int zrows_for(int G) {
  int target = G >= 128 ? 8 : (G >= 48 ? 16 : 32);
  int p = 1;
  while (p * 2 <= target && (G * G) % (p * 2) == 0) p *= 2;
  return p;
}
__host__ __device__ constexpr int ztile_for(int G) { return G >= 16 ? 16 : G; }

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// zin: (nvec, 2, G³) → W (nvec, 3, G³), curl into the Cartesian frame + inverse z-FFT
template <int G, int KIND>
__global__ __launch_bounds__(kThreads) void zin_kernel(const float2* __restrict__ in,
                                                       float2* __restrict__ ws, int nvec,
                                                       FrameTables ft,
                                                       const float2* __restrict__ twid_g) {
  using P = Plan<G>;
  constexpr int R = zrows_for(G);
  constexpr long long G2 = (long long)G * G, G3 = G2 * G;
  __shared__ float2 buf[3 * R * P::kStride];
  __shared__ float2 twid[G];
  for (int i = threadIdx.x; i < G; i += kThreads) twid[i] = twid_g[i];
  const long long row0 = (long long)blockIdx.x * R;  // flattened (v, x, y)
  const int v = int(row0 / G2);
  const long long xy0 = row0 - (long long)v * G2;
  const float2* a_in = in + (long long)v * 2 * G3;
  const float2* b_in = a_in + G3;
  for (int e = threadIdx.x; e < R * G; e += kThreads) {
    const int r = e / G, z = e - r * G;
    const long long xy = xy0 + r;
    const int x = int(xy / G), y = int(xy - (long long)x * G);
    const long long vox = xy * G + z;
    const float2 a = a_in[vox], b = b_in[vox];
    const Frame f = make_frame(ft, x, y, z, vox);
    float2 c1, c2;
    if (KIND == kTheta) {  // c1 = −i·kn·b, c2 = i·kn·a
      c1 = make_float2(f.kn * b.y, -f.kn * b.x);
      c2 = make_float2(-f.kn * a.y, f.kn * a.x);
    } else {               // c1 = i·b/knr, c2 = −i·a/knr
      const float inv = 1.0f / fmaxf(f.kn, ft.kn_floor);
      c1 = make_float2(-b.y * inv, b.x * inv);
      c2 = make_float2(a.y * inv, -a.x * inv);
    }
    const int p = P::pos(z);
    buf[(0 * R + r) * P::kStride + p] = make_float2(fmaf(c1.x, f.t1x, c2.x * f.t2x), fmaf(c1.y, f.t1x, c2.y * f.t2x));
    buf[(1 * R + r) * P::kStride + p] = make_float2(fmaf(c1.x, f.t1y, c2.x * f.t2y), fmaf(c1.y, f.t1y, c2.y * f.t2y));
    buf[(2 * R + r) * P::kStride + p] = make_float2(fmaf(c1.x, f.t1z, c2.x * f.t2z), fmaf(c1.y, f.t1z, c2.y * f.t2z));
  }
  __syncthreads();
  lds_fft<G, +1, 3 * R, kThreads>(buf, twid);
  float2* w = ws + (long long)v * 3 * G3;
  for (int e = threadIdx.x; e < R * G; e += kThreads) {
    const int r = e / G, z = e - r * G;
    const long long vox = (xy0 + r) * G + z;
    const int p = P::pos(z);
#pragma unroll
    for (int c = 0; c < 3; ++c) w[c * G3 + vox] = buf[(c * R + r) * P::kStride + p];
  }
}

// ---------------------------------------------------------------------------
// Strided pass over W (nvec·3 fields) in place.  AXIS 1 = y (stride G), 0 = x
// (stride G²).  MID: inverse FFT, multiply by rmul (real, per voxel), forward FFT.
template <int G, int AXIS, int S, bool MID>
__global__ __launch_bounds__(kThreads) void strided_kernel(float2* __restrict__ ws,
                                                           const float* __restrict__ rmul,
                                                           const float2* __restrict__ twid_g) {
  using P = Plan<G>;
  constexpr int ZT = ztile_for(G);
  constexpr long long G2 = (long long)G * G, G3 = G2 * G;
  constexpr long long stride = AXIS == 1 ? G : G2;
  __shared__ float2 buf[ZT * P::kStride];
  __shared__ float2 twid[G];
  for (int i = threadIdx.x; i < G; i += kThreads) twid[i] = twid_g[i];
  constexpr int nzt = (G + ZT - 1) / ZT;
  const long long fo = blockIdx.x / nzt;  // (field, other-axis index)
  const int zt = blockIdx.x - int(fo * nzt);
  const long long field = fo / G;
  const int o = int(fo - field * G);
  const int z0 = zt * ZT;
  float2* base = ws + field * G3 + (AXIS == 1 ? (long long)o * G2 : (long long)o * G) + z0;
  for (int e = threadIdx.x; e < ZT * G; e += kThreads) {
    const int n = e / ZT, z = e - n * ZT;
    if (z0 + z < G) buf[z * P::kStride + P::pos(n)] = base[n * stride + z];
  }
  __syncthreads();
  if (!MID) {
    lds_fft<G, S, ZT, kThreads>(buf, twid);
  } else {
    lds_fft<G, +1, ZT, kThreads>(buf, twid);
    for (int e = threadIdx.x; e < ZT * G; e += kThreads) {
      const int n = e / ZT, z = e - n * ZT;
      if (z0 + z < G) {
        // x-pass: voxel (x = n, y = o, z)
        const float s = rmul[(long long)n * G2 + (long long)o * G + z0 + z];
        float2& d = buf[z * P::kStride + P::pos(n)];
        d = make_float2(d.x * s, d.y * s);
      }
    }
    __syncthreads();
    lds_fft<G, -1, ZT, kThreads>(buf, twid);
  }
  for (int e = threadIdx.x; e < ZT * G; e += kThreads) {
    const int n = e / ZT, z = e - n * ZT;
    if (z0 + z < G) base[n * stride + z] = buf[z * P::kStride + P::pos(n)];
  }
}

// ---------------------------------------------------------------------------
struct EpiArgs {
  float alpha, beta;        // store: out = alpha·(op x) + beta·x
  float ck;                 // Chebyshev: Y += ck·T_{k+1}
  const float2* x;          // the operator's input block (T_k)
  float2* out;              // store: output; Cheb/KPM: T_{k−1} in, T_{k+1} out
  float2* y;                // Chebyshev accumulator
  const float2* z;          // kDot: probe block for Re⟨z, out⟩
  double* partial;          // KPM/kDot: per-block partial sums, 2 per block
};

__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kThreads / 64; ++i) s += red[i];
  }
  return s;  // valid in thread 0
}

// zout: W (nvec,3,G³) → forward z-FFT → project + curl back → epilogue
template <int G, int KIND, int EPI>
__global__ __launch_bounds__(kThreads) void zout_kernel(const float2* __restrict__ ws, int nvec,
                                                        FrameTables ft,
                                                        const float2* __restrict__ twid_g,
                                                        EpiArgs ea) {
  using P = Plan<G>;
  constexpr int R = zrows_for(G);
  constexpr long long G2 = (long long)G * G, G3 = G2 * G;
  __shared__ float2 buf[3 * R * P::kStride];
  __shared__ float2 twid[G];
  __shared__ double red[2 * kThreads / 64];
  for (int i = threadIdx.x; i < G; i += kThreads) twid[i] = twid_g[i];
  const long long row0 = (long long)blockIdx.x * R;
  const int v = int(row0 / G2);
  const long long xy0 = row0 - (long long)v * G2;
  const float2* w = ws + (long long)v * 3 * G3;
  for (int e = threadIdx.x; e < R * G; e += kThreads) {
    const int r = e / G, z = e - r * G;
    const long long vox = (xy0 + r) * G + z;
    const int p = P::pos(z);
#pragma unroll
    for (int c = 0; c < 3; ++c) buf[(c * R + r) * P::kStride + p] = w[c * G3 + vox];
  }
  __syncthreads();
  lds_fft<G, -1, 3 * R, kThreads>(buf, twid);
  const long long vbase = (long long)v * 2 * G3;
  double s0 = 0.0, s1 = 0.0;
  for (int e = threadIdx.x; e < R * G; e += kThreads) {
    const int r = e / G, z = e - r * G;
    const long long xy = xy0 + r;
    const int x = int(xy / G), y = int(xy - (long long)x * G);
    const long long vox = xy * G + z;
    const int p = P::pos(z);
    const float2 d0 = buf[(0 * R + r) * P::kStride + p];
    const float2 d1 = buf[(1 * R + r) * P::kStride + p];
    const float2 d2 = buf[(2 * R + r) * P::kStride + p];
    const Frame f = make_frame(ft, x, y, z, vox);
    const float2 f1 = make_float2(fmaf(f.t1x, d0.x, fmaf(f.t1y, d1.x, f.t1z * d2.x)),
                                  fmaf(f.t1x, d0.y, fmaf(f.t1y, d1.y, f.t1z * d2.y)));
    const float2 f2 = make_float2(fmaf(f.t2x, d0.x, fmaf(f.t2y, d1.x, f.t2z * d2.x)),
                                  fmaf(f.t2x, d0.y, fmaf(f.t2y, d1.y, f.t2z * d2.y)));
    float2 o0, o1;
    if (KIND == kTheta) {  // (−i·kn·f2, i·kn·f1)
      o0 = make_float2(f.kn * f2.y, -f.kn * f2.x);
      o1 = make_float2(-f.kn * f1.y, f.kn * f1.x);
    } else {               // (i·f2/knr, −i·f1/knr)
      const float inv = 1.0f / fmaxf(f.kn, ft.kn_floor);
      o0 = make_float2(-f2.y * inv, f2.x * inv);
      o1 = make_float2(f1.y * inv, -f1.x * inv);
    }
    const long long i0 = vbase + vox, i1 = vbase + G3 + vox;
    if (EPI == kStore) {
      if (ea.beta != 0.f) {
        const float2 x0 = ea.x[i0], x1 = ea.x[i1];
        o0 = make_float2(fmaf(ea.alpha, o0.x, ea.beta * x0.x), fmaf(ea.alpha, o0.y, ea.beta * x0.y));
        o1 = make_float2(fmaf(ea.alpha, o1.x, ea.beta * x1.x), fmaf(ea.alpha, o1.y, ea.beta * x1.y));
      } else if (ea.alpha != 1.f) {
        o0 = cscale(o0, ea.alpha);
        o1 = cscale(o1, ea.alpha);
      }
      ea.out[i0] = o0;
      ea.out[i1] = o1;
    } else {
      // T_{k+1} = 2·(alpha·Θ T_k − T_k) − T_{k−1}   (interior.py:341-348, exp_kpm_dos.py:148)
      const float2 x0 = ea.x[i0], x1 = ea.x[i1];
      const float2 p0 = ea.out[i0], p1 = ea.out[i1];
      const float2 n0 = make_float2(2.f * fmaf(ea.alpha, o0.x, -x0.x) - p0.x, 2.f * fmaf(ea.alpha, o0.y, -x0.y) - p0.y);
      const float2 n1 = make_float2(2.f * fmaf(ea.alpha, o1.x, -x1.x) - p1.x, 2.f * fmaf(ea.alpha, o1.y, -x1.y) - p1.y);
      ea.out[i0] = n0;
      ea.out[i1] = n1;
      if (EPI == kCheb) {
        const float2 y0 = ea.y[i0], y1 = ea.y[i1];
        ea.y[i0] = make_float2(fmaf(ea.ck, n0.x, y0.x), fmaf(ea.ck, n0.y, y0.y));
        ea.y[i1] = make_float2(fmaf(ea.ck, n1.x, y1.x), fmaf(ea.ck, n1.y, y1.y));
      } else if (EPI == kKpm) {
        // ⟨T_k, T_k⟩ and Re⟨T_{k+1}, T_k⟩ in FP64 (products of FP32 are exact in FP64)
        s0 += (double)x0.x * x0.x + (double)x0.y * x0.y + (double)x1.x * x1.x + (double)x1.y * x1.y;
        s1 += (double)n0.x * x0.x + (double)n0.y * x0.y + (double)n1.x * x1.x + (double)n1.y * x1.y;
      } else {  // kDot: Re⟨z, T_{k+1}⟩
        const float2 z0 = ea.z[i0], z1 = ea.z[i1];
        s0 += (double)z0.x * n0.x + (double)z0.y * n0.y + (double)z1.x * n1.x + (double)z1.y * n1.y;
      }
    }
  }
  if (EPI == kKpm || EPI == kDot) {
    const double t0 = block_sum(s0, red);
    const double t1 = block_sum(s1, red + kThreads / 64);
    if (threadIdx.x == 0) {
      ea.partial[2 * (long long)blockIdx.x] = t0;
      ea.partial[2 * (long long)blockIdx.x + 1] = t1;
    }
  }
}

}  // namespace efn
