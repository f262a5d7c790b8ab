// Split three-pass Θ·H = ∇×(ε⁻¹∇×H) for grids whose (y, z) plane does not fit
// one CU (reference eigenfns/operator.py:82-95 and :104-127):
//   xin_split   (m,2,G³) → W (m,3,G³): curl into the Cartesian frame, inverse
//               x-FFT and the first radix-2 stage of the inverse z-FFT
//   yzlds       W in place, one (vector, component, x, half) half-plane at a
//               time per persistent block: inverse z-FFT of length G/2 and
//               inverse y-FFT, × ε⁻¹/G³ (or ε/G³), forward y-FFT and forward
//               z-FFT of length G/2
//   xout_split  W → (m,2,G³): last radix-2 stage of the forward z-FFT, forward
//               x-FFT, projection on t̂₁, t̂₂, curl back and the fused epilogue
// Compulsory traffic 16+24 | 24+24(+ε) | 24+16 = 132 B per voxel and vector, as
// for the whole-plane three passes, with half the plane on chip.
//
// The z split (H = G/2, ω = exp(2πi/G), k ∈ [0, H), real z = 2m + h):
//   inverse  f(2m + h) = Σ_k e_h(k)·ω^{2km},  e_0(k) = c(k) + c(k+H),
//            e_1(k) = (c(k) − c(k+H))·ω^{k}  — two independent H-point inverse
//            DFTs, stored at W slots k (h = 0) and H + k (h = 1);
//   forward  F_h(k) = Σ_m g(2m + h)·ω^{−2km} (H-point, slot hH + k), and
//            F(k) = F_0(k) + ω^{−k}F_1(k),  F(k+H) = F_0(k) − ω^{−k}F_1(k).
// The radix-2 stages are pointwise in x and y, so the x passes do them: a tile
// holds 16 slots of each half at the same k, the two halves in lanes l and
// l ^ 16 of one wave, exchanged with a lane swizzle.
#pragma once

#include "theta3_kernels.hpp"

namespace efn {

// grids whose Θ three-pass form runs on half-planes
__host__ __device__ constexpr bool split_pass_grid(int G) { return G == 192; }

// x-pass tiles: kSplitRun slots of each half at the same k (128-B runs), 32
// lanes per step task row; one 512-thread block per tile (measured faster
// than 8-slot tiles of 64-B runs on 256-thread blocks, 3,758 against 3,541
// vector-applies/s at 192³, even with those tiles' line-sharing pairs dealt
// to one XCD).
constexpr int kSplitRun = 16;
template <int G>
__host__ __device__ constexpr int split_tiles() { return (G / 2) / kSplitRun; }
template <int G>
__host__ __device__ constexpr int xin_split_threads() { return 2 * kSplitRun * Plan<G>::N2; }
template <int G>
__host__ __device__ constexpr int xout_split_threads() { return 2 * kSplitRun * Plan<G, true>::N1; }

// block → (vector, y, z0) of an x-pass tile
template <int G>
__device__ __forceinline__ void split_tile_of(int b, int& v, int& y, int& z0) {
  constexpr int nt = split_tiles<G>();
  v = b / (G * nt);
  const int rem = b - v * (G * nt);
  y = rem / nt;
  z0 = (rem - y * nt) * kSplitRun;
}

// W slot (= wavenumber k) of lane position zz ∈ [0, 2·kSplitRun) of the tile at z0
template <int G>
__device__ __forceinline__ int split_slot(int z0, int zz) {
  return zz < kSplitRun ? z0 + zz : G / 2 + z0 + (zz - kSplitRun);
}

// value of lane ^ kSplitRun (both lanes active): ds_swizzle in bit-mask mode,
// and-mask 0x1f, or-mask 0, xor-mask kSplitRun
__device__ __forceinline__ float lane_swap_half(float v) {
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1F | (kSplitRun << 10)));
}
__device__ __forceinline__ float2 lane_swap_half(float2 v) {
  return make_float2(lane_swap_half(v.x), lane_swap_half(v.y));
}

// xin_split: one thread per step-1 task (zz, n2) of the x four-step, as xin;
// after step 3 the two halves' values of one (x, k) meet in lanes l, l ^ 8.
// The x passes hold the three components of a thread's points (36 complex at
// 192³); left at 168 VGPRs they run one 512-thread block per CU, which
// measured faster than 128 VGPRs with the spills that forces (Θ at 192³:
// 3,704 against 3,583 with xout at 128, 3,208 with both).
template <int G, int KIND>
__global__ __launch_bounds__(xin_split_threads<G>()) __attribute__((amdgpu_waves_per_eu(3)))
void xin_split_kernel(const float2* __restrict__ in,
                                                                         float2* __restrict__ ws, VecMap vm,
                                                                         long long v0,
                                                                         const float2* __restrict__ twid_g) {
  using P = Plan<G>;
  constexpr int N1 = P::N1, N2 = P::N2;
  constexpr int ZT = 2 * kSplitRun;
  constexpr int NT = xin_split_threads<G>();
  constexpr int G2 = G * G;
  constexpr long long G3 = (long long)G2 * G;
  static_assert((ZT * N1) % 64 == 0, "step-3 tasks fill whole waves (lane swizzle partners active)");
  __shared__ float2 buf[ZT * P::kStride];
  __shared__ float2 twid[G];
  int v, y, z0;
  split_tile_of<G>(blockIdx.x, v, y, z0);
  const VecRef vr = map_vector(vm, v0 + v);
  const FrameTables ft = vr.desc->ft;
  const int t = threadIdx.x;
  const int zz = t % ZT, n2 = t / ZT;
  const int zg = split_slot<G>(z0, zz);
  const float kz = gld(ft.kz + zg);
  const float ky = gld(ft.ky + y);
  float kx[N1];
  uint32_t swp[N1];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) {
    const int x = N2 * n1 + n2;
    kx[n1] = gld(ft.kx + x);
    swp[n1] = gld(ft.swap + (((x * G + y) * G + zg) >> 5));
  }
  __builtin_amdgcn_sched_barrier(0);
  const float2* a_in = in + vr.in_vec * 2 * G3 + (long long)y * G + zg;
  const float2* b_in = a_in + G3;
  float2 av[N1], bv[N1];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) {
    const long long off = (long long)(N2 * n1 + n2) * G2;
    av[n1] = a_in[off];
    bv[n1] = b_in[off];
  }
  stage_twiddles<NT, G>(twid, twid_g);
  float2 o[3][N1];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) {
    const int x = N2 * n1 + n2;
    const int vox = (x * G + y) * G + zg;
    RowFrame rf;
    rf.Kx = kx[n1];
    rf.Ky = ky;
    rf.rho2 = fmaf(rf.Kx, rf.Kx, ky * ky);
    rf.rrho = frame_rsqrt(rf.rho2);
    const Frame f = frame_at(rf, kz, swap_bit(swp[n1], vox), vox == ft.sing_voxel);
    float2 oc[3];
    curl_in<G, KIND>(f, av[n1], bv[n1], ft.kn_floor, oc);
    o[0][n1] = oc[0];
    o[1][n1] = oc[1];
    o[2][n1] = oc[2];
  }
  __syncthreads();
  float2* wbase = ws + (long long)v * 3 * G3 + (long long)y * G;
  // step-3 task s = (zs, k1), zs = s mod ZT: lanes l and l ^ 8 hold the same
  // (x, k) of the two halves
  const int zs = t % ZT, k1 = t / ZT;
  const bool hi = zs >= kSplitRun;
  const int slot = split_slot<G>(z0, zs);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    __builtin_amdgcn_sched_barrier(0);
    const float2* tw = twid + opaque_zero();
    step1_store_lean<G, +1, P>(o[c], n2, buf + zz * P::kStride, tw);
    __syncthreads();
    if (t < ZT * N1) {
      float2 wz = make_float2(1.f, 0.f);  // ω^{+k} on the upper half's difference
      if (hi) {
        wz = tw[z0 + (zs - kSplitRun)];
        wz.y = -wz.y;
      }
      float2 r[N2];
      step3_load<G, +1, P>(r, k1, buf + zs * P::kStride);
      float2* dst = wbase + c * G3 + slot;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const float2 p = lane_swap_half(r[k2]);
        const float2 d = hi ? csub(p, r[k2]) : cadd(r[k2], p);
        dst[(long long)(k1 + N1 * k2) * G2] = cmul(d, wz);
      }
    }
    __syncthreads();
  }
}

// xout_split: step-1 tasks (zs, n2) read one component at a time at their slot
// (the next one's loads in flight while the current one is transformed); the
// radix-2 combine of the two halves (lanes l, l ^ 8) precedes step 1, so from
// there on lane position zs stands for wavenumber split_slot(z0, zs).
template <int G, int KIND, int EPI>
__global__ __launch_bounds__(xout_split_threads<G>()) __attribute__((amdgpu_waves_per_eu(3)))
void xout_split_kernel(const float2* __restrict__ ws,
                                                                           VecMap vm, long long v0,
                                                                           const float2* __restrict__ twid_g,
                                                                           EpiArgs ea) {
  using P = Plan<G, true>;
  constexpr int N1 = P::N1, N2 = P::N2;
  constexpr int ZT = 2 * kSplitRun;
  constexpr int NT = xout_split_threads<G>();
  constexpr int G2 = G * G;
  constexpr long long G3 = (long long)G2 * G;
  constexpr int T1 = ZT * N2;  // step-1 tasks (zs, n2)
  static_assert(T1 % 64 == 0, "step-1 tasks fill whole waves (lane swizzle partners active)");
  __shared__ float2 buf[ZT * P::kStride];
  __shared__ float2 twid[G];
  __shared__ double red[NT / 64];
  int v, y, z0;
  split_tile_of<G>(blockIdx.x, v, y, z0);
  const VecRef vr = map_vector(vm, v0 + v);
  const FrameTables ft = vr.desc->ft;
  const int t = threadIdx.x;
  const int zs = t % ZT, n2s = t / ZT;
  const bool hi = zs >= kSplitRun;
  const float2* src = ws + (long long)v * 3 * G3 + (long long)y * G + split_slot<G>(z0, zs) +
                      (long long)n2s * G2;
  float2 cur[N1], nxt[N1];
  if (t < T1) {
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) cur[n1] = src[(long long)(N2 * n1) * G2];
  }
  stage_twiddles<NT, G>(twid, twid_g);
  __syncthreads();
  const int z = t % ZT, k1 = t / ZT;
  float2 d[3][N2];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    __builtin_amdgcn_sched_barrier(0);
    const float2* tw = twid + opaque_zero();
    if (t < T1) {
      if (c < 2) {
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) nxt[n1] = src[(c + 1) * G3 + (long long)(N2 * n1) * G2];
      }
      // ω^{−k} on the upper half before the exchange: lower lanes then add,
      // upper lanes subtract
      const float2 wz = hi ? tw[z0 + (zs - kSplitRun)] : make_float2(1.f, 0.f);
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) {
        const float2 u = cmul(cur[n1], wz);
        const float2 q = lane_swap_half(u);
        cur[n1] = hi ? csub(q, u) : cadd(u, q);
      }
      step1_store_lean<G, -1, P>(cur, n2s, buf + zs * P::kStride, tw);
    }
    __syncthreads();
    step3_load<G, -1, P>(d[c], k1, buf + z * P::kStride);
    __syncthreads();
    if (c < 2) {
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) cur[n1] = nxt[n1];
    }
  }
  // the epilogue's frame at the task's points (z = split_slot(z0, t mod ZT))
  const int zg = split_slot<G>(z0, z);
  const float kz = gld(ft.kz + zg);
  const float ky = gld(ft.ky + y);
  float kx[N2];
  uint32_t swp[N2];
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) {
    const int x = k1 + N1 * k2;
    kx[k2] = gld(ft.kx + x);
    swp[k2] = gld(ft.swap + (((x * G + y) * G + zg) >> 5));
  }
  const long long vbase = vr.out_vec * 2 * G3;
  const long long xbase = vr.in_vec * 2 * G3;
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) {
    const int x = k1 + N1 * k2;
    const int vox = (x * G + y) * G + zg;
    RowFrame rf;
    rf.Kx = kx[k2];
    rf.Ky = ky;
    rf.rho2 = fmaf(rf.Kx, rf.Kx, ky * ky);
    rf.rrho = frame_rsqrt(rf.rho2);
    const Frame f = frame_at(rf, kz, swap_bit(swp[k2], vox), vox == ft.sing_voxel);
    zout_point<KIND, EPI>(f, d[0][k2], d[1][k2], d[2][k2], ft.kn_floor, ea, vbase + vox, vbase + G3 + vox,
                          xbase + vox, xbase + G3 + vox, s0, s1);
  }
  if (EPI == kKpm || EPI == kDot) {
    const double t0 = block_sum_nt<NT>(s0, red);
    const double t1 = block_sum_nt<NT>(s1, red);
    if (threadIdx.x == 0) {
      ea.partial[2 * (long long)blockIdx.x] = t0;
      ea.partial[2 * (long long)blockIdx.x + 1] = t1;
    }
  }
}

// ---------------------------------------------------------------------------
// The persistent half-plane pass: each block (one per CU) walks a contiguous
// range of half-planes (v, x, h, c), c fastest.  A half-plane (G rows y × H
// columns) sits in LDS with an odd row stride (H + 1) while the next one's
// loads are in flight in registers, so its HBM traffic overlaps this one's
// transforms.  Phases (NT threads, all in place in LDS):
//   z⁻¹  H = 8·12: step 1 DFT-8 over z = 12·n1 + n2, step 3 DFT-12 runs; real
//        m then sits at column p = 12·(m mod 8) + m / 8
//   y⁻¹  G = 12·n1 + n2: step 1 DFT-16 over n1; middle stage per (p, k1):
//        DFT-12 over n2 (real y' = k1 + 16·k2), × ε, forward DFT-12, twiddle;
//        forward step 3 DFT-16 over k1 → natural rows
//   z    m = a + 8·b: DFT-12 over b (the column run 12a..12a+11), twiddle, then
//        DFT-8 over a, written to HBM as wavenumber k1 + 12·k2.
// Tasks of the z phases run over consecutive rows in consecutive lanes (the
// odd row stride puts them in distinct banks; the last one, which stores to
// HBM, over consecutive wavenumbers), those of the y phases over consecutive
// columns.
template <int G>
struct LdsHalf {
  static constexpr int GZ = G / 2;   // 96
  static constexpr int RS = GZ + 1;  // LDS row stride (odd)
  static constexpr int NT = 768;
  static constexpr int Z1 = 8, Z2 = 12;   // GZ = Z1·Z2
  static constexpr int Y1 = 16, Y2 = 12;  // G = Y1·Y2, y = Y2·n1 + n2
  static constexpr int PER = G * GZ / NT; // elements per thread (24)
  static constexpr int MT = GZ * Y1 / NT; // middle-stage tasks per thread (2)
  static_assert(Z1 * Z2 == GZ && Y1 * Y2 == G && (G * GZ) % NT == 0 && NT % GZ == 0, "LDS half-plane shape");
  static_assert((G * Z2) % NT == 0 && (G * Z1) % NT == 0 && (GZ * Y1) % NT == 0, "whole tasks per thread");
};

template <int G>
__global__ __launch_bounds__(LdsHalf<G>::NT) void yzlds_kernel(float2* __restrict__ ws, VecMap vm,
                                                               long long v0, int use_eps,
                                                               const float2* __restrict__ twid_g,
                                                               long long items) {
  using L = LdsHalf<G>;
  constexpr int GZ = L::GZ, RS = L::RS, NT = L::NT, Z1 = L::Z1, Z2 = L::Z2, Y1 = L::Y1, Y2 = L::Y2,
                PER = L::PER, MT = L::MT;
  constexpr int G2 = G * G;
  constexpr long long G3 = (long long)G2 * G;
  __shared__ float2 P[G * RS];
  __shared__ float2 twid[G];  // exp(−2πi e/G)
  __shared__ float2 twh[GZ];  // exp(−2πi e/(G/2))
  const int tid = threadIdx.x;
  const long long nb = gridDim.x;
  const long long it0 = (long long)blockIdx.x * items / nb;
  const long long it1 = (long long)(blockIdx.x + 1) * items / nb;
  if (it0 >= it1) return;
  stage_twiddles<NT, G>(twid, twid_g);
  for (int i = tid; i < GZ; i += NT) twh[i] = twid_g[2 * i];
  // item → (v, c, x, h) and its plane base (row stride G)
  auto plane_of = [&](long long it, int& c, long long& v, int& xh) {
    c = int(it % 3);
    const long long u = it / 3;
    v = u / (2 * G);
    xh = int(u - v * 2 * G);
    return ws + (v * 3 + c) * G3 + (long long)(xh >> 1) * G2 + (xh & 1) * GZ;
  };
  // element e = tid + NT·i of a half-plane is row r0 + (NT/GZ)·i, column c0
  constexpr int RSTEP = NT / GZ;
  const int r0 = tid / GZ, c0 = tid % GZ;
  float2 pre[PER];
  {
    int c, xh;
    long long v;
    const float2* src = plane_of(it0, c, v, xh) + r0 * G + c0;
#pragma unroll
    for (int i = 0; i < PER; ++i) pre[i] = src[RSTEP * i * G];
  }
  for (long long it = it0; it < it1; ++it) {
    // per-item copy of the thread index: the compiler would otherwise hoist
    // every phase's LDS addresses out of the loop and keep them all live
    const int tid = threadIdx.x + opaque_zero();
    int c, xh;
    long long v;
    float2* pl = plane_of(it, c, v, xh);
    const OpDesc* dsc = map_vector(vm, v0 + v).desc;
    const float* eps = (use_eps ? dsc->eps_h : dsc->inv_eps_h) + (long long)xh * G * GZ;
    __syncthreads();  // the previous item's last LDS reads are done
    {
      // (offsets rebuilt per item: the compiler would otherwise keep all PER
      // addresses live across the loop)
      float2* dst = P + r0 * RS + c0 + opaque_zero();
#pragma unroll
      for (int i = 0; i < PER; ++i) dst[RSTEP * i * RS] = pre[i];
    }
    if (it + 1 < it1) {
      int c1, xh1;
      long long v1;
      const float2* src = plane_of(it + 1, c1, v1, xh1) + r0 * G + c0 + opaque_zero();
#pragma unroll
      for (int i = 0; i < PER; ++i) pre[i] = src[RSTEP * i * G];
    }
    const int pm = c0, k1m = r0;  // this thread's middle-stage task (p, k1)
    __syncthreads();
    // ---- z⁻¹ step 1: tasks (y, n2), DFT-Z1 over z = Z2·n1 + n2
#pragma unroll
    for (int j = 0; j < G * Z2 / NT; ++j) {
      __builtin_amdgcn_sched_barrier(0);
      const int s = tid + NT * j;
      const int n2 = s / G, y = s - n2 * G;
      float2* row = P + y * RS + n2;
      float2 w[Z1];
#pragma unroll
      for (int n1 = 0; n1 < Z1; ++n1) w[n1] = row[Z2 * n1];
      DFT<Z1, +1>::run(w);
#pragma unroll
      for (int k1 = 1; k1 < Z1; ++k1) {
        float2 t = twh[n2 * k1];
        t.y = -t.y;
        w[k1] = cmul(w[k1], t);
      }
#pragma unroll
      for (int k1 = 0; k1 < Z1; ++k1) row[Z2 * k1] = w[k1];
    }
    __syncthreads();
    // ---- z⁻¹ step 3: tasks (y, k1), DFT-Z2 over the run Z2·k1 .. Z2·k1 + Z2 − 1
#pragma unroll
    for (int j = 0; j < G * Z1 / NT; ++j) {
      __builtin_amdgcn_sched_barrier(0);
      const int s = tid + NT * j;
      const int k1 = s / G, y = s - k1 * G;
      float2* run = P + y * RS + Z2 * k1;
      float2 w[Z2];
#pragma unroll
      for (int n2 = 0; n2 < Z2; ++n2) w[n2] = run[n2];
      DFT<Z2, +1>::run(w);
#pragma unroll
      for (int k2 = 0; k2 < Z2; ++k2) run[k2] = w[k2];
    }
    __syncthreads();
    // this item's ε factors for the middle stage, in flight during y⁻¹ step 1
    // (middle task j of this thread: column pm, k1 = k1m + RSTEP·j)
    float ev[MT][Y2];
    {
      const float* ep = eps + (long long)k1m * GZ + Z1 * (pm % Z2) + pm / Z2 + opaque_zero();
#pragma unroll
      for (int j = 0; j < MT; ++j)
#pragma unroll
        for (int k2 = 0; k2 < Y2; ++k2) ev[j][k2] = gld(ep + (RSTEP * j + Y1 * k2) * GZ);
    }
    // ---- y⁻¹ step 1: tasks (p, n2), DFT-Y1 over rows y = Y2·n1 + n2
    for (int s = tid; s < GZ * Y2; s += NT) {
      __builtin_amdgcn_sched_barrier(0);
      const int n2 = s / GZ, p = s - n2 * GZ;
      float2* col = P + n2 * RS + p;
      float2 w[Y1];
#pragma unroll
      for (int n1 = 0; n1 < Y1; ++n1) w[n1] = col[Y2 * n1 * RS];
      DFT<Y1, +1>::run(w);
      const float2* tw = twid + opaque_zero();
#pragma unroll
      for (int k0 = 1; k0 < Y1; k0 += 4) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k1 = k0; k1 < k0 + 4 && k1 < Y1; ++k1) {
          float2 t = tw[n2 * k1];
          t.y = -t.y;
          w[k1] = cmul(w[k1], t);
        }
      }
#pragma unroll
      for (int k1 = 0; k1 < Y1; ++k1) col[Y2 * k1 * RS] = w[k1];
    }
    __syncthreads();
    // ---- middle: tasks (p, k1): DFT-Y2 over n2 (rows Y2·k1 + n2), × ε, forward
    // DFT-Y2, twiddle ω^{−k1·f1}
#pragma unroll
    for (int j = 0; j < MT; ++j) {
      __builtin_amdgcn_sched_barrier(0);
      const int k1 = k1m + RSTEP * j;
      float2* col = P + Y2 * k1 * RS + pm;
      float2 w[Y2];
#pragma unroll
      for (int n2 = 0; n2 < Y2; ++n2) w[n2] = col[n2 * RS];
      DFT<Y2, +1>::run(w);
#pragma unroll
      for (int k2 = 0; k2 < Y2; ++k2) w[k2] = cscale(w[k2], ev[j][k2]);
      DFT<Y2, -1>::run(w);
      const float2* tw = twid + opaque_zero();
#pragma unroll
      for (int f0 = 1; f0 < Y2; f0 += 4) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int f1 = f0; f1 < f0 + 4 && f1 < Y2; ++f1) w[f1] = cmul(w[f1], tw[k1 * f1]);
      }
#pragma unroll
      for (int f1 = 0; f1 < Y2; ++f1) col[f1 * RS] = w[f1];
    }
    __syncthreads();
    // ---- forward y step 3: tasks (p, f1), DFT-Y1 over k1 → row f1 + Y2·f2
    for (int s = tid; s < GZ * Y2; s += NT) {
      __builtin_amdgcn_sched_barrier(0);
      const int f1 = s / GZ, p = s - f1 * GZ;
      float2* col = P + f1 * RS + p;
      float2 w[Y1];
#pragma unroll
      for (int k1 = 0; k1 < Y1; ++k1) w[k1] = col[Y2 * k1 * RS];
      DFT<Y1, -1>::run(w);
#pragma unroll
      for (int f2 = 0; f2 < Y1; ++f2) col[Y2 * f2 * RS] = w[f2];
    }
    __syncthreads();
    // ---- forward z step 1: tasks (y, a), DFT-Z2 over the run of m = a + Z1·b
    // (columns Z2·a + b), twiddle ω_H^{−k1·a}
#pragma unroll
    for (int j = 0; j < G * Z1 / NT; ++j) {
      __builtin_amdgcn_sched_barrier(0);
      const int s = tid + NT * j;
      const int a = s / G, y = s - a * G;
      float2* run = P + y * RS + Z2 * a;
      float2 w[Z2];
#pragma unroll
      for (int b = 0; b < Z2; ++b) w[b] = run[b];
      DFT<Z2, -1>::run(w);
#pragma unroll
      for (int k1 = 1; k1 < Z2; ++k1) w[k1] = cmul(w[k1], twh[k1 * a]);
#pragma unroll
      for (int k1 = 0; k1 < Z2; ++k1) run[k1] = w[k1];
    }
    __syncthreads();
    // ---- forward z step 3: tasks (y, k1), DFT-Z1 over a → wavenumber k1 + Z2·k2
#pragma unroll
    for (int j = 0; j < G * Z2 / NT; ++j) {
      __builtin_amdgcn_sched_barrier(0);
      const int s = tid + NT * j;
      const int y = s / Z2, k1 = s - y * Z2;  // k1 fastest: 96-B store runs
      const float2* col = P + y * RS + k1;
      float2 w[Z1];
#pragma unroll
      for (int a = 0; a < Z1; ++a) w[a] = col[Z2 * a];
      DFT<Z1, -1>::run(w);
      float2* dst = pl + (long long)y * G + k1;
#pragma unroll
      for (int k2 = 0; k2 < Z1; ++k2) dst[Z2 * k2] = w[k2];
    }
  }
}

}  // namespace efn
