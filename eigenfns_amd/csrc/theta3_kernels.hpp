// Three-pass Θ·H = ∇×(ε⁻¹∇×H) (reference eigenfns/operator.py:82-95) and the
// transverse-projection preconditioner (:104-127):
//   xin    (m,2,G³) → W (m,3,G³): curl into the Cartesian frame, inverse x-FFT
//   yzmid  W in place, one (vector, component, x) plane per block: inverse z-
//          and y-FFTs, × ε⁻¹/G³ (or ε/G³), forward y- and z-FFTs
//   xout   W → (m,2,G³): forward x-FFT, projection on t̂₁, t̂₂, curl back and
//          the fused epilogue (store / Chebyshev / KPM / KPM-count)
// Compulsory traffic 16+24 | 24+24(+ε) | 24+16 = 132 B per voxel and vector,
// against 228 B for the five-pass form (zin | y⁻¹ | x⁻¹·ε·x | y | zout).
//
// The plane pass holds a whole G×G plane of one component on chip: the y-FFT's
// middle stage (inverse step 3, the real-space multiply, forward step 1) keeps
// it in registers, while the z-FFTs and the first/last y stages stream it
// through LDS in batches of whole row groups.  Inverse y: y = Y2·n1 + n2,
// step 1 is a DFT-Y1 over n1 (rows ≡ n2 mod Y2), step 3 a DFT-Y2 over n2 giving
// y' = k1 + Y1·k2; the forward y-FFT starts from that layout (DFT-Y2 over k2,
// twiddle, DFT-Y1 over k1) and ends with rows ≡ f1 mod Y2 — the same row groups,
// so the inverse and forward batches cover the same rows.
#pragma once

#include "operator_kernels.hpp"

namespace efn {

// ---------------------------------------------------------------------------
// x passes: tiles of 16 consecutive z (128-B runs) × all x at one (vector, y).
constexpr int kX3Tile = 16;
template <int G>
__host__ __device__ constexpr int xin_threads() { return kX3Tile * Plan<G>::N2; }
template <int G>
__host__ __device__ constexpr int xout_threads() { return kX3Tile * Plan<G, true>::N1; }
template <int G>
__host__ __device__ constexpr int x3_tiles() { return G / kX3Tile; }

// A zero the compiler cannot see through: offsetting the twiddle table by it
// in each batch keeps the loads of one batch from being merged with (and held
// live until) another batch's — in the register-bound plane pass the merged
// twiddles of every stage stayed live across the whole kernel.
__device__ __forceinline__ int opaque_zero() {
  int z = 0;
  asm volatile("" : "+v"(z));
  return z;
}

template <int NT, int G>
__device__ __forceinline__ void stage_twiddles(float2* __restrict__ twid, const float2* __restrict__ twid_g) {
  for (int i = threadIdx.x; i < G; i += NT) twid[i] = twid_g[i];
}

// Sum of one double over a block of NT threads (result in thread 0).
template <int NT>
__device__ __forceinline__ double block_sum_nt(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += red[i];
  }
  return s;
}

// xin: one thread per step-1 task (z, n2) of the x four-step (x = N2·n1 + n2):
// it loads both transverse components at its N1 points, forms the three
// Cartesian components there, and runs step 1 for each; step 3 writes W.
template <int G, int KIND>
__global__ __launch_bounds__(xin_threads<G>()) void xin_kernel(const float2* __restrict__ in,
                                                             float2* __restrict__ ws, VecMap vm,
                                                             long long v0,
                                                             const float2* __restrict__ twid_g) {
  using P = Plan<G>;
  constexpr int N1 = P::N1, N2 = P::N2;
  constexpr int ZT = kX3Tile;
  constexpr int NT = xin_threads<G>();
  constexpr int nzt = x3_tiles<G>();
  constexpr int G2 = G * G;
  constexpr long long G3 = (long long)G2 * G;
  __shared__ float2 buf[ZT * P::kStride];
  __shared__ float2 twid[G];
  const int b = blockIdx.x;
  const int v = b / (G * nzt);
  const int rem = b - v * (G * nzt);
  const int y = rem / nzt;
  const int z0 = (rem - y * nzt) * ZT;
  const VecRef vr = map_vector(vm, v0 + v);
  const FrameTables ft = vr.desc->ft;
  const int t = threadIdx.x;
  const int z = t % ZT, n2 = t / ZT;
  const int zg = z0 + z;
  // frame tables ahead of the data (vector loads complete in issue order)
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
  float2* wbase = ws + (long long)v * 3 * G3 + (long long)y * G + z0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    step1_store<G, +1, P>(o[c], n2, buf + z * P::kStride, twid);
    __syncthreads();
    for (int s = t; s < ZT * N1; s += NT) {
      const int zz = s % ZT, k1 = s / ZT;
      float2 r[N2];
      step3_load<G, +1, P>(r, k1, buf + zz * P::kStride);
      float2* dst = wbase + c * G3 + zz;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) dst[(long long)(k1 + N1 * k2) * G2] = r[k2];
    }
    __syncthreads();
  }
}

// xout: the x four-step with the factors swapped (step-3 tasks (z, k1) hold
// x = k1 + N1·k2); each thread runs step 3 for all three components of its
// task, then projects, curls back and applies the epilogue at its N2 points.
template <int G, int KIND, int EPI>
__global__ __launch_bounds__(xout_threads<G>()) void xout_kernel(const float2* __restrict__ ws, VecMap vm,
                                                               long long v0,
                                                               const float2* __restrict__ twid_g,
                                                               EpiArgs ea) {
  using P = Plan<G, true>;
  constexpr int N1 = P::N1, N2 = P::N2;
  constexpr int ZT = kX3Tile;
  constexpr int NT = xout_threads<G>();
  constexpr int nzt = x3_tiles<G>();
  constexpr int G2 = G * G;
  constexpr long long G3 = (long long)G2 * G;
  constexpr int T1 = ZT * N2;  // step-1 tasks (z, n2)
  __shared__ float2 buf[ZT * P::kStride];
  __shared__ float2 twid[G];
  __shared__ double red[NT / 64];
  const int b = blockIdx.x;
  const int v = b / (G * nzt);
  const int rem = b - v * (G * nzt);
  const int y = rem / nzt;
  const int z0 = (rem - y * nzt) * ZT;
  const VecRef vr = map_vector(vm, v0 + v);
  const FrameTables ft = vr.desc->ft;
  const int t = threadIdx.x;
  const float2* wbase = ws + (long long)v * 3 * G3 + (long long)y * G + z0;
  // step-1 inputs of all three components first
  float2 in1[3][N1];
  const int zs = t % ZT, n2s = t / ZT;
  if (t < T1) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) in1[c][n1] = wbase[c * G3 + (long long)(N2 * n1 + n2s) * G2 + zs];
  }
  // the epilogue's frame tables
  const int z = t % ZT, k1 = t / ZT;
  const int zg = z0 + z;
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
  stage_twiddles<NT, G>(twid, twid_g);
  __syncthreads();
  float2 d[3][N2];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (t < T1) step1_store<G, -1, P>(in1[c], n2s, buf + zs * P::kStride, twid);
    __syncthreads();
    step3_load<G, -1, P>(d[c], k1, buf + z * P::kStride);
    __syncthreads();
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
// The plane pass.  MidCfg<G>: z plans for the inverse (PZ: step-1 loads of
// 16 consecutive z) and forward (PZF: step-3 stores of 16 consecutive z)
// transforms, the y plan PY (Y1 = PY::N1 step-1 size, Y2 = PY::N2 the middle
// radix), the block size NT, the row groups per batch BG (NT = G·BG: one
// y step-1 / last-stage task per thread per batch) and WPE, the waves per SIMD
// the register allocation must leave room for.
template <int G>
struct MidCfg;
template <>
struct MidCfg<128> {
  using PZ = Plan<128, false>;  // 8 × 16
  using PZF = Plan<128, true>;  // 16 × 8
  using PY = Plan<128, true>;   // Y1 = 16, Y2 = 8
  static constexpr int NT = 512, BG = 4;
  // 4 waves per SIMD (≤ 128 VGPRs) lets two 512-thread blocks share a CU, so
  // one block's loads and stores overlap the other's exchanges: +8.5 % on Θ
  // at 128³ over the unconstrained 139 VGPRs (one block per CU), although 10
  // VGPRs then spill (44 B per lane of scratch, served from cache)
  static constexpr int WPE = 4;
};
template <>
struct MidCfg<96> {
  using PZ = Plan<96, false>;   // 8 × 12
  using PZF = Plan<96, true>;   // 12 × 8
  // Y1 = 12, Y2 = 8: the register-resident middle stage runs DFT-8s instead
  // of DFT-12s (111 instead of 137 VGPRs, four waves per SIMD): +21 % on Θ at
  // 96³ (35,178 against 29,086 vector-applies/s at 9-vector chunks)
  using PY = Plan<96, true>;
  static constexpr int NT = 384, BG = 4;
  static constexpr int WPE = 1;
};
template <>
struct MidCfg<48> {
  using PZ = Plan<48, false>;   // 6 × 8
  using PZF = Plan<48, true>;   // 8 × 6
  using PY = Plan<48, false>;   // Y1 = 6, Y2 = 8
  static constexpr int NT = 96, BG = 2;  // (G·Y1 = 288 rows of tasks: 96-thread blocks)
  static constexpr int WPE = 1;
};
template <>
struct MidCfg<32> {
  using PZ = Plan<32, false>;   // 4 × 8
  using PZF = Plan<32, true>;   // 8 × 4
  using PY = Plan<32, false>;   // Y1 = 4, Y2 = 8
  static constexpr int NT = 128, BG = 4;
  static constexpr int WPE = 1;
};
template <>
struct MidCfg<64> {
  using PZ = Plan<64, false>;   // 8 × 8
  using PZF = Plan<64, true>;
  using PY = Plan<64, false>;   // Y1 = 8, Y2 = 8
  static constexpr int NT = 256, BG = 4;
  static constexpr int WPE = 1;
};

template <int G>
__host__ __device__ constexpr int mid_threads() { return MidCfg<G>::NT; }
template <int G>
__host__ __device__ constexpr int mid_waves_per_eu() { return MidCfg<G>::WPE; }

template <int G>
struct MidShape {
  using C = MidCfg<G>;
  static constexpr int Y1 = C::PY::N1, Y2 = C::PY::N2, BG = C::BG, NT = C::NT;
  static constexpr int RT = G * Y1 / NT;  // resident (z, k1) tasks per thread
  static constexpr int NB = Y2 / BG;      // batches
  static constexpr int R = BG * Y1;       // rows per batch
  static constexpr int KS = C::PZ::kStride > C::PZF::kStride ? C::PZ::kStride : C::PZF::kStride;
  static_assert(Y1 * Y2 == G && Y2 % BG == 0 && G * BG == NT && (G * Y1) % NT == 0, "plane pass shape");
  static_assert(NT % G == 0, "resident tasks keep z = thread mod G");
  static_assert(G * BG * Y1 <= R * KS, "exchange fits the row buffer");
};

// W plane (v, c, x) in place; blocks b ≡ xcd (mod 8) run on one XCD, and the
// three components of a plane are dealt to one XCD back to back, so its ε
// plane comes from HBM once and from that XCD's L2 for the other two.
template <int G>
__global__ __launch_bounds__(mid_threads<G>()) __attribute__((amdgpu_waves_per_eu(mid_waves_per_eu<G>())))
void yzmid_kernel(float2* __restrict__ ws, VecMap vm,
                                                                long long v0, int use_eps,
                                                                const float2* __restrict__ twid_g,
                                                                long long planes) {
  using C = MidCfg<G>;
  using S = MidShape<G>;
  using PZ = typename C::PZ;
  using PZF = typename C::PZF;
  constexpr int Y1 = S::Y1, Y2 = S::Y2, BG = S::BG, NT = S::NT, RT = S::RT, R = S::R, KS = S::KS;
  constexpr int G2 = G * G;
  constexpr long long G3 = (long long)G2 * G;
  __shared__ float2 buf[R * KS];
  __shared__ float2 twid[G];
  const long long b = blockIdx.x;
  long long plane;
  int c;
  if ((planes & 7) == 0) {
    const long long q = b >> 3;
    c = int(q % 3);
    plane = (q / 3) * 8 + (b & 7);
  } else {
    c = int(b % 3);
    plane = b / 3;
  }
  const long long v = plane / G;
  const int x = int(plane - v * G);
  float2* pl = ws + (v * 3 + c) * G3 + (long long)x * G2;
  const OpDesc* dsc = map_vector(vm, v0 + v).desc;
  const float* eps = (use_eps ? dsc->eps : dsc->inv_eps) + (long long)x * G2;
  const int tid = threadIdx.x;
  stage_twiddles<NT, G>(twid, twid_g);
  __syncthreads();
  // resident tasks (z, k1): τ = tid + j·NT, z = τ mod G, k1 = τ / G
  const int zr = tid % G;
  const int k1r = tid / G;  // + j·NT/G
  float2 res[RT][Y2];
  // local row r of batch bb ↔ plane row y = Y2·(r mod Y1) + bb·BG + r / Y1
  auto row_of = [](int bb, int r) { return Y2 * (r % Y1) + bb * BG + r / Y1; };
  // natural z after the inverse z step 3 done in place in each run
  auto zpos = [](int zz) { return (PZ::N2 + 1) * (zz % PZ::N1) + zz / PZ::N1; };
  // the one y-stage task of this thread in each batch: (z, g)
  const int zy = tid % G, gy = tid / G;

  // ---- inverse z and y step 1, batch by batch --------------------------------
#pragma unroll
  for (int bb = 0; bb < S::NB; ++bb) {
    const float2* tw = twid + opaque_zero();
    for (int s = tid; s < R * PZ::N2; s += NT) {
      const int r = s / PZ::N2, n2 = s - r * PZ::N2;
      const float2* src = pl + (long long)row_of(bb, r) * G;
      float2 vv[PZ::N1];
#pragma unroll
      for (int n1 = 0; n1 < PZ::N1; ++n1) vv[n1] = src[PZ::N2 * n1 + n2];
      step1_store<G, +1, PZ>(vv, n2, buf + r * KS, tw);
    }
    __syncthreads();
    for (int s = tid; s < R * PZ::N1; s += NT) {
      const int r = s / PZ::N1, k1 = s - r * PZ::N1;
      float2 o[PZ::N2];
      step3_load<G, +1, PZ>(o, k1, buf + r * KS);
      float2* run = buf + r * KS + (PZ::N2 + 1) * k1;
#pragma unroll
      for (int k2 = 0; k2 < PZ::N2; ++k2) run[k2] = o[k2];
    }
    __syncthreads();
    float2 yv[Y1];
#pragma unroll
    for (int n1 = 0; n1 < Y1; ++n1) yv[n1] = buf[(gy * Y1 + n1) * KS + zpos(zy)];
    __syncthreads();
    DFT<Y1, +1>::run(yv);
    {
      const int n2 = bb * BG + gy;
#pragma unroll
      for (int k1 = 0; k1 < Y1; ++k1) {
        if (k1 > 0) {
          float2 w = tw[n2 * k1];
          w.y = -w.y;
          yv[k1] = cmul(yv[k1], w);
        }
        buf[(gy * Y1 + k1) * G + zy] = yv[k1];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RT; ++j) {
      const int k1 = k1r + j * (NT / G);
#pragma unroll
      for (int g = 0; g < BG; ++g) res[j][bb * BG + g] = buf[(g * Y1 + k1) * G + zr];
    }
    __syncthreads();
  }

  // ---- y step 3 (inverse), × ε, forward y step 1, in registers ----------------
  // (factors and twiddles fetched four at a time: this stage runs with the
  // whole plane resident, so its own working set is kept small)
#pragma unroll
  for (int j = 0; j < RT; ++j) {
    __builtin_amdgcn_sched_barrier(0);
    const float2* tw = twid + opaque_zero();
    const int k1 = k1r + j * (NT / G);
    const float* ej = eps + (long long)k1 * G + zr;
    DFT<Y2, +1>::run(res[j]);
#pragma unroll
    for (int k0 = 0; k0 < Y2; k0 += 4) {
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k2 = k0; k2 < k0 + 4 && k2 < Y2; ++k2)
        res[j][k2] = cscale(res[j][k2], gld(ej + (long long)Y1 * k2 * G));
    }
    DFT<Y2, -1>::run(res[j]);
#pragma unroll
    for (int f0 = 1; f0 < Y2; f0 += 4) {
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int f1 = f0; f1 < f0 + 4 && f1 < Y2; ++f1) res[j][f1] = cmul(res[j][f1], tw[k1 * f1]);
    }
    // the first forward batch's values go to the (idle) exchange area now, so
    // only the later batches' stay in registers through the forward phase
#pragma unroll
    for (int g = 0; g < BG; ++g) buf[(g * Y1 + k1) * G + zr] = res[j][g];
  }

  // ---- forward y step 3 and z, batch by batch --------------------------------
  auto zfpos = [](int zz) { return PZF::pos(zz); };
#pragma unroll
  for (int bb = 0; bb < S::NB; ++bb) {
    __builtin_amdgcn_sched_barrier(0);
    const float2* tw = twid + opaque_zero();
    if (bb > 0) {
#pragma unroll
      for (int j = 0; j < RT; ++j) {
        const int k1 = k1r + j * (NT / G);
#pragma unroll
        for (int g = 0; g < BG; ++g) buf[(g * Y1 + k1) * G + zr] = res[j][bb * BG + g];
      }
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
    float2 fv[Y1];
#pragma unroll
    for (int k1 = 0; k1 < Y1; ++k1) fv[k1] = buf[(gy * Y1 + k1) * G + zy];
    __syncthreads();
    DFT<Y1, -1>::run(fv);
#pragma unroll
    for (int f2 = 0; f2 < Y1; ++f2) buf[(gy * Y1 + f2) * KS + zfpos(zy)] = fv[f2];
    __syncthreads();
    for (int s = tid; s < R * PZF::N2; s += NT) {
      const int r = s / PZF::N2, n2 = s - r * PZF::N2;
      float2* pen = buf + r * KS;
      float2 vv[PZF::N1];
#pragma unroll
      for (int n1 = 0; n1 < PZF::N1; ++n1) vv[n1] = pen[(PZF::N2 + 1) * n1 + n2];
      step1_store<G, -1, PZF>(vv, n2, pen, tw);
    }
    __syncthreads();
    for (int s = tid; s < R * PZF::N1; s += NT) {
      const int r = s / PZF::N1, k1 = s - r * PZF::N1;
      float2 o[PZF::N2];
      step3_load<G, -1, PZF>(o, k1, buf + r * KS);
      float2* dst = pl + (long long)row_of(bb, r) * G;
#pragma unroll
      for (int k2 = 0; k2 < PZF::N2; ++k2) dst[k1 + PZF::N1 * k2] = o[k2];
    }
    __syncthreads();
  }
}

}  // namespace efn
