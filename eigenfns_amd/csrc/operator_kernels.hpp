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
// bit-packed FP64 reference-axis switch computed on the host.  The x-pass
// multiply maps the three components of a tile to one XCD so ε is fetched once.
#pragma once

#include "fft.hpp"

#ifndef EFN_FRAME_RSQ
#define EFN_FRAME_RSQ 1
#endif
#ifndef EFN_ZWAVE_ROWS
#define EFN_ZWAVE_ROWS 1
#endif
#ifndef EFN_ZIN_SWAP
#define EFN_ZIN_SWAP 1
#endif
#ifndef EFN_XEPS_TILED
#define EFN_XEPS_TILED 1
#endif
#ifndef EFN_XMID_PAIR
#define EFN_XMID_PAIR 1
#endif
#ifndef EFN_XMID_EPS_EARLY
#define EFN_XMID_EPS_EARLY 1
#endif

namespace efn {

enum OpKind { kTheta = 0, kPrecond = 1, kCartH = 2 };
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

// Per-operator device descriptor (one per context, in device memory).
struct OpDesc {
  FrameTables ft;
  const float* inv_eps;     // ε⁻¹/G³ per voxel
  const float* eps;         // ε/G³ per voxel
  const float* inv_eps_x;   // the same two tables in the x pass's tile order:
  const float* eps_x;       // [y][z / ZT][x][z mod ZT], ZT = xtile_for(G)
  const float* inv_eps_h;   // the same two tables de-interleaved in z for the
  const float* eps_h;       // half-plane pass: [x][h][y][m] = (x, y, z = 2m + h)
};

// Vector map of one launch: launch-local vector v belongs to slot v / nrow and
// is row v % nrow of that slot (nrow > 0), or — ragged launches, nrow = 0 — to
// the slot s with row0[s] ≤ v < row0[s + 1], row v − row0[s]; the slot's
// operator, input and output blocks are given per slot (offsets in vectors
// from the launch's base pointers).  One slot is the plain single-operator
// call; several slots batch operators of one grid size (k-points, structures,
// requests, of equal or different block widths) into the same passes.
constexpr int kMaxSlots = 64;
struct VecMap {
  const OpDesc* desc[kMaxSlots];
  long long in_off[kMaxSlots];
  long long out_off[kMaxSlots];
  int row0[kMaxSlots + 1];
  int nrow;
  int nslot;
};

struct VecRef {
  const OpDesc* desc;
  long long in_vec, out_vec;  // absolute vector indices from the base pointers
};

__device__ __forceinline__ VecRef map_vector(const VecMap& vm, long long gv) {
  // a launch covers fewer than 2³¹ vectors: 32-bit arithmetic
  int slot;
  long long row;
  if (vm.nrow > 0) {
    slot = int(gv) / vm.nrow;
    row = int(gv) - slot * vm.nrow;
  } else {  // binary search of the (block-uniform) vector in the row prefix table
    int lo = 0, hi = vm.nslot - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (vm.row0[mid] <= int(gv)) lo = mid;
      else hi = mid - 1;
    }
    slot = lo;
    row = int(gv) - vm.row0[lo];
  }
  VecRef r;
  r.desc = vm.desc[slot];
  r.in_vec = vm.in_off[slot] + row;
  r.out_vec = vm.out_off[slot] + row;
  return r;
}

// The transverse frame of operator.py:53-71 (t̂₁ = K×ref/|K×ref|,
// t̂₂ = K×t̂₁/|K×t̂₁|, zero at K = 0; ref = x̂ where the FP64 switch bit is set,
// else ẑ) in branch-free closed forms, algebraically identical, in FP32:
//   reference axis ẑ:  t̂₁ = (K_y, −K_x, 0)/ρ,   t̂₂ = (K_x·K_z/ρ, K_y·K_z/ρ, −ρ)/|K|,  ρ = √(K_x²+K_y²)
//   reference axis x̂:  t̂₁ = (0, K_z, −K_y)/σ,   t̂₂ = (−σ, K_x·K_y/σ, K_x·K_z/σ)/|K|,  σ = √(K_y²+K_z²)
// ρ > 0 wherever ẑ is used and σ > 0 wherever x̂ is used (away from K = 0).
struct RowFrame {
  float Kx, Ky, rho2, rrho;
};

// Loads through pointers that were themselves read from memory (the operator
// descriptor's tables) compile to flat loads, which count against the LDS
// counter too and so make every LDS wait also wait for them; these tables are
// always in global memory, so load them as such.
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

// 1/√x of the frame: the hardware reciprocal square root (≤ 1 ulp) in place
// of a correctly rounded square root and division (about 20 instructions).
__device__ __forceinline__ float frame_rsqrt(float x) {
#if EFN_FRAME_RSQ
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}

__device__ __forceinline__ RowFrame row_frame(const FrameTables& ft, int x, int y) {
  RowFrame rf;
  rf.Kx = gld(ft.kx + x);
  rf.Ky = gld(ft.ky + y);
  rf.rho2 = fmaf(rf.Kx, rf.Kx, rf.Ky * rf.Ky);
  rf.rrho = frame_rsqrt(rf.rho2);
  return rf;
}

// Branch-free: the reference axis picks operands, one product per component;
// at K = 0 both scale factors are zero, which zeroes every component.
__device__ __forceinline__ Frame frame_at(const RowFrame& rf, float Kz, bool sw, bool sing) {
  const float Kx = rf.Kx, Ky = rf.Ky;
  const float kn2 = fmaf(Kz, Kz, rf.rho2);
  const float rkn = frame_rsqrt(kn2);
  const float sig2 = fmaf(Ky, Ky, Kz * Kz);
  const float rsig = frame_rsqrt(sig2);
  const float r1 = sing ? 0.f : (sw ? rsig : rf.rrho);  // 1/ρ (axis ẑ) or 1/σ (axis x̂)
  const float r2 = sing ? 0.f : r1 * rkn;
  Frame f;
  f.kn = sing ? 0.f : kn2 * rkn;
  f.t1x = (sw ? 0.f : Ky) * r1;
  f.t1y = (sw ? Kz : -Kx) * r1;
  f.t1z = (sw ? -Ky : 0.f) * r1;
  f.t2x = (sw ? -sig2 : Kx * Kz) * r2;
  f.t2y = (Ky * (sw ? Kx : Kz)) * r2;
  f.t2z = (sw ? Kx * Kz : -rf.rho2) * r2;
  return f;
}

__device__ __forceinline__ bool swap_bit(uint32_t word, long long vox) { return (word >> (vox & 31)) & 1u; }

__device__ __forceinline__ Frame point_frame(const FrameTables& ft, const RowFrame& rf, int z, long long vox) {
  return frame_at(rf, gld(ft.kz + z), swap_bit(gld(ft.swap + (vox >> 5)), vox), vox == ft.sing_voxel);
}

// z-tile width of the strided (y, x) passes: 16 complex = 128 B runs.
__host__ __device__ constexpr int ztile_for(int G) { return G >= 16 ? 16 : G; }

// The y pass (AXIS 1) at 128³ and 192³: tiles of EFN_YTILE z on 512-thread
// blocks (32 z: 256-B runs, 6 instead of 5 waves per SIMD) measured +1.4 % on
// Θ at 128³ and +0.1 % on Θ, +0.4 % on the fused Chebyshev/KPM steps at 192³.
#ifndef EFN_YTILE
#define EFN_YTILE 32
#endif
#ifndef EFN_YTILE_MAXG
#define EFN_YTILE_MAXG 224
#endif
template <int G, int AXIS>
__host__ __device__ constexpr int strided_tile() {
  return AXIS == 1 && G >= 128 && G < EFN_YTILE_MAXG ? EFN_YTILE : ztile_for(G);
}
template <int G, int AXIS>
__host__ __device__ constexpr int strided_threads() { return strided_tile<G, AXIS>() > 16 ? 512 : 256; }

// z-tile width of the x pass with the real-space multiply (xmid): 32 z (256-B
// runs per x) on 512-thread blocks from 64³ up, measured +1.2 % on Θ at 192³
// and 128³, +0.4-0.5 % at 64³-96³ and −2 % at 48³ against 16 z.
#ifndef EFN_XTILE
#define EFN_XTILE 32
#endif

__host__ __device__ constexpr int xtile_for(int G) { return G >= 64 ? EFN_XTILE : (G >= 16 ? 16 : G); }
// x-pass block size: one step-1 task per thread (512 threads for 32-z tiles)
template <int G>
__host__ __device__ constexpr int xmid_threads() { return xtile_for(G) > 16 ? 512 : 256; }

// Rows per zin/zout block when all three components are in LDS at once: a
// power of two dividing G² (a block never spans two vectors).  Measured on
// MI355X (Θ, 32-vector blocks): 4 rows at G ≥ 160 and 8 at 64 ≤ G < 160 beat
// 8/16 by 7 % at 192³ and 14 % at 96³ — more, smaller blocks per CU hide the
// HBM latency of the direct stage-1/stage-3 accesses better than fuller ones
// (8 rows over 4 at 128³: +0.8 %, at 192³: −1.5 %); below 64, 16 rows beat 8
// and 32 by 5-6 % at 48³ and 32³ (64-96 vectors).
#ifndef EFN_ZROWS_BIG
#define EFN_ZROWS_BIG 4
#endif
#ifndef EFN_ZROWS_MID
#define EFN_ZROWS_MID 8
#endif
#ifndef EFN_ZROWS_SMALL
#define EFN_ZROWS_SMALL 16
#endif
__host__ __device__ constexpr int zrows3_for(int G) {
  int target = G >= 160 ? EFN_ZROWS_BIG : (G >= 64 ? EFN_ZROWS_MID : EFN_ZROWS_SMALL);
  int p = 1;
  while (p * 2 <= target && (G * G) % (p * 2) == 0) p *= 2;
  return p;
}
template <int G>
__host__ __device__ constexpr int zrows3() { return zrows3_for(G); }

constexpr int kThreads = 256;

// The pass kernels' twiddle table, staged per block in LDS once the block's
// first data loads are in flight (read from the global table instead, which
// leaves the LDS to the pencils and allows more blocks per CU, it measured
// 7 % slower).
#define EFN_TWID_DECL(G) __shared__ float2 twid[G]
#define EFN_TWID_STAGE(G) for (int i = threadIdx.x; i < (G); i += kThreads) twid[i] = twid_g[i]

// ---------------------------------------------------------------------------
// The curl into the Cartesian frame at one voxel (zin's pointwise step).
template <int G, int KIND>
__device__ __forceinline__ void curl_in(const Frame& f, float2 a, float2 b, float kn_floor, float2 (&o)[3]) {
  float2 c1, c2;
  if (KIND == kTheta) {  // c1 = −i·kn·b, c2 = i·kn·a
    c1 = make_float2(f.kn * b.y, -f.kn * b.x);
    c2 = make_float2(-f.kn * a.y, f.kn * a.x);
  } else if (KIND == kCartH) {  // Cartesian H = a·t̂₁ + b·t̂₂ (operator.py:169-172)
    c1 = a;
    c2 = b;
  } else {               // c1 = i·b/knr, c2 = −i·a/knr
    const float inv = 1.0f / fmaxf(f.kn, kn_floor);
    c1 = make_float2(-b.y * inv, b.x * inv);
    c2 = make_float2(a.y * inv, -a.x * inv);
  }
  o[0] = make_float2(fmaf(c1.x, f.t1x, c2.x * f.t2x), fmaf(c1.y, f.t1x, c2.y * f.t2x));
  o[1] = make_float2(fmaf(c1.x, f.t1y, c2.x * f.t2y), fmaf(c1.y, f.t1y, c2.y * f.t2y));
  o[2] = make_float2(fmaf(c1.x, f.t1z, c2.x * f.t2z), fmaf(c1.y, f.t1z, c2.y * f.t2z));
}

// Wave-row layout of the z passes' pointwise phase: when a row is a whole
// number of 64-point runs and the R rows split evenly over the four waves,
// wave w owns rows w·R/4 .. (w+1)·R/4 − 1, so (x, y) and the row frame are
// uniform across the wave and each lane's points are z = lane + 64·j.
template <int G>
__host__ __device__ constexpr bool z_wave_rows() {
  return EFN_ZWAVE_ROWS && G % 64 == 0 && zrows3<G>() % (kThreads / 64) == 0;
}

// zin: (nvec, 2, G³) → W (nvec, 3, G³), curl into the Cartesian frame + inverse
// z-FFT.  A coalesced prologue reads each input point once and writes the three
// Cartesian components in natural order to LDS; the four-step FFT runs on the
// 3·R pencils; step-3 outputs go straight to HBM.
template <int G, int KIND>
__global__ __launch_bounds__(kThreads) void zin_kernel(const float2* __restrict__ in,
                                                       float2* __restrict__ ws, VecMap vm,
                                                       long long v0,
                                                       const float2* __restrict__ twid_g) {
  // the swapped factors (wider step-3 store runs) measured +0.7 % on Θ at 192³
  // and +1.1 % at 48³, −0.3 % and −0.6 % at 128³ and 96³
  using P = Plan<G, EFN_ZIN_SWAP != 0 && (G >= 160 || G < 64)>;
  constexpr int N1 = P::N1, N2 = P::N2;
  constexpr int R = zrows3<G>();
  constexpr long long G2 = (long long)G * G, G3 = G2 * G;
  __shared__ float2 buf[3 * R * P::kStride];
  EFN_TWID_DECL(G);
  const long long row0 = (long long)blockIdx.x * R;  // flattened (v, x, y)
  const int v = int(row0 / G2);
  const long long xy0 = row0 - (long long)v * G2;
  const VecRef vr = map_vector(vm, v0 + v);
  const FrameTables ft = vr.desc->ft;
  const float2* a_in = in + vr.in_vec * 2 * G3;
  const float2* b_in = a_in + G3;
  if constexpr (z_wave_rows<G>()) {
    constexpr int RW = R / (kThreads / 64);  // rows per wave
    constexpr int ZJ = G / 64;               // points per lane per row
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int rw0 = w * RW;
    // frame tables first, then the data: vector loads complete in issue order,
    // so the frame arithmetic can start while the data is still in flight
    float kz[ZJ];
    uint32_t swp[RW][ZJ];
#pragma unroll
    for (int j = 0; j < ZJ; ++j) kz[j] = gld(ft.kz + lane + 64 * j);
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
      for (int j = 0; j < ZJ; ++j) {
        const int vox = int((xy0 + rw0 + r) * G) + lane + 64 * j;
        swp[r][j] = gld(ft.swap + (vox >> 5));
      }
    __builtin_amdgcn_sched_barrier(0);  // keep the table loads ahead of the data
    float2 apre[RW][ZJ], bpre[RW][ZJ];
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
      for (int j = 0; j < ZJ; ++j) {
        const int vox = int((xy0 + rw0 + r) * G) + lane + 64 * j;
        apre[r][j] = a_in[vox];
        bpre[r][j] = b_in[vox];
      }
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int xy = int(xy0) + rw0 + r;
      const int x = xy / G, y = xy - x * G;
      const RowFrame rf = row_frame(ft, x, y);
#pragma unroll
      for (int j = 0; j < ZJ; ++j) {
        const int z = lane + 64 * j;
        const int vox = xy * G + z;
        const Frame f = frame_at(rf, kz[j], swap_bit(swp[r][j], vox), vox == ft.sing_voxel);
        float2 o[3];
        curl_in<G, KIND>(f, apre[r][j], bpre[r][j], ft.kn_floor, o);
        const int p = P::pos(z);
#pragma unroll
        for (int c = 0; c < 3; ++c) buf[(c * R + rw0 + r) * P::kStride + p] = o[c];
      }
    }
  } else {
    // every input point of this thread loaded before any frame arithmetic, so
    // the loads are in flight together (measured +1.3 % on Θ at 192³)
    constexpr int NE = (R * G + kThreads - 1) / kThreads;
    float2 apre[NE], bpre[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int e = threadIdx.x + k * kThreads;
      if ((R * G) % kThreads == 0 || e < R * G) {
        const long long vox = xy0 * G + e;
        apre[k] = a_in[vox];
        bpre[k] = b_in[vox];
      }
    }
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int e = threadIdx.x + k * kThreads;
      if (!((R * G) % kThreads == 0 || e < R * G)) continue;
      const int r = e / G, z = e - r * G;
      const long long xy = xy0 + r;
      const int x = int(xy / G), y = int(xy - (long long)x * G);
      const long long vox = xy * G + z;
      const Frame f = point_frame(ft, row_frame(ft, x, y), z, vox);
      float2 o[3];
      curl_in<G, KIND>(f, apre[k], bpre[k], ft.kn_floor, o);
      const int p = P::pos(z);
#pragma unroll
      for (int c = 0; c < 3; ++c) buf[(c * R + r) * P::kStride + p] = o[c];
    }
  }
  // staged after the data loads are issued, so its round trip overlaps theirs
  EFN_TWID_STAGE(G);
  __syncthreads();
  for (int t = threadIdx.x; t < 3 * R * N2; t += kThreads) {
    const int p = t / N2, n2 = t - p * N2;
    float2* pen = buf + p * P::kStride;
    float2 vv[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) vv[n1] = pen[(N2 + 1) * n1 + n2];  // pos(N2·n1 + n2)
    step1_store<G, +1, P>(vv, n2, pen, twid);
  }
  __syncthreads();
  float2* w = ws + (long long)v * 3 * G3;
  // step 3 straight to HBM: task k1 writes X[k1 + N1·k2], runs of N1 consecutive
  // outputs (with the swapped factors, 16 at 192³: whole 128-B lines)
  for (int t = threadIdx.x; t < 3 * R * N1; t += kThreads) {
    const int p = t / N1, k1 = t - p * N1;
    const int c = p / R, r = p - c * R;
    float2 o[N2];
    step3_load<G, +1, P>(o, k1, buf + p * P::kStride);
    float2* dst = w + c * G3 + (xy0 + r) * G;
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) dst[k1 + N1 * k2] = o[k2];
  }
}

// ---------------------------------------------------------------------------
// Strided pass over W (nvec·3 fields) in place.  AXIS 1 = y (stride G),
// 0 = x (stride G²).  Tile = ZT consecutive z × all G positions on the axis;
// HBM → registers (step 1) → LDS → registers (step 3) → HBM.
template <int G, int AXIS, int S>
__global__ __launch_bounds__((strided_threads<G, AXIS>())) void strided_kernel(float2* __restrict__ ws,
                                                           const float2* __restrict__ twid_g) {
  using P = Plan<G>;
  constexpr int N1 = P::N1, N2 = P::N2;
  constexpr int ZT = strided_tile<G, AXIS>();
  constexpr int NT = strided_threads<G, AXIS>();
  constexpr long long G2 = (long long)G * G, G3 = G2 * G;
  constexpr long long stride = AXIS == 1 ? G : G2;
  __shared__ float2 buf[ZT * P::kStride];
  EFN_TWID_DECL(G);
  constexpr int nzt = (G + ZT - 1) / ZT;
  const long long fo = blockIdx.x / nzt;  // (field, other-axis index)
  const int zt = blockIdx.x - int(fo * nzt);
  const long long field = fo / G;
  const int o = int(fo - field * G);
  const int z0 = zt * ZT;
  float2* base = ws + field * G3 + (AXIS == 1 ? (long long)o * G2 : (long long)o * G) + z0;
  auto active = [&](int t) { return t < ZT * N2 && z0 + t % ZT < G; };
  auto load1 = [&](int t, float2 (&v)[N1]) {
    const int z = t % ZT, n2 = t / ZT;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = base[(N2 * n1 + n2) * stride + z];
  };
  int t = threadIdx.x;
  {  // the first step-1 inputs are in flight before the twiddle table is staged
    float2 v[N1];
    if (active(t)) load1(t, v);
    EFN_TWID_STAGE(G);
    __syncthreads();
    if (active(t)) step1_store<G, S>(v, t / ZT, buf + (t % ZT) * P::kStride, twid);
    t += NT;
  }
  for (; t < ZT * N2; t += NT) {
    if (active(t)) {
      float2 v[N1];
      load1(t, v);
      step1_store<G, S>(v, t / ZT, buf + (t % ZT) * P::kStride, twid);
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ZT * N1; t += NT) {
    const int z = t % ZT, k1 = t / ZT;
    if (z0 + z < G) {
      float2 r[N2];
      step3_load<G, S>(r, k1, buf + z * P::kStride);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) base[(k1 + N1 * k2) * stride + z] = r[k2];
    }
  }
}

// x-pass with the real-space multiply: inverse x-FFT, ×rmul (ε⁻¹/G³ or ε/G³),
// forward x-FFT.  The three Cartesian components of one tile run on blocks of
// the same XCD, back to back, so the rmul tile is fetched from HBM once and
// served from that XCD's L2 to the other two.
template <int G>
__global__ __launch_bounds__((xmid_threads<G>())) void xmid_kernel(float2* __restrict__ ws, VecMap vm,
                                                        long long v0, int use_eps,
                                                        const float2* __restrict__ twid_g,
                                                        long long tiles) {
  using P = Plan<G>;
  constexpr int N1 = P::N1, N2 = P::N2;
  constexpr int ZT = xtile_for(G);
  constexpr int NT = xmid_threads<G>();
  constexpr long long G2 = (long long)G * G, G3 = G2 * G;
  constexpr int nzt = (G + ZT - 1) / ZT;
  // forward-transform rows (register-resident variant): (N1+1)·N2 slots, odd stride
  constexpr int kFStrideRaw = (N1 + 1) * N2;
  constexpr int kFStride = (kFStrideRaw % 2 == 0) ? kFStrideRaw + 1 : kFStrideRaw;
  constexpr int kBufStride = kFStride > P::kStride ? kFStride : P::kStride;
  __shared__ float2 buf[ZT * kBufStride];
  EFN_TWID_DECL(G);
  const long long b = blockIdx.x;
  long long tile;
  int c;
  if (EFN_XMID_PAIR && G > 64 && G < 160 && (tiles & 15) == 0 && nzt % 2 == 0) {
    // as below, and each XCD takes the two z-tiles of a 32-z span in turn: their
    // ε rows (16 floats = 64 B per x) share 128-B lines, which one XCD's L2
    // then fetches once instead of two XCDs fetching each (x-pass fetches −13 %
    // at 192³; Θ +1.4 % at 96³ and +0.8 % at 128³, but −0.8 % at 192³ and
    // −0.1 % at 64³, so there the tiles are dealt out singly)
    const long long q = b >> 3, sq = q / 3;
    c = int(q - 3 * sq);
    tile = (sq >> 1) * 16 + (b & 7) * 2 + (sq & 1);
  } else if ((tiles & 7) == 0) {  // XCD-aware: blocks b ≡ xcd (mod 8) run on one XCD
    const long long q = b >> 3;
    c = int(q % 3);
    tile = (q / 3) * 8 + (b & 7);
  } else {
    c = int(b % 3);
    tile = b / 3;
  }
  const long long v = tile / (G * nzt);
  const int rem = int(tile - v * (G * nzt));
  const int o = rem / nzt;  // y
  const int z0 = (rem - o * nzt) * ZT;
  float2* base = ws + (v * 3 + c) * G3 + (long long)o * G + z0;
  const OpDesc* dsc = map_vector(vm, v0 + v).desc;
#if EFN_XEPS_TILED
  // a tile's factors are one contiguous run of G·ZT floats (whole 128-B lines:
  // in the voxel order a tile row is 64 B, half a line shared with the next
  // z-tile, which another XCD fetches again)
  const float* eps = (use_eps ? dsc->eps_x : dsc->inv_eps_x) + ((long long)o * nzt + z0 / ZT) * G * ZT;
  constexpr long long kEpsX = ZT;  // between consecutive x
#else
  const float* eps = (use_eps ? dsc->eps : dsc->inv_eps) + (long long)o * G + z0;
  constexpr long long kEpsX = G2;
#endif
  constexpr int T3 = ZT * N1;
  constexpr int TPT = (T3 + NT - 1) / NT;
  // inverse step 1: HBM → registers → LDS
  auto active = [&](int t) { return t < ZT * N2 && z0 + t % ZT < G; };
  auto load1 = [&](int t, float2 (&vv)[N1]) {
    const int z = t % ZT, n2 = t / ZT;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) vv[n1] = base[(N2 * n1 + n2) * G2 + z];
  };
  // this thread's ε⁻¹ (or ε) factors for its step-3 tasks, fetched once the
  // step-1 inputs have arrived and consumed after step 3, so their round trip
  // overlaps the first half of the transform (measured +0.6 % on Θ at 192³,
  // −0.3 to −0.8 % at 48³-128³, so only on the large grids)
  constexpr bool kEpsEarly = EFN_XMID_EPS_EARLY && G >= 160;
  float epre[TPT][N2];
  auto load_eps = [&]() {
#pragma unroll
    for (int j = 0; j < TPT; ++j) {
      const int t = threadIdx.x + j * NT;
      const int z = t % ZT, k1 = t / ZT;
      if (t < T3 && z0 + z < G) {
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) epre[j][k2] = gld(eps + (long long)(k1 + N1 * k2) * kEpsX + z);
      }
    }
  };
  int t1 = threadIdx.x;
  {  // the first step-1 inputs are in flight before the twiddle table is staged
    float2 vv[N1];
    if (active(t1)) load1(t1, vv);
    EFN_TWID_STAGE(G);
    __syncthreads();
    if constexpr (kEpsEarly) load_eps();
    if (active(t1)) step1_store<G, +1>(vv, t1 / ZT, buf + (t1 % ZT) * P::kStride, twid);
    t1 += NT;
  }
  for (; t1 < ZT * N2; t1 += NT) {
    if (active(t1)) {
      float2 vv[N1];
      load1(t1, vv);
      step1_store<G, +1>(vv, t1 / ZT, buf + (t1 % ZT) * P::kStride, twid);
    }
  }
  __syncthreads();
  // inverse step 3, real-space multiply, back to LDS in natural order
  float2 r[TPT][N2];
#pragma unroll
  for (int j = 0; j < TPT; ++j) {
    const int t = threadIdx.x + j * NT;
    const int z = t % ZT, k1 = t / ZT;
    if (t < T3 && z0 + z < G) {
      step3_load<G, +1>(r[j], k1, buf + z * P::kStride);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const float sc = kEpsEarly ? epre[j][k2] : gld(eps + (long long)(k1 + N1 * k2) * kEpsX + z);
        r[j][k2] = cscale(r[j][k2], sc);
      }
    }
  }
  __syncthreads();
  // Forward x-FFT straight from registers: a task holds x = k1 + N1·k2
  // (k2 = 0..N2−1), the input of the transposed four-step x = n2' + N1·n1'
  // (n2' = k1, n1' = k2): DFT-N2 over k2 → f1', twiddle ω^{k1·f1'}, then
  // DFT-N1 over k1 per f1' → f = f1' + N2·f2'.  LDS rows hold element
  // (n2', f1') at n2' + (N1+1)·f1' (odd stride for the strided writes).
  float2* fbuf = buf;  // reused: every inverse step-3 read has completed (barrier above)
#pragma unroll
  for (int j = 0; j < TPT; ++j) {
    const int t = threadIdx.x + j * NT;
    const int z = t % ZT, k1 = t / ZT;
    if (t < T3 && z0 + z < G) {
      DFT<N2, -1>::run(r[j]);
#pragma unroll
      for (int f1 = 1; f1 < N2; ++f1) r[j][f1] = cmul(r[j][f1], twid[k1 * f1]);
#pragma unroll
      for (int f1 = 0; f1 < N2; ++f1) fbuf[z * kFStride + k1 + (N1 + 1) * f1] = r[j][f1];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ZT * N2; t += NT) {
    const int z = t % ZT, f1 = t / ZT;
    if (z0 + z < G) {
      float2 o2[N1];
#pragma unroll
      for (int n2 = 0; n2 < N1; ++n2) o2[n2] = fbuf[z * kFStride + n2 + (N1 + 1) * f1];
      DFT<N1, -1>::run(o2);
#pragma unroll
      for (int f2 = 0; f2 < N1; ++f2) base[(long long)(f1 + N2 * f2) * G2 + z] = o2[f2];
    }
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

// The pointwise epilogue of zout for one voxel: project the three Cartesian
// components d on t̂₁, t̂₂, curl back, and apply the fused update at output
// index i0/i1 (components 0/1) and input index j0/j1.
template <int KIND, int EPI>
__device__ __forceinline__ void zout_point(const Frame& f, float2 d0, float2 d1, float2 d2, float kn_floor,
                                           const EpiArgs& ea, long long i0, long long i1, long long j0,
                                           long long j1, double& s0, double& s1) {
  const float2 f1 = make_float2(fmaf(f.t1x, d0.x, fmaf(f.t1y, d1.x, f.t1z * d2.x)),
                                fmaf(f.t1x, d0.y, fmaf(f.t1y, d1.y, f.t1z * d2.y)));
  const float2 f2 = make_float2(fmaf(f.t2x, d0.x, fmaf(f.t2y, d1.x, f.t2z * d2.x)),
                                fmaf(f.t2x, d0.y, fmaf(f.t2y, d1.y, f.t2z * d2.y)));
  const float kn = f.kn;
  float2 o0, o1;
  if (KIND == kTheta) {  // (−i·kn·f2, i·kn·f1)
    o0 = make_float2(kn * f2.y, -kn * f2.x);
    o1 = make_float2(-kn * f1.y, kn * f1.x);
  } else {               // (i·f2/knr, −i·f1/knr)
    const float inv = 1.0f / fmaxf(kn, kn_floor);
    o0 = make_float2(-f2.y * inv, f2.x * inv);
    o1 = make_float2(f1.y * inv, -f1.x * inv);
  }
  if (EPI == kStore) {
    if (ea.beta != 0.f) {
      const float2 x0 = ea.x[j0], x1 = ea.x[j1];
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
    const float2 x0 = ea.x[j0], x1 = ea.x[j1];
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

// zout: W (nvec,3,G³) → forward z-FFT → project on t̂₁, t̂₂ → curl back →
// epilogue.  Step-1 inputs of the three components come straight from HBM;
// step 3 writes back to LDS in natural order; the epilogue reads the three
// components per point and applies the fused update with coalesced HBM access.
template <int G, int KIND, int EPI>
__global__ __launch_bounds__(kThreads) void zout_kernel(const float2* __restrict__ ws, VecMap vm,
                                                        long long v0,
                                                        const float2* __restrict__ twid_g,
                                                        EpiArgs ea) {
  using P = Plan<G>;
  constexpr int N1 = P::N1, N2 = P::N2;
  constexpr int R = zrows3<G>();
  constexpr long long G2 = (long long)G * G, G3 = G2 * G;
  constexpr bool kWaveRows = z_wave_rows<G>();
  constexpr int RW = kWaveRows ? R / (kThreads / 64) : 1;  // rows per wave
  constexpr int ZJ = kWaveRows ? G / 64 : 1;               // points per lane per row
  __shared__ float2 buf[3 * R * P::kStride];
  EFN_TWID_DECL(G);
  __shared__ double red[2 * kThreads / 64];
  const long long row0 = (long long)blockIdx.x * R;
  const int v = int(row0 / G2);
  const long long xy0 = row0 - (long long)v * G2;
  const VecRef vr = map_vector(vm, v0 + v);
  const FrameTables ft = vr.desc->ft;
  const float2* w = ws + (long long)v * 3 * G3;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // wave-row layout: the epilogue's frame tables are fetched first of all
  float kz[ZJ];
  uint32_t swp[RW][ZJ];
  if constexpr (kWaveRows) {
#pragma unroll
    for (int j = 0; j < ZJ; ++j) kz[j] = gld(ft.kz + lane + 64 * j);
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
      for (int j = 0; j < ZJ; ++j) {
        const int vox = int((xy0 + wv * RW + r) * G) + lane + 64 * j;
        swp[r][j] = gld(ft.swap + (vox >> 5));
      }
    __builtin_amdgcn_sched_barrier(0);
  }
  constexpr int T1 = 3 * R * N2;
  auto step1_src = [&](int t, float2 (&vv)[N1]) {
    const int p = t / N2, n2 = t - p * N2;
    const int c = p / R, r = p - c * R;
    const float2* src = w + c * G3 + (xy0 + r) * G;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) vv[n1] = src[N2 * n1 + n2];
  };
  // the first step-1 inputs are in flight before the twiddle table is staged
  {
    float2 vv[N1];
    const int t = threadIdx.x;
    if (t < T1) step1_src(t, vv);
    EFN_TWID_STAGE(G);
    __syncthreads();
    if (t < T1) step1_store<G, -1>(vv, t % N2, buf + (t / N2) * P::kStride, twid);
  }
  for (int t = threadIdx.x + kThreads; t < T1; t += kThreads) {
    float2 vv[N1];
    step1_src(t, vv);
    step1_store<G, -1>(vv, t % N2, buf + (t / N2) * P::kStride, twid);
  }
  __syncthreads();
  // step 3 in place: the task (pencil, k1) reads its own run and writes its
  // results X[k1 + N1·k2] back into that run at (N2 + 1)·k1 + k2 (no other task
  // touches the run, so no barrier between the reads and the writes); the
  // epilogue finds z at zpos(z) = (N2 + 1)·(z mod N1) + z / N1
  for (int t = threadIdx.x; t < 3 * R * N1; t += kThreads) {
    const int p = t / N1, k1 = t - p * N1;
    float2 r[N2];
    step3_load<G, -1>(r, k1, buf + p * P::kStride);
    float2* run = buf + p * P::kStride + (N2 + 1) * k1;
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) run[k2] = r[k2];
  }
  __syncthreads();
  auto zpos = [](int z) { return (N2 + 1) * (z % N1) + z / N1; };
  const long long vbase = vr.out_vec * 2 * G3;   // out, y, z blocks
  const long long xbase = vr.in_vec * 2 * G3;    // the operator's input block
  double s0 = 0.0, s1 = 0.0;
  if constexpr (kWaveRows) {
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int rr = wv * RW + r;
      const int xy = int(xy0) + rr;
      const int x = xy / G, y = xy - x * G;
      const RowFrame rf = row_frame(ft, x, y);
#pragma unroll
      for (int j = 0; j < ZJ; ++j) {
        const int z = lane + 64 * j;
        const int vox = xy * G + z;
        const int p = zpos(z);
        const Frame f = frame_at(rf, kz[j], swap_bit(swp[r][j], vox), vox == ft.sing_voxel);
        zout_point<KIND, EPI>(f, buf[(0 * R + rr) * P::kStride + p], buf[(1 * R + rr) * P::kStride + p],
                              buf[(2 * R + rr) * P::kStride + p], ft.kn_floor, ea, vbase + vox,
                              vbase + G3 + vox, xbase + vox, xbase + G3 + vox, s0, s1);
      }
    }
  } else {
    // a fixed trip count, unrolled, so the epilogue loads of all of a thread's
    // points can be issued together (+0.8 % on Θ, +1.4-1.8 % fused steps at 192³)
    constexpr int NE = (R * G + kThreads - 1) / kThreads;
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int e = threadIdx.x + k * kThreads;
      if ((R * G) % kThreads == 0 || e < R * G) {
        const int r = e / G, z = e - r * G;
        const long long xy = xy0 + r;
        const int x = int(xy / G), y = int(xy - (long long)x * G);
        const long long vox = xy * G + z;
        const int p = zpos(z);
        const Frame f = point_frame(ft, row_frame(ft, x, y), z, vox);
        zout_point<KIND, EPI>(f, buf[(0 * R + r) * P::kStride + p], buf[(1 * R + r) * P::kStride + p],
                              buf[(2 * R + r) * P::kStride + p], ft.kn_floor, ea, vbase + vox,
                              vbase + G3 + vox, xbase + vox, xbase + G3 + vox, s0, s1);
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

// ---------------------------------------------------------------------------
// Real-space fields from the inverse-transformed workspace W (3 Cartesian
// components, unnormalised inverse FFT) of the vectors v0 .. v0+nv−1:
//   kRealH:  H(r) = W/G³                                  (operator.py:174-176)
//   kRealE:  E(r) = (ε⁻¹/G³)·W/√λ                         (operator.py:178-184)
//   kRealEd: ε|E|² summed over the components, FP32        (run_modes.py:120-122)
// out: (m,3,G³) complex64 for H and E, (m,G³) float32 for ε|E|²; rows by the
// vector map's output offsets; lam: eigenvalue per input vector (NULL: 1).
enum RealMode { kRealH = 0, kRealE = 1, kRealEd = 2 };

template <int MODE>
__global__ __launch_bounds__(kThreads) void real_epilogue_kernel(const float2* __restrict__ ws, VecMap vm,
                                                                 long long v0, long long count,
                                                                 long long G3, float inv_n,
                                                                 float n_f, const float* __restrict__ lam,
                                                                 void* __restrict__ out) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= count) return;
  const long long v = i / G3, vox = i - v * G3;
  const VecRef vr = map_vector(vm, v0 + v);
  const float2* w = ws + v * 3 * G3 + vox;
  const float2 d0 = w[0], d1 = w[G3], d2 = w[2 * G3];
  if (MODE == kRealH) {
    float2* o = static_cast<float2*>(out) + vr.out_vec * 3 * G3 + vox;
    o[0] = cscale(d0, inv_n);
    o[G3] = cscale(d1, inv_n);
    o[2 * G3] = cscale(d2, inv_n);
    return;
  }
  const float rs = lam ? 1.0f / sqrtf(lam[vr.in_vec]) : 1.0f;
  const float sc = gld(vr.desc->inv_eps + vox) * rs;
  const float2 e0 = cscale(d0, sc), e1 = cscale(d1, sc), e2 = cscale(d2, sc);
  if (MODE == kRealE) {
    float2* o = static_cast<float2*>(out) + vr.out_vec * 3 * G3 + vox;
    o[0] = e0;
    o[G3] = e1;
    o[2 * G3] = e2;
  } else {
    const float eps = gld(vr.desc->eps + vox) * n_f;  // ε (the table holds ε/G³)
    const float s2 = (e0.x * e0.x + e0.y * e0.y) + (e1.x * e1.x + e1.y * e1.y) +
                     (e2.x * e2.x + e2.y * e2.y);
    static_cast<float*>(out)[vr.out_vec * G3 + vox] = eps * s2;
  }
}

}  // namespace efn
