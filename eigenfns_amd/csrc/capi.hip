// C ABI of libeigenfns_hip.so: operator contexts and kernel dispatch.
// Declarations and reference citations: include/eigenfns_hip.h.
#include <hip/hip_runtime.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/eigenfns_hip.h"
#include "blas_kernels.hpp"
#include "eig_kernels.hpp"
#include "heevql_kernels.hpp"
#include "cholqr_kernels.hpp"
#include "operator_kernels.hpp"
#include "raster_kernels.hpp"
#include "theta3s_kernels.hpp"


namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define EFN_HIP_CHECK(expr)                                                              \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return fail(EFN_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));          \
  } while (0)

#define EFN_LAUNCH_CHECK()                                                               \
  do {                                                                                   \
    hipError_t _e = hipGetLastError();                                                   \
    if (_e != hipSuccess)                                                                \
      return fail(EFN_EHIP, std::string("kernel launch: ") + hipGetErrorString(_e));     \
  } while (0)

using efn::EpiArgs;
using efn::FrameTables;

struct OpRequest {
  int kind;              // efn::kTheta / kPrecond
  int epi;               // efn::Epi
  const float2* in;      // operator input base (T_k)
  int m;                 // vectors in the launch (slots × rows)
  efn::VecMap vm;        // per-slot operator and block offsets
  EpiArgs ea;            // output / epilogue base pointers
  int reduce_mode;       // 0 none, 1 KPM moments, 2 KPM count
  double* mom;           // KPM: moments base
  int ld;
  int k;
  double* est;           // KPM count
  double ck64;
};

}  // namespace

struct efn_ctx {
  int G = 0;
  int passes = 0;                // 0: the engine's choice; 3 or 5: this context's (efn_ctx_set_passes)
  double L = 0.0;
  double k[3] = {0, 0, 0};
  int chunk = 0;
  long long sing_voxel = -1;
  float kn_floor = 0.f;
  double eps_mean_inv = 0.0;
  float* d_axes = nullptr;       // kx | ky | kz  (3G floats)
  uint32_t* d_swap = nullptr;    // G³ bits
  float* d_inv_eps = nullptr;    // ε⁻¹/G³
  float* d_eps = nullptr;        // ε/G³
  float* d_kn2 = nullptr;        // |k+G|² (for the diagonal preconditioner)
  float* d_xeps = nullptr;       // ε⁻¹/G³ then ε/G³ in the x pass's tile order
  float* d_heps = nullptr;       // ε⁻¹/G³ then ε/G³ de-interleaved in z (half-plane grids)
  float2* d_twid = nullptr;      // exp(−2πi e/G)
  float2* d_ws = nullptr;        // chunk·3·G³ complex64
  double* d_partial = nullptr;   // chunk·(G²/R)·2
  double2* d_red = nullptr;      // chunk·2 doubles of reduced sums
  efn::OpDesc* d_desc = nullptr; // this operator's descriptor for the pass kernels
  size_t ws_bytes = 0;

  int (*run)(efn_ctx*, const OpRequest&, hipStream_t) = nullptr;
  int (*real)(efn_ctx*, const OpRequest&, int, const float*, void*, hipStream_t) = nullptr;

  FrameTables tables() const {
    FrameTables ft;
    ft.kx = d_axes;
    ft.ky = d_axes + G;
    ft.kz = d_axes + 2 * G;
    ft.swap = d_swap;
    ft.sing_voxel = sing_voxel;
    ft.kn_floor = kn_floor;
    return ft;
  }
};

namespace {

// Process-wide Θ pass structure (efn_theta_engine): EFN_THETA_AUTO takes the
// per-grid choice of theta_passes_auto, 3 and 5 force one form where compiled.
std::atomic<int> g_theta_engine{EFN_THETA_AUTO};

// The plane pass holds one G×G plane of a component on chip: it fits one CU's
// registers and LDS up to 128³.  At 192³ the plane (295 KB) does not, and the
// three passes run on half-planes (theta3s_kernels.hpp); 256³ keeps five.
template <int G>
constexpr bool has_three_pass() {
  return G == 32 || G == 48 || G == 64 || G == 96 || G == 128 || efn::split_pass_grid(G);
}

// The pass structure used for grid G when the engine is EFN_THETA_AUTO: three
// passes wherever whole planes fit (Θ on a 32-vector block, 10,000-point
// network, default chunk: 32³ 818,490 vs 580,530 vector-applies/s, 48³
// 292,861 vs 231,907, 64³ 128,026 vs 105,982, 96³ 29,107 vs 28,800, 128³
// 14,413 vs 13,342), five at 192³, where the half-plane three passes move 137
// instead of 229 B per voxel but issue-bound measured 3,758 against 3,860
// vector-applies/s.
template <int G>
constexpr int theta_passes_auto() { return has_three_pass<G>() && !efn::split_pass_grid(G) ? 3 : 5; }

// A forced engine (3 or 5) wins; under EFN_THETA_AUTO a context's own choice
// (ctx_passes, e.g. from the tuning table) wins over the compiled default.
template <int G>
int theta_passes(int ctx_passes = 0) {
  const int e = g_theta_engine.load();
  const int want = e != EFN_THETA_AUTO ? e : (ctx_passes ? ctx_passes : theta_passes_auto<G>());
  return (want == 3 && has_three_pass<G>()) ? 3 : 5;
}

template <int G>
int theta_passes_default() { return theta_passes<G>(0); }

template <int G>
bool three_pass_compiled() { return has_three_pass<G>(); }

// reduce per-block partials of one chunk and apply the KPM formulas
__global__ void reduce_partials_kernel(const double* __restrict__ partial, int nb, int nvec,
                                       int mode, double* __restrict__ mom, int ld, int k,
                                       double* __restrict__ est, double ck) {
  __shared__ double red[8];
  const int v = blockIdx.x;
  double s0 = 0.0, s1 = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    s0 += partial[2 * ((long long)v * nb + b)];
    s1 += partial[2 * ((long long)v * nb + b) + 1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = s0;
    red[4 + (threadIdx.x >> 6)] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = red[0] + red[1] + red[2] + red[3];
    double t1 = red[4] + red[5] + red[6] + red[7];
    if (mode == 1) {
      double* row = mom + (long long)v * ld;
      row[2 * k] = 2.0 * t0 - row[0];
      row[2 * k + 1] = 2.0 * t1 - row[1];
    } else if (mode == 2) {
      est[v] += ck * t0;
    }
  }
}

// xout with the request's operator kind and epilogue
template <int G>
void launch_xout(const OpRequest& rq, const EpiArgs& ea, dim3 grid, long long v0, const float2* ws,
                 const float2* twid, hipStream_t s) {
  using namespace efn;
  const dim3 blk(xout_threads<G>());
  if (rq.kind == kPrecond) {
    hipLaunchKernelGGL((xout_kernel<G, kPrecond, kStore>), grid, blk, 0, s, ws, rq.vm, v0, twid, ea);
    return;
  }
  switch (rq.epi) {
    case kStore:
      hipLaunchKernelGGL((xout_kernel<G, kTheta, kStore>), grid, blk, 0, s, ws, rq.vm, v0, twid, ea);
      break;
    case kCheb:
      hipLaunchKernelGGL((xout_kernel<G, kTheta, kCheb>), grid, blk, 0, s, ws, rq.vm, v0, twid, ea);
      break;
    case kKpm:
      hipLaunchKernelGGL((xout_kernel<G, kTheta, kKpm>), grid, blk, 0, s, ws, rq.vm, v0, twid, ea);
      break;
    default:
      hipLaunchKernelGGL((xout_kernel<G, kTheta, kDot>), grid, blk, 0, s, ws, rq.vm, v0, twid, ea);
      break;
  }
}

// compute units of the current device (persistent launches size their grid by it)
long long cu_count() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
    n = 256;
  return n;
}

// xout_split with the request's operator kind and epilogue
template <int G>
void launch_xout_split(const OpRequest& rq, const EpiArgs& ea, dim3 grid, long long v0, const float2* ws,
                       const float2* twid, hipStream_t s) {
  using namespace efn;
  const dim3 blk(xout_split_threads<G>());
  if (rq.kind == kPrecond) {
    hipLaunchKernelGGL((xout_split_kernel<G, kPrecond, kStore>), grid, blk, 0, s, ws, rq.vm, v0, twid, ea);
    return;
  }
  switch (rq.epi) {
    case kStore:
      hipLaunchKernelGGL((xout_split_kernel<G, kTheta, kStore>), grid, blk, 0, s, ws, rq.vm, v0, twid, ea);
      break;
    case kCheb:
      hipLaunchKernelGGL((xout_split_kernel<G, kTheta, kCheb>), grid, blk, 0, s, ws, rq.vm, v0, twid, ea);
      break;
    case kKpm:
      hipLaunchKernelGGL((xout_split_kernel<G, kTheta, kKpm>), grid, blk, 0, s, ws, rq.vm, v0, twid, ea);
      break;
    default:
      hipLaunchKernelGGL((xout_split_kernel<G, kTheta, kDot>), grid, blk, 0, s, ws, rq.vm, v0, twid, ea);
      break;
  }
}

// Split three passes per chunk: xin_split | yzlds | xout_split (theta3s_kernels.hpp).
template <int G>
int run_op3s(efn_ctx* c, const OpRequest& rq, hipStream_t s) {
  using namespace efn;
  constexpr int nb = G * split_tiles<G>();  // xin/xout blocks per vector
  const int use_eps = rq.kind == kTheta ? 0 : 1;
  for (int v0 = 0; v0 < rq.m; v0 += c->chunk) {
    const int nv = rq.m - v0 < c->chunk ? rq.m - v0 : c->chunk;
    float2* const ws = c->d_ws;
    const long long halves = (long long)nv * G * 2;
    const dim3 gx((unsigned)(nv * nb));
    if (rq.kind == kTheta)
      hipLaunchKernelGGL((xin_split_kernel<G, kTheta>), gx, dim3(xin_split_threads<G>()), 0, s, rq.in, ws,
                         rq.vm, (long long)v0, c->d_twid);
    else
      hipLaunchKernelGGL((xin_split_kernel<G, kPrecond>), gx, dim3(xin_split_threads<G>()), 0, s, rq.in, ws,
                         rq.vm, (long long)v0, c->d_twid);
    EFN_LAUNCH_CHECK();
    // persistent: one block per CU walks its range of half-planes
    const long long items = 3 * halves;
    const long long nblk = items < cu_count() ? items : cu_count();
    hipLaunchKernelGGL((yzlds_kernel<G>), dim3((unsigned)nblk), dim3(LdsHalf<G>::NT), 0, s, ws, rq.vm,
                       (long long)v0, use_eps, c->d_twid, items);
    EFN_LAUNCH_CHECK();
    EpiArgs ea = rq.ea;
    ea.x = rq.in;
    ea.partial = c->d_partial;
    launch_xout_split<G>(rq, ea, gx, (long long)v0, ws, c->d_twid, s);
    EFN_LAUNCH_CHECK();
    if (rq.reduce_mode) {
      hipLaunchKernelGGL(reduce_partials_kernel, dim3(nv), dim3(256), 0, s, c->d_partial, nb, nv,
                         rq.reduce_mode, rq.mom ? rq.mom + (long long)v0 * rq.ld : nullptr, rq.ld,
                         rq.k, rq.est ? rq.est + v0 : nullptr, rq.ck64);
      EFN_LAUNCH_CHECK();
    }
  }
  return EFN_OK;
}

// Three passes per chunk: xin | yzmid | xout (theta3_kernels.hpp).
template <int G>
int run_op3(efn_ctx* c, const OpRequest& rq, hipStream_t s) {
  using namespace efn;
  if constexpr (split_pass_grid(G)) return run_op3s<G>(c, rq, s);
  else {
  constexpr int nb = G * x3_tiles<G>();  // xin/xout blocks per vector
  const int use_eps = rq.kind == kTheta ? 0 : 1;
  for (int v0 = 0; v0 < rq.m; v0 += c->chunk) {
    const int nv = rq.m - v0 < c->chunk ? rq.m - v0 : c->chunk;
    float2* const ws = c->d_ws;
    const dim3 gx((unsigned)(nv * nb)), gm((unsigned)(3 * nv * G));
    if (rq.kind == kTheta)
      hipLaunchKernelGGL((xin_kernel<G, kTheta>), gx, dim3(xin_threads<G>()), 0, s, rq.in, ws, rq.vm,
                         (long long)v0, c->d_twid);
    else
      hipLaunchKernelGGL((xin_kernel<G, kPrecond>), gx, dim3(xin_threads<G>()), 0, s, rq.in, ws, rq.vm,
                         (long long)v0, c->d_twid);
    EFN_LAUNCH_CHECK();
    hipLaunchKernelGGL((yzmid_kernel<G>), gm, dim3(mid_threads<G>()), 0, s, ws, rq.vm, (long long)v0,
                       use_eps, c->d_twid, (long long)nv * G);
    EFN_LAUNCH_CHECK();
    EpiArgs ea = rq.ea;
    ea.x = rq.in;
    ea.partial = c->d_partial;
    launch_xout<G>(rq, ea, gx, (long long)v0, ws, c->d_twid, s);
    EFN_LAUNCH_CHECK();
    if (rq.reduce_mode) {
      hipLaunchKernelGGL(reduce_partials_kernel, dim3(nv), dim3(256), 0, s, c->d_partial, nb, nv,
                         rq.reduce_mode, rq.mom ? rq.mom + (long long)v0 * rq.ld : nullptr, rq.ld,
                         rq.k, rq.est ? rq.est + v0 : nullptr, rq.ck64);
      EFN_LAUNCH_CHECK();
    }
  }
  return EFN_OK;
  }
}

template <int G>
int run_op(efn_ctx* c, const OpRequest& rq, hipStream_t s) {
  using namespace efn;
  if constexpr (has_three_pass<G>()) {
    if (theta_passes<G>(c->passes) == 3) return run_op3<G>(c, rq, s);
  }
  constexpr int R = zrows3<G>();
  constexpr int nyt = (G + strided_tile<G, 1>() - 1) / strided_tile<G, 1>();
  constexpr int nxt = (G + xtile_for(G) - 1) / xtile_for(G);
  constexpr int nb = G * G / R;  // zin/zout blocks per vector
  const int use_eps = rq.kind == kTheta ? 0 : 1;
  for (int v0 = 0; v0 < rq.m; v0 += c->chunk) {
    const int nv = rq.m - v0 < c->chunk ? rq.m - v0 : c->chunk;
    float2* const ws = c->d_ws;
    double* const partial = c->d_partial;
    const long long tiles = (long long)nv * G * nxt;  // x-pass tiles per component
    const dim3 gz((unsigned)(nv * nb)), gs((unsigned)(nv * 3 * G * nyt)), gx((unsigned)(3 * tiles)),
        blk(kThreads), blky(strided_threads<G, 1>());
    if (rq.kind == kTheta)
      hipLaunchKernelGGL((zin_kernel<G, kTheta>), gz, blk, 0, s, rq.in, ws, rq.vm, (long long)v0, c->d_twid);
    else
      hipLaunchKernelGGL((zin_kernel<G, kPrecond>), gz, blk, 0, s, rq.in, ws, rq.vm, (long long)v0, c->d_twid);
    EFN_LAUNCH_CHECK();
    hipLaunchKernelGGL((strided_kernel<G, 1, +1>), gs, blky, 0, s, ws, c->d_twid);
    hipLaunchKernelGGL((xmid_kernel<G>), gx, dim3(xmid_threads<G>()), 0, s, ws, rq.vm, (long long)v0, use_eps,
                       c->d_twid, tiles);
    hipLaunchKernelGGL((strided_kernel<G, 1, -1>), gs, blky, 0, s, ws, c->d_twid);
    EFN_LAUNCH_CHECK();
    EpiArgs ea = rq.ea;
    ea.x = rq.in;
    ea.partial = partial;
    if (rq.kind == kPrecond) {
      hipLaunchKernelGGL((zout_kernel<G, kPrecond, kStore>), gz, blk, 0, s, ws, rq.vm, (long long)v0, c->d_twid, ea);
    } else {
      switch (rq.epi) {
        case kStore:
          hipLaunchKernelGGL((zout_kernel<G, kTheta, kStore>), gz, blk, 0, s, ws, rq.vm, (long long)v0, c->d_twid, ea);
          break;
        case kCheb:
          hipLaunchKernelGGL((zout_kernel<G, kTheta, kCheb>), gz, blk, 0, s, ws, rq.vm, (long long)v0, c->d_twid, ea);
          break;
        case kKpm:
          hipLaunchKernelGGL((zout_kernel<G, kTheta, kKpm>), gz, blk, 0, s, ws, rq.vm, (long long)v0, c->d_twid, ea);
          break;
        default:
          hipLaunchKernelGGL((zout_kernel<G, kTheta, kDot>), gz, blk, 0, s, ws, rq.vm, (long long)v0, c->d_twid, ea);
          break;
      }
    }
    EFN_LAUNCH_CHECK();
    if (rq.reduce_mode) {
      hipLaunchKernelGGL(reduce_partials_kernel, dim3(nv), dim3(256), 0, s, partial, nb, nv,
                         rq.reduce_mode, rq.mom ? rq.mom + (long long)v0 * rq.ld : nullptr, rq.ld,
                         rq.k, rq.est ? rq.est + v0 : nullptr, rq.ck64);
      EFN_LAUNCH_CHECK();
    }
  }
  return EFN_OK;
}

struct GridEntry {
  int G;
  int (*run)(efn_ctx*, const OpRequest&, hipStream_t);
  int (*real)(efn_ctx*, const OpRequest&, int, const float*, void*, hipStream_t);
  int zrows;  // z-pencils per zin/zout block
  int (*passes)(int);     // pass structure Θ uses now (3 or 5) for a context's choice
  bool (*has3)();         // the plane pass is compiled for this grid
};

// Real-space fields (modes: efn::RealMode): zin forms the Cartesian field
// (H, or the curl for E), then inverse y and x passes and the epilogue.
template <int G>
int run_real(efn_ctx* c, const OpRequest& rq, int mode, const float* lam, void* out, hipStream_t s) {
  using namespace efn;
  constexpr int R = zrows3<G>();
  constexpr int nzt = (G + ztile_for(G) - 1) / ztile_for(G);
  constexpr int nyt = (G + strided_tile<G, 1>() - 1) / strided_tile<G, 1>();
  constexpr int nb = G * G / R;
  constexpr long long G3 = (long long)G * G * G;
  for (int v0 = 0; v0 < rq.m; v0 += c->chunk) {
    const int nv = rq.m - v0 < c->chunk ? rq.m - v0 : c->chunk;
    const dim3 gz((unsigned)(nv * nb)), gs((unsigned)(nv * 3 * G * nzt)), gsy((unsigned)(nv * 3 * G * nyt)),
        blk(kThreads), blky(strided_threads<G, 1>());
    if (mode == kRealH)
      hipLaunchKernelGGL((zin_kernel<G, kCartH>), gz, blk, 0, s, rq.in, c->d_ws, rq.vm, (long long)v0, c->d_twid);
    else
      hipLaunchKernelGGL((zin_kernel<G, kTheta>), gz, blk, 0, s, rq.in, c->d_ws, rq.vm, (long long)v0, c->d_twid);
    EFN_LAUNCH_CHECK();
    hipLaunchKernelGGL((strided_kernel<G, 1, +1>), gsy, blky, 0, s, c->d_ws, c->d_twid);
    hipLaunchKernelGGL((strided_kernel<G, 0, +1>), gs, blk, 0, s, c->d_ws, c->d_twid);
    EFN_LAUNCH_CHECK();
    const long long count = (long long)nv * G3;
    const dim3 ge((unsigned)((count + kThreads - 1) / kThreads));
    const float inv_n = 1.0f / (float)G3, n_f = (float)G3;
    if (mode == kRealH)
      hipLaunchKernelGGL((real_epilogue_kernel<kRealH>), ge, blk, 0, s, c->d_ws, rq.vm, (long long)v0, count, G3, inv_n, n_f, lam, out);
    else if (mode == kRealE)
      hipLaunchKernelGGL((real_epilogue_kernel<kRealE>), ge, blk, 0, s, c->d_ws, rq.vm, (long long)v0, count, G3, inv_n, n_f, lam, out);
    else
      hipLaunchKernelGGL((real_epilogue_kernel<kRealEd>), ge, blk, 0, s, c->d_ws, rq.vm, (long long)v0, count, G3, inv_n, n_f, lam, out);
    EFN_LAUNCH_CHECK();
  }
  return EFN_OK;
}

#define EFN_GRID(g) {g, &run_op<g>, &run_real<g>, efn::zrows3<g>(), &theta_passes<g>, &three_pass_compiled<g>}
const GridEntry kGrids[] = {EFN_GRID(4),   EFN_GRID(6),   EFN_GRID(8),   EFN_GRID(12),
                            EFN_GRID(16),  EFN_GRID(24),  EFN_GRID(32),  EFN_GRID(48),
                            EFN_GRID(64),  EFN_GRID(96),  EFN_GRID(128), EFN_GRID(192),
                            EFN_GRID(256)};
#undef EFN_GRID

const GridEntry* find_grid(int G) {
  for (const auto& e : kGrids)
    if (e.G == G) return &e;
  return nullptr;
}



template <class T>
void free_dev(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void destroy(efn_ctx* c) {
  free_dev(c->d_axes);
  free_dev(c->d_swap);
  free_dev(c->d_inv_eps);
  free_dev(c->d_eps);
  free_dev(c->d_kn2);
  free_dev(c->d_xeps);
  free_dev(c->d_heps);
  free_dev(c->d_twid);
  free_dev(c->d_ws);
  free_dev(c->d_partial);
  free_dev(c->d_red);
  free_dev(c->d_desc);
  delete c;
}

// Splits of the k range depend on n only, so a problem's Gram is summed in the
// same order whether it runs alone or batched with others (and whatever its
// row counts): results are bitwise batch-invariant.
int gram_nsplit(int, int, long long n) {
  long long ns = (n + 4095) / 4096;  // at least 4096 entries per split
  if (ns > 256) ns = 256;
  return (int)(ns < 1 ? 1 : ns);
}

std::atomic<int> g_gram_engine{EFN_GRAM_MFMA_F64};
std::atomic<int> g_combine_mfma{1};
// largest m_out for which combine runs as one row tile on either engine
// (vector ALU up to 64 rows per tile, matrix cores 16·NA ≤ 64), so the
// outputs may lie in S
constexpr int kCombineInPlaceMax = 64;

int rowreduce_nsplit(long long n) {
  long long ns = (n + 65535) / 65536;
  return (int)(ns < 1 ? 1 : (ns > 256 ? 256 : ns));
}

}  // namespace

extern "C" {

int efn_version(void) { return 400; }  // 0.4.0: ragged launches, selected-row statistics, efn_heev_ql_batched

const char* efn_last_error(void) { return g_last_error.c_str(); }

int efn_grid_supported(int grid_size) { return find_grid(grid_size) != nullptr; }

int efn_ctx_create(int G, double L, double kx, double ky, double kz, const float* eps_host,
                   int max_chunk, efn_ctx** out) {
  if (!out) return fail(EFN_EINVAL, "out is NULL");
  *out = nullptr;
  if (!eps_host) return fail(EFN_EINVAL, "eps is NULL");
  if (!(L > 0.0)) return fail(EFN_EINVAL, "box_size must be positive");
  const GridEntry* ge = find_grid(G);
  if (!ge) return fail(EFN_EUNSUPPORTED, "grid size " + std::to_string(G) + " has no compiled kernels");
  efn_ctx* c = new efn_ctx();
  c->G = G;
  c->L = L;
  c->k[0] = kx; c->k[1] = ky; c->k[2] = kz;
  c->run = ge->run;
  c->real = ge->real;
  const long long G3 = (long long)G * G * G;
  // --- FP64 axis values exactly as operator.py:47-49 (2π·fftfreq(G, L/G) + 2π·k/L)
  std::vector<double> ax[3];
  const double val = 1.0 / (G * (L / G));
  for (int a = 0; a < 3; ++a) {
    ax[a].resize(G);
    const double k0 = 2 * M_PI * c->k[a] / L;
    for (int i = 0; i < G; ++i) {
      const int f = i < (G + 1) / 2 ? i : i - G;  // fftfreq order
      ax[a][i] = 2 * M_PI * (f * val) + k0;
    }
  }
  std::vector<float> axes(3 * G);
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < G; ++i) axes[a * G + i] = (float)ax[a][i];
  // --- swap bits (operator.py:57, FP64) and singular voxel; FP32 |k+G| for kn_floor
  std::vector<uint32_t> swap((G3 + 31) / 32, 0u);
  std::vector<float> kn2f(G3);
  float knmin = INFINITY;
  long long sing = -1;
  for (int x = 0; x < G; ++x)
    for (int y = 0; y < G; ++y)
      for (int z = 0; z < G; ++z) {
        const long long v = ((long long)x * G + y) * G + z;
        const double KX = ax[0][x], KY = ax[1][y], KZ = ax[2][z];
        const double kn = std::sqrt((KX * KX + KY * KY) + KZ * KZ);
        const double knc = kn > 1e-300 ? kn : 1e-300;
        if (std::fabs(KZ) > 0.9 * knc) swap[v >> 5] |= 1u << (v & 31);
        if (KX == 0.0 && KY == 0.0 && KZ == 0.0) sing = v;
        const float knf = (float)kn;
        kn2f[v] = knf * knf;
        if (knf > 0.f && knf < knmin) knmin = knf;
      }
  c->sing_voxel = sing;
  c->kn_floor = 0.5f * knmin;
  // --- ε tables
  std::vector<float> inv_eps(G3), epsv(G3);
  double sum_inv = 0.0;
  const double inv_g3 = 1.0 / (double)G3;
  for (long long v = 0; v < G3; ++v) {
    const double e = (double)eps_host[v];
    if (!(e > 0.0)) {
      destroy(c);
      return fail(EFN_EINVAL, "eps must be positive everywhere");
    }
    const float inv32 = 1.0f / eps_host[v];   // FP32 division as operator.py:146
    sum_inv += 1.0 / e;
    inv_eps[v] = (float)((double)inv32 * inv_g3);
    epsv[v] = (float)(e * inv_g3);
  }
  c->eps_mean_inv = sum_inv / (double)G3;
  // the same tables in the x pass's tile order [y][z / ZT][x][z mod ZT]
  const int xzt = efn::xtile_for(G), xnzt = (G + xzt - 1) / xzt;
  const size_t nx = (size_t)G * xnzt * G * xzt;
  std::vector<float> xeps(2 * nx, 0.f);
  for (int x = 0; x < G; ++x)
    for (int y = 0; y < G; ++y)
      for (int z = 0; z < G; ++z) {
        const size_t v = ((size_t)x * G + y) * G + z;
        const size_t t = (((size_t)y * xnzt + z / xzt) * G + x) * xzt + z % xzt;
        xeps[t] = inv_eps[v];
        xeps[nx + t] = epsv[v];
      }
  // and de-interleaved in z for the half-plane pass [x][h][y][m], z = 2m + h
  std::vector<float> heps;
  if (efn::split_pass_grid(G)) {
    const int H = G / 2;
    heps.assign(2 * (size_t)G3, 0.f);
    for (int x = 0; x < G; ++x)
      for (int h = 0; h < 2; ++h)
        for (int y = 0; y < G; ++y)
          for (int m = 0; m < H; ++m) {
            const size_t v = ((size_t)x * G + y) * G + 2 * m + h;
            const size_t t = (((size_t)x * 2 + h) * G + y) * H + m;
            heps[t] = inv_eps[v];
            heps[(size_t)G3 + t] = epsv[v];
          }
  }
  std::vector<float2> twid(G);
  for (int e = 0; e < G; ++e)
    twid[e] = make_float2((float)std::cos(2 * M_PI * e / G), (float)-std::sin(2 * M_PI * e / G));
  // --- chunk: by default the three-component workspace stays within ~192 MB so
  // the passes of one chunk find their intermediate in the 256 MB Infinity Cache
  // (one vector at 192³ = 170 MB); measured +16% over 1 GB chunks at 192³.
  const size_t per_vec = (size_t)3 * G3 * sizeof(float2);
  int chunk = max_chunk;
  if (chunk < 0) {
    chunk = 0;  // table-only context: usable in multi launches that bring a workspace
  } else if (chunk == 0) {
    chunk = (int)((size_t(192) << 20) / per_vec);
    if (chunk < 1) chunk = 1;
    if (chunk > 1024) chunk = 1024;
  }
  c->chunk = chunk;
  c->ws_bytes = per_vec * chunk;
  // per-block partials: zout blocks (G²/rows) or xout blocks (G·G/16), whichever is more
  const int nb = std::max(G * G / ge->zrows, G * std::max(1, G / efn::kX3Tile));
#define EFN_ALLOC(ptr, bytes)                                                         \
  do {                                                                                \
    if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                           \
      (void)hipGetLastError();                                                        \
      destroy(c);                                                                     \
      return fail(EFN_ENOMEM, "hipMalloc of " + std::to_string((size_t)(bytes)) + " bytes failed"); \
    }                                                                                 \
  } while (0)
  EFN_ALLOC(c->d_axes, axes.size() * sizeof(float));
  EFN_ALLOC(c->d_swap, swap.size() * sizeof(uint32_t));
  EFN_ALLOC(c->d_inv_eps, G3 * sizeof(float));
  EFN_ALLOC(c->d_eps, G3 * sizeof(float));
  EFN_ALLOC(c->d_kn2, G3 * sizeof(float));
  EFN_ALLOC(c->d_xeps, xeps.size() * sizeof(float));
  if (!heps.empty()) EFN_ALLOC(c->d_heps, heps.size() * sizeof(float));
  EFN_ALLOC(c->d_twid, G * sizeof(float2));
  if (chunk > 0) {
    EFN_ALLOC(c->d_ws, c->ws_bytes);
    EFN_ALLOC(c->d_partial, (size_t)chunk * nb * 2 * sizeof(double));
    EFN_ALLOC(c->d_red, (size_t)chunk * sizeof(double2));
  }
  EFN_ALLOC(c->d_desc, sizeof(efn::OpDesc));
#undef EFN_ALLOC
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMemcpy(c->d_axes, axes.data(), axes.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_swap, swap.data(), swap.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_inv_eps, inv_eps.data(), G3 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_eps, epsv.data(), G3 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_kn2, kn2f.data(), G3 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_xeps, xeps.data(), xeps.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess && !heps.empty())
    e = hipMemcpy(c->d_heps, heps.data(), heps.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_twid, twid.data(), G * sizeof(float2), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    efn::OpDesc d;
    d.ft = c->tables();
    d.inv_eps = c->d_inv_eps;
    d.eps = c->d_eps;
    d.inv_eps_x = c->d_xeps;
    d.eps_x = c->d_xeps + nx;
    d.inv_eps_h = c->d_heps;
    d.eps_h = c->d_heps ? c->d_heps + G3 : nullptr;
    e = hipMemcpy(c->d_desc, &d, sizeof(d), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    destroy(c);
    return fail(EFN_EHIP, std::string("upload of operator tables: ") + hipGetErrorString(e));
  }
  *out = c;
  return EFN_OK;
}

int efn_ctx_destroy(efn_ctx* ctx) {
  if (ctx) destroy(ctx);
  return EFN_OK;
}

int efn_ctx_query(const efn_ctx* c, double* kn_floor, double* eps_mean_inv, int* chunk,
                  size_t* workspace_bytes) {
  if (!c) return fail(EFN_EINVAL, "ctx is NULL");
  if (kn_floor) *kn_floor = c->kn_floor;
  if (eps_mean_inv) *eps_mean_inv = c->eps_mean_inv;
  if (chunk) *chunk = c->chunk;
  if (workspace_bytes) *workspace_bytes = c->ws_bytes;
  return EFN_OK;
}

}  // extern "C"

namespace {
__global__ void basis_kernel(efn::FrameTables ft, int G, float* __restrict__ kn, float* __restrict__ t) {
  const long long G3 = (long long)G * G * G;
  for (long long v = (long long)blockIdx.x * 256 + threadIdx.x; v < G3; v += (long long)gridDim.x * 256) {
    const int x = (int)(v / ((long long)G * G));
    const int y = (int)((v / G) % G);
    const int z = (int)(v % G);
    const efn::Frame f = efn::point_frame(ft, efn::row_frame(ft, x, y), z, v);
    if (kn) kn[v] = f.kn;
    if (t) {
      t[0 * G3 + v] = f.t1x; t[1 * G3 + v] = f.t1y; t[2 * G3 + v] = f.t1z;
      t[3 * G3 + v] = f.t2x; t[4 * G3 + v] = f.t2y; t[5 * G3 + v] = f.t2z;
    }
  }
}

__global__ void precond_simple_kernel(const float2* __restrict__ in, const float* __restrict__ kn2,
                                      long long G3, long long count, float scale,
                                      float2* __restrict__ out) {
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long long)gridDim.x * 256) {
    const float d = 1.0f / (kn2[e % G3] + scale);
    const float2 v = in[e];
    out[e] = make_float2(v.x * d, v.y * d);
  }
}

unsigned grid_for(long long count) {
  long long b = (count + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}
}  // namespace

extern "C" {

int efn_ctx_basis(const efn_ctx* c, float* kn_dev, float* t_dev, void* stream) {
  if (!c) return fail(EFN_EINVAL, "ctx is NULL");
  const long long G3 = (long long)c->G * c->G * c->G;
  hipLaunchKernelGGL(basis_kernel, dim3(grid_for(G3)), dim3(256), 0, (hipStream_t)stream,
                     c->tables(), c->G, kn_dev, t_dev);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

static efn::VecMap single_map(const efn_ctx* c, int m) {
  efn::VecMap vm;
  memset(&vm, 0, sizeof(vm));
  vm.desc[0] = c->d_desc;
  vm.in_off[0] = 0;
  vm.out_off[0] = 0;
  vm.nrow = m > 0 ? m : 1;
  vm.nslot = 1;
  return vm;
}

static int check_block(const efn_ctx* c, const void* p, int m, const char* what) {
  if (!c) return fail(EFN_EINVAL, "ctx is NULL");
  if (m < 0) return fail(EFN_EINVAL, "negative block size");
  if (m > 0 && !p) return fail(EFN_EINVAL, std::string(what) + " is NULL");
  return EFN_OK;
}

// operator passes need the context's own workspace
static int check_ws(const efn_ctx* c) {
  if (c->chunk <= 0) return fail(EFN_EINVAL, "context was created without a workspace (max_chunk < 0)");
  return EFN_OK;
}

int efn_theta(efn_ctx* c, const void* in, void* out, int m, float alpha, float beta, void* stream) {
  int rc;
  if ((rc = check_block(c, in, m, "in")) || (rc = check_block(c, out, m, "out")) || (rc = check_ws(c)))
    return rc;
  if (m == 0) return EFN_OK;
  if (in == out) return fail(EFN_EINVAL, "in and out must not alias");
  OpRequest rq{};
  rq.kind = efn::kTheta;
  rq.epi = efn::kStore;
  rq.in = (const float2*)in;
  rq.m = m;
  rq.vm = single_map(c, m);
  rq.ea.alpha = alpha;
  rq.ea.beta = beta;
  rq.ea.out = (float2*)out;
  return c->run(c, rq, (hipStream_t)stream);
}

int efn_realspace(efn_ctx* c, int mode, const void* in, const float* lam_dev, void* out, int m,
                  void* stream) {
  int rc;
  if ((rc = check_block(c, in, m, "in")) || (rc = check_block(c, out, m, "out")) || (rc = check_ws(c)))
    return rc;
  if (mode < efn::kRealH || mode > efn::kRealEd) return fail(EFN_EINVAL, "mode must be 0 (H), 1 (E) or 2 (eps|E|^2)");
  if (m == 0) return EFN_OK;
  if (in == out) return fail(EFN_EINVAL, "in and out must not alias");
  OpRequest rq{};
  rq.kind = mode == efn::kRealH ? efn::kCartH : efn::kTheta;
  rq.in = (const float2*)in;
  rq.m = m;
  rq.vm = single_map(c, m);
  return c->real(c, rq, mode, lam_dev, out, (hipStream_t)stream);
}

int efn_precondition(efn_ctx* c, const void* in, void* out, int m, void* stream) {
  int rc;
  if ((rc = check_block(c, in, m, "in")) || (rc = check_block(c, out, m, "out")) || (rc = check_ws(c)))
    return rc;
  if (m == 0) return EFN_OK;
  if (in == out) return fail(EFN_EINVAL, "in and out must not alias");
  OpRequest rq{};
  rq.kind = efn::kPrecond;
  rq.epi = efn::kStore;
  rq.in = (const float2*)in;
  rq.m = m;
  rq.vm = single_map(c, m);
  rq.ea.alpha = 1.f;
  rq.ea.beta = 0.f;
  rq.ea.out = (float2*)out;
  return c->run(c, rq, (hipStream_t)stream);
}

int efn_precondition_simple(efn_ctx* c, const void* in, void* out, int m, float scale, void* stream) {
  int rc;
  if ((rc = check_block(c, in, m, "in")) || (rc = check_block(c, out, m, "out"))) return rc;
  if (m == 0) return EFN_OK;
  const long long G3 = (long long)c->G * c->G * c->G;
  const long long count = (long long)m * 2 * G3;
  hipLaunchKernelGGL(precond_simple_kernel, dim3(grid_for(count)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)in, c->d_kn2, G3, count, scale, (float2*)out);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

int efn_cheb_step(efn_ctx* c, const void* T1, void* T0, void* Y, float alpha, float ck, int m,
                  void* stream) {
  int rc;
  if ((rc = check_block(c, T1, m, "T1")) || (rc = check_block(c, T0, m, "T0")) ||
      (rc = check_block(c, Y, m, "Y")) || (rc = check_ws(c)))
    return rc;
  if (m == 0) return EFN_OK;
  if (T1 == T0 || T1 == Y || T0 == Y) return fail(EFN_EINVAL, "T1, T0 and Y must be distinct");
  OpRequest rq{};
  rq.kind = efn::kTheta;
  rq.epi = efn::kCheb;
  rq.in = (const float2*)T1;
  rq.m = m;
  rq.vm = single_map(c, m);
  rq.ea.alpha = alpha;
  rq.ea.ck = ck;
  rq.ea.out = (float2*)T0;
  rq.ea.y = (float2*)Y;
  return c->run(c, rq, (hipStream_t)stream);
}

int efn_kpm_step(efn_ctx* c, const void* T1, void* T0, double* mom, int ld, int k, float alpha,
                 int m, void* stream) {
  int rc;
  if ((rc = check_block(c, T1, m, "T1")) || (rc = check_block(c, T0, m, "T0")) || (rc = check_ws(c)))
    return rc;
  if (m == 0) return EFN_OK;
  if (!mom || k < 1 || 2 * k + 1 >= ld) return fail(EFN_EINVAL, "bad moment buffer / index");
  if (T1 == T0) return fail(EFN_EINVAL, "T1 and T0 must be distinct");
  OpRequest rq{};
  rq.kind = efn::kTheta;
  rq.epi = efn::kKpm;
  rq.in = (const float2*)T1;
  rq.m = m;
  rq.vm = single_map(c, m);
  rq.ea.alpha = alpha;
  rq.ea.out = (float2*)T0;
  rq.reduce_mode = 1;
  rq.mom = mom;
  rq.ld = ld;
  rq.k = k;
  return c->run(c, rq, (hipStream_t)stream);
}

int efn_kpm_count_step(efn_ctx* c, const void* T1, void* T0, const void* Z, double* est, double ck,
                       float alpha, int m, void* stream) {
  int rc;
  if ((rc = check_block(c, T1, m, "T1")) || (rc = check_block(c, T0, m, "T0")) ||
      (rc = check_block(c, Z, m, "Z")) || (rc = check_ws(c)))
    return rc;
  if (m == 0) return EFN_OK;
  if (!est) return fail(EFN_EINVAL, "est is NULL");
  if (T1 == T0) return fail(EFN_EINVAL, "T1 and T0 must be distinct");
  OpRequest rq{};
  rq.kind = efn::kTheta;
  rq.epi = efn::kDot;
  rq.in = (const float2*)T1;
  rq.m = m;
  rq.vm = single_map(c, m);
  rq.ea.alpha = alpha;
  rq.ea.out = (float2*)T0;
  rq.ea.z = (const float2*)Z;
  rq.reduce_mode = 2;
  rq.est = est;
  rq.ck64 = ck;
  return c->run(c, rq, (hipStream_t)stream);
}

static int build_multi(efn_ctx* ws, int nslot, efn_ctx* const* ctxs, const void* in_base,
                       const int64_t* in_off, void* out_base, const int64_t* out_off, int nrow,
                       OpRequest& rq) {
  if (!ws) return fail(EFN_EINVAL, "workspace ctx is NULL");
  if (ws->chunk <= 0) return fail(EFN_EINVAL, "workspace ctx was created without a workspace");
  if (nslot < 0 || nslot > efn::kMaxSlots)
    return fail(EFN_EINVAL, "nslot must be in [0, " + std::to_string(efn::kMaxSlots) + "]");
  if (nrow < 0) return fail(EFN_EINVAL, "negative rows per slot");
  if (nslot == 0 || nrow == 0) return 1;  // nothing to do
  if (!ctxs || !in_off || !out_off || !in_base || !out_base) return fail(EFN_EINVAL, "NULL pointer");
  memset(&rq.vm, 0, sizeof(rq.vm));
  for (int i = 0; i < nslot; ++i) {
    if (!ctxs[i]) return fail(EFN_EINVAL, "NULL ctx in slot list");
    if (ctxs[i]->G != ws->G) return fail(EFN_EINVAL, "all slots must share the workspace grid size");
    if (in_off[i] < 0 || out_off[i] < 0) return fail(EFN_EINVAL, "negative slot offset");
    rq.vm.desc[i] = ctxs[i]->d_desc;
    rq.vm.in_off[i] = in_off[i];
    rq.vm.out_off[i] = out_off[i];
  }
  rq.vm.nrow = nrow;
  rq.vm.nslot = nslot;
  rq.in = (const float2*)in_base;
  rq.m = nslot * nrow;
  rq.ea.out = (float2*)out_base;
  return 0;
}

// ragged slots: slot i holds rows[i] vectors (0 allowed); its launch vectors
// follow those of slot i − 1.  Returns 1 when there is nothing to do.
static int build_ragged(efn_ctx* ws, int nslot, efn_ctx* const* ctxs, const void* in_base,
                        const int64_t* in_off, void* out_base, const int64_t* out_off, const int* rows,
                        OpRequest& rq) {
  if (!ws) return fail(EFN_EINVAL, "workspace ctx is NULL");
  if (ws->chunk <= 0) return fail(EFN_EINVAL, "workspace ctx was created without a workspace");
  if (nslot < 0 || nslot > efn::kMaxSlots)
    return fail(EFN_EINVAL, "nslot must be in [0, " + std::to_string(efn::kMaxSlots) + "]");
  if (nslot == 0) return 1;
  if (!ctxs || !in_off || !out_off || !rows || !in_base || !out_base) return fail(EFN_EINVAL, "NULL pointer");
  memset(&rq.vm, 0, sizeof(rq.vm));
  long long total = 0;
  for (int i = 0; i < nslot; ++i) {
    if (!ctxs[i]) return fail(EFN_EINVAL, "NULL ctx in slot list");
    if (ctxs[i]->G != ws->G) return fail(EFN_EINVAL, "all slots must share the workspace grid size");
    if (in_off[i] < 0 || out_off[i] < 0 || rows[i] < 0) return fail(EFN_EINVAL, "negative slot offset or rows");
    rq.vm.desc[i] = ctxs[i]->d_desc;
    rq.vm.in_off[i] = in_off[i];
    rq.vm.out_off[i] = out_off[i];
    rq.vm.row0[i] = (int)total;
    total += rows[i];
    if (total >= (1LL << 30)) return fail(EFN_EINVAL, "too many vectors in one launch");
  }
  rq.vm.row0[nslot] = (int)total;
  rq.vm.nrow = 0;
  rq.vm.nslot = nslot;
  if (total == 0) return 1;
  rq.in = (const float2*)in_base;
  rq.m = (int)total;
  rq.ea.out = (float2*)out_base;
  return 0;
}

static bool ragged_overlap(int nslot, const int64_t* in_off, const int64_t* out_off, const int* rows) {
  for (int i = 0; i < nslot; ++i)
    for (int j = 0; j < nslot; ++j)
      if (rows[i] > 0 && rows[j] > 0 && in_off[i] < out_off[j] + rows[j] && out_off[j] < in_off[i] + rows[i])
        return true;
  return false;
}

int efn_theta_multi(efn_ctx* ws, int nslot, efn_ctx* const* ctxs, const void* in_base,
                    const int64_t* in_off, void* out_base, const int64_t* out_off, int nrow,
                    float alpha, float beta, void* stream) {
  OpRequest rq{};
  rq.kind = efn::kTheta;
  rq.epi = efn::kStore;
  const int rc = build_multi(ws, nslot, ctxs, in_base, in_off, out_base, out_off, nrow, rq);
  if (rc) return rc > 0 ? EFN_OK : rc;
  if (in_base == out_base) {
    for (int i = 0; i < nslot; ++i)
      for (int j = 0; j < nslot; ++j)
        if (in_off[i] < out_off[j] + nrow && out_off[j] < in_off[i] + nrow)
          return fail(EFN_EINVAL, "input and output rows overlap");
  }
  rq.ea.alpha = alpha;
  rq.ea.beta = beta;
  return ws->run(ws, rq, (hipStream_t)stream);
}

int efn_precondition_multi(efn_ctx* ws, int nslot, efn_ctx* const* ctxs, const void* in_base,
                           const int64_t* in_off, void* out_base, const int64_t* out_off, int nrow,
                           void* stream) {
  OpRequest rq{};
  rq.kind = efn::kPrecond;
  rq.epi = efn::kStore;
  const int rc = build_multi(ws, nslot, ctxs, in_base, in_off, out_base, out_off, nrow, rq);
  if (rc) return rc > 0 ? EFN_OK : rc;
  if (in_base == out_base) {
    for (int i = 0; i < nslot; ++i)
      for (int j = 0; j < nslot; ++j)
        if (in_off[i] < out_off[j] + nrow && out_off[j] < in_off[i] + nrow)
          return fail(EFN_EINVAL, "input and output rows overlap");
  }
  rq.ea.alpha = 1.f;
  rq.ea.beta = 0.f;
  return ws->run(ws, rq, (hipStream_t)stream);
}

int efn_theta_ragged(efn_ctx* ws, int nslot, efn_ctx* const* ctxs, const void* in_base,
                     const int64_t* in_off, void* out_base, const int64_t* out_off, const int* rows,
                     float alpha, float beta, void* stream) {
  OpRequest rq{};
  rq.kind = efn::kTheta;
  rq.epi = efn::kStore;
  const int rc = build_ragged(ws, nslot, ctxs, in_base, in_off, out_base, out_off, rows, rq);
  if (rc) return rc > 0 ? EFN_OK : rc;
  if (in_base == out_base && ragged_overlap(nslot, in_off, out_off, rows))
    return fail(EFN_EINVAL, "input and output rows overlap");
  rq.ea.alpha = alpha;
  rq.ea.beta = beta;
  return ws->run(ws, rq, (hipStream_t)stream);
}

int efn_precondition_ragged(efn_ctx* ws, int nslot, efn_ctx* const* ctxs, const void* in_base,
                            const int64_t* in_off, void* out_base, const int64_t* out_off, const int* rows,
                            void* stream) {
  OpRequest rq{};
  rq.kind = efn::kPrecond;
  rq.epi = efn::kStore;
  const int rc = build_ragged(ws, nslot, ctxs, in_base, in_off, out_base, out_off, rows, rq);
  if (rc) return rc > 0 ? EFN_OK : rc;
  if (in_base == out_base && ragged_overlap(nslot, in_off, out_off, rows))
    return fail(EFN_EINVAL, "input and output rows overlap");
  rq.ea.alpha = 1.f;
  rq.ea.beta = 0.f;
  return ws->run(ws, rq, (hipStream_t)stream);
}

int efn_combine_engine(int engine) {
  if (engine < 0) return g_combine_mfma.load();
  if (engine > 1) return fail(EFN_EINVAL, "unknown combine engine");
  return g_combine_mfma.exchange(engine);
}

int efn_theta_engine(int engine) {
  if (engine < 0) return g_theta_engine.load();
  if (engine != EFN_THETA_AUTO && engine != 3 && engine != 5)
    return fail(EFN_EINVAL, "theta engine must be EFN_THETA_AUTO (0), 3 or 5");
  return g_theta_engine.exchange(engine);
}

int efn_theta_passes(int grid_size) {
  const GridEntry* ge = find_grid(grid_size);
  if (!ge) return fail(EFN_EUNSUPPORTED, "grid size " + std::to_string(grid_size) + " has no compiled kernels");
  return ge->passes(0);
}

int efn_theta_three_pass(int grid_size) {
  const GridEntry* ge = find_grid(grid_size);
  if (!ge) return fail(EFN_EUNSUPPORTED, "grid size " + std::to_string(grid_size) + " has no compiled kernels");
  return ge->has3() ? 1 : 0;
}

int efn_ctx_set_passes(efn_ctx* ctx, int passes) {
  if (!ctx) return fail(EFN_EINVAL, "ctx is NULL");
  const GridEntry* ge = find_grid(ctx->G);
  if (passes < 0) return ge->passes(ctx->passes);
  if (passes != 0 && passes != 3 && passes != 5) return fail(EFN_EINVAL, "passes must be 0, 3 or 5");
  if (passes == 3 && !ge->has3())
    return fail(EFN_EUNSUPPORTED, "grid size " + std::to_string(ctx->G) + " has no three-pass kernels");
  ctx->passes = passes;
  return ge->passes(passes);
}

int efn_gram_engine(int engine) {
  if (engine < 0) return g_gram_engine.load();
  if (engine > EFN_GRAM_MFMA_F32) return fail(EFN_EINVAL, "unknown Gram engine");
  return g_gram_engine.exchange(engine);
}

size_t efn_gram_work_bytes(int ma, int mb, int64_t n) {
  if (ma <= 0 || mb <= 0 || n <= 0) return 0;
  return (size_t)gram_nsplit(ma, mb, n) * ma * mb * sizeof(double2);
}

size_t efn_gram_multi_work_bytes(int nprob, int ma, int mb, int64_t n) {
  return nprob <= 0 ? 0 : (size_t)nprob * efn_gram_work_bytes(ma, mb, n);
}

}  // extern "C"

namespace {
// ma_p / mb_p (both or neither): per-problem sizes of a ragged launch, at most
// ma / mb; the output keeps ma × mb per problem, zero outside each problem's own.
int gram_multi_impl(int nprob, const void* const* A, const void* const* B, int ma, int mb, int64_t n,
                    void* out, void* work, size_t work_bytes, bool upper, void* stream,
                    const int* ma_p = nullptr, const int* mb_p = nullptr) {
  if (nprob < 0 || nprob > efn::kMaxProb)
    return fail(EFN_EINVAL, "nprob must be in [0, " + std::to_string(efn::kMaxProb) + "]");
  if (ma < 0 || mb < 0 || n < 0) return fail(EFN_EINVAL, "negative size");
  if (nprob == 0 || ma == 0 || mb == 0) return EFN_OK;
  if (!A || !B || !out) return fail(EFN_EINVAL, "NULL pointer");
  efn::ProbDims pd;
  memset(&pd, 0, sizeof(pd));
  if (ma_p) {
    pd.ragged = 1;
    for (int i = 0; i < nprob; ++i) {
      if (ma_p[i] < 0 || mb_p[i] < 0 || ma_p[i] > ma || mb_p[i] > mb)
        return fail(EFN_EINVAL, "ragged sizes must lie in [0, launch size]");
      if (upper && ma_p[i] != mb_p[i]) return fail(EFN_EINVAL, "a Hermitian product needs square blocks");
      pd.ma[i] = (short)ma_p[i];
      pd.mb[i] = (short)mb_p[i];
    }
  }
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    const long long cnt = (long long)nprob * ma * mb;
    hipLaunchKernelGGL(efn::zero_double2_kernel, dim3((unsigned)((cnt + 255) / 256 < 4096 ? (cnt + 255) / 256 : 4096)),
                       dim3(256), 0, s, (double2*)out, cnt);
    EFN_LAUNCH_CHECK();
    return EFN_OK;
  }
  efn::PairPtrs pp;
  memset(&pp, 0, sizeof(pp));
  for (int i = 0; i < nprob; ++i) {
    if ((!A[i] || !B[i]) && !(pd.ragged && (pd.ma[i] == 0 || pd.mb[i] == 0)))
      return fail(EFN_EINVAL, "NULL operand pointer");
    pp.a[i] = (const float2*)A[i];
    pp.b[i] = (const float2*)B[i];
  }
  const size_t need = efn_gram_multi_work_bytes(nprob, ma, mb, n);
  if (!work || work_bytes < need) return fail(EFN_EINVAL, "gram work buffer too small");
  const int ns = gram_nsplit(ma, mb, n);
  long long kper = (n + ns - 1) / ns;
  kper = (kper + efn::kGK - 1) / efn::kGK * efn::kGK;
  bool aligned = (n % 2) == 0;
  for (int i = 0; i < nprob; ++i)
    aligned = aligned && ((uintptr_t)A[i] % 16 == 0) && ((uintptr_t)B[i] % 16 == 0);
  int engine = aligned ? g_gram_engine.load() : EFN_GRAM_VALU_F64;
  if (pd.ragged) {  // per-problem sizes are implemented by the matrix-core engines
    if (!aligned) return fail(EFN_EUNSUPPORTED, "ragged Gram needs 16-byte aligned blocks of even length");
    if (engine == EFN_GRAM_VALU_F64) engine = EFN_GRAM_MFMA_F64;
  }
  const int edge = engine == EFN_GRAM_VALU_F64 ? efn::kGT : efn::kMWaveTile;
  const int ti = (ma + edge - 1) / edge, tj = (mb + edge - 1) / edge;
  const int ntile = upper ? ti * (ti + 1) / 2 : ti * tj;
  const int tiles_j = upper ? -ti : tj;
  const dim3 grid((unsigned)ns, (unsigned)ntile, (unsigned)nprob);
  if (engine == EFN_GRAM_VALU_F64)
    hipLaunchKernelGGL(efn::gram_partial_kernel, grid, dim3(256), 0, s, pp, ma, mb, tiles_j, (long long)n,
                       kper, (double2*)work);
  else if (engine == EFN_GRAM_MFMA_F64)
    hipLaunchKernelGGL(efn::gram_mfma_kernel<efn::GramF64>, dim3((unsigned)(ns * ntile), 1, (unsigned)nprob),
                       dim3(256), 0, s, pp, pd, ma, mb, tiles_j, ntile, ns, (long long)n, kper, (double2*)work);
  else
    hipLaunchKernelGGL(efn::gram_mfma_kernel<efn::GramF32>, dim3((unsigned)(ns * ntile), 1, (unsigned)nprob),
                       dim3(256), 0, s, pp, pd, ma, mb, tiles_j, ntile, ns, (long long)n, kper, (double2*)work);
  EFN_LAUNCH_CHECK();
  const long long count = (long long)ma * mb;
  hipLaunchKernelGGL(efn::sum_partials_kernel, dim3((unsigned)((count + 255) / 256), (unsigned)nprob),
                     dim3(256), 0, s, (const double2*)work, ns, count, mb, upper ? 1 : 0, pd, (double2*)out);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}
}  // namespace

extern "C" {

int efn_gram_multi(int nprob, const void* const* A, const void* const* B, int ma, int mb, int64_t n,
                   void* out, void* work, size_t work_bytes, void* stream) {
  return gram_multi_impl(nprob, A, B, ma, mb, n, out, work, work_bytes, false, stream);
}

int efn_gram_multi_herm(int nprob, const void* const* A, const void* const* B, int m, int64_t n,
                        void* out, void* work, size_t work_bytes, void* stream) {
  return gram_multi_impl(nprob, A, B, m, m, n, out, work, work_bytes, true, stream);
}

int efn_gram(const void* A, int ma, const void* B, int mb, int64_t n, void* out, void* work,
             size_t work_bytes, void* stream) {
  return gram_multi_impl(1, &A, &B, ma, mb, n, out, work, work_bytes, false, stream);
}

int efn_gram_ragged(int nprob, const void* const* A, const void* const* B, const int* ma, const int* mb,
                    int64_t n, int hermitian, void* out, void* work, size_t work_bytes, void* stream) {
  if (nprob < 0 || nprob > efn::kMaxProb)
    return fail(EFN_EINVAL, "nprob must be in [0, " + std::to_string(efn::kMaxProb) + "]");
  if (nprob == 0) return EFN_OK;
  if (!ma || !mb) return fail(EFN_EINVAL, "NULL size list");
  int mam = 0, mbm = 0;
  for (int i = 0; i < nprob; ++i) {
    mam = std::max(mam, ma[i]);
    mbm = std::max(mbm, mb[i]);
  }
  return gram_multi_impl(nprob, A, B, mam, mbm, n, out, work, work_bytes, hermitian != 0, stream, ma, mb);
}

}  // extern "C"

namespace {
// The split partials of efn_gram_noalloc live in this file-scope array: each
// device that loads the library holds its own copy (EFN_GRAM_NOALLOC_MAX_BYTES)
// until the process ends.  Host code never uses the symbol itself as an address
// (on the host it is only a stand-in): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) unsigned char scratch[EFN_GRAM_NOALLOC_MAX_BYTES];
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_addr;  // indexed by device ordinal

int scratch_address(void** out) {
  int dev = 0;
  EFN_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_addr.size() <= dev) g_scratch_addr.resize(dev + 1, nullptr);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(EFN_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return EFN_OK;
}
}  // namespace

extern "C" {

int efn_gram_noalloc(const void* A, int ma, const void* B, int mb, int64_t n, void* out,
                     void* stream) {
  if (ma < 0 || mb < 0 || n < 0) return fail(EFN_EINVAL, "negative size");
  const size_t need = efn_gram_work_bytes(ma, mb, n);
  if (need > (size_t)EFN_GRAM_NOALLOC_MAX_BYTES)
    return fail(EFN_EINVAL, "efn_gram_noalloc: " + std::to_string(ma) + "x" + std::to_string(mb) +
                                " over n = " + std::to_string((long long)n) + " needs " +
                                std::to_string(need) + " bytes of split partials; the static scratch holds " +
                                std::to_string((size_t)EFN_GRAM_NOALLOC_MAX_BYTES) +
                                " (use efn_gram with a work buffer)");
  void* work = nullptr;
  int rc;
  if ((rc = scratch_address(&work))) return rc;
  // every partial the sum kernel reads is written by the partial kernels of the
  // same call, so the scratch needs no clearing between calls
  return gram_multi_impl(1, &A, &B, ma, mb, n, out, work, (size_t)EFN_GRAM_NOALLOC_MAX_BYTES, false,
                         stream);
}

namespace {
// min_p / mout_p / split_p (all or none): per-problem sizes of a ragged
// launch, at most m_in / m_out; C keeps m_in × m_out per problem.
int combine_impl(int nprob, const void* C, const void* const* S, int m_in, int m_out, int64_t n,
                 const void* const* X, float alpha, void* const* out, void* const* out2, int split,
                 void* stream, const int* min_p = nullptr, const int* mout_p = nullptr,
                 const int* split_p = nullptr) {
  if (nprob < 0 || nprob > efn::kMaxProb)
    return fail(EFN_EINVAL, "nprob must be in [0, " + std::to_string(efn::kMaxProb) + "]");
  if (m_in < 0 || m_out < 0 || n < 0) return fail(EFN_EINVAL, "negative size");
  if (nprob == 0 || m_out == 0 || n == 0) return EFN_OK;
  if (!out || (m_in > 0 && (!C || !S))) return fail(EFN_EINVAL, "NULL pointer");
  if (split < 0 || split > m_out) return fail(EFN_EINVAL, "split must be in [0, m_out]");
  if (split < m_out && !out2 && !min_p) return fail(EFN_EINVAL, "NULL second output");
  efn::CombinePtrs cp;
  memset(&cp, 0, sizeof(cp));
  cp.split = split;
  if (min_p) {
    cp.ragged = 1;
    for (int i = 0; i < nprob; ++i) {
      if (min_p[i] < 0 || mout_p[i] < 0 || min_p[i] > m_in || mout_p[i] > m_out || split_p[i] < 0 ||
          split_p[i] > mout_p[i])
        return fail(EFN_EINVAL, "ragged sizes must lie in [0, launch size] and split in [0, m_out]");
      if (split_p[i] < mout_p[i] && !out2) return fail(EFN_EINVAL, "NULL second output");
      cp.m_in_p[i] = (short)min_p[i];
      cp.m_out_p[i] = (short)mout_p[i];
      cp.split_p[i] = (short)split_p[i];
    }
  }
  // this problem's sizes
  auto pin = [&](int i) { return min_p ? min_p[i] : m_in; };
  auto pout = [&](int i) { return mout_p ? mout_p[i] : m_out; };
  auto psplit = [&](int i) { return split_p ? split_p[i] : split; };
  auto al16 = [](const void* q) { return ((uintptr_t)q & 15) == 0; };
  // [p, p + rows·n) of complex64 against S's [S, S + m_in·n)
  auto overlaps_s = [&](const void* p, long long rows, const void* s) {
    if (!p || !s || rows <= 0) return false;
    const uintptr_t a0 = (uintptr_t)p, a1 = a0 + (uintptr_t)(rows * n) * sizeof(float2);
    const uintptr_t b0 = (uintptr_t)s, b1 = b0 + (uintptr_t)((long long)m_in * n) * sizeof(float2);
    return a0 < b1 && b0 < a1;
  };
  int vec = (n & 1) == 0;
  for (int i = 0; i < nprob; ++i) {
    const int mi = pin(i), mo = pout(i), sp = psplit(i);
    if (mo == 0) continue;  // nothing written for this problem
    if ((sp > 0 && !out[i]) || (sp < mo && !out2[i]) || (mi > 0 && !S[i]))
      return fail(EFN_EINVAL, "NULL block pointer");
    // in place needs one row tile: every reader of a column is its writer, so
    // an output block that overlaps S anywhere (not only at S itself) is
    // limited to one row tile of outputs
    if (mi > 0 && mo > kCombineInPlaceMax &&
        (overlaps_s(sp > 0 ? out[i] : nullptr, sp, S[i]) ||
         overlaps_s(out2 && sp < mo ? out2[i] : nullptr, mo - sp, S[i])))
      return fail(EFN_EINVAL, "outputs may overlap S only for m_out <= " + std::to_string(kCombineInPlaceMax));
    cp.s[i] = mi > 0 ? (const float2*)S[i] : nullptr;
    cp.x[i] = X ? (const float2*)X[i] : nullptr;
    cp.out[i] = (float2*)out[i];
    cp.out2[i] = out2 ? (float2*)out2[i] : nullptr;
    if ((cp.out[i] && !al16(cp.out[i])) || (cp.out2[i] && !al16(cp.out2[i])) ||
        (cp.s[i] && !al16(cp.s[i])) || (cp.x[i] && !al16(cp.x[i])))
      vec = 0;
  }
  hipStream_t s = (hipStream_t)stream;
  const unsigned gx = (unsigned)((n + 255) / 256);
  const float2* Cc = (const float2*)C;
  // matrix cores: 64 columns per wave, 16·NA output rows per block row (one
  // row tile up to 64 outputs, so S is read once); ragged launches always
  if ((g_combine_mfma.load() && m_out > 8) || cp.ragged) {
    if (m_out <= 16)
      hipLaunchKernelGGL((efn::combine_mfma_kernel<1>), dim3(gx, 1, nprob), dim3(256), 0, s, Cc, m_in,
                         m_out, cp, (long long)n, alpha, vec);
    else if (m_out <= 32)
      hipLaunchKernelGGL((efn::combine_mfma_kernel<2>), dim3(gx, 1, nprob), dim3(256), 0, s, Cc, m_in,
                         m_out, cp, (long long)n, alpha, vec);
    else if (m_out <= 48)
      hipLaunchKernelGGL((efn::combine_mfma_kernel<3>), dim3(gx, 1, nprob), dim3(256), 0, s, Cc, m_in,
                         m_out, cp, (long long)n, alpha, vec);
    else
      hipLaunchKernelGGL((efn::combine_mfma_kernel<4>), dim3(gx, (m_out + 63) / 64, nprob), dim3(256),
                         0, s, Cc, m_in, m_out, cp, (long long)n, alpha, vec);
    EFN_LAUNCH_CHECK();
    return EFN_OK;
  }
  // vector ALU: one row tile up to 64 outputs, then tiles of 64
  if (m_out <= 8)
    hipLaunchKernelGGL((efn::combine_kernel<8>), dim3(gx, 1, nprob), dim3(256), 0, s, Cc, m_in, m_out, cp,
                       (long long)n, alpha);
  else if (m_out <= 16)
    hipLaunchKernelGGL((efn::combine_kernel<16>), dim3(gx, 1, nprob), dim3(256), 0, s, Cc, m_in, m_out, cp,
                       (long long)n, alpha);
  else if (m_out <= 32)
    hipLaunchKernelGGL((efn::combine_kernel<32>), dim3(gx, 1, nprob), dim3(256), 0, s, Cc, m_in, m_out, cp,
                       (long long)n, alpha);
  else
    hipLaunchKernelGGL((efn::combine_kernel<64>), dim3(gx, (m_out + 63) / 64, nprob), dim3(256), 0, s, Cc,
                       m_in, m_out, cp, (long long)n, alpha);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}
}  // namespace

int efn_combine_multi(int nprob, const void* C, const void* const* S, int m_in, int m_out, int64_t n,
                      const void* const* X, float alpha, void* const* out, void* stream) {
  return combine_impl(nprob, C, S, m_in, m_out, n, X, alpha, out, nullptr, m_out, stream);
}

int efn_combine_multi_split(int nprob, const void* C, const void* const* S, int m_in, int m_out,
                            int64_t n, void* const* out, void* const* out2, int split, void* stream) {
  return combine_impl(nprob, C, S, m_in, m_out, n, nullptr, 0.f, out, out2, split, stream);
}

int efn_combine_ragged(int nprob, const void* C, const void* const* S, const int* m_in, const int* m_out,
                       int64_t n, const void* const* X, float alpha, void* const* out, void* const* out2,
                       const int* split, void* stream) {
  if (nprob < 0 || nprob > efn::kMaxProb)
    return fail(EFN_EINVAL, "nprob must be in [0, " + std::to_string(efn::kMaxProb) + "]");
  if (nprob == 0) return EFN_OK;
  if (!m_in || !m_out || !split) return fail(EFN_EINVAL, "NULL size list");
  int mim = 0, mom = 0;
  for (int i = 0; i < nprob; ++i) {
    mim = std::max(mim, m_in[i]);
    mom = std::max(mom, m_out[i]);
  }
  return combine_impl(nprob, C, S, mim, mom, n, X, alpha, out, out2, mom, stream, m_in, m_out, split);
}

int efn_heev_batched(int nprob, int n, const void* A, double* w, void* V, int max_sweeps, void* stream) {
  if (nprob < 0 || n < 0) return fail(EFN_EINVAL, "negative size");
  if (n > efn::kEigMaxN) return fail(EFN_EINVAL, "n must be <= " + std::to_string(efn::kEigMaxN));
  if (nprob == 0 || n == 0) return EFN_OK;
  if (!A || !w || !V) return fail(EFN_EINVAL, "NULL pointer");
  if (A == V) return fail(EFN_EINVAL, "A and V must not alias");
  if (max_sweeps <= 0) max_sweeps = 30;
  const int P = (n + (n & 1)) / 2;
  const size_t lds = (size_t)n * (n + 1) * sizeof(double2) + 3 * (size_t)P * sizeof(double2) +
                     (size_t)n * sizeof(double);
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute((const void*)efn::heev_jacobi_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail(EFN_EHIP, std::string("heev LDS size: ") + hipGetErrorString(e));
  }
  hipLaunchKernelGGL(efn::heev_jacobi_kernel, dim3((unsigned)nprob), dim3(efn::kEigThreads), lds,
                     (hipStream_t)stream, (const double2*)A, n, max_sweeps, w, (double2*)V, (const int*)nullptr);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

namespace {
// Shared checks and launch of the one-workgroup-per-matrix eigensolvers
// (efn_heev_ql_batched, efn_heevx_batched): same arguments, same layouts.
using HeevKernel = void (*)(const double2*, int, efn::ql::Dims, int, double*, double2*, int*);
int heev_small_launch(HeevKernel kernel, const char* name, size_t lds, int nprob, int n, const int* n_each,
                      const void* A, int lowest, double* w, void* V, int* info, void* stream) {
  namespace ql = efn::ql;
  if (nprob < 0 || n < 0) return fail(EFN_EINVAL, "negative size");
  if (n > ql::kMaxN) return fail(EFN_EINVAL, "n must be <= " + std::to_string(ql::kMaxN));
  if (nprob == 0 || n == 0) return EFN_OK;
  if (lowest < 1 || lowest > n) return fail(EFN_EINVAL, "lowest must be in [1, n]");
  if (!A || !w || !V || !info) return fail(EFN_EINVAL, "NULL pointer");
  const uintptr_t a0 = (uintptr_t)A, a1 = a0 + (uintptr_t)nprob * n * n * sizeof(double2);
  const uintptr_t v0 = (uintptr_t)V, v1 = v0 + (uintptr_t)nprob * n * lowest * sizeof(double2);
  if (a0 < v1 && v0 < a1) return fail(EFN_EINVAL, "A and V must not overlap");
  ql::Dims dims;
  memset(&dims, 0, sizeof(dims));
  if (n_each) {
    if (nprob > ql::kMaxRagged)
      return fail(EFN_EINVAL, "per-problem sizes for at most " + std::to_string(ql::kMaxRagged) + " problems");
    dims.ragged = 1;
    for (int p = 0; p < nprob; ++p) {
      if (n_each[p] < 0 || n_each[p] > n) return fail(EFN_EINVAL, "per-problem sizes must lie in [0, n]");
      dims.n[p] = (short)n_each[p];
    }
  }
  if (lds > 65536) {
    const hipError_t e =
        hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail(EFN_EHIP, std::string(name) + " LDS size: " + hipGetErrorString(e));
  }
  hipLaunchKernelGGL(kernel, dim3((unsigned)nprob), dim3(ql::kThreads), lds, (hipStream_t)stream,
                     (const double2*)A, n, dims, lowest, w, (double2*)V, info);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}
}  // namespace

int efn_heev_ql_batched(int nprob, int n, const int* n_each, const void* A, int lowest, double* w, void* V,
                        int* info, void* stream) {
  const size_t lds = n > 0 && n <= efn::ql::kMaxN ? (size_t)efn::ql::Layout(n).bytes() : 0;
  return heev_small_launch(efn::ql::heev_ql_kernel, "heev_ql", lds, nprob, n, n_each, A, lowest, w, V, info,
                           stream);
}

int efn_heevx_batched(int nprob, int n, const int* n_each, const void* A, int lowest, double* w, void* V,
                      int* info, void* stream) {
  const size_t lds = n > 0 && n <= efn::ql::kMaxN ? (size_t)efn::ql::heevx_lds_bytes(n) : 0;
  return heev_small_launch(efn::ql::heevx_kernel, "heevx", lds, nprob, n, n_each, A, lowest, w, V, info, stream);
}

int efn_cholqr_batched(int nprob, int n, const void* G, void* C, double drop, int* rank, void* stream) {
  if (nprob < 0 || n < 0) return fail(EFN_EINVAL, "negative size");
  if (n > efn::kCholMaxN) return fail(EFN_EINVAL, "n must be <= " + std::to_string(efn::kCholMaxN));
  if (!(drop >= 0.0 && drop < 1.0)) return fail(EFN_EINVAL, "drop must be in [0, 1)");
  if (nprob == 0 || n == 0) return EFN_OK;
  if (!G || !C || !rank) return fail(EFN_EINVAL, "NULL pointer");
  if (G == C) return fail(EFN_EINVAL, "G and C must not alias");
  const size_t lds = (size_t)n * (n + 1) * sizeof(double2) + (size_t)(n + 2) * sizeof(int);
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute((const void*)efn::cholqr_coeffs_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail(EFN_EHIP, std::string("cholqr LDS size: ") + hipGetErrorString(e));
  }
  hipLaunchKernelGGL(efn::cholqr_coeffs_kernel, dim3((unsigned)nprob), dim3(efn::kCholThreads), lds,
                     (hipStream_t)stream, (const double2*)G, n, drop, (double2*)C, rank);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

}  // extern "C"

namespace {
// In place: each n×n block M ← Mᴴ (rocSOLVER's column-major eigenvectors of
// conj(A) read row-major are Vᵀ with V the eigenvectors of A; conjugating the
// transpose gives V with column j the eigenvector of w[j]).
__global__ void conj_transpose_kernel(double2* __restrict__ M, int n) {
  double2* a = M + (long long)blockIdx.x * n * n;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - i * n;
    if (i < j) {
      const double2 x = a[i * n + j], y = a[j * n + i];
      a[i * n + j] = make_double2(y.x, -y.y);
      a[j * n + i] = make_double2(x.x, -x.y);
    } else if (i == j) {
      a[i * n + i].y = -a[i * n + i].y;
    }
  }
}

// Per matrix: 1 if every entry is exactly zero (a block without P rows gives
// an all-zero Gram matrix, which the divide and conquer does not return as
// w = 0, V = I).
__global__ void zero_matrix_flags_kernel(const double2* __restrict__ A, int n, int* __restrict__ flag) {
  __shared__ int nz;
  if (threadIdx.x == 0) nz = 0;
  __syncthreads();
  const double2* a = A + (long long)blockIdx.x * n * n;
  int any = 0;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) any |= (a[e].x != 0.0 || a[e].y != 0.0);
  if (any) atomicOr(&nz, 1);
  __syncthreads();
  if (threadIdx.x == 0) flag[blockIdx.x] = nz ? 0 : 1;
}

// Per matrix: 1 if any of its eigenvalues or eigenvector entries is not finite.
__global__ void nonfinite_flags_kernel(const double* __restrict__ w, const double2* __restrict__ V, int n,
                                       int* __restrict__ flag) {
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const double2* v = V + (long long)blockIdx.x * n * n;
  int any = 0;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) any |= !(isfinite(v[e].x) && isfinite(v[e].y));
  for (int i = threadIdx.x; i < n; i += blockDim.x) any |= !isfinite(w[(long long)blockIdx.x * n + i]);
  if (any) atomicOr(&bad, 1);
  __syncthreads();
  if (threadIdx.x == 0) flag[blockIdx.x] = bad;
}

__global__ void zero_matrix_fix_kernel(const int* __restrict__ flag, int n, double* __restrict__ w,
                                       double2* __restrict__ V) {
  if (!flag[blockIdx.x]) return;
  double2* v = V + (long long)blockIdx.x * n * n;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - i * n;
    v[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) w[(long long)blockIdx.x * n + i] = 0.0;
}

// Per matrix C = op(X)·Y, n × n complex128, row-major.  mode 0: X is a
// Hermitian matrix of which only the upper triangle (and the real diagonal)
// is read; 1: X as stored; 2: Xᴴ.  strides in elements (0: one matrix for
// the whole batch).
__global__ void small_zgemm_kernel(const double2* __restrict__ X, long long sx, int mode,
                                   const double2* __restrict__ Y, long long sy,
                                   double2* __restrict__ C, long long sc, int n) {
  const double2* x = X + blockIdx.x * sx;
  const double2* y = Y + blockIdx.x * sy;
  double2* c = C + blockIdx.x * sc;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - i * n;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < n; ++k) {
      double2 a;
      if (mode == 0) {
        a = i < k ? x[i * n + k] : (i == k ? make_double2(x[i * n + i].x, 0.0)
                                           : make_double2(x[k * n + i].x, -x[k * n + i].y));
      } else if (mode == 1) {
        a = x[i * n + k];
      } else {
        a = make_double2(x[k * n + i].x, -x[k * n + i].y);
      }
      const double2 b = y[k * n + j];
      re = fma(a.x, b.x, fma(-a.y, b.y, re));
      im = fma(a.x, b.y, fma(a.y, b.x, im));
    }
    c[e] = make_double2(re, im);
  }
}

// A fixed unitary n × n matrix per (device, n): the orthonormalised columns
// of a seeded complex Gaussian matrix (modified Gram–Schmidt in FP64, twice),
// uploaded on first use and kept for the life of the process.
int fixed_unitary(int n, const double2** out) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, double2*> cache;
  int dev = 0;
  EFN_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair(dev, n);
  auto it = cache.find(key);
  if (it == cache.end()) {
    std::vector<double2> q((size_t)n * n);
    uint64_t st = 0x5EEDull + (uint64_t)n;
    auto next = [&st]() {  // splitmix64 → uniform (0, 1)
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (auto& v : q) {
      const double r = std::sqrt(-2.0 * std::log(next())), t = 2.0 * M_PI * next();
      v = make_double2(r * std::cos(t), r * std::sin(t));
    }
    for (int pass = 0; pass < 2; ++pass)
      for (int j = 0; j < n; ++j) {  // column j against columns < j, then normalised
        for (int k = 0; k < j; ++k) {
          double pr = 0.0, pi = 0.0;  // ⟨q_k, q_j⟩
          for (int i = 0; i < n; ++i) {
            const double2 a = q[(size_t)i * n + k], b = q[(size_t)i * n + j];
            pr += a.x * b.x + a.y * b.y;
            pi += a.x * b.y - a.y * b.x;
          }
          for (int i = 0; i < n; ++i) {
            const double2 a = q[(size_t)i * n + k];
            q[(size_t)i * n + j].x -= pr * a.x - pi * a.y;
            q[(size_t)i * n + j].y -= pr * a.y + pi * a.x;
          }
        }
        double nr = 0.0;
        for (int i = 0; i < n; ++i) nr += q[(size_t)i * n + j].x * q[(size_t)i * n + j].x + q[(size_t)i * n + j].y * q[(size_t)i * n + j].y;
        nr = 1.0 / std::sqrt(nr);
        for (int i = 0; i < n; ++i) { q[(size_t)i * n + j].x *= nr; q[(size_t)i * n + j].y *= nr; }
      }
    double2* d = nullptr;
    EFN_HIP_CHECK(hipMalloc((void**)&d, q.size() * sizeof(double2)));
    EFN_HIP_CHECK(hipMemcpy(d, q.data(), q.size() * sizeof(double2), hipMemcpyHostToDevice));
    it = cache.emplace(key, d).first;
  }
  *out = it->second;
  return EFN_OK;
}

// One rocBLAS handle per (host thread, device): creating one costs
// milliseconds; the stream is set on every call.
int rocblas_handle_for(rocblas_handle* out) {
  thread_local std::map<int, rocblas_handle> handles;
  int dev = 0;
  EFN_HIP_CHECK(hipGetDevice(&dev));
  auto it = handles.find(dev);
  if (it == handles.end()) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success || !h)
      return fail(EFN_EHIP, "rocblas_create_handle failed");
    it = handles.emplace(dev, h).first;
  }
  *out = it->second;
  return EFN_OK;
}
// rocSOLVER's own scratch for the batched eigensolve, owned by this library:
// sized once per (n, nprob) with rocBLAS's size query and handed to the handle
// as a user workspace, so a call never allocates — which is what lets a
// stream capture (hipGraph) record it.  A size not seen before is queried and
// allocated outside any capture (the first, eager call).
struct RocWorkspace {
  void* ptr = nullptr;
  size_t bytes = 0;
  std::map<std::pair<int, int>, size_t> need;  // (n, nprob) → bytes
};

int rocsolver_workspace_for(rocblas_handle h, int n, int nprob, hipStream_t s, double2* V, double* w,
                            double* E, rocblas_int* info) {
  thread_local std::map<int, RocWorkspace> spaces;
  int dev = 0;
  EFN_HIP_CHECK(hipGetDevice(&dev));
  RocWorkspace& ws = spaces[dev];
  auto it = ws.need.find({n, nprob});
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  EFN_HIP_CHECK(hipStreamIsCapturing(s, &cap));
  if (it == ws.need.end()) {
    if (cap != hipStreamCaptureStatusNone)
      return fail(EFN_EINVAL, "heevd_batched: size " + std::to_string(n) + " x " + std::to_string(nprob) +
                                  " first seen under stream capture; run it once uncaptured");
    size_t bytes = 0;
    if (rocblas_start_device_memory_size_query(h) != rocblas_status_success)
      return fail(EFN_EHIP, "rocblas_start_device_memory_size_query failed");
    (void)rocsolver_zheevd_strided_batched(h, rocblas_evect_original, rocblas_fill_lower, n,
                                           (rocblas_double_complex*)V, n, (rocblas_stride)n * n, w,
                                           (rocblas_stride)n, E, (rocblas_stride)n, info, nprob);
    if (rocblas_stop_device_memory_size_query(h, &bytes) != rocblas_status_success)
      return fail(EFN_EHIP, "rocblas_stop_device_memory_size_query failed");
    it = ws.need.emplace(std::make_pair(n, nprob), bytes).first;
  }
  if (it->second > ws.bytes) {
    if (cap != hipStreamCaptureStatusNone)
      return fail(EFN_EINVAL, "heevd_batched: workspace must grow under stream capture; run it once uncaptured");
    EFN_HIP_CHECK(hipStreamSynchronize(s));
    if (ws.ptr) (void)hipFree(ws.ptr);
    ws.ptr = nullptr;
    ws.bytes = 0;
    const size_t bytes = ((it->second + 4095) / 4096) * 4096;
    if (hipMalloc(&ws.ptr, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(EFN_ENOMEM, "rocSOLVER workspace of " + std::to_string(bytes) + " bytes");
    }
    ws.bytes = bytes;
  }
  if (ws.bytes && rocblas_set_workspace(h, ws.ptr, ws.bytes) != rocblas_status_success)
    return fail(EFN_EHIP, "rocblas_set_workspace failed");
  return EFN_OK;
}
}  // namespace

extern "C" {

size_t efn_heevd_work_bytes(int nprob, int n) {
  if (nprob <= 0 || n <= 0) return 0;
  const size_t e = (size_t)nprob * n * sizeof(double);
  const size_t t = (size_t)nprob * n * n * sizeof(double2);
  return ((e + 255) / 256) * 256 + ((3 * (size_t)nprob * sizeof(int) + 255) / 256) * 256 + t;
}

int efn_heevd_batched(int nprob, int n, const void* A, double* w, void* V, void* work, size_t work_bytes,
                      void* stream) {
  if (nprob < 0 || n < 0) return fail(EFN_EINVAL, "negative size");
  if (nprob == 0 || n == 0) return EFN_OK;
  if (!A || !w || !V) return fail(EFN_EINVAL, "NULL pointer");
  if (A == V) return fail(EFN_EINVAL, "A and V must not alias");
  if (!work || work_bytes < efn_heevd_work_bytes(nprob, n)) return fail(EFN_EINVAL, "heevd work buffer too small");
  hipStream_t s = (hipStream_t)stream;
  rocblas_handle h;
  int rc;
  if ((rc = rocblas_handle_for(&h))) return rc;
  if (rocblas_set_stream(h, s) != rocblas_status_success) return fail(EFN_EHIP, "rocblas_set_stream failed");
  const size_t e_bytes = (((size_t)nprob * n * sizeof(double) + 255) / 256) * 256;
  const size_t i_bytes = ((3 * (size_t)nprob * sizeof(int) + 255) / 256) * 256;
  double* E = (double*)work;
  rocblas_int* info = (rocblas_int*)((char*)work + e_bytes);
  if ((rc = rocsolver_workspace_for(h, n, nprob, s, (double2*)V, w, E, info))) return rc;
  int* zero = (int*)(info + nprob);
  int* redo = zero + nprob;
  double2* T = (double2*)((char*)work + e_bytes + i_bytes);
  hipLaunchKernelGGL(zero_matrix_flags_kernel, dim3((unsigned)nprob), dim3(256), 0, s, (const double2*)A, n, zero);
  EFN_LAUNCH_CHECK();
  // B = Qᴴ·A·Q with a fixed unitary Q: the same eigenvalues, and no exactly
  // zero column (the dead rows of a Rayleigh–Ritz matrix are decoupled from
  // the rest, which made the divide and conquer return non-finite vectors)
  const double2* Q;
  if ((rc = fixed_unitary(n, &Q))) return rc;
  const long long nn = (long long)n * n;
  hipLaunchKernelGGL(small_zgemm_kernel, dim3((unsigned)nprob), dim3(256), 0, s, (const double2*)A, nn, 0, Q,
                     0LL, T, nn, n);
  hipLaunchKernelGGL(small_zgemm_kernel, dim3((unsigned)nprob), dim3(256), 0, s, Q, 0LL, 2, (const double2*)T,
                     nn, (double2*)V, nn, n);
  EFN_LAUNCH_CHECK();
  // the row-major B read column-major is Bᵀ = conj(B); its lower triangle is B's upper one
  const rocblas_status st = rocsolver_zheevd_strided_batched(
      h, rocblas_evect_original, rocblas_fill_lower, n, (rocblas_double_complex*)V, n,
      (rocblas_stride)n * n, w, (rocblas_stride)n, E, (rocblas_stride)n, info, nprob);
  if (st != rocblas_status_success)
    return fail(EFN_EHIP, "rocsolver_zheevd_strided_batched: status " + std::to_string((int)st));
  hipLaunchKernelGGL(conj_transpose_kernel, dim3((unsigned)nprob), dim3(256), 0, s, (double2*)V, n);
  // eigenvectors of A: Q·U
  hipLaunchKernelGGL(small_zgemm_kernel, dim3((unsigned)nprob), dim3(256), 0, s, Q, 0LL, 1, (const double2*)V,
                     nn, T, nn, n);
  EFN_HIP_CHECK(hipMemcpyAsync(V, T, (size_t)nprob * nn * sizeof(double2), hipMemcpyDeviceToDevice, s));
  EFN_LAUNCH_CHECK();
  hipLaunchKernelGGL(zero_matrix_fix_kernel, dim3((unsigned)nprob), dim3(256), 0, s, zero, n, w, (double2*)V);
  EFN_LAUNCH_CHECK();
  // safety net: a matrix whose decomposition came back non-finite (the divide
  // and conquer failed on exactly degenerate clusters) is solved again by the
  // Jacobi kernel, which skips every other matrix
  if (n <= efn::kEigMaxN) {
    hipLaunchKernelGGL(nonfinite_flags_kernel, dim3((unsigned)nprob), dim3(256), 0, s, w, (const double2*)V, n, redo);
    const int P = (n + (n & 1)) / 2;
    const size_t lds = (size_t)n * (n + 1) * sizeof(double2) + 3 * (size_t)P * sizeof(double2) +
                       (size_t)n * sizeof(double);
    if (lds > 65536) {
      const hipError_t e = hipFuncSetAttribute((const void*)efn::heev_jacobi_kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return fail(EFN_EHIP, std::string("heev LDS size: ") + hipGetErrorString(e));
    }
    hipLaunchKernelGGL(efn::heev_jacobi_kernel, dim3((unsigned)nprob), dim3(efn::kEigThreads), lds, s,
                       (const double2*)A, n, 30, w, (double2*)V, (const int*)redo);
    EFN_LAUNCH_CHECK();
  }
  return EFN_OK;
}

int efn_combine(const void* C, int m_in, int m_out, const void* S, int64_t n, const void* X,
                float alpha, void* out, void* stream) {
  void* o = out;
  return efn_combine_multi(1, C, &S, m_in, m_out, n, X ? &X : nullptr, alpha, &o, stream);
}

size_t efn_rowreduce_work_bytes(int m, int64_t n) {
  if (m <= 0 || n <= 0) return 0;
  return (size_t)rowreduce_nsplit(n) * m * sizeof(double2);
}

int efn_rowreduce_strided(int op, const void* A, int64_t stride_a, const void* B, int64_t stride_b,
                          int nrow, int m, int64_t n, void* out, void* work, size_t work_bytes,
                          void* stream) {
  if (op != 0 && op != 1) return fail(EFN_EINVAL, "op must be 0 or 1");
  if (m < 0 || n < 0) return fail(EFN_EINVAL, "negative size");
  if (m == 0) return EFN_OK;
  if (nrow <= 0 || stride_a < nrow || (op == 1 && stride_b < nrow))
    return fail(EFN_EINVAL, "row map needs 0 < nrow <= stride");
  if (!A || !out || (op == 1 && !B)) return fail(EFN_EINVAL, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    hipLaunchKernelGGL(efn::zero_double2_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s,
                       (double2*)out, (long long)m);
    EFN_LAUNCH_CHECK();
    return EFN_OK;
  }
  const size_t need = efn_rowreduce_work_bytes(m, n);
  if (!work || work_bytes < need) return fail(EFN_EINVAL, "rowreduce work buffer too small");
  const int ns = rowreduce_nsplit(n);
  const long long kper = (n + ns - 1) / ns;
  if (op == 0)
    hipLaunchKernelGGL((efn::rowreduce_partial_kernel<0>), dim3(ns, m), dim3(256), 0, s,
                       (const float2*)A, (long long)stride_a, (const float2*)A, (long long)stride_a,
                       nrow, (long long)n, kper, (double2*)work);
  else
    hipLaunchKernelGGL((efn::rowreduce_partial_kernel<1>), dim3(ns, m), dim3(256), 0, s,
                       (const float2*)A, (long long)stride_a, (const float2*)B, (long long)stride_b,
                       nrow, (long long)n, kper, (double2*)work);
  EFN_LAUNCH_CHECK();
  hipLaunchKernelGGL(efn::rowreduce_final_kernel, dim3((m + 255) / 256), dim3(256), 0, s,
                     (const double2*)work, ns, m, (double2*)out);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

size_t efn_rowstats_work_bytes(int m, int64_t n) {
  return 2 * efn_rowreduce_work_bytes(m, n);
}

static int rowstats_impl(const void* A, int64_t stride_a, const void* B, int64_t stride_b, const int* sel,
                         int nrow, int nslot, int m, int64_t n, double* out4, void* work, size_t work_bytes,
                         void* stream) {
  if (m < 0 || n <= 0) return fail(EFN_EINVAL, "need m >= 0 and n > 0");
  if (m == 0) return EFN_OK;
  if (nrow <= 0 || stride_a < nrow || stride_b < nrow) return fail(EFN_EINVAL, "row map needs 0 < nrow <= stride");
  if (!A || !B || !out4) return fail(EFN_EINVAL, "NULL pointer");
  if (!work || work_bytes < efn_rowstats_work_bytes(m, n)) return fail(EFN_EINVAL, "rowstats work buffer too small");
  hipStream_t s = (hipStream_t)stream;
  const int ns = rowreduce_nsplit(n);
  const long long kper = (n + ns - 1) / ns;
  hipLaunchKernelGGL(efn::rowstats_partial_kernel, dim3(ns, m), dim3(256), 0, s, (const float2*)A,
                     (long long)stride_a, (const float2*)B, (long long)stride_b, efn::RowSel{sel, nrow, nslot},
                     (long long)n, kper, (double2*)work);
  EFN_LAUNCH_CHECK();
  hipLaunchKernelGGL(efn::rowstats_final_kernel, dim3((m + 255) / 256), dim3(256), 0, s,
                     (const double2*)work, ns, m, out4);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

int efn_rowstats_strided(const void* A, int64_t stride_a, const void* B, int64_t stride_b, int nrow,
                         int m, int64_t n, double* out4, void* work, size_t work_bytes, void* stream) {
  return rowstats_impl(A, stride_a, B, stride_b, nullptr, nrow, 0, m, n, out4, work, work_bytes, stream);
}

// selected rows: sel (device, m entries) = slot << 16 | r with r < nrow < 65536 and slot < nslot
static const char* sel_check(const int* sel, int nrow, int nslot) {
  if (!sel) return "NULL row selection";
  if (nrow <= 0 || nrow > 0xFFFF) return "selected rows need 0 < nrow < 65536";
  if (nslot <= 0 || nslot > 0x7FFF) return "selected rows need 0 < nslot < 32768";
  return nullptr;
}

int efn_rowstats_sel(const void* A, int64_t stride_a, const void* B, int64_t stride_b, const int* sel,
                     int nrow, int nslot, int m, int64_t n, double* out4, void* work, size_t work_bytes,
                     void* stream) {
  if (const char* e = sel_check(sel, nrow, nslot)) return fail(EFN_EINVAL, e);
  return rowstats_impl(A, stride_a, B, stride_b, sel, nrow, nslot, m, n, out4, work, work_bytes, stream);
}

static int row_axpby_norm_impl(const void* X, int64_t stride_x, const double* a, const void* Y,
                               int64_t stride_y, const double* b, const int* sel, int nrow, int nslot, int m,
                               int64_t n, void* out, int64_t stride_out, double* norm2_c128, void* work,
                               size_t work_bytes, void* stream) {
  if (m < 0 || n <= 0) return fail(EFN_EINVAL, "need m >= 0 and n > 0");
  if (m == 0) return EFN_OK;
  if (nrow <= 0 || stride_x < nrow || stride_out < nrow || (Y && stride_y < nrow))
    return fail(EFN_EINVAL, "row map needs 0 < nrow <= stride");
  if (!X || !a || !out || !norm2_c128 || (Y && !b)) return fail(EFN_EINVAL, "NULL pointer");
  if (out == X || (Y && out == Y)) return fail(EFN_EINVAL, "out must not alias X or Y");
  if (!work || work_bytes < efn_rowreduce_work_bytes(m, n)) return fail(EFN_EINVAL, "work buffer too small");
  hipStream_t s = (hipStream_t)stream;
  const int ns = rowreduce_nsplit(n);
  const long long kper = (n + ns - 1) / ns;
  hipLaunchKernelGGL(efn::row_axpby_norm_kernel, dim3(ns, m), dim3(256), 0, s, (const float2*)X,
                     (long long)stride_x, a, (const float2*)Y, (long long)stride_y, b,
                     efn::RowSel{sel, nrow, nslot}, (long long)n, kper, (float2*)out, (long long)stride_out,
                     (double2*)work);
  EFN_LAUNCH_CHECK();
  hipLaunchKernelGGL(efn::rowreduce_final_kernel, dim3((m + 255) / 256), dim3(256), 0, s,
                     (const double2*)work, ns, m, (double2*)norm2_c128);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

int efn_row_axpby_norm_strided(const void* X, int64_t stride_x, const double* a, const void* Y,
                               int64_t stride_y, const double* b, int nrow, int m, int64_t n, void* out,
                               int64_t stride_out, double* norm2_c128, void* work, size_t work_bytes,
                               void* stream) {
  return row_axpby_norm_impl(X, stride_x, a, Y, stride_y, b, nullptr, nrow, 0, m, n, out, stride_out,
                             norm2_c128, work, work_bytes, stream);
}

int efn_row_axpby_norm_sel(const void* X, int64_t stride_x, const double* a, const void* Y, int64_t stride_y,
                           const double* b, const int* sel, int nrow, int nslot, int m, int64_t n, void* out,
                           int64_t stride_out, double* norm2_c128, void* work, size_t work_bytes, void* stream) {
  if (const char* e = sel_check(sel, nrow, nslot)) return fail(EFN_EINVAL, e);
  return row_axpby_norm_impl(X, stride_x, a, Y, stride_y, b, sel, nrow, nslot, m, n, out, stride_out,
                             norm2_c128, work, work_bytes, stream);
}

int efn_rowreduce(int op, const void* A, const void* B, int m, int64_t n, void* out, void* work,
                  size_t work_bytes, void* stream) {
  const int nrow = m > 0 ? m : 1;
  return efn_rowreduce_strided(op, A, nrow, B, nrow, nrow, m, n, out, work, work_bytes, stream);
}

int efn_row_axpby_strided(const void* X, int64_t stride_x, const double* a, const void* Y,
                          int64_t stride_y, const double* b, int nrow, int m, int64_t n, void* out,
                          int64_t stride_out, void* stream) {
  if (m < 0 || n < 0) return fail(EFN_EINVAL, "negative size");
  if (m == 0 || n == 0) return EFN_OK;
  if (!X || !a || !out || (Y && !b)) return fail(EFN_EINVAL, "NULL pointer");
  if (nrow <= 0 || stride_x < nrow || stride_out < nrow || (Y && stride_y < nrow))
    return fail(EFN_EINVAL, "row map needs 0 < nrow <= stride");
  long long gx = (n + 255) / 256;
  if (gx > 4096) gx = 4096;
  hipLaunchKernelGGL(efn::row_axpby_kernel, dim3((unsigned)gx, (unsigned)m), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)X, (long long)stride_x, a, (const float2*)Y,
                     (long long)stride_y, b, nrow, (long long)n, (float2*)out, (long long)stride_out);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

int efn_row_axpby(const void* X, const double* a, const void* Y, const double* b, int m, int64_t n,
                  void* out, void* stream) {
  const int nrow = m > 0 ? m : 1;
  return efn_row_axpby_strided(X, nrow, a, Y, nrow, b, nrow, m, n, out, nrow, stream);
}

int efn_axpby(const void* X, float a, const void* Y, float b, int64_t count, void* out, void* stream) {
  if (count < 0) return fail(EFN_EINVAL, "negative size");
  if (count == 0) return EFN_OK;
  if (!X || !out) return fail(EFN_EINVAL, "NULL pointer");
  hipLaunchKernelGGL(efn::axpby_kernel, dim3(grid_for(count)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)X, a, (const float2*)Y, b, (long long)count, (float2*)out);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

int efn_copy_multi(int nprob, const void* const* src, void* const* dst, int64_t count, void* stream) {
  if (nprob < 0 || nprob > efn::kMaxProb)
    return fail(EFN_EINVAL, "nprob must be in [0, " + std::to_string(efn::kMaxProb) + "]");
  if (count < 0) return fail(EFN_EINVAL, "negative size");
  if (nprob == 0 || count == 0) return EFN_OK;
  if (!dst) return fail(EFN_EINVAL, "NULL pointer");
  efn::CopyPtrs cp;
  memset(&cp, 0, sizeof(cp));
  for (int i = 0; i < nprob; ++i) {
    if (!dst[i]) return fail(EFN_EINVAL, "NULL destination");
    cp.src[i] = src ? (const float2*)src[i] : nullptr;
    cp.dst[i] = (float2*)dst[i];
  }
  hipLaunchKernelGGL(efn::copy_multi_kernel, dim3(grid_for(count), (unsigned)nprob), dim3(256), 0,
                     (hipStream_t)stream, cp, (long long)count);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

int efn_random_block(const efn_ctx* c, int kind, uint64_t seed, int64_t row0, int m, void* out,
                     void* stream) {
  int rc;
  if ((rc = check_block(c, out, m, "out"))) return rc;
  if (kind != 0 && kind != 1) return fail(EFN_EINVAL, "kind must be 0 (normal) or 1 (rademacher)");
  if (row0 < 0) return fail(EFN_EINVAL, "negative row offset");
  if (m == 0) return EFN_OK;
  const long long G3 = (long long)c->G * c->G * c->G;
  const long long count = (long long)m * 2 * G3;
  const uint64_t key = [](uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }(seed);
  hipLaunchKernelGGL(efn::random_block_kernel, dim3(grid_for(count)), dim3(256), 0, (hipStream_t)stream,
                     kind, key, (long long)row0, G3, count, c->sing_voxel, (float2*)out);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}


int efn_rasterize(int grid_size, double box_size, double aspect, double bb, const double* rodf_dev,
                  const int* rodi_dev, int nrods, int periodic, float eps_rod, float eps_bg,
                  float* eps_dev, void* stream) {
  if (grid_size <= 0 || grid_size > 4096) return fail(EFN_EINVAL, "grid_size must be in [1, 4096]");
  if (!(box_size > 0.0) || !(aspect > 0.0) || !(bb >= 0.0))
    return fail(EFN_EINVAL, "box_size and aspect must be positive, bb non-negative");
  if (nrods < 0) return fail(EFN_EINVAL, "nrods must be non-negative");
  if (!eps_dev || (nrods > 0 && (!rodf_dev || !rodi_dev))) return fail(EFN_EINVAL, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  const long long count = (long long)grid_size * grid_size * grid_size;
  hipLaunchKernelGGL(efn::fill_f32_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                     eps_dev, count, eps_bg);
  EFN_LAUNCH_CHECK();
  if (nrods == 0) return EFN_OK;
  hipLaunchKernelGGL(efn::raster_kernel, dim3((unsigned)nrods), dim3(256), 0, s, rodf_dev, rodi_dev,
                     grid_size, box_size, aspect, bb, periodic ? 1 : 0, eps_rod, eps_dev);
  EFN_LAUNCH_CHECK();
  return EFN_OK;
}

}  // extern "C"
