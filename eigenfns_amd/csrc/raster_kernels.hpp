// Binary pen-like rasterisation of rod networks on the device.
//
// Reference: eigenfns/structure.py:51-151 (rasterize_penlike).  Each rod is a
// cylinder of radius b in the unwarped space z' = z/s; a voxel centre c is
// inside when, with R = (c_x − p1u_x, c_y − p1u_y, c_z/s − p1u_z),
//   t = (R_x·n_x + R_y·n_y) + R_z·n_z,   0 ≤ t ≤ L_u,
//   |R − t·n|² = (r_x² + r_y²) + r_z² ≤ b².
// The per-rod FP64 parameters (p1u, n, L_u, index box) are computed on the
// host with the reference's own NumPy expressions; this kernel evaluates the
// per-voxel test with the same FP64 operations in the same order and no
// multiply-add contraction, so the mask is bit-identical.  Voxel centres are
// (g + 0.5)·dx − L/2 for the (unwrapped, when periodic) index g.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace efn {

constexpr int kRodF = 8;  // doubles per rod: p1u[3], n[3], L_u, unused
constexpr int kRodI = 6;  // ints per rod: first/last voxel index along x, y, z

__global__ void fill_f32_kernel(float* __restrict__ out, long long count, float value) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = value;
}

// grid: (rods); each block walks its rod's voxel box, z fastest.  Voxels inside
// several rods are written by several blocks with the same value.
__global__ __launch_bounds__(256) void raster_kernel(const double* __restrict__ rodf,
                                                     const int* __restrict__ rodi, int G,
                                                     double box_size, double s, double bb,
                                                     int periodic, float eps_rod,
                                                     float* __restrict__ eps) {
#pragma clang fp contract(off)
  const double* p = rodf + (long long)blockIdx.x * kRodF;
  const int* r = rodi + (long long)blockIdx.x * kRodI;
  const double dx = box_size / G;
  const double half = box_size / 2.0;
  const double p0 = p[0], p1 = p[1], p2 = p[2], n0 = p[3], n1 = p[4], n2 = p[5], Lu = p[6];
  const int x0 = r[0], y0 = r[2], z0 = r[4];
  const int nx = r[1] - x0 + 1, ny = r[3] - y0 + 1, nz = r[5] - z0 + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  const long long nyz = (long long)ny * nz;
  const long long nvox = (long long)nx * nyz;
  for (long long v = threadIdx.x; v < nvox; v += blockDim.x) {
    const int ix = (int)(v / nyz);
    const long long rem = v - (long long)ix * nyz;
    const int iy = (int)(rem / nz);
    const int iz = (int)(rem - (long long)iy * nz);
    const int gx = x0 + ix, gy = y0 + iy, gz = z0 + iz;
    const double cx = ((double)gx + 0.5) * dx - half;
    const double cy = ((double)gy + 0.5) * dx - half;
    const double cz = ((double)gz + 0.5) * dx - half;
    const double RX = cx - p0;
    const double RY = cy - p1;
    const double RZ = cz / s - p2;
    double t = RX * n0 + RY * n1;
    t = t + RZ * n2;
    const double rX = RX - t * n0;
    const double rY = RY - t * n1;
    const double rZ = RZ - t * n2;
    double d2 = rX * rX + rY * rY;
    d2 = d2 + rZ * rZ;
    if (t >= 0.0 && t <= Lu && d2 <= bb) {
      int ox = gx, oy = gy, oz = gz;
      if (periodic) {
        ox = ((gx % G) + G) % G;
        oy = ((gy % G) + G) % G;
        oz = ((gz % G) + G) % G;
      } else if (gx < 0 || gx >= G || gy < 0 || gy >= G || gz < 0 || gz >= G) {
        continue;  // clipped boxes lie inside the grid; never write outside it
      }
      eps[((long long)ox * G + oy) * G + oz] = eps_rod;
    }
  }
}

}  // namespace efn
