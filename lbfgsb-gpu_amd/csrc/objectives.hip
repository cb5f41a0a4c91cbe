// Objective kernels for the benchmark workloads (the callback side of the
// boundary).  They run on the callback's stream, write g in the bulk type and
// per-block FP64 partials of f, which one workgroup finishes on the device.
//   GRID_QUAD  f = 0.5 sum_edges (v_a - v_b)^2 - h^2 sum b v  (zero ring)
//   GRID_COS   f = 0.5 sum_edges (dv)^2 + h^2 sum (c cos v - b v)
//   ROSEN      f = sum_i 100 (x_{2i} - x_{2i-1}^2)^2 + (1 - x_{2i-1})^2
//   PENTA      f = 0.5 x'Ax - b'x, A symmetric pentadiagonal
// The shape follows the reference's example callback
// (examples/dsscfg/dsscfg_cuda.cu:43-133): a device kernel for g and an
// energy reduction whose value is handed back as the host f.
#include "common.hpp"
#include "launch.hpp"
#include "objectives.hpp"

namespace lbfgsb {

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_obj(ObjDesc o, long long n, const T* __restrict__ x, T* __restrict__ g, double* part) {
  __shared__ double sh[kWavesPerBlock];
  const T* b = (const T*)o.b;
  const T* c = (const T*)o.c;
  const T* a0 = (const T*)o.a0;
  const T* a1 = (const T*)o.a1;
  const T* a2 = (const T*)o.a2;
  double acc[1] = {0.0};
  const double h2 = o.h * o.h;
  const long long nx = o.nx, ny = o.ny;
  const long long count = (o.kind == OBJK_ROSEN) ? n / 2 : n;
  for (long long k = blockIdx.x * (long long)kBlock + threadIdx.x; k < count;
       k += (long long)gridDim.x * kBlock) {
    if (o.kind == OBJK_GRID_QUAD || o.kind == OBJK_GRID_COS) {
      const long long ix = k % nx, iy = k / nx;
      const double v = (double)x[k];
      const double vl = ix > 0 ? (double)x[k - 1] : 0.0;
      const double vr = ix < nx - 1 ? (double)x[k + 1] : 0.0;
      const double vd = iy > 0 ? (double)x[k - nx] : 0.0;
      const double vu = iy < ny - 1 ? (double)x[k + nx] : 0.0;
      const double bk = (double)b[k];
      double gk = (4.0 * v - vl - vr - vd - vu) - h2 * bk;
      const double er = vr - v, eu = vu - v;
      double e = er * er + eu * eu;
      if (ix == 0) e += v * v;
      if (iy == 0) e += v * v;
      double fk = 0.5 * e - h2 * bk * v;
      if (o.kind == OBJK_GRID_COS) {
        const double ck = (double)c[k];
        fk += h2 * ck * cos(v);
        gk -= h2 * ck * sin(v);
      }
      g[k] = (T)gk;
      acc[0] += fk;
    } else if (o.kind == OBJK_ROSEN) {
      const double a = (double)x[2 * k], bb = (double)x[2 * k + 1];
      const double t1 = bb - a * a, t2 = 1.0 - a;
      acc[0] += 100.0 * t1 * t1 + t2 * t2;
      g[2 * k] = (T)(-400.0 * t1 * a - 2.0 * t2);
      g[2 * k + 1] = (T)(200.0 * t1);
    } else {
      const double xi = (double)x[k];
      double ax = (double)a0[k] * xi;
      if (k + 1 < n) ax += (double)a1[k] * (double)x[k + 1];
      if (k + 2 < n) ax += (double)a2[k] * (double)x[k + 2];
      if (k >= 1) ax += (double)a1[k - 1] * (double)x[k - 1];
      if (k >= 2) ax += (double)a2[k - 2] * (double)x[k - 2];
      acc[0] += 0.5 * ax * xi - (double)b[k] * xi;
      g[k] = (T)(ax - (double)b[k]);
    }
  }
  block_sum_store<1>(acc, sh, part + blockIdx.x, gridDim.x);
}

__global__ void s_obj_finish(const double* part, int G, double* f_out) {
  const int lane = threadIdx.x & 63;
  if (threadIdx.x >= 64) return;
  double s = 0.0, c = 0.0;
  for (int b = lane; b < G; b += 64) {
    double e;
    two_sum(s, part[b], s, e);
    c += e;
  }
  for (int o = 32; o > 0; o >>= 1) {
    double os = __shfl_xor(s, o, kWave), oc = __shfl_xor(c, o, kWave), e;
    two_sum(s, os, s, e);
    c += oc + e;
  }
  if (lane == 0) *f_out = s + c;
}

template <typename T>
void launch_objective(hipStream_t s, const ObjDesc& o, long long n, const T* x, T* g,
                      double* part, int G, double* f_dev) {
  hipLaunchKernelGGL(k_obj<T>, dim3(G), dim3(kBlock), 0, s, o, n, x, g, part);
  hipLaunchKernelGGL(s_obj_finish, dim3(1), dim3(64), 0, s, part, G, f_dev);
}

template void launch_objective<float>(hipStream_t, const ObjDesc&, long long, const float*,
                                      float*, double*, int, double*);
template void launch_objective<double>(hipStream_t, const ObjDesc&, long long, const double*,
                                       double*, double*, int, double*);

}  // namespace lbfgsb
