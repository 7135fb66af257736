// primitive.hip -- the closest point on a triangle for a batch of points.
//
// Replaces bvh::kokkos::closest_point (primitive/kokkos_closest_point.hpp:
// 55-133): the region tests of the closest-point-on-triangle method
// (vertex, edge and face regions from the dot products of the edge vectors
// with the vectors to the point), evaluated with the reference's expressions
// in its order (dot = (x0 y0 + x1 y1) + x2 y2, vector expressions left to
// right, no contraction), so the results are bit for bit the reference's.
// One thread per point; the points and results are n x 3 arrays.
#include "device.hpp"
#include "kernels.hpp"

namespace dbvh {

// m::dot: the products summed left to right from 0 (the sign of an exact
// zero follows 0 + x, which the compiler keeps without fast-math)
__device__ __forceinline__ double dot3(const double* u, const double* v) {
  double s = 0.0;
  s = s + u[0] * v[0];
  s = s + u[1] * v[1];
  s = s + u[2] * v[2];
  return s;
}

__global__ __launch_bounds__(256) void k_closest_point(const double* __restrict__ pts, uint64_t n, Tri tri,
                                                       double* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ab[3], ac[3], bc[3], p[3], ap[3], bp[3], cp[3];
  for (int k = 0; k < 3; ++k) {
    ab[k] = tri.b[k] - tri.a[k];
    ac[k] = tri.c[k] - tri.a[k];
    bc[k] = tri.c[k] - tri.b[k];
    p[k] = pts[3 * i + k];
    ap[k] = p[k] - tri.a[k];
    bp[k] = p[k] - tri.b[k];
    cp[k] = p[k] - tri.c[k];
  }
  double* o = out + 3 * i;
  // vertex region a
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) {
    for (int k = 0; k < 3; ++k) o[k] = tri.a[k];
    return;
  }
  // vertex region b
  const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) {
    for (int k = 0; k < 3; ++k) o[k] = tri.b[k];
    return;
  }
  // edge region ab
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    const double v = d1 / (d1 - d3);
    for (int k = 0; k < 3; ++k) o[k] = tri.a[k] + v * ab[k];
    return;
  }
  // vertex region c
  const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) {
    for (int k = 0; k < 3; ++k) o[k] = tri.c[k];
    return;
  }
  // edge region ac
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
    for (int k = 0; k < 3; ++k) o[k] = tri.a[k] + w * ac[k];
    return;
  }
  // edge region bc
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int k = 0; k < 3; ++k) o[k] = tri.b[k] + w * bc[k];
    return;
  }
  // interior
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
  for (int k = 0; k < 3; ++k) o[k] = tri.a[k] + ab[k] * v + ac[k] * w;
}

hipError_t launch_closest_point(const double* pts, uint64_t n, const Tri& tri, double* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (!pts || !out) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_closest_point, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pts, n, tri, out);
  return hipGetLastError();
}

}  // namespace dbvh
