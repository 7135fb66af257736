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


// ---- exact primitive tests beyond k-DOP overlap ---------------------------------------
// The reference's primitive layer is the triangle type (primitive/triangle.hpp:
// 40-52) and closest_point; it has no intersection test.  These decide, for
// candidate pairs of the k-DOP narrowphase, whether the primitives themselves
// meet, by one declared rule:
//  * closed sets: triangles (balls) that only touch intersect;
//  * each decision is the sign of an fp64 expression evaluated as written
//    (no contraction: -ffp-contract=off), so ties -- zero determinants, equal
//    distances -- are broken the same way on every run and by the CPU
//    restatement; on integer coordinates up to 2^15 every determinant is
//    exact and the rule is the exact geometric answer;
//  * a triangle with an exactly zero fp64 normal is degenerate and meets
//    nothing.
// Triangle-triangle: both plane-side rejections, then the edge-crossing
// formulation (two closed triangles meet iff an edge of one meets the other
// triangle): a segment crosses a triangle iff the orient3d signed volumes of
// the segment with the triangle's three edges agree in sign (zeros allowed),
// and segments in the triangle's plane -- and coplanar triangles -- are
// decided by orient2d tests in the coordinate plane that drops the largest
// normal component.  One thread per pair.
struct V3 {
  double x, y, z;
};
struct V2 {
  double u, v;
};

__device__ __forceinline__ double orient3d(V3 a, V3 b, V3 c, V3 d) {
  const double adx = a.x - d.x, ady = a.y - d.y, adz = a.z - d.z;
  const double bdx = b.x - d.x, bdy = b.y - d.y, bdz = b.z - d.z;
  const double cdx = c.x - d.x, cdy = c.y - d.y, cdz = c.z - d.z;
  return adx * (bdy * cdz - bdz * cdy) + bdx * (cdy * adz - cdz * ady) + cdx * (ady * bdz - adz * bdy);
}

__device__ __forceinline__ V3 tri_normal(V3 a, V3 b, V3 c) {
  const double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
  const double vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
  return V3{uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx};
}

// the coordinate dropped by the 2D tests: largest |n| component, the first on ties
__device__ __forceinline__ int drop_axis(V3 n) {
  const double ax = fabs(n.x), ay = fabs(n.y), az = fabs(n.z);
  if (ax >= ay && ax >= az) return 0;
  return ay >= az ? 1 : 2;
}

__device__ __forceinline__ V2 proj2(V3 p, int drop) {
  return drop == 0 ? V2{p.y, p.z} : (drop == 1 ? V2{p.x, p.z} : V2{p.x, p.y});
}

__device__ __forceinline__ double orient2d(V2 a, V2 b, V2 c) {
  return (a.u - c.u) * (b.v - c.v) - (a.v - c.v) * (b.u - c.u);
}

__device__ __forceinline__ bool on_seg2(V2 a, V2 b, V2 p) {
  const double ulo = a.u < b.u ? a.u : b.u, uhi = a.u < b.u ? b.u : a.u;
  const double vlo = a.v < b.v ? a.v : b.v, vhi = a.v < b.v ? b.v : a.v;
  return ulo <= p.u && p.u <= uhi && vlo <= p.v && p.v <= vhi;
}

__device__ __forceinline__ bool seg_seg2(V2 p1, V2 p2, V2 q1, V2 q2) {
  const double d1 = orient2d(q1, q2, p1), d2 = orient2d(q1, q2, p2);
  const double d3 = orient2d(p1, p2, q1), d4 = orient2d(p1, p2, q2);
  if (((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0)) && ((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0))) return true;
  if (d1 == 0 && on_seg2(q1, q2, p1)) return true;
  if (d2 == 0 && on_seg2(q1, q2, p2)) return true;
  if (d3 == 0 && on_seg2(p1, p2, q1)) return true;
  if (d4 == 0 && on_seg2(p1, p2, q2)) return true;
  return false;
}

__device__ __forceinline__ bool in_tri2(V2 t0, V2 t1, V2 t2, V2 p) {
  const double d1 = orient2d(t0, t1, p), d2 = orient2d(t1, t2, p), d3 = orient2d(t2, t0, p);
  return (d1 >= 0 && d2 >= 0 && d3 >= 0) || (d1 <= 0 && d2 <= 0 && d3 <= 0);
}

// closed segment [a, b] in the plane of triangle t
__device__ __noinline__ bool seg_tri2(V3 a, V3 b, const V3* t, int drop) {
  const V2 a2 = proj2(a, drop), b2 = proj2(b, drop);
  const V2 t0 = proj2(t[0], drop), t1 = proj2(t[1], drop), t2 = proj2(t[2], drop);
  if (in_tri2(t0, t1, t2, a2) || in_tri2(t0, t1, t2, b2)) return true;
  return seg_seg2(a2, b2, t0, t1) || seg_seg2(a2, b2, t1, t2) || seg_seg2(a2, b2, t2, t0);
}

// closed segment [a, b] against the closed, non-degenerate triangle t
__device__ __forceinline__ bool seg_tri(V3 a, V3 b, const V3* t, int drop) {
  const double oa = orient3d(t[0], t[1], t[2], a), ob = orient3d(t[0], t[1], t[2], b);
  if ((oa > 0 && ob > 0) || (oa < 0 && ob < 0)) return false;
  if (oa == 0 && ob == 0) return seg_tri2(a, b, t, drop);
  const double s1 = orient3d(a, b, t[0], t[1]), s2 = orient3d(a, b, t[1], t[2]), s3 = orient3d(a, b, t[2], t[0]);
  return (s1 >= 0 && s2 >= 0 && s3 >= 0) || (s1 <= 0 && s2 <= 0 && s3 <= 0);
}

__device__ __noinline__ bool coplanar_tri_tri(const V3* P, const V3* Q, int drop) {
  V2 p2[3], q2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    p2[i] = proj2(P[i], drop);
    q2[i] = proj2(Q[i], drop);
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (seg_seg2(p2[i], p2[(i + 1) % 3], q2[j], q2[(j + 1) % 3])) return true;
  return in_tri2(q2[0], q2[1], q2[2], p2[0]) || in_tri2(p2[0], p2[1], p2[2], q2[0]);
}

__device__ bool tri_tri(const V3* P, const V3* Q) {
  const V3 np = tri_normal(P[0], P[1], P[2]), nq = tri_normal(Q[0], Q[1], Q[2]);
  if ((np.x == 0 && np.y == 0 && np.z == 0) || (nq.x == 0 && nq.y == 0 && nq.z == 0)) return false;
  double dq[3], dp[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) dq[i] = orient3d(P[0], P[1], P[2], Q[i]);
  if ((dq[0] > 0 && dq[1] > 0 && dq[2] > 0) || (dq[0] < 0 && dq[1] < 0 && dq[2] < 0)) return false;
#pragma unroll
  for (int i = 0; i < 3; ++i) dp[i] = orient3d(Q[0], Q[1], Q[2], P[i]);
  if ((dp[0] > 0 && dp[1] > 0 && dp[2] > 0) || (dp[0] < 0 && dp[1] < 0 && dp[2] < 0)) return false;
  const int dropp = drop_axis(np), dropq = drop_axis(nq);
  if ((dq[0] == 0 && dq[1] == 0 && dq[2] == 0) || (dp[0] == 0 && dp[1] == 0 && dp[2] == 0))
    return coplanar_tri_tri(P, Q, dropp);
  for (int i = 0; i < 3; ++i)
    if (seg_tri(P[i], P[(i + 1) % 3], Q, dropq)) return true;
  for (int i = 0; i < 3; ++i)
    if (seg_tri(Q[i], Q[(i + 1) % 3], P, dropp)) return true;
  return false;
}

__device__ __forceinline__ bool sphere_sphere(const double* a, const double* b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  const double rs = a[3] + b[3];
  return dx * dx + dy * dy + dz * dz <= rs * rs;
}

// a triangle of element e of SoA vertex planes (kPoly, 3 vertices)
__device__ __forceinline__ void load_tri(const ElemInput& in, uint32_t e, V3* t) {
  const uint64_t n = in.n;
#pragma unroll
  for (int v = 0; v < 3; ++v)
    t[v] = V3{in.data[(uint64_t)(3 * v) * n + e], in.data[(uint64_t)(3 * v + 1) * n + e],
              in.data[(uint64_t)(3 * v + 2) * n + e]};
}

__device__ __forceinline__ void load_sphere(const ElemInput& in, uint32_t e, double* s) {
  const uint64_t n = in.n;
#pragma unroll
  for (int c = 0; c < 4; ++c) s[c] = in.data[(uint64_t)c * n + e];
}

// n pairs of AoS triangles (9 doubles each) or balls (4 doubles)
__global__ __launch_bounds__(256) void k_tri_pairs(const double* __restrict__ p, const double* __restrict__ q,
                                                   uint64_t n, uint8_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    V3 P[3], Q[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      P[v] = V3{p[9 * i + 3 * v], p[9 * i + 3 * v + 1], p[9 * i + 3 * v + 2]};
      Q[v] = V3{q[9 * i + 3 * v], q[9 * i + 3 * v + 1], q[9 * i + 3 * v + 2]};
    }
    out[i] = tri_tri(P, Q) ? 1 : 0;
  }
}

__global__ __launch_bounds__(256) void k_sphere_pairs(const double* __restrict__ a, const double* __restrict__ b,
                                                      uint64_t n, uint8_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = sphere_sphere(a + 4 * i, b + 4 * i) ? 1 : 0;
}

// Result pairs (input index of a << 32 | input index of b) whose primitives
// meet: flag[i] in {0, 1} as a 32-bit word (then scanned into offsets).
template <bool TRI>
__global__ __launch_bounds__(256) void k_pair_flags(const uint64_t* __restrict__ pairs, uint64_t count, ElemInput qa,
                                                    ElemInput tb, uint32_t* __restrict__ flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = pairs[i];
    const uint32_t a = (uint32_t)(key >> 32), b = (uint32_t)key;
    bool hit;
    if constexpr (TRI) {
      V3 P[3], Q[3];
      load_tri(qa, a, P);
      load_tri(tb, b, Q);
      hit = tri_tri(P, Q);
    } else {
      double sa[4], sb[4];
      load_sphere(qa, a, sa);
      load_sphere(tb, b, sb);
      hit = sphere_sphere(sa, sb);
    }
    flag[i] = hit ? 1u : 0u;
  }
}

// stable compaction: pair i goes to out[off[i]] when it was kept; the last
// thread also writes the kept count (off[n-1] + its own flag)
__global__ __launch_bounds__(256) void k_pair_compact(const uint64_t* __restrict__ pairs, uint64_t count,
                                                      const ElemInput qa, const ElemInput tb, int tri,
                                                      const uint32_t* __restrict__ off, uint64_t* __restrict__ out,
                                                      unsigned long long* __restrict__ kept) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t o = off[i];
    const uint32_t next = (i + 1 < count) ? off[i + 1] : 0u;
    bool keep;
    if (i + 1 < count) {
      keep = next != o;  // the scan advanced past pair i
    } else {
      // the last pair: decide it again (its flag was overwritten by the scan)
      const uint64_t key = pairs[i];
      const uint32_t a = (uint32_t)(key >> 32), b = (uint32_t)key;
      if (tri) {
        V3 P[3], Q[3];
        load_tri(qa, a, P);
        load_tri(tb, b, Q);
        keep = tri_tri(P, Q);
      } else {
        double sa[4], sb[4];
        load_sphere(qa, a, sa);
        load_sphere(tb, b, sb);
        keep = sphere_sphere(sa, sb);
      }
      *kept = (unsigned long long)o + (keep ? 1ull : 0ull);
    }
    if (keep) out[o] = pairs[i];
  }
}

static unsigned grid_for(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (unsigned)(b > 8192 ? 8192 : (b ? b : 1));
}

hipError_t launch_tri_pairs(const double* p, const double* q, uint64_t n, uint8_t* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tri_pairs, dim3(grid_for(n)), dim3(256), 0, s, p, q, n, out);
  return hipGetLastError();
}

hipError_t launch_sphere_pairs(const double* a, const double* b, uint64_t n, uint8_t* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sphere_pairs, dim3(grid_for(n)), dim3(256), 0, s, a, b, n, out);
  return hipGetLastError();
}

hipError_t launch_pair_flags(const uint64_t* pairs, uint64_t count, const ElemInput& qa, const ElemInput& tb, int tri,
                             uint32_t* flag, hipStream_t s) {
  if (count == 0) return hipSuccess;
  if (tri) hipLaunchKernelGGL(k_pair_flags<true>, dim3(grid_for(count)), dim3(256), 0, s, pairs, count, qa, tb, flag);
  else hipLaunchKernelGGL(k_pair_flags<false>, dim3(grid_for(count)), dim3(256), 0, s, pairs, count, qa, tb, flag);
  return hipGetLastError();
}

hipError_t launch_pair_compact(const uint64_t* pairs, uint64_t count, const ElemInput& qa, const ElemInput& tb,
                               int tri, const uint32_t* off, uint64_t* out, unsigned long long* kept, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_compact, dim3(grid_for(count)), dim3(256), 0, s, pairs, count, qa, tb, tri, off, out,
                     kept);
  return hipGetLastError();
}

}  // namespace dbvh
