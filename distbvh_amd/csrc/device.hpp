// device.hpp -- shared gfx950 device code: 26-DOP arithmetic, Morton codes,
// node layout, wave helpers.
//
// All fp64 code is compiled with -ffp-contract=off so that every expression
// is evaluated as written (bit-exact with the CPU restatement).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dbvh {

constexpr int kAxes = 13;
constexpr int kSlabs = 26;
constexpr uint32_t kEnd = 0xffffffffu;     // "no node" / end of traversal
constexpr uint32_t kMixed = 0xffffffffu;   // node spans several patches
constexpr int kWave = 64;

// The 13 positive normals of dop_26 (kdop.hpp:657-677).
#define DBVH_R2 0.7071067811865475
#define DBVH_R3 0.5773502691896258

// project(v, axis) = v0*a0 + v1*a1 + v2*a2 left to right (kdop.hpp:223-227).
// The normal components are compile-time constants; multiplication by 1.0
// folds to the identity (exact), the 0.0 and +-R terms are kept.
template <int A>
__device__ __forceinline__ double project(double x, double y, double z) {
  constexpr double N[13][3] = {
      {1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.},
      {DBVH_R3, DBVH_R3, DBVH_R3}, {DBVH_R3, -DBVH_R3, DBVH_R3},
      {DBVH_R3, DBVH_R3, -DBVH_R3}, {DBVH_R3, -DBVH_R3, -DBVH_R3},
      {DBVH_R2, DBVH_R2, 0.0}, {DBVH_R2, 0.0, DBVH_R2}, {0.0, DBVH_R2, DBVH_R2},
      {DBVH_R2, -DBVH_R2, 0.0}, {DBVH_R2, 0.0, -DBVH_R2}, {0.0, DBVH_R2, -DBVH_R2}};
  return x * N[A][0] + y * N[A][1] + z * N[A][2];
}

// std::min / std::max (kdop.hpp:258-259) and m::min / m::max
// (math/common.hpp:86-96) as selects; never v_min_f64 (different +-0 / NaN
// rules).
__device__ __forceinline__ double sel_min(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double sel_max(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ float sel_minf(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float sel_maxf(float a, float b) { return (a < b) ? b : a; }

struct Kdop {
  double s[kSlabs];  // min0, max0, ..., min12, max12
};

template <int A = 0>
__device__ __forceinline__ void proj_all(double x, double y, double z, double* p) {
  p[A] = project<A>(x, y, z);
  if constexpr (A + 1 < kAxes) proj_all<A + 1>(x, y, z, p);
}

// from_vertices, kdop.hpp:238-264.
template <int NV>
__device__ __forceinline__ void kdop_from_vertices(const double (&v)[NV][3], double eps, double* s) {
  double p[kAxes];
  proj_all(v[0][0], v[0][1], v[0][2], p);
#pragma unroll
  for (int a = 0; a < kAxes; ++a) {
    s[2 * a] = p[a] - eps;
    s[2 * a + 1] = p[a] + eps;
  }
#pragma unroll
  for (int j = 1; j < NV; ++j) {
    proj_all(v[j][0], v[j][1], v[j][2], p);
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      s[2 * a] = sel_min(s[2 * a], p[a] - eps);
      s[2 * a + 1] = sel_max(s[2 * a + 1], p[a] + eps);
    }
  }
}

// from_sphere, kdop.hpp:273-288.
__device__ __forceinline__ void kdop_from_sphere(double x, double y, double z, double r, double* s) {
  double p[kAxes];
  proj_all(x, y, z, p);
#pragma unroll
  for (int a = 0; a < kAxes; ++a) {
    s[2 * a] = p[a] - r;
    s[2 * a + 1] = p[a] + r;
  }
}

// overlap(kdop, kdop), kdop.hpp:120-125 / 538-549 (strict).
__device__ __forceinline__ bool overlap(const double* a, const double* b) {
  bool o = true;
#pragma unroll
  for (int i = 0; i < kAxes; ++i) o = o & (a[2 * i] < b[2 * i + 1]) & (b[2 * i] < a[2 * i + 1]);
  return o;
}
__device__ __forceinline__ bool overlapf(const float* a, const float* b) {
  bool o = true;
#pragma unroll
  for (int i = 0; i < kAxes; ++i) o = o & (a[2 * i] < b[2 * i + 1]) & (b[2 * i] < a[2 * i + 1]);
  return o;
}

// union_with, kdop.hpp:317-325 (m::min / m::max).
__device__ __forceinline__ void kdop_union(double* d, const double* s) {
#pragma unroll
  for (int i = 0; i < kAxes; ++i) {
    d[2 * i] = (d[2 * i] < s[2 * i]) ? d[2 * i] : s[2 * i];
    d[2 * i + 1] = (d[2 * i + 1] > s[2 * i + 1]) ? d[2 * i + 1] : s[2 * i + 1];
  }
}

// ---- Morton (hash.hpp) ------------------------------------------------------
__device__ __forceinline__ uint32_t expand32(uint32_t v) {
  v = (v * 0x00010001u) & 0xff0000ffu;
  v = (v * 0x00000101u) & 0x0f00f00fu;
  v = (v * 0x00000011u) & 0xc30c30c3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
__device__ __forceinline__ uint64_t expand64(uint64_t v) {
  v = (v * 0x1000001ull) & 0xfff000000fffull;
  v = (v * 0x1001ull) & 0xfc003f000fc003full;
  v = (v * 0x41ull) & 0x01c0e070381c0e07ull;
  v = (v * 0x11ull) & 0x10c86432190c8643ull;
  v = (v * 0x5ull) & 0x1249249249249249ull;
  return v;
}
// The reference's static_cast<uintN_t>(double) as its x86-64 build evaluates
// it: truncation to int64 and reinterpretation; NaN, +-inf and |s| >= 2^63
// give 0 (what remains after the mask).  A degenerate axis (0 * inf = NaN,
// hash.hpp:151-155) therefore quantizes to 0.
__device__ __forceinline__ uint64_t float_to_uint(double s) {
  if (!(s > -9223372036854775808.0 && s < 9223372036854775808.0)) return 0;
  return (uint64_t)(int64_t)s;
}

struct Bounds {
  double min[3];
  double inv[3];
  double max[3];
  double pad;
};

// quantize64 + morton (hash.hpp:177-198, 121-130; contact_entity.hpp:112-117)
__device__ __forceinline__ uint64_t morton63(double x, double y, double z, const Bounds& b) {
  uint64_t q[3];
  const double p[3] = {x, y, z};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double norm = (p[i] - b.min[i]) * b.inv[i];
    q[i] = float_to_uint(norm * 2097152.0) & 0x1fffffull;
  }
  return (expand64(q[2]) << 2) | (expand64(q[1]) << 1) | expand64(q[0]);
}
// quantize32 + morton (hash.hpp:146-167, 109-113)
__device__ __forceinline__ uint32_t morton30(double x, double y, double z, const Bounds& b) {
  uint32_t q[3];
  const double p[3] = {x, y, z};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double norm = (p[i] - b.min[i]) * b.inv[i];
    q[i] = (uint32_t)float_to_uint(norm * 1024.0) & 0x3ffu;
  }
  return (expand32(q[2]) << 2) + (expand32(q[1]) << 1) + expand32(q[0]);
}

// Work-group b runs on XCD b % 8 (round-robin dispatch).  Tile t = xcd_tile(b)
// gives each XCD a contiguous range of tiles (a bijection of [0, nb)), so
// tiles that are adjacent in the sorted order (and touch the same lines:
// neighbouring parts of the tree, elements near each other in the input)
// share one L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
  const uint32_t x = b & 7u, j = b >> 3, per = nb >> 3, rem = nb & 7u;
  return x * per + (x < rem ? x : rem) + j;
}

// ---- tree records ---------------------------------------------------------------
// Internal node k (k in [0, n-2]) is the split between sorted positions k and
// k+1; leaf p is node n-1+p.
// One record per internal node holding both children's slabs, so a
// traversal step tests two nodes with one memory round trip.  A node's own
// slabs live in its parent's record (the root's in a separate array).
// fp32 form: 256 B, read with scalar loads; fp64 form: 448 B, staged through
// LDS by a wave-wide vector load.
// Rec32: two 128-byte halves, one per child, each written as seven 16-byte
// stores (slabs, child id, element) by the child that completes: a half is
// the child's ordered fp32 slab images (104 B), its node id and, for a leaf,
// its input index; half 0 also holds this node's range and patch (one 16-byte
// store by the node's completer).
struct alignas(256) Rec32 {
  struct alignas(16) Half {
    int32_t box[kSlabs];  // ordered images of the child's fp32 slabs
    uint32_t child;       // child node id (leaf p = n-1+p)
    uint32_t elem;        // input index of a leaf child
    uint32_t first, last; // half 0: this node's range of sorted positions
    uint32_t patch;       // half 0: patch id, or kMixed
    uint32_t pad;
  } h[2];
};
static_assert(sizeof(Rec32) == 256 && sizeof(Rec32::Half) == 128, "record layout");

struct alignas(64) Rec64 {
  double box[2][kSlabs];  // left, right child (416 B)
  uint32_t child[2];
  uint32_t elem[2];
  uint32_t first, last;
  uint32_t patch;
  uint32_t pad;
};
static_assert(sizeof(Rec64) == 448, "record layout");

// field access common to both record formats
__host__ __device__ __forceinline__ uint32_t rec_child(const Rec64& r, int s) { return r.child[s]; }
__host__ __device__ __forceinline__ uint32_t rec_child(const Rec32& r, int s) { return r.h[s].child; }
__host__ __device__ __forceinline__ uint32_t rec_elem(const Rec64& r, int s) { return r.elem[s]; }
__host__ __device__ __forceinline__ uint32_t rec_elem(const Rec32& r, int s) { return r.h[s].elem; }
__host__ __device__ __forceinline__ uint32_t rec_patch(const Rec64& r) { return r.patch; }
__host__ __device__ __forceinline__ uint32_t rec_patch(const Rec32& r) { return r.h[0].patch; }
__host__ __device__ __forceinline__ uint32_t rec_first(const Rec64& r) { return r.first; }
__host__ __device__ __forceinline__ uint32_t rec_first(const Rec32& r) { return r.h[0].first; }
__host__ __device__ __forceinline__ uint32_t rec_last(const Rec64& r) { return r.last; }
__host__ __device__ __forceinline__ uint32_t rec_last(const Rec32& r) { return r.h[0].last; }
__host__ __device__ __forceinline__ const double* rec_box(const Rec64& r, int s) { return r.box[s]; }
__host__ __device__ __forceinline__ const int32_t* rec_box(const Rec32& r, int s) { return r.h[s].box; }
// this node's range and patch, written by its completer
__device__ __forceinline__ void rec_set_range(Rec64& r, uint32_t first, uint32_t last, uint32_t patch) {
  r.first = first;
  r.last = last;
  r.patch = patch;
}
__device__ __forceinline__ void rec_set_range(Rec32& r, uint32_t first, uint32_t last, uint32_t patch) {
  *(uint4*)&r.h[0].first = make_uint4(first, last, patch, 0u);
}

// Where leaf p's slabs live: the half `side` of internal node `parent`'s
// record.  Leaf p is the left child of internal node p when delta(p) <
// delta(p-1) (or p == 0), else the right child of internal node p-1 (the
// bottom-up build's rule for a one-position node).  Requires n >= 2.
__device__ __forceinline__ void leaf_slot(const uint8_t* delta, uint64_t n, uint64_t p, uint32_t& parent,
                                          int& side) {
  const bool is_left = (p == 0) || (p != n - 1 && delta[p] < delta[p - 1]);
  parent = (uint32_t)(is_left ? p : p - 1);
  side = is_left ? 0 : 1;
}

template <bool FP32>
struct Form {
  using T = double;
  using Rec = Rec64;
};
template <>
struct Form<true> {
  using T = int32_t;  // ord32 image of an outward-rounded fp32 slab
  using Rec = Rec32;
};


// Order-preserving map of a double to an unsigned 64-bit integer (x < y ==>
// ordered_u64(x) < ordered_u64(y) for non-NaN values), so integer atomic
// min / max act as double min / max.
__host__ __device__ __forceinline__ uint64_t ordered_u64(double x) {
  uint64_t u;
  __builtin_memcpy(&u, &x, 8);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double from_ordered_u64(uint64_t u) {
  u = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
  double x;
  __builtin_memcpy(&x, &u, 8);
  return x;
}

// ---- element data ----------------------------------------------------------------
enum ElemKind : int { kPoly = 0, kSphere = 1, kSnapshot = 2 };

struct ElemInput {
  int kind;
  int nv;                   // vertices per element (kPoly)
  double eps;               // from_vertices epsilon (kPoly)
  const double* data;       // planes (kPoly / kSphere) or n x 26 kdops (kSnapshot)
  const double* centroids;  // n x 3 (kSnapshot)
  uint64_t n;
};

// next representable float below / above a finite f (IEEE-754 bit stepping)
__device__ __forceinline__ float next_down(float f) {
  if (!(f == f) || f == -__builtin_inff()) return f;
  if (f == 0.0f) return __uint_as_float(0x80000001u);
  const uint32_t u = __float_as_uint(f);
  return __uint_as_float(f > 0.0f ? u - 1u : u + 1u);
}
__device__ __forceinline__ float next_up(float f) {
  if (!(f == f) || f == __builtin_inff()) return f;
  if (f == 0.0f) return __uint_as_float(0x00000001u);
  const uint32_t u = __float_as_uint(f);
  return __uint_as_float(f > 0.0f ? u + 1u : u - 1u);
}
__device__ __forceinline__ float round_down_1ulp(double x) { return next_down((float)x); }
__device__ __forceinline__ float round_up_1ulp(double x) { return next_up((float)x); }

template <int KIND, int NV>
__device__ __forceinline__ void element_kdop(const ElemInput& in, uint32_t e, double* kd) {
  const uint64_t n = in.n;
  if constexpr (KIND == kPoly) {
    double v[NV][3];
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) v[j][c] = in.data[(uint64_t)(3 * j + c) * n + e];
    kdop_from_vertices<NV>(v, in.eps, kd);
  } else if constexpr (KIND == kSphere) {
    kdop_from_sphere(in.data[e], in.data[n + e], in.data[2 * n + e], in.data[3 * n + e], kd);
  } else {
    const double2* k = (const double2*)(in.data + (uint64_t)e * kSlabs);
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      double2 t = k[a];
      kd[2 * a] = t.x;
      kd[2 * a + 1] = t.y;
    }
  }
}

// element_kdop with the kind and vertex count chosen at run time (a
// wave-uniform switch)
__device__ __forceinline__ void element_kdop_any(const ElemInput& in, uint32_t e, double* kd) {
  if (in.kind == kSphere) {
    element_kdop<kSphere, 1>(in, e, kd);
  } else if (in.kind == kSnapshot) {
    element_kdop<kSnapshot, 1>(in, e, kd);
  } else {
    switch (in.nv) {
      case 1: element_kdop<kPoly, 1>(in, e, kd); break;
      case 2: element_kdop<kPoly, 2>(in, e, kd); break;
      case 3: element_kdop<kPoly, 3>(in, e, kd); break;
      case 4: element_kdop<kPoly, 4>(in, e, kd); break;
      case 5: element_kdop<kPoly, 5>(in, e, kd); break;
      case 6: element_kdop<kPoly, 6>(in, e, kd); break;
      case 7: element_kdop<kPoly, 7>(in, e, kd); break;
      default: element_kdop<kPoly, 8>(in, e, kd); break;
    }
  }
}

// fp32 outward slabs relative to the object origin: the fp64 slab minus the
// origin's projection (fp64), rounded to nearest fp32 and moved one ulp out.
// Strict overlap of element e (from element data) with the k-DOP kd, one
// axis at a time so that e's 26 slabs are never live together.  Each slab is
// computed with the operations and order of element_kdop, so the result is
// overlap(kd, element_kdop(e)).
template <int KIND, int NV, int A = 0>
__device__ __forceinline__ bool overlap_axes(const double (&v)[NV][3], double eps, const double* kd) {
  double lo, hi;
  const double p0 = project<A>(v[0][0], v[0][1], v[0][2]);
  lo = p0 - eps;
  hi = p0 + eps;
  if constexpr (KIND == kPoly) {
#pragma unroll
    for (int j = 1; j < NV; ++j) {
      const double pj = project<A>(v[j][0], v[j][1], v[j][2]);
      lo = sel_min(lo, pj - eps);
      hi = sel_max(hi, pj + eps);
    }
  }
  const bool o = (kd[2 * A] < hi) & (lo < kd[2 * A + 1]);
  if constexpr (A + 1 < kAxes) return o & overlap_axes<KIND, NV, A + 1>(v, eps, kd);
  else return o;
}

template <int KIND, int NV>
__device__ __forceinline__ bool element_overlap(const ElemInput& in, uint32_t e, const double* kd) {
  const uint64_t n = in.n;
  if constexpr (KIND == kPoly) {
    double v[NV][3];
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) v[j][c] = in.data[(uint64_t)(3 * j + c) * n + e];
    return overlap_axes<kPoly, NV>(v, in.eps, kd);
  } else if constexpr (KIND == kSphere) {
    const double v[1][3] = {{in.data[e], in.data[n + e], in.data[2 * n + e]}};
    return overlap_axes<kSphere, 1>(v, in.data[3 * n + e], kd);
  } else {
    const double2* k = (const double2*)(in.data + (uint64_t)e * kSlabs);
    bool o = true;
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      const double2 t = k[a];
      o = o & (kd[2 * a] < t.y) & (t.x < kd[2 * a + 1]);
    }
    return o;
  }
}

__device__ __forceinline__ bool element_overlap_any(const ElemInput& in, uint32_t e, const double* kd) {
  if (in.kind == kSphere) return element_overlap<kSphere, 1>(in, e, kd);
  if (in.kind == kSnapshot) return element_overlap<kSnapshot, 1>(in, e, kd);
  switch (in.nv) {
    case 1: return element_overlap<kPoly, 1>(in, e, kd);
    case 2: return element_overlap<kPoly, 2>(in, e, kd);
    case 3: return element_overlap<kPoly, 3>(in, e, kd);
    case 4: return element_overlap<kPoly, 4>(in, e, kd);
    case 5: return element_overlap<kPoly, 5>(in, e, kd);
    case 6: return element_overlap<kPoly, 6>(in, e, kd);
    case 7: return element_overlap<kPoly, 7>(in, e, kd);
    default: return element_overlap<kPoly, 8>(in, e, kd);
  }
}

// Ordered-integer image of a float: monotone, consecutive floats differ by
// 1, ord(-0) = ord(+0) = 0, so integer compares and min / max act as float
// compares and min / max.  The fp32 form stores its slabs this way; a NaN
// lower slab becomes INT32_MAX and a NaN upper slab INT32_MIN, so that (like
// the fp64 compare with NaN) no overlap test involving it holds.
__device__ __forceinline__ int32_t ord32(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u >> 31) ? -(int32_t)(u & 0x7fffffffu) : (int32_t)u;
}
__device__ __forceinline__ void ord32_slabs(const float* kf, int32_t* ko) {
#pragma unroll
  for (int a = 0; a < kAxes; ++a) {
    const float lo = kf[2 * a], hi = kf[2 * a + 1];
    ko[2 * a] = (lo == lo) ? ord32(lo) : 0x7fffffff;
    ko[2 * a + 1] = (hi == hi) ? ord32(hi) : (int32_t)0x80000000;
  }
}

// Inverse of ord32 (NaN images excluded).
__device__ __forceinline__ float from_ord32(int32_t k) {
  const uint32_t u = k >= 0 ? (uint32_t)k : (0x80000000u | (uint32_t)(-(int64_t)k));
  return __uint_as_float(u);
}

// The ordered images q of one element's outward fp32 slabs in the frame of
// origin oq, moved to the frame of origin ot and still outward.  With pq, pt
// the origins' fp64 projections, the exact slab x has x - pq >= f (lower
// slab), so x - pt >= f + (pq - pt).  (s, e) is the exact TwoSum of
// pq - pt; (f + s) + e, rounded to nearest in fp64, errs by far less than
// half an fp32 ulp of the result, and moving its fp32 rounding one ulp out
// keeps the bound.  Upper slabs likewise.  NaN images (INT32_MAX lower,
// INT32_MIN upper) stay as they are.  (ox, oy, oz) and (tx, ty, tz) are the
// two origins.
template <int A = 0>
__device__ __forceinline__ void reframe32(int32_t* q, double ox, double oy, double oz, double tx, double ty,
                                          double tz) {
  // one axis at a time, so that only this axis's projections are live
  const double pq = project<A>(ox, oy, oz);
  const double pt = project<A>(tx, ty, tz);
  const double s = pq - pt;
  const double bb = s - pq;
  const double e = (pq - (s - bb)) + (-pt - bb);
  const int32_t klo = q[2 * A], khi = q[2 * A + 1];
  if (klo != 0x7fffffff) {
    const float f = round_down_1ulp(((double)from_ord32(klo) + s) + e);
    q[2 * A] = (f == f) ? ord32(f) : 0x7fffffff;
  }
  if (khi != (int32_t)0x80000000) {
    const float f = round_up_1ulp(((double)from_ord32(khi) + s) + e);
    q[2 * A + 1] = (f == f) ? ord32(f) : (int32_t)0x80000000;
  }
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (A + 1 < kAxes) reframe32<A + 1>(q, ox, oy, oz, tx, ty, tz);
}

__device__ __forceinline__ void to_fp32_outward(const double* kd, const double* po, float* kf) {
#pragma unroll
  for (int a = 0; a < kAxes; ++a) {
    kf[2 * a] = round_down_1ulp(kd[2 * a] - po[a]);
    kf[2 * a + 1] = round_up_1ulp(kd[2 * a + 1] - po[a]);
  }
}


__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

}  // namespace dbvh
