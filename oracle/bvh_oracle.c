/*
 * bvh_oracle.c -- CPU restatement of distBVH's single-rank collision path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker and the timed CPU baseline.  The
 * product (distbvh_amd/) never links or calls this file.
 *
 * Floating-point semantics: compiled with -O3 -ffp-contract=off and without
 * -ffast-math, so every expression is evaluated exactly in the written order
 * (the reference's own release flags add -ffast-math, CMakeLists.txt:65; the
 * canonical semantics declared for this build are the strict ones).
 */
#include "bvh_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* k-DOP normals: dop_26<double>::normals(), kdop.hpp:657-694.  Only the first
 * 13 (positive) normals are used by from_vertices / from_sphere
 * (kdop.hpp:249, 279). */
#define R2 0.7071067811865475
#define R3 0.5773502691896258
static const double k_normals[26][3] = {
    {1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.},
    /* corners */
    {R3, R3, R3}, {R3, -R3, R3}, {R3, R3, -R3}, {R3, -R3, -R3},
    /* edges */
    {R2, R2, 0.0}, {R2, 0.0, R2}, {0.0, R2, R2}, {R2, -R2, 0.0}, {R2, 0.0, -R2}, {0.0, R2, -R2},
    /* negative cardinal */
    {-1., 0., 0.}, {0., -1., 0.}, {0., 0., -1.},
    /* negative corners */
    {-R3, -R3, -R3}, {-R3, R3, -R3}, {-R3, -R3, R3}, {-R3, R3, R3},
    /* negative edges */
    {-R2, -R2, 0.0}, {-R2, 0.0, -R2}, {0.0, -R2, -R2}, {-R2, R2, 0.0}, {-R2, 0.0, R2}, {0.0, -R2, R2}};

double orc_normal(int axis, int comp) { return k_normals[axis][comp]; }

/* kdop_base::project, kdop.hpp:223-227 / 460-465:
 * v0*a0 + v1*a1 + v2*a2, evaluated left to right. */
double orc_project(const double *v, int axis) {
  const double *a = k_normals[axis];
  return v[0] * a[0] + v[1] * a[1] + v[2] * a[2];
}

/* std::min / std::max selects (kdop.hpp:258-259) */
static inline double std_min(double a, double b) { return (b < a) ? b : a; }
static inline double std_max(double a, double b) { return (a < b) ? b : a; }
/* m::min / m::max selects (math/common.hpp:86-96) */
static inline double m_min(double a, double b) { return (a < b) ? a : b; }
static inline double m_max(double a, double b) { return (a > b) ? a : b; }

/* default extent: min = DBL_MAX, max = lowest (kdop.hpp:59-60) */
void orc_kdop_empty(double *k) {
  for (int a = 0; a < ORC_NAXIS; ++a) {
    k[2 * a] = DBL_MAX;
    k[2 * a + 1] = -DBL_MAX;
  }
}

/* kdop_base::from_vertices, kdop.hpp:238-264 */
void orc_kdop_from_vertices(const double *v, int nv, double eps, double *k) {
  if (nv <= 0) {
    orc_kdop_empty(k);
    return;
  }
  for (int i = 0; i < ORC_NAXIS; ++i) {
    double proj = orc_project(v, i);
    double mn = proj - eps;
    double mx = proj + eps;
    for (int j = 1; j < nv; ++j) {
      proj = orc_project(v + 3 * j, i);
      mn = std_min(mn, proj - eps);
      mx = std_max(mx, proj + eps);
    }
    k[2 * i] = mn;
    k[2 * i + 1] = mx;
  }
}

/* kdop_base::from_sphere, kdop.hpp:273-288 */
void orc_kdop_from_sphere(const double *c, double r, double *k) {
  for (int i = 0; i < ORC_NAXIS; ++i) {
    double center = orc_project(c, i);
    k[2 * i] = center - r;
    k[2 * i + 1] = center + r;
  }
}

/* overlap(extent), kdop.hpp:120-125; overlap(kdop), kdop.hpp:538-549 */
int orc_overlap(const double *a, const double *b) {
  for (int i = 0; i < ORC_NAXIS; ++i) {
    if (!(a[2 * i] < b[2 * i + 1] && b[2 * i] < a[2 * i + 1])) return 0;
  }
  return 1;
}

/* kdop_base::union_with, kdop.hpp:317-325 (m::min / m::max) */
void orc_kdop_union(double *dst, const double *src) {
  for (int i = 0; i < ORC_NAXIS; ++i) {
    dst[2 * i] = m_min(dst[2 * i], src[2 * i]);
    dst[2 * i + 1] = m_max(dst[2 * i + 1], src[2 * i + 1]);
  }
}

/* kdop_base::longest_axis, kdop.hpp:354-362: std::max_element by length
 * (first maximum wins). */
int orc_longest_axis(const double *k) {
  int best = 0;
  double bl = k[1] - k[0];
  for (int i = 1; i < ORC_NAXIS; ++i) {
    double l = k[2 * i + 1] - k[2 * i];
    if (bl < l) {
      best = i;
      bl = l;
    }
  }
  return best;
}

/* kdop_base(first, last) / from_kdops, kdop.hpp:174-200 */
static void kdop_from_range(const double *kdops, size_t b, size_t e, double *out) {
  if (e <= b) {
    orc_kdop_empty(out);
    return;
  }
  for (int a = 0; a < ORC_NAXIS; ++a) {
    double mn = kdops[b * ORC_KDOP + 2 * a];
    double mx = kdops[b * ORC_KDOP + 2 * a + 1];
    for (size_t i = b + 1; i < e; ++i) {
      mn = m_min(kdops[i * ORC_KDOP + 2 * a], mn);
      mx = m_max(kdops[i * ORC_KDOP + 2 * a + 1], mx);
    }
    out[2 * a] = mn;
    out[2 * a + 1] = mx;
  }
}

/* Element k-DOPs for polytope elements (e.g. the triangle / hex test element,
 * tests/TestCommon.hpp:67-73 computes kdop = from_vertices(vertices)) and the
 * element centroid = accumulate(vertices, zeros) / count
 * (tests/TestCommon.hpp:62-65). */
void orc_elements_poly(const double *verts, int nv, size_t n, double eps, double *kdops,
                       double *centroids) {
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < n; ++e) {
    double v[8 * 3];
    double s[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < nv; ++j)
      for (int c = 0; c < 3; ++c) {
        v[3 * j + c] = verts[(size_t)(3 * j + c) * n + e];
        s[c] = s[c] + v[3 * j + c];
      }
    orc_kdop_from_vertices(v, nv, eps, kdops + e * ORC_KDOP);
    for (int c = 0; c < 3; ++c) centroids[e * 3 + c] = s[c] / (double)nv;
  }
}

void orc_elements_sphere(const double *xyzr, size_t n, double *kdops, double *centroids) {
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < n; ++e) {
    double c[3] = {xyzr[e], xyzr[n + e], xyzr[2 * n + e]};
    orc_kdop_from_sphere(c, xyzr[3 * n + e], kdops + e * ORC_KDOP);
    for (int k = 0; k < 3; ++k) centroids[e * 3 + k] = c[k];
  }
}

/* compute_bounds(elements, min_inv_diag_bounds), contact_entity.hpp:136-148
 * with min_diag_bounds_union (contact_entity.hpp:57-102).  The reference's
 * join drops the partner's max (contact_entity.hpp:90); the serial result, and
 * this restatement, take the true max.  width = max - min, inv = 1.0 / width. */
void orc_compute_bounds(const double *kdops, size_t n, double *min3, double *inv_diag3) {
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (size_t e = 0; e < n; ++e) {
    for (int i = 0; i < 3; ++i) {
      double cmin = kdops[e * ORC_KDOP + 2 * i];
      double cmax = kdops[e * ORC_KDOP + 2 * i + 1];
      mn[i] = (cmin < mn[i]) ? cmin : mn[i];
      mx[i] = (mx[i] < cmax) ? cmax : mx[i];
    }
  }
  for (int i = 0; i < 3; ++i) {
    double width = mx[i] - mn[i];
    min3[i] = mn[i];
    inv_diag3[i] = 1.0 / width;
  }
}

/* detail::expand32, hash.hpp:50-58 */
uint32_t orc_expand32(uint32_t v) {
  v = (v * 0x00010001u) & 0xff0000ffu;
  v = (v * 0x00000101u) & 0x0f00f00fu;
  v = (v * 0x00000011u) & 0xc30c30c3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

/* detail::expand64, hash.hpp:72-81 */
uint64_t orc_expand64(uint64_t v) {
  v = (v * 0x1000001u) & 0xfff000000fffull;
  v = (v * 0x1001u) & 0xfc003f000fc003full;
  v = (v * 0x41u) & 0x01c0e070381c0e07ull;
  v = (v * 0x11u) & 0x10c86432190c8643ull;
  v = (v * 0x5u) & 0x1249249249249249ull;
  return v;
}

/* Portable parallel bit deposit (what _pdep_u64 computes, hash.hpp:84-91);
 * used by the tests as an independent check of expand64. */
uint64_t orc_pdep64(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  for (uint64_t bb = 1; mask; bb <<= 1) {
    uint64_t low = mask & (~mask + 1);
    if (v & bb) r |= low;
    mask &= mask - 1;
  }
  return r;
}

/* morton(u32,u32,u32), hash.hpp:109-113 */
uint32_t orc_morton32(uint32_t x, uint32_t y, uint32_t z) {
  return (orc_expand32(z) << 2) + (orc_expand32(y) << 1) + orc_expand32(x);
}

/* morton(u64,u64,u64) -> detail::morton64, hash.hpp:100-103, 121-130 */
uint64_t orc_morton64(uint64_t x, uint64_t y, uint64_t z) {
  return (orc_expand64(z) << 2u) | (orc_expand64(y) << 1u) | orc_expand64(x);
}

/* static_cast<uintN_t>(double) as the reference's x86-64 build evaluates it:
 * truncation toward zero into a signed 64-bit integer, reinterpretation, and
 * (by the caller) the mask.  NaN, +-inf and |s| >= 2^63 give 0 (what the
 * cvttsd2si sequence leaves after the mask).  The GPU uses the same canonical
 * definition, so a degenerate axis (0 * inf = NaN, hash.hpp:151-155 /
 * 182-186) quantizes to 0. */
uint64_t orc_float_to_uint(double s) {
  if (!(s > -9223372036854775808.0 && s < 9223372036854775808.0)) return 0;
  return (uint64_t)(int64_t)s;
}

/* quantize32(p, min, inv_diag), hash.hpp:146-167 */
void orc_quantize32(const double *p, const double *min3, const double *inv3, uint32_t *q) {
  for (int i = 0; i < 3; ++i) {
    double norm = (p[i] - min3[i]) * inv3[i];
    double scaled = norm * 1024.0;
    q[i] = (uint32_t)orc_float_to_uint(scaled) & 0x3ffu;
  }
}

/* quantize64(p, min, inv_diag), hash.hpp:177-198 */
void orc_quantize64(const double *p, const double *min3, const double *inv3, uint64_t *q) {
  for (int i = 0; i < 3; ++i) {
    double norm = (p[i] - min3[i]) * inv3[i];
    double scaled = norm * 2097152.0;
    q[i] = orc_float_to_uint(scaled) & 0x1fffffull;
  }
}

/* detail::morton_impl, contact_entity.hpp:104-118 (element centroid) */
void orc_morton_codes32(const double *centroids, size_t n, const double *min3, const double *inv3,
                        uint32_t *out) {
  for (size_t e = 0; e < n; ++e) {
    uint32_t q[3];
    orc_quantize32(centroids + 3 * e, min3, inv3, q);
    out[e] = orc_morton32(q[0], q[1], q[2]);
  }
}

void orc_morton_codes64(const double *centroids, size_t n, const double *min3, const double *inv3,
                        uint64_t *out) {
  for (size_t e = 0; e < n; ++e) {
    uint64_t q[3];
    orc_quantize64(centroids + 3 * e, min3, inv3, q);
    out[e] = orc_morton64(q[0], q[1], q[2]);
  }
}

/* prefix_sum, util/prefix_sum.hpp:41-54 (exclusive) */
void orc_prefix_sum_u32(uint32_t *v, size_t n) {
  uint32_t x = 0;
  for (size_t i = 0; i < n; ++i) {
    uint32_t val = v[i];
    v[i] = x;
    x += val;
  }
}

/* radix_sorter<T, I>::operator() / step, util/sort.hpp:100-136: one stable
 * split pass per bit, num_bits = 8*sizeof(T) passes. */
#define DEFINE_RADIX(NAME, T)                                                       \
  void NAME(T *keys, uint32_t *idx, size_t n) {                                     \
    if (n == 0) return;                                                             \
    T *ks = (T *)malloc(n * sizeof(T));                                             \
    uint32_t *is = (uint32_t *)malloc(n * sizeof(uint32_t));                        \
    uint32_t *bits = (uint32_t *)malloc(n * sizeof(uint32_t));                      \
    uint32_t *scan = (uint32_t *)malloc(n * sizeof(uint32_t));                      \
    T *hin = keys, *hout = ks;                                                      \
    uint32_t *iin = idx, *iout = is;                                                \
    const unsigned nb = 8 * sizeof(T);                                              \
    for (unsigned s = 0; s < nb; ++s) {                                             \
      for (size_t i = 0; i < n; ++i) {                                              \
        T h = hin[i] >> s;                                                          \
        bits[i] = (uint32_t)(~h & 1u);                                              \
        scan[i] = bits[i];                                                          \
      }                                                                             \
      orc_prefix_sum_u32(scan, n);                                                  \
      const size_t total = (size_t)scan[n - 1] + bits[n - 1];                       \
      for (size_t i = 0; i < n; ++i) {                                              \
        size_t t = i - scan[i] + total;                                             \
        size_t ni = bits[i] ? scan[i] : t;                                          \
        iout[ni] = iin[i];                                                          \
        hout[ni] = hin[i];                                                          \
      }                                                                             \
      T *tk = hin; hin = hout; hout = tk;                                           \
      uint32_t *ti = iin; iin = iout; iout = ti;                                    \
    }                                                                               \
    /* an even pass count leaves the result in the caller's buffers */             \
    free(ks); free(is); free(bits); free(scan);                                     \
  }
DEFINE_RADIX(orc_radix_sort32, uint32_t)
DEFINE_RADIX(orc_radix_sort64, uint64_t)

/* clz, util/bits.hpp:88-92 (Kokkos::countl_zero; 0 -> width) */
int orc_clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }
int orc_clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

/* morton_cluster::operator(), split/cluster.hpp:89-168, generalized to the
 * code width: depth(i) = clz(h[i] ^ h[i+1]) or UINT_MAX when equal; stable
 * sort of (depth, i); the first `nsplits` entries, re-sorted by position. */
static void cluster_splits_depths(const uint32_t *depths, size_t n, size_t nsplits, uint64_t *splits) {
  if (n == 0) {
    for (size_t k = 0; k < nsplits; ++k) splits[k] = 0;
    return;
  }
  const size_t max_idx = n - 1;
  if (nsplits > max_idx)
    for (size_t k = 0; k < nsplits; ++k) splits[k] = max_idx;
  if (n < 2) return;
  size_t m = n - 1;
  uint32_t *d = (uint32_t *)malloc(m * sizeof(uint32_t));
  uint32_t *di = (uint32_t *)malloc(m * sizeof(uint32_t));
  uint32_t *re = (uint32_t *)malloc(m * sizeof(uint32_t));
  memcpy(d, depths, m * sizeof(uint32_t));
  for (size_t i = 0; i < m; ++i) di[i] = (uint32_t)i;
  orc_radix_sort32(d, di, m); /* m_depth_sorter, split/cluster.hpp:153 */
  for (size_t i = 0; i < m; ++i) re[di[i]] = (i < nsplits) ? 1u : 0u;
  orc_prefix_sum_u32(re, m);
  for (size_t i = 0; i < m; ++i)
    if (i < nsplits) splits[re[di[i]]] = di[i];
  free(d); free(di); free(re);
}

void orc_cluster_splits32(const uint32_t *h, size_t n, size_t nsplits, uint64_t *splits) {
  uint32_t *dep = (uint32_t *)malloc((n > 1 ? n - 1 : 1) * sizeof(uint32_t));
  for (size_t i = 0; i + 1 < n; ++i) {
    uint32_t mask = h[i] ^ h[i + 1];
    dep[i] = mask ? (uint32_t)orc_clz32(mask) : 0xffffffffu;
  }
  cluster_splits_depths(dep, n, nsplits, splits);
  free(dep);
}

void orc_cluster_splits64(const uint64_t *h, size_t n, size_t nsplits, uint64_t *splits) {
  uint32_t *dep = (uint32_t *)malloc((n > 1 ? n - 1 : 1) * sizeof(uint32_t));
  for (size_t i = 0; i + 1 < n; ++i) {
    uint64_t mask = h[i] ^ h[i + 1];
    dep[i] = mask ? (uint32_t)orc_clz64(mask) : 0xffffffffu;
  }
  cluster_splits_depths(dep, n, nsplits, splits);
  free(dep);
}

/* set_entity_data_impl, collision_object.cpp:134-141: patch k covers
 * [k == 0 ? 0 : splits[k-1], k == od-1 ? n : splits[k]); patch<>::set_elements
 * / compute_kdops_and_centroid, patch.hpp:80-117: kdop = from_kdops, centroid
 * = accumulate(centroids, zeros) / size; an empty patch keeps the default
 * kdop and a zero centroid. */
void orc_patches(const uint64_t *splits, size_t od, size_t n, const double *sk, const double *sc,
                 uint64_t *begin, uint64_t *end, double *pk, double *pc) {
  for (size_t k = 0; k < od; ++k) {
    size_t b = (k == 0) ? 0 : splits[k - 1];
    size_t e = (k == od - 1) ? n : splits[k];
    begin[k] = b;
    end[k] = e;
    if (e > b) {
      kdop_from_range(sk, b, e, pk + k * ORC_KDOP);
      double s[3] = {0.0, 0.0, 0.0};
      for (size_t i = b; i < e; ++i)
        for (int c = 0; c < 3; ++c) s[c] = s[c] + sc[i * 3 + c];
      for (int c = 0; c < 3; ++c) pc[k * 3 + c] = s[c] / (double)(e - b);
    } else {
      orc_kdop_empty(pk + k * ORC_KDOP);
      for (int c = 0; c < 3; ++c) pc[k * 3 + c] = 0.0;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* top-down tree: top_down_builder::builder::build, bvh_build.hpp:80-133 with
 * split_in_place<split::mean> (split/split.hpp:54-75, split/mean.hpp:49-65). */
typedef struct {
  orc_tree *t;
  size_t cap;
  size_t leaf_size;
} tbuild;

static size_t tree_new_node(tbuild *b, const double *kd) {
  orc_tree *t = b->t;
  if (t->nnodes == b->cap) {
    b->cap = b->cap ? 2 * b->cap : 64;
    t->kdop = (double *)realloc(t->kdop, b->cap * ORC_KDOP * sizeof(double));
    t->left = (int64_t *)realloc(t->left, b->cap * sizeof(int64_t));
    t->right = (int64_t *)realloc(t->right, b->cap * sizeof(int64_t));
    t->first = (uint64_t *)realloc(t->first, b->cap * sizeof(uint64_t));
    t->last = (uint64_t *)realloc(t->last, b->cap * sizeof(uint64_t));
  }
  size_t id = t->nnodes++;
  memcpy(t->kdop + id * ORC_KDOP, kd, ORC_KDOP * sizeof(double));
  t->left[id] = -1;
  t->right[id] = -1;
  t->first[id] = t->last[id] = 0;
  return id;
}

static void leaf_swap(orc_tree *t, size_t a, size_t b) {
  double tk[ORC_KDOP], tc[3];
  memcpy(tk, t->leaf_kdop + a * ORC_KDOP, sizeof(tk));
  memcpy(t->leaf_kdop + a * ORC_KDOP, t->leaf_kdop + b * ORC_KDOP, sizeof(tk));
  memcpy(t->leaf_kdop + b * ORC_KDOP, tk, sizeof(tk));
  memcpy(tc, t->leaf_centroid + a * 3, sizeof(tc));
  memcpy(t->leaf_centroid + a * 3, t->leaf_centroid + b * 3, sizeof(tc));
  memcpy(t->leaf_centroid + b * 3, tc, sizeof(tc));
  uint64_t g = t->leaf_gid[a];
  t->leaf_gid[a] = t->leaf_gid[b];
  t->leaf_gid[b] = g;
}

/* std::partition as libstdc++ implements it for bidirectional iterators */
static size_t partition_mean(orc_tree *t, size_t first, size_t last, int axis, double split) {
#define PRED(i) (orc_project(t->leaf_centroid + 3 * (i), axis) < split)
  for (;;) {
    for (;;) {
      if (first == last) return first;
      if (PRED(first)) ++first;
      else break;
    }
    --last;
    for (;;) {
      if (first == last) return first;
      if (!PRED(last)) --last;
      else break;
    }
    leaf_swap(t, first, last);
    ++first;
  }
#undef PRED
}

static void tree_build_rec(tbuild *b, size_t node, size_t eb, size_t ee) {
  orc_tree *t = b->t;
  t->first[node] = eb;
  t->last[node] = ee;
  if (ee - eb <= b->leaf_size) return;
  int axis = orc_longest_axis(t->kdop + node * ORC_KDOP);
  /* split::mean::split_point: accumulate projections from 0 / size */
  double sum = 0.0;
  for (size_t i = eb; i < ee; ++i) sum = sum + orc_project(t->leaf_centroid + 3 * i, axis);
  double split_point = sum / (double)(ee - eb);
  size_t split = partition_mean(t, eb, ee, axis, split_point);
  if (split == eb || split == ee) split = eb + (ee - eb) / 2;
  double kd[ORC_KDOP];
  if (split != eb) {
    kdop_from_range(t->leaf_kdop, eb, split, kd);
    size_t l = tree_new_node(b, kd);
    t->left[node] = (int64_t)l;
    tree_build_rec(b, l, eb, split);
  }
  if (split != ee) {
    kdop_from_range(t->leaf_kdop, split, ee, kd);
    size_t r = tree_new_node(b, kd);
    t->right[node] = (int64_t)r;
    tree_build_rec(b, r, split, ee);
  }
}

/* rebuild_tree, tree_build.hpp:54-71 (leaf size is 1 there; parameterized) */
orc_tree *orc_tree_build_top_down(const double *kdops, const double *centroids, const uint64_t *gids,
                                  size_t n, size_t leaf_size) {
  orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
  t->nleaves = n;
  if (n == 0) return t;
  t->leaf_kdop = (double *)malloc(n * ORC_KDOP * sizeof(double));
  t->leaf_centroid = (double *)malloc(n * 3 * sizeof(double));
  t->leaf_gid = (uint64_t *)malloc(n * sizeof(uint64_t));
  memcpy(t->leaf_kdop, kdops, n * ORC_KDOP * sizeof(double));
  memcpy(t->leaf_centroid, centroids, n * 3 * sizeof(double));
  memcpy(t->leaf_gid, gids, n * sizeof(uint64_t));
  tbuild b = {t, 0, leaf_size ? leaf_size : 1};
  double kd[ORC_KDOP];
  kdop_from_range(t->leaf_kdop, 0, n, kd);
  size_t root = tree_new_node(&b, kd);
  tree_build_rec(&b, root, 0, n);
  return t;
}

void orc_tree_free(orc_tree *t) {
  if (!t) return;
  free(t->kdop); free(t->left); free(t->right); free(t->first); free(t->last);
  free(t->leaf_kdop); free(t->leaf_centroid); free(t->leaf_gid);
  free(t);
}

orc_pairs *orc_pairs_new(void) { return (orc_pairs *)calloc(1, sizeof(orc_pairs)); }
void orc_pairs_free(orc_pairs *p) {
  if (!p) return;
  free(p->data);
  free(p);
}
void orc_pairs_push(orc_pairs *p, uint64_t a, uint64_t b) {
  if (p->n == p->cap) {
    p->cap = p->cap ? 2 * p->cap : 256;
    p->data = (uint64_t *)realloc(p->data, p->cap * 2 * sizeof(uint64_t));
  }
  p->data[2 * p->n] = a;
  p->data[2 * p->n + 1] = b;
  p->n++;
}
static void pairs_append(orc_pairs *dst, const orc_pairs *src) {
  for (size_t i = 0; i < src->n; ++i) orc_pairs_push(dst, src->data[2 * i], src->data[2 * i + 1]);
}

static int is_leaf(const orc_tree *t, size_t node) { return t->left[node] < 0 && t->right[node] < 0; }

/* detail::query_node_impl / query_tree, collision_query.hpp:166-213 */
static void query_rec(const orc_tree *t, size_t node, const double *kd, uint64_t gid, orc_pairs *out) {
  if (!orc_overlap(kd, t->kdop + node * ORC_KDOP)) return;
  if (is_leaf(t, node)) {
    for (size_t i = t->first[node]; i < t->last[node]; ++i) orc_pairs_push(out, gid, t->leaf_gid[i]);
  } else {
    if (t->left[node] >= 0) query_rec(t, (size_t)t->left[node], kd, gid, out);
    if (t->right[node] >= 0) query_rec(t, (size_t)t->right[node], kd, gid, out);
  }
}

void orc_query_tree(const orc_tree *t, const double *kdop, uint64_t gid, orc_pairs *out) {
  if (t->nnodes == 0) return;
  query_rec(t, 0, kdop, gid, out);
}

/* query_tree followed by the exact element test of the narrowphase: a leaf
 * holding several elements yields them all as candidates
 * (collision_query.hpp:178-181); each candidate is kept only if its own
 * k-DOP overlaps (overlap(ak, bk), narrowphase/kokkos.hpp:221). */
static void query_rec_exact(const orc_tree *t, size_t node, const double *kd, uint64_t gid, orc_pairs *out) {
  if (!orc_overlap(kd, t->kdop + node * ORC_KDOP)) return;
  if (is_leaf(t, node)) {
    for (size_t i = t->first[node]; i < t->last[node]; ++i)
      if (orc_overlap(kd, t->leaf_kdop + i * ORC_KDOP)) orc_pairs_push(out, gid, t->leaf_gid[i]);
  } else {
    if (t->left[node] >= 0) query_rec_exact(t, (size_t)t->left[node], kd, gid, out);
    if (t->right[node] >= 0) query_rec_exact(t, (size_t)t->right[node], kd, gid, out);
  }
}

/* detail::get_overlapping_indices, collision_query.hpp:293-329 */
static void overlap_rec(const orc_tree *l, int64_t ln, const orc_tree *r, int64_t rn, int as_written,
                        orc_pairs *out) {
  if (ln < 0 || rn < 0) return;
  if (!orc_overlap(l->kdop + ln * ORC_KDOP, r->kdop + rn * ORC_KDOP)) return;
  if (is_leaf(l, (size_t)ln)) {
    if (is_leaf(r, (size_t)rn)) {
      for (size_t i = l->first[ln]; i < l->last[ln]; ++i)
        for (size_t j = r->first[rn]; j < r->last[rn]; ++j) {
          if (as_written) /* (_right_leafs[i], _left_leafs[j]), lines 310-312 */
            orc_pairs_push(out, r->leaf_gid[i], l->leaf_gid[j]);
          else
            orc_pairs_push(out, l->leaf_gid[i], r->leaf_gid[j]);
        }
    } else {
      overlap_rec(l, ln, r, r->left[rn], as_written, out);
      overlap_rec(l, ln, r, r->right[rn], as_written, out);
    }
  } else if (is_leaf(r, (size_t)rn)) {
    overlap_rec(l, l->left[ln], r, rn, as_written, out);
    overlap_rec(l, l->right[ln], r, rn, as_written, out);
  } else {
    overlap_rec(l, l->left[ln], r, r->left[rn], as_written, out);
    overlap_rec(l, l->left[ln], r, r->right[rn], as_written, out);
    overlap_rec(l, l->right[ln], r, r->left[rn], as_written, out);
    overlap_rec(l, l->right[ln], r, r->right[rn], as_written, out);
  }
}

/* detail::get_self_colliding_indices, collision_query.hpp:331-357 */
static void self_rec(const orc_tree *t, int64_t node, orc_pairs *out) {
  if (node < 0) return;
  if (is_leaf(t, (size_t)node)) {
    if (t->last[node] > t->first[node]) {
      for (size_t i = t->first[node]; i < t->last[node] - 1; ++i)
        for (size_t j = i; j < t->last[node]; ++j) orc_pairs_push(out, t->leaf_gid[i], t->leaf_gid[j]);
    }
  } else {
    self_rec(t, t->left[node], out);
    self_rec(t, t->right[node], out);
    if (t->left[node] >= 0 && t->right[node] >= 0)
      overlap_rec(t, t->left[node], t, t->right[node], 1, out);
  }
}

void orc_self_collision_set(const orc_tree *t, orc_pairs *out) {
  if (t->nnodes) self_rec(t, 0, out);
}

void orc_potential_collision_set(const orc_tree *l, const orc_tree *r, int as_written, orc_pairs *out) {
  if (l->nnodes && r->nnodes) overlap_rec(l, 0, r, 0, as_written, out);
}

/* max_variant_axis, narrowphase/kokkos.hpp:119-161 */
int orc_max_variant_axis(const double *ca, size_t na, const double *cb, size_t nb) {
  double s[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
  for (size_t i = 0; i < na; ++i)
    for (int j = 0; j < 3; ++j) {
      s[j] += ca[3 * i + j];
      s2[j] += ca[3 * i + j] * ca[3 * i + j];
    }
  for (size_t i = 0; i < nb; ++i)
    for (int j = 0; j < 3; ++j) {
      s[j] += cb[3 * i + j];
      s2[j] += cb[3 * i + j] * cb[3 * i + j];
    }
  double var[3];
  for (int i = 0; i < 3; ++i) var[i] = s2[i] - s[i] * s[i] / (double)(na + nb);
  int axis = 0;
  if (var[1] > var[0]) axis = 1;
  if (var[2] > var[axis]) axis = 2;
  return axis;
}

/* sort_and_sweep_local, narrowphase/kokkos.hpp:164-248.  The quantized
 * uint32 values use the same canonical float->integer conversion as the
 * Morton quantizer (low 32 bits of the truncated int64). */
void orc_sort_and_sweep(const double *pa, const double *ka, const uint64_t *ga, size_t na,
                        const double *pb, const double *kb, const uint64_t *gb, size_t nb, int axis,
                        orc_pairs *out) {
  double gmin = std_min(pa[2 * axis], pb[2 * axis]);
  double gmax = std_max(pa[2 * axis + 1], pb[2 * axis + 1]);
  double conv = (double)UINT32_MAX / (gmax - gmin);
  uint32_t *bmin = (uint32_t *)malloc((nb ? nb : 1) * sizeof(uint32_t));
  uint32_t *bi = (uint32_t *)malloc((nb ? nb : 1) * sizeof(uint32_t));
  for (size_t i = 0; i < nb; ++i) {
    bi[i] = (uint32_t)i;
    bmin[i] = (uint32_t)orc_float_to_uint((kb[i * ORC_KDOP + 2 * axis] - gmin) * conv);
  }
  orc_radix_sort32(bmin, bi, nb);
  for (size_t i = 0; i < na; ++i) {
    const double *ak = ka + i * ORC_KDOP;
    for (size_t j = 0; j < nb; ++j) {
      uint32_t amin = (uint32_t)orc_float_to_uint((ak[2 * axis] - gmin) * conv);
      uint32_t amax = (uint32_t)orc_float_to_uint((ak[2 * axis + 1] - gmin) * conv);
      const double *bk = kb + (size_t)bi[j] * ORC_KDOP;
      uint32_t bm = bmin[j];
      double bmax = (bk[2 * axis + 1] - gmin) * conv;
      if (bmax < (double)amin) continue;
      if (bm > amax) break;
      if (orc_overlap(ak, bk)) orc_pairs_push(out, ga[i], gb[bi[j]]);
    }
  }
  free(bmin);
  free(bi);
}

/* brute force all pairs: the partition-invariant master pair set */
orc_pairs *orc_brute_pairs(const double *ka, size_t na, const double *kb, size_t nb, int nthreads) {
  orc_pairs *res = orc_pairs_new();
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
  int nt = nthreads;
#else
  int nt = 1;
#endif
  orc_pairs **loc = (orc_pairs **)calloc((size_t)nt, sizeof(orc_pairs *));
  for (int t = 0; t < nt; ++t) loc[t] = orc_pairs_new();
  /* contiguous row blocks per thread keep the (i, j) order after concatenation */
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    size_t chunk = (na + (size_t)nt - 1) / (size_t)nt;
    size_t lo = (size_t)tid * chunk, hi = lo + chunk < na ? lo + chunk : na;
    for (size_t i = lo; i < hi; ++i)
      for (size_t j = 0; j < nb; ++j)
        if (orc_overlap(ka + i * ORC_KDOP, kb + j * ORC_KDOP)) orc_pairs_push(loc[tid], i, j);
  }
  for (int t = 0; t < nt; ++t) {
    pairs_append(res, loc[t]);
    orc_pairs_free(loc[t]);
  }
  free(loc);
  return res;
}

orc_pairs *orc_brute_self(const double *k, size_t n, int nthreads) {
  orc_pairs *res = orc_pairs_new();
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
  int nt = nthreads;
#else
  int nt = 1;
#endif
  orc_pairs **loc = (orc_pairs **)calloc((size_t)nt, sizeof(orc_pairs *));
  for (int t = 0; t < nt; ++t) loc[t] = orc_pairs_new();
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    size_t chunk = (n + (size_t)nt - 1) / (size_t)nt;
    size_t lo = (size_t)tid * chunk, hi = lo + chunk < n ? lo + chunk : n;
    for (size_t i = lo; i < hi; ++i)
      for (size_t j = i + 1; j < n; ++j)
        if (orc_overlap(k + i * ORC_KDOP, k + j * ORC_KDOP)) orc_pairs_push(loc[tid], i, j);
  }
  for (int t = 0; t < nt; ++t) {
    pairs_append(res, loc[t]);
    orc_pairs_free(loc[t]);
  }
  free(loc);
  return res;
}

/* ------------------------------------------------------------------------ */
double orc_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return (x > y) - (x < y);
}
void orc_sort_u64(uint64_t *v, size_t n) { qsort(v, n, sizeof(uint64_t), cmp_u64); }
void orc_free(void *p) { free(p); }

/* One object through set_entity_data(clustering) + set_entity_data_impl:
 * bounds -> Morton -> 1-bit LSD radix sort -> cluster splits -> snapshots in
 * sorted order -> patches.  The snapshot order uses the permutation as a
 * gather map for both the patch k-DOPs and the narrowphase elements (the
 * reference scatters snapshots with it, collision_object.hpp:239-242, and
 * gathers the narrowphase bytes, collision_object/impl.hpp:112-118). */
void orc_object_build(orc_object *o, const orc_pipeline_cfg *cfg, orc_pipeline_times *tm) {
  const size_t n = o->n, od = cfg->od;
  double t0 = orc_now();
  double mn[3], inv[3];
  orc_compute_bounds(o->kdops, n, mn, inv);
  double t1 = orc_now();
  for (size_t i = 0; i < n; ++i) o->perm[i] = (uint32_t)i;
  if (cfg->split == ORC_SPLIT_GEOM_AXIS || cfg->split == ORC_SPLIT_ML_GEOM_AXIS) {
    /* set_entity_data_geom_axis / _ml_geom_axis (collision_object.hpp:133-162):
     * depth = bit_log2(od); od must be a power of two (collision_object.cpp:
     * 122-124 asserts splits + 1 == od) */
    int depth = 0;
    while (((size_t)1 << (depth + 1)) <= od) ++depth;
    if (cfg->split == ORC_SPLIT_GEOM_AXIS) orc_split_geom_axis(o->kdops, o->centroids, n, depth, o->perm, o->splits);
    else orc_split_ml_geom_axis(o->kdops, o->centroids, n, depth, o->perm, o->splits);
    for (size_t i = 0; i < n; ++i) o->codes[i] = 0;
    if (tm) tm->t_cluster += orc_now() - t1;
  } else if (cfg->morton_bits == 30) {
    uint32_t *c32 = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    orc_morton_codes32(o->centroids, n, mn, inv, c32);
    double t2 = orc_now();
    orc_radix_sort32(c32, o->perm, n);
    double t3 = orc_now();
    orc_cluster_splits32(c32, n, od - 1, o->splits);
    double t4 = orc_now();
    for (size_t i = 0; i < n; ++i) o->codes[i] = c32[i];
    free(c32);
    if (tm) { tm->t_morton += t2 - t1; tm->t_sort += t3 - t2; tm->t_cluster += t4 - t3; }
  } else {
    orc_morton_codes64(o->centroids, n, mn, inv, o->codes);
    double t2 = orc_now();
    orc_radix_sort64(o->codes, o->perm, n);
    double t3 = orc_now();
    orc_cluster_splits64(o->codes, n, od - 1, o->splits);
    double t4 = orc_now();
    if (tm) { tm->t_morton += t2 - t1; tm->t_sort += t3 - t2; tm->t_cluster += t4 - t3; }
  }
  double t5 = orc_now();
  double *sk = (double *)malloc((n ? n : 1) * ORC_KDOP * sizeof(double));
  double *sc = (double *)malloc((n ? n : 1) * 3 * sizeof(double));
  for (size_t p = 0; p < n; ++p) {
    memcpy(sk + p * ORC_KDOP, o->kdops + (size_t)o->perm[p] * ORC_KDOP, ORC_KDOP * sizeof(double));
    memcpy(sc + p * 3, o->centroids + (size_t)o->perm[p] * 3, 3 * sizeof(double));
  }
  orc_patches(o->splits, od, n, sk, sc, o->pbegin, o->pend, o->patch_kdops, o->patch_centroids);
  free(sk);
  free(sc);
  double t6 = orc_now();
  if (tm) { tm->t_bounds += t1 - t0; tm->t_patches += t6 - t5; }
}

/* ---- geom_axis / ml_geom_axis element splitters ---------------------------------- */

/* from_kdops over idx[b..e) (kdop.hpp:174-214: slab-wise m::min / m::max in
 * order) and axis::longest (split/axis.hpp:42-49, kdop.hpp:354-362). */
static int split_axis(const double *kdops, const uint32_t *idx, size_t b, size_t e) {
  double k[ORC_KDOP];
  orc_kdop_empty(k);
  for (size_t i = b; i < e; ++i) {
    const double *s = kdops + (size_t)idx[i] * ORC_KDOP;
    for (int a = 0; a < ORC_NAXIS; ++a) {
      k[2 * a] = m_min(k[2 * a], s[2 * a]);
      k[2 * a + 1] = m_max(k[2 * a + 1], s[2 * a + 1]);
    }
  }
  return orc_longest_axis(k);
}

/* libstdc++ std::partition (bidirectional iterators) of idx[b..e) by
 * project(centroid, axis) < split; returns the partition point. */
static size_t partition_idx(uint32_t *idx, const double *centroids, size_t first, size_t last, int axis,
                            double split) {
#define PRED(i) (orc_project(centroids + 3 * (size_t)idx[i], axis) < split)
  for (;;) {
    for (;;) {
      if (first == last) return first;
      if (PRED(first)) ++first;
      else break;
    }
    --last;
    for (;;) {
      if (first == last) return first;
      if (!PRED(last)) --last;
      else break;
    }
    uint32_t t = idx[first];
    idx[first] = idx[last];
    idx[last] = t;
    ++first;
  }
#undef PRED
}

typedef struct {
  const double *kdops, *centroids;
  size_t n;
  uint32_t *idx;
  uint64_t *splits;
  size_t nsplits;
  int axis;        /* geom_axis: one axis and split point for every call */
  double split;
} split_state;

/* split_permutations_recursive_impl (split/split.hpp:220-243) with
 * split_permutation (split/split.hpp:177-195): the axis and the split point
 * are taken over all elements at every level (the element span is never
 * narrowed there). */
static void geom_rec(split_state *st, int depth, size_t b, size_t e) {
  size_t sp = partition_idx(st->idx, st->centroids, b, e, st->axis, st->split);
  if (sp == b || sp == e) sp = b + (e - b) / 2;
  if (depth > 0) geom_rec(st, depth - 1, b, sp);
  st->splits[st->nsplits++] = sp;
  if (depth > 0) geom_rec(st, depth - 1, sp, e);
}

/* split_permutations<split::mean, axis::longest> (split/split.hpp:309-323):
 * idx = iota, then 2^depth - 1 split offsets in recursion order.  The split
 * point is split::mean over the elements in input order (split/mean.hpp:
 * 49-65: std::accumulate from 0, then / size). */
void orc_split_geom_axis(const double *kdops, const double *centroids, size_t n, int depth, uint32_t *idx,
                         uint64_t *splits) {
  for (size_t i = 0; i < n; ++i) idx[i] = (uint32_t)i;
  if (depth <= 0) return;
  split_state st = {kdops, centroids, n, idx, splits, 0, 0, 0.0};
  st.axis = split_axis(kdops, idx, 0, n);
  double sum = 0.0;
  for (size_t i = 0; i < n; ++i) sum = sum + orc_project(centroids + 3 * i, st.axis);
  st.split = sum / (double)n;
  geom_rec(&st, depth - 1, 0, n);
}

/* split_permutation_ml (split/split.hpp:121-158): fewer than 2 elements give
 * offset 0; the mean is split::mean's span overload (split/mean.hpp:67-79),
 * a parallel_reduce restated in element order (Kokkos::Serial); the
 * partition of the (element, index) pairs moves elements and indices
 * together; an empty side gives size / 2. */
static size_t split_ml(split_state *st, size_t b, size_t e, int axis) {
  if (e - b < 2) return 0;
  double sum = 0.0;
  for (size_t i = b; i < e; ++i) sum += orc_project(st->centroids + 3 * (size_t)st->idx[i], axis);
  const double split = sum / (double)(e - b);
  size_t sp = partition_idx(st->idx, st->centroids, b, e, axis, split);
  size_t delta = sp - b;
  if (sp == b || sp == e) delta = (e - b) / 2;
  return delta;
}

/* split_permutations_recursive_impl_ml (split/split.hpp:186-218): the axis
 * from the current range (axis 0 for an empty one). */
static void ml_rec(split_state *st, int depth, size_t b, size_t e) {
  int axis = 0;
  if (e > b) axis = split_axis(st->kdops, st->idx, b, e);
  const size_t sp = b + split_ml(st, b, e, axis);
  if (depth > 0) ml_rec(st, depth - 1, b, sp);
  st->splits[st->nsplits++] = sp;
  if (depth > 0) ml_rec(st, depth - 1, sp, e);
}

/* split_permutations_ml<split::mean, axis::longest> (split/split.hpp:285-307). */
void orc_split_ml_geom_axis(const double *kdops, const double *centroids, size_t n, int depth, uint32_t *idx,
                            uint64_t *splits) {
  for (size_t i = 0; i < n; ++i) idx[i] = (uint32_t)i;
  if (depth <= 0) return;
  split_state st = {kdops, centroids, n, idx, splits, 0, 0, 0.0};
  ml_rec(&st, depth - 1, 0, n);
}

/* ---- closest point on a triangle ----------------------------------------------------- */

/* m::dot (math/vec.hpp:247-250, haddv_op math/ops/vec_base_ops.hpp:248-263):
 * the products summed left to right from 0 */
static double dot3(const double *u, const double *v) {
  double s = 0.0;
  s += u[0] * v[0];
  s += u[1] * v[1];
  s += u[2] * v[2];
  return s;
}

/* bvh::kokkos::closest_point (primitive/kokkos_closest_point.hpp:55-133), one
 * point after the other; points and out are n x 3, tri = a, b, c */
void orc_closest_point(const double *points, size_t n, const double *tri, double *out) {
  const double *a = tri, *b = tri + 3, *c = tri + 6;
  double ab[3], ac[3], bc[3];
  for (int k = 0; k < 3; ++k) {
    ab[k] = b[k] - a[k];
    ac[k] = c[k] - a[k];
    bc[k] = c[k] - b[k];
  }
  for (size_t i = 0; i < n; ++i) {
    const double *p = points + 3 * i;
    double *o = out + 3 * i;
    double ap[3], bp[3], cp[3];
    for (int k = 0; k < 3; ++k) {
      ap[k] = p[k] - a[k];
      bp[k] = p[k] - b[k];
      cp[k] = p[k] - c[k];
    }
    double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { for (int k = 0; k < 3; ++k) o[k] = a[k]; continue; }
    double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { for (int k = 0; k < 3; ++k) o[k] = b[k]; continue; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
      double v = d1 / (d1 - d3);
      for (int k = 0; k < 3; ++k) o[k] = a[k] + v * ab[k];
      continue;
    }
    double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) { for (int k = 0; k < 3; ++k) o[k] = c[k]; continue; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
      double w = d2 / (d2 - d6);
      for (int k = 0; k < 3; ++k) o[k] = a[k] + w * ac[k];
      continue;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
      double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      for (int k = 0; k < 3; ++k) o[k] = b[k] + w * bc[k];
      continue;
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom, w = vc * denom;
    for (int k = 0; k < 3; ++k) o[k] = a[k] + ab[k] * v + ac[k] * w;
  }
}

typedef struct {
  uint64_t pp, ep, pq, eq;
} rec4;

static int cmp_rec(const void *a, const void *b) {
  const rec4 *x = (const rec4 *)a, *y = (const rec4 *)b;
  if (x->ep != y->ep) return (x->ep > y->ep) - (x->ep < y->ep);
  if (x->eq != y->eq) return (x->eq > y->eq) - (x->eq < y->eq);
  if (x->pp != y->pp) return (x->pp > y->pp) - (x->pp < y->pp);
  return (x->pq > y->pq) - (x->pq < y->pq);
}

/* Gathered per-patch element arrays of an object (sorted order). */
static void gather_patch(const orc_object *o, size_t k, double *kd, double *ce, uint64_t *gi) {
  size_t b = o->pbegin[k], e = o->pend[k];
  for (size_t p = b; p < e; ++p) {
    size_t src = o->perm[p];
    memcpy(kd + (p - b) * ORC_KDOP, o->kdops + src * ORC_KDOP, ORC_KDOP * sizeof(double));
    memcpy(ce + (p - b) * 3, o->centroids + src * 3, 3 * sizeof(double));
    gi[p - b] = o->gids[src];
  }
}

/* broadphase: the patch tree of b (top-down over the non-empty patches,
 * collision_object/top_down.cpp:113-127) queried with every non-empty patch of
 * a (collision_object/broadphase.cpp:68-119); narrowphase per patch pair: the
 * element tree of patch q (build_snapshot_tree_top_down, tree_build.hpp:
 * 159-185) queried with every element of patch p (query_tree), which is the
 * element-level route a narrowphase functor takes in the reference. */
orc_pairs *orc_collide(const orc_object *a, const orc_object *b, const orc_pipeline_cfg *cfg,
                       orc_pipeline_times *tm) {
  const size_t od = cfg->od;
  double t0 = orc_now();
  /* patch snapshots of b */
  size_t nb = 0;
  double *bk = (double *)malloc(od * ORC_KDOP * sizeof(double));
  double *bc = (double *)malloc(od * 3 * sizeof(double));
  uint64_t *bg = (uint64_t *)malloc(od * sizeof(uint64_t));
  for (size_t k = 0; k < od; ++k)
    if (b->pend[k] > b->pbegin[k]) {
      memcpy(bk + nb * ORC_KDOP, b->patch_kdops + k * ORC_KDOP, ORC_KDOP * sizeof(double));
      memcpy(bc + nb * 3, b->patch_centroids + k * 3, 3 * sizeof(double));
      bg[nb] = k;
      nb++;
    }
  orc_tree *ptree = orc_tree_build_top_down(bk, bc, bg, nb, 1);
  orc_pairs *pp = orc_pairs_new();
  for (size_t k = 0; k < od; ++k)
    if (a->pend[k] > a->pbegin[k]) orc_query_tree(ptree, a->patch_kdops + k * ORC_KDOP, k, pp);
  orc_tree_free(ptree);
  free(bk); free(bc); free(bg);
  double t1 = orc_now();
  if (tm) tm->patch_pairs = pp->n;

  /* deterministic sample of patch pairs (every stride-th) for bounded timing */
  size_t stride = 1;
  if (cfg->sample > 0.0 && cfg->sample < 1.0) stride = (size_t)(1.0 / cfg->sample + 0.5);
  if (stride < 1) stride = 1;

  /* element trees for the patches of b that appear in sampled pairs */
  orc_tree **etrees = (orc_tree **)calloc(od, sizeof(orc_tree *));
  char *need = (char *)calloc(od, 1);
  for (size_t i = 0; i < pp->n; i += stride) need[pp->data[2 * i + 1]] = 1;
  int nt = cfg->nthreads > 0 ? cfg->nthreads : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (size_t k = 0; k < od; ++k) {
    if (!need[k]) continue;
    size_t m = b->pend[k] - b->pbegin[k];
    double *kd = (double *)malloc((m ? m : 1) * ORC_KDOP * sizeof(double));
    double *ce = (double *)malloc((m ? m : 1) * 3 * sizeof(double));
    uint64_t *gi = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
    gather_patch(b, k, kd, ce, gi);
    etrees[k] = orc_tree_build_top_down(kd, ce, gi, m, cfg->leaf_size);
    free(kd); free(ce); free(gi);
  }
  size_t npairs_done = 0;
  orc_pairs **loc = (orc_pairs **)calloc((size_t)nt, sizeof(orc_pairs *));
  for (int t = 0; t < nt; ++t) loc[t] = orc_pairs_new();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(+ : npairs_done)
  for (size_t i = 0; i < pp->n; i += stride) {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    size_t p = pp->data[2 * i], q = pp->data[2 * i + 1];
    const orc_tree *et = etrees[q];
    orc_pairs *tmp = orc_pairs_new();
    for (size_t pos = a->pbegin[p]; pos < a->pend[p]; ++pos) {
      size_t src = a->perm[pos];
      if (et->nnodes) query_rec_exact(et, 0, a->kdops + src * ORC_KDOP, a->gids[src], tmp);
    }
    for (size_t r = 0; r < tmp->n; ++r) {
      orc_pairs_push(loc[tid], p, tmp->data[2 * r]);
      orc_pairs_push(loc[tid], q, tmp->data[2 * r + 1]);
    }
    orc_pairs_free(tmp);
    npairs_done += 1;
  }
  orc_pairs *res = orc_pairs_new();
  for (int t = 0; t < nt; ++t) {
    pairs_append(res, loc[t]);
    orc_pairs_free(loc[t]);
  }
  free(loc);
  for (size_t k = 0; k < od; ++k) orc_tree_free(etrees[k]);
  free(etrees);
  free(need);
  orc_pairs_free(pp);
  /* records sorted by (element_p, element_q) */
  qsort(res->data, res->n / 2, sizeof(rec4), cmp_rec);
  double t2 = orc_now();
  if (tm) {
    tm->t_broadphase += t1 - t0;
    tm->t_narrowphase += t2 - t1;
    tm->patch_pairs_done = npairs_done;
  }
  return res;
}

/* ---- exact primitive tests (no reference counterpart) --------------------------------
 * The reference's primitive layer is the triangle type
 * (primitive/triangle.hpp:40-52) and closest_point; it has no intersection
 * test.  The rule stated here is the GPU primitive narrowphase's:
 *  * closed sets: triangles (or balls) that only touch intersect;
 *  * every decision is the sign of an fp64 expression evaluated in the order
 *    written (no contraction), so the CPU and the GPU break ties (zero
 *    determinants, equal distances) identically.  For integer coordinates of
 *    magnitude <= 2^15 every determinant below is exact, so the rule is then
 *    the exact geometric answer;
 *  * a triangle whose fp64 normal (b - a) x (c - a) is exactly zero is
 *    degenerate and intersects nothing.
 * Triangle-triangle: the two plane-side tests first (all three vertices of
 * one triangle strictly on one side of the other's plane: disjoint); then
 * the edge-crossing formulation -- two closed triangles meet iff an edge of
 * one meets the other triangle -- with each segment-triangle test made of
 * orient3d signs (the line through the segment crosses the triangle iff the
 * three signed volumes of the segment with the triangle's edges agree in
 * sign, zeros allowed) and, for a segment in the triangle's plane or for
 * coplanar triangles, orient2d tests in the coordinate plane that drops the
 * largest component of the normal. */
typedef struct { double x, y, z; } orc_v3;

static orc_v3 orc_v(const double *p) {
  orc_v3 r = {p[0], p[1], p[2]};
  return r;
}

/* sign source: det [a - d; b - d; c - d] */
static double orc_orient3d(orc_v3 a, orc_v3 b, orc_v3 c, orc_v3 d) {
  const double adx = a.x - d.x, ady = a.y - d.y, adz = a.z - d.z;
  const double bdx = b.x - d.x, bdy = b.y - d.y, bdz = b.z - d.z;
  const double cdx = c.x - d.x, cdy = c.y - d.y, cdz = c.z - d.z;
  return adx * (bdy * cdz - bdz * cdy) + bdx * (cdy * adz - cdz * ady) + cdx * (ady * bdz - adz * bdy);
}

static orc_v3 orc_normal3(orc_v3 a, orc_v3 b, orc_v3 c) {
  const double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
  const double vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
  orc_v3 n = {uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx};
  return n;
}

/* the axis dropped by the 2D tests: largest |n| component, the first on ties */
static int orc_drop_axis(orc_v3 n) {
  const double ax = fabs(n.x), ay = fabs(n.y), az = fabs(n.z);
  if (ax >= ay && ax >= az) return 0;
  return ay >= az ? 1 : 2;
}

typedef struct { double u, v; } orc_v2;

static orc_v2 orc_proj(orc_v3 p, int drop) {
  orc_v2 r;
  if (drop == 0) { r.u = p.y; r.v = p.z; }
  else if (drop == 1) { r.u = p.x; r.v = p.z; }
  else { r.u = p.x; r.v = p.y; }
  return r;
}

static double orc_orient2d(orc_v2 a, orc_v2 b, orc_v2 c) {
  return (a.u - c.u) * (b.v - c.v) - (a.v - c.v) * (b.u - c.u);
}

/* p on the closed segment [a, b], given that the three are collinear */
static int orc_on_seg(orc_v2 a, orc_v2 b, orc_v2 p) {
  const double ulo = a.u < b.u ? a.u : b.u, uhi = a.u < b.u ? b.u : a.u;
  const double vlo = a.v < b.v ? a.v : b.v, vhi = a.v < b.v ? b.v : a.v;
  return ulo <= p.u && p.u <= uhi && vlo <= p.v && p.v <= vhi;
}

/* closed segments [p1, p2] and [q1, q2] in the plane */
static int orc_seg_seg2(orc_v2 p1, orc_v2 p2, orc_v2 q1, orc_v2 q2) {
  const double d1 = orc_orient2d(q1, q2, p1), d2 = orc_orient2d(q1, q2, p2);
  const double d3 = orc_orient2d(p1, p2, q1), d4 = orc_orient2d(p1, p2, q2);
  if (((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0)) && ((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0))) return 1;
  if (d1 == 0 && orc_on_seg(q1, q2, p1)) return 1;
  if (d2 == 0 && orc_on_seg(q1, q2, p2)) return 1;
  if (d3 == 0 && orc_on_seg(p1, p2, q1)) return 1;
  if (d4 == 0 && orc_on_seg(p1, p2, q2)) return 1;
  return 0;
}

/* p in the closed triangle t (plane) */
static int orc_in_tri2(const orc_v2 *t, orc_v2 p) {
  const double d1 = orc_orient2d(t[0], t[1], p), d2 = orc_orient2d(t[1], t[2], p), d3 = orc_orient2d(t[2], t[0], p);
  return (d1 >= 0 && d2 >= 0 && d3 >= 0) || (d1 <= 0 && d2 <= 0 && d3 <= 0);
}

/* closed segment [a, b] lying in the plane of triangle t, in 2D */
static int orc_seg_tri2(orc_v3 a, orc_v3 b, const orc_v3 *t, int drop) {
  const orc_v2 a2 = orc_proj(a, drop), b2 = orc_proj(b, drop);
  const orc_v2 t2[3] = {orc_proj(t[0], drop), orc_proj(t[1], drop), orc_proj(t[2], drop)};
  if (orc_in_tri2(t2, a2) || orc_in_tri2(t2, b2)) return 1;
  for (int i = 0; i < 3; ++i)
    if (orc_seg_seg2(a2, b2, t2[i], t2[(i + 1) % 3])) return 1;
  return 0;
}

/* closed segment [a, b] against the closed (non-degenerate) triangle t */
static int orc_seg_tri(orc_v3 a, orc_v3 b, const orc_v3 *t, int drop) {
  const double oa = orc_orient3d(t[0], t[1], t[2], a), ob = orc_orient3d(t[0], t[1], t[2], b);
  if ((oa > 0 && ob > 0) || (oa < 0 && ob < 0)) return 0;
  if (oa == 0 && ob == 0) return orc_seg_tri2(a, b, t, drop);
  const double s1 = orc_orient3d(a, b, t[0], t[1]), s2 = orc_orient3d(a, b, t[1], t[2]);
  const double s3 = orc_orient3d(a, b, t[2], t[0]);
  return (s1 >= 0 && s2 >= 0 && s3 >= 0) || (s1 <= 0 && s2 <= 0 && s3 <= 0);
}

int orc_tri_tri(const double *pp, const double *qq) {
  const orc_v3 P[3] = {orc_v(pp), orc_v(pp + 3), orc_v(pp + 6)};
  const orc_v3 Q[3] = {orc_v(qq), orc_v(qq + 3), orc_v(qq + 6)};
  const orc_v3 np = orc_normal3(P[0], P[1], P[2]), nq = orc_normal3(Q[0], Q[1], Q[2]);
  if ((np.x == 0 && np.y == 0 && np.z == 0) || (nq.x == 0 && nq.y == 0 && nq.z == 0)) return 0;
  double dq[3], dp[3];
  for (int i = 0; i < 3; ++i) dq[i] = orc_orient3d(P[0], P[1], P[2], Q[i]);
  if ((dq[0] > 0 && dq[1] > 0 && dq[2] > 0) || (dq[0] < 0 && dq[1] < 0 && dq[2] < 0)) return 0;
  for (int i = 0; i < 3; ++i) dp[i] = orc_orient3d(Q[0], Q[1], Q[2], P[i]);
  if ((dp[0] > 0 && dp[1] > 0 && dp[2] > 0) || (dp[0] < 0 && dp[1] < 0 && dp[2] < 0)) return 0;
  const int dropp = orc_drop_axis(np), dropq = orc_drop_axis(nq);
  if ((dq[0] == 0 && dq[1] == 0 && dq[2] == 0) || (dp[0] == 0 && dp[1] == 0 && dp[2] == 0)) {
    /* coplanar: an edge crossing, or one triangle inside the other (P's plane) */
    orc_v2 p2[3], q2[3];
    for (int i = 0; i < 3; ++i) {
      p2[i] = orc_proj(P[i], dropp);
      q2[i] = orc_proj(Q[i], dropp);
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        if (orc_seg_seg2(p2[i], p2[(i + 1) % 3], q2[j], q2[(j + 1) % 3])) return 1;
    return orc_in_tri2(q2, p2[0]) || orc_in_tri2(p2, q2[0]);
  }
  for (int i = 0; i < 3; ++i)
    if (orc_seg_tri(P[i], P[(i + 1) % 3], Q, dropq)) return 1;
  for (int i = 0; i < 3; ++i)
    if (orc_seg_tri(Q[i], Q[(i + 1) % 3], P, dropp)) return 1;
  return 0;
}

int orc_sphere_sphere(const double *a, const double *b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  const double rs = a[3] + b[3];
  return dx * dx + dy * dy + dz * dz <= rs * rs;
}

void orc_tri_tri_batch(const double *p, const double *q, size_t n, uint8_t *out) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) out[i] = (uint8_t)orc_tri_tri(p + 9 * i, q + 9 * i);
}

void orc_sphere_sphere_batch(const double *a, const double *b, size_t n, uint8_t *out) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) out[i] = (uint8_t)orc_sphere_sphere(a + 4 * i, b + 4 * i);
}
