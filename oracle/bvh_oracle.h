/*
 * bvh_oracle.h -- CPU restatement of distBVH's single-rank collision path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in distbvh_amd/ and the timed CPU baseline in bench.py.
 * Nothing in the product links, loads or calls it.
 *
 * Every function restates a reference function in plain C (file:line cited
 * at the definition, paths relative to the distBVH source tree).  The k-DOP
 * layout is the reference's `extents` array flattened: kdop[2*a] = min of
 * axis a, kdop[2*a+1] = max of axis a, a = 0..12 (26 doubles).
 *
 * Parity status (how each function is pinned; see tests/test_oracle_golden.py):
 *   kdop from vertices / sphere, project, overlap, longest_axis:
 *       KDopTest.cpp known answers + closed form of the normals.
 *   bounds, quantize, morton32/64, expand:  hash_test.cpp / snapshot_test.cpp
 *       known answers, BMI2-style bit-deposit brute force.
 *   radix sort 1-bit LSD: sort_test.cpp known answer + numpy stable argsort.
 *   prefix sum: UtilTest.cpp known answer.   clz: bits_test.cpp.
 *   cluster splits: closed form (sort of (clz depth, index), first c) checked
 *       against a brute-force python restatement on random codes.
 *   patches, top-down tree, query_tree, self/potential collision sets,
 *   sort_and_sweep, pipeline: CollisionTest.cpp / KokkosTest.cpp counts and
 *       brute-force all-pairs overlap (the pair set is partition invariant).
 *   triangle-triangle, sphere-sphere: separating-axis tests in exact rational
 *       arithmetic and closed-form configurations (no reference counterpart).
 */
#ifndef BVH_ORACLE_H
#define BVH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NAXIS 13
#define ORC_KDOP 26

/* ---- k-DOP (kdop.hpp) ---------------------------------------------------- */
double orc_normal(int axis, int comp);
double orc_project(const double *v, int axis);
void orc_kdop_empty(double *k);
void orc_kdop_from_vertices(const double *v, int nv, double eps, double *k);
void orc_kdop_from_sphere(const double *c, double r, double *k);
int orc_overlap(const double *a, const double *b);
void orc_kdop_union(double *dst, const double *src);
int orc_longest_axis(const double *k);

/* Batch element k-DOPs and centroids in input order.
 * verts: SoA planes, plane (v*3 + c) of length n holds coordinate c of vertex v. */
void orc_elements_poly(const double *verts, int nv, size_t n, double eps,
                       double *kdops, double *centroids);
/* spheres: SoA planes x, y, z, r */
void orc_elements_sphere(const double *xyzr, size_t n, double *kdops, double *centroids);

/* ---- bounds / Morton (contact_entity.hpp, hash.hpp) ---------------------- */
void orc_compute_bounds(const double *kdops, size_t n, double *min3, double *inv_diag3);
uint32_t orc_expand32(uint32_t v);
uint64_t orc_expand64(uint64_t v);
uint64_t orc_pdep64(uint64_t v, uint64_t mask);
uint32_t orc_morton32(uint32_t x, uint32_t y, uint32_t z);
uint64_t orc_morton64(uint64_t x, uint64_t y, uint64_t z);
uint64_t orc_float_to_uint(double s);
void orc_quantize32(const double *p, const double *min3, const double *inv3, uint32_t *q);
void orc_quantize64(const double *p, const double *min3, const double *inv3, uint64_t *q);
void orc_morton_codes32(const double *centroids, size_t n, const double *min3,
                        const double *inv3, uint32_t *out);
void orc_morton_codes64(const double *centroids, size_t n, const double *min3,
                        const double *inv3, uint64_t *out);

/* ---- sort / scan / bits (util/) ------------------------------------------ */
void orc_radix_sort32(uint32_t *keys, uint32_t *idx, size_t n);
void orc_radix_sort64(uint64_t *keys, uint32_t *idx, size_t n);
void orc_prefix_sum_u32(uint32_t *v, size_t n);
int orc_clz32(uint32_t v);
int orc_clz64(uint64_t v);

/* ---- cluster splits + patches (split/cluster.hpp, collision_object.cpp) -- */
void orc_cluster_splits32(const uint32_t *sorted_codes, size_t n, size_t nsplits, uint64_t *splits);
void orc_cluster_splits64(const uint64_t *sorted_codes, size_t n, size_t nsplits, uint64_t *splits);
/* Patch k covers sorted positions [begin[k], end[k]).  kdops/centroids are in
 * sorted order (gathered by the permutation).  Outputs per patch: size, kdop
 * (26), centroid (3). */
void orc_patches(const uint64_t *splits, size_t od, size_t n, const double *sorted_kdops,
                 const double *sorted_centroids, uint64_t *begin, uint64_t *end,
                 double *patch_kdops, double *patch_centroids);

/* ---- top-down tree (bvh_build.hpp, split/split.hpp, split/mean.hpp) ------ */
typedef struct {
  size_t nnodes;
  size_t nleaves;
  double *kdop;      /* nnodes * 26 */
  int64_t *left;     /* child node index or -1 */
  int64_t *right;
  uint64_t *first;   /* leaf range [first, last) into leaf arrays */
  uint64_t *last;
  /* leaves after partitioning (the tree's own copy, reordered) */
  double *leaf_kdop;     /* nleaves * 26 */
  double *leaf_centroid; /* nleaves * 3 */
  uint64_t *leaf_gid;
} orc_tree;

orc_tree *orc_tree_build_top_down(const double *kdops, const double *centroids,
                                  const uint64_t *gids, size_t n, size_t leaf_size);
void orc_tree_free(orc_tree *t);

/* growable pair buffer: pairs stored as (first, second) uint64 */
typedef struct {
  uint64_t *data;
  size_t n;
  size_t cap;
} orc_pairs;
orc_pairs *orc_pairs_new(void);
void orc_pairs_free(orc_pairs *p);
void orc_pairs_push(orc_pairs *p, uint64_t a, uint64_t b);

void orc_query_tree(const orc_tree *t, const double *kdop, uint64_t gid, orc_pairs *out);
void orc_self_collision_set(const orc_tree *t, orc_pairs *out);
/* as_written != 0 reproduces collision_query.hpp:310-312 (right/left leaf
 * arrays swapped); 0 emits (left leaf, right leaf). */
void orc_potential_collision_set(const orc_tree *l, const orc_tree *r, int as_written, orc_pairs *out);

/* ---- element narrowphase (narrowphase/kokkos.hpp) ------------------------- */
int orc_max_variant_axis(const double *ca, size_t na, const double *cb, size_t nb);
void orc_sort_and_sweep(const double *pa_kdop, const double *ka, const uint64_t *ga, size_t na,
                        const double *pb_kdop, const double *kb, const uint64_t *gb, size_t nb,
                        int axis, orc_pairs *out);

/* ---- brute force (the master pair-set oracle) ---------------------------- */
orc_pairs *orc_brute_pairs(const double *ka, size_t na, const double *kb, size_t nb, int nthreads);
orc_pairs *orc_brute_self(const double *k, size_t n, int nthreads);

/* ---- full single-rank pipeline ------------------------------------------- */
typedef struct {
  size_t od;            /* overdecomposition factor (patches per object) */
  int morton_bits;      /* 30 (reference clustering path) or 63 */
  size_t leaf_size;     /* leaf size of the per-patch element trees */
  int nthreads;         /* OpenMP threads for the narrowphase (1 = serial) */
  double sample;        /* fraction of broadphase patch pairs narrowphased (1 = all) */
  int split;            /* ORC_SPLIT_*: split_algorithm of set_entity_data */
} orc_pipeline_cfg;

/* split_algorithm, types.hpp:52-57 */
#define ORC_SPLIT_GEOM_AXIS 0
#define ORC_SPLIT_ML_GEOM_AXIS 1
#define ORC_SPLIT_CLUSTERING 2

/* bvh::kokkos::closest_point: n x 3 points, tri = a, b, c (9 doubles) */
void orc_closest_point(const double *points, size_t n, const double *tri, double *out);

/* ---- exact primitive tests beyond k-DOP overlap ----------------------------
 * The reference has a triangle type (primitive/triangle.hpp:40-52) and the
 * closest point kernel, but no primitive intersection test; these state the
 * rule the GPU's primitive narrowphase implements (closed sets: touching
 * counts as intersecting; fp64 evaluation in a fixed order, so CPU and GPU
 * decide ties identically; a triangle whose fp64 normal is exactly zero is
 * degenerate and intersects nothing).  Pinned against exact rational
 * arithmetic (separating axes over Fractions) on integer and random inputs
 * and against closed-form configurations (tests/test_oracle_primitive.py).
 * tri: 9 doubles a, b, c; sphere: 4 doubles x, y, z, r. */
int orc_tri_tri(const double *p, const double *q);
int orc_sphere_sphere(const double *a, const double *b);
/* n pairs: p, q are n x 9 (n x 4), out[i] = 0 / 1 */
void orc_tri_tri_batch(const double *p, const double *q, size_t n, uint8_t *out);
void orc_sphere_sphere_batch(const double *a, const double *b, size_t n, uint8_t *out);

/* split_permutations / split_permutations_ml with split::mean and
 * axis::longest: idx (n) and 2^depth - 1 split offsets */
void orc_split_geom_axis(const double *kdops, const double *centroids, size_t n, int depth, uint32_t *idx,
                         uint64_t *splits);
void orc_split_ml_geom_axis(const double *kdops, const double *centroids, size_t n, int depth, uint32_t *idx,
                            uint64_t *splits);

typedef struct {
  /* per object */
  double t_bounds, t_morton, t_sort, t_cluster, t_patches, t_tree;
  /* per step */
  double t_broadphase, t_narrowphase;
  size_t patch_pairs, patch_pairs_done;
} orc_pipeline_times;

/* One object's build: outputs are caller-allocated (n or od sized). */
typedef struct {
  size_t n;
  double *kdops;      /* n*26 input order */
  double *centroids;  /* n*3  input order */
  uint64_t *gids;     /* n */
  /* results */
  uint64_t *codes;    /* n sorted codes (64-bit container for both widths) */
  uint32_t *perm;     /* n: sorted position -> input index */
  uint64_t *splits;   /* od-1 */
  uint64_t *pbegin, *pend; /* od */
  double *patch_kdops;     /* od*26 */
  double *patch_centroids; /* od*3 */
} orc_object;

void orc_object_build(orc_object *o, const orc_pipeline_cfg *cfg, orc_pipeline_times *tm);

/* Broadphase (patch tree query) + narrowphase (per patch pair element tree
 * query) between built objects a and b.  Returns records
 * {patch_p, element_p, patch_q, element_q} as 4 consecutive uint64 per pair
 * (stored as two orc_pairs entries), sorted by (element_p, element_q). */
orc_pairs *orc_collide(const orc_object *a, const orc_object *b, const orc_pipeline_cfg *cfg,
                       orc_pipeline_times *tm);

/* utilities */
double orc_now(void);
void orc_sort_u64(uint64_t *v, size_t n);
/* generic wrappers for python callers that free buffers */
void orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
