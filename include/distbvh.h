/*
 * distbvh.h -- C ABI of the MI355X (gfx950) collision pipeline.
 *
 * The entry points below are the drop-in boundary for distBVH's single-rank
 * collision path.  Each one names the reference interface it replaces
 * (paths relative to the distBVH source tree).  Plain pointers and sizes
 * only; no exceptions cross this boundary; every call returns a bvh_status.
 *
 * Threading: one host thread drives a world.  All device work of a world is
 * ordered on one HIP stream (the caller's, or one the world creates with the
 * default priority), with one exception: the build that set_entity_data
 * enqueues runs on the object's own stream (default priority), after all
 * work enqueued on the world stream before the call, so that the builds of
 * several objects run concurrently; the world stream waits for it at the
 * next call on the world that is not a set_entity_data or init_broadphase
 * (bvh_broadphase(a, b) lets a's upper tree levels run beside its walk and
 * waits for them before it returns; BVH_FLAG_SERIAL_BUILDS keeps every build
 * on the world stream).  Until then the caller's own work on the world
 * stream is not ordered after it.
 *
 * Element data: polytope vertices are SoA planes, plane (v*3 + c) of length
 * n holds coordinate c (x, y, z) of vertex v of every element; spheres are
 * the four planes x, y, z, r.  Pointers may be host or device memory
 * (BVH_MEM_HOST / BVH_MEM_DEVICE).  Host data is uploaded during the call and
 * not retained.  Device data is read in place and, like the element pointer
 * the reference keeps until the narrowphase (collision_object.cpp:128-129),
 * must stay valid until the object's next set_entity_data.
 */
#ifndef DISTBVH_H
#define DISTBVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BVH_ABI_VERSION 1

typedef enum bvh_status {
  BVH_OK = 0,
  BVH_E_INVALID = 1,     /* bad argument; od not a power of two (collision_world.cpp:59-65) */
  BVH_E_OOM = 2,         /* device allocation failed */
  BVH_E_OVERFLOW = 3,    /* a caller buffer is too small; the needed size is reported */
  BVH_E_HIP = 4,         /* HIP runtime error (text in bvh_last_error) */
  BVH_E_STATE = 5,       /* call out of order (e.g. broadphase before set_entity_data) */
  BVH_E_UNSUPPORTED = 6  /* option not built */
} bvh_status;

/* split_algorithm, types.hpp:52-57 */
typedef enum bvh_split_algorithm {
  BVH_SPLIT_GEOM_AXIS = 0,
  BVH_SPLIT_ML_GEOM_AXIS = 1,
  BVH_SPLIT_CLUSTERING = 2,
  /* no reference counterpart: patches given by the caller
   * (bvh_object_set_patch_offsets), e.g. one patch per mesh of a batch */
  BVH_SPLIT_GIVEN = 3
} bvh_split_algorithm;

typedef enum bvh_memory { BVH_MEM_HOST = 0, BVH_MEM_DEVICE = 1 } bvh_memory;

typedef enum bvh_precision {
  BVH_PREC_FP64 = 0,          /* the fp64 pair set (bit-exact with the CPU path): tree bounds
                                 are fp32 rounded outward; a leaf candidate is kept when its
                                 fp32 margin proves the fp64 overlap, else decided on fp64
                                 k-DOPs; element k-DOPs are exported in fp64 */
  BVH_PREC_FP32_OUTWARD = 1,  /* the same walk and pairs, fp32 slabs only (no fp64 element
                                 k-DOP export) */
  BVH_PREC_FP64_RECORDS = 2   /* fp64 slabs in every tree record (the all-fp64 walk) */
} bvh_precision;

/* bvh_world_config.flags */
#define BVH_FLAG_SORT_RESULTS 0x1u /* sort element pairs by (element_p, element_q) */
#define BVH_FLAG_TIMING 0x2u       /* record per-stage HIP event timings */
#define BVH_FLAG_SERIAL_BUILDS 0x4u /* build on the world stream, one object after another */

/* world_config, collision_world.hpp:48-52 (log levels replaced by device
 * options) */
typedef struct bvh_world_config {
  uint32_t morton_bits; /* 63 (21 bits/axis, default) or 30 (10 bits/axis, the
                           reference clustering path, split/cluster.hpp:139) */
  uint32_t leaf_size;   /* 1..8, accepted for configuration compatibility: the trees
                           hold one element per leaf, as the reference's do
                           (tree_build.hpp:69); a leaf's slabs sit in its parent's
                           record, so the walk tests leaves two at a time */
  uint32_t precision;   /* bvh_precision */
  uint32_t flags;
  void *stream;         /* hipStream_t, or NULL for a world-owned stream */
  int32_t device;       /* HIP device ordinal (default 0) */
  uint32_t reserved[7];
} bvh_world_config;

typedef struct bvh_world_s *bvh_world;
typedef struct bvh_object_s *bvh_object;
typedef struct bvh_graph_s *bvh_graph;

/* detailed_narrowphase_result, tests/collision_object_test.cpp:318-324: the
 * 32-byte record the built-in narrowphase emits, read back through
 * for_each_result<R> (collision_object.hpp:177-187). */
typedef struct bvh_narrowphase_record {
  uint64_t patch_p;
  uint64_t element_p;
  uint64_t patch_q;
  uint64_t element_q;
} bvh_narrowphase_record;

/* patch<>, patch.hpp:55-123 (global id, size, 26-DOP, centroid) */
typedef struct bvh_patch_info {
  uint64_t global_id;
  uint64_t size;
  double kdop[26]; /* min0, max0, ..., min12, max12 */
  double centroid[3];
} bvh_patch_info;

/* ---- status ---------------------------------------------------------------- */
const char *bvh_status_string(bvh_status s);
/* text of the last error on this thread ("" when none) */
const char *bvh_last_error(void);
uint32_t bvh_abi_version(void);

/* ---- world (collision_world.hpp:54-114) -------------------------------------- */
void bvh_world_config_init(bvh_world_config *cfg);
/* collision_world(std::size_t od, const world_config &), collision_world.hpp:61 */
bvh_status bvh_world_create(size_t overdecomposition, const bvh_world_config *cfg, bvh_world *out);
/* ~collision_world() */
bvh_status bvh_world_destroy(bvh_world w);
/* overdecomposition_factor(), collision_world.hpp:73 */
bvh_status bvh_world_overdecomposition(bvh_world w, size_t *od);
/* create_collision_object(), collision_world.hpp:70 (the world owns it) */
bvh_status bvh_create_collision_object(bvh_world w, bvh_object *out);
/* num_collision_objects(), collision_world.hpp:72 */
bvh_status bvh_num_collision_objects(bvh_world w, size_t *n);
/* start_iteration() / finish_iteration(), collision_world.hpp:96-97 */
bvh_status bvh_start_iteration(bvh_world w);
bvh_status bvh_finish_iteration(bvh_world w);
/* wait for all device work of the world */
bvh_status bvh_synchronize(bvh_world w);
/* the world's HIP stream (hipStream_t): callers that write element data on
 * another stream make this one wait for that work (an event) before passing
 * the data in (no reference counterpart: the reference's Kokkos views are
 * ordered by Kokkos::fence) */
bvh_status bvh_world_stream(bvh_world w, void **stream);
/* Per-stage device times (ms) recorded (BVH_FLAG_TIMING) since the previous
 * call of this function, in call order; reading clears the record (at most
 * 512 stages are kept).  names[i] gets a static string. */
bvh_status bvh_stage_times(bvh_world w, const char **names, double *ms, size_t cap, size_t *count);

/* ---- collision object (collision_object.hpp:55-283) --------------------------- */
/* id(), collision_object.hpp:176 */
bvh_status bvh_object_id(bvh_object o, size_t *id);
/* set_entity_data(View<T*>, split_algorithm), collision_object.hpp:69-84 with
 * T a polytope element whose kdop() is from_vertices (kdop.hpp:238-264) and
 * whose centroid() is the vertex mean (tests/TestCommon.hpp:62-73).
 * verts_per_element in 1..8; global_ids may be NULL (ids 0..n-1). */
bvh_status bvh_set_entity_data_poly(bvh_object o, const double *verts, uint32_t verts_per_element,
                                    uint64_t n, const uint64_t *global_ids, double epsilon,
                                    uint32_t mem, uint32_t split_algorithm);
/* Patches for set_entity_data(..., BVH_SPLIT_GIVEN): patch k holds the
 * elements of input indices [offsets[k-1], offsets[k]) (offsets[-1] = 0,
 * offsets[od-1] = n), count = od - 1 non-decreasing offsets (host memory),
 * kept for the object's later builds.  Inside each patch the tree orders
 * elements by Morton code, as for geom_axis.  No reference counterpart: the
 * reference's split_algorithm (types.hpp:52-57) always computes its patches;
 * a batch of independent meshes names its own. */
bvh_status bvh_object_set_patch_offsets(bvh_object o, const uint64_t *offsets, size_t count);
/* set_entity_data for sphere elements: kdop = from_sphere (kdop.hpp:273-288),
 * centroid = center.  xyzr: planes x, y, z, r. */
bvh_status bvh_set_entity_data_sphere(bvh_object o, const double *xyzr, uint64_t n,
                                      const uint64_t *global_ids, uint32_t mem,
                                      uint32_t split_algorithm);
/* set_entity_data for a generic ContactEntity (traits.hpp:114-135): the
 * caller supplies each element's 26-DOP (n x 26, AoS) and centroid (n x 3),
 * i.e. the reference's entity_snapshot (snapshot.hpp:68-149). */
bvh_status bvh_set_entity_data_snapshot(bvh_object o, const double *kdops, const double *centroids,
                                        uint64_t n, const uint64_t *global_ids, uint32_t mem,
                                        uint32_t split_algorithm);
/* Refit (no reference counterpart: the reference rebuilds on every
 * set_entity_data).  New element data of the same kind and count as the last
 * set_entity_data (vertex planes, sphere planes, or n x 26 k-DOPs) on the
 * topology of the last build: the permutation, patches and tree shape are
 * kept and every node's k-DOP is recomputed bottom-up, so queries stay exact.
 * Morton codes, bounds and patch centroids stay those of the last build. */
bvh_status bvh_refit(bvh_object o, const double *data, uint32_t mem);
/* init_broadphase(), collision_object.hpp:87 / collision_object.cpp:148-215 */
bvh_status bvh_init_broadphase(bvh_object o);
/* broadphase(collision_object &other), collision_object.hpp:170 /
 * collision_object.cpp:248-299: patch broadphase + the built-in element
 * narrowphase (strict 26-DOP overlap, kdop.hpp:538-549).  Results land in
 * `o` (impl.cpp:328-348 sends both halves to the initiating object). */
bvh_status bvh_broadphase(bvh_object o, bvh_object other);
/* the broadphase half of broadphase(): patch pairs only, for callers that
 * run their own narrowphase per patch pair (the reference's narrowphase
 * functor, collision_world.hpp:75-94); the result count is then 0. */
bvh_status bvh_broadphase_patches(bvh_object o, bvh_object other);
/* broadphase(other) for k pairs at once, objects[i] x others[i] (k distinct
 * receiving objects, no object with itself): the coalesced step of a
 * serving loop.  Each pair's patch broadphase is its own small launch; the
 * element narrowphases of all pairs are one launch (every form; the forms
 * with fp32-outward records add one fp64 re-check launch per pair; with
 * sort_results the pairs run one after another); one synchronization reads
 * every pair's counts.  Results land in objects[i] as after
 * broadphase.  No reference counterpart (the reference runs one broadphase
 * per call, collision_object.hpp:170). */
bvh_status bvh_broadphase_batch(bvh_object *objects, bvh_object *others, uint64_t k);
/* self_collision_set(tree), collision_query.hpp:360-368 at element level:
 * unordered element pairs (i < j) of one object with overlapping 26-DOPs. */
bvh_status bvh_self_collision(bvh_object o);
/* self_collision restricted to pairs whose elements lie in different
 * patches (i < j): for an object holding a batch of independent meshes, one
 * patch per mesh (BVH_SPLIT_GIVEN), the pairs between meshes -- what the
 * reference's broadphase gives between separate collision objects, whose
 * narrowphase never pairs an object with itself
 * (collision_object/impl.cpp:229-235).  The walk skips every subtree inside
 * the querying element's own patch. */
bvh_status bvh_self_collision_cross_patch(bvh_object o);
/* end_phase(), collision_object.hpp:172 */
bvh_status bvh_end_phase(bvh_object o);

/* for_each_result<R>(f), collision_object.hpp:177-187: number of records and
 * the 32-byte records themselves (host or device destination). */
bvh_status bvh_result_count(bvh_object o, uint64_t *count);
bvh_status bvh_result_records(bvh_object o, bvh_narrowphase_record *out, uint64_t cap, uint32_t mem);
/* compact pairs (input index of element_p << 32 | input index of element_q) */
bvh_status bvh_result_pairs(bvh_object o, uint64_t *out, uint64_t cap, uint32_t mem);
/* order-independent digest of the compact pairs, computed on the device:
 * digest[0] = sum of mix(key), digest[1] = xor of mix'(key), with mix the
 * splitmix64 finalizer.  Equal pair sets give equal digests in any order (no
 * reference counterpart; for comparing result sets too large to copy). */
bvh_status bvh_result_digest(bvh_object o, uint64_t digest[2]);
/* device pointer to the compact pairs (valid until the next broadphase) */
bvh_status bvh_result_pairs_device(bvh_object o, const uint64_t **pairs, uint64_t *count);
/* counters of the last query: tree nodes visited (summed over waves of 64
 * queries), broadphase patch pairs, candidate pairs (fp32-outward records:
 * pairs whose outward-rounded fp32 slabs overlap; fp64 records: the pair
 * count), rechecked pairs (candidates decided on fp64 k-DOPs: with
 * fp32-outward records those whose fp32 gap proves nothing, every candidate
 * between two objects, whose frames differ; 0 with fp64 records) and the
 * largest fp32 slab error in ulps of the two objects'
 * leaf slabs (0 with fp64 records; measured by this call from the element
 * data, which must still be valid).  Any pointer may be NULL. */
bvh_status bvh_result_stats(bvh_object o, uint64_t *node_visits, uint64_t *patch_pairs,
                            uint64_t *candidate_pairs, uint64_t *rechecked_pairs,
                            double *max_slab_error_ulps);
/* broadphase patch pairs (query_tree emits, broadphase.cpp:93-118):
 * out[2k] = patch of o, out[2k+1] = patch of other. */
bvh_status bvh_patch_pair_count(bvh_object o, uint64_t *count);
bvh_status bvh_patch_pairs(bvh_object o, uint64_t *out, uint64_t cap);

/* local_patches(), collision_object.hpp:189; overdecomposition_factor() */
bvh_status bvh_local_patches(bvh_object o, bvh_patch_info *out, size_t cap);
bvh_status bvh_object_overdecomposition(bvh_object o, size_t *od);
/* device memory the object holds, in bytes (its build and result buffers) */
bvh_status bvh_object_device_bytes(bvh_object o, uint64_t *bytes);
/* Frees the object's device memory (no reference counterpart: the reference's
 * objects live as long as their world).  The handle stays valid; the object is
 * empty until its next set_entity_data.  For serving loops that keep many
 * objects resident under a device-memory budget. */
bvh_status bvh_object_release(bvh_object o);
/* the root node's 26-DOP (fp32 form: the outward fp32 slabs shifted back to
 * world coordinates); for refit-or-rebuild decisions */
bvh_status bvh_object_root_kdop(bvh_object o, double *kdop);
/* the input indices of the elements of one patch in patch order (the bytes
 * a narrowphase functor receives, collision_object/impl.hpp:112-118) */
bvh_status bvh_patch_elements(bvh_object o, size_t patch, uint64_t *indices, uint64_t cap,
                              uint64_t *count);

/* ---- parity hooks (stage outputs of the last set_entity_data) ------------------ */
/* compute_bounds(min_inv_diag_bounds), contact_entity.hpp:136-148 */
bvh_status bvh_object_bounds(bvh_object o, double *min3, double *inv_diag3);
/* sorted Morton codes (always widened to 64 bits) and the sort permutation
 * (sorted position -> input index), util/sort.hpp:100-115 */
bvh_status bvh_object_sorted(bvh_object o, uint64_t *codes, uint32_t *perm, uint32_t mem);
/* cluster splits (od - 1 values), split/cluster.hpp:89-168 */
bvh_status bvh_object_splits(bvh_object o, uint64_t *splits, size_t cap);
/* element 26-DOPs in sorted order (n x 26), the snapshots' kdops */
bvh_status bvh_object_sorted_kdops(bvh_object o, double *kdops, uint32_t mem);
/* for_each_tree (collision_object.hpp:164-168, collision_object.cpp:217-229)
 * hands the patch-level tree (a snapshot_tree over the non-empty patches,
 * tree.hpp:75-452, count() = its leaves).  Here: the levels of the tree above
 * the patch roots, in pre-order, as bvh_node (node.hpp:318-322): offsets
 * relative to the node (0: none / no parent), the exact fp64 k-DOP (a leaf:
 * its patch's; an inner node: the union of its children) and the range
 * [first, last) of its leaves in leaf_patches (patch ids, left to right).
 * Either output may be null to query the sizes. */
typedef struct bvh_tree_node {
  int64_t parent_offset;
  int64_t child_offsets[2];
  double kdop[26];
  uint64_t entity_range[2];
} bvh_tree_node;
bvh_status bvh_object_patch_tree(bvh_object o, bvh_tree_node *nodes, size_t cap, size_t *nnodes,
                                 uint64_t *leaf_patches, size_t leaf_cap, size_t *nleaves);
/* the element-level tree (no reference counterpart: the reference's element
 * trees are per patch pair): node count,
 * then per node kdop (26), children (-1 for none) and leaf range [first, last].
 * With fp64 records the k-DOPs are the exact slab unions; with fp32-outward
 * records (BVH_PREC_FP64, BVH_PREC_FP32_OUTWARD) they are those bounds in
 * world coordinates, each containing the node's exact k-DOP. */
bvh_status bvh_object_tree(bvh_object o, double *kdops, int64_t *children, uint64_t *ranges,
                           size_t cap, size_t *nnodes, size_t *root);

/* ---- stand-alone device primitives (device pointers, world's stream) ---------- */
/* radix_sorter<T, I>, util/sort.hpp:75-144: stable sort of keys with a
 * uint32 payload (payload may be NULL), in place. */
bvh_status bvh_sort_u64(bvh_world w, uint64_t *keys, uint32_t *vals, uint64_t n);
bvh_status bvh_sort_u32(bvh_world w, uint32_t *keys, uint32_t *vals, uint64_t n);
/* bvh_sort_u64 without any device allocation; no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * The scratch is one static device array per GPU (26 MB, the library's own
 * code-object memory, held until the process ends), so n is at most
 * 1,048,576 (2^20) keys; larger n gives BVH_E_INVALID.  Every call on one
 * GPU shares that array: calls that overlap on one GPU (two host threads,
 * or two worlds on the same device) race and must be serialized by the
 * caller.  Same result as bvh_sort_u64; waits for it. */
bvh_status bvh_sort_u64_noalloc(bvh_world w, uint64_t *keys, uint32_t *vals, uint64_t n);
/* prefix_sum, util/prefix_sum.hpp:41-54 (exclusive), in place */
bvh_status bvh_exclusive_scan_u32(bvh_world w, uint32_t *vals, uint64_t n);

/* bvh::kokkos::closest_point(points, triangle, closest)
 * (primitive/kokkos_closest_point.hpp:55-133): the closest point of the
 * triangle tri = {a, b, c} (9 doubles, host memory) to each of n points;
 * points and closest are n x 3 doubles in `mem` (BVH_MEM_HOST or
 * BVH_MEM_DEVICE).  Bit for bit the reference's expressions. */
bvh_status bvh_closest_point(bvh_world w, const double *points, uint64_t n, const double *tri,
                             double *closest, uint32_t mem);

/* Exact primitive tests beyond k-DOP overlap.  The reference's primitive
 * layer is the triangle type (primitive/triangle.hpp:40-52) and closest_point
 * (primitive/kokkos_closest_point.hpp:55-133); it has no intersection test,
 * so these state one rule: closed sets (touching counts as intersecting);
 * every decision is the sign of an fp64 expression evaluated in a fixed
 * order (orient3d / orient2d determinants, squared distances), so ties are
 * broken the same way on every run (on integer coordinates up to 2^15 every
 * determinant is exact and so is the answer); a triangle whose fp64 normal is
 * exactly zero meets nothing.  Triangle-triangle: plane-side rejection, then
 * the edge-crossing formulation (an edge of one triangle meets the other)
 * with orient3d sign tests, coplanar cases in 2D.
 * n pairs: p, q are n x 9 doubles (triangle vertices a, b, c) or n x 4
 * (x, y, z, r), out[i] = 1 if pair i intersects; all in `mem`. */
bvh_status bvh_triangle_pairs_intersect(bvh_world w, const double *p, const double *q, uint64_t n, uint32_t mem,
                                        uint8_t *out);
bvh_status bvh_sphere_pairs_intersect(bvh_world w, const double *a, const double *b, uint64_t n, uint32_t mem,
                                      uint8_t *out);
/* The primitive narrowphase on the object's results: keeps the pairs (of the
 * last broadphase / self_collision) whose triangles (3-vertex polytope
 * elements on both objects) or spheres intersect by the rule above, in their
 * order; the records and counts then describe the kept pairs.  *kept (may be
 * NULL) is their number.  The element data must still be valid.  No
 * reference counterpart: it is a narrowphase functor's job in the reference
 * (collision_world.hpp:75-94). */
bvh_status bvh_result_filter_exact(bvh_object o, uint64_t *kept);

/* kdop::from_vertices (kdop.hpp:238-264) for n polytopes: SoA vertex planes
 * (as for set_entity_data) in, n x 26 k-DOPs (min, max per direction) out,
 * both in `mem`; the same values the build computes.  For callers that sweep
 * elements outside an object (sort_and_sweep_local). */
bvh_status bvh_element_kdops(bvh_world w, const double *verts, uint32_t verts_per_element, uint64_t n,
                             double epsilon, uint32_t mem, double *kdops);
/* max_variant_axis(a, b) (narrowphase/kokkos.hpp:119-161): the axis (0, 1, 2)
 * of largest variance of the centroids of A and B together (na x 3 and
 * nb x 3 doubles in `mem`), var = sum(c^2) - sum(c)^2 / (na + nb), the first
 * axis on ties.  The sums are a fixed-order parallel reduction; the
 * reference's parallel_reduce has no fixed order either.  *axis: host. */
bvh_status bvh_max_variant_axis(bvh_world w, const double *ca, uint64_t na, const double *cb, uint64_t nb,
                                uint32_t mem, int32_t *axis);
/* sort_and_sweep_local(pa, a, pb, b, axis, f) (narrowphase/kokkos.hpp:164-248):
 * the pairs of elements of patch A and patch B whose 26-DOPs overlap, found
 * by sweeping B sorted by its slab minima on direction `axis` (0..12), scaled
 * to uint32 over the union of the patches' slabs pa, pb (26 doubles each,
 * host memory) and truncated.  ka / kb are na x 26 / nb x 26 element k-DOPs
 * (min, max per direction), ga / gb their global ids (NULL: the local
 * indices), all in `mem`; pairs (2 x uint64 per pair: id in A, id in B, in
 * `mem`) come in the order the reference's callback sees them in for each
 * element of A: A ascending, then B in sorted order.  *npairs is the count;
 * pairs may be NULL to ask for it; BVH_E_OVERFLOW if cap is below it.
 * nb < 2^30, at most 2^32 - 1 pairs. */
bvh_status bvh_sort_and_sweep(bvh_world w, const double *pa, const double *ka, const uint64_t *ga, uint64_t na,
                              const double *pb, const double *kb, const uint64_t *gb, uint64_t nb, int32_t axis,
                              uint32_t mem, uint64_t *pairs, uint64_t cap, uint64_t *npairs);

/* ---- wire formats (the byte layouts the reference sends between ranks) ------------
 * Packed and unpacked on the device (or from / to host memory), for a
 * multi-GPU driver to move with RCCL.  Each layout is the field order of the
 * reference's serialize() with every field as its in-memory bytes (DARMA
 * checkpoint copies byte-copyable types whole,
 * serialization/bvh_serialize.hpp:66-79); checkpoint's own message framing
 * is not reproduced. */
typedef enum bvh_element_kind {
  BVH_ELEM_POLY = 0,     /* polytope vertices (3 * verts_per_element doubles) */
  BVH_ELEM_SPHERE = 1,   /* x, y, z, r */
  BVH_ELEM_SNAPSHOT = 2  /* 26 k-DOP slabs, then 3 centroid coordinates */
} bvh_element_kind;
#define BVH_PATCH_RECORD_BYTES 256u
#define BVH_RESULT_RECORD_BYTES 32u
/* patch<>::serialize (patch.hpp:93-97): per patch of o, in patch order, one
 * 256-byte record: u64 global_id, u64 size, the 13 extents {f64 min, f64 max}
 * of the 26-DOP (kdop.hpp:57-60; an empty patch keeps {DBL_MAX, -DBL_MAX}),
 * the centroid as vec3<double> storage (4 doubles: x, y, z, 0;
 * math/storage/vec_storage.hpp:64-71).  *bytes = od * 256; out may be NULL to
 * ask for it.  bvh_unpack_patches reads such records back (*count of them). */
bvh_status bvh_pack_patches(bvh_object o, void *out, uint64_t cap, uint64_t *bytes, uint32_t mem);
bvh_status bvh_unpack_patches(bvh_world w, const void *in, uint64_t bytes, uint32_t mem, bvh_patch_info *out,
                              uint64_t cap, uint64_t *count);
/* The element payload of patches (the bytes of a patch a rank sends to the
 * ranks that ghost it: narrowphase_patch_collection_type::bytes,
 * collision_object/types.hpp:187-201, gathered in patch order as
 * prepare_local_patch_for_sending does, collision_object/impl.hpp:90-124):
 * for each listed patch in list order, each element in patch order, one
 * record of `stride` bytes: u64 global id, then the element's doubles
 * (bvh_element_kind).  bvh_element_record_bytes gives the stride of o's
 * elements; out may be NULL to ask for *bytes.  At most 65535 patches. */
bvh_status bvh_element_record_bytes(bvh_object o, uint64_t *stride);
bvh_status bvh_pack_patch_elements(bvh_object o, const uint64_t *patches, uint64_t npatches, void *out,
                                   uint64_t cap, uint64_t *bytes, uint32_t mem);
/* count element records back into set_entity_data's layout: SoA planes
 * (BVH_ELEM_POLY: 3 * verts_per_element planes; BVH_ELEM_SPHERE: x, y, z, r
 * planes) or, for snapshots, count x 26 k-DOPs and count x 3 centroids;
 * gids (may be NULL) gets the global ids.  All arrays in `mem`. */
bvh_status bvh_unpack_elements(bvh_world w, const void *in, uint64_t count, uint32_t kind,
                               uint32_t verts_per_element, double *data, double *centroids, uint64_t *gids,
                               uint32_t mem);
/* result_msg::serialize (collision_object/types.hpp:109-138): u64 stride
 * (32), u64 byte count, then the 32-byte records of o's results.  *bytes =
 * 16 + 32 * results; out may be NULL to ask for it.  bvh_unpack_results
 * checks the stride and byte count and copies the records out (*count). */
bvh_status bvh_pack_results(bvh_object o, void *out, uint64_t cap, uint64_t *bytes, uint32_t mem);
bvh_status bvh_unpack_results(bvh_world w, const void *in, uint64_t bytes, uint32_t mem, bvh_narrowphase_record *out,
                              uint64_t cap, uint64_t *count);

/* ---- autotuner ------------------------------------------------------------------ */
typedef struct bvh_tuning {
  uint32_t batch;       /* fp64 walk: 1 = one record per visit in scalar registers, 2 (default)
                           or 4 = that many records per memory round trip loaded into LDS */
  uint32_t block_size;  /* narrowphase threads per block: 64, 128 or 256 */
  uint32_t xcd_remap;   /* 1: each XCD walks a contiguous range of query tiles */
  uint32_t tried;       /* configurations timed (entries of tried_ms) */
  double ms;            /* best query time */
  double tried_ms[24];  /* in the order batch (1, 2, 4) x block (64, 128, 256) x remap (off, on) */
} bvh_tuning;
/* Times the query o x other (self collision when o == other, cross-patch if
 * that was o's last query) over the batch
 * (1, 2, 4; fp64 form only: the fp32 walk takes one record per round trip),
 * block size (64, 128, 256) and XCD remap (off, on), best of 3 runs each,
 * keeps the fastest setting for the world and leaves that query's results in
 * o.  The pair set does not depend on the setting.
 * (No reference counterpart: the walk here is wave-wide and record-based,
 * and these are its tuning parameters.) */
bvh_status bvh_autotune(bvh_object o, bvh_object other, bvh_tuning *best);
/* Sets / reads the world's narrowphase launch shape (batch, block_size,
 * xcd_remap of a bvh_tuning; the other fields are ignored / zero).  Default:
 * batch 1, block 256, remap on.  No reference counterpart. */
bvh_status bvh_set_launch(bvh_world w, const bvh_tuning *shape);
bvh_status bvh_get_launch(bvh_world w, bvh_tuning *shape);

/* ---- batching form: a whole step as one HIP graph ----------------------------- */
/* No reference counterpart (the reference launches per call).  Between
 * bvh_capture_begin and bvh_capture_end the world's stream is captured:
 * set_entity_data (device data, no global ids), bvh_refit, init_broadphase,
 * broadphase, broadphase_patches and self_collision record their device work
 * without synchronizing.  Run the same step once before capturing so that
 * every buffer has its size: a call that would grow a buffer fails with
 * BVH_E_STATE inside a capture, and any later (re)allocation in the process
 * makes bvh_graph_launch refuse the graph.  A launched graph re-reads the
 * element data in place; its results are read with the usual calls, which
 * synchronize; a result list that outgrew its buffer reports
 * BVH_E_OVERFLOW (run the step once outside the graph and capture again).
 * sort_results cannot be captured.  Builds of objects built before the
 * capture fork onto their build streams inside it, so the graph has a branch
 * per build.  A graph may outlive its world only to be destroyed. */
bvh_status bvh_capture_begin(bvh_world w);
bvh_status bvh_capture_end(bvh_world w, bvh_graph *graph);
bvh_status bvh_graph_launch(bvh_graph graph);
bvh_status bvh_graph_destroy(bvh_graph graph);

#ifdef __cplusplus
}
#endif
#endif
