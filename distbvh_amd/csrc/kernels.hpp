// kernels.hpp -- host-side launchers of the device kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace dbvh {

struct SortPassInfo {
  uint32_t active, in_sel, out_sel, first;
};
struct SortCtl {
  SortPassInfo pass[8];
  uint32_t tile_counter[8];
  uint32_t offsets[8 * 256];
};
// Buffers of one sort: keys in k0 (not modified), result in ka / va; kb / vb
// are ping-pong scratch.  v0 == nullptr means payload = input index; va ==
// nullptr means keys only.
template <typename K>
struct SortArgs {
  const K* k0;
  K* ka;
  K* kb;
  const uint32_t* v0;
  uint32_t* va;
  uint32_t* vb;
  uint64_t n;
  uint64_t ntiles;
  SortCtl* ctl;
  uint32_t* lookback;  // npass * ntiles * 256 words, zero before the sort
};

uint64_t sort_tiles(uint64_t n);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
hipError_t launch_iota_u32(uint32_t* p, uint64_t n, hipStream_t s);
template <typename T>
hipError_t launch_copy(T* dst, const T* src, uint64_t n, hipStream_t s);
template <typename K>
hipError_t launch_sort_histogram(const K* keys, uint64_t n, int npass, uint32_t* hist, hipStream_t s);
template <typename K>
hipError_t launch_sort_passes(const SortArgs<K>& a, int npass, const uint32_t* hist, hipStream_t s);
// the same launches one at a time: the plan, then pass p (0 .. npass-1) in order
// The build's digit histograms (the Morton kernel's 8 x 256 words of sort
// digits, then the 256 split depths of the cluster kernel): kHistCopies
// copies, block b adding into copy b mod kHistCopies, so that the blocks'
// global adds spread over more addresses; their readers sum the copies.
constexpr int kHistCopies = 8;
constexpr uint32_t kHistWords = 9 * 256;
template <typename K>
hipError_t launch_sort_plan(const SortArgs<K>& a, int npass, const uint32_t* hist, hipStream_t s, int copies = 1,
                            uint32_t copy_stride = 0);
template <typename K>
hipError_t launch_sort_pass(const SortArgs<K>& a, int p, hipStream_t s);
// exclusive scan in place; lb: sort_tiles(n) zeroed 64-bit words, counter: zeroed
hipError_t launch_scan_u32(uint32_t* v, uint64_t n, uint64_t* lb, uint32_t* counter, hipStream_t s);

// ---- object build ---------------------------------------------------------------
// A node handed from one build phase to the next, after it has written its
// half of its parent's record: its range [l, r], its patch and the split
// heights on both sides (d = delta[l-1] | delta[r] << 8, 0 at the ends).
// l == kEnd marks an empty entry.
struct alignas(16) FrontNode {
  uint32_t l, r, patch, d;
};

// leaves per workgroup of the first tree-build phase
constexpr uint32_t kTreeLeafChunk = 128;

struct BuildBufs {
  // original order
  double* centroid;        // 3 planes of n (kPoly)
  double* partials;        // per-block cardinal min/max (6 per block)
  Bounds* bounds;
  // sort
  void* codes0;            // Morton codes, input order (u64 or u32)
  void* codes_a;           // sorted codes
  void* codes_b;
  uint32_t* perm;          // va: sorted position -> input index
  uint32_t* perm_b;
  uint32_t* hist;          // 8 * 256
  SortCtl* ctl;
  uint32_t* lookback;
  // cluster
  uint32_t* depth0;        // n-1 depth keys
  uint32_t* depth_a;
  uint32_t* depth_b;
  uint32_t* depth_idx;
  uint32_t* depth_idx_b;
  uint32_t* splits;        // od - 1
  // tree
  uint8_t* delta;          // n-1
  uint32_t* patch_of_pos;  // n
  uint32_t* patch_of_input;// n
  // n-1 arrival words: 0 before any child arrives, then the first arriver's
  // far end + 1 (high word) and its patch (low word)
  unsigned long long* counters;
  uint32_t* patch_root;    // od
  uint32_t* root;          // 1
  Rec64* rec64;            // fp64 form: n-1 child-pair records
  double* rootbox64;       // fp64 form: the root's slabs (26)
  Rec32* rec32;            // fp32 form: n-1 child-pair records
  int32_t* rootbox32;      // fp32 form: the root's slabs (26, ord32 images)
  double* origin;          // fp32 form: the object frame (origin, 3 doubles; its 13 axis projections)
  uint32_t* status;        // device error bits (kErr*)
  FrontNode* front_a;      // 2 per chunk of kTreeLeafChunk leaves: nodes handed to the next build phase
  FrontNode* front_b;
  unsigned long long* patch_k64;  // od x 26 exact (fp64) patch slabs, ordered_u64 encoding
  // fp32 form: the patch roots' slabs as their records hold them (ord32
  // images), slab-major: slab a of patch k at a * od + k; written where the
  // climb finds each patch root, read by the walks that start at the patches
  int32_t* patch_box32;
  uint64_t od;
};

// device error bits: set instead of faulting when an invariant breaks
constexpr uint32_t kErrTree = 1u;      // inconsistent tree link during the build
constexpr uint32_t kErrTraverse = 2u;  // traversal exceeded its step bound
constexpr uint32_t kErrIndex = 4u;     // element index out of range

hipError_t launch_prep(const ElemInput& in, const BuildBufs& b, hipStream_t s);
hipError_t launch_morton(const ElemInput& in, const BuildBufs& b, int bits, hipStream_t s);
hipError_t launch_cluster(const BuildBufs& b, uint64_t n, int bits, uint64_t nsplits, hipStream_t s);
// k_delta (patch ids, split heights, cleared counters)
hipError_t launch_delta(const ElemInput& in, const BuildBufs& b, int bits, uint64_t od, hipStream_t s);
// k_tree_leaves + k_tree_up (leaf slabs in sorted order + bottom-up build
// with refit); the two parts on their own: the leaf kernel writes every
// leaf's half of its parent's record, the upper phases the nodes above the
// leaf kernel's chunks
hipError_t launch_tree(const ElemInput& in, const BuildBufs& b, uint64_t od, int fp32, hipStream_t s);
hipError_t launch_tree_leaves(const ElemInput& in, const BuildBufs& b, uint64_t od, int fp32, hipStream_t s);
hipError_t launch_tree_up(uint64_t n, const BuildBufs& b, int fp32, hipStream_t s);
// patch_k64 set to the empty k-DOP before the leaf kernel
hipError_t launch_patch_k64_init(unsigned long long* patch_k64, uint64_t od, hipStream_t s);
// fp64 form: the leaves' slabs in sorted order (n x 26), read from their
// parents' records
hipError_t launch_sorted_kdops(const Rec64* rec, const double* rootbox, const uint8_t* delta, uint64_t n, double* out,
                               hipStream_t s);
// exact fp64 patch slabs of an object with fp32-outward records (pk64 first
// set by launch_patch_k64_init), from element data in sorted order
hipError_t launch_patch_slabs(const ElemInput& in, const uint32_t* perm, const uint32_t* patch_of_pos,
                              unsigned long long* pk64, hipStream_t s);
// largest fp32-outward slab error of an object's elements in fp32 ulps (atomic
// max of the float bits into *out, which the caller zeroes)
hipError_t launch_slab_error(const ElemInput& in, const double* org, uint32_t* out, hipStream_t s);
// every leaf's fp64 slabs recomputed from its element, sorted order (n x 26)
hipError_t launch_sorted_elem_kdops(const ElemInput& in, const uint32_t* perm, double* out, hipStream_t s);
uint64_t prep_blocks(uint64_t n);
// the patch-level tree in pre-order: 4 words per node (parent, left, right:
// output indices or -1; patch or -1), *count nodes; stack: 4 od words
hipError_t launch_patch_tree(const BuildBufs& b, uint64_t n, uint64_t od, int fp32, int64_t* out, uint32_t* count,
                             uint64_t* stack, hipStream_t s);

// ---- geom_axis / ml_geom_axis splitters (split.hip) -------------------------------
// element centroids: 3 planes of n (cplanes), or n x 3 (caos)
struct SplitInput {
  const double* cplanes;
  const double* caos;
  uint64_t n;
};
struct SplitBufs {
  uint32_t* rp;        // n: position -> input index (the reference order)
  uint32_t* rp_b;      // n
  uint32_t* scan;      // n + 1: test bits, then their exclusive scan
  uint32_t* tmpf;      // n
  uint32_t* tmpt;      // n
  uint32_t* seg;       // 2 od: range begin at [s], end at [od + s]
  uint32_t* seg_next;  // 2 od
  unsigned long long* seg_k64;  // od x 26: union of the range's element k-DOPs (ordered u64)
  double* seg_acc;     // od x 5: sum, sum of magnitudes, interval lo, hi, exact split point
  int32_t* seg_axis;   // od
  uint32_t* seg_flag;  // od x 2: undecided elements, partition point (kEnd: nothing moves)
  uint64_t* scan_lb;   // sort_tiles(n + 1) words
  uint32_t* scan_ctr;  // 1
  double* xbuf;        // n: projections in range order (ranges that need the exact split point)
};
// The permutation (*result, one of sb.rp / sb.rp_b), the od - 1 split
// offsets in the reference's order and the patch of every input element;
// od a power of two.
hipError_t launch_split(const ElemInput& in, const SplitInput& si, SplitBufs sb, uint64_t od, int ml,
                        uint32_t* splits, uint32_t* patch_key, uint32_t** result, hipStream_t s);
// caller-given patches: patch k = input indices [given[k-1], given[k]); the
// reference order is the input order (*result = sb.rp, the identity)
hipError_t launch_given_split(SplitBufs sb, uint64_t n, uint64_t od, const uint32_t* given, uint32_t* splits,
                              uint32_t* patch_key, uint32_t** result, hipStream_t s);
template <typename T>
hipError_t launch_gather_by(const T* src, const uint32_t* idx, uint64_t n, T* dst, hipStream_t s);

// ---- primitives (primitive.hip) ----------------------------------------------------
struct Tri {
  double a[3], b[3], c[3];
};
// closest point on the triangle for each of n points (n x 3 in, n x 3 out)
hipError_t launch_closest_point(const double* pts, uint64_t n, const Tri& tri, double* out, hipStream_t s);
// closed triangle / ball intersection of n pairs: p, q n x 9 (triangles a, b,
// c) or n x 4 (x, y, z, r), out[i] in {0, 1}
hipError_t launch_tri_pairs(const double* p, const double* q, uint64_t n, uint8_t* out, hipStream_t s);
hipError_t launch_sphere_pairs(const double* a, const double* b, uint64_t n, uint8_t* out, hipStream_t s);
// result pairs whose primitives meet (tri: triangles of 3-vertex polytope
// planes, else spheres): flag[i] = 0 / 1, then after an exclusive scan of
// the flags (off), the kept pairs in order into out and their count into *kept
hipError_t launch_pair_flags(const uint64_t* pairs, uint64_t count, const ElemInput& qa, const ElemInput& tb, int tri,
                             uint32_t* flag, hipStream_t s);
hipError_t launch_pair_compact(const uint64_t* pairs, uint64_t count, const ElemInput& qa, const ElemInput& tb,
                               int tri, const uint32_t* off, uint64_t* out, unsigned long long* kept, hipStream_t s);

// ---- wire formats (wire.hip) ---------------------------------------------------------
constexpr uint64_t kPatchRecordBytes = 256;  // u64 id, u64 size, 26 f64 slabs, 4 f64 centroid storage
constexpr uint64_t kResultRecordBytes = 32;  // detailed_narrowphase_result
hipError_t launch_pack_patches(const unsigned long long* patch_k64, const uint32_t* splits, const uint32_t* roots,
                               const double* cent, uint64_t n, uint64_t od, void* out, hipStream_t s);
// src[offset_k + i] = ref_perm[begin_k + i] for each listed patch k (range:
// {begin, end, offset} per patch); max_len = the longest range
hipError_t launch_patch_positions(const uint32_t* ref_perm, const uint64_t* range, uint64_t npatch,
                                  uint64_t max_len, uint32_t* src, hipStream_t s);
// element records {u64 global id, `planes` doubles} of the elements src[0..count)
hipError_t launch_pack_elements(const ElemInput& in, const uint64_t* gids, const uint32_t* src, uint64_t count,
                                uint32_t planes, void* out, hipStream_t s);
hipError_t launch_unpack_elements(const void* in, uint64_t count, uint32_t planes, int snapshot, double* data,
                                  double* cent, uint64_t* gids, hipStream_t s);
// result message header: stride, byte count
hipError_t launch_result_header(void* out, uint64_t count, hipStream_t s);

// ---- sort and sweep between two patches (sweep.hip) --------------------------------
struct SweepArgs {
  const double* ka;            // A's k-DOPs, na x 26
  const uint64_t* ga;          // A's global ids (null: local indices)
  uint64_t na;
  const double* kbs;           // B's k-DOPs in sorted order, nb x 26
  const double* bmax;          // B's scaled maxima on the axis, sorted order
  const uint32_t* bq;          // B's truncated minima, sorted
  const uint32_t* bidx;        // sorted position -> B index
  const uint64_t* gb;          // B's global ids (null: local indices)
  uint64_t nb;
  int axis;
  double gmin, conv;
  uint32_t* count;             // per a: count (count pass), exclusive offset (write pass)
  unsigned long long* total;   // zeroed before the count pass
  uint64_t* out;               // 2 words per pair
};
// max_variant_axis: partials (variant_axis_partials() x 6 doubles), axis out
hipError_t launch_variant_axis(const double* ca, uint64_t na, const double* cb, uint64_t nb, double* part,
                               int32_t* axis, hipStream_t s);
uint32_t variant_axis_partials();
hipError_t launch_sweep_prepare(const double* kb, uint64_t nb, int axis, double gmin, double conv, uint32_t* bq,
                                uint32_t* bidx, hipStream_t s);
hipError_t launch_sweep_gather(const double* kb, const uint32_t* bidx, uint64_t nb, int axis, double gmin,
                               double conv, double* kbs, double* bmax, hipStream_t s);
hipError_t launch_sweep(const SweepArgs& a, bool write, hipStream_t s);
// 26-DOPs of in.n polytopes of in.nv vertices (kPoly planes), n x 26
hipError_t launch_element_kdops(const ElemInput& in, double* out, hipStream_t s);

// ---- queries ----------------------------------------------------------------------
struct TreeView {
  const Rec64* rec64;
  const double* rootbox64;
  const Rec32* rec32;
  const int32_t* rootbox32;
  const uint8_t* delta;  // split heights (locate a leaf's slabs in its parent's record)
  const uint32_t* root;
  const uint32_t* patch_root;
  const uint32_t* patch_of_input;
  const uint32_t* perm;
  uint64_t n;
  uint64_t od;
  uint32_t* status;
  const double* origin;  // fp32 form: frame origin (3 doubles)
  ElemInput elem;        // the object's element data (fp32 form re-checks in fp64)
  const unsigned long long* patch_k64;  // exact patch slabs (ordered_u64)
  const uint32_t* splits;               // od - 1 patch starts in tree-order positions
  const int32_t* patch_box32;           // fp32 form: the patch roots' slabs, slab-major (a * od + k)
};

struct PairOut {
  uint64_t* pairs;      // (input index a << 32) | input index b
  uint64_t cap;
  unsigned long long* count;
  unsigned long long* visits;      // optional: node visits summed over waves
  unsigned long long* candidates;  // optional: fp32 candidate pairs before the fp64 check
  // fp32 form: candidates too close to call in fp32, re-checked in fp64.
  // The list is kRecheckStripes stripes of rcap entries; wave w appends to
  // stripe w % kRecheckStripes through that stripe's counter,
  // rcount[stripe * kRecheckStride], so no single counter takes every
  // flush.  The re-check kernel adds rtotal[0] = sum and rtotal[1] = largest
  // of the stripe counts (a stripe count above rcap is an overflow).
  uint64_t* recheck;
  uint64_t rcap;
  unsigned long long* rcount;
  unsigned long long* rtotal;
};

constexpr uint32_t kRecheckStripes = 64;
constexpr uint32_t kRecheckStride = 16;  // counter words: one 128-byte line per stripe

// launch shape of the narrowphase (tuned by the autotuner): threads per
// block (64, 128 or 256), whether blocks are remapped so each XCD walks a
// contiguous range of query tiles, and the fp64 walk's batch: 1 = one
// record per visit in scalar registers (k_narrow64s), 2 or 4 = that many
// records per memory round trip loaded into LDS (k_narrow64; 2 by default,
// the fastest on dense contact).  The fp32 walk always takes one record in
// scalar registers.
struct LaunchShape {
  unsigned block = 256;
  bool xcd_remap = true;
  unsigned batch = 2;
};

// Tree forms: fp64 slabs in every record (kFormRecords64); fp32-outward
// records (kFormFp32 and kFormMixed, the default fp64 form): the same walk,
// which keeps the leaf candidates whose fp32 gap proves the fp64 overlap and
// decides the rest on fp64 k-DOPs, so both give the fp64 pair set; the fp64
// form also serves exact fp64 element k-DOPs (bvh_object_sorted_kdops).
constexpr int kFormRecords64 = 0;
constexpr int kFormFp32 = 1;
constexpr int kFormMixed = 2;
// element narrowphase: every leaf of `q` (queries, in its sorted order)
// against tree `t`; self = 1 emits each unordered pair of one tree once,
// self = 2 only those whose elements lie in different patches.
hipError_t launch_narrowphase(const TreeView& q, const TreeView& t, int leaf_size, int self,
                              int form, const LaunchShape& shape, const PairOut& out, hipStream_t s);
// One query of a batched narrowphase launch: tree views, outputs and the
// blocks (256 threads, one query element each) [block0, block0 + blocks)
struct BatchQuery {
  TreeView q;
  TreeView t;
  PairOut out;
  uint32_t block0;
  uint32_t blocks;
};
// the pair walks of nq queries (no self, no XCD remap) in one launch of
// `blocks` blocks of 256 threads; bq in device memory; batch 2 or 4 for the
// walk over fp64 records.  The fp32-outward forms leave their re-check
// lists to launch_recheck_query, one launch per query.
// a query's ncount counter words, then the two objects' status words (4 x
// u32 each, 8-byte aligned), into host (pinned, device-accessible) memory
hipError_t launch_readback(const unsigned long long* count, int ncount, const uint32_t* st_a, const uint32_t* st_b,
                           unsigned long long* host, hipStream_t s);
hipError_t launch_narrowphase_batch(const BatchQuery* bq, uint32_t nq, uint32_t blocks, int batch, int form,
                                    hipStream_t s);
hipError_t launch_recheck_query(const TreeView& q, const TreeView& t, const PairOut& out, hipStream_t s);
// patch broadphase: patch roots of `q` against the patch level of `t`
hipError_t launch_patch_broadphase(const TreeView& q, const TreeView& t, int fp32, const PairOut& out,
                                   hipStream_t s);
// records {patch_p, element_p, patch_q, element_q}
hipError_t launch_records(const uint64_t* pairs, uint64_t count, const uint32_t* patch_a,
                          const uint64_t* gid_a, const uint32_t* patch_b, const uint64_t* gid_b,
                          uint64_t patch_offset_a, uint64_t patch_offset_b, uint64_t* out,
                          hipStream_t s);
// out[0] += sum of mix(key), out[1] ^= xor of mix'(key) (out zeroed by the caller)
hipError_t launch_digest(const uint64_t* pairs, uint64_t count, unsigned long long* out, hipStream_t s);
hipError_t launch_patch_centroids(const double* centroid_planes, const double* sphere_planes,
                                  const double* snap_centroids, const uint32_t* perm,
                                  const uint32_t* splits, uint64_t n, uint64_t od, double* out,
                                  hipStream_t s);

}  // namespace dbvh
