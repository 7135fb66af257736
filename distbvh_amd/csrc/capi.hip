// capi.hip -- the C ABI (include/distbvh.h): worlds, collision objects and
// the stream-ordered orchestration of the device stages.
//
// Reference call stacks this replaces (single process, one GPU):
//  * collision_object::set_entity_data_clustering (collision_object.hpp:89-131)
//    + set_entity_data_impl (collision_object.cpp:108-146)
//      -> prep -> morton -> sort -> cluster -> tree
//  * init_broadphase (collision_object.cpp:148-215): the patch tree is part of
//    the object tree, so this only validates state
//  * broadphase + narrowphase (collision_object.cpp:248-397)
//      -> patch broadphase kernel + element narrowphase kernel, results in
//         the initiating object (impl.cpp:328-348)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>
#include <atomic>
#include <mutex>

#include "../../include/distbvh.h"
#include "kernels.hpp"

using namespace dbvh;

namespace {

thread_local std::string g_err;

bvh_status fail(bvh_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

// On failure the runtime's last-error state is cleared too, so that the
// error is reported once, by this call, and not again by the caller's next
// runtime call (torch checks hipGetLastError after its own launches).
#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      (void)hipGetLastError();                                                              \
      return fail(_e == hipErrorOutOfMemory ? BVH_E_OOM : BVH_E_HIP,                        \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                        \
    }                                                                                       \
  } while (0)

#define BVH_TRY(expr)                    \
  do {                                   \
    bvh_status _s = (expr);              \
    if (_s != BVH_OK) return _s;         \
  } while (0)

// grow-only device buffer
// Every device (re)allocation bumps the generation of the world that owns
// the buffer; a graph captured at an older generation of its world may hold
// freed pointers and is refused.  Buffers of no world (temporaries of one
// call) count on a process-wide generation that no graph reads.
std::atomic<uint64_t> g_alloc_gen{0};
thread_local bool t_capturing = false;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  std::atomic<uint64_t>* gen = &g_alloc_gen;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  // frees the memory (a later ensure allocates again); graphs captured on it
  // become stale
  void release() {
    if (!p) return;
    gen->fetch_add(1);
    (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  bvh_status ensure(size_t want) {
    if (want <= bytes && p) return BVH_OK;
    if (t_capturing)
      return fail(BVH_E_STATE, "a device buffer would grow inside a capture: run the step once before capturing");
    gen->fetch_add(1);
    if (p) {
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    size_t b = want ? want : 256;
    b = (b + 255) & ~size_t(255);
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(BVH_E_OOM, "hipMalloc(" + std::to_string(b) + "): " + hipGetErrorString(e));
    }
    bytes = b;
    return BVH_OK;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Stage {
  const char* name;
  hipEvent_t a, b;
  hipStream_t s;  // the stream its events are recorded on
  bool open;      // begun, not ended
};

}  // namespace

struct bvh_object_s;
// one or a few launches of a build, run on the object's stream
using BuildStep = std::function<bvh_status()>;

// The host-side state of a query launched inside a capture: which partner
// and kind, the list capacities its kernels were launched with and where its
// pairs land.  A replay of the graph restores it in the object, so results
// read after the replay describe the captured query even if the object ran
// other queries eagerly since the capture.
struct CapturedQuery {
  bvh_object_s* o;
  bvh_object_s* other;
  bool self, narrow, self_result, cross;
  uint64_t pcap, cap, rcap;
  const uint64_t* result_ptr;
};

struct bvh_world_s {
  size_t od = 1;
  bvh_world_config cfg{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<std::unique_ptr<bvh_object_s>> objects;
  // timing
  std::vector<Stage> stages;
  std::vector<hipEvent_t> event_pool;
  size_t events_used = 0;
  // primitive-call scratch
  DevBuf s_kb, s_ka, s_va, s_vb, s_hist, s_ctl, s_lb;
  DevBuf prim_in, prim_out;  // host-memory staging of bvh_closest_point
  // bvh_max_variant_axis / bvh_sort_and_sweep: staging, sorted B, counts
  DevBuf sw_ka, sw_kb, sw_ga, sw_gb, sw_bq, sw_bidx, sw_kbs, sw_bmax, sw_count, sw_total, sw_out, sw_part;
  LaunchShape shape;  // narrowphase launch shape (autotuned)
  // stream capture (batching form): the queries captured so far, with the
  // state a replay restores in their objects
  bool capturing = false;
  std::vector<CapturedQuery> captured;
  std::vector<bvh_graph> graphs;  // live graphs captured on this world
  unsigned long long* h_count = nullptr;  // pinned, 16 words: query counters and status words
  bool iteration_open = false;
  // allocation generation of this world's buffers (its objects' and its own)
  std::atomic<uint64_t> alloc_gen{0};
  // batched queries: the table on the device and its pinned host staging
  DevBuf bq_dev;
  BatchQuery* bq_host = nullptr;
  size_t bq_host_cap = 0;

  // Only buffers that captured work can reference count on the world's
  // generation (its objects' buffers).  The stand-alone primitives' and the
  // batched query's buffers serve calls that are never captured, so their
  // growth leaves the world's graphs launchable.

  hipEvent_t take_event() {
    if (events_used == event_pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      event_pool.push_back(e);
    }
    return event_pool[events_used++];
  }
  void reset_stages() {
    stages.clear();
    events_used = 0;
  }
  void stage_begin(const char* name) {
    if (!(cfg.flags & BVH_FLAG_TIMING) || capturing || stages.size() >= 512) return;
    Stage s{name, take_event(), take_event(), stream, true};
    if (s.a && s.b) {
      (void)hipEventRecord(s.a, stream);
      stages.push_back(s);
    }
  }
  // ends the newest open stage of the current stream (the pending builds of
  // several objects interleave their stages, each on its own stream)
  void stage_end() {
    if (!(cfg.flags & BVH_FLAG_TIMING) || capturing) return;
    for (size_t i = stages.size(); i-- > 0;)
      if (stages[i].open && stages[i].s == stream) {
        (void)hipEventRecord(stages[i].b, stream);
        stages[i].open = false;
        return;
      }
  }
};

struct bvh_object_s {
  bvh_world w = nullptr;
  size_t id = 0;
  // input
  ElemInput in{};
  bool have_data = false;
  bool built = false;
  bool has_gids = false;
  DevBuf input;  // uploaded host elements (or snapshot kdops)
  DevBuf input_c;  // uploaded snapshot centroids
  DevBuf gids;
  // build buffers
  DevBuf centroid, partials, bounds, codes0, codes_a, codes_b, perm, perm_b, hist, ctl,
      lookback, depth0, depth_a, depth_b, depth_idx, depth_idx_b, splits, delta, patch_of_pos,
      patch_of_input, counters, patch_root, root, rec64, rootbox64, rec32, rootbox32, status,
      front_a, front_b, patch_k64, patch_box32;
  DevBuf frame;  // fp32-outward records: the frame origin of this object's slabs (3 doubles), its 13 projections
  DevBuf slab_err;  // the slab-error statistic (one word)
  BuildBufs bb{};
  int bits = 63;
  int fp32 = 0;  // fp32-outward tree records (forms kFormFp32 and kFormMixed)
  int form = kFormMixed;
  // geom_axis / ml_geom_axis: the splitter's buffers and the reference order
  // (position -> input index) that patch centroids and patch elements follow;
  // the tree's own order (perm) is patch-major, Morton order inside a patch
  uint32_t split = BVH_SPLIT_CLUSTERING;
  // BVH_SPLIT_GIVEN: the caller's patch offsets (od - 1, input order)
  DevBuf given;
  std::vector<uint64_t> given_h;
  DevBuf sp_rp, sp_rp_b, sp_scan, sp_tmpf, sp_tmpt, sp_seg, sp_seg_next, sp_k64, sp_acc, sp_axis, sp_flag, sp_lb,
      sp_ctr, sp_key, key0, key_a, key_b, sp_hist, sp_x;
  const uint32_t* ref_rp = nullptr;
  const uint32_t* ref_perm() const { return split == BVH_SPLIT_CLUSTERING ? perm.as<uint32_t>() : ref_rp; }
  // results
  DevBuf pairs, pairs_a, pairs_b, records, ppairs, sort_hist, sort_ctl, sort_lb, count, recheck, digest;
  // exact primitive filter of the results: flags / offsets, kept pairs, scan words
  DevBuf ex_flag, ex_out, ex_lb, ex_ctr;
  // wire formats: patch centroids, patch ranges, element sources, host staging
  DevBuf wire_cent, wire_range, wire_src, wire_out;
  uint64_t npairs = 0;
  uint64_t npatch_pairs = 0;
  // exact fp64 patch slabs valid for the current build (the fp64 records'
  // build writes them; the fp32-outward records' are reduced on demand), and
  // the last query's patch pairs computed (on demand after a narrowphase)
  bool patch_slabs_ready = false;
  bool patch_pairs_ready = false;
  uint64_t nvisits = 0;
  uint64_t ncandidates = 0;
  uint64_t nrecheck = 0;      // fp32 candidates re-checked in fp64
  uint64_t nrecheck_max = 0;  // the fullest re-check stripe
  const uint64_t* result_ptr = nullptr;  // device pointer to the (sorted) pairs
  bvh_object_s* result_other = nullptr;
  bool self_result = false;
  bool have_results = false;
  bool results_pending = false;  // a captured query: counters not read yet
  // the last launched query: kind and the list capacities it ran with
  bool q_self = false, q_narrow = false, q_cross = false;
  uint64_t q_pcap = 0, q_cap = 0, q_rcap = 0;
  // Build stream: a build forks from the world stream onto it (so the builds
  // of several objects run concurrently) and the world stream waits for it
  // (ev_done) at the next call that is not a set_entity_data.
  hipStream_t bs = nullptr;
  hipEvent_t ev_fork = nullptr, ev_done = nullptr;
  bool pending = false;
  // A forked build enqueues its launches later, at the next call that reads
  // an object (build_pending, steps: the launches before the tree stage):
  // the pending builds' steps interleaved, then each tree stage after the
  // steps of all of them (ev_sorted), so that the latency-bound sorts of
  // several objects share the GPU and no sort waits behind another object's
  // tree kernel.
  std::vector<BuildStep> steps;
  size_t steps_done = 0;  // steps already launched (the first ones go at the call)
  hipEvent_t ev_sorted = nullptr;
  bool build_pending = false;
  // The flush runs the pending builds' leaf kernels one after another (each
  // object's upper tree phases overlap the next one's leaf kernel); ev_leaves
  // follows this build's leaf kernel (leaves_event: recorded for the pending
  // build), which is all a query walk needs of its query object.
  hipEvent_t ev_leaves = nullptr;
  bool leaves_event = false;

  bvh_object_s() = default;
  bvh_object_s(const bvh_object_s&) = delete;
  bvh_object_s& operator=(const bvh_object_s&) = delete;
  ~bvh_object_s() {
    if (bs) {
      (void)hipStreamSynchronize(bs);
      (void)hipStreamDestroy(bs);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_done) (void)hipEventDestroy(ev_done);
    if (ev_sorted) (void)hipEventDestroy(ev_sorted);
    if (ev_leaves) (void)hipEventDestroy(ev_leaves);
  }

  // every device buffer of the object
  template <typename F>
  void for_each_buf(F f) {
    DevBuf* all[] = {&input, &input_c, &gids, &centroid, &partials, &bounds, &codes0, &codes_a, &codes_b,
                     &perm, &perm_b, &hist, &ctl, &lookback, &depth0, &depth_a, &depth_b, &depth_idx,
                     &depth_idx_b, &splits, &delta, &patch_of_pos, &patch_of_input, &counters, &patch_root,
                     &root, &rec64, &rootbox64, &rec32, &rootbox32, &status, &frame, &slab_err,
                     &front_a, &front_b, &patch_k64, &patch_box32, &pairs, &pairs_a, &pairs_b, &records, &ppairs,
                     &sort_hist, &sort_ctl, &sort_lb, &count, &recheck, &digest, &sp_rp, &sp_rp_b,
                     &sp_scan, &sp_tmpf, &sp_tmpt, &sp_seg, &sp_seg_next, &sp_k64, &sp_acc, &sp_axis,
                     &sp_flag, &sp_lb, &sp_ctr, &sp_key, &key0, &key_a, &key_b, &sp_hist, &sp_x,
                     &ex_flag, &ex_out, &ex_lb, &ex_ctr, &wire_cent, &wire_range, &wire_src, &wire_out,
                     &given};
    for (DevBuf* d : all) f(*d);
  }
  size_t device_bytes() {
    size_t t = 0;
    for_each_buf([&](DevBuf& d) { t += d.bytes; });
    return t;
  }
};

// ---- helpers ---------------------------------------------------------------------------
// The world stream waits for o's build (if one is still pending).
static bvh_status join(bvh_object_s* o) {
  if (!o->pending) return BVH_OK;
  o->pending = false;
  HIP_TRY(hipStreamWaitEvent(o->w->stream, o->ev_done, 0));
  return BVH_OK;
}
static bvh_status flush_builds(bvh_world w, const bvh_object_s* first = nullptr);
// ... for every object of the world, after enqueuing the tree stages still
// pending: every call that reads objects' device data or may be followed by
// the caller's use of it starts with this.
static bvh_status join_all(bvh_world w) {
  BVH_TRY(flush_builds(w));
  for (auto& o : w->objects) BVH_TRY(join(o.get()));
  return BVH_OK;
}
// Points the world's stream at another stream for the scope (a build enqueued
// on an object's build stream); restores it on every exit path.
struct StreamSwap {
  bvh_world w;
  hipStream_t keep;
  StreamSwap(bvh_world w_, hipStream_t s) : w(w_), keep(w_->stream) { w->stream = s; }
  ~StreamSwap() { w->stream = keep; }
};

static TreeView view_of(const bvh_object_s* o) {
  TreeView v{};
  v.rec64 = o->rec64.as<Rec64>();
  v.rootbox64 = o->rootbox64.as<double>();
  v.rec32 = o->rec32.as<Rec32>();
  v.delta = o->delta.as<uint8_t>();
  v.rootbox32 = o->rootbox32.as<int32_t>();
  v.root = o->root.as<uint32_t>();
  v.patch_root = o->patch_root.as<uint32_t>();
  v.patch_of_input = o->patch_of_input.as<uint32_t>();
  v.perm = o->perm.as<uint32_t>();
  v.n = o->in.n;
  v.od = o->w->od;
  v.status = o->status.as<uint32_t>();
  v.origin = o->fp32 ? o->frame.as<double>() : nullptr;
  v.elem = o->in;
  v.patch_k64 = o->patch_k64.as<unsigned long long>();
  v.patch_box32 = o->fp32 ? o->patch_box32.as<int32_t>() : nullptr;
  v.splits = o->splits.as<uint32_t>();
  return v;
}

static bool is_pow2(size_t v) { return v && !(v & (v - 1)); }

static bvh_status copy_out(void* dst, const void* src, size_t bytes, uint32_t mem, hipStream_t s) {
  if (bytes == 0) return BVH_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, mem == BVH_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return BVH_OK;
}

static bvh_status object_alloc(bvh_object_s* o) {
  const uint64_t n = o->in.n;
  const uint64_t od = o->w->od;
  const uint64_t n1 = n > 0 ? n : 1;
  const uint64_t tiles = sort_tiles(n1);
  const size_t kw = o->bits == 63 ? 8 : 4;
  BVH_TRY(o->centroid.ensure(o->in.kind == kPoly ? 3 * n1 * sizeof(double) : 256));
  BVH_TRY(o->partials.ensure(prep_blocks(n1) * 6 * sizeof(double)));
  BVH_TRY(o->bounds.ensure(sizeof(Bounds)));
  BVH_TRY(o->codes0.ensure(n1 * kw));
  BVH_TRY(o->codes_a.ensure(n1 * kw));
  BVH_TRY(o->codes_b.ensure(n1 * kw));
  BVH_TRY(o->perm.ensure(n1 * 4));
  BVH_TRY(o->perm_b.ensure(n1 * 4));
  BVH_TRY(o->hist.ensure(kHistCopies * kHistWords * 4));
  BVH_TRY(o->ctl.ensure(2 * sizeof(SortCtl)));
  BVH_TRY(o->lookback.ensure(9 * tiles * 256 * 4));
  BVH_TRY(o->depth0.ensure(n1 * 4));
  BVH_TRY(o->depth_a.ensure(n1 * 4));
  BVH_TRY(o->depth_b.ensure(n1 * 4));
  BVH_TRY(o->depth_idx.ensure(n1 * 4));
  BVH_TRY(o->depth_idx_b.ensure(n1 * 4));
  BVH_TRY(o->splits.ensure((od > 1 ? od - 1 : 1) * 4));
  BVH_TRY(o->delta.ensure(n1));
  BVH_TRY(o->patch_of_pos.ensure(n1 * 4));
  BVH_TRY(o->patch_of_input.ensure(n1 * 4));
  BVH_TRY(o->counters.ensure(n1 * sizeof(unsigned long long)));
  BVH_TRY(o->patch_root.ensure(od * 4));
  BVH_TRY(o->root.ensure(4));
  BVH_TRY(o->status.ensure(4 * sizeof(uint32_t)));  // [error bits, -, max fp32 slab error, -]
  const uint64_t nchunks = (n1 + kTreeLeafChunk - 1) / kTreeLeafChunk;
  BVH_TRY(o->front_a.ensure((2 * nchunks + 2) * sizeof(FrontNode)));
  BVH_TRY(o->front_b.ensure((2 * nchunks + 2) * sizeof(FrontNode)));
  BVH_TRY(o->patch_k64.ensure(od * kSlabs * sizeof(unsigned long long)));
  if (o->fp32) {
    BVH_TRY(o->rec32.ensure((n1 > 1 ? n1 - 1 : 1) * sizeof(Rec32)));
    BVH_TRY(o->rootbox32.ensure(kSlabs * sizeof(int32_t)));
    BVH_TRY(o->frame.ensure(16 * sizeof(double)));
    BVH_TRY(o->patch_box32.ensure(od * kSlabs * sizeof(int32_t)));
  } else {
    BVH_TRY(o->rec64.ensure((n1 > 1 ? n1 - 1 : 1) * sizeof(Rec64)));
    BVH_TRY(o->rootbox64.ensure(kSlabs * sizeof(double)));
  }
  BuildBufs& b = o->bb;
  b.centroid = o->centroid.as<double>();
  b.partials = o->partials.as<double>();
  b.bounds = o->bounds.as<Bounds>();
  b.codes0 = o->codes0.p;
  b.codes_a = o->codes_a.p;
  b.codes_b = o->codes_b.p;
  b.perm = o->perm.as<uint32_t>();
  b.perm_b = o->perm_b.as<uint32_t>();
  b.hist = o->hist.as<uint32_t>();
  b.ctl = o->ctl.as<SortCtl>();
  b.lookback = o->lookback.as<uint32_t>();
  b.depth0 = o->depth0.as<uint32_t>();
  b.depth_a = o->depth_a.as<uint32_t>();
  b.depth_b = o->depth_b.as<uint32_t>();
  b.depth_idx = o->depth_idx.as<uint32_t>();
  b.depth_idx_b = o->depth_idx_b.as<uint32_t>();
  b.splits = o->splits.as<uint32_t>();
  b.delta = o->delta.as<uint8_t>();
  b.patch_of_pos = o->patch_of_pos.as<uint32_t>();
  b.patch_of_input = o->patch_of_input.as<uint32_t>();
  b.counters = o->counters.as<unsigned long long>();
  b.patch_root = o->patch_root.as<uint32_t>();
  b.root = o->root.as<uint32_t>();
  b.rec64 = o->fp32 ? nullptr : o->rec64.as<Rec64>();
  b.rootbox64 = o->fp32 ? nullptr : o->rootbox64.as<double>();
  b.rec32 = o->fp32 ? o->rec32.as<Rec32>() : nullptr;
  b.rootbox32 = o->fp32 ? o->rootbox32.as<int32_t>() : nullptr;
  b.origin = o->fp32 ? o->frame.as<double>() : nullptr;
  b.status = o->status.as<uint32_t>();
  b.front_a = o->front_a.as<FrontNode>();
  b.front_b = o->front_b.as<FrontNode>();
  b.patch_k64 = o->patch_k64.as<unsigned long long>();
  b.patch_box32 = o->fp32 ? o->patch_box32.as<int32_t>() : nullptr;
  b.od = od;
  if (o->split != BVH_SPLIT_CLUSTERING) {
    BVH_TRY(o->sp_rp.ensure(n1 * 4));
    BVH_TRY(o->sp_rp_b.ensure(n1 * 4));
    BVH_TRY(o->sp_scan.ensure((n1 + 1) * 4));
    BVH_TRY(o->sp_tmpf.ensure(n1 * 4));
    BVH_TRY(o->sp_tmpt.ensure(n1 * 4));
    BVH_TRY(o->sp_seg.ensure(2 * od * 4));
    BVH_TRY(o->sp_seg_next.ensure(2 * od * 4));
    BVH_TRY(o->sp_k64.ensure(od * kSlabs * 8));
    BVH_TRY(o->sp_acc.ensure(od * 5 * sizeof(double)));
    BVH_TRY(o->sp_axis.ensure(od * 4));
    BVH_TRY(o->sp_flag.ensure(od * 2 * 4));
    BVH_TRY(o->sp_lb.ensure(sort_tiles(n1 + 1) * 8));
    BVH_TRY(o->sp_ctr.ensure(4));
    BVH_TRY(o->sp_key.ensure(n1 * 4));
    BVH_TRY(o->key0.ensure(n1 * 4));
    BVH_TRY(o->key_a.ensure(n1 * 4));
    BVH_TRY(o->key_b.ensure(n1 * 4));
    BVH_TRY(o->sp_hist.ensure(2 * 256 * 4));
    BVH_TRY(o->sp_x.ensure(n1 * sizeof(double)));
  }
  return BVH_OK;
}

// geom_axis / ml_geom_axis / given patches: the reference order and split
// offsets from the splitter (given: the caller's offsets, input order), then
// the tree order: the Morton order (already in perm and
// codes_a) stably sorted by patch, and the codes gathered into that order, so
// that every patch is a contiguous, Morton-sorted range of the tree order.
static bvh_status split_and_order(bvh_object_s* o) {
  bvh_world w = o->w;
  hipStream_t s = w->stream;
  const uint64_t n = o->in.n, od = w->od;
  SplitInput si{};
  si.n = n;
  if (o->in.kind == kPoly) si.cplanes = o->centroid.as<double>();
  else if (o->in.kind == kSphere) si.cplanes = o->in.data;  // planes x, y, z, r
  else si.caos = o->in.centroids;
  SplitBufs sb{};
  sb.rp = o->sp_rp.as<uint32_t>();
  sb.rp_b = o->sp_rp_b.as<uint32_t>();
  sb.scan = o->sp_scan.as<uint32_t>();
  sb.tmpf = o->sp_tmpf.as<uint32_t>();
  sb.tmpt = o->sp_tmpt.as<uint32_t>();
  sb.seg = o->sp_seg.as<uint32_t>();
  sb.seg_next = o->sp_seg_next.as<uint32_t>();
  sb.seg_k64 = o->sp_k64.as<unsigned long long>();
  sb.seg_acc = o->sp_acc.as<double>();
  sb.seg_axis = o->sp_axis.as<int32_t>();
  sb.seg_flag = o->sp_flag.as<uint32_t>();
  sb.scan_lb = o->sp_lb.as<uint64_t>();
  sb.scan_ctr = o->sp_ctr.as<uint32_t>();
  sb.xbuf = o->sp_x.as<double>();
  uint32_t* rp = nullptr;
  uint32_t* patch_key = o->sp_key.as<uint32_t>();
  if (o->split == BVH_SPLIT_GIVEN)
    HIP_TRY(launch_given_split(sb, n, od, o->given.as<uint32_t>(), o->bb.splits, patch_key, &rp, s));
  else
    HIP_TRY(launch_split(o->in, si, sb, od, o->split == BVH_SPLIT_ML_GEOM_AXIS ? 1 : 0, o->bb.splits, patch_key, &rp,
                         s));
  o->ref_rp = rp;
  // stable sort of the Morton order by patch
  HIP_TRY(launch_gather_by<uint32_t>(patch_key, o->bb.perm, n, o->key0.as<uint32_t>(), s));
  const int npass = od <= 256 ? 1 : (od <= 65536 ? 2 : 4);
  HIP_TRY(launch_fill_u32(o->sp_hist.as<uint32_t>(), 2 * 256, 0u, s));
  HIP_TRY(launch_fill_u32(o->bb.lookback, (uint64_t)(npass > 2 ? 2 : npass) * sort_tiles(n) * 256, 0u, s));
  if (npass > 2) return fail(BVH_E_UNSUPPORTED, "geom_axis splitting supports at most 65536 patches");
  HIP_TRY(launch_sort_histogram<uint32_t>(o->key0.as<uint32_t>(), n, npass, o->sp_hist.as<uint32_t>(), s));
  SortArgs<uint32_t> a{};
  a.k0 = o->key0.as<uint32_t>(); a.ka = o->key_a.as<uint32_t>(); a.kb = o->key_b.as<uint32_t>();
  a.v0 = o->bb.perm; a.va = o->bb.perm_b; a.vb = sb.tmpf;  // tmpf is free once the splitter is done
  a.n = n; a.ntiles = sort_tiles(n); a.ctl = o->bb.ctl + 1; a.lookback = o->bb.lookback;
  HIP_TRY(launch_sort_passes<uint32_t>(a, npass, o->sp_hist.as<uint32_t>(), s));
  HIP_TRY(launch_copy<uint32_t>(o->bb.perm, o->bb.perm_b, n, s));
  if (o->bits == 63)
    HIP_TRY(launch_gather_by<uint64_t>((const uint64_t*)o->bb.codes0, o->bb.perm, n, (uint64_t*)o->bb.codes_a, s));
  else
    HIP_TRY(launch_gather_by<uint32_t>((const uint32_t*)o->bb.codes0, o->bb.perm, n, (uint32_t*)o->bb.codes_a, s));
  return BVH_OK;
}

static bvh_status object_build_tree(bvh_object_s* o);

// steps of a deferred build launched by its set_entity_data call itself
constexpr size_t kEagerSteps = 2;

// The launches of a build up to (not including) its tree stage, as steps run
// in order on the object's stream: the bounds, the Morton codes, the sort's
// plan and each of its passes, the patch splits, the frame and the split
// heights.  Host state that later calls read is set when the steps are made.
static void object_build_steps(bvh_object_s* o, std::vector<BuildStep>& st) {
  bvh_world w = o->w;
  const hipStream_t s = w->stream;
  const uint64_t n = o->in.n;
  st.clear();
  st.push_back([=] {
    // (the bounds kernel also clears the status words)
    w->stage_begin("bounds");
    HIP_TRY(launch_prep(o->in, o->bb, s));
    w->stage_end();
    return BVH_OK;
  });
  st.push_back([=] {
    w->stage_begin("morton");
    HIP_TRY(launch_morton(o->in, o->bb, o->bits, s));
    w->stage_end();
    return BVH_OK;
  });
  const int npass = o->bits == 63 ? 8 : 4;
  for (int p = -1; p < npass; ++p) {
    st.push_back([=] {
      if (p < 0) w->stage_begin("sort");
      if (o->bits == 63) {
        SortArgs<uint64_t> a{};
        a.k0 = (const uint64_t*)o->bb.codes0; a.ka = (uint64_t*)o->bb.codes_a; a.kb = (uint64_t*)o->bb.codes_b;
        a.v0 = nullptr; a.va = o->bb.perm; a.vb = o->bb.perm_b;
        a.n = n; a.ntiles = sort_tiles(n); a.ctl = o->bb.ctl; a.lookback = o->bb.lookback;
        HIP_TRY(p < 0 ? launch_sort_plan<uint64_t>(a, npass, o->bb.hist, s, kHistCopies, kHistWords)
                      : launch_sort_pass<uint64_t>(a, p, s));
      } else {
        SortArgs<uint32_t> a{};
        a.k0 = (const uint32_t*)o->bb.codes0; a.ka = (uint32_t*)o->bb.codes_a; a.kb = (uint32_t*)o->bb.codes_b;
        a.v0 = nullptr; a.va = o->bb.perm; a.vb = o->bb.perm_b;
        a.n = n; a.ntiles = sort_tiles(n); a.ctl = o->bb.ctl; a.lookback = o->bb.lookback;
        HIP_TRY(p < 0 ? launch_sort_plan<uint32_t>(a, npass, o->bb.hist, s, kHistCopies, kHistWords)
                      : launch_sort_pass<uint32_t>(a, p, s));
      }
      if (p == npass - 1) w->stage_end();
      return BVH_OK;
    });
  }
  st.push_back([=] {
    if (o->split == BVH_SPLIT_CLUSTERING) {
      w->stage_begin("cluster");
      HIP_TRY(launch_cluster(o->bb, n, o->bits, w->od - 1, s));
      w->stage_end();
    } else {
      w->stage_begin("split");
      BVH_TRY(split_and_order(o));
      w->stage_end();
    }
    // (the fp32 frame, the Morton bounds minimum of this build, is written
    // with the bounds)
    if (!o->fp32) HIP_TRY(launch_patch_k64_init(o->bb.patch_k64, w->od, s));  // written at the patch roots
    w->stage_begin("delta");
    HIP_TRY(launch_delta(o->in, o->bb, o->bits, w->od, s));
    w->stage_end();
    return BVH_OK;
  });
  o->patch_slabs_ready = !o->fp32;
}

// The whole object build, stream ordered, no host synchronization.  With
// defer its launches are left to flush_builds (the object's build_pending),
// which interleaves them with the other pending builds'.
static bvh_status object_build(bvh_object_s* o, bool defer) {
  bvh_world w = o->w;
  hipStream_t s = w->stream;
  const uint64_t n = o->in.n;
  BVH_TRY(object_alloc(o));
  o->steps.clear();
  if (n == 0) {
    HIP_TRY(launch_fill_u32(o->bb.status, 4, 0u, s));
    // splits are all 0 (split/cluster.hpp:101-105), every patch empty
    HIP_TRY(launch_fill_u32(o->bb.splits, w->od > 1 ? w->od - 1 : 0, 0u, s));
    HIP_TRY(launch_fill_u32(o->bb.patch_root, w->od, kEnd, s));
    HIP_TRY(launch_fill_u32(o->bb.root, 1, kEnd, s));
    o->patch_slabs_ready = false;  // every patch empty: the init alone
    o->built = true;
    return BVH_OK;
  }
  object_build_steps(o, o->steps);
  o->built = true;
  if (defer) {
    // the first steps (bounds, Morton codes) are launched at once, so the GPU
    // works while the caller makes its next calls; the rest wait for the flush
    o->steps_done = 0;
    for (; o->steps_done < kEagerSteps && o->steps_done < o->steps.size(); ++o->steps_done)
      BVH_TRY(o->steps[o->steps_done]());
    o->build_pending = true;
    return BVH_OK;
  }
  for (BuildStep& f : o->steps) BVH_TRY(f());
  o->steps.clear();
  return object_build_tree(o);
}

// The tree stage of a build: leaf slabs and the bottom-up climb into records.
static bvh_status object_build_tree(bvh_object_s* o) {
  bvh_world w = o->w;
  w->stage_begin("tree");
  HIP_TRY(launch_tree(o->in, o->bb, w->od, o->fp32, w->stream));
  w->stage_end();
  return BVH_OK;
}

// Enqueues the world's pending (forked) builds: their steps interleaved, one
// step of each build in turn, each on its build stream, so that every build
// has work on the GPU from the start and the latency-bound sort passes of
// several builds run side by side; then the tree stages, after the steps of
// all of them (no sort pass waits behind another build's tree kernel); and
// moves each build's ev_done behind its tree.  The trees' leaf kernels run
// one after another (two side by side took longer than one after the other,
// the vector memory units being their bound), `first` (a query's target)
// first, and each build's upper tree phases, latency-bound and narrow,
// overlap the next build's leaf kernel.
static bvh_status flush_builds(bvh_world w, const bvh_object_s* first) {
  std::vector<bvh_object_s*> p;
  for (auto& o : w->objects)
    if (o->build_pending) p.push_back(o.get());
  if (p.empty()) return BVH_OK;
  for (size_t j = 1; j < p.size(); ++j)
    if (p[j] == first) std::swap(p[0], p[j]);
  std::vector<bvh_status> sa(p.size(), BVH_OK);
  size_t nsteps = 0;
  for (bvh_object_s* a : p) nsteps = std::max(nsteps, a->steps.size() - a->steps_done);
  for (size_t i = 0; i < nsteps; ++i)
    for (size_t j = 0; j < p.size(); ++j) {
      const size_t k = p[j]->steps_done + i;
      if (k < p[j]->steps.size() && sa[j] == BVH_OK) {
        StreamSwap swap(w, p[j]->bs);
        sa[j] = p[j]->steps[k]();
      }
    }
  for (bvh_object_s* a : p) HIP_TRY(hipEventRecord(a->ev_sorted, a->bs));
  // (a query's target starts its tree as soon as its own steps are done: the
  // other builds' leaf kernels come after its leaf kernel anyway)
  for (bvh_object_s* a : p)
    for (bvh_object_s* b : p)
      if (a != b && a != first) HIP_TRY(hipStreamWaitEvent(a->bs, b->ev_sorted, 0));
  bvh_status st = BVH_OK;
  const bvh_object_s* prev = nullptr;  // the build whose leaf kernel runs before the next one's
  for (size_t j = 0; j < p.size(); ++j) {
    bvh_object_s* a = p[j];
    a->build_pending = false;
    a->steps.clear();
    if (sa[j] == BVH_OK) {
      StreamSwap swap(w, a->bs);
      if (prev) HIP_TRY(hipStreamWaitEvent(a->bs, prev->ev_leaves, 0));
      w->stage_begin("tree");
      hipError_t e = launch_tree_leaves(a->in, a->bb, w->od, a->fp32, a->bs);
      if (e == hipSuccess) e = hipEventRecord(a->ev_leaves, a->bs);
      if (e == hipSuccess) e = launch_tree_up(a->in.n, a->bb, a->fp32, a->bs);
      w->stage_end();
      if (e != hipSuccess) {
        sa[j] = fail(BVH_E_HIP, std::string("tree stage: ") + hipGetErrorString(e));
        (void)hipGetLastError();
      } else {
        a->leaves_event = true;
        prev = a;
      }
    }
    // recorded also after a failure, so that nothing enqueued is left unordered
    HIP_TRY(hipEventRecord(a->ev_done, a->bs));
    if (st == BVH_OK) st = sa[j];
  }
  return st;
}

// The world stream before a query of o against another object: every pending
// build flushed (the target's tree first), the target's and every other
// build joined, and o's joined only up to its leaf kernel when that was
// flushed (the walk reads o's leaf slabs, positions and permutation, all in
// place by then, and nothing its upper tree phases write): those phases run
// beside the walk, and run_query joins them before it reads the counters.
static bvh_status join_for_query(bvh_object_s* o, bvh_object_s* other) {
  bvh_world w = o->w;
  BVH_TRY(flush_builds(w, other));
  for (auto& x : w->objects) {
    bvh_object_s* a = x.get();
    if (a == o && a != other && a->pending && a->leaves_event)
      HIP_TRY(hipStreamWaitEvent(w->stream, a->ev_leaves, 0));  // o stays pending
    else
      BVH_TRY(join(a));
  }
  return BVH_OK;
}

// Refit: new element k-DOPs on the topology of the last build (same
// permutation, patches and split heights); arrival words and roots are
// cleared and the bottom-up pass runs again.
static bvh_status object_refit(bvh_object_s* o) {
  bvh_world w = o->w;
  hipStream_t s = w->stream;
  const uint64_t n = o->in.n;
  HIP_TRY(launch_fill_u32(o->bb.status, 4, 0u, s));
  if (n == 0) return BVH_OK;
  w->stage_begin("refit");
  if (n > 1) HIP_TRY(launch_fill_u32((uint32_t*)o->bb.counters, 2 * (n - 1), 0u, s));
  HIP_TRY(launch_fill_u32(o->bb.patch_root, w->od, kEnd, s));
  HIP_TRY(launch_fill_u32(o->bb.root, 1, kEnd, s));
  if (!o->fp32) HIP_TRY(launch_patch_k64_init(o->bb.patch_k64, w->od, s));
  o->patch_slabs_ready = !o->fp32;
  HIP_TRY(launch_tree(o->in, o->bb, w->od, o->fp32, s));
  w->stage_end();
  return BVH_OK;
}

static bvh_status set_entity_build(bvh_object_s* o, int kind, int nv, double eps, const double* data,
                                   const double* cent, uint64_t n, const uint64_t* gids, uint32_t mem,
                                   uint32_t split, hipStream_t bstream, bool defer);

static bvh_status set_entity_common(bvh_object_s* o, int kind, int nv, double eps, const double* data,
                                    const double* cent, uint64_t n, const uint64_t* gids, uint32_t mem,
                                    uint32_t split) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  if (split != BVH_SPLIT_CLUSTERING && split != BVH_SPLIT_GEOM_AXIS && split != BVH_SPLIT_ML_GEOM_AXIS &&
      split != BVH_SPLIT_GIVEN)
    return fail(BVH_E_INVALID, "unknown split algorithm");
  if (split == BVH_SPLIT_GIVEN) {
    if (o->given_h.size() + 1 != o->w->od)
      return fail(BVH_E_STATE, "BVH_SPLIT_GIVEN needs bvh_object_set_patch_offsets first");
    if (!o->given_h.empty() && o->given_h.back() > n)
      return fail(BVH_E_INVALID, "a patch offset lies beyond the element count");
  }
  // geom_axis splits log2(od) times: od must be a power of two
  // (collision_object.cpp:122-124 asserts splits + 1 == od)
  if ((split == BVH_SPLIT_GEOM_AXIS || split == BVH_SPLIT_ML_GEOM_AXIS) && !is_pow2(o ? o->w->od : 0))
    return fail(BVH_E_INVALID, "geom_axis / ml_geom_axis need a power-of-two overdecomposition factor");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  if (n > 0 && !data) return fail(BVH_E_INVALID, "null element data");
  if (n >= (1ull << 30)) return fail(BVH_E_INVALID, "at most 2^30 - 1 elements per object");
  bvh_world w = o->w;
  if (o->pending) {
    // a build still pending is superseded by this one (nothing read it; its
    // launches not made yet are dropped); the previous build may still use the
    // buffers this one may
    // reallocate (inside a capture nothing is reallocated, and nothing may
    // synchronize)
    o->build_pending = false;
    o->steps.clear();
    if (!w->capturing) HIP_TRY(hipStreamSynchronize(o->bs));
    BVH_TRY(join(o));
  }
  // The build runs on the object's own stream, after everything enqueued on
  // the world stream so far (inside a capture the build stream joins the
  // capture there and the graph gets a branch per build), unless builds are
  // serial.
  const bool fork = !(w->cfg.flags & BVH_FLAG_SERIAL_BUILDS) &&
                    !(w->capturing && !o->bs);  // no stream is created inside a capture
  if (fork) {
    if (!o->bs) {
      HIP_TRY(hipStreamCreateWithFlags(&o->bs, hipStreamNonBlocking));
      HIP_TRY(hipEventCreateWithFlags(&o->ev_fork, hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&o->ev_done, hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&o->ev_sorted, hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&o->ev_leaves, hipEventDisableTiming));
    }
    HIP_TRY(hipEventRecord(o->ev_fork, w->stream));
    HIP_TRY(hipStreamWaitEvent(o->bs, o->ev_fork, 0));
  }
  // Inside a capture the build is not deferred: deferral makes the build
  // streams wait on each other directly, and hipStreamEndCapture crashed in
  // the host runtime on a graph holding such waits between two builds.  The
  // capture keeps the fork / join pattern only (a graph branch per build).
  const bool defer = fork && !w->capturing;
  o->leaves_event = false;  // recorded again by the flush of a deferred build
  const bvh_status st =
      set_entity_build(o, kind, nv, eps, data, cent, n, gids, mem, split, fork ? o->bs : w->stream, defer);
  if (fork) {
    if (st != BVH_OK) {
      o->build_pending = false;
      o->steps.clear();
    }
    // the world stream waits for this at the next call that needs it (also
    // after a failed build, so that nothing it enqueued is left unordered)
    HIP_TRY(hipEventRecord(o->ev_done, o->bs));
    o->pending = true;
  }
  return st;
}

static bvh_status set_entity_build(bvh_object_s* o, int kind, int nv, double eps, const double* data,
                                   const double* cent, uint64_t n, const uint64_t* gids, uint32_t mem,
                                   uint32_t split, hipStream_t bstream, bool defer) {
  bvh_world w = o->w;
  StreamSwap swap(w, bstream);
  hipStream_t s = w->stream;
  o->have_results = false;
  o->in.kind = kind;
  o->in.nv = nv;
  o->in.eps = eps;
  o->in.n = n;
  o->bits = (int)w->cfg.morton_bits;
  o->form = w->cfg.precision == BVH_PREC_FP32_OUTWARD ? kFormFp32
                                                       : (w->cfg.precision == BVH_PREC_FP64_RECORDS ? kFormRecords64
                                                                                                    : kFormMixed);
  o->fp32 = o->form != kFormRecords64 ? 1 : 0;
  o->split = split;
  size_t planes = kind == kPoly ? (size_t)nv * 3 : (kind == kSphere ? 4 : kSlabs);
  size_t bytes = planes * n * sizeof(double);
  if (w->capturing && (mem == BVH_MEM_HOST || gids))
    return fail(BVH_E_STATE, "a capture takes device element data without global ids");
  if (mem == BVH_MEM_HOST && n > 0) {
    w->stage_begin("upload");
    BVH_TRY(o->input.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(o->input.p, data, bytes, hipMemcpyHostToDevice, s));
    o->in.data = o->input.as<double>();
    if (kind == kSnapshot) {
      BVH_TRY(o->input_c.ensure(3 * n * sizeof(double)));
      HIP_TRY(hipMemcpyAsync(o->input_c.p, cent, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
      o->in.centroids = o->input_c.as<double>();
    }
    w->stage_end();
  } else {
    // device data is read in place; like the reference's element pointer
    // (collision_object.cpp:128-129) it must stay valid until the object's
    // next set_entity_data
    o->in.data = data;
    o->in.centroids = cent;
  }
  o->has_gids = gids != nullptr;
  if (gids && n > 0) {
    BVH_TRY(o->gids.ensure(n * sizeof(uint64_t)));
    HIP_TRY(hipMemcpyAsync(o->gids.p, gids, n * sizeof(uint64_t),
                           mem == BVH_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
  }
  o->have_data = true;
  o->built = false;
  o->build_pending = false;
  o->steps.clear();
  return object_build(o, defer);
}

// Counter words of a query (device, o->count): pairs, node visits, fp32
// candidates, fp32 re-checks (sum and largest stripe), patch pairs; from word
// kCntStripes on, the fp32 re-check list's stripe counters, one 128-byte line
// each.
constexpr int kCntPairs = 0, kCntVisits = 1, kCntCand = 2, kCntRecheck = 3, kCntRecheckMax = 4, kCntPatch = 5,
              kCntWords = 6;
constexpr uint64_t kCntStripes = 16;
constexpr uint64_t kCntTotal = kCntStripes + (uint64_t)kRecheckStripes * kRecheckStride;

// Reads the counters of the last query of `o` and the device error words of
// both objects (one synchronization).  Sets *overflow when a list outgrew
// its buffer (the query must run again with larger buffers).
static bvh_status read_counts(bvh_object_s* o, bvh_object_s* other, bool* overflow) {
  bvh_world w = o->w;
  hipStream_t s = w->stream;
  const unsigned long long* dcount = o->count.as<unsigned long long>();
  HIP_TRY(launch_readback(dcount, kCntWords, o->status.as<uint32_t>(), other->status.as<uint32_t>(), w->h_count, s));
  HIP_TRY(hipStreamSynchronize(s));
  o->npairs = o->q_narrow ? w->h_count[kCntPairs] : 0;
  o->nvisits = w->h_count[kCntVisits];
  o->ncandidates = w->h_count[kCntCand];
  o->nrecheck = w->h_count[kCntRecheck];
  o->nrecheck_max = w->h_count[kCntRecheckMax];
  o->npatch_pairs = o->q_self ? 0 : w->h_count[kCntPatch];
  uint32_t sw[8];
  std::memcpy(sw, &w->h_count[kCntWords], sizeof(sw));
  const uint32_t st = sw[0] | sw[4];
  if (st) return fail(BVH_E_HIP, "device invariant check failed (status bits " + std::to_string(st) + ")");
  *overflow = o->npairs > o->q_cap || o->npatch_pairs > o->q_pcap || (o->fp32 && o->nrecheck_max > o->q_rcap);
  return BVH_OK;
}

static bvh_status ensure_pairs(bvh_object_s* o, uint64_t cap) {
  BVH_TRY(o->pairs.ensure(cap * sizeof(uint64_t)));
  return BVH_OK;
}

// Stable sort of the result pairs by (element_p, element_q).
static bvh_status sort_results(bvh_object_s* o) {
  hipStream_t s = o->w->stream;
  const uint64_t n = o->npairs;
  if (n < 2) return BVH_OK;
  // the sort's look-back words carry 30-bit counts
  if (n >= (1ull << 30)) return fail(BVH_E_UNSUPPORTED, "sort_results holds at most 2^30 - 1 pairs");
  BVH_TRY(o->pairs_a.ensure(n * 8));
  BVH_TRY(o->pairs_b.ensure(n * 8));
  BVH_TRY(o->sort_hist.ensure(8 * 256 * 4));
  BVH_TRY(o->sort_ctl.ensure(sizeof(SortCtl)));
  const uint64_t tiles = sort_tiles(n);
  BVH_TRY(o->sort_lb.ensure(8 * tiles * 256 * 4));
  HIP_TRY(launch_fill_u32(o->sort_hist.as<uint32_t>(), 8 * 256, 0u, s));
  HIP_TRY(launch_fill_u32(o->sort_lb.as<uint32_t>(), 8 * tiles * 256, 0u, s));
  HIP_TRY(launch_sort_histogram<uint64_t>(o->pairs.as<uint64_t>(), n, 8, o->sort_hist.as<uint32_t>(), s));
  SortArgs<uint64_t> a{};
  a.k0 = o->pairs.as<uint64_t>(); a.ka = o->pairs_a.as<uint64_t>(); a.kb = o->pairs_b.as<uint64_t>();
  a.v0 = nullptr; a.va = nullptr; a.vb = nullptr;
  a.n = n; a.ntiles = tiles; a.ctl = o->sort_ctl.as<SortCtl>(); a.lookback = o->sort_lb.as<uint32_t>();
  HIP_TRY(launch_sort_passes<uint64_t>(a, 8, o->sort_hist.as<uint32_t>(), s));
  o->result_ptr = o->pairs_a.as<uint64_t>();
  return BVH_OK;
}

// The exact fp64 patch slabs of the current build, reduced from element data
// when the tree holds fp32-outward records (stream ordered; the element data
// must still be valid).
static bvh_status ensure_patch_slabs(bvh_object_s* o) {
  if (o->patch_slabs_ready) return BVH_OK;
  hipStream_t s = o->w->stream;
  HIP_TRY(launch_patch_k64_init(o->bb.patch_k64, o->w->od, s));
  if (o->in.n > 0)
    HIP_TRY(launch_patch_slabs(o->in, o->perm.as<uint32_t>(), o->patch_of_pos.as<uint32_t>(),
                               o->patch_k64.as<unsigned long long>(), s));
  o->patch_slabs_ready = true;
  return BVH_OK;
}

// The patch pairs of o's last query (its patch broadphase), computed when
// something asks for them: the built-in narrowphase walks every element and
// does not need them.  Synchronizes.
static bvh_status ensure_patch_pairs(bvh_object_s* o) {
  if (!o->have_results || o->patch_pairs_ready) return BVH_OK;
  bvh_object_s* t = o->result_other;
  if (o->self_result || !t) {
    o->npatch_pairs = 0;
    o->patch_pairs_ready = true;
    return BVH_OK;
  }
  if (o->w->capturing) return fail(BVH_E_STATE, "patch pairs are read after the capture ends");
  if (!t->built) return fail(BVH_E_STATE, "the other object was released or rebuilt");
  hipStream_t s = o->w->stream;
  BVH_TRY(ensure_patch_slabs(o));
  BVH_TRY(ensure_patch_slabs(t));
  unsigned long long* dcount = o->count.as<unsigned long long>();
  for (int attempt = 0; attempt < 2; ++attempt) {
    o->q_pcap = o->ppairs.bytes / 8;
    HIP_TRY(launch_fill_u32((uint32_t*)(dcount + kCntPatch), 2, 0u, s));
    PairOut po{o->ppairs.as<uint64_t>(), o->q_pcap, dcount + kCntPatch, nullptr, nullptr, nullptr, 0, nullptr, nullptr};
    HIP_TRY(launch_patch_broadphase(view_of(o), view_of(t), o->fp32, po, s));
    unsigned long long cnt = 0;
    BVH_TRY(copy_out(&cnt, dcount + kCntPatch, sizeof(cnt), BVH_MEM_HOST, s));
    o->npatch_pairs = cnt;
    if (cnt <= o->q_pcap) break;
    if (attempt == 1) return fail(BVH_E_OVERFLOW, "patch pair list still overflows after growing");
    BVH_TRY(o->ppairs.ensure(cnt * 8 + 1024 * 8));
  }
  o->patch_pairs_ready = true;
  return BVH_OK;
}

// Launches a query with the current buffer sizes; no synchronization (so it
// can be captured into a graph).  The counters stay on the device.
static bvh_status launch_query(bvh_object_s* o, bvh_object_s* other, bool self, bool narrow, bool cross) {
  bvh_world w = o->w;
  hipStream_t s = w->stream;
  unsigned long long* dcount = o->count.as<unsigned long long>();
  const TreeView qv = view_of(o), tv = view_of(other);
  o->q_self = self;
  o->q_narrow = narrow;
  o->q_cross = self && cross;
  o->q_pcap = o->ppairs.bytes / 8;
  o->q_cap = o->pairs.bytes / 8;
  o->q_rcap = o->recheck.bytes / 8 / kRecheckStripes;  // per stripe
  HIP_TRY(launch_fill_u32((uint32_t*)dcount, 2 * kCntTotal, 0u, s));
  // the patch broadphase runs here only for a broadphase without the element
  // narrowphase (the patch pairs are its result); after a narrowphase it
  // runs when its pairs are asked for (ensure_patch_pairs)
  o->patch_pairs_ready = self;
  if (!self && !narrow) {
    w->stage_begin("broadphase");
    BVH_TRY(ensure_patch_slabs(o));
    BVH_TRY(ensure_patch_slabs(other));
    PairOut po{o->ppairs.as<uint64_t>(), o->q_pcap, dcount + kCntPatch, nullptr, nullptr, nullptr, 0, nullptr, nullptr};
    HIP_TRY(launch_patch_broadphase(qv, tv, o->fp32, po, s));
    o->patch_pairs_ready = true;
    w->stage_end();
  }
  if (narrow) {
    w->stage_begin("narrowphase");
    PairOut po{o->pairs.as<uint64_t>(), o->q_cap, dcount + kCntPairs, dcount + kCntVisits, dcount + kCntCand,
               o->recheck.as<uint64_t>(), o->q_rcap, dcount + kCntStripes, dcount + kCntRecheck};
    HIP_TRY(launch_narrowphase(qv, tv, (int)w->cfg.leaf_size, self ? (cross ? 2 : 1) : 0, o->form, w->shape, po, s));
    w->stage_end();
  }
  o->result_ptr = o->pairs.as<uint64_t>();
  o->result_other = other;
  o->self_result = self;
  return BVH_OK;
}

// After a query whose lists overflowed: every list grown to what the query
// needs.  Candidates the re-check could not hold were not decided, and each
// of them can add one pair, so the pair list is sized for them too.
static bvh_status grow_query_lists(bvh_object_s* o) {
  const bool rover = o->fp32 && o->nrecheck_max > o->q_rcap;
  const uint64_t need = o->npairs + (rover ? o->nrecheck : 0);
  if (o->npatch_pairs > o->q_pcap) BVH_TRY(o->ppairs.ensure(o->npatch_pairs * 8 + 1024 * 8));
  if (need > o->q_cap) BVH_TRY(ensure_pairs(o, need + (need >> 3) + 1024));
  if (rover)  // every stripe as large as the fullest one needs
    BVH_TRY(o->recheck.ensure((o->nrecheck_max + (o->nrecheck_max >> 3) + 64) * kRecheckStripes * 8));
  return BVH_OK;
}

static bvh_status run_query(bvh_object_s* o, bvh_object_s* other, bool self, bool narrow = true, bool cross = false) {
  bvh_world w = o->w;
  if (!o->built || !other->built) return fail(BVH_E_STATE, "set_entity_data must precede broadphase");
  o->have_results = false;
  o->results_pending = false;
  if (w->capturing && (w->cfg.flags & BVH_FLAG_SORT_RESULTS))
    return fail(BVH_E_STATE, "sort_results cannot be captured (it needs the pair count on the host)");
  // buffer sizes: grown here, outside any capture (a captured query uses them as they are)
  BVH_TRY(o->count.ensure(kCntTotal * sizeof(unsigned long long)));
  BVH_TRY(o->ppairs.ensure(std::max<uint64_t>(1024, 4 * w->od) * 8));
  if (narrow) {
    BVH_TRY(ensure_pairs(o, std::max<uint64_t>(1u << 20, 8 * o->in.n)));
    if (o->fp32) BVH_TRY(o->recheck.ensure(std::max<uint64_t>(1u << 16, o->in.n) * 8));
  }
  if (w->capturing) {
    BVH_TRY(launch_query(o, other, self, narrow, cross));
    // nothing to read until the graph has run (bvh_graph_launch sets these)
    o->have_results = false;
    o->results_pending = false;
    w->captured.push_back(CapturedQuery{o, other, self, narrow, self, self && cross, o->q_pcap, o->q_cap, o->q_rcap,
                                        o->result_ptr});
    return BVH_OK;
  }
  for (int attempt = 0; attempt < 3; ++attempt) {
    BVH_TRY(launch_query(o, other, self, narrow, cross));
    BVH_TRY(join(o));  // (a query object's upper tree phases may have run beside the walk)
    bool overflow = false;
    BVH_TRY(read_counts(o, other, &overflow));
    if (!overflow) break;
    if (attempt == 2) return fail(BVH_E_OVERFLOW, "result lists still overflow after growing");
    BVH_TRY(grow_query_lists(o));
  }
  if (narrow && (w->cfg.flags & BVH_FLAG_SORT_RESULTS)) {
    w->stage_begin("sort_results");
    BVH_TRY(sort_results(o));
    w->stage_end();
  }
  o->have_results = true;
  return BVH_OK;
}

// Several object pairs o[i] x other[i] in one coalesced step: their counters
// cleared and patch broadphases run per pair, then every element
// narrowphase in one launch (k_narrow64_batch / k_narrow32_batch), then one synchronization
// for all their counters.  A pair whose lists overflowed runs again on its
// own with grown buffers.  The fp32-outward forms add one re-check launch per
// pair after the walk.
static bvh_status run_query_batch(bvh_object_s* const* os, bvh_object_s* const* others, uint64_t k) {
  bvh_world w = os[0]->w;
  hipStream_t s = w->stream;
  for (uint64_t i = 0; i < k; ++i) {
    bvh_object_s* o = os[i];
    if (!o->built || !others[i]->built) return fail(BVH_E_STATE, "set_entity_data must precede broadphase");
    o->have_results = false;
    o->results_pending = false;
    BVH_TRY(o->count.ensure(kCntTotal * sizeof(unsigned long long)));
    BVH_TRY(o->ppairs.ensure(std::max<uint64_t>(1024, 4 * w->od) * 8));
    BVH_TRY(ensure_pairs(o, std::max<uint64_t>(1u << 20, 8 * o->in.n)));
    if (o->fp32) BVH_TRY(o->recheck.ensure(std::max<uint64_t>(1u << 16, o->in.n) * 8));
  }
  if (k > w->bq_host_cap) {
    if (w->bq_host) (void)hipHostFree(w->bq_host);
    w->bq_host = nullptr;
    w->bq_host_cap = 0;
    void* hp = nullptr;
    HIP_TRY(hipHostMalloc(&hp, k * sizeof(BatchQuery), hipHostMallocDefault));
    w->bq_host = static_cast<BatchQuery*>(hp);
    w->bq_host_cap = k;
  }
  BVH_TRY(w->bq_dev.ensure(k * sizeof(BatchQuery)));
  // the previous batch's table may still be in flight to the device
  HIP_TRY(hipStreamSynchronize(s));
  uint32_t blocks = 0;
  uint32_t nq = 0;
  for (uint64_t i = 0; i < k; ++i) {
    bvh_object_s* o = os[i];
    bvh_object_s* t = others[i];
    unsigned long long* dcount = o->count.as<unsigned long long>();
    o->q_self = false;
    o->q_narrow = true;
    o->q_cross = false;
    o->q_pcap = o->ppairs.bytes / 8;
    o->q_cap = o->pairs.bytes / 8;
    o->q_rcap = o->fp32 ? o->recheck.bytes / 8 / kRecheckStripes : 0;  // per stripe
    o->result_ptr = o->pairs.as<uint64_t>();
    o->result_other = t;
    o->self_result = false;
    HIP_TRY(launch_fill_u32((uint32_t*)dcount, 2 * kCntTotal, 0u, s));
    const TreeView qv = view_of(o), tv = view_of(t);
    o->patch_pairs_ready = false;  // computed when asked for (ensure_patch_pairs)
    if (o->in.n == 0 || t->in.n == 0) continue;
    BatchQuery& e = w->bq_host[nq++];
    e.q = qv;
    e.t = tv;
    e.out = o->fp32 ? PairOut{o->pairs.as<uint64_t>(), o->q_cap, dcount + kCntPairs, dcount + kCntVisits,
                              dcount + kCntCand, o->recheck.as<uint64_t>(), o->q_rcap, dcount + kCntStripes,
                              dcount + kCntRecheck}
                    : PairOut{o->pairs.as<uint64_t>(), o->q_cap, dcount + kCntPairs, dcount + kCntVisits,
                              dcount + kCntCand, nullptr, 0, nullptr, nullptr};
    e.block0 = blocks;
    e.blocks = (uint32_t)((o->in.n + 255) / 256);
    blocks += e.blocks;
  }
  if (nq) {
    w->stage_begin("narrowphase_batch");
    HIP_TRY(hipMemcpyAsync(w->bq_dev.p, w->bq_host, nq * sizeof(BatchQuery), hipMemcpyHostToDevice, s));
    HIP_TRY(launch_narrowphase_batch(w->bq_dev.as<BatchQuery>(), nq, blocks, w->shape.batch == 4 ? 4 : 2,
                                     os[0]->form, s));
    // the fp32-outward forms: each query's re-check list, decided in fp64
    for (uint32_t j = 0; j < nq && os[0]->fp32; ++j)
      HIP_TRY(launch_recheck_query(w->bq_host[j].q, w->bq_host[j].t, w->bq_host[j].out, s));
    w->stage_end();
  }
  // every pair's counters and both objects' error words, one synchronization
  std::vector<unsigned long long> h(12 * k);
  for (uint64_t i = 0; i < k; ++i) {
    HIP_TRY(hipMemcpyAsync(&h[12 * i], os[i]->count.p, kCntWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&h[12 * i + kCntWords], os[i]->status.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&h[12 * i + kCntWords + 2], others[i]->status.p, 4 * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  for (uint64_t i = 0; i < k; ++i) {
    bvh_object_s* o = os[i];
    const unsigned long long* c = &h[12 * i];
    o->npairs = c[kCntPairs];
    o->nvisits = c[kCntVisits];
    o->ncandidates = c[kCntCand];
    o->nrecheck = o->fp32 ? c[kCntRecheck] : 0;
    o->nrecheck_max = o->fp32 ? c[kCntRecheckMax] : 0;
    o->npatch_pairs = c[kCntPatch];
    uint32_t sw[8];
    std::memcpy(sw, &c[kCntWords], sizeof(sw));
      if (sw[0] | sw[4])
      return fail(BVH_E_HIP, "device invariant check failed (status bits " + std::to_string(sw[0] | sw[4]) + ")");
    if (o->npairs > o->q_cap || o->npatch_pairs > o->q_pcap || (o->fp32 && o->nrecheck_max > o->q_rcap)) {
      // this pair alone again, with the lists grown to its counts
      BVH_TRY(grow_query_lists(o));
      BVH_TRY(run_query(o, others[i], false));
      continue;
    }
    o->have_results = true;
  }
  return BVH_OK;
}

// Results of a captured query become known on the host at the first access
// after the graph ran.
static bvh_status results_ready(bvh_object_s* o) {
  if (!o->have_results || !o->results_pending) return BVH_OK;
  if (o->w->capturing) return fail(BVH_E_STATE, "results are read after the capture ends");
  bool overflow = false;
  BVH_TRY(read_counts(o, o->result_other, &overflow));
  if (overflow) {
    o->have_results = false;
    return fail(BVH_E_OVERFLOW, "a captured query outgrew its buffers; run the step once outside the graph");
  }
  o->results_pending = false;
  return BVH_OK;
}

// ---- stand-alone primitives ------------------------------------------------------------------
// Scratch of one stand-alone sort: key and value ping-pong buffers (n keys,
// n values each), the digit histograms of every pass (8 x 256 words), the
// sort control block and the look-back words (npass x tiles x 256 words).
struct SortScratch {
  void* ka;
  void* kb;
  uint32_t* va;
  uint32_t* vb;
  uint32_t* hist;
  SortCtl* ctl;
  uint32_t* lb;
};

// Stable sort of n keys (with optional values) in place, in the scratch
// `sc`; the histograms and look-back words are cleared here by fill kernels,
// the control block by the sort's plan kernel.  Waits for the result.
template <typename K>
static bvh_status sort_in_place(bvh_world w, K* keys, uint32_t* vals, uint64_t n, const SortScratch& sc) {
  hipStream_t s = w->stream;
  const int npass = sizeof(K);
  const uint64_t tiles = sort_tiles(n);
  HIP_TRY(launch_fill_u32(sc.hist, 8 * 256, 0u, s));
  HIP_TRY(launch_fill_u32(sc.lb, npass * tiles * 256, 0u, s));
  HIP_TRY(launch_sort_histogram<K>(keys, n, npass, sc.hist, s));
  SortArgs<K> a{};
  a.k0 = keys; a.ka = static_cast<K*>(sc.ka); a.kb = static_cast<K*>(sc.kb);
  a.v0 = vals; a.va = vals ? sc.va : nullptr; a.vb = vals ? sc.vb : nullptr;
  a.n = n; a.ntiles = tiles; a.ctl = sc.ctl; a.lookback = sc.lb;
  HIP_TRY(launch_sort_passes<K>(a, npass, sc.hist, s));
  HIP_TRY(hipMemcpyAsync(keys, a.ka, n * sizeof(K), hipMemcpyDeviceToDevice, s));
  if (vals) HIP_TRY(hipMemcpyAsync(vals, a.va, n * 4, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  return BVH_OK;
}

template <typename K>
static bvh_status prim_sort(bvh_world w, K* keys, uint32_t* vals, uint64_t n) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  if (w->capturing) return fail(BVH_E_STATE, "sort inside a capture");
  if (n == 0) return BVH_OK;
  if (!keys) return fail(BVH_E_INVALID, "null keys");
  if (n >= (1ull << 30)) return fail(BVH_E_INVALID, "at most 2^30 - 1 keys");
  const int npass = sizeof(K);
  const uint64_t tiles = sort_tiles(n);
  BVH_TRY(w->s_ka.ensure(n * sizeof(K)));
  BVH_TRY(w->s_kb.ensure(n * sizeof(K)));
  BVH_TRY(w->s_va.ensure(n * 4));
  BVH_TRY(w->s_vb.ensure(n * 4));
  BVH_TRY(w->s_hist.ensure(8 * 256 * 4));
  BVH_TRY(w->s_ctl.ensure(sizeof(SortCtl)));
  BVH_TRY(w->s_lb.ensure(npass * tiles * 256 * 4));
  const SortScratch sc{w->s_ka.p, w->s_kb.p, w->s_va.as<uint32_t>(), w->s_vb.as<uint32_t>(),
                       w->s_hist.as<uint32_t>(), w->s_ctl.as<SortCtl>(), w->s_lb.as<uint32_t>()};
  return sort_in_place<K>(w, keys, vals, n, sc);
}

// ---- the sort that allocates nothing (bvh_sort_u64_noalloc) -----------------------------------
// Its scratch is one file-scope device array: the code object's own memory on
// every GPU the library runs on, held until the process ends, so a call makes
// no device allocation.  Carved, 256-byte aligned, for at most
// kNoAllocMaxKeys 64-bit keys: key and value ping-pong buffers, the digit
// histograms, the control block and the look-back words of 8 passes.
// In host code `scratch` names a host-side stand-in; the device address of
// each GPU's copy comes from hipGetSymbolAddress (noalloc_scratch).
constexpr uint64_t kNoAllocMaxKeys = 1ull << 20;
constexpr uint64_t kNoAllocTiles = (kNoAllocMaxKeys + 4095) / 4096;  // the sort's 4,096-key tiles
constexpr uint64_t na_align(uint64_t b) { return (b + 255) & ~255ull; }
constexpr uint64_t kNaKa = 0;
constexpr uint64_t kNaKb = kNaKa + na_align(kNoAllocMaxKeys * 8);
constexpr uint64_t kNaVa = kNaKb + na_align(kNoAllocMaxKeys * 8);
constexpr uint64_t kNaVb = kNaVa + na_align(kNoAllocMaxKeys * 4);
constexpr uint64_t kNaHist = kNaVb + na_align(kNoAllocMaxKeys * 4);
constexpr uint64_t kNaCtl = kNaHist + na_align(8 * 256 * 4);
constexpr uint64_t kNaLb = kNaCtl + na_align(sizeof(SortCtl));
constexpr uint64_t kNoAllocBytes = kNaLb + na_align(8 * kNoAllocTiles * 256 * 4);

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
std::mutex g_na_mutex;
void* g_na_ptr[kNoAllocDevices] = {};

// Device address of `scratch` on the current device (set up once per device,
// with that device current).
static bvh_status noalloc_setup_current() {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) return fail(BVH_E_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lk(g_na_mutex);
  if (g_na_ptr[dev]) return BVH_OK;
  void* p = nullptr;
  HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(BVH_E_HIP, "hipGetSymbolAddress(scratch) gave a null address");
  g_na_ptr[dev] = p;
  return BVH_OK;
}

// The scratch of the world's device; set up by bvh_world_create, or here
// with the world's device made current for the call.
static bvh_status noalloc_scratch(bvh_world w, void** out) {
  const int dev = (int)w->cfg.device;
  if (dev < 0 || dev >= kNoAllocDevices) return fail(BVH_E_UNSUPPORTED, "device index above 63");
  {
    std::lock_guard<std::mutex> lk(g_na_mutex);
    if (g_na_ptr[dev]) {
      *out = g_na_ptr[dev];
      return BVH_OK;
    }
  }
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != dev) HIP_TRY(hipSetDevice(dev));
  const bvh_status st = noalloc_setup_current();
  if (cur != dev) HIP_TRY(hipSetDevice(cur));
  BVH_TRY(st);
  std::lock_guard<std::mutex> lk(g_na_mutex);
  *out = g_na_ptr[dev];
  return BVH_OK;
}

// device copy of a caller array: the array itself (device memory) or the
// staging buffer it was uploaded into (host memory)
template <typename T>
static bvh_status stage_in(DevBuf& buf, const T* src, uint64_t count, uint32_t mem, hipStream_t s, const T** dev) {
  if (mem == BVH_MEM_DEVICE || !src || count == 0) {
    *dev = src;
    return BVH_OK;
  }
  BVH_TRY(buf.ensure(count * sizeof(T)));
  HIP_TRY(hipMemcpyAsync(buf.p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  *dev = buf.as<T>();
  return BVH_OK;
}

// ============================================================================================
extern "C" {

const char* bvh_status_string(bvh_status s) {
  switch (s) {
    case BVH_OK: return "ok";
    case BVH_E_INVALID: return "invalid argument";
    case BVH_E_OOM: return "out of device memory";
    case BVH_E_OVERFLOW: return "buffer too small";
    case BVH_E_HIP: return "HIP error";
    case BVH_E_STATE: return "call out of order";
    case BVH_E_UNSUPPORTED: return "unsupported";
  }
  return "unknown";
}

const char* bvh_last_error(void) { return g_err.c_str(); }
uint32_t bvh_abi_version(void) { return BVH_ABI_VERSION; }

void bvh_world_config_init(bvh_world_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->morton_bits = 63;
  cfg->leaf_size = 1;
  cfg->precision = BVH_PREC_FP64;
  cfg->flags = 0;
  cfg->stream = nullptr;
  cfg->device = 0;
}

bvh_status bvh_world_create(size_t od, const bvh_world_config* cfg, bvh_world* out) {
  if (!out) return fail(BVH_E_INVALID, "null output");
  *out = nullptr;
  if (!is_pow2(od)) return fail(BVH_E_INVALID, "overdecomposition factor must be a power of 2");
  if (od > 16384) return fail(BVH_E_INVALID, "overdecomposition factor above 16384");
  bvh_world_config c;
  bvh_world_config_init(&c);
  if (cfg) c = *cfg;
  if (c.morton_bits != 63 && c.morton_bits != 30) return fail(BVH_E_INVALID, "morton_bits must be 63 or 30");
  if (c.leaf_size < 1 || c.leaf_size > 8) return fail(BVH_E_INVALID, "leaf_size must be in 1..8");
  if (c.precision != BVH_PREC_FP64 && c.precision != BVH_PREC_FP32_OUTWARD && c.precision != BVH_PREC_FP64_RECORDS)
    return fail(BVH_E_INVALID, "unknown precision");
  HIP_TRY(hipSetDevice(c.device));
  auto w = std::make_unique<bvh_world_s>();
  w->od = od;
  w->cfg = c;
  if (c.stream) {
    w->stream = (hipStream_t)c.stream;
  } else {
    HIP_TRY(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    w->own_stream = true;
  }
  BVH_TRY(noalloc_setup_current());  // the device address of bvh_sort_u64_noalloc's scratch
  void* hc = nullptr;
  HIP_TRY(hipHostMalloc(&hc, 16 * sizeof(unsigned long long), hipHostMallocDefault));
  w->h_count = (unsigned long long*)hc;
  *out = w.release();
  return BVH_OK;
}

static void orphan_graphs(bvh_world w);

bvh_status bvh_world_destroy(bvh_world w) {
  if (!w) return BVH_OK;
  for (auto& o : w->objects) o->build_pending = false;  // builds nobody reads
  (void)join_all(w);
  (void)hipStreamSynchronize(w->stream);
  // graphs outlive their world only as handles to destroy
  orphan_graphs(w);
  w->objects.clear();
  for (hipEvent_t e : w->event_pool) (void)hipEventDestroy(e);
  if (w->h_count) (void)hipHostFree(w->h_count);
  if (w->bq_host) (void)hipHostFree(w->bq_host);
  if (w->own_stream) (void)hipStreamDestroy(w->stream);
  delete w;
  return BVH_OK;
}

bvh_status bvh_world_overdecomposition(bvh_world w, size_t* od) {
  if (!w || !od) return fail(BVH_E_INVALID, "null argument");
  *od = w->od;
  return BVH_OK;
}

bvh_status bvh_create_collision_object(bvh_world w, bvh_object* out) {
  if (!w || !out) return fail(BVH_E_INVALID, "null argument");
  auto o = std::make_unique<bvh_object_s>();
  o->w = w;
  o->for_each_buf([&](DevBuf& d) { d.gen = &w->alloc_gen; });
  // outputs of calls that are never captured (records, digests, the exact
  // filter, wire formats): their growth leaves the world's graphs launchable
  DevBuf* outside[] = {&o->records, &o->digest, &o->slab_err, &o->ex_flag, &o->ex_out, &o->ex_lb, &o->ex_ctr,
                       &o->wire_cent, &o->wire_range, &o->wire_src, &o->wire_out};
  for (DevBuf* d : outside) d->gen = &g_alloc_gen;
  o->id = w->objects.size();
  *out = o.get();
  w->objects.push_back(std::move(o));
  return BVH_OK;
}

bvh_status bvh_num_collision_objects(bvh_world w, size_t* n) {
  if (!w || !n) return fail(BVH_E_INVALID, "null argument");
  *n = w->objects.size();
  return BVH_OK;
}

bvh_status bvh_start_iteration(bvh_world w) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  w->iteration_open = true;
  return BVH_OK;
}

bvh_status bvh_finish_iteration(bvh_world w) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  for (auto& o : w->objects) BVH_TRY(bvh_end_phase(o.get()));
  w->iteration_open = false;
  HIP_TRY(hipStreamSynchronize(w->stream));
  return BVH_OK;
}

bvh_status bvh_synchronize(bvh_world w) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  HIP_TRY(hipStreamSynchronize(w->stream));
  return BVH_OK;
}

bvh_status bvh_world_stream(bvh_world w, void** stream) {
  if (!w || !stream) return fail(BVH_E_INVALID, "null argument");
  *stream = (void*)w->stream;
  return BVH_OK;
}

bvh_status bvh_stage_times(bvh_world w, const char** names, double* ms, size_t cap, size_t* count) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  HIP_TRY(hipStreamSynchronize(w->stream));
  size_t k = 0;
  for (const Stage& st : w->stages) {
    if (st.open) continue;  // never ended (a failed call)
    if (k < cap) {
      float t = 0.f;
      HIP_TRY(hipEventElapsedTime(&t, st.a, st.b));
      if (names) names[k] = st.name;
      if (ms) ms[k] = t;
    }
    ++k;
  }
  if (count) *count = k;
  w->reset_stages();
  return BVH_OK;
}

bvh_status bvh_object_id(bvh_object o, size_t* id) {
  if (!o || !id) return fail(BVH_E_INVALID, "null argument");
  *id = o->id;
  return BVH_OK;
}

bvh_status bvh_object_device_bytes(bvh_object o, uint64_t* bytes) {
  if (!o || !bytes) return fail(BVH_E_INVALID, "null argument");
  *bytes = o->device_bytes();
  return BVH_OK;
}

bvh_status bvh_object_release(bvh_object o) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  bvh_world w = o->w;
  if (w->capturing) return fail(BVH_E_STATE, "release inside a capture");
  o->build_pending = false;  // nobody reads this build any more
  BVH_TRY(join_all(w));
  // queued work of the world and of the object's build stream may still use
  // the buffers
  if (o->bs) HIP_TRY(hipStreamSynchronize(o->bs));
  HIP_TRY(hipStreamSynchronize(w->stream));
  o->for_each_buf([](DevBuf& d) { d.release(); });
  // results of other objects against this one read its buffers: gone too
  for (auto& x : w->objects)
    if (x->result_other == o) {
      x->have_results = false;
      x->results_pending = false;
      x->result_other = nullptr;
    }
  o->built = false;
  o->have_data = false;
  o->have_results = false;
  o->results_pending = false;
  o->patch_slabs_ready = o->patch_pairs_ready = false;
  o->npairs = o->npatch_pairs = o->nvisits = o->ncandidates = o->nrecheck = o->nrecheck_max = 0;
  o->result_ptr = nullptr;
  o->result_other = nullptr;
  o->ref_rp = nullptr;
  o->in = ElemInput{};
  return BVH_OK;
}

// fp32-outward slabs (ordered-integer images, relative to the frame origin
// `org`) back in world coordinates as fp64, each moved one fp64 ulp outward
// so that the box still contains the exact one after the rounding of the sum.
static void decode_box32(const int32_t* ko, const double* org, double* out) {
  static const double N[13][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                  {0.5773502691896258, 0.5773502691896258, 0.5773502691896258},
                                  {0.5773502691896258, -0.5773502691896258, 0.5773502691896258},
                                  {0.5773502691896258, 0.5773502691896258, -0.5773502691896258},
                                  {0.5773502691896258, -0.5773502691896258, -0.5773502691896258},
                                  {0.7071067811865475, 0.7071067811865475, 0}, {0.7071067811865475, 0, 0.7071067811865475},
                                  {0, 0.7071067811865475, 0.7071067811865475}, {0.7071067811865475, -0.7071067811865475, 0},
                                  {0.7071067811865475, 0, -0.7071067811865475}, {0, 0.7071067811865475, -0.7071067811865475}};
  for (int a = 0; a < kAxes; ++a) {
    const double po = org[0] * N[a][0] + org[1] * N[a][1] + org[2] * N[a][2];
    for (int h = 0; h < 2; ++h) {
      const int32_t k = ko[2 * a + h];
      // inverse of the ordered-integer image
      const uint32_t u = k >= 0 ? (uint32_t)k : (0x80000000u | (uint32_t)(-(int64_t)k));
      float f;
      std::memcpy(&f, &u, 4);
      const double v = (double)f + po;
      out[2 * a + h] = h ? std::nextafter(v, HUGE_VAL) : std::nextafter(v, -HUGE_VAL);
    }
  }
}

bvh_status bvh_object_root_kdop(bvh_object o, double* kdop) {
  if (!o || !kdop) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(o->w));
  if (!o->built) return fail(BVH_E_STATE, "object not built");
  for (int i = 0; i < 26; ++i) kdop[i] = (i & 1) ? -1.7976931348623157e308 : 1.7976931348623157e308;
  if (o->in.n == 0) return BVH_OK;
  hipStream_t s = o->w->stream;
  if (o->fp32) {
    int32_t ko[kSlabs];
    double org[3];
    BVH_TRY(copy_out(ko, o->rootbox32.p, sizeof(ko), BVH_MEM_HOST, s));
    BVH_TRY(copy_out(org, o->frame.p, sizeof(org), BVH_MEM_HOST, s));
    decode_box32(ko, org, kdop);
    return BVH_OK;
  }
  return copy_out(kdop, o->rootbox64.p, 26 * sizeof(double), BVH_MEM_HOST, s);
}

bvh_status bvh_object_overdecomposition(bvh_object o, size_t* od) {
  if (!o || !od) return fail(BVH_E_INVALID, "null argument");
  *od = o->w->od;
  return BVH_OK;
}

bvh_status bvh_object_set_patch_offsets(bvh_object o, const uint64_t* offsets, size_t count) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  BVH_TRY(join_all(o->w));
  if (o->w->capturing) return fail(BVH_E_STATE, "patch offsets inside a capture");
  if (count + 1 != o->w->od) return fail(BVH_E_INVALID, "od - 1 patch offsets expected");
  if (count && !offsets) return fail(BVH_E_INVALID, "null offsets");
  std::vector<uint32_t> dev(count ? count : 1, 0);
  for (size_t k = 0; k < count; ++k) {
    if (k && offsets[k] < offsets[k - 1]) return fail(BVH_E_INVALID, "patch offsets must not decrease");
    if (offsets[k] >= (1ull << 30)) return fail(BVH_E_INVALID, "patch offset beyond 2^30 - 1 elements");
    dev[k] = (uint32_t)offsets[k];
  }
  BVH_TRY(o->given.ensure(dev.size() * 4));
  HIP_TRY(hipMemcpyAsync(o->given.p, dev.data(), dev.size() * 4, hipMemcpyHostToDevice, o->w->stream));
  HIP_TRY(hipStreamSynchronize(o->w->stream));
  o->given_h.assign(offsets, offsets + count);
  return BVH_OK;
}

bvh_status bvh_set_entity_data_poly(bvh_object o, const double* verts, uint32_t nv, uint64_t n,
                                    const uint64_t* gids, double eps, uint32_t mem, uint32_t split) {
  if (nv < 1 || nv > 8) return fail(BVH_E_INVALID, "verts_per_element must be in 1..8");
  return set_entity_common(o, kPoly, (int)nv, eps, verts, nullptr, n, gids, mem, split);
}

bvh_status bvh_set_entity_data_sphere(bvh_object o, const double* xyzr, uint64_t n, const uint64_t* gids,
                                      uint32_t mem, uint32_t split) {
  return set_entity_common(o, kSphere, 1, 0.0, xyzr, nullptr, n, gids, mem, split);
}

bvh_status bvh_set_entity_data_snapshot(bvh_object o, const double* kdops, const double* centroids, uint64_t n,
                                        const uint64_t* gids, uint32_t mem, uint32_t split) {
  if (n > 0 && !centroids) return fail(BVH_E_INVALID, "null centroids");
  return set_entity_common(o, kSnapshot, 1, 0.0, kdops, centroids, n, gids, mem, split);
}

bvh_status bvh_refit(bvh_object o, const double* data, uint32_t mem) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  BVH_TRY(join_all(o->w));
  if (!o->built || !o->have_data) return fail(BVH_E_STATE, "refit before set_entity_data");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  const uint64_t n = o->in.n;
  if (n > 0 && !data) return fail(BVH_E_INVALID, "null element data");
  if (o->w->capturing && mem == BVH_MEM_HOST) return fail(BVH_E_STATE, "a capture takes device element data");
  o->have_results = false;
  if (n > 0) {
    const size_t planes = o->in.kind == kPoly ? (size_t)o->in.nv * 3 : (o->in.kind == kSphere ? 4 : kSlabs);
    const size_t bytes = planes * n * sizeof(double);
    if (mem == BVH_MEM_HOST) {
      BVH_TRY(o->input.ensure(bytes));
      HIP_TRY(hipMemcpyAsync(o->input.p, data, bytes, hipMemcpyHostToDevice, o->w->stream));
      o->in.data = o->input.as<double>();
    } else {
      o->in.data = data;
    }
  }
  return object_refit(o);
}

bvh_status bvh_init_broadphase(bvh_object o) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  if (!o->built) return fail(BVH_E_STATE, "init_broadphase before set_entity_data");
  o->have_results = false;
  return BVH_OK;
}

bvh_status bvh_broadphase(bvh_object o, bvh_object other) {
  if (!o || !other) return fail(BVH_E_INVALID, "null object");
  if (o->w != other->w) {
    BVH_TRY(join_all(o->w));
    return fail(BVH_E_INVALID, "objects of different worlds");
  }
  if (o == other) {
    BVH_TRY(join_all(o->w));
    // the reference skips self collisions in the narrowphase
    // (collision_object/impl.cpp:229-235): no results
    o->npairs = 0;
    o->npatch_pairs = 0;
    o->have_results = true;
    o->result_ptr = o->pairs.as<uint64_t>();
    o->result_other = o;
    o->self_result = true;
    return BVH_OK;
  }
  BVH_TRY(join_for_query(o, other));
  return run_query(o, other, false);
}

bvh_status bvh_broadphase_patches(bvh_object o, bvh_object other) {
  if (!o || !other) return fail(BVH_E_INVALID, "null object");
  BVH_TRY(join_all(o->w));
  if (o->w != other->w) return fail(BVH_E_INVALID, "objects of different worlds");
  if (o == other) {
    o->npairs = 0;
    o->npatch_pairs = 0;
    o->have_results = true;
    o->result_ptr = o->pairs.as<uint64_t>();
    o->result_other = o;
    o->self_result = true;
    return BVH_OK;
  }
  return run_query(o, other, false, false);
}

bvh_status bvh_broadphase_batch(bvh_object* os, bvh_object* others, uint64_t k) {
  if (k == 0) return BVH_OK;
  if (!os || !others) return fail(BVH_E_INVALID, "null object list");
  bvh_world w = os[0] ? os[0]->w : nullptr;
  if (!w) return fail(BVH_E_INVALID, "null object");
  for (uint64_t i = 0; i < k; ++i) {
    if (!os[i] || !others[i]) return fail(BVH_E_INVALID, "null object");
    if (os[i]->w != w || others[i]->w != w) return fail(BVH_E_INVALID, "objects of different worlds");
    if (os[i] == others[i]) return fail(BVH_E_INVALID, "a pair of one object (self collision is not batched)");
    for (uint64_t j = 0; j < i; ++j)
      if (os[j] == os[i]) return fail(BVH_E_INVALID, "an object receives the results of two pairs");
  }
  BVH_TRY(join_all(w));
  if (w->capturing) return fail(BVH_E_STATE, "broadphase_batch inside a capture");
  if (w->cfg.flags & BVH_FLAG_SORT_RESULTS) {
    for (uint64_t i = 0; i < k; ++i) BVH_TRY(run_query(os[i], others[i], false));
    return BVH_OK;
  }
  return run_query_batch(os, others, k);
}

bvh_status bvh_self_collision(bvh_object o) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  BVH_TRY(join_all(o->w));
  return run_query(o, o, true);
}

bvh_status bvh_self_collision_cross_patch(bvh_object o) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  BVH_TRY(join_all(o->w));
  return run_query(o, o, true, true, true);
}

bvh_status bvh_end_phase(bvh_object o) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  return BVH_OK;
}

bvh_status bvh_result_count(bvh_object o, uint64_t* count) {
  if (!o || !count) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(o->w));
  BVH_TRY(results_ready(o));
  *count = o->have_results ? o->npairs : 0;
  return BVH_OK;
}

bvh_status bvh_result_pairs_device(bvh_object o, const uint64_t** pairs, uint64_t* count) {
  if (!o || !pairs || !count) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(o->w));
  BVH_TRY(results_ready(o));
  *pairs = o->have_results ? o->result_ptr : nullptr;
  *count = o->have_results ? o->npairs : 0;
  return BVH_OK;
}

bvh_status bvh_result_digest(bvh_object o, uint64_t digest[2]) {
  if (!o || !digest) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(o->w));
  BVH_TRY(results_ready(o));
  const uint64_t n = o->have_results ? o->npairs : 0;
  hipStream_t s = o->w->stream;
  BVH_TRY(o->digest.ensure(2 * sizeof(unsigned long long)));
  unsigned long long* d = o->digest.as<unsigned long long>();
  HIP_TRY(launch_fill_u32((uint32_t*)d, 4, 0u, s));
  HIP_TRY(launch_digest(o->result_ptr, n, d, s));
  return copy_out(digest, d, 2 * sizeof(uint64_t), BVH_MEM_HOST, s);
}

bvh_status bvh_result_pairs(bvh_object o, uint64_t* out, uint64_t cap, uint32_t mem) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  BVH_TRY(join_all(o->w));
  BVH_TRY(results_ready(o));
  const uint64_t n = o->have_results ? o->npairs : 0;
  if (cap < n) return fail(BVH_E_OVERFLOW, "result buffer holds " + std::to_string(cap) + " of " + std::to_string(n));
  if (n && !out) return fail(BVH_E_INVALID, "null output");
  return copy_out(out, o->result_ptr, n * 8, mem, o->w->stream);
}

bvh_status bvh_result_records(bvh_object o, bvh_narrowphase_record* out, uint64_t cap, uint32_t mem) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  BVH_TRY(join_all(o->w));
  BVH_TRY(results_ready(o));
  const uint64_t n = o->have_results ? o->npairs : 0;
  if (cap < n) return fail(BVH_E_OVERFLOW, "record buffer holds " + std::to_string(cap) + " of " + std::to_string(n));
  if (n == 0) return BVH_OK;
  if (!out) return fail(BVH_E_INVALID, "null output");
  hipStream_t s = o->w->stream;
  bvh_object_s* q = o->result_other;
  BVH_TRY(o->records.ensure(n * sizeof(bvh_narrowphase_record)));
  HIP_TRY(launch_records(o->result_ptr, n, o->patch_of_input.as<uint32_t>(),
                         o->has_gids ? o->gids.as<uint64_t>() : nullptr, q->patch_of_input.as<uint32_t>(),
                         q->has_gids ? q->gids.as<uint64_t>() : nullptr, 0, 0, o->records.as<uint64_t>(), s));
  return copy_out(out, o->records.p, n * sizeof(bvh_narrowphase_record), mem, s);
}

// Largest distance of an object's fp32-outward leaf slabs from their fp64
// values, in fp32 ulps (float bits), recomputed from its element data (which
// must still be valid); 0 for fp64 records.
static bvh_status object_slab_error(bvh_object_s* o, uint32_t* bits) {
  *bits = 0;
  if (!o || !o->fp32 || !o->built || o->in.n == 0) return BVH_OK;
  hipStream_t s = o->w->stream;
  BVH_TRY(o->slab_err.ensure(sizeof(uint32_t)));
  HIP_TRY(launch_fill_u32(o->slab_err.as<uint32_t>(), 1, 0u, s));
  HIP_TRY(launch_slab_error(o->in, o->frame.as<double>(), o->slab_err.as<uint32_t>(), s));
  return copy_out(bits, o->slab_err.p, sizeof(uint32_t), BVH_MEM_HOST, s);
}

bvh_status bvh_result_stats(bvh_object o, uint64_t* node_visits, uint64_t* patch_pairs, uint64_t* candidates,
                            uint64_t* rechecked, double* max_slab_error_ulps) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  BVH_TRY(join_all(o->w));
  BVH_TRY(results_ready(o));
  if (node_visits) *node_visits = o->have_results ? o->nvisits : 0;
  if (patch_pairs) {
    BVH_TRY(ensure_patch_pairs(o));
    *patch_pairs = o->have_results ? o->npatch_pairs : 0;
  }
  if (candidates) *candidates = o->have_results ? (o->fp32 ? o->ncandidates : o->npairs) : 0;
  if (rechecked) *rechecked = o->have_results && o->fp32 ? o->nrecheck : 0;
  if (max_slab_error_ulps) {
    // the queried objects' slabs, measured now (not part of the build)
    uint32_t b0 = 0, b1 = 0;
    if (o->have_results && o->fp32) {
      BVH_TRY(object_slab_error(o, &b0));
      if (o->result_other && o->result_other != o) BVH_TRY(object_slab_error(o->result_other, &b1));
    }
    const uint32_t bits = b0 > b1 ? b0 : b1;
    float f;
    std::memcpy(&f, &bits, sizeof(f));
    *max_slab_error_ulps = (double)f;
  }
  return BVH_OK;
}

bvh_status bvh_patch_pair_count(bvh_object o, uint64_t* count) {
  if (!o || !count) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(o->w));
  BVH_TRY(results_ready(o));
  BVH_TRY(ensure_patch_pairs(o));
  *count = o->have_results ? o->npatch_pairs : 0;
  return BVH_OK;
}

bvh_status bvh_patch_pairs(bvh_object o, uint64_t* out, uint64_t cap) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  BVH_TRY(join_all(o->w));
  BVH_TRY(results_ready(o));
  BVH_TRY(ensure_patch_pairs(o));
  const uint64_t n = o->have_results ? o->npatch_pairs : 0;
  if (cap < n) return fail(BVH_E_OVERFLOW, "patch pair buffer too small");
  if (n == 0) return BVH_OK;
  std::vector<uint64_t> tmp(n);
  BVH_TRY(copy_out(tmp.data(), o->ppairs.p, n * 8, BVH_MEM_HOST, o->w->stream));
  std::sort(tmp.begin(), tmp.end());
  for (uint64_t i = 0; i < n; ++i) {
    out[2 * i] = tmp[i] >> 32;
    out[2 * i + 1] = tmp[i] & 0xffffffffu;
  }
  return BVH_OK;
}

bvh_status bvh_object_bounds(bvh_object o, double* min3, double* inv3) {
  if (!o || !o->built) return fail(BVH_E_STATE, "object not built");
  BVH_TRY(join_all(o->w));
  Bounds b{};
  if (o->in.n > 0) BVH_TRY(copy_out(&b, o->bounds.p, sizeof(Bounds), BVH_MEM_HOST, o->w->stream));
  for (int c = 0; c < 3; ++c) {
    if (min3) min3[c] = b.min[c];
    if (inv3) inv3[c] = b.inv[c];
  }
  return BVH_OK;
}

bvh_status bvh_object_sorted(bvh_object o, uint64_t* codes, uint32_t* perm, uint32_t mem) {
  if (!o || !o->built) return fail(BVH_E_STATE, "object not built");
  BVH_TRY(join_all(o->w));
  const uint64_t n = o->in.n;
  if (n == 0) return BVH_OK;
  hipStream_t s = o->w->stream;
  if (codes) {
    if (o->bits == 63) {
      BVH_TRY(copy_out(codes, o->codes_a.p, n * 8, mem, s));
    } else {
      std::vector<uint32_t> tmp(n);
      BVH_TRY(copy_out(tmp.data(), o->codes_a.p, n * 4, BVH_MEM_HOST, s));
      if (mem == BVH_MEM_DEVICE) return fail(BVH_E_UNSUPPORTED, "30-bit codes are returned to host memory only");
      for (uint64_t i = 0; i < n; ++i) codes[i] = tmp[i];
    }
  }
  if (perm) BVH_TRY(copy_out(perm, o->perm.p, n * 4, mem, s));
  return BVH_OK;
}

bvh_status bvh_object_splits(bvh_object o, uint64_t* splits, size_t cap) {
  if (!o || !o->built) return fail(BVH_E_STATE, "object not built");
  BVH_TRY(join_all(o->w));
  const size_t ns = o->w->od - 1;
  if (cap < ns) return fail(BVH_E_OVERFLOW, "splits buffer too small");
  if (ns == 0) return BVH_OK;
  std::vector<uint32_t> tmp(ns);
  BVH_TRY(copy_out(tmp.data(), o->splits.p, ns * 4, BVH_MEM_HOST, o->w->stream));
  for (size_t i = 0; i < ns; ++i) splits[i] = tmp[i];
  return BVH_OK;
}

bvh_status bvh_object_sorted_kdops(bvh_object o, double* kdops, uint32_t mem) {
  if (!o || !o->built) return fail(BVH_E_STATE, "object not built");
  BVH_TRY(join_all(o->w));
  if (o->form == kFormFp32) return fail(BVH_E_UNSUPPORTED, "the fp32 form stores fp32 slabs only");
  const uint64_t n = o->in.n;
  if (n == 0) return BVH_OK;
  hipStream_t s = o->w->stream;
  if (o->form == kFormMixed) {
    // the records hold fp32 bounds: each leaf's fp64 slabs from its element
    if (mem == BVH_MEM_DEVICE) {
      HIP_TRY(launch_sorted_elem_kdops(o->in, o->perm.as<uint32_t>(), kdops, s));
      HIP_TRY(hipStreamSynchronize(s));
      return BVH_OK;
    }
    DevBuf tmp;
    BVH_TRY(tmp.ensure(n * kSlabs * sizeof(double)));
    HIP_TRY(launch_sorted_elem_kdops(o->in, o->perm.as<uint32_t>(), tmp.as<double>(), s));
    return copy_out(kdops, tmp.p, n * kSlabs * sizeof(double), BVH_MEM_HOST, s);
  }
  // each leaf's slabs from its parent's record
  if (mem == BVH_MEM_DEVICE) {
    HIP_TRY(launch_sorted_kdops(o->rec64.as<Rec64>(), o->rootbox64.as<double>(), o->delta.as<uint8_t>(), n, kdops, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BVH_OK;
  }
  DevBuf tmp;
  BVH_TRY(tmp.ensure(n * kSlabs * sizeof(double)));
  HIP_TRY(launch_sorted_kdops(o->rec64.as<Rec64>(), o->rootbox64.as<double>(), o->delta.as<uint8_t>(), n,
                              tmp.as<double>(), s));
  return copy_out(kdops, tmp.p, n * kSlabs * sizeof(double), BVH_MEM_HOST, s);
}

bvh_status bvh_local_patches(bvh_object o, bvh_patch_info* out, size_t cap) {
  if (!o || !o->built) return fail(BVH_E_STATE, "object not built");
  BVH_TRY(join_all(o->w));
  if (o->w->capturing) return fail(BVH_E_STATE, "local_patches inside a capture");
  BVH_TRY(ensure_patch_slabs(o));
  const size_t od = o->w->od;
  if (cap < od) return fail(BVH_E_OVERFLOW, "patch buffer too small");
  hipStream_t s = o->w->stream;
  const uint64_t n = o->in.n;
  std::vector<uint32_t> roots(od, kEnd), splits(od > 1 ? od - 1 : 1, 0);
  BVH_TRY(copy_out(roots.data(), o->patch_root.p, od * 4, BVH_MEM_HOST, s));
  if (od > 1) BVH_TRY(copy_out(splits.data(), o->splits.p, (od - 1) * 4, BVH_MEM_HOST, s));
  std::vector<double> cent(3 * od, 0.0);
  if (n > 0) {
    DevBuf dc;
    BVH_TRY(dc.ensure(3 * od * sizeof(double)));
    const double* cpl = o->in.kind == kPoly ? o->centroid.as<double>() : nullptr;
    const double* spl = o->in.kind == kSphere ? o->in.data : nullptr;
    const double* cao = o->in.kind == kSnapshot ? o->in.centroids : nullptr;
    HIP_TRY(launch_patch_centroids(cpl, spl, cao, o->ref_perm(), o->splits.as<uint32_t>(), n, od,
                                   dc.as<double>(), s));
    BVH_TRY(copy_out(cent.data(), dc.p, 3 * od * sizeof(double), BVH_MEM_HOST, s));
  }
  for (size_t k = 0; k < od; ++k) {
    bvh_patch_info& pi = out[k];
    const uint64_t b = (k == 0) ? 0 : splits[k - 1];
    const uint64_t e = (k == od - 1) ? n : splits[k];
    pi.global_id = k;
    pi.size = e - b;
    for (int a = 0; a < 13; ++a) {
      pi.kdop[2 * a] = 1.7976931348623157e308;
      pi.kdop[2 * a + 1] = -1.7976931348623157e308;
    }
    if (roots[k] != kEnd) {
      uint64_t enc[kSlabs];
      BVH_TRY(copy_out(enc, o->patch_k64.as<unsigned long long>() + k * kSlabs, sizeof(enc), BVH_MEM_HOST, s));
      for (int i = 0; i < 26; ++i) pi.kdop[i] = from_ordered_u64(enc[i]);
    }
    for (int c = 0; c < 3; ++c) pi.centroid[c] = cent[3 * k + c];
  }
  return BVH_OK;
}

bvh_status bvh_patch_elements(bvh_object o, size_t patch, uint64_t* indices, uint64_t cap, uint64_t* count) {
  if (!o || !o->built) return fail(BVH_E_STATE, "object not built");
  BVH_TRY(join_all(o->w));
  const size_t od = o->w->od;
  if (patch >= od) return fail(BVH_E_INVALID, "patch index out of range");
  const uint64_t n = o->in.n;
  std::vector<uint32_t> splits(od > 1 ? od - 1 : 1, 0);
  if (od > 1) BVH_TRY(copy_out(splits.data(), o->splits.p, (od - 1) * 4, BVH_MEM_HOST, o->w->stream));
  const uint64_t b = (patch == 0) ? 0 : splits[patch - 1];
  const uint64_t e = (patch == od - 1) ? n : splits[patch];
  if (count) *count = e - b;
  if (cap < e - b) return fail(BVH_E_OVERFLOW, "index buffer too small");
  if (e == b) return BVH_OK;
  std::vector<uint32_t> tmp(e - b);
  BVH_TRY(copy_out(tmp.data(), o->ref_perm() + b, (e - b) * 4, BVH_MEM_HOST, o->w->stream));
  for (uint64_t i = 0; i < e - b; ++i) indices[i] = tmp[i];
  return BVH_OK;
}

bvh_status bvh_object_patch_tree(bvh_object o, bvh_tree_node* nodes, size_t cap, size_t* nnodes,
                                 uint64_t* leaf_patches, size_t leaf_cap, size_t* nleaves) {
  if (!o || !o->built) return fail(BVH_E_STATE, "object not built");
  BVH_TRY(join_all(o->w));
  if (o->w->capturing) return fail(BVH_E_STATE, "patch tree export inside a capture");
  BVH_TRY(ensure_patch_slabs(o));
  bvh_world w = o->w;
  hipStream_t s = w->stream;
  const uint64_t n = o->in.n, od = w->od;
  DevBuf dout, dcnt, dstack;
  BVH_TRY(dout.ensure(4 * 2 * od * sizeof(int64_t)));
  BVH_TRY(dcnt.ensure(sizeof(uint32_t)));
  BVH_TRY(dstack.ensure(4 * od * sizeof(uint64_t)));
  HIP_TRY(launch_patch_tree(o->bb, n, od, o->fp32, dout.as<int64_t>(), dcnt.as<uint32_t>(), dstack.as<uint64_t>(), s));
  uint32_t cnt = 0, st = 0;
  BVH_TRY(copy_out(&cnt, dcnt.p, sizeof(cnt), BVH_MEM_HOST, s));
  BVH_TRY(copy_out(&st, o->status.p, sizeof(st), BVH_MEM_HOST, s));
  if (st & kErrTree) return fail(BVH_E_HIP, "inconsistent patch tree");
  std::vector<int64_t> topo(4 * (size_t)cnt);
  if (cnt) BVH_TRY(copy_out(topo.data(), dout.p, topo.size() * sizeof(int64_t), BVH_MEM_HOST, s));
  std::vector<uint64_t> enc(od * kSlabs);
  BVH_TRY(copy_out(enc.data(), o->patch_k64.p, enc.size() * sizeof(uint64_t), BVH_MEM_HOST, s));
  // leaves are met left to right in pre-order; ranges and k-DOPs bottom-up
  std::vector<bvh_tree_node> t(cnt);
  std::vector<uint64_t> leaves;
  for (uint32_t k = 0; k < cnt; ++k) {
    bvh_tree_node& nd = t[k];
    const int64_t par = topo[4 * k], l = topo[4 * k + 1], r = topo[4 * k + 2], pv = topo[4 * k + 3];
    nd.parent_offset = par < 0 ? 0 : par - (int64_t)k;
    nd.child_offsets[0] = l < 0 ? 0 : l - (int64_t)k;
    nd.child_offsets[1] = r < 0 ? 0 : r - (int64_t)k;
    if (pv >= 0) {
      nd.entity_range[0] = leaves.size();
      nd.entity_range[1] = leaves.size() + 1;
      for (int i = 0; i < kSlabs; ++i) nd.kdop[i] = from_ordered_u64(enc[(size_t)pv * kSlabs + i]);
      leaves.push_back((uint64_t)pv);
    }
  }
  for (uint32_t k = cnt; k-- > 0;) {
    bvh_tree_node& nd = t[k];
    if (topo[4 * k + 3] >= 0) continue;
    const bvh_tree_node& a = t[k + nd.child_offsets[0]];
    const bvh_tree_node& b = t[k + nd.child_offsets[1]];
    nd.entity_range[0] = a.entity_range[0];
    nd.entity_range[1] = b.entity_range[1];
    for (int i = 0; i < kAxes; ++i) {  // union_with: m::min / m::max (kdop.hpp:317-325)
      nd.kdop[2 * i] = (a.kdop[2 * i] < b.kdop[2 * i]) ? a.kdop[2 * i] : b.kdop[2 * i];
      nd.kdop[2 * i + 1] = (a.kdop[2 * i + 1] > b.kdop[2 * i + 1]) ? a.kdop[2 * i + 1] : b.kdop[2 * i + 1];
    }
  }
  if (nnodes) *nnodes = cnt;
  if (nleaves) *nleaves = leaves.size();
  if (nodes) {
    if (cap < cnt) return fail(BVH_E_OVERFLOW, "node buffer too small");
    std::copy(t.begin(), t.end(), nodes);
  }
  if (leaf_patches) {
    if (leaf_cap < leaves.size()) return fail(BVH_E_OVERFLOW, "leaf buffer too small");
    std::copy(leaves.begin(), leaves.end(), leaf_patches);
  }
  return BVH_OK;
}

bvh_status bvh_object_tree(bvh_object o, double* kdops, int64_t* children, uint64_t* ranges, size_t cap,
                           size_t* nnodes, size_t* root) {
  if (!o || !o->built) return fail(BVH_E_STATE, "object not built");
  BVH_TRY(join_all(o->w));
  const uint64_t n = o->in.n;
  const size_t nn = n ? 2 * n - 1 : 0;
  if (nnodes) *nnodes = nn;
  uint32_t r = kEnd;
  if (n) BVH_TRY(copy_out(&r, o->root.p, 4, BVH_MEM_HOST, o->w->stream));
  if (root) *root = (r == kEnd) ? (size_t)-1 : r;
  if (!kdops && !children && !ranges) return BVH_OK;
  if (cap < nn) return fail(BVH_E_OVERFLOW, "tree buffer too small");
  if (nn == 0) return BVH_OK;
  // a node's slabs are in its parent's record (the root's in rootbox); the
  // fp32-outward records give their bounds in world coordinates (each
  // contains the node's exact k-DOP)
  std::vector<Rec64> recs(n - 1);
  double rootk[kSlabs];
  if (o->fp32) {
    std::vector<Rec32> r32(n - 1);
    int32_t rk[kSlabs];
    double org[3];
    if (n > 1) BVH_TRY(copy_out(r32.data(), o->rec32.p, (n - 1) * sizeof(Rec32), BVH_MEM_HOST, o->w->stream));
    BVH_TRY(copy_out(rk, o->rootbox32.p, sizeof(rk), BVH_MEM_HOST, o->w->stream));
    BVH_TRY(copy_out(org, o->frame.p, sizeof(org), BVH_MEM_HOST, o->w->stream));
    decode_box32(rk, org, rootk);
    for (size_t x = 0; x + 1 < n; ++x) {
      Rec64& R = recs[x];
      const Rec32& Q = r32[x];
      for (int side = 0; side < 2; ++side) {
        R.child[side] = rec_child(Q, side);
        R.elem[side] = rec_elem(Q, side);
        if (kdops) decode_box32(rec_box(Q, side), org, R.box[side]);
      }
      R.first = rec_first(Q);
      R.last = rec_last(Q);
      R.patch = rec_patch(Q);
    }
  } else {
    if (n > 1) BVH_TRY(copy_out(recs.data(), o->rec64.p, (n - 1) * sizeof(Rec64), BVH_MEM_HOST, o->w->stream));
    BVH_TRY(copy_out(rootk, o->rootbox64.p, sizeof(rootk), BVH_MEM_HOST, o->w->stream));
  }
  if (kdops && r < nn) std::memcpy(kdops + 26 * (size_t)r, rootk, sizeof(rootk));
  for (size_t i = n - 1; i < nn; ++i) {
    if (children) children[2 * i] = children[2 * i + 1] = -1;
    if (ranges) ranges[2 * i] = ranges[2 * i + 1] = i - (n - 1);
  }
  for (size_t x = 0; x + 1 < n; ++x) {
    const Rec64& R = recs[x];
    if (children) {
      children[2 * x] = R.child[0];
      children[2 * x + 1] = R.child[1];
    }
    if (ranges) {
      ranges[2 * x] = R.first;
      ranges[2 * x + 1] = R.last;
    }
    for (int side = 0; side < 2; ++side) {
      const uint32_t c = R.child[side];
      if (c >= nn) return fail(BVH_E_HIP, "tree record holds an invalid child id");
      if (kdops) std::memcpy(kdops + 26 * (size_t)c, R.box[side], 26 * sizeof(double));
    }
  }
  return BVH_OK;
}

bvh_status bvh_exclusive_scan_u32(bvh_world w, uint32_t* vals, uint64_t n) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  if (w->capturing) return fail(BVH_E_STATE, "scan inside a capture");
  if (n == 0) return BVH_OK;
  if (!vals) return fail(BVH_E_INVALID, "null values");
  hipStream_t s = w->stream;
  const uint64_t tiles = sort_tiles(n);
  BVH_TRY(w->s_lb.ensure(tiles * 8 + 256));
  BVH_TRY(w->s_ctl.ensure(sizeof(SortCtl)));
  HIP_TRY(launch_fill_u32(w->s_lb.as<uint32_t>(), 2 * tiles, 0u, s));
  HIP_TRY(launch_fill_u32(w->s_ctl.as<uint32_t>(), 1, 0u, s));
  HIP_TRY(launch_scan_u32(vals, n, w->s_lb.as<uint64_t>(), w->s_ctl.as<uint32_t>(), s));
  HIP_TRY(hipStreamSynchronize(s));
  return BVH_OK;
}

bvh_status bvh_closest_point(bvh_world w, const double* points, uint64_t n, const double* tri, double* closest,
                             uint32_t mem) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  if (w->capturing) return fail(BVH_E_STATE, "closest_point inside a capture");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  if (!tri) return fail(BVH_E_INVALID, "null triangle");
  if (n == 0) return BVH_OK;
  if (!points || !closest) return fail(BVH_E_INVALID, "null points or output");
  hipStream_t s = w->stream;
  Tri t;
  std::memcpy(&t, tri, sizeof(t));
  const size_t bytes = n * 3 * sizeof(double);
  if (mem == BVH_MEM_DEVICE) {
    HIP_TRY(launch_closest_point(points, n, t, closest, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BVH_OK;
  }
  BVH_TRY(w->prim_in.ensure(bytes));
  BVH_TRY(w->prim_out.ensure(bytes));
  HIP_TRY(hipMemcpyAsync(w->prim_in.p, points, bytes, hipMemcpyHostToDevice, s));
  HIP_TRY(launch_closest_point(w->prim_in.as<double>(), n, t, w->prim_out.as<double>(), s));
  return copy_out(closest, w->prim_out.p, bytes, BVH_MEM_HOST, s);
}

// ---- exact primitive tests -------------------------------------------------------------
static bvh_status prim_pairs(bvh_world w, const double* p, const double* q, uint64_t n, uint32_t mem, uint8_t* out,
                             int width, bool tri) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  if (w->capturing) return fail(BVH_E_STATE, "primitive tests inside a capture");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  if (n == 0) return BVH_OK;
  if (!p || !q || !out) return fail(BVH_E_INVALID, "null input or output");
  hipStream_t s = w->stream;
  const double *dp = nullptr, *dq = nullptr;
  BVH_TRY(stage_in(w->sw_ka, p, n * width, mem, s, &dp));
  BVH_TRY(stage_in(w->sw_kb, q, n * width, mem, s, &dq));
  uint8_t* dout = out;
  if (mem == BVH_MEM_HOST) {
    BVH_TRY(w->sw_out.ensure(n));
    dout = w->sw_out.as<uint8_t>();
  }
  if (tri) HIP_TRY(launch_tri_pairs(dp, dq, n, dout, s));
  else HIP_TRY(launch_sphere_pairs(dp, dq, n, dout, s));
  if (mem == BVH_MEM_HOST) return copy_out(out, dout, n, BVH_MEM_HOST, s);
  HIP_TRY(hipStreamSynchronize(s));
  return BVH_OK;
}

bvh_status bvh_triangle_pairs_intersect(bvh_world w, const double* p, const double* q, uint64_t n, uint32_t mem,
                                        uint8_t* out) {
  return prim_pairs(w, p, q, n, mem, out, 9, true);
}

bvh_status bvh_sphere_pairs_intersect(bvh_world w, const double* a, const double* b, uint64_t n, uint32_t mem,
                                      uint8_t* out) {
  return prim_pairs(w, a, b, n, mem, out, 4, false);
}

bvh_status bvh_result_filter_exact(bvh_object o, uint64_t* kept) {
  if (!o) return fail(BVH_E_INVALID, "null object");
  BVH_TRY(join_all(o->w));
  if (o->w->capturing) return fail(BVH_E_STATE, "result_filter_exact inside a capture");
  BVH_TRY(results_ready(o));
  if (!o->have_results) return fail(BVH_E_STATE, "no results to filter: run broadphase or self_collision first");
  bvh_object_s* t = o->result_other;
  const uint64_t n = o->npairs;
  if (n == 0 || !t) {
    if (kept) *kept = 0;
    return BVH_OK;
  }
  const ElemInput& a = o->in;
  const ElemInput& b = t->in;
  int tri;
  if (a.kind == kPoly && a.nv == 3 && b.kind == kPoly && b.nv == 3) tri = 1;
  else if (a.kind == kSphere && b.kind == kSphere) tri = 0;
  else return fail(BVH_E_UNSUPPORTED, "exact primitive tests take triangles (3-vertex polytopes) or spheres on both objects");
  if (n >= (1ull << 32)) return fail(BVH_E_UNSUPPORTED, "result_filter_exact holds at most 2^32 - 1 pairs");
  hipStream_t s = o->w->stream;
  const uint64_t tiles = sort_tiles(n);
  BVH_TRY(o->ex_flag.ensure(n * 4));
  BVH_TRY(o->ex_out.ensure(n * 8));
  BVH_TRY(o->ex_lb.ensure(tiles * 8));
  BVH_TRY(o->ex_ctr.ensure(256));
  uint32_t* flag = o->ex_flag.as<uint32_t>();
  unsigned long long* dkept = reinterpret_cast<unsigned long long*>(o->ex_ctr.as<char>() + 128);
  HIP_TRY(launch_fill_u32(o->ex_lb.as<uint32_t>(), 2 * tiles, 0u, s));
  HIP_TRY(launch_fill_u32(o->ex_ctr.as<uint32_t>(), 64, 0u, s));
  HIP_TRY(launch_pair_flags(o->result_ptr, n, a, b, tri, flag, s));
  HIP_TRY(launch_scan_u32(flag, n, o->ex_lb.as<uint64_t>(), o->ex_ctr.as<uint32_t>(), s));
  HIP_TRY(launch_pair_compact(o->result_ptr, n, a, b, tri, flag, o->ex_out.as<uint64_t>(), dkept, s));
  unsigned long long k = 0;
  BVH_TRY(copy_out(&k, dkept, sizeof(k), BVH_MEM_HOST, s));
  o->result_ptr = o->ex_out.as<uint64_t>();
  o->npairs = k;
  if (kept) *kept = k;
  return BVH_OK;
}

// ---- wire formats ------------------------------------------------------------------------
// doubles of one element record of an element kind (after its u64 id)
static uint32_t element_doubles(int kind, int nv) {
  return kind == kPoly ? (uint32_t)(3 * nv) : (kind == kSphere ? 4u : (uint32_t)(kSlabs + 3));
}

// the patch's range [begin, end) of positions in the reference order
static bvh_status patch_ranges(bvh_object_s* o, std::vector<uint64_t>& b, std::vector<uint64_t>& e) {
  const size_t od = o->w->od;
  const uint64_t n = o->in.n;
  std::vector<uint32_t> splits(od > 1 ? od - 1 : 1, 0);
  if (od > 1 && n > 0) BVH_TRY(copy_out(splits.data(), o->splits.p, (od - 1) * 4, BVH_MEM_HOST, o->w->stream));
  b.resize(od);
  e.resize(od);
  for (size_t k = 0; k < od; ++k) {
    b[k] = (k == 0) ? 0 : splits[k - 1];
    e[k] = (k == od - 1) ? n : splits[k];
  }
  return BVH_OK;
}

bvh_status bvh_pack_patches(bvh_object o, void* out, uint64_t cap, uint64_t* bytes, uint32_t mem) {
  if (!o || !bytes) return fail(BVH_E_INVALID, "null argument");
  if (!o->built) return fail(BVH_E_STATE, "object not built");
  BVH_TRY(join_all(o->w));
  if (o->w->capturing) return fail(BVH_E_STATE, "pack inside a capture");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  const uint64_t od = o->w->od, n = o->in.n;
  *bytes = od * kPatchRecordBytes;
  if (!out) return BVH_OK;
  if (cap < *bytes) return fail(BVH_E_OVERFLOW, "patch buffer holds " + std::to_string(cap) + " of " +
                                                    std::to_string(*bytes) + " bytes");
  BVH_TRY(ensure_patch_slabs(o));
  hipStream_t s = o->w->stream;
  BVH_TRY(o->wire_cent.ensure(3 * od * sizeof(double)));
  double* cent = o->wire_cent.as<double>();
  HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(cent), 6 * od, 0u, s));
  if (n > 0) {
    const double* cpl = o->in.kind == kPoly ? o->centroid.as<double>() : nullptr;
    const double* spl = o->in.kind == kSphere ? o->in.data : nullptr;
    const double* cao = o->in.kind == kSnapshot ? o->in.centroids : nullptr;
    HIP_TRY(launch_patch_centroids(cpl, spl, cao, o->ref_perm(), o->splits.as<uint32_t>(), n, od, cent, s));
  }
  void* dst = out;
  if (mem == BVH_MEM_HOST) {
    BVH_TRY(o->wire_out.ensure(*bytes));
    dst = o->wire_out.p;
  }
  HIP_TRY(launch_pack_patches(o->patch_k64.as<unsigned long long>(), o->splits.as<uint32_t>(),
                              o->patch_root.as<uint32_t>(), cent, n, od, dst, s));
  if (mem == BVH_MEM_HOST) return copy_out(out, dst, *bytes, BVH_MEM_HOST, s);
  HIP_TRY(hipStreamSynchronize(s));
  return BVH_OK;
}

bvh_status bvh_unpack_patches(bvh_world w, const void* in, uint64_t bytes, uint32_t mem, bvh_patch_info* out,
                              uint64_t cap, uint64_t* count) {
  if (!w || !count) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(w));
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  if (bytes % kPatchRecordBytes) return fail(BVH_E_INVALID, "not a whole number of 256-byte patch records");
  const uint64_t k = bytes / kPatchRecordBytes;
  *count = k;
  if (!out) return BVH_OK;
  if (cap < k) return fail(BVH_E_OVERFLOW, "patch info buffer too small");
  if (k == 0) return BVH_OK;
  if (!in) return fail(BVH_E_INVALID, "null input");
  std::vector<uint64_t> host(bytes / 8);
  if (mem == BVH_MEM_DEVICE) BVH_TRY(copy_out(host.data(), in, bytes, BVH_MEM_HOST, w->stream));
  else std::memcpy(host.data(), in, bytes);
  for (uint64_t i = 0; i < k; ++i) {
    const uint64_t* r = host.data() + i * (kPatchRecordBytes / 8);
    bvh_patch_info& pi = out[i];
    pi.global_id = r[0];
    pi.size = r[1];
    std::memcpy(pi.kdop, r + 2, sizeof(pi.kdop));
    std::memcpy(pi.centroid, r + 2 + kSlabs, sizeof(pi.centroid));
  }
  return BVH_OK;
}

bvh_status bvh_element_record_bytes(bvh_object o, uint64_t* stride) {
  if (!o || !stride) return fail(BVH_E_INVALID, "null argument");
  if (!o->have_data) return fail(BVH_E_STATE, "no element data");
  *stride = 8 + 8 * (uint64_t)element_doubles(o->in.kind, o->in.nv);
  return BVH_OK;
}

bvh_status bvh_pack_patch_elements(bvh_object o, const uint64_t* patches, uint64_t npatches, void* out, uint64_t cap,
                                   uint64_t* bytes, uint32_t mem) {
  if (!o || !bytes) return fail(BVH_E_INVALID, "null argument");
  if (!o->built) return fail(BVH_E_STATE, "object not built");
  BVH_TRY(join_all(o->w));
  if (o->w->capturing) return fail(BVH_E_STATE, "pack inside a capture");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  if (npatches && !patches) return fail(BVH_E_INVALID, "null patch list");
  if (npatches > 65535) return fail(BVH_E_INVALID, "at most 65535 patches per pack");
  const size_t od = o->w->od;
  std::vector<uint64_t> b, e;
  BVH_TRY(patch_ranges(o, b, e));
  std::vector<uint64_t> range(3 * (npatches ? npatches : 1));
  uint64_t total = 0, longest = 0;
  for (uint64_t k = 0; k < npatches; ++k) {
    if (patches[k] >= od) return fail(BVH_E_INVALID, "patch id out of range");
    range[3 * k] = b[patches[k]];
    range[3 * k + 1] = e[patches[k]];
    range[3 * k + 2] = total;
    total += e[patches[k]] - b[patches[k]];
    longest = std::max(longest, e[patches[k]] - b[patches[k]]);
  }
  const uint64_t stride = 8 + 8 * (uint64_t)element_doubles(o->in.kind, o->in.nv);
  *bytes = total * stride;
  if (!out || total == 0) return BVH_OK;
  if (cap < *bytes) return fail(BVH_E_OVERFLOW, "element buffer holds " + std::to_string(cap) + " of " +
                                                    std::to_string(*bytes) + " bytes");
  hipStream_t s = o->w->stream;
  BVH_TRY(o->wire_range.ensure(range.size() * 8));
  BVH_TRY(o->wire_src.ensure(total * 4));
  HIP_TRY(hipMemcpyAsync(o->wire_range.p, range.data(), range.size() * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(launch_patch_positions(o->ref_perm(), o->wire_range.as<uint64_t>(), npatches, longest,
                                 o->wire_src.as<uint32_t>(), s));
  void* dst = out;
  if (mem == BVH_MEM_HOST) {
    BVH_TRY(o->wire_out.ensure(*bytes));
    dst = o->wire_out.p;
  }
  HIP_TRY(launch_pack_elements(o->in, o->has_gids ? o->gids.as<uint64_t>() : nullptr, o->wire_src.as<uint32_t>(),
                               total, element_doubles(o->in.kind, o->in.nv), dst, s));
  if (mem == BVH_MEM_HOST) return copy_out(out, dst, *bytes, BVH_MEM_HOST, s);
  // the range table above is pageable host memory read by the copy engine:
  // keep it alive until the stream has passed it
  HIP_TRY(hipStreamSynchronize(s));
  return BVH_OK;
}

bvh_status bvh_unpack_elements(bvh_world w, const void* in, uint64_t count, uint32_t kind, uint32_t nv, double* data,
                               double* centroids, uint64_t* gids, uint32_t mem) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  if (w->capturing) return fail(BVH_E_STATE, "unpack inside a capture");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  if (kind != kPoly && kind != kSphere && kind != kSnapshot) return fail(BVH_E_INVALID, "unknown element kind");
  if (kind == kPoly && (nv < 1 || nv > 8)) return fail(BVH_E_INVALID, "verts_per_element must be in 1..8");
  if (count == 0) return BVH_OK;
  if (!in || !data || (kind == kSnapshot && !centroids)) return fail(BVH_E_INVALID, "null input or output");
  const uint32_t planes = element_doubles((int)kind, (int)nv);
  const uint64_t stride = 8 + 8 * (uint64_t)planes;
  hipStream_t s = w->stream;
  const uint64_t nd = (kind == kSnapshot ? (uint64_t)kSlabs : planes) * count;
  if (mem == BVH_MEM_DEVICE) {
    HIP_TRY(launch_unpack_elements(in, count, planes, kind == kSnapshot, data, centroids, gids, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BVH_OK;
  }
  BVH_TRY(w->prim_in.ensure(count * stride));
  BVH_TRY(w->prim_out.ensure(nd * 8 + (kind == kSnapshot ? 3 * count * 8 : 0) + count * 8));
  char* o = w->prim_out.as<char>();
  double* dd = reinterpret_cast<double*>(o);
  double* dc = kind == kSnapshot ? dd + nd : nullptr;
  uint64_t* dg = reinterpret_cast<uint64_t*>(o + nd * 8 + (kind == kSnapshot ? 3 * count * 8 : 0));
  HIP_TRY(hipMemcpyAsync(w->prim_in.p, in, count * stride, hipMemcpyHostToDevice, s));
  HIP_TRY(launch_unpack_elements(w->prim_in.p, count, planes, kind == kSnapshot, dd, dc, dg, s));
  BVH_TRY(copy_out(data, dd, nd * 8, BVH_MEM_HOST, s));
  if (dc) BVH_TRY(copy_out(centroids, dc, 3 * count * 8, BVH_MEM_HOST, s));
  if (gids) BVH_TRY(copy_out(gids, dg, count * 8, BVH_MEM_HOST, s));
  return BVH_OK;
}

bvh_status bvh_pack_results(bvh_object o, void* out, uint64_t cap, uint64_t* bytes, uint32_t mem) {
  if (!o || !bytes) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(o->w));
  if (o->w->capturing) return fail(BVH_E_STATE, "pack inside a capture");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  BVH_TRY(results_ready(o));
  const uint64_t n = o->have_results ? o->npairs : 0;
  *bytes = 16 + n * kResultRecordBytes;
  if (!out) return BVH_OK;
  if (cap < *bytes) return fail(BVH_E_OVERFLOW, "result message buffer holds " + std::to_string(cap) + " of " +
                                                    std::to_string(*bytes) + " bytes");
  hipStream_t s = o->w->stream;
  void* dst = out;
  if (mem == BVH_MEM_HOST) {
    BVH_TRY(o->wire_out.ensure(*bytes));
    dst = o->wire_out.p;
  }
  HIP_TRY(launch_result_header(dst, n, s));
  if (n > 0) {
    bvh_object_s* q = o->result_other;
    HIP_TRY(launch_records(o->result_ptr, n, o->patch_of_input.as<uint32_t>(),
                           o->has_gids ? o->gids.as<uint64_t>() : nullptr, q->patch_of_input.as<uint32_t>(),
                           q->has_gids ? q->gids.as<uint64_t>() : nullptr, 0, 0,
                           reinterpret_cast<uint64_t*>(static_cast<char*>(dst) + 16), s));
  }
  if (mem == BVH_MEM_HOST) return copy_out(out, dst, *bytes, BVH_MEM_HOST, s);
  HIP_TRY(hipStreamSynchronize(s));
  return BVH_OK;
}

bvh_status bvh_unpack_results(bvh_world w, const void* in, uint64_t bytes, uint32_t mem, bvh_narrowphase_record* out,
                              uint64_t cap, uint64_t* count) {
  if (!w || !count) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(w));
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  if (!in || bytes < 16) return fail(BVH_E_INVALID, "a result message holds at least its 16-byte header");
  uint64_t hdr[2];
  if (mem == BVH_MEM_DEVICE) BVH_TRY(copy_out(hdr, in, 16, BVH_MEM_HOST, w->stream));
  else std::memcpy(hdr, in, 16);
  if (hdr[0] != kResultRecordBytes)
    return fail(BVH_E_INVALID, "result stride " + std::to_string(hdr[0]) + ", expected 32");
  if (hdr[1] % kResultRecordBytes || hdr[1] > bytes - 16)
    return fail(BVH_E_INVALID, "result byte count inconsistent with the message size");
  const uint64_t k = hdr[1] / kResultRecordBytes;
  *count = k;
  if (!out || k == 0) return BVH_OK;
  if (cap < k) return fail(BVH_E_OVERFLOW, "record buffer too small");
  const char* src = static_cast<const char*>(in) + 16;
  if (mem == BVH_MEM_HOST) {
    std::memcpy(out, src, hdr[1]);
    return BVH_OK;
  }
  return copy_out(out, src, hdr[1], BVH_MEM_DEVICE, w->stream);
}

// ---- sort and sweep between two patches ---------------------------------------------
bvh_status bvh_element_kdops(bvh_world w, const double* verts, uint32_t nv, uint64_t n, double eps, uint32_t mem,
                             double* kdops) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  if (w->capturing) return fail(BVH_E_STATE, "element_kdops inside a capture");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  if (nv < 1 || nv > 8) return fail(BVH_E_INVALID, "verts_per_element must be in 1..8");
  if (n == 0) return BVH_OK;
  if (!verts || !kdops) return fail(BVH_E_INVALID, "null vertices or output");
  if (n >= (1ull << 30)) return fail(BVH_E_INVALID, "at most 2^30 - 1 elements");
  hipStream_t s = w->stream;
  ElemInput in{};
  in.kind = kPoly;
  in.nv = (int)nv;
  in.eps = eps;
  in.n = n;
  const double* dv = nullptr;
  BVH_TRY(stage_in(w->sw_ka, verts, (uint64_t)nv * 3 * n, mem, s, &dv));
  in.data = dv;
  if (mem == BVH_MEM_DEVICE) {
    HIP_TRY(launch_element_kdops(in, kdops, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BVH_OK;
  }
  BVH_TRY(w->sw_kbs.ensure(n * 26 * sizeof(double)));
  HIP_TRY(launch_element_kdops(in, w->sw_kbs.as<double>(), s));
  return copy_out(kdops, w->sw_kbs.p, n * 26 * sizeof(double), BVH_MEM_HOST, s);
}

bvh_status bvh_max_variant_axis(bvh_world w, const double* ca, uint64_t na, const double* cb, uint64_t nb,
                                uint32_t mem, int32_t* axis) {
  if (!w || !axis) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(w));
  if (w->capturing) return fail(BVH_E_STATE, "max_variant_axis inside a capture");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  if ((na && !ca) || (nb && !cb)) return fail(BVH_E_INVALID, "null centroids");
  hipStream_t s = w->stream;
  const double *da = nullptr, *db = nullptr;
  BVH_TRY(stage_in(w->sw_ka, ca, na * 3, mem, s, &da));
  BVH_TRY(stage_in(w->sw_kb, cb, nb * 3, mem, s, &db));
  BVH_TRY(w->sw_part.ensure((size_t)variant_axis_partials() * 6 * sizeof(double) + 256));
  int32_t* dax = reinterpret_cast<int32_t*>(w->sw_part.as<char>() + (size_t)variant_axis_partials() * 6 * sizeof(double));
  HIP_TRY(launch_variant_axis(da, na, db, nb, w->sw_part.as<double>(), dax, s));
  return copy_out(axis, dax, sizeof(int32_t), BVH_MEM_HOST, s);
}

bvh_status bvh_sort_and_sweep(bvh_world w, const double* pa, const double* ka, const uint64_t* ga, uint64_t na,
                              const double* pb, const double* kb, const uint64_t* gb, uint64_t nb, int32_t axis,
                              uint32_t mem, uint64_t* pairs, uint64_t cap, uint64_t* npairs) {
  if (!w || !npairs) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(w));
  *npairs = 0;
  if (w->capturing) return fail(BVH_E_STATE, "sort_and_sweep inside a capture");
  if (mem != BVH_MEM_HOST && mem != BVH_MEM_DEVICE) return fail(BVH_E_INVALID, "bad memory kind");
  if (!pa || !pb) return fail(BVH_E_INVALID, "null patch k-DOP");
  if (axis < 0 || axis >= 13) return fail(BVH_E_INVALID, "axis must be a k-DOP direction, 0..12");
  if ((na && !ka) || (nb && !kb)) return fail(BVH_E_INVALID, "null element k-DOPs");
  if (na == 0 || nb == 0) return BVH_OK;
  if (nb >= (1ull << 30)) return fail(BVH_E_INVALID, "at most 2^30 - 1 elements of B");
  hipStream_t s = w->stream;
  // the sweep range: the union of the two patches' slabs on the axis, in the
  // reference's std::min / std::max and conversion factor (kokkos.hpp:175-178)
  const double lo_a = pa[2 * axis], lo_b = pb[2 * axis], hi_a = pa[2 * axis + 1], hi_b = pb[2 * axis + 1];
  const double gmin = (lo_b < lo_a) ? lo_b : lo_a;
  const double gmax = (hi_a < hi_b) ? hi_b : hi_a;
  const double conv = (double)UINT32_MAX / (gmax - gmin);
  SweepArgs a{};
  BVH_TRY(stage_in(w->sw_ka, ka, na * 26, mem, s, &a.ka));
  BVH_TRY(stage_in(w->sw_ga, ga, na, mem, s, &a.ga));
  const double* dkb = nullptr;
  BVH_TRY(stage_in(w->sw_kb, kb, nb * 26, mem, s, &dkb));
  BVH_TRY(stage_in(w->sw_gb, gb, nb, mem, s, &a.gb));
  BVH_TRY(w->sw_bq.ensure(nb * 4));
  BVH_TRY(w->sw_bidx.ensure(nb * 4));
  BVH_TRY(w->sw_kbs.ensure(nb * 26 * sizeof(double)));
  BVH_TRY(w->sw_bmax.ensure(nb * sizeof(double)));
  BVH_TRY(w->sw_count.ensure(na * 4));
  BVH_TRY(w->sw_total.ensure(sizeof(unsigned long long)));
  HIP_TRY(launch_sweep_prepare(dkb, nb, axis, gmin, conv, w->sw_bq.as<uint32_t>(), w->sw_bidx.as<uint32_t>(), s));
  BVH_TRY(prim_sort<uint32_t>(w, w->sw_bq.as<uint32_t>(), w->sw_bidx.as<uint32_t>(), nb));
  HIP_TRY(launch_sweep_gather(dkb, w->sw_bidx.as<uint32_t>(), nb, axis, gmin, conv, w->sw_kbs.as<double>(),
                              w->sw_bmax.as<double>(), s));
  a.na = na;
  a.nb = nb;
  a.kbs = w->sw_kbs.as<double>();
  a.bmax = w->sw_bmax.as<double>();
  a.bq = w->sw_bq.as<uint32_t>();
  a.bidx = w->sw_bidx.as<uint32_t>();
  a.axis = axis;
  a.gmin = gmin;
  a.conv = conv;
  a.count = w->sw_count.as<uint32_t>();
  a.total = w->sw_total.as<unsigned long long>();
  HIP_TRY(launch_fill_u32(w->sw_total.as<uint32_t>(), 2, 0u, s));
  HIP_TRY(launch_sweep(a, false, s));
  unsigned long long total = 0;
  BVH_TRY(copy_out(&total, a.total, sizeof(total), BVH_MEM_HOST, s));
  if (total >= (1ull << 32)) return fail(BVH_E_INVALID, "more than 2^32 - 1 pairs between two patches");
  *npairs = total;
  if (!pairs || total == 0) return BVH_OK;
  if (cap < total) return fail(BVH_E_OVERFLOW, "pair buffer holds " + std::to_string(cap) + " pairs, " +
                                                   std::to_string(total) + " needed");
  BVH_TRY(bvh_exclusive_scan_u32(w, a.count, na));
  if (mem == BVH_MEM_DEVICE) {
    a.out = pairs;
    HIP_TRY(launch_sweep(a, true, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BVH_OK;
  }
  BVH_TRY(w->sw_out.ensure(total * 2 * sizeof(uint64_t)));
  a.out = w->sw_out.as<uint64_t>();
  HIP_TRY(launch_sweep(a, true, s));
  return copy_out(pairs, a.out, total * 2 * sizeof(uint64_t), BVH_MEM_HOST, s);
}

// ---- autotuner ---------------------------------------------------------------------------
bvh_status bvh_autotune(bvh_object o, bvh_object other, bvh_tuning* best) {
  if (!o || !other || !best) return fail(BVH_E_INVALID, "null argument");
  BVH_TRY(join_all(o->w));
  if (o->w != other->w) return fail(BVH_E_INVALID, "objects of different worlds");
  bvh_world w = o->w;
  if (w->capturing) return fail(BVH_E_STATE, "autotune inside a capture");
  const bool self = o == other;
  // a self query is tuned in the kind last run on o (all pairs or cross-patch)
  const bool cross = self && o->q_cross;
  static const uint32_t kBatch[3] = {1, 2, 4};
  static const uint32_t kBlock[3] = {64, 128, 256};
  const LaunchShape keep_shape = w->shape;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  bvh_tuning b{};
  b.ms = -1.0;
  bvh_status st = BVH_OK;
  uint32_t k = 0;
  // the fp32 walk takes one record per round trip (scalar registers)
  const int nbatch = o->fp32 ? 1 : 3;
  for (int ki = 0; ki < nbatch && st == BVH_OK; ++ki)
    for (int bi = 0; bi < 3 && st == BVH_OK; ++bi)
      for (int xi = 0; xi < 2 && st == BVH_OK; ++xi) {
        w->shape.batch = o->fp32 ? keep_shape.batch : kBatch[ki];
        w->shape.block = kBlock[bi];
        w->shape.xcd_remap = xi == 1;
        st = run_query(o, other, self, true, cross);  // sizes the buffers
        double t = 0.0;
        for (int r = 0; r < 3 && st == BVH_OK; ++r) {
          (void)hipEventRecord(e0, w->stream);
          st = run_query(o, other, self, true, cross);
          (void)hipEventRecord(e1, w->stream);
          (void)hipEventSynchronize(e1);
          float f = 0.0f;
          (void)hipEventElapsedTime(&f, e0, e1);
          t = (r == 0 || f < t) ? f : t;
        }
        if (k < 24) b.tried_ms[k] = t;
        ++k;
        if (st == BVH_OK && (b.ms < 0.0 || t < b.ms)) {
          b.ms = t;
          b.batch = w->shape.batch;
          b.block_size = w->shape.block;
          b.xcd_remap = w->shape.xcd_remap ? 1 : 0;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != BVH_OK) {
    w->shape = keep_shape;
    return st;
  }
  b.tried = k;
  w->shape.batch = b.batch;
  w->shape.block = b.block_size;
  w->shape.xcd_remap = b.xcd_remap != 0;
  *best = b;
  return run_query(o, other, self, true, cross);  // leave the results of the chosen setting
}

bvh_status bvh_set_launch(bvh_world w, const bvh_tuning* shape) {
  if (!w || !shape) return fail(BVH_E_INVALID, "null argument");
  if (w->capturing) return fail(BVH_E_STATE, "launch shape changed inside a capture");
  const uint32_t bs = shape->block_size, kb = shape->batch;
  if ((bs != 64 && bs != 128 && bs != 256) || (kb != 1 && kb != 2 && kb != 4) || shape->xcd_remap > 1)
    return fail(BVH_E_INVALID, "block_size must be 64, 128 or 256, batch 1, 2 or 4, xcd_remap 0 or 1");
  w->shape.block = bs;
  w->shape.batch = kb;
  w->shape.xcd_remap = shape->xcd_remap != 0;
  return BVH_OK;
}

bvh_status bvh_get_launch(bvh_world w, bvh_tuning* shape) {
  if (!w || !shape) return fail(BVH_E_INVALID, "null argument");
  *shape = bvh_tuning{};
  shape->block_size = w->shape.block;
  shape->batch = w->shape.batch;
  shape->xcd_remap = w->shape.xcd_remap ? 1 : 0;
  return BVH_OK;
}

// ---- batching form: whole steps captured into a HIP graph ------------------------------
struct bvh_graph_s {
  bvh_world w = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t gen = 0;                     // the world's allocation generation at capture
  std::vector<CapturedQuery> queries;   // the captured queries, the last one per object
};

static void orphan_graphs(bvh_world w) {
  for (bvh_graph g : w->graphs) g->w = nullptr;
  w->graphs.clear();
}

bvh_status bvh_capture_begin(bvh_world w) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  if (w->capturing || t_capturing) return fail(BVH_E_STATE, "a capture is already open on this thread");
  HIP_TRY(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
  w->capturing = true;
  t_capturing = true;
  w->captured.clear();
  return BVH_OK;
}

bvh_status bvh_capture_end(bvh_world w, bvh_graph* out) {
  if (!w || !out) return fail(BVH_E_INVALID, "null argument");
  *out = nullptr;
  if (!w->capturing) return fail(BVH_E_STATE, "no capture open");
  // builds forked inside the capture rejoin the world stream before it ends
  const bvh_status js = join_all(w);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(w->stream, &graph);
  w->capturing = false;
  t_capturing = false;
  if (e != hipSuccess || js != BVH_OK) {
    w->captured.clear();
    if (graph) (void)hipGraphDestroy(graph);
    if (js != BVH_OK) return js;
    return fail(BVH_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  auto g = std::make_unique<bvh_graph_s>();
  g->w = w;
  g->graph = graph;
  g->gen = w->alloc_gen.load();
  // the last query captured on an object is the one whose results it holds
  for (const CapturedQuery& c : w->captured) {
    auto it = std::find_if(g->queries.begin(), g->queries.end(), [&](const CapturedQuery& x) { return x.o == c.o; });
    if (it != g->queries.end()) *it = c;
    else g->queries.push_back(c);
  }
  w->captured.clear();
  const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail(BVH_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  w->graphs.push_back(g.get());
  *out = g.release();
  return BVH_OK;
}

bvh_status bvh_graph_launch(bvh_graph g) {
  if (!g) return fail(BVH_E_INVALID, "null graph");
  if (!g->w) return fail(BVH_E_STATE, "the graph's world was destroyed");
  BVH_TRY(join_all(g->w));
  if (g->w->capturing) return fail(BVH_E_STATE, "graph launch inside a capture");
  if (g->gen != g->w->alloc_gen.load())
    return fail(BVH_E_STATE, "device buffers of this world were reallocated since the capture; capture again");
  HIP_TRY(hipGraphLaunch(g->exec, g->w->stream));
  // the replayed builds leave no exact patch slabs in fp32-outward records
  for (auto& ob : g->w->objects)
    if (ob->fp32) ob->patch_slabs_ready = false;
  // the captured queries' state, whatever the objects ran since the capture
  for (const CapturedQuery& c : g->queries) {
    bvh_object_s* o = c.o;
    o->result_other = c.other;
    o->q_self = c.self;
    o->q_narrow = c.narrow;
    o->q_cross = c.cross;
    o->self_result = c.self_result;
    o->q_pcap = c.pcap;
    o->q_cap = c.cap;
    o->q_rcap = c.rcap;
    o->result_ptr = c.result_ptr;
    o->have_results = true;
    o->results_pending = true;
    o->patch_pairs_ready = c.self || !c.narrow;  // a patch-only broadphase captured its patch pairs
  }
  return BVH_OK;
}

bvh_status bvh_graph_destroy(bvh_graph g) {
  if (!g) return BVH_OK;
  if (g->w) {
    (void)hipStreamSynchronize(g->w->stream);
    auto& v = g->w->graphs;
    v.erase(std::remove(v.begin(), v.end(), g), v.end());
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return BVH_OK;
}

bvh_status bvh_sort_u64(bvh_world w, uint64_t* keys, uint32_t* vals, uint64_t n) { return prim_sort<uint64_t>(w, keys, vals, n); }

bvh_status bvh_sort_u64_noalloc(bvh_world w, uint64_t* keys, uint32_t* vals, uint64_t n) {
  if (!w) return fail(BVH_E_INVALID, "null world");
  BVH_TRY(join_all(w));
  if (n > kNoAllocMaxKeys)
    return fail(BVH_E_INVALID, "bvh_sort_u64_noalloc sorts at most " + std::to_string(kNoAllocMaxKeys) +
                                   " keys (" + std::to_string(n) + " given); use bvh_sort_u64");
  if (w->capturing) return fail(BVH_E_STATE, "sort inside a capture");
  if (n == 0) return BVH_OK;
  if (!keys) return fail(BVH_E_INVALID, "null keys");
  if (sort_tiles(n) > kNoAllocTiles) return fail(BVH_E_INVALID, "sort tiling does not fit the scratch");
  void* base = nullptr;
  BVH_TRY(noalloc_scratch(w, &base));
  unsigned char* b = static_cast<unsigned char*>(base);
  const SortScratch sc{b + kNaKa, b + kNaKb, (uint32_t*)(b + kNaVa), (uint32_t*)(b + kNaVb),
                       (uint32_t*)(b + kNaHist), (SortCtl*)(b + kNaCtl), (uint32_t*)(b + kNaLb)};
  return sort_in_place<uint64_t>(w, keys, vals, n, sc);
}
bvh_status bvh_sort_u32(bvh_world w, uint32_t* keys, uint32_t* vals, uint64_t n) { return prim_sort<uint32_t>(w, keys, vals, n); }

}  // extern "C"
