// query.hip -- broadphase and narrowphase traversal on gfx950.
//
//  * k_narrow64s / k_narrow64 / k_narrow32: one lane per query element (the
//    query object's leaves in its sorted order, so a wave holds 64 spatially
//    adjacent elements).  The wave walks the other object's tree with a
//    wave-uniform node: each internal node's record holds both children's
//    slabs, so one memory round trip tests two nodes; leaf children are
//    tested at once and internal children that some lane overlaps go on a
//    per-wave LDS stack.  Records arrive in scalar registers (k_narrow32,
//    k_narrow64s) or, K at a time, staged in LDS (k_narrow64).
//    Hits are staged per wave in LDS and appended with one atomic per flush.
//    Replaces query_tree / query_node_impl (collision_query.hpp:166-213) as
//    used by a narrowphase functor per patch pair, and self_collision_set
//    (collision_query.hpp:331-368) in self mode (each unordered pair once).
//  * k_patch_pairs: one lane per pair of patches, testing the two patch
//    k-DOPs (start_broadphase, collision_object/broadphase.cpp:68-119, which
//    walks the other object's patch-level tree to the same set).
#include "device.hpp"
#include "kernels.hpp"

namespace dbvh {

__device__ __forceinline__ void emit(const PairOut& out, bool pred, uint64_t key) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  const int leader = __ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(out.count, (unsigned long long)__popcll(m));
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)base, leader);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(base >> 32), leader);
  base = ((unsigned long long)hi << 32) | lo;
  if (pred) {
    const unsigned long long idx = base + (unsigned long long)__popcll(m & lanemask_lt());
    if (idx < out.cap) out.pairs[idx] = key;
  }
}

__device__ __forceinline__ uint64_t pair_key(uint32_t a, uint32_t b, bool self) {
  if (self && b < a) {
    const uint32_t t = a;
    a = b;
    b = t;
  }
  return ((uint64_t)a << 32) | b;
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ void empty_slabs(double* k) {
#pragma unroll
  for (int a = 0; a < kAxes; ++a) {
    k[2 * a] = __builtin_inf();
    k[2 * a + 1] = -__builtin_inf();
  }
}

// Per-wave pair staging in LDS: hits are appended with a ballot (all lanes of
// the wave are converged in the uniform traversal loop) and written to
// global memory in coalesced runs with one atomic per flush.
constexpr uint32_t kWaveBuf = 512;

__device__ __forceinline__ void wave_flush(uint64_t* buf, uint32_t& fill, const PairOut& out) {
  if (fill == 0) return;
  __builtin_amdgcn_wave_barrier();
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(out.count, (unsigned long long)fill);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)base);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(base >> 32));
  base = ((unsigned long long)hi << 32) | lo;
  for (uint32_t i = lane_id(); i < fill; i += kWave)
    if (base + i < out.cap) out.pairs[base + i] = buf[i];
  __builtin_amdgcn_wave_barrier();
  fill = 0;
}

// End of a kernel whose waves all reach this point together (every thread of
// the block calls it): the block's waves append their staged pairs with one
// atomic for the whole block (the pair counter is one address; its atomics
// are serialized, about 12 ns each).
__device__ __forceinline__ void block_flush(uint64_t* buf, uint32_t fill, const PairOut& out) {
  __shared__ uint32_t s_fill[256 / kWave];
  __shared__ unsigned long long s_base;
  const uint32_t wv = threadIdx.x / kWave, nw = blockDim.x / kWave;
  if (lane_id() == 0) s_fill[wv] = fill;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t v = 0; v < nw; ++v) tot += s_fill[v];
    s_base = tot ? atomicAdd(out.count, (unsigned long long)tot) : 0ull;
  }
  __syncthreads();
  unsigned long long base = s_base;
  for (uint32_t v = 0; v < wv; ++v) base += s_fill[v];
  for (uint32_t i = lane_id(); i < fill; i += kWave)
    if (base + i < out.cap) out.pairs[base + i] = buf[i];
}

__device__ __forceinline__ void wave_emit(uint64_t* buf, uint32_t& fill, const PairOut& out, bool pred,
                                          uint64_t key) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  const uint32_t cnt = (uint32_t)__popcll(m);
  if (fill + cnt > kWaveBuf) wave_flush(buf, fill, out);
  if (pred) buf[fill + (uint32_t)__popcll(m & lanemask_lt())] = key;
  fill += cnt;
}

// fp32-outward form.  Node slabs are fp32 in the frame of their object's
// origin P, rounded outward, so an fp32 miss proves an fp64 miss: with q, t the exact slabs and f(.) the rounded ones,
// q_lo < t_hi implies f(q_lo) <= q_lo - P < t_hi - P <= f(t_hi).  The query's
// slabs are its own fp32 leaf slabs.
//
// A leaf that passes the fp32 test is a candidate.  The rounding is also
// bounded from the inside: a leaf slab f = next_down(rn(fl64(x - P))) has
// x - P < next_up(next_up(f)) (rn moves by at most half an ulp and fl64 by far
// less), and likewise for upper slabs.  The records hold the ordered-integer
// images ord32() of the fp32 slabs, where next_up adds one and integer
// compares order like float compares, so ord(f(t_hi)) - ord(f(q_lo)) > 6 on
// a slab pair proves q_lo < t_hi in fp64, and both the overlap test and this
// bound are integer compares against scalar registers.  Candidates proven on every
// slab are emitted; the rest go to a re-check list whose pairs are tested on
// fp64 k-DOPs recomputed from element data (k_recheck).  The pairs are
// exactly those of the fp64 form.
//
// The tree is stored as child-pair records: visiting an internal node loads
// both children's slabs with one round of scalar loads, tests both, handles
// leaf children at once and pushes the internal children some lane overlaps
// on a per-wave LDS stack (left on top: depth-first order).
typedef __attribute__((address_space(4))) const Rec32 CRec32;
typedef __attribute__((address_space(4))) const int32_t CInt;

constexpr int32_t kInnerUlps = 6;

// stack entries per wave: above the height of any radix tree over the
// augmented keys (14 patch bits + 63 code bits + 30 position bits)
constexpr uint32_t kStack = 128;

// A walk of the fp32 records over a tree of 2..kFrontierMax patches starts at
// the patch roots whose slabs overlap the union of the wave's query slabs
// (one 26-compare test per patch, each lane testing every 64th patch) instead
// of at the root: the visits above the patch roots are saved (on 2 x 2M
// triangles in 256 patches, 32 of 200 per wave).  A wave's
// queries are 64 consecutive positions of their own tree, so the union is
// small and overlaps few patches.
constexpr uint32_t kFrontierMax = 256;
constexpr uint32_t kFrontierWords = kFrontierMax / kWave;
constexpr uint32_t kFrontierPush = 32;

// Minimum / maximum of v over the wave (every lane active), in a scalar
// register: DPP within each row of 16 lanes (quad_perm [1,0,3,2], [2,3,0,1],
// row_ror 4, 8), then row_bcast 15 and 31 carry the rows' results to lane 63.
template <bool MAX>
__device__ __forceinline__ int32_t wave_reduce_i32(int32_t v) {
  auto op = [](int32_t x, int32_t y) { return MAX ? max(x, y) : min(x, y); };
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x124, 0xF, 0xF, false));
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x128, 0xF, 0xF, false));
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));
  return __builtin_amdgcn_readlane(v, 63);
}


// Smallest gap over the 13 slab pairs between query q and box b, in ordered
// fp32 images: min(b_hi - q_lo, q_hi - b_lo) with saturating subtraction
// (which keeps every sign).  The boxes overlap strictly iff the gap is > 0,
// and the fp64 overlap of the exact slabs is proven iff it is > kInnerUlps.
template <typename P>
__device__ __forceinline__ int32_t min_gap(const int32_t* q, P b) {
  int32_t m = 0x7fffffff;
#pragma unroll
  for (int a = 0; a < kAxes; ++a) {
    const int32_t d1 = __builtin_elementwise_sub_sat(b[2 * a + 1], q[2 * a]);
    const int32_t d2 = __builtin_elementwise_sub_sat(q[2 * a + 1], b[2 * a]);
    m = min(m, min(d1, d2));
  }
  return m;
}

// The same v_cmpx chain on ordered fp32 images (int32 compares): the strict
// overlap of the lane's query slabs with one wave-uniform box, for internal
// children of the fp32 walk (a leaf child needs the gap, min_gap).  Returns
// the mask of the lanes that hit (EXEC after the chain; every lane of the
// wave is active in the walk), which stays in scalar registers: an internal
// child only needs to know whether any lane hits it.
__device__ __forceinline__ uint64_t box_mask32(const int32_t* q, const int32_t* b) {
  unsigned long long res, save, m;
  asm volatile(
      "s_mov_b64 %[save], exec\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q0], %[b1]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b0], %[q1]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q2], %[b3]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b2], %[q3]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q4], %[b5]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b4], %[q5]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q6], %[b7]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b6], %[q7]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q8], %[b9]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b8], %[q9]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q10], %[b11]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b10], %[q11]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q12], %[b13]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b12], %[q13]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q14], %[b15]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b14], %[q15]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q16], %[b17]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b16], %[q17]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q18], %[b19]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b18], %[q19]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q20], %[b21]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b20], %[q21]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q22], %[b23]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b22], %[q23]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[q24], %[b25]\n\t"
      "v_cmpx_lt_i32_e64 %[m], %[b24], %[q25]\n\t"
      "s_nop 4\n\t"
      "s_mov_b64 %[res], exec\n\t"
      "s_mov_b64 exec, %[save]\n\t"
      : [res] "=&s"(res), [save] "=&s"(save), [m] "=&s"(m)
      : [q0] "v"(q[0]), [q1] "v"(q[1]), [q2] "v"(q[2]), [q3] "v"(q[3]), [q4] "v"(q[4]), [q5] "v"(q[5]), [q6] "v"(q[6]), [q7] "v"(q[7]), [q8] "v"(q[8]), [q9] "v"(q[9]), [q10] "v"(q[10]), [q11] "v"(q[11]), [q12] "v"(q[12]), [q13] "v"(q[13]), [q14] "v"(q[14]), [q15] "v"(q[15]), [q16] "v"(q[16]), [q17] "v"(q[17]), [q18] "v"(q[18]), [q19] "v"(q[19]), [q20] "v"(q[20]), [q21] "v"(q[21]), [q22] "v"(q[22]), [q23] "v"(q[23]), [q24] "v"(q[24]), [q25] "v"(q[25]),
        [b0] "s"(b[0]), [b1] "s"(b[1]), [b2] "s"(b[2]), [b3] "s"(b[3]), [b4] "s"(b[4]), [b5] "s"(b[5]), [b6] "s"(b[6]), [b7] "s"(b[7]), [b8] "s"(b[8]), [b9] "s"(b[9]), [b10] "s"(b[10]), [b11] "s"(b[11]), [b12] "s"(b[12]), [b13] "s"(b[13]), [b14] "s"(b[14]), [b15] "s"(b[15]), [b16] "s"(b[16]), [b17] "s"(b[17]), [b18] "s"(b[18]), [b19] "s"(b[19]), [b20] "s"(b[20]), [b21] "s"(b[21]), [b22] "s"(b[22]), [b23] "s"(b[23]), [b24] "s"(b[24]), [b25] "s"(b[25]));
  return res;
}

// Per-wave LDS staging of the candidates queued for the fp64 re-check: one
// atomic on the wave's stripe counter per flush (mesh self collision queues
// every pair of elements that share a vertex, so these are many).
constexpr uint32_t kRecheckBuf = 128;

__device__ __forceinline__ void recheck_flush(uint64_t* buf, uint32_t& fill, const PairOut& out, uint32_t stripe) {
  if (fill == 0) return;
  __builtin_amdgcn_wave_barrier();
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(out.rcount + (uint64_t)stripe * kRecheckStride, (unsigned long long)fill);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)base);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(base >> 32));
  base = ((unsigned long long)hi << 32) | lo;
  uint64_t* list = out.recheck + (uint64_t)stripe * out.rcap;
  for (uint32_t i = lane_id(); i < fill; i += kWave)
    if (base + i < out.rcap) list[base + i] = buf[i];
  __builtin_amdgcn_wave_barrier();
  fill = 0;
}

// The self walks take CROSS = true for cross-patch self collision: a lane
// skips every child whose position range lies inside its own patch's range
// [pb, pe) (patches are contiguous position ranges of the tree order), so
// pairs within one patch are neither tested nor emitted.
__device__ __forceinline__ void own_patch_range(const TreeView& q, uint32_t qa, bool valid, uint32_t n, uint32_t& pb,
                                                uint32_t& pe) {
  pb = 0;
  pe = 0;
  if (!valid) return;
  const uint32_t k = q.patch_of_input[qa];
  pb = k == 0 ? 0u : q.splits[k - 1];
  pe = (uint64_t)k + 1 >= q.od ? n : q.splits[k];
}

// The fp32-outward record walk of one tile of blockDim.x queries (the body of
// k_narrow32 and of its batched form k_narrow32_batch); buf / rbuf / stk are
// the wave's LDS pair staging, re-check staging and stack, rstripe the
// wave's re-check stripe.  Leaf candidates with a gap above `inner` are
// emitted, the rest queued for the fp64 re-check.
template <bool SELF, bool CROSS>
__device__ __forceinline__ void walk32(const TreeView& q, const TreeView& t, uint32_t tile, const PairOut& out,
                                       int32_t inner, uint64_t* buf, uint64_t* rbuf, uint32_t* stk,
                                       uint32_t rstripe) {
  uint32_t fill = 0, rfill = 0;
  const uint64_t p = (uint64_t)tile * blockDim.x + threadIdx.x;
  const bool valid = p < q.n;
  if (q.origin != t.origin) inner = 0x7fffffff;
  int32_t qf[kSlabs];  // ord32 images of the query's outward fp32 slabs
  uint32_t qa = 0;
  if (valid) {
    // this leaf's slabs: its half of its parent's record (the root's array if n = 1)
    const int32_t* src = q.rootbox32;
    if (q.n > 1) {
      uint32_t parent;
      int side;
      leaf_slot(q.delta, q.n, p, parent, side);
      src = q.rec32[parent].h[side].box;
    }
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      const int2 v = ((const int2*)src)[a];
      qf[2 * a] = v.x;
      qf[2 * a + 1] = v.y;
    }
    qa = q.perm[p];
    // another object's tree has its own frame: the slabs moved to it, still
    // outward (their fp32 rounding then bounds nothing from the inside, so
    // every candidate of such a query is decided in fp64)
    if (q.origin != t.origin)
      reframe32(qf, q.origin[0], q.origin[1], q.origin[2], t.origin[0], t.origin[1], t.origin[2]);
  } else {
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {  // the empty box overlaps nothing
      qf[2 * a] = 0x7fffffff;
      qf[2 * a + 1] = (int32_t)0x80000000;
    }
  }
  const uint32_t n = (uint32_t)t.n;
  const uint32_t tleaf0 = n - 1;
  const uint32_t lane = lane_id();
  const uint32_t p32 = valid ? (uint32_t)p : 0u;  // positions fit 32 bits; invalid lanes hit nothing
  uint32_t pb = 0, pe = 0;
  if (CROSS) own_patch_range(q, qa, valid, n, pb, pe);
  CRec32* recs = (CRec32*)t.rec32;
  uint32_t sink = 0, pf = 0;  // record prefetch (below)
  uint32_t steps = 0;
  uint32_t ncand = 0;
  // fp32 candidates at one (wave-uniform) leaf, gap g per lane: emit what
  // the inner bound proves, queue the rest for the fp64 re-check
  auto leaf_check = [&](int32_t g, uint32_t elem, bool h) {
    const uint64_t m = __ballot(h);
    if (m == 0) return;
    ncand += (uint32_t)__popcll(m);
    const bool sure = h & (g > inner);
    const uint64_t key = pair_key(qa, elem, SELF);
    wave_emit(buf, fill, out, sure, key);
    const bool unsure = h & !sure;
    const uint64_t mu = __ballot(unsure);
    if (mu != 0) {
      const uint32_t cnt = (uint32_t)__popcll(mu);
      if (rfill + cnt > kRecheckBuf) recheck_flush(rbuf, rfill, out, rstripe);
      if (unsure) rbuf[rfill + (uint32_t)__popcll(mu & lanemask_lt())] = key;
      rfill += cnt;
    }
  };
  // the node to visit next (kEnd: pop the stack)
  uint32_t x = kEnd;
  uint32_t sp = 0;
  const uint32_t od = (uint32_t)t.od;
  bool from_root = true;
  // (not in the cross-patch self walk: its workloads have many patches, and
  // the start's registers would cost that kernel a wave per SIMD)
  if (!CROSS && t.patch_box32 != nullptr && n > 1 && od > 1 && od <= kFrontierMax) {
    // self: patches that end at or before the wave's first position hold no
    // pair of it
    const uint32_t pmin = __builtin_amdgcn_readlane(p32, 0);
    // lane l tests patches l, l + 64, ... (clamped to the last patch, with
    // the test off); the union is reduced one axis at a time, so that it
    // takes two scalar registers and each axis's loads go out together
    uint32_t kk[kFrontierWords];
    bool h[kFrontierWords];
#pragma unroll
    for (uint32_t w = 0; w < kFrontierWords; ++w) {
      const uint32_t k = w * kWave + lane;
      kk[w] = (k < od ? k : od - 1) & (kFrontierMax - 1);  // (the mask: a 32-bit load offset)
      h[w] = (k < od) & (t.patch_root[kk[w]] != kEnd);
      if (SELF) h[w] = h[w] & ((kk[w] + 1 >= od ? n : t.splits[kk[w]]) > pmin + 1);
    }
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      const int32_t ulo = wave_reduce_i32<false>(qf[2 * a]), uhi = wave_reduce_i32<true>(qf[2 * a + 1]);
      const int32_t* plo = t.patch_box32 + (uint64_t)(2 * a) * od;  // this axis's slabs of every patch
      const int32_t* phi = plo + od;
#pragma unroll
      for (uint32_t w = 0; w < kFrontierWords; ++w) {
        if (w * kWave >= od) break;
        h[w] = h[w] & (ulo < phi[kk[w]]) & (plo[kk[w]] < uhi);
      }
    }
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t w = 0; w < kFrontierWords; ++w) cnt += (uint32_t)__popcll(__ballot(h[w]));
    // the patch roots hit go on the stack; within one patch's subtree the
    // walk stacks fewer than 96 entries (its height is below the 93 key bits
    // under the patch bits), so kFrontierPush roots leave the stack room; a
    // wave hitting more patches starts at the root
    if (cnt <= kFrontierPush) {
      from_root = false;
      uint32_t leafm = 0;  // bit w: this lane's patch of round w is one element, hit
#pragma unroll
      for (uint32_t w = 0; w < kFrontierWords; ++w) {
        const uint32_t r = h[w] ? t.patch_root[kk[w]] : kEnd;
        const bool in = h[w] & (r < tleaf0);
        const uint64_t mi = __ballot(in);
        if (in) stk[sp + (uint32_t)__popcll(mi & lanemask_lt())] = r;
        sp += (uint32_t)__popcll(mi);
        leafm |= (uint32_t)(h[w] & (r >= tleaf0)) << w;
      }
      // a patch of one element: its leaf is tested here
      uint64_t lm = __ballot(leafm != 0);
      while (lm != 0) {
        const uint32_t j = (uint32_t)__builtin_ctzll(lm);
        lm &= lm - 1;
        uint32_t bits = __builtin_amdgcn_readlane(leafm, j);
        while (bits != 0) {
          const uint32_t k = (uint32_t)__builtin_ctz(bits) * kWave + j;
          bits &= bits - 1;
          const uint32_t pos = uniform(t.patch_root[k]) - tleaf0;
          if (pos >= n) {
            if (lane == 0) atomicOr(t.status, kErrTraverse);
            continue;
          }
          int32_t bx[kSlabs];
#pragma unroll
          for (int a = 0; a < kSlabs; ++a) bx[a] = t.patch_box32[(uint64_t)a * od + k];
          const int32_t g = min_gap(qf, bx);
          bool hit = g > 0;
          if (SELF) hit = hit & (pos > p32);
          leaf_check(g, uniform(t.perm[pos]), hit);
        }
      }
    }
  }
  if (from_root) {
    const int32_t g = min_gap(qf, (CInt*)t.rootbox32);
    bool h = g > 0;
    if (SELF) h = h & ((n - 1) > p32);  // the root's last position
    if (n == 1) {
      if (!SELF) leaf_check(g, uniform(t.perm[0]), h);
    } else if (__ballot(h) != 0) {
      x = uniform(*t.root);
    }
  }
  for (;;) {
    if (x == kEnd) {
      if (sp == 0) break;
      x = uniform(stk[--sp]);
    }
    if (x >= tleaf0 || ++steps > n) {
      if (lane == 0) atomicOr(t.status, kErrTraverse);
      break;
    }
    // both halves' slabs, child ids and leaf elements (the first 112 bytes of
    // each 128-byte half) and `last`: every scalar load issued before any is
    // used, one memory round trip per visit
    CInt* rp = (CInt*)(recs + x);
    int32_t rw[56];
#pragma unroll
    for (int k = 0; k < 28; ++k) rw[k] = rp[k];
#pragma unroll
    for (int k = 0; k < 28; ++k) rw[28 + k] = rp[32 + k];
    const uint32_t rlast = SELF ? recs[x].h[0].last : 0u;
    const uint32_t rfirst = CROSS ? recs[x].h[0].first : 0u;
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t cl = (uint32_t)rw[26], cr = (uint32_t)rw[54];
    const uint32_t el = (uint32_t)rw[27], er = (uint32_t)rw[55];
    // internal children: overlap only (v_cmpx chain); leaf children: the gap,
    // which also decides whether fp32 proves the fp64 overlap
    const bool leaf_l = cl >= tleaf0, leaf_r = cr >= tleaf0;
    // The next record the walk loads is one of these children's (or one it
    // pushed earlier).  The lanes load a word of each 128-byte line of both
    // children's records (lanes 4j + 0, 1: the left child, 4j + 2, 3: the
    // right; a leaf child's lanes read this record again), so the lines are
    // on their way to L2 while this record is tested; the words feed a sink
    // that nothing reads.  Every lane loads: no branch around the load.
    {
      sink += pf;
      uint32_t c = (lane & 2) ? cr : cl;
      c = c < tleaf0 ? c : x;
      pf = ((const uint32_t*)(recs + c))[(lane & 1) * 32];
    }
    // per lane, whether a child may hold its pairs at all: position bounds
    // (self) and the own patch (cross), non-short-circuit &: no branch
    bool okl = true, okr = true;
    if (SELF) {
      okl = x > p32;  // the left child ends at position x
      okr = rlast > p32;
    }
    if (CROSS) {  // children [rfirst, x] and [x + 1, rlast] inside the lane's own patch
      okl = okl & !((rfirst >= pb) & (x < pe));
      okr = okr & !((x + 1 >= pb) & (rlast < pe));
    }
    // the lanes hitting each child, as scalar masks (the v_cmpx chain tests
    // the same strict overlap as gap > 0); a leaf child that some lane hits
    // also gets each lane's gap for the candidate test (both gaps before
    // either test, so the record's registers are free by then)
    int32_t gl = 0, gr = 0;
    uint64_t ml = box_mask32(qf, rw), mr = box_mask32(qf, rw + 28);
    if (SELF || CROSS) {
      ml &= __ballot(okl);
      mr &= __ballot(okr);
    }
    if (leaf_l && ml != 0) gl = min_gap(qf, rw);
    if (leaf_r && mr != 0) gr = min_gap(qf, rw + 28);
    const bool any_l = ml != 0, any_r = mr != 0;
    if (any_l && leaf_l) leaf_check(gl, el, (gl > 0) & okl & (!SELF || (cl - tleaf0 > p32)));
    if (any_r && leaf_r) leaf_check(gr, er, (gr > 0) & okr & (!SELF || (cr - tleaf0 > p32)));
    const bool go_l = any_l && !leaf_l, go_r = any_r && !leaf_r;
    // depth first: continue with the left child, keep the right one
    if (go_l && go_r) {
      if (sp >= kStack) {
        if (lane == 0) atomicOr(t.status, kErrTraverse);
        break;
      }
      stk[sp++] = cr;
      x = cl;
    } else {
      x = go_l ? cl : (go_r ? cr : kEnd);
    }
  }
  wave_flush(buf, fill, out);
  recheck_flush(rbuf, rfill, out, rstripe);
  sink += pf;
  asm volatile("; record prefetch sink %0" ::"v"(sink));  // keeps the prefetch loads
  if (lane == 0) {
    if (out.visits) atomicAdd(out.visits, (unsigned long long)steps);
    if (out.candidates) atomicAdd(out.candidates, (unsigned long long)ncand);
  }
}

template <bool SELF, bool CROSS = false>
__global__ __launch_bounds__(256) void k_narrow32(TreeView q, TreeView t, int xcd, PairOut out, int32_t inner) {
  __shared__ uint64_t s_buf[256 / kWave][kWaveBuf];
  __shared__ uint64_t s_rbuf[256 / kWave][kRecheckBuf];
  __shared__ uint32_t s_stack[256 / kWave][kStack];
  const uint32_t wv = threadIdx.x / kWave;
  const uint32_t rstripe = (blockIdx.x * (blockDim.x / kWave) + wv) % kRecheckStripes;
  const uint32_t tile = xcd ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
  walk32<SELF, CROSS>(q, t, tile, out, inner, s_buf[wv], s_rbuf[wv], s_stack[wv], rstripe);
}

// Several queries (object pairs) in one launch of the fp32-outward walk:
// block b walks tile b - block0 of the query whose [block0, block0 +
// blocks) holds b (a wave-uniform binary search over the table).
__global__ __launch_bounds__(256) void k_narrow32_batch(const BatchQuery* __restrict__ bq, uint32_t nq,
                                                        int32_t inner) {
  __shared__ uint64_t s_buf[256 / kWave][kWaveBuf];
  __shared__ uint64_t s_rbuf[256 / kWave][kRecheckBuf];
  __shared__ uint32_t s_stack[256 / kWave][kStack];
  const uint32_t wv = threadIdx.x / kWave;
  uint32_t lo = 0, hi = nq;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (bq[mid].block0 <= blockIdx.x) lo = mid;
    else hi = mid;
  }
  const BatchQuery& e = bq[lo];
  if (blockIdx.x - e.block0 >= e.blocks) return;
  const uint32_t tile = blockIdx.x - e.block0;
  const uint32_t rstripe = (tile * (blockDim.x / kWave) + wv) % kRecheckStripes;
  walk32<false, false>(e.q, e.t, tile, e.out, inner, s_buf[wv], s_rbuf[wv], s_stack[wv], rstripe);
}

// Strict overlap of the lane's query slabs q with one box b whose 26 slabs
// are wave-uniform (scalar registers), as the AND of 26 fp64 compares chained
// with v_cmpx: each compare clears the EXEC bit of the lanes it fails (a
// lane already cleared stays cleared), so the AND accumulates in EXEC with no
// scalar instruction per compare.  EXEC is saved first and restored last;
// the caller runs it with every lane of the wave active.  Returns the wave's
// hit mask; *hit is the lane's bit.  `touch` are loads whose only purpose is
// to bring cache lines in: they are consumed here so that they are kept.
__device__ __forceinline__ uint64_t box_hits(const double* q, const double* b, uint32_t touch0, uint32_t touch1,
                                             bool* hit) {
  unsigned long long res, save, m;
  int v;
  asm volatile(
      "s_mov_b64 %[save], exec\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q0], %[b1]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b0], %[q1]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q2], %[b3]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b2], %[q3]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q4], %[b5]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b4], %[q5]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q6], %[b7]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b6], %[q7]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q8], %[b9]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b8], %[q9]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q10], %[b11]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b10], %[q11]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q12], %[b13]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b12], %[q13]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q14], %[b15]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b14], %[q15]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q16], %[b17]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b16], %[q17]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q18], %[b19]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b18], %[q19]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q20], %[b21]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b20], %[q21]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q22], %[b23]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b22], %[q23]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[q24], %[b25]\n\t"
      "v_cmpx_lt_f64_e64 %[m], %[b24], %[q25]\n\t"
      "s_nop 4\n\t"
      "s_mov_b64 %[res], exec\n\t"
      "s_mov_b64 exec, %[save]\n\t"
      "v_cndmask_b32_e64 %[v], 0, 1, %[res]\n\t"
      : [res] "=&s"(res), [save] "=&s"(save), [m] "=&s"(m), [v] "=&v"(v)
      : [q0] "v"(q[0]), [q1] "v"(q[1]), [q2] "v"(q[2]), [q3] "v"(q[3]), [q4] "v"(q[4]), [q5] "v"(q[5]), [q6] "v"(q[6]), [q7] "v"(q[7]), [q8] "v"(q[8]), [q9] "v"(q[9]), [q10] "v"(q[10]), [q11] "v"(q[11]), [q12] "v"(q[12]), [q13] "v"(q[13]), [q14] "v"(q[14]), [q15] "v"(q[15]), [q16] "v"(q[16]), [q17] "v"(q[17]), [q18] "v"(q[18]), [q19] "v"(q[19]), [q20] "v"(q[20]), [q21] "v"(q[21]), [q22] "v"(q[22]), [q23] "v"(q[23]), [q24] "v"(q[24]), [q25] "v"(q[25]),
        [b0] "s"(b[0]), [b1] "s"(b[1]), [b2] "s"(b[2]), [b3] "s"(b[3]), [b4] "s"(b[4]), [b5] "s"(b[5]), [b6] "s"(b[6]), [b7] "s"(b[7]), [b8] "s"(b[8]), [b9] "s"(b[9]), [b10] "s"(b[10]), [b11] "s"(b[11]), [b12] "s"(b[12]), [b13] "s"(b[13]), [b14] "s"(b[14]), [b15] "s"(b[15]), [b16] "s"(b[16]), [b17] "s"(b[17]), [b18] "s"(b[18]), [b19] "s"(b[19]), [b20] "s"(b[20]), [b21] "s"(b[21]), [b22] "s"(b[22]), [b23] "s"(b[23]), [b24] "s"(b[24]), [b25] "s"(b[25]),
        [t0] "s"(touch0), [t1] "s"(touch1));
  *hit = v != 0;
  return res;
}

// Both boxes of an LDS-staged fp64 record against the lane's query slabs,
// one axis at a time: the scheduling barrier after each axis keeps four of
// the record's doubles in registers, which holds the walk at 88-94 VGPRs
// (5 waves per SIMD) where loading both boxes at once needs over 128.
__device__ __forceinline__ void overlap_pair(const double* qk, const Rec64& r, bool& h0, bool& h1) {
  bool o0 = true, o1 = true;
#pragma unroll
  for (int a = 0; a < kAxes; ++a) {
    o0 = o0 & (qk[2 * a] < r.box[0][2 * a + 1]) & (r.box[0][2 * a] < qk[2 * a + 1]);
    o1 = o1 & (qk[2 * a] < r.box[1][2 * a + 1]) & (r.box[1][2 * a] < qk[2 * a + 1]);
    __builtin_amdgcn_sched_barrier(0);
  }
  h0 = o0;
  h1 = o1;
}

// fp64 form: the same record walk, K records per memory round trip.  Two
// fp64 boxes (416 B) do not fit the scalar registers, so each 448-byte record
// is loaded into LDS by a direct-to-LDS load (28 lanes x 16 bytes, no VGPR
// staging, which keeps the walk at 80 VGPRs and six waves per SIMD for K = 2)
// and every lane reads both boxes as LDS broadcasts.  Each iteration pops up to K nodes
// off the wave's stack and issues all their loads before it tests any: the
// set of visited nodes does not depend on K (a node is visited when some lane
// overlaps it and its parent was visited), only the number of dependent
// round trips does.  Leaf tests are exact, so hits are emitted directly.
// The walk of one tile of blockDim.x queries (the body of k_narrow64 and of
// its batched form k_narrow64_batch); buf / stk / R are the wave's LDS pair
// staging, stack and record slots.
template <bool SELF, int K, bool CROSS>
__device__ __forceinline__ void walk64(const TreeView& q, const TreeView& t, uint32_t tile, const PairOut& out,
                                       uint64_t* buf, uint32_t* stk, Rec64* R) {
  uint32_t fill = 0;
  const uint64_t p = (uint64_t)tile * blockDim.x + threadIdx.x;
  const bool valid = p < q.n;
  double qk[kSlabs];
  uint32_t qa = 0;
  if (valid) {
    // this leaf's slabs: its half of its parent's record (the root's array if n = 1)
    const double* src = q.rootbox64;
    if (q.n > 1) {
      uint32_t parent;
      int side;
      leaf_slot(q.delta, q.n, p, parent, side);
      src = q.rec64[parent].box[side];
    }
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      const double2 v = ((const double2*)src)[a];
      qk[2 * a] = v.x;
      qk[2 * a + 1] = v.y;
    }
    qa = q.perm[p];
  } else {
    empty_slabs(qk);
  }
  const uint32_t n = (uint32_t)t.n;
  const uint32_t tleaf0 = n - 1;
  const uint32_t lane = lane_id();
  const uint32_t p32 = valid ? (uint32_t)p : 0u;  // positions fit 32 bits; invalid lanes hit nothing
  uint32_t pb = 0, pe = 0;
  if (CROSS) own_patch_range(q, qa, valid, n, pb, pe);
  uint32_t steps = 0;
  uint32_t sp = 0;
  {
    // the root's slabs through scalar loads (wave-uniform)
    const __attribute__((address_space(4))) double* rb = (const __attribute__((address_space(4))) double*)t.rootbox64;
    bool h = true;
#pragma unroll
    for (int a = 0; a < kAxes; ++a) h = h & (qk[2 * a] < rb[2 * a + 1]) & (rb[2 * a] < qk[2 * a + 1]);
    if (SELF) h = h && (uint64_t)(n - 1) > p;  // the root's last position
    if (n == 1) {
      if (!SELF) wave_emit(buf, fill, out, h, pair_key(qa, uniform(t.perm[0]), SELF));
    } else if (__ballot(h) != 0) {
      stk[0] = uniform(*t.root);
      sp = 1;
    }
  }
  while (sp != 0) {
    const uint32_t m = sp < (uint32_t)K ? sp : (uint32_t)K;
    uint32_t xs[K];
#pragma unroll
    for (int k = 0; k < K; ++k) xs[k] = (uint32_t)k < m ? uniform(stk[sp - 1 - k]) : 0u;
    sp -= m;
    steps += m;
    bool bad = steps > n;
#pragma unroll
    for (int k = 0; k < K; ++k) bad = bad || ((uint32_t)k < m && xs[k] >= tleaf0);
    if (bad) {
      if (lane == 0) atomicOr(t.status, kErrTraverse);
      break;
    }
    {
      // every record straight from memory into its LDS slot (the loads
      // write LDS directly, so no VGPRs hold them in between): lanes 0..27
      // each bring 16 bytes, which land at the slot + 16 x lane
#pragma unroll
      for (int k = 0; k < K; ++k)
        if ((uint32_t)k < m && lane < sizeof(Rec64) / 16)
          __builtin_amdgcn_global_load_lds((const void*)((const char*)(t.rec64 + xs[k]) + lane * 16),
                                           (void __attribute__((address_space(3)))*)&R[k], 16, 0, 0);
      __builtin_amdgcn_s_waitcnt(0);  // the LDS writes of those loads are done
      asm volatile("" ::: "memory");
      __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if ((uint32_t)k < m) {
        const Rec64& Rk = R[k];
        const uint32_t cl = uniform(Rk.child[0]), cr = uniform(Rk.child[1]);
        const uint32_t el = uniform(Rk.elem[0]), er = uniform(Rk.elem[1]);
        bool hl, hr;
        overlap_pair(qk, Rk, hl, hr);
        if (SELF) {
          // position bounds (non-short-circuit &: no divergent branch)
          hl = hl & (xs[k] > p32);  // the left child ends at position x
          hr = hr & (uniform(Rk.last) > p32);
        }
        if (CROSS) {  // children inside the lane's own patch
          const uint32_t f = uniform(Rk.first), l = uniform(Rk.last);
          hl = hl & !((f >= pb) & (xs[k] < pe));
          hr = hr & !((xs[k] + 1 >= pb) & (l < pe));
        }
        // ballots and readfirstlane values: every branch below is wave-uniform,
        // so the stack pointer and counters stay in scalar registers
        const bool any_l = __ballot(hl) != 0, any_r = __ballot(hr) != 0;
        const bool leaf_l = cl >= tleaf0, leaf_r = cr >= tleaf0;
        if (any_l && leaf_l)
          wave_emit(buf, fill, out, SELF ? (hl & (cl - tleaf0 > p32)) : hl, pair_key(qa, el, SELF));
        if (any_r && leaf_r)
          wave_emit(buf, fill, out, SELF ? (hr & (cr - tleaf0 > p32)) : hr, pair_key(qa, er, SELF));
        // the left child ends on top of the stack
        if (sp + 2 > kStack) {
          bad = true;
        } else {
          if (any_r && !leaf_r) stk[sp++] = cr;
          if (any_l && !leaf_l) stk[sp++] = cl;
        }
      }
      // one record's boxes in registers at a time
      __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_wave_barrier();  // every lane has read the records
    if (bad) {
      if (lane == 0) atomicOr(t.status, kErrTraverse);
      break;
    }
  }
  wave_flush(buf, fill, out);
  if (lane == 0 && out.visits) atomicAdd(out.visits, (unsigned long long)steps);
}

template <bool SELF, int K, bool CROSS = false>
__global__ __launch_bounds__(256) void k_narrow64(TreeView q, TreeView t, int xcd, PairOut out) {
  __shared__ uint64_t s_buf[256 / kWave][kWaveBuf];
  __shared__ uint32_t s_stack[256 / kWave][kStack];
  __shared__ Rec64 s_rec[256 / kWave][K];
  const uint32_t wv = threadIdx.x / kWave;
  const uint32_t tile = xcd ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
  walk64<SELF, K, CROSS>(q, t, tile, out, s_buf[wv], s_stack[wv], s_rec[wv]);
}

// Several queries (object pairs) in one launch: block b walks tile
// b - block0 of the query whose [block0, block0 + tiles) holds b (a
// wave-uniform binary search over the table).
template <int K>
__global__ __launch_bounds__(256) void k_narrow64_batch(const BatchQuery* __restrict__ bq, uint32_t nq) {
  __shared__ uint64_t s_buf[256 / kWave][kWaveBuf];
  __shared__ uint32_t s_stack[256 / kWave][kStack];
  __shared__ Rec64 s_rec[256 / kWave][K];
  const uint32_t wv = threadIdx.x / kWave;
  uint32_t lo = 0, hi = nq;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (bq[mid].block0 <= blockIdx.x) lo = mid;
    else hi = mid;
  }
  const BatchQuery& e = bq[lo];
  if (blockIdx.x - e.block0 >= e.blocks) return;
  walk64<false, K, false>(e.q, e.t, blockIdx.x - e.block0, e.out, s_buf[wv], s_stack[wv], s_rec[wv]);
}

// fp64 form, one record per visit in scalar registers (batch 1): box 0
// with the child ids and the cache lines of box 1 in one round of scalar
// loads, then box 1 from the scalar cache; each box is tested with one
// v_cmpx chain (box_hits), so a visit issues no LDS instruction and no
// scalar instruction per compare.  The next node is kept in a register, the
// right child of a two-child visit goes on the per-wave LDS stack.
typedef __attribute__((address_space(4))) const double CDbl;

template <bool SELF, bool CROSS = false>
__global__ __launch_bounds__(256) void k_narrow64s(TreeView q, TreeView t, int xcd, PairOut out) {
  __shared__ uint64_t s_buf[256 / kWave][kWaveBuf];
  __shared__ uint32_t s_stack[256 / kWave][kStack];
  const uint32_t wv = threadIdx.x / kWave;
  uint64_t* buf = s_buf[wv];
  uint32_t* stk = s_stack[wv];
  uint32_t fill = 0;
  const uint32_t tile = xcd ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint64_t p = (uint64_t)tile * blockDim.x + threadIdx.x;
  const bool valid = p < q.n;
  double qk[kSlabs];
  uint32_t qa = 0;
  if (valid) {
    const double* src = q.rootbox64;
    if (q.n > 1) {
      uint32_t parent;
      int side;
      leaf_slot(q.delta, q.n, p, parent, side);
      src = q.rec64[parent].box[side];
    }
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      const double2 v = ((const double2*)src)[a];
      qk[2 * a] = v.x;
      qk[2 * a + 1] = v.y;
    }
    qa = q.perm[p];
  } else {
    empty_slabs(qk);
  }
  const uint32_t n = (uint32_t)t.n;
  const uint32_t tleaf0 = n - 1;
  const uint32_t lane = lane_id();
  const uint32_t p32 = valid ? (uint32_t)p : 0u;  // positions fit 32 bits; invalid lanes hit nothing
  uint32_t pb = 0, pe = 0;
  if (CROSS) own_patch_range(q, qa, valid, n, pb, pe);
  uint32_t steps = 0;
  uint32_t sp = 0;
  uint32_t x = kEnd;  // the node to visit next (kEnd: pop the stack)
  {
    const CDbl* rb = (const CDbl*)t.rootbox64;
    bool h = true;
#pragma unroll
    for (int a = 0; a < kAxes; ++a) h = h & (qk[2 * a] < rb[2 * a + 1]) & (rb[2 * a] < qk[2 * a + 1]);
    if (SELF) h = h & ((uint32_t)(n - 1) > p32);  // the root's last position
    if (n == 1) {
      if (!SELF) wave_emit(buf, fill, out, h, pair_key(qa, uniform(t.perm[0]), SELF));
    } else if (__ballot(h) != 0) {
      x = uniform(*t.root);
    }
  }
  for (;;) {
    if (x == kEnd) {
      if (sp == 0) break;
      x = uniform(stk[--sp]);
    }
    if (x >= tleaf0 || ++steps > n) {
      if (lane == 0) atomicOr(t.status, kErrTraverse);
      break;
    }
    const CDbl* rp = (const CDbl*)(t.rec64 + x);
    const CInt* ip = (const CInt*)(t.rec64 + x);
    double bx[kSlabs];
#pragma unroll
    for (int k = 0; k < kSlabs; ++k) bx[k] = rp[k];
    // child ids, leaf elements (and `last` in self mode): dwords 104..109;
    // dwords 64 and 80 only bring box 1's cache lines into the scalar cache
    const uint32_t cl = (uint32_t)ip[104], cr = (uint32_t)ip[105];
    const uint32_t el = (uint32_t)ip[106], er = (uint32_t)ip[107];
    const uint32_t rlast = SELF ? (uint32_t)ip[109] : 0u;
    const uint32_t rfirst = CROSS ? (uint32_t)ip[108] : 0u;
    const uint32_t t0 = (uint32_t)ip[64], t1 = (uint32_t)ip[80];
    __builtin_amdgcn_sched_barrier(0);
    bool hl, hr;
    box_hits(qk, bx, t0, t1, &hl);
#pragma unroll
    for (int k = 0; k < kSlabs; ++k) bx[k] = rp[kSlabs + k];
    box_hits(qk, bx, 0u, 0u, &hr);
    if (SELF) {
      hl = hl & (x > p32);  // the left child ends at position x
      hr = hr & (rlast > p32);
    }
    if (CROSS) {  // children inside the lane's own patch
      hl = hl & !((rfirst >= pb) & (x < pe));
      hr = hr & !((x + 1 >= pb) & (rlast < pe));
    }
    const bool any_l = __ballot(hl) != 0, any_r = __ballot(hr) != 0;
    const bool leaf_l = cl >= tleaf0, leaf_r = cr >= tleaf0;
    if (any_l && leaf_l)
      wave_emit(buf, fill, out, SELF ? (hl & (cl - tleaf0 > p32)) : hl, pair_key(qa, el, SELF));
    if (any_r && leaf_r)
      wave_emit(buf, fill, out, SELF ? (hr & (cr - tleaf0 > p32)) : hr, pair_key(qa, er, SELF));
    const bool go_l = any_l && !leaf_l, go_r = any_r && !leaf_r;
    // depth first: continue with the left child, keep the right one
    if (go_l && go_r) {
      if (sp >= kStack) {
        if (lane == 0) atomicOr(t.status, kErrTraverse);
        break;
      }
      stk[sp++] = cr;
      x = cl;
    } else {
      x = go_l ? cl : (go_r ? cr : kEnd);
    }
  }
  wave_flush(buf, fill, out);
  if (lane == 0 && out.visits) atomicAdd(out.visits, (unsigned long long)steps);
}

// fp64 re-check of the fp32 candidates that the inner bound left open: both
// elements' k-DOPs recomputed from element data, the pair emitted on overlap.
// The stripe lengths are read on the device, so no host round trip separates
// the two kernels.  The grid is kRecheckStripes x (blocks per stripe); block
// b works on stripe b % kRecheckStripes.  Hits are staged per wave in LDS
// (one atomic on the pair counter per flush, not per wave that hits).  Block
// 0 also adds the stripe counts' sum and maximum to rtotal for the host.
// KIND / NV: both objects' element type (the common case, with registers
// sized for it); kGeneric: any mix, chosen at run time.
constexpr int kGeneric = -1;

// blocks per re-check stripe (2,048 in all): each block strides over its
// stripe's list, so a block stages many pairs per flush
constexpr uint64_t kRecheckBlocksPerStripe = 32;

template <int KIND, int NV>
__global__ __launch_bounds__(256) void k_recheck(TreeView q, TreeView t, PairOut out) {
  __shared__ uint64_t s_buf[256 / kWave][kWaveBuf];
  uint64_t* buf = s_buf[threadIdx.x / kWave];
  uint32_t fill = 0;
  if (blockIdx.x == 0 && threadIdx.x < kRecheckStripes) {
    const unsigned long long c = out.rcount[(uint64_t)threadIdx.x * kRecheckStride];
    atomicAdd(out.rtotal, c);
    atomicMax(out.rtotal + 1, c);
  }
  const uint32_t stripe = blockIdx.x % kRecheckStripes;
  const uint32_t per = gridDim.x / kRecheckStripes;
  const unsigned long long cnt0 = out.rcount[(uint64_t)stripe * kRecheckStride];
  const uint64_t cnt = cnt0 < out.rcap ? cnt0 : out.rcap;
  const uint64_t* list = out.recheck + (uint64_t)stripe * out.rcap;
  const uint64_t stride = (uint64_t)per * blockDim.x;
  for (uint64_t base = (uint64_t)(blockIdx.x / kRecheckStripes) * blockDim.x; base < cnt; base += stride) {
    const uint64_t i = base + threadIdx.x;
    bool hit = false;
    uint64_t key = 0;
    if (i < cnt) {
      key = list[i];
      double ka[kSlabs];
      if constexpr (KIND == kGeneric) {
        element_kdop_any(q.elem, (uint32_t)(key >> 32), ka);
        hit = element_overlap_any(t.elem, (uint32_t)key, ka);
      } else {
        element_kdop<KIND, NV>(q.elem, (uint32_t)(key >> 32), ka);
        hit = element_overlap<KIND, NV>(t.elem, (uint32_t)key, ka);
      }
    }
    wave_emit(buf, fill, out, hit, key);
  }
  block_flush(buf, fill, out);
}

static void launch_recheck(const TreeView& q, const TreeView& t, const PairOut& out, unsigned blocks, hipStream_t s) {
  const ElemInput& a = q.elem;
  const ElemInput& b = t.elem;
  const dim3 g(blocks), th(256);
  if (a.kind == b.kind && (a.kind != kPoly || a.nv == b.nv)) {
    if (a.kind == kSphere) {
      hipLaunchKernelGGL((k_recheck<kSphere, 1>), g, th, 0, s, q, t, out);
      return;
    }
    if (a.kind == kSnapshot) {
      hipLaunchKernelGGL((k_recheck<kSnapshot, 1>), g, th, 0, s, q, t, out);
      return;
    }
    switch (a.nv) {
      case 3: hipLaunchKernelGGL((k_recheck<kPoly, 3>), g, th, 0, s, q, t, out); return;
      case 4: hipLaunchKernelGGL((k_recheck<kPoly, 4>), g, th, 0, s, q, t, out); return;
      case 8: hipLaunchKernelGGL((k_recheck<kPoly, 8>), g, th, 0, s, q, t, out); return;
      default: break;
    }
  }
  hipLaunchKernelGGL((k_recheck<kGeneric, 0>), g, th, 0, s, q, t, out);
}

// exact patch slabs (the union of the patch's element k-DOPs)
__device__ __forceinline__ void patch_slabs(const TreeView& v, uint64_t k, double* out) {
  const unsigned long long* e = v.patch_k64 + k * kSlabs;
#pragma unroll
  for (int i = 0; i < kSlabs; ++i) out[i] = from_ordered_u64(e[i]);
}

// Patch broadphase over all patch pairs: one lane per (patch of q, patch of
// t), t fastest, so a wave shares its q patch (a broadcast load) and reads 64
// consecutive t patches; both objects' patch slabs stay in L2.  Empty patches
// never overlap (broadphase.cpp:79-80).
__global__ __launch_bounds__(256) void k_patch_pairs(TreeView q, TreeView t, PairOut out) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t qp = k / t.od, tp = k - qp * t.od;
  bool hit = false;
  if (qp < q.od) {
    const uint32_t rq = q.patch_root[qp], rt = t.patch_root[tp];
    if (rq != kEnd && rq < 2 * q.n - 1 && rt != kEnd && rt < 2 * t.n - 1) {
      double qk[kSlabs], tk[kSlabs];
      patch_slabs(q, qp, qk);
      patch_slabs(t, tp, tk);
      hit = overlap(qk, tk);
    }
  }
  emit(out, hit, (qp << 32) | tp);
}

__global__ __launch_bounds__(256) void k_records(const uint64_t* __restrict__ pairs, uint64_t count,
                                                 const uint32_t* __restrict__ patch_a, const uint64_t* __restrict__ gid_a,
                                                 const uint32_t* __restrict__ patch_b, const uint64_t* __restrict__ gid_b,
                                                 uint64_t off_a, uint64_t off_b, uint64_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = pairs[i];
    const uint32_t a = (uint32_t)(key >> 32), b = (uint32_t)key;
    ulonglong4 r;
    r.x = patch_a[a] + off_a;
    r.y = gid_a ? gid_a[a] : a;
    r.z = patch_b[b] + off_b;
    r.w = gid_b ? gid_b[b] : b;
    ((ulonglong4*)out)[i] = r;
  }
}

// patch centroid = accumulate(element centroids in patch order, zeros) / size
// (patch.hpp:113-116), summed sequentially per patch.
__global__ __launch_bounds__(64) void k_patch_centroids(const double* __restrict__ cplanes,
                                                        const double* __restrict__ splanes,
                                                        const double* __restrict__ caos,
                                                        const uint32_t* __restrict__ perm,
                                                        const uint32_t* __restrict__ splits, uint64_t n,
                                                        uint64_t od, double* __restrict__ out) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= od) return;
  const uint64_t b = (k == 0) ? 0 : splits[k - 1];
  const uint64_t e = (k == od - 1) ? n : splits[k];
  double s[3] = {0.0, 0.0, 0.0};
  for (uint64_t pos = b; pos < e; ++pos) {
    const uint32_t src = perm[pos];
    for (int c = 0; c < 3; ++c) {
      double v;
      if (cplanes) v = cplanes[(uint64_t)c * n + src];
      else if (splanes) v = splanes[(uint64_t)c * n + src];
      else v = caos[3 * (uint64_t)src + c];
      s[c] = s[c] + v;
    }
  }
  for (int c = 0; c < 3; ++c) out[3 * k + c] = (e > b) ? s[c] / (double)(e - b) : 0.0;
}

// Order-independent digest of a pair list: the sum and the xor of two
// independent 64-bit mixes (splitmix64 finalizer) of every key.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_digest(const uint64_t* __restrict__ pairs, uint64_t count,
                                                unsigned long long* out) {
  uint64_t s = 0, x = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = pairs[i];
    s += mix64(k);
    x ^= mix64(k ^ 0x9e3779b97f4a7c15ull);
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, 64);
    x ^= __shfl_xor(x, o, 64);
  }
  // one pair of atomics per block (the two addresses serialize them)
  __shared__ uint64_t ss[256 / kWave], sx[256 / kWave];
  if (lane_id() == 0) {
    ss[threadIdx.x / kWave] = s;
    sx[threadIdx.x / kWave] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t v = 1; v < blockDim.x / kWave; ++v) {
      s += ss[v];
      x ^= sx[v];
    }
    if ((s | x) != 0) {
      atomicAdd(out, (unsigned long long)s);
      atomicXor(out + 1, (unsigned long long)x);
    }
  }
}

// ---- launchers ---------------------------------------------------------------------
template <int K>
static void narrow64_dispatch(const TreeView& q, const TreeView& t, int self, int xcd, unsigned blocks, unsigned bs,
                              const PairOut& out, hipStream_t s) {
  if (self == 2) hipLaunchKernelGGL((k_narrow64<true, K, true>), dim3(blocks), dim3(bs), 0, s, q, t, xcd, out);
  else if (self) hipLaunchKernelGGL((k_narrow64<true, K>), dim3(blocks), dim3(bs), 0, s, q, t, xcd, out);
  else hipLaunchKernelGGL((k_narrow64<false, K>), dim3(blocks), dim3(bs), 0, s, q, t, xcd, out);
}

hipError_t launch_narrowphase(const TreeView& q, const TreeView& t, int leaf_size, int self, int form,
                              const LaunchShape& shape, const PairOut& out, hipStream_t s) {
  (void)leaf_size;  // the record walk tests every leaf from its parent's record
  if (q.n == 0 || t.n == 0) return hipSuccess;
  const unsigned bs = shape.block;
  if (bs != 64 && bs != 128 && bs != 256) return hipErrorInvalidValue;
  if (!q.perm || !t.perm || !t.root || !t.status || !out.pairs || !out.count) return hipErrorInvalidValue;
  if (self == 2 && (!q.patch_of_input || (q.od > 1 && !q.splits))) return hipErrorInvalidValue;
  const unsigned blocks = (unsigned)((q.n + bs - 1) / bs);
  const int xcd = shape.xcd_remap ? 1 : 0;
  if (form == kFormFp32 || form == kFormMixed) {
    if (!q.rec32 || !q.rootbox32 || !q.delta || !t.rec32 || !t.rootbox32 || !q.origin || !t.origin)
      return hipErrorInvalidValue;
    if (!out.recheck || !out.rcount || !out.rtotal) return hipErrorInvalidValue;
    // the fp32 walk tests leaves from their parents' records: no leaf clusters
    // one record per round trip: it arrives in scalar registers (the batch
    // applies to the fp64 walk).  Both forms keep the leaf candidates whose
    // fp32 gap proves the fp64 overlap and re-check the rest in fp64.
    const int32_t inner = kInnerUlps;
    if (self == 2) hipLaunchKernelGGL((k_narrow32<true, true>), dim3(blocks), dim3(bs), 0, s, q, t, xcd, out, inner);
    else if (self) hipLaunchKernelGGL(k_narrow32<true>, dim3(blocks), dim3(bs), 0, s, q, t, xcd, out, inner);
    else hipLaunchKernelGGL(k_narrow32<false>, dim3(blocks), dim3(bs), 0, s, q, t, xcd, out, inner);
    // in self mode both keys index the one object (q == t)
    uint64_t per = (out.rcap + 255) / 256;  // blocks per stripe
    if (per > kRecheckBlocksPerStripe) per = kRecheckBlocksPerStripe;
    if (per == 0) per = 1;
    launch_recheck(q, t, out, (unsigned)(per * kRecheckStripes), s);
    return hipGetLastError();
  }
  if (!q.rec64 || !q.rootbox64 || !q.delta || !t.rec64 || !t.rootbox64) return hipErrorInvalidValue;
  switch (shape.batch) {
    case 1:
      if (self == 2) hipLaunchKernelGGL((k_narrow64s<true, true>), dim3(blocks), dim3(bs), 0, s, q, t, xcd, out);
      else if (self) hipLaunchKernelGGL(k_narrow64s<true>, dim3(blocks), dim3(bs), 0, s, q, t, xcd, out);
      else hipLaunchKernelGGL(k_narrow64s<false>, dim3(blocks), dim3(bs), 0, s, q, t, xcd, out);
      break;
    case 2: narrow64_dispatch<2>(q, t, self, xcd, blocks, bs, out, s); break;
    case 4: narrow64_dispatch<4>(q, t, self, xcd, blocks, bs, out, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The counters of a query and both objects' status words (4 x u32 each) in
// one launch, straight into pinned host memory: one small kernel instead of
// three device-to-host copies.
__global__ void k_readback(const unsigned long long* __restrict__ count, int ncount,
                           const unsigned long long* __restrict__ st_a, const unsigned long long* __restrict__ st_b,
                           unsigned long long* __restrict__ host) {
  const int i = (int)threadIdx.x;
  if (i < ncount) host[i] = count[i];
  else if (i < ncount + 2) host[i] = st_a[i - ncount];
  else if (i < ncount + 4) host[i] = st_b[i - ncount - 2];
  __threadfence_system();  // the words are in host memory before the kernel ends
}

hipError_t launch_readback(const unsigned long long* count, int ncount, const uint32_t* st_a, const uint32_t* st_b,
                           unsigned long long* host, hipStream_t s) {
  if (ncount < 0 || ncount + 4 > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_readback, dim3(1), dim3(64), 0, s, count, ncount, (const unsigned long long*)st_a,
                     (const unsigned long long*)st_b, host);
  return hipGetLastError();
}

hipError_t launch_narrowphase_batch(const BatchQuery* bq, uint32_t nq, uint32_t blocks, int batch, int form,
                                    hipStream_t s) {
  if (nq == 0 || blocks == 0) return hipSuccess;
  if (form == kFormFp32 || form == kFormMixed) {
    hipLaunchKernelGGL(k_narrow32_batch, dim3(blocks), dim3(256), 0, s, bq, nq, kInnerUlps);
  } else if (batch == 4) {
    hipLaunchKernelGGL(k_narrow64_batch<4>, dim3(blocks), dim3(256), 0, s, bq, nq);
  } else {
    hipLaunchKernelGGL(k_narrow64_batch<2>, dim3(blocks), dim3(256), 0, s, bq, nq);
  }
  return hipGetLastError();
}

hipError_t launch_recheck_query(const TreeView& q, const TreeView& t, const PairOut& out, hipStream_t s) {
  if (q.n == 0 || t.n == 0) return hipSuccess;
  if (!out.recheck || !out.rcount || !out.rtotal) return hipErrorInvalidValue;
  uint64_t per = (out.rcap + 255) / 256;  // blocks per stripe
  if (per > kRecheckBlocksPerStripe) per = kRecheckBlocksPerStripe;
  if (per == 0) per = 1;
  launch_recheck(q, t, out, (unsigned)(per * kRecheckStripes), s);
  return hipGetLastError();
}

hipError_t launch_patch_broadphase(const TreeView& q, const TreeView& t, int fp32, const PairOut& out,
                                   hipStream_t s) {
  if (q.od == 0 || t.od == 0 || t.n == 0 || q.n == 0) return hipSuccess;
  if (!q.patch_k64 || !t.patch_k64) return hipErrorInvalidValue;
  (void)fp32;
  const uint64_t total = q.od * t.od;
  hipLaunchKernelGGL(k_patch_pairs, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, q, t, out);
  return hipGetLastError();
}

hipError_t launch_records(const uint64_t* pairs, uint64_t count, const uint32_t* patch_a, const uint64_t* gid_a,
                          const uint32_t* patch_b, const uint64_t* gid_b, uint64_t off_a, uint64_t off_b,
                          uint64_t* out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_records, dim3((unsigned)blocks), dim3(256), 0, s, pairs, count, patch_a, gid_a, patch_b,
                     gid_b, off_a, off_b, out);
  return hipGetLastError();
}

hipError_t launch_digest(const uint64_t* pairs, uint64_t count, unsigned long long* out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_digest, dim3((unsigned)blocks), dim3(256), 0, s, pairs, count, out);
  return hipGetLastError();
}

hipError_t launch_patch_centroids(const double* cplanes, const double* splanes, const double* caos,
                                  const uint32_t* perm, const uint32_t* splits, uint64_t n, uint64_t od,
                                  double* out, hipStream_t s) {
  if (od == 0) return hipSuccess;
  hipLaunchKernelGGL(k_patch_centroids, dim3((unsigned)((od + 63) / 64)), dim3(64), 0, s, cplanes, splanes, caos,
                     perm, splits, n, od, out);
  return hipGetLastError();
}

}  // namespace dbvh
