// build.hip -- object build on gfx950: element bounds, Morton codes, cluster
// splits, and the bottom-up tree with refit.
//
// Stage map (reference file:line -> kernel):
//  * per-element 26-DOP + centroid, kdop.hpp:238-288 / TestCommon.hpp:62-73
//      -> k_prep (centroid, cardinal bounds), k_tree (slabs, in sorted order)
//  * compute_bounds(min_inv_diag), contact_entity.hpp:136-148
//      -> k_prep block reduction + last-block finalize
//  * morton(view, bounds, codes), contact_entity.hpp:157-191, hash.hpp
//      -> k_morton (+ digit histograms of the sort)
//  * radix_sorter, util/sort.hpp:75-144 -> sort.hip
//  * morton_cluster splits, split/cluster.hpp:143-168
//      -> k_cluster_depth + one 8-bit sort pass + k_cluster_select
//  * patches (collision_object.cpp:134-141) and the tree
//      -> k_delta + k_tree (Apetrei-style bottom-up build: each leaf climbs,
//         the second child to reach a parent merges both children's slabs)
//
// The tree is the binary radix tree of the augmented keys
// (patch id, Morton code, sorted position): every patch is a subtree, the
// levels above the patch subtrees form the patch tree, and duplicate codes
// are split by position.
#include "device.hpp"
#include "kernels.hpp"

namespace dbvh {

constexpr int kPrepBlock = 256;
constexpr int kPrepMaxBlocks = 2048;

uint64_t prep_blocks(uint64_t n) {
  uint64_t b = (n + kPrepBlock - 1) / kPrepBlock;
  if (b < 1) b = 1;
  if (b > kPrepMaxBlocks) b = kPrepMaxBlocks;
  return b;
}

struct ZeroRegion {
  uint32_t* p;
  uint64_t n;
  uint32_t v;
};
struct ZeroList {
  ZeroRegion r[6];
  int count;
};

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = sel_min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = sel_max(v, __shfl_xor(v, o, 64));
  return v;
}

// Centroid (poly: accumulate(vertices, zeros) / nv) and the cardinal slab
// bounds as block partials (k_prep_final folds them).
template <int KIND, int NV>
__global__ __launch_bounds__(kPrepBlock) void k_prep(ElemInput in, BuildBufs b, ZeroList zl) {
  const uint64_t n = in.n;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // clear the scratch the following stages accumulate into
  for (int r = 0; r < zl.count; ++r)
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < zl.r[r].n; i += stride)
      zl.r[r].p[i] = zl.r[r].v;

  double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
  double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    double lo[3], hi[3];
    if constexpr (KIND == kPoly) {
      double v[NV][3];
      double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int j = 0; j < NV; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          v[j][c] = in.data[(uint64_t)(3 * j + c) * n + e];
          s[c] = s[c] + v[j][c];
        }
#pragma unroll
      for (int c = 0; c < 3; ++c) b.centroid[(uint64_t)c * n + e] = s[c] / (double)NV;
      // cardinal slabs of from_vertices (axes 0..2)
      double p0 = project<0>(v[0][0], v[0][1], v[0][2]);
      double p1 = project<1>(v[0][0], v[0][1], v[0][2]);
      double p2 = project<2>(v[0][0], v[0][1], v[0][2]);
      lo[0] = p0 - in.eps; hi[0] = p0 + in.eps;
      lo[1] = p1 - in.eps; hi[1] = p1 + in.eps;
      lo[2] = p2 - in.eps; hi[2] = p2 + in.eps;
#pragma unroll
      for (int j = 1; j < NV; ++j) {
        p0 = project<0>(v[j][0], v[j][1], v[j][2]);
        p1 = project<1>(v[j][0], v[j][1], v[j][2]);
        p2 = project<2>(v[j][0], v[j][1], v[j][2]);
        lo[0] = sel_min(lo[0], p0 - in.eps); hi[0] = sel_max(hi[0], p0 + in.eps);
        lo[1] = sel_min(lo[1], p1 - in.eps); hi[1] = sel_max(hi[1], p1 + in.eps);
        lo[2] = sel_min(lo[2], p2 - in.eps); hi[2] = sel_max(hi[2], p2 + in.eps);
      }
    } else if constexpr (KIND == kSphere) {
      const double x = in.data[e], y = in.data[n + e], z = in.data[2 * n + e], r = in.data[3 * n + e];
      double p0 = project<0>(x, y, z), p1 = project<1>(x, y, z), p2 = project<2>(x, y, z);
      lo[0] = p0 - r; hi[0] = p0 + r;
      lo[1] = p1 - r; hi[1] = p1 + r;
      lo[2] = p2 - r; hi[2] = p2 + r;
    } else {
      const double* k = in.data + e * kSlabs;
      lo[0] = k[0]; hi[0] = k[1];
      lo[1] = k[2]; hi[1] = k[3];
      lo[2] = k[4]; hi[2] = k[5];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      mn[c] = (lo[c] < mn[c]) ? lo[c] : mn[c];
      mx[c] = (mx[c] < hi[c]) ? hi[c] : mx[c];
    }
  }
  __shared__ double s_part[kPrepBlock / 64][6];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double a = wave_min(mn[c]);
    double z = wave_max(mx[c]);
    if (lane == 0) {
      s_part[w][c] = a;
      s_part[w][3 + c] = z;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[6];
    for (int c = 0; c < 6; ++c) r[c] = s_part[0][c];
    for (int ww = 1; ww < kPrepBlock / 64; ++ww)
      for (int c = 0; c < 3; ++c) {
        r[c] = sel_min(r[c], s_part[ww][c]);
        r[3 + c] = sel_max(r[3 + c], s_part[ww][3 + c]);
      }
    for (int c = 0; c < 6; ++c) b.partials[blockIdx.x * 6 + c] = r[c];
  }
}

// Fold of the block partials into min / inv_diag / max (one block).  A
// separate launch instead of a last-block fold: on gfx950 the device-scope
// fence that fold needs writes back the whole L2 in every block.
constexpr int kPrepFold = 8;
__global__ __launch_bounds__(kPrepBlock) void k_prep_final(BuildBufs b, uint32_t nblocks) {
  __shared__ double s_part[kPrepBlock / 64][6];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double r[6];
  for (int c = 0; c < 3; ++c) {
    r[c] = 1.7976931348623157e308;
    r[3 + c] = -1.7976931348623157e308;
  }
  // kPrepFold partials per thread and round, their loads issued together
  for (unsigned blk0 = threadIdx.x; blk0 < nblocks; blk0 += kPrepFold * blockDim.x) {
    double pv[kPrepFold][6];
#pragma unroll
    for (int u = 0; u < kPrepFold; ++u) {
      const unsigned blk = blk0 + u * blockDim.x;
      const unsigned bl = blk < nblocks ? blk : blk0;  // (blk0 is in range)
#pragma unroll
      for (int c = 0; c < 6; ++c) pv[u][c] = b.partials[bl * 6 + c];
    }
#pragma unroll
    for (int u = 0; u < kPrepFold; ++u)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        r[c] = sel_min(r[c], pv[u][c]);
        r[3 + c] = sel_max(r[3 + c], pv[u][3 + c]);
      }
  }
  for (int c = 0; c < 3; ++c) {
    r[c] = wave_min(r[c]);
    r[3 + c] = wave_max(r[3 + c]);
  }
  if (lane == 0)
    for (int c = 0; c < 6; ++c) s_part[w][c] = r[c];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int c = 0; c < 6; ++c) r[c] = s_part[0][c];
    for (int ww = 1; ww < kPrepBlock / 64; ++ww)
      for (int c = 0; c < 3; ++c) {
        r[c] = sel_min(r[c], s_part[ww][c]);
        r[3 + c] = sel_max(r[3 + c], s_part[ww][3 + c]);
      }
    Bounds bd;
    for (int c = 0; c < 3; ++c) {
      bd.min[c] = r[c];
      bd.max[c] = r[3 + c];
      double width = r[3 + c] - r[c];
      bd.inv[c] = 1.0 / width;
    }
    bd.pad = 0.0;
    *b.bounds = bd;
    if (b.origin) {
      // fp32-outward records: the object's frame, its origin (the Morton
      // bounds minimum, non-finite components 0) and the origin's projections
      // on the 13 axes (what the leaf kernel subtracts); a refit keeps it
      double o[3];
      for (int c = 0; c < 3; ++c) {
        const double v = bd.min[c];
        o[c] = (v - v == 0.0) ? v : 0.0;  // finite, else 0
        b.origin[c] = o[c];
      }
      proj_all(o[0], o[1], o[2], b.origin + 3);
    }
  }
}

// Morton codes in input order + digit histograms of every sort pass.
constexpr int kMortonBatch = 4;
template <int KIND, int BITS>
__global__ __launch_bounds__(256) void k_morton(ElemInput in, BuildBufs b) {
  constexpr int NP = (BITS == 63) ? 8 : 4;
  __shared__ uint32_t sh[NP * 256];
  for (int i = threadIdx.x; i < NP * 256; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  const Bounds bd = *b.bounds;
  const uint64_t n = in.n;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // kMortonBatch elements per lane and round, their centroid loads issued
  // together (one memory round trip per round, not per element)
  for (uint64_t e0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e0 < n; e0 += kMortonBatch * stride) {
    double xs[kMortonBatch], ys[kMortonBatch], zs[kMortonBatch];
#pragma unroll
    for (int u = 0; u < kMortonBatch; ++u) {
      const uint64_t e = e0 + u * stride;
      xs[u] = ys[u] = zs[u] = 0.0;
      if (e < n) {
        if constexpr (KIND == kPoly) {
          xs[u] = b.centroid[e]; ys[u] = b.centroid[n + e]; zs[u] = b.centroid[2 * n + e];
        } else if constexpr (KIND == kSphere) {
          xs[u] = in.data[e]; ys[u] = in.data[n + e]; zs[u] = in.data[2 * n + e];
        } else {
          xs[u] = in.centroids[3 * e]; ys[u] = in.centroids[3 * e + 1]; zs[u] = in.centroids[3 * e + 2];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kMortonBatch; ++u) {
      const uint64_t e = e0 + u * stride;
      const bool valid = e < n;
      uint64_t code = 0;
      if (valid) {
        if constexpr (BITS == 63) {
          code = morton63(xs[u], ys[u], zs[u], bd);
          ((uint64_t*)b.codes0)[e] = code;
        } else {
          const uint32_t c32 = morton30(xs[u], ys[u], zs[u], bd);
          ((uint32_t*)b.codes0)[e] = c32;
          code = c32;
        }
      }
      // a wave whose valid lanes share the digit adds once
      const uint64_t active = __ballot(valid);
      if (active == 0) continue;
      const uint32_t leader = (uint32_t)(__ffsll((long long)active) - 1);
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const uint32_t d = (uint32_t)((code >> (8 * p)) & 255);
        const uint32_t d0 = __builtin_amdgcn_readlane(d, leader);
        if (__ballot(valid & (d != d0)) == 0) {
          if (lane_id() == leader) atomicAdd(&sh[p * 256 + d0], (uint32_t)__popcll(active));
        } else if (valid) {
          atomicAdd(&sh[p * 256 + d], 1u);
        }
      }
    }
  }
  __syncthreads();
  uint32_t* hc = b.hist + (blockIdx.x % kHistCopies) * kHistWords;
  for (int i = threadIdx.x; i < NP * 256; i += blockDim.x)
    if (sh[i]) atomicAdd(&hc[i], sh[i]);
}

// depth(i) = clz(h[i] ^ h[i+1]) (255 for equal codes; the reference uses
// UINT_MAX, split/cluster.hpp:148) + histogram of the single 8-bit digit.
template <typename K>
__global__ __launch_bounds__(256) void k_cluster_depth(const K* __restrict__ codes, uint64_t n,
                                                       uint32_t* __restrict__ depth, uint32_t* __restrict__ hist) {
  __shared__ uint32_t sh[256];
  sh[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t m = n - 1;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // kMortonBatch boundaries per lane and round, their code loads issued together
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < m; i0 += kMortonBatch * stride) {
    K lo[kMortonBatch], hi[kMortonBatch];
#pragma unroll
    for (int u = 0; u < kMortonBatch; ++u) {
      const uint64_t i = i0 + u * stride;
      lo[u] = hi[u] = 0;
      if (i < m) {
        lo[u] = codes[i];
        hi[u] = codes[i + 1];
      }
    }
#pragma unroll
    for (int u = 0; u < kMortonBatch; ++u) {
      const uint64_t i = i0 + u * stride;
      if (i >= m) break;
      const K x = lo[u] ^ hi[u];
      uint32_t d;
      if constexpr (sizeof(K) == 8) d = x ? (uint32_t)__clzll((long long)x) : 255u;
      else d = x ? (uint32_t)__clz((int)x) : 255u;
      depth[i] = d;
      atomicAdd(&sh[d], 1u);
    }
  }
  __syncthreads();
  if (sh[threadIdx.x]) atomicAdd(&hist[(blockIdx.x % kHistCopies) * kHistWords + threadIdx.x], sh[threadIdx.x]);
}

// The first c entries of the (depth, index) order, sorted by position
// (split/cluster.hpp:157-168), without sorting the depths: with cum(d) the
// number of boundaries of depth <= d (the histogram of k_cluster_depth), the
// threshold D is the smallest depth with cum(D) >= c.  Every boundary of
// depth < D is taken, and the first need = c - cum(D - 1) boundaries of depth
// D in position order.  A boundary's output slot is the number of taken
// boundaries before it: (#depth < D before it) + min(#depth == D before it,
// need).  Remaining split slots take n - 1 (split/cluster.hpp:101-116).
constexpr uint32_t kClusterItems = 16;  // boundaries per thread of the count / emit passes
constexpr uint32_t kClusterTile = 256 * kClusterItems;

struct ClusterCut {
  uint32_t D, need;
};

// threshold of the histogram, computed by every block (256 bins in LDS)
__device__ __forceinline__ ClusterCut cluster_cut(const uint32_t* __restrict__ hist, uint64_t c) {
  __shared__ uint32_t h[256];
  __shared__ ClusterCut cut;
  uint32_t hv[kHistCopies];  // (all loads issued before the sum)
#pragma unroll
  for (int c = 0; c < kHistCopies; ++c) hv[c] = hist[c * kHistWords + threadIdx.x];
  uint32_t v = 0;
#pragma unroll
  for (int c = 0; c < kHistCopies; ++c) v += hv[c];
  h[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t cum = 0;
    uint32_t d = 0;
    for (; d < 255; ++d) {
      if (cum + h[d] >= c) break;
      cum += h[d];
    }
    cut = ClusterCut{d, (uint32_t)(c - cum)};
  }
  __syncthreads();
  return cut;
}

// per tile: boundaries of depth < D and of depth == D
__global__ __launch_bounds__(256) void k_cluster_count(const uint32_t* __restrict__ depth, uint64_t m,
                                                       const uint32_t* __restrict__ hist, uint64_t c,
                                                       uint32_t* __restrict__ counts) {
  const ClusterCut cut = cluster_cut(hist, c);
  const uint64_t base = (uint64_t)blockIdx.x * kClusterTile;
  uint32_t lt = 0, eq = 0;
#pragma unroll
  for (uint32_t k = 0; k < kClusterItems; ++k) {
    const uint64_t i = base + (uint64_t)k * 256 + threadIdx.x;
    if (i < m) {
      const uint32_t d = depth[i];
      lt += d < cut.D;
      eq += d == cut.D;
    }
  }
  // block sums: wave reductions, then LDS
  for (int o = 32; o > 0; o >>= 1) {
    lt += __shfl_xor(lt, o, 64);
    eq += __shfl_xor(eq, o, 64);
  }
  __shared__ uint32_t s[2][4];
  if (lane_id() == 0) {
    s[0][threadIdx.x / 64] = lt;
    s[1][threadIdx.x / 64] = eq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    counts[2 * blockIdx.x] = s[0][0] + s[0][1] + s[0][2] + s[0][3];
    counts[2 * blockIdx.x + 1] = s[1][0] + s[1][1] + s[1][2] + s[1][3];
  }
}

// exclusive scan of the tiles' counts in place (one block)
__global__ __launch_bounds__(1024) void k_cluster_scan(uint32_t* __restrict__ counts, uint32_t ntiles) {
  __shared__ uint32_t s[2][1024];
  uint32_t run[2] = {0, 0};
  for (uint32_t base = 0; base < ntiles; base += 1024) {
    const uint32_t t = base + threadIdx.x;
    const uint32_t a = t < ntiles ? counts[2 * t] : 0u, b = t < ntiles ? counts[2 * t + 1] : 0u;
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
      const uint32_t x = threadIdx.x >= o ? s[0][threadIdx.x - o] : 0u;
      const uint32_t y = threadIdx.x >= o ? s[1][threadIdx.x - o] : 0u;
      __syncthreads();
      s[0][threadIdx.x] += x;
      s[1][threadIdx.x] += y;
      __syncthreads();
    }
    if (t < ntiles) {
      counts[2 * t] = run[0] + s[0][threadIdx.x] - a;
      counts[2 * t + 1] = run[1] + s[1][threadIdx.x] - b;
    }
    run[0] += s[0][1023];
    run[1] += s[1][1023];
    __syncthreads();
  }
}

// the taken boundaries in position order; slots c .. nsplits - 1 take n - 1
__global__ __launch_bounds__(256) void k_cluster_emit(const uint32_t* __restrict__ depth, uint64_t m,
                                                      const uint32_t* __restrict__ hist, uint64_t c,
                                                      const uint32_t* __restrict__ offs, uint64_t nsplits,
                                                      uint32_t* __restrict__ splits) {
  const ClusterCut cut = cluster_cut(hist, c);
  const uint64_t base = (uint64_t)blockIdx.x * kClusterTile;
  __shared__ uint32_t wl[4], we[4];
  uint32_t lt0 = offs[2 * blockIdx.x], eq0 = offs[2 * blockIdx.x + 1];
  const uint32_t wv = threadIdx.x / 64;
  for (uint32_t k = 0; k < kClusterItems; ++k) {
    const uint64_t i = base + (uint64_t)k * 256 + threadIdx.x;
    const uint32_t d = i < m ? depth[i] : 0xffffffffu;
    const bool isl = i < m && d < cut.D, ise = i < m && d == cut.D;
    const uint64_t ml = __ballot(isl), me = __ballot(ise);
    if (lane_id() == 0) {
      wl[wv] = (uint32_t)__popcll(ml);
      we[wv] = (uint32_t)__popcll(me);
    }
    __syncthreads();
    uint32_t lb = lt0, eb = eq0;
    for (uint32_t v = 0; v < wv; ++v) {
      lb += wl[v];
      eb += we[v];
    }
    const uint32_t lr = lb + (uint32_t)__popcll(ml & lanemask_lt());
    const uint32_t er = eb + (uint32_t)__popcll(me & lanemask_lt());
    if (isl || (ise && er < cut.need)) splits[lr + (er < cut.need ? er : cut.need)] = (uint32_t)i;
    lt0 += wl[0] + wl[1] + wl[2] + wl[3];
    eq0 += we[0] + we[1] + we[2] + we[3];
    __syncthreads();
  }
  if (blockIdx.x == 0) {
    const uint32_t fillv = (uint32_t)m;  // n - 1
    for (uint64_t i = c + threadIdx.x; i < nsplits; i += 256) splits[i] = fillv;
  }
}

constexpr int kSelectMax = 16384;
__global__ __launch_bounds__(1024) void k_cluster_select(const uint32_t* __restrict__ depth_idx, uint64_t n,
                                                         uint64_t nsplits, uint32_t* __restrict__ splits) {
  __shared__ uint32_t s[kSelectMax];
  const uint64_t c = (n == 0) ? 0 : ((nsplits < n - 1) ? nsplits : n - 1);
  uint32_t m = 1;
  while (m < c) m <<= 1;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) s[i] = (i < c) ? depth_idx[i] : 0xffffffffu;
  __syncthreads();
  for (uint32_t k = 2; k <= m; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          const uint32_t a = s[i], bb = s[ixj];
          if ((a > bb) == up) {
            s[i] = bb;
            s[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  const uint32_t fillv = (n == 0) ? 0u : (uint32_t)(n - 1);
  for (uint64_t i = threadIdx.x; i < nsplits; i += blockDim.x) splits[i] = (i < c) ? s[i] : fillv;
}

__device__ __forceinline__ uint32_t patch_of(const uint32_t* __restrict__ splits, uint32_t ns, uint32_t i) {
  // number of split values <= i (patch k covers [splits[k-1], splits[k]))
  uint32_t lo = 0, hi = ns;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (splits[mid] <= i) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t msb32(uint32_t x) { return 31u - (uint32_t)__clz((int)x); }
__device__ __forceinline__ uint32_t msb64(uint64_t x) { return 63u - (uint32_t)__clzll((long long)x); }

// patch of every sorted position, the split "height" delta(i) of the
// augmented key between positions i and i+1, zeroed arrival counters.  A
// block covers kDeltaTile consecutive positions [base, base + kDeltaTile):
// the split values that fall in [base, base + kDeltaTile] (usually none or
// one) are staged in LDS, so a position's patch is the block's first patch
// plus a search over those few values instead of the whole split list.
constexpr uint32_t kDeltaItems = 8;
constexpr uint32_t kDeltaTile = 256 * kDeltaItems;
constexpr uint32_t kDeltaLds = 1024;  // staged split values per block

template <typename K>
__global__ __launch_bounds__(256) void k_delta(const K* __restrict__ codes, uint64_t n, const uint32_t* __restrict__ splits,
                                               uint32_t nsplits, uint64_t od, BuildBufs b) {
  __shared__ uint32_t s_split[kDeltaLds];
  __shared__ uint32_t s_lo, s_cnt;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < od; i += (uint64_t)gridDim.x * blockDim.x)
    b.patch_root[i] = kEnd;
  if (blockIdx.x == 0 && threadIdx.x == 0) *b.root = kEnd;  // set again by the build; stays empty if it fails
  const uint64_t base = (uint64_t)blockIdx.x * kDeltaTile;
  if (base >= n) return;
  if (threadIdx.x == 0) {
    // splits[lo, hi) are the values in [base, base + kDeltaTile]
    const uint32_t lo = base ? patch_of(splits, nsplits, (uint32_t)(base - 1)) : 0u;
    const uint64_t last = base + kDeltaTile < n ? base + kDeltaTile : n;
    const uint32_t hi = patch_of(splits, nsplits, (uint32_t)last);
    s_lo = lo;
    s_cnt = hi - lo;
  }
  __syncthreads();
  const uint32_t lo = s_lo, cnt = s_cnt;
  const bool staged = cnt <= kDeltaLds;
  if (staged)
    for (uint32_t k = threadIdx.x; k < cnt; k += 256) s_split[k] = splits[lo + k];
  __syncthreads();
  // patch of position i >= base: lo + the staged values <= i
  auto patch_at = [&](uint32_t i) -> uint32_t {
    if (!staged) return patch_of(splits, nsplits, i);
    uint32_t a = 0, z = cnt;
    while (a < z) {
      const uint32_t mid = (a + z) >> 1;
      if (s_split[mid] <= i) a = mid + 1;
      else z = mid;
    }
    return lo + a;
  };
  for (uint32_t k = 0; k < kDeltaItems; ++k) {
    const uint64_t i = base + (uint64_t)k * 256 + threadIdx.x;
    if (i >= n) break;
    const uint32_t pi = patch_at((uint32_t)i);
    b.patch_of_pos[i] = pi;
    if (i + 1 < n) {
      b.counters[i] = 0;
      const uint32_t pj = patch_at((uint32_t)(i + 1));
      uint32_t d;
      if (pi != pj) {
        d = 128u + msb32(pi ^ pj);
      } else {
        const uint64_t x = (uint64_t)(codes[i] ^ codes[i + 1]);
        d = x ? 64u + msb64(x) : msb32((uint32_t)i ^ (uint32_t)(i + 1));
      }
      b.delta[i] = (uint8_t)d;
    }
  }
}

__device__ __forceinline__ uint32_t patch_begin(const uint32_t* splits, uint64_t od, uint32_t k) {
  return k == 0 ? 0u : splits[k - 1];
}
__device__ __forceinline__ uint32_t patch_end(const uint32_t* splits, uint64_t od, uint64_t n, uint32_t k) {
  return (uint64_t)k == od - 1 ? (uint32_t)n : splits[k];
}

// ---- bottom-up build with refit -------------------------------------------------
// A node [l, r] is the left child of internal node r when delta(r) <
// delta(l-1) (or l == 0), else the right child of internal node l-1.  Each
// leaf climbs; at a parent, the first child to arrive stops and the second
// merges both children's slabs, writes the parent and continues.  Every
// synchronization is workgroup-local:
//  * phase 1 (k_tree_leaves): a workgroup owns kTreeChunk (128) consecutive leaves and
//    counts arrivals in LDS.  A child whose parent splits exactly at a chunk
//    boundary (at most the left child of internal node c1-1 and the right
//    child of internal node c0-1) is handed to the next phase; a first arriver
//    whose sibling extends past the chunk leaves its arrival word in global
//    memory.
//  * phase k >= 2 (k_tree_up): chunks kTreeChunk times larger; a workgroup
//    continues the handed-over nodes of its kTreeChunk sub-chunks.  An arrival is one
//    workgroup-scope atomic exchange on the parent's arrival word: the first
//    arriver deposits its far end (l for a left child, r for a right child)
//    and its patch, so the second learns the parent's range and patch from
//    the exchange itself and reads only the two slab halves.  No other
//    workgroup touches those words in that launch; earlier launches are
//    ordered by the kernel boundary.  The last phase is one chunk and finds
//    the root.
//
// Both forms store one record per internal node with both children's slabs
// (Rec32 / Rec64).  Each node writes its slabs into its half of the parent's
// record before it counts its arrival, and the second arriver reads the
// sibling's half; the root's slabs go to a separate array.

template <bool FP32>
using SlabT = typename Form<FP32>::T;
template <bool FP32>
using RecT = typename Form<FP32>::Rec;

template <bool FP32>
__device__ __forceinline__ RecT<FP32>* recs_of(const BuildBufs& b) {
  if constexpr (FP32) return b.rec32;
  else return b.rec64;
}

template <typename T>
__device__ __forceinline__ void slab_union(T* d, const T* s) {
#pragma unroll
  for (int i = 0; i < kAxes; ++i) {
    d[2 * i] = (d[2 * i] < s[2 * i]) ? d[2 * i] : s[2 * i];
    d[2 * i + 1] = (d[2 * i + 1] > s[2 * i + 1]) ? d[2 * i + 1] : s[2 * i + 1];
  }
}

template <bool FP32>
__device__ __forceinline__ void rec_store_half(const BuildBufs& b, uint32_t parent, int side, const SlabT<FP32>* k,
                                               uint32_t child, uint32_t elem) {
  RecT<FP32>* R = recs_of<FP32>(b) + parent;
  if constexpr (FP32) {
    // 128-byte half: slabs, child id and element as seven 16-byte stores
    int4* d = (int4*)R->h[side].box;
#pragma unroll
    for (int q = 0; q < 6; ++q) d[q] = make_int4(k[4 * q], k[4 * q + 1], k[4 * q + 2], k[4 * q + 3]);
    d[6] = make_int4(k[24], k[25], (int32_t)child, (int32_t)elem);
  } else {
    double2* d = (double2*)R->box[side];  // 208-byte halves: 16-byte aligned
#pragma unroll
    for (int q = 0; q < kAxes; ++q) d[q] = make_double2(k[2 * q], k[2 * q + 1]);
    R->child[side] = child;
    R->elem[side] = elem;
  }
}

template <bool FP32>
__device__ __forceinline__ void rec_load_half(const BuildBufs& b, uint32_t parent, int side, SlabT<FP32>* k) {
  const RecT<FP32>* R = recs_of<FP32>(b) + parent;
  if constexpr (FP32) {
    const int4* d = (const int4*)R->h[side].box;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int4 v = d[q];
      k[4 * q] = v.x;
      k[4 * q + 1] = v.y;
      k[4 * q + 2] = v.z;
      k[4 * q + 3] = v.w;
    }
    const int2 v = *(const int2*)(R->h[side].box + 24);
    k[24] = v.x;
    k[25] = v.y;
  } else {
    const double2* d = (const double2*)R->box[side];
#pragma unroll
    for (int q = 0; q < kAxes; ++q) {
      const double2 v = d[q];
      k[2 * q] = v.x;
      k[2 * q + 1] = v.y;
    }
    // keep all loads in flight: one memory round trip per tree level
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <bool FP32>
__device__ __forceinline__ void store_rootbox(const BuildBufs& b, const SlabT<FP32>* k) {
#pragma unroll
  for (int i = 0; i < kSlabs; ++i) {
    if constexpr (FP32) b.rootbox32[i] = k[i];
    else b.rootbox64[i] = k[i];
  }
}

// fp64 form: a patch root's slabs are the patch's exact slabs (one writer
// per patch); the fp32 form accumulates them from the elements instead
template <bool FP32>
__device__ __forceinline__ void store_patch_box(const BuildBufs& b, uint32_t patch, const SlabT<FP32>* k) {
  if constexpr (!FP32) {
    unsigned long long* d = b.patch_k64 + (uint64_t)patch * kSlabs;
#pragma unroll
    for (int i = 0; i < kSlabs; ++i) d[i] = ordered_u64(k[i]);
  } else {
    // the outward fp32 slabs, as the patch root's record half holds them
#pragma unroll
    for (int i = 0; i < kSlabs; ++i) b.patch_box32[(uint64_t)i * b.od + patch] = k[i];
  }
}

// a one-patch node is its patch's root when a patch boundary (delta >= 128)
// lies on both sides
__device__ __forceinline__ bool is_patch_root(uint32_t patch, uint32_t l, uint32_t r, uint64_t n, uint32_t dl,
                                              uint32_t dr) {
  return patch != kMixed && (l == 0 || dl >= 128) && (r == n - 1 || dr >= 128);
}

// leaves per phase-1 workgroup (and the growth factor of the upper phases):
// 128 leaves keep the phase-1 LDS at 30 KB, three workgroups per CU (measured:
// 612 us per 2M-triangle tree instead of 653 at 256)
constexpr uint32_t kTreeChunk = kTreeLeafChunk;

// Upper-phase climb (phases >= 2), one thread per handed-over node of the
// workgroup's chunk [c0, c1).  Node [l, r] of patch `mypatch` with slabs kd
// (its half of its parent's record is written); dl = delta[l-1] and dr =
// delta[r] (0 at the ends).  Every arrival deposits the node's slabs, far end
// and patch in the thread's LDS slot, then exchanges the parent's arrival
// word for bit 63 | slot: the first arriver stops; the second finds its
// sibling in that LDS slot, or, for a word left by phase 1 (far end + 1 in
// the high half, patch in the low half; below 2^62 since n < 2^30), in the
// sibling's half of the parent's record, written by an earlier launch.  So a
// level costs one workgroup-scope atomic and LDS latency; record halves are
// stored for the queries and the next phase, and nothing in this launch
// waits for those stores.  The climb stops at the root or at a child whose
// parent splits on the chunk boundary (handed to the next phase).  A first
// arrival returns its parent in *pend (else kEnd): if its sibling arrives
// only in a later launch, k_tree_up rewrites the word in the phase-1 format.
template <bool FP32>
__device__ __forceinline__ void climb(const BuildBufs& b, uint64_t n, uint32_t l, uint32_t r, uint32_t mypatch,
                                      uint32_t dl, uint32_t dr, SlabT<FP32>* kd, uint32_t c0, uint32_t c1,
                                      uint32_t slot, SlabT<FP32> (*s_slab)[kSlabs], uint32_t* s_far,
                                      uint32_t* s_patch, FrontNode* s_front, uint32_t* s_nfront, uint32_t* pend) {
  *pend = kEnd;
  RecT<FP32>* recs = recs_of<FP32>(b);
  for (;;) {
    const bool is_left = (l == 0) || (r != n - 1 && dr < dl);
    const uint32_t parent = is_left ? r : l - 1;
    if (parent < c0 || parent + 1 >= c1) {
      const uint32_t k = atomicAdd(s_nfront, 1u);
      if (k < 2) s_front[k] = FrontNode{l, r, mypatch, dl | (dr << 8)};
      else atomicOr(b.status, kErrTree);
      return;
    }
#pragma unroll
    for (int i = 0; i < kSlabs; ++i) s_slab[slot][i] = kd[i];
    s_far[slot] = is_left ? l : r;
    s_patch[slot] = mypatch;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    const unsigned long long old = __hip_atomic_exchange(&b.counters[parent], (1ull << 63) | slot,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    if (old == 0) {  // first arriver: the sibling finishes the parent
      *pend = parent;
      return;
    }
    uint32_t far, spatch;
    SlabT<FP32> ks[kSlabs];
    if (old >> 63) {
      const uint32_t ss = (uint32_t)old & 0x7fffffffu;
      if (ss >= 2 * kTreeChunk) {
        atomicOr(b.status, kErrTree);
        return;
      }
      far = s_far[ss];
      spatch = s_patch[ss];
#pragma unroll
      for (int i = 0; i < kSlabs; ++i) ks[i] = s_slab[ss][i];
    } else {
      far = (uint32_t)(old >> 32) - 1u;
      spatch = (uint32_t)old;
      rec_load_half<FP32>(b, parent, is_left ? 1 : 0, ks);
    }
    // a left child's sibling ends past r, a right child's starts before l
    if (is_left ? (far <= r || far >= n) : (far >= l)) {
      atomicOr(b.status, kErrTree);
      return;
    }
    if (is_left) {
      r = far;
      dr = (r != n - 1) ? b.delta[r] : 0u;
    } else {
      l = far;
      dl = l ? b.delta[l - 1] : 0u;
    }
    slab_union(kd, ks);
    mypatch = (mypatch == spatch) ? mypatch : kMixed;
    rec_set_range(recs[parent], l, r, mypatch);
    if (is_patch_root(mypatch, l, r, n, dl, dr)) {
      b.patch_root[mypatch] = parent;
      store_patch_box<FP32>(b, mypatch, kd);
    }
    if (l == 0 && r == n - 1) {
      *b.root = parent;
      store_rootbox<FP32>(b, kd);
      return;
    }
    const bool up_left = (l == 0) || (r != n - 1 && dr < dl);
    rec_store_half<FP32>(b, up_left ? r : l - 1, up_left ? 0 : 1, kd, parent, kEnd);
  }
}


__device__ __forceinline__ void empty_kdop(double* kd) {
#pragma unroll
  for (int a = 0; a < kAxes; ++a) {
    kd[2 * a] = __builtin_inf();
    kd[2 * a + 1] = -__builtin_inf();
  }
}

// Exact (fp64) patch slabs, the union of the patch's element k-DOPs, for the
// patch broadphase and local_patches (fp32 form; the fp64 form takes them
// from its patch roots).  Called by every lane of the wave; lanes without an
// element pass pk = kEnd and the empty k-DOP.  A wave whose elements all lie
// in one patch reduces in registers and lane 0 leaves the partial in its LDS
// slot (part, ppatch = the patch) for patch_flush; otherwise each lane issues
// its own atomics and the slot stays empty (ppatch = kEnd).
__device__ __forceinline__ void patch_partial(unsigned long long* pk64, uint32_t pk, const double* kd, double* part,
                                              uint32_t* ppatch) {
  const uint64_t m = __ballot(pk != kEnd);
  const uint32_t pk0 = m ? __builtin_amdgcn_readlane(pk, __ffsll((long long)m) - 1) : kEnd;
  if (m != 0 && __ballot(pk != kEnd && pk != pk0) == 0) {
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      // one slab pair at a time: the reductions are not interleaved, which
      // keeps the register count of the leaf kernel down
      const double lo = wave_min(kd[2 * a]);
      const double hi = wave_max(kd[2 * a + 1]);
      if (lane_id() == 0) {
        part[2 * a] = lo;
        part[2 * a + 1] = hi;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    if (lane_id() == 0) *ppatch = pk0;
  } else {
    if (lane_id() == 0) *ppatch = kEnd;
    if (pk != kEnd) {
      unsigned long long* d = pk64 + (uint64_t)pk * kSlabs;
#pragma unroll
      for (int a = 0; a < kAxes; ++a) {
        atomicMin(d + 2 * a, (unsigned long long)ordered_u64(kd[2 * a]));
        atomicMax(d + 2 * a + 1, (unsigned long long)ordered_u64(kd[2 * a + 1]));
      }
    }
  }
}

// After a barrier: lane i < 26 of the first wave folds slab i of the waves'
// partials, patch by patch, and issues one atomic per patch and slab (a
// quarter of the atomics of one set per wave, issued 26 lanes at a time).
template <int NW>
__device__ __forceinline__ void patch_flush(unsigned long long* pk64, const double (*part)[kSlabs],
                                            const uint32_t* ppatch) {
  const uint32_t i = threadIdx.x;
  if (i >= kSlabs) return;
  const bool hi = i & 1;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const uint32_t P = ppatch[w];
    bool first = P != kEnd;
#pragma unroll
    for (int v = 0; v < w; ++v) first = first && ppatch[v] != P;
    if (!first) continue;
    // folded in the ordered-integer domain of the atomics, so the result is
    // the one that one atomic per wave gives
    unsigned long long x = ordered_u64(part[w][i]);
#pragma unroll
    for (int v = w + 1; v < NW; ++v) {
      if (ppatch[v] != P) continue;
      const unsigned long long y = ordered_u64(part[v][i]);
      x = hi ? (y > x ? y : x) : (y < x ? y : x);
    }
    unsigned long long* d = pk64 + (uint64_t)P * kSlabs + i;
    if (hi) atomicMax(d, x);
    else atomicMin(d, x);
  }
}

// Exact fp64 patch slabs (the union of each patch's element k-DOPs) of an
// object whose tree holds fp32-outward records: one thread per sorted
// position, the element's k-DOP recomputed from its data, reduced per wave
// when the wave lies in one patch (else per lane with atomics), the block's
// waves folded in LDS.  pk64 starts at the ordered images of the empty slabs.
template <int KIND, int NV>
__global__ __launch_bounds__(256) void k_patch_slabs(ElemInput in, const uint32_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ patch_of_pos,
                                                     unsigned long long* __restrict__ pk64) {
  __shared__ double s_part[256 / kWave][kSlabs];
  __shared__ uint32_t s_ppatch[256 / kWave];
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t e = p < in.n ? perm[p] : kEnd;
  const bool act = p < in.n && e < in.n;
  double kd[kSlabs];
  uint32_t pk = kEnd;
  if (act) {
    element_kdop<KIND, NV>(in, e, kd);
    pk = patch_of_pos[p];
  } else {
    empty_kdop(kd);
  }
  patch_partial(pk64, pk, kd, s_part[threadIdx.x / kWave], &s_ppatch[threadIdx.x / kWave]);
  __syncthreads();
  patch_flush<256 / kWave>(pk64, s_part, s_ppatch);
}

__global__ void k_patch_k64_init(unsigned long long* pk64, uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) pk64[i] = (i & 1) ? 0ull : ~0ull;
}

// Phase-1 state of one chunk of kTreeChunk leaves, all in LDS.  Slot j holds the newest
// finished node whose range starts at position c0 + j (its slabs, last
// position, patch); endslot[j] is the slot of the newest finished node that
// ends at c0 + j.  A parent is written to its first position's slot, so a
// slot is only overwritten by the thread that has just read it.  far[j] and
// fpatch[j] are the far end and patch of the first child to arrive at
// internal node c0 + j.
template <bool FP32>
struct ChunkLds {
  SlabT<FP32> slab[kSlabs][kTreeChunk + 1];  // SoA, rows padded: conflict-free stores, and column reads
  uint32_t cnt[kTreeChunk];      // arrivals at internal node c0 + j
  uint32_t far[kTreeChunk];
  uint32_t fpatch[kTreeChunk];
  uint32_t last[kTreeChunk];
  uint32_t patch[kTreeChunk];
  uint32_t endslot[kTreeChunk];
  uint8_t delta[kTreeChunk + 1];  // delta[c0 - 1 + j]
  FrontNode front[2];
  uint32_t nfront;
  uint32_t order[kTreeChunk];  // fp32 climb: per wave, the lanes that formed a node at this level
};

template <bool FP32>
__device__ __forceinline__ void publish(ChunkLds<FP32>& L, uint32_t slot, uint32_t endslot, const SlabT<FP32>* kd,
                                        uint32_t last, uint32_t patch) {
#pragma unroll
  for (int i = 0; i < kSlabs; ++i) L.slab[i][slot] = kd[i];
  L.last[slot] = last;
  L.patch[slot] = patch;
  L.endslot[endslot] = slot;
}

// Phase-1 climb inside the chunk [c0, c1) from node [l, r] of patch
// `mypatch`, whose slabs are in LDS slot l - c0 and in its half of its
// parent's record.  Arrivals are LDS atomics whose fences order LDS only, so
// a level costs LDS latency.  The second arriver forms the parent's slabs
// from the two children's LDS slots, so no slabs are carried from one level
// to the next (which keeps the register count down); record halves go to
// global memory with stores that nothing in this kernel waits for.
template <bool FP32>
__device__ __forceinline__ void climb_chunk(const BuildBufs& b, uint64_t n, uint32_t l, uint32_t r, uint32_t mypatch,
                                            uint32_t c0, uint32_t c1, ChunkLds<FP32>& L, SlabT<FP32>* own) {
  RecT<FP32>* recs = recs_of<FP32>(b);
  for (;;) {
    // delta[l - 1] = L.delta[l - c0], delta[r] = L.delta[r - c0 + 1] (0 at the ends)
    const uint32_t dl = L.delta[l - c0], dr = L.delta[r - c0 + 1];
    const bool is_left = (l == 0) || (r != n - 1 && dr < dl);
    const uint32_t parent = is_left ? r : l - 1;
    if (parent < c0 || parent + 1 >= c1) {
      const uint32_t k = atomicAdd(&L.nfront, 1u);
      if (k < 2) L.front[k] = FrontNode{l, r, mypatch, dl | (dr << 8)};
      else atomicOr(b.status, kErrTree);
      return;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    const uint32_t old = __hip_atomic_fetch_add(&L.cnt[parent - c0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    if (old == 0) {
      // first arriver: the sibling finishes the parent, in this phase or,
      // through the parent's arrival word, in a later one
      L.far[parent - c0] = is_left ? l : r;
      L.fpatch[parent - c0] = mypatch;
      return;
    }
    if (old != 1) {
      atomicOr(b.status, kErrTree);
      return;
    }
    // the sibling starts at r + 1 (left child) or ends at l - 1 (right child)
    const uint32_t ss = is_left ? (r + 1 - c0) : L.endslot[l - 1 - c0];
    if (ss >= kTreeChunk) {
      atomicOr(b.status, kErrTree);
      return;
    }
    const uint32_t ms = l - c0;
    const uint32_t slast = L.last[ss], spatch = L.patch[ss];
    // fp32 slabs (32-bit images) stay in registers from one level to the
    // next; the fp64 ones are read back from the node's own slot
    SlabT<FP32> kdl[FP32 ? 1 : kSlabs];
    SlabT<FP32>* kd = FP32 ? own : kdl;
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      const SlabT<FP32> lo0 = FP32 ? own[2 * a] : L.slab[2 * a][ms], lo1 = L.slab[2 * a][ss];
      const SlabT<FP32> hi0 = FP32 ? own[2 * a + 1] : L.slab[2 * a + 1][ms], hi1 = L.slab[2 * a + 1][ss];
      kd[2 * a] = (lo0 < lo1) ? lo0 : lo1;
      kd[2 * a + 1] = (hi0 > hi1) ? hi0 : hi1;
    }
    if (is_left) r = slast;
    else l = c0 + ss;
    mypatch = (mypatch == spatch) ? mypatch : kMixed;
    rec_set_range(recs[parent], l, r, mypatch);
    const uint32_t ndl = L.delta[l - c0], ndr = L.delta[r - c0 + 1];
    if (is_patch_root(mypatch, l, r, n, ndl, ndr)) {
      b.patch_root[mypatch] = parent;
      store_patch_box<FP32>(b, mypatch, kd);
    }
    if (l == 0 && r == n - 1) {
      *b.root = parent;
      store_rootbox<FP32>(b, kd);
      return;
    }
    const bool up_left = (l == 0) || (r != n - 1 && ndr < ndl);
    rec_store_half<FP32>(b, up_left ? r : l - 1, up_left ? 0 : 1, kd, parent, kEnd);
    publish<FP32>(L, l - c0, r - c0, kd, r, mypatch);
  }
}

// The record stores of one level of the fp32 phase-1 climb, made by the
// whole wave in address order.  A lane that formed node `node` this level
// (else kEnd) stores 8 pieces of 16 bytes: its half `half` of its parent's
// record (7 pieces: slabs from its LDS slot `sl`, its node id, no element;
// half kEnd for the root) and the range piece of its own record (first,
// last, patch).  The forming lanes are ranked in lane order in the wave's
// part of L.order; lane l of store j writes piece q = k mod 8 of the node of
// rank k / 8 (k = 64 j + l), so one store covers eight nodes' pieces rather
// than one piece of up to 64 scattered nodes.  Called by all lanes of the
// wave, after the forming lanes published their slabs in their LDS slots.
__device__ __forceinline__ void store_level32(const BuildBufs& b, ChunkLds<true>& L, uint32_t c0, uint32_t node,
                                              uint32_t half, uint32_t sl) {
  const uint64_t mask = __ballot(node != kEnd);
  if (mask == 0) return;
  int4* base = (int4*)b.rec32;
  const uint32_t lane = lane_id(), w0 = threadIdx.x - lane;
  if (node != kEnd) L.order[w0 + (uint32_t)__popcll(mask & lanemask_lt())] = lane;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  const uint32_t pieces = 8u * (uint32_t)__popcll(mask);
  for (uint32_t k0 = 0; k0 < pieces; k0 += kWave) {  // wave-uniform
    const uint32_t k = k0 + lane;
    const uint32_t i = k < pieces ? k >> 3 : 0u, q = k & 7u;
    const int src = (int)L.order[w0 + i];
    const uint32_t nd = (uint32_t)__shfl((int)node, src, kWave);
    const uint32_t hf = (uint32_t)__shfl((int)half, src, kWave);
    const uint32_t s = (uint32_t)__shfl((int)sl, src, kWave);
    if (k >= pieces) continue;
    if (q == 7) {  // this node's range and patch (its record's half 0)
      base[(uint64_t)nd * 16u + 7u] = make_int4((int32_t)(c0 + s), (int32_t)L.last[s], (int32_t)L.patch[s], 0);
    } else if (hf != kEnd) {
      // piece q of the half: slabs 4q .. 4q+3; piece 6: slabs 24, 25, node, no element
      const uint32_t r0 = 4u * q;
      const int32_t x = L.slab[r0][s], y = L.slab[r0 + 1][s];
      const int32_t z = q < 6 ? L.slab[q < 6 ? r0 + 2 : 0][s] : (int32_t)nd;
      const int32_t w = q < 6 ? L.slab[q < 6 ? r0 + 3 : 0][s] : (int32_t)kEnd;
      base[(uint64_t)hf * 8u + q] = make_int4(x, y, z, w);
    }
  }
}

// The fp32 phase-1 climb (climb_chunk's steps), wave-synchronous: every lane
// of the wave runs each level, lanes whose climb has ended idle, so that the
// level's record stores go out together (store_level32).  alive: the lane
// holds leaf l = r (its slabs in own and in its LDS slot).
__device__ __forceinline__ void climb_chunk32(const BuildBufs& b, uint64_t n, bool alive, uint32_t l, uint32_t r,
                                              uint32_t mypatch, uint32_t c0, uint32_t c1, ChunkLds<true>& L,
                                              int32_t* own) {
  while (__ballot(alive) != 0) {
    uint32_t node = kEnd, half = kEnd;
    if (alive) {
      const uint32_t dl = L.delta[l - c0], dr = L.delta[r - c0 + 1];
      const bool is_left = (l == 0) || (r != n - 1 && dr < dl);
      const uint32_t parent = is_left ? r : l - 1;
      if (parent < c0 || parent + 1 >= c1) {
        const uint32_t k = atomicAdd(&L.nfront, 1u);
        if (k < 2) L.front[k] = FrontNode{l, r, mypatch, dl | (dr << 8)};
        else atomicOr(b.status, kErrTree);
        alive = false;
      } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        const uint32_t old =
            __hip_atomic_fetch_add(&L.cnt[parent - c0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        // the sibling starts at r + 1 (left child) or ends at l - 1 (right child)
        const uint32_t ss = (old == 1) ? (is_left ? (r + 1 - c0) : L.endslot[l - 1 - c0]) : 0u;
        if (old == 0) {
          // first arriver: the sibling finishes the parent, in this phase or,
          // through the parent's arrival word, in a later one
          L.far[parent - c0] = is_left ? l : r;
          L.fpatch[parent - c0] = mypatch;
          alive = false;
        } else if (old != 1 || ss >= kTreeChunk) {
          atomicOr(b.status, kErrTree);
          alive = false;
        } else {
          const uint32_t slast = L.last[ss], spatch = L.patch[ss];
#pragma unroll
          for (int a = 0; a < kAxes; ++a) {
            const int32_t lo1 = L.slab[2 * a][ss], hi1 = L.slab[2 * a + 1][ss];
            own[2 * a] = (own[2 * a] < lo1) ? own[2 * a] : lo1;
            own[2 * a + 1] = (own[2 * a + 1] > hi1) ? own[2 * a + 1] : hi1;
          }
          if (is_left) r = slast;
          else l = c0 + ss;
          mypatch = (mypatch == spatch) ? mypatch : kMixed;
          const uint32_t ndl = L.delta[l - c0], ndr = L.delta[r - c0 + 1];
          if (is_patch_root(mypatch, l, r, n, ndl, ndr)) {
            b.patch_root[mypatch] = parent;
            store_patch_box<true>(b, mypatch, own);
          }
          node = parent;
          if (l == 0 && r == n - 1) {
            *b.root = parent;
            store_rootbox<true>(b, own);
            alive = false;
          } else {
            const bool up_left = (l == 0) || (r != n - 1 && ndr < ndl);
            half = 2u * (up_left ? r : l - 1) + (up_left ? 0u : 1u);
          }
          publish<true>(L, l - c0, r - c0, own, r, mypatch);
        }
      }
    }
    store_level32(b, L, c0, node, half, l - c0);
  }
}

// The fp32-outward leaf slabs, one axis at a time: each axis of the
// element's fp64 k-DOP is formed (the operations and order of element_kdop),
// folded into the exact patch slabs, rounded outward in the object's frame
// and measured (error in fp32 ulps) before the next axis is formed, so only
// one axis is live in fp64 (the whole k-DOP held 52 more registers).
template <int KIND, int NV>
struct LeafSource {
  double v[NV][3];     // kPoly: the vertices; kSphere: v[0] the centre
  double eps;          // kPoly: from_vertices epsilon; kSphere: the radius
  const double* kd;    // kSnapshot: the element's 26 slabs
};

template <int KIND, int NV>
__device__ __forceinline__ void leaf_source(const ElemInput& in, uint32_t e, bool act, LeafSource<KIND, NV>& s) {
  const uint64_t n = in.n;
  const uint64_t ee = act ? e : 0u;
  if constexpr (KIND == kPoly) {
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) s.v[j][c] = act ? in.data[(uint64_t)(3 * j + c) * n + ee] : 0.0;
    s.eps = in.eps;
  } else if constexpr (KIND == kSphere) {
#pragma unroll
    for (int c = 0; c < 3; ++c) s.v[0][c] = act ? in.data[(uint64_t)c * n + ee] : 0.0;
    s.eps = act ? in.data[3 * n + ee] : 0.0;
  } else {
    s.kd = in.data + ee * kSlabs;
  }
}

template <int KIND, int NV, int A>
__device__ __forceinline__ void leaf_axis(const LeafSource<KIND, NV>& s, bool act, double& lo, double& hi) {
  if (!act) {
    lo = __builtin_inf();
    hi = -__builtin_inf();
    return;
  }
  if constexpr (KIND == kPoly) {
    const double p0 = project<A>(s.v[0][0], s.v[0][1], s.v[0][2]);
    lo = p0 - s.eps;
    hi = p0 + s.eps;
#pragma unroll
    for (int j = 1; j < NV; ++j) {
      const double pj = project<A>(s.v[j][0], s.v[j][1], s.v[j][2]);
      lo = sel_min(lo, pj - s.eps);
      hi = sel_max(hi, pj + s.eps);
    }
  } else if constexpr (KIND == kSphere) {
    const double pc = project<A>(s.v[0][0], s.v[0][1], s.v[0][2]);
    lo = pc - s.eps;
    hi = pc + s.eps;
  } else {
    lo = s.kd[2 * A];
    hi = s.kd[2 * A + 1];
  }
}

// distance of an fp32 slab from its fp64 value, in fp32 ulps: an ulp is a
// power of two 2^k (frexp exponent k + 1), so the division is the exponent
// shift ldexp(d, 1 - (k + 1)), exact
__device__ __forceinline__ float slab_error_ulps(float f, double exact) {
  const float mag = __builtin_fabsf(f);
  const float ulp = next_up(mag) - mag;
  const int e = __builtin_amdgcn_frexp_expf(ulp);
  return (float)__builtin_ldexp(__builtin_fabs((double)f - exact), 1 - e);
}

// Each axis's outward slabs go straight to the leaf's LDS slot (column
// `slot` of slab), so no slab stays in a register while the next axis is
// formed.
template <int KIND, int NV, int A = 0>
__device__ __forceinline__ void leaf_axes32(const LeafSource<KIND, NV>& s, bool act, const double* po,
                                            int32_t (*slab)[kTreeLeafChunk + 1], uint32_t slot) {
  double lo, hi;
  leaf_axis<KIND, NV, A>(s, act, lo, hi);
  if (act) {
    const float fl = round_down_1ulp(lo - po[A]), fh = round_up_1ulp(hi - po[A]);
    const int32_t kl = (fl == fl) ? ord32(fl) : 0x7fffffff;
    const int32_t kh = (fh == fh) ? ord32(fh) : (int32_t)0x80000000;
    slab[2 * A][slot] = kl;
    slab[2 * A + 1][slot] = kh;
  }
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (A + 1 < kAxes) leaf_axes32<KIND, NV, A + 1>(s, act, po, slab, slot);
}

// Largest distance of an object's fp32-outward leaf slabs from their fp64
// values, in fp32 ulps (atomic max of the float bits into *out): the leaf
// slabs recomputed from element data exactly as the build rounds them.  A
// diagnostic, launched when the statistic is asked for.
template <int KIND, int NV>
__global__ __launch_bounds__(256) void k_slab_error(ElemInput in, const double* __restrict__ org,
                                                    uint32_t* __restrict__ out) {
  float err = 0.0f;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < in.n; e += (uint64_t)gridDim.x * blockDim.x) {
    double kd[kSlabs], po[kAxes];
    element_kdop<KIND, NV>(in, (uint32_t)e, kd);
    proj_all(org[0], org[1], org[2], po);
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      const double xl = kd[2 * a] - po[a], xh = kd[2 * a + 1] - po[a];
      const float el = slab_error_ulps(round_down_1ulp(xl), xl), eh = slab_error_ulps(round_up_1ulp(xh), xh);
      err = (el > err) ? el : err;  // NaN distances are skipped
      err = (eh > err) ? eh : err;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float x = __shfl_xor(err, o, 64);
    err = x > err ? x : err;
  }
  if (lane_id() == 0 && err > 0.0f) atomicMax(out, __float_as_uint(err));
}

// The fp32 leaf halves of a wave (slabs from the leaves' LDS slots, leaf id,
// input index) stored as 16-byte pieces in address order: lane l of store j
// writes piece q = k mod 7 of the wave's leaf k / 7 (k = 64 j + l), so the 64
// lanes of one store cover about nine consecutive halves rather than 64
// scattered ones.  half: the lane's record half (2 x parent + side), or kEnd
// for a lane without a leaf.  Called by all 64 lanes of the wave.
__device__ __forceinline__ void store_leaf_halves32(const BuildBufs& b, const ChunkLds<true>& L, uint32_t t,
                                                    uint32_t half, uint32_t e, uint32_t c0, uint64_t n) {
  int4* base = (int4*)b.rec32;
  const uint32_t lane = t & (kWave - 1), w0 = t - lane;
  const uint32_t leaf0 = (uint32_t)(n - 1) + c0 + w0;
#pragma unroll
  for (uint32_t j = 0; j < 7; ++j) {
    const uint32_t k = kWave * j + lane;
    const uint32_t i = k / 7u, q = k - 7u * i;
    const uint32_t h = (uint32_t)__shfl((int)half, (int)i, kWave);
    const uint32_t ei = (uint32_t)__shfl((int)e, (int)i, kWave);
    const uint32_t sl = w0 + i;
    // piece q holds slabs 4q .. 4q+3; piece 6 slabs 24, 25, the leaf id and
    // its input index
    const int32_t x = L.slab[4 * q][sl], y = L.slab[4 * q + 1][sl];
    const int32_t z = q < 6 ? L.slab[4 * q + 2][sl] : (int32_t)(leaf0 + i);
    const int32_t w = q < 6 ? L.slab[4 * q + 3][sl] : (int32_t)ei;
    if (h != kEnd) base[(uint64_t)h * 8u + q] = make_int4(x, y, z, w);
  }
}

template <int KIND, int NV, bool FP32>
__global__ __launch_bounds__(kTreeChunk) void k_tree_leaves(ElemInput in, BuildBufs b, uint64_t od) {
  __shared__ ChunkLds<FP32> L;
  const uint64_t n = in.n;
  const uint32_t t = threadIdx.x;
  // chunks adjacent in the sorted order on one XCD: they gather elements from
  // the same lines of the element planes
  const uint32_t chunk = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t c0 = chunk * kTreeChunk;
  const uint32_t c1 = (uint32_t)((uint64_t)c0 + kTreeChunk < n ? c0 + kTreeChunk : n);
  // the frame origin's projections (uniform, loaded before any store)
  double po[FP32 ? kAxes : 1];
  if constexpr (FP32) {
#pragma unroll
    for (int a = 0; a < kAxes; ++a) po[a] = b.origin[3 + a];
  }
  L.cnt[t] = 0;
  {
    // delta[c0 - 1 + t] for t in [0, 256], where the index is in [0, n - 2]
    const uint64_t di = (uint64_t)c0 + t;  // = (c0 - 1 + t) + 1
    L.delta[t] = (di >= 1 && di - 1 < n - 1) ? b.delta[di - 1] : 0;
    if (t == 0) {
      const uint64_t dj = (uint64_t)c0 + kTreeChunk;
      L.delta[kTreeChunk] = (dj - 1 < n - 1) ? b.delta[dj - 1] : 0;
      L.nfront = 0;
    }
  }
  const uint32_t p = c0 + t;
  const uint32_t e = p < c1 ? b.perm[p] : kEnd;
  if (p < c1 && e >= n) atomicOr(b.status, kErrIndex);
  const bool act = p < c1 && e < n;
  uint32_t pk = act ? b.patch_of_pos[p] : kEnd;
  const uint32_t leaf = (uint32_t)(n - 1 + p);
  SlabT<FP32> own[FP32 ? kSlabs : 1];  // fp32: the climbing node's slabs
  // the leaf's half of its parent's record (the root's array if n = 1)
  const uint32_t dl = (act && p) ? b.delta[p - 1] : 0u, dr = (act && p != n - 1) ? b.delta[p] : 0u;
  const bool is_left = (p == 0) || (p != n - 1 && dr < dl);
  const uint32_t parent = is_left ? p : p - 1;
  const int side = is_left ? 0 : 1;
  if constexpr (FP32) {
    // (the exact patch slabs of the fp32-outward records are reduced by
    // k_patch_slabs when something asks for them)
    LeafSource<KIND, NV> src;
    leaf_source<KIND, NV>(in, e, act, src);
    leaf_axes32<KIND, NV>(src, act, po, L.slab, t);
  }
  if (act) {
    b.patch_of_input[e] = pk;
    const bool single = patch_begin(b.splits, od, pk) == p && patch_end(b.splits, od, n, pk) == p + 1;
    if (single) b.patch_root[pk] = leaf;
    if constexpr (FP32) {
      L.last[t] = p;
      L.patch[t] = pk;
      L.endslot[t] = t;
      // the leaf's half of its parent's record (the root's array if n = 1),
      // from its LDS slot, all its stores together; the slabs stay in
      // registers for the climb
#pragma unroll
      for (int i = 0; i < kSlabs; ++i) own[i] = L.slab[i][t];
      if (single) store_patch_box<true>(b, pk, own);
      if (n == 1) {
        *b.root = leaf;
        store_rootbox<true>(b, own);
      }
      // (with n > 1 the wave stores the leaf halves together, below)
    } else {
      double kd[kSlabs];
      element_kdop<KIND, NV>(in, e, kd);
      if (single) store_patch_box<false>(b, pk, kd);
      publish<false>(L, t, t, kd, p, pk);
      if (n == 1) {
        *b.root = leaf;
        store_rootbox<false>(b, kd);
      } else {
        rec_store_half<false>(b, parent, side, kd, leaf, e);
      }
    }
  }
  if constexpr (FP32) {
    if (n > 1) store_leaf_halves32(b, L, t, act ? 2 * parent + (uint32_t)side : kEnd, e, c0, n);
  }
  __syncthreads();  // counters, deltas and leaf slots ready
  if constexpr (FP32) climb_chunk32(b, n, act && n > 1, p, p, pk, c0, c1, L, own);
  else if (act && n > 1) climb_chunk<FP32>(b, n, p, p, pk, c0, c1, L, own);
  __syncthreads();
  // a first arrival whose sibling lies outside the chunk leaves its arrival word
  if (c0 + t + 1 < c1 && L.cnt[t] == 1)
    b.counters[c0 + t] = ((unsigned long long)(L.far[t] + 1u) << 32) | L.fpatch[t];
  if (t < 2) b.front_a[2 * chunk + t] = (t < L.nfront) ? L.front[t] : FrontNode{kEnd, kEnd, kEnd, 0u};
}

// Upper phase: a block of 2 x kTreeChunk threads per chunk of kTreeChunk sub-chunks, one
// thread per handed-over node (two per sub-chunk).  A handed-over node's
// slabs are in its half of its parent's record; thread t's LDS slot is t.
template <bool FP32>
__global__ __launch_bounds__(2 * kTreeChunk) void k_tree_up(BuildBufs b, uint64_t n, uint64_t chunk, uint64_t nsub,
                                                            const FrontNode* __restrict__ front_in,
                                                            FrontNode* __restrict__ front_out) {
  __shared__ SlabT<FP32> s_slab[2 * kTreeChunk][kSlabs];
  __shared__ uint32_t s_far[2 * kTreeChunk];
  __shared__ uint32_t s_patch[2 * kTreeChunk];
  __shared__ FrontNode s_front[2];
  __shared__ uint32_t s_nfront;
  const uint32_t t = threadIdx.x;
  const uint64_t c0l = (uint64_t)blockIdx.x * chunk;
  const uint32_t c0 = (uint32_t)c0l;
  const uint32_t c1 = (uint32_t)(c0l + chunk < n ? c0l + chunk : n);
  if (t == 0) s_nfront = 0;
  __syncthreads();
  const uint64_t slot = (uint64_t)blockIdx.x * (2 * kTreeChunk) + t;  // = 2 * sub + i
  uint32_t pend = kEnd;
  if (slot < 2 * nsub) {
    const FrontNode f = front_in[slot];
    if (f.l != kEnd) {
      if (f.l > f.r || f.r >= n || (f.l == 0 && f.r == n - 1)) {
        atomicOr(b.status, kErrTree);  // never the root
      } else {
        const uint32_t dl = f.d & 0xffu, dr = f.d >> 8;
        const bool is_left = (f.l == 0) || (f.r != n - 1 && dr < dl);
        SlabT<FP32> kd[kSlabs];
        rec_load_half<FP32>(b, is_left ? f.r : f.l - 1, is_left ? 0 : 1, kd);
        climb<FP32>(b, n, f.l, f.r, f.patch, dl, dr, kd, c0, c1, t, s_slab, s_far, s_patch, s_front, &s_nfront,
                    &pend);
      }
    }
  }
  __syncthreads();
  // a first arrival whose deposit is still in the word: its sibling arrives in
  // a later launch, which needs the far end and patch from the word itself
  // (the slabs are in the record half)
  if (pend != kEnd) {
    const unsigned long long w =
        __hip_atomic_load(&b.counters[pend], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (w == ((1ull << 63) | t)) b.counters[pend] = ((unsigned long long)(s_far[t] + 1u) << 32) | s_patch[t];
  }
  if (t < 2) front_out[2 * blockIdx.x + t] = (t < s_nfront) ? s_front[t] : FrontNode{kEnd, kEnd, kEnd, 0u};
}

// fp64 form: leaf p's slabs from its parent's record (n x 26, sorted order)
__global__ __launch_bounds__(256) void k_sorted_kdops(const Rec64* __restrict__ rec, const double* __restrict__ rootbox,
                                                      const uint8_t* __restrict__ delta, uint64_t n,
                                                      double* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * kSlabs;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t p = i / kSlabs, k = i - p * kSlabs;
    if (n == 1) {
      out[i] = rootbox[k];
    } else {
      uint32_t parent;
      int side;
      leaf_slot(delta, n, p, parent, side);
      out[i] = rec[parent].box[side][k];
    }
  }
}

// fp64 form with fp32 records: leaf p's exact slabs recomputed from its
// element (the build's fp64 k-DOP, bit for bit), n x 26 in sorted order
__global__ __launch_bounds__(256) void k_sorted_elem_kdops(ElemInput in, const uint32_t* __restrict__ perm,
                                                           double* __restrict__ out) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < in.n; p += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t e = perm[p];
    double kd[kSlabs];
    if (e < in.n) {
      element_kdop_any(in, e, kd);
    } else {
      empty_kdop(kd);
    }
    double2* d = (double2*)(out + p * kSlabs);
#pragma unroll
    for (int a = 0; a < kAxes; ++a) d[a] = make_double2(kd[2 * a], kd[2 * a + 1]);
  }
}

// The patch-level tree (the levels of the tree above the patch roots) in
// pre-order, for for_each_tree: per output node its parent and children
// (output indices, -1 for none) and its patch (a leaf) or -1.  One thread;
// `stack` holds 2 x 2 od words.  A node is a patch root when its range lies
// in one patch (its record's patch, or the leaf's position's patch).
template <bool FP32>
__global__ void k_patch_tree(BuildBufs b, uint64_t n, uint64_t od, int64_t* __restrict__ out,
                             uint32_t* __restrict__ count, uint64_t* __restrict__ stack) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  *count = 0;
  if (n == 0) return;
  const RecT<FP32>* recs = recs_of<FP32>(b);
  const uint64_t cap = 2 * od;
  uint64_t* snode = stack;
  uint64_t* spar = stack + cap;
  uint64_t sp = 0, cnt = 0;
  snode[sp] = *b.root;
  spar[sp] = ~0ull;
  ++sp;
  while (sp > 0) {
    --sp;
    const uint32_t x = (uint32_t)snode[sp];
    const uint64_t par = spar[sp];
    if (cnt >= cap || x >= 2 * n - 1) {
      atomicOr(b.status, kErrTree);
      return;
    }
    const uint64_t k = cnt++;
    out[4 * k] = par == ~0ull ? -1 : (int64_t)(par >> 1);
    out[4 * k + 1] = -1;
    out[4 * k + 2] = -1;
    if (par != ~0ull) out[4 * (par >> 1) + 1 + (par & 1)] = (int64_t)k;
    const uint32_t pv = x >= n - 1 ? b.patch_of_pos[x - (n - 1)] : rec_patch(recs[x]);
    if (pv != kMixed) {
      out[4 * k + 3] = pv;
    } else {
      out[4 * k + 3] = -1;
      if (sp + 2 > cap) {
        atomicOr(b.status, kErrTree);
        return;
      }
      snode[sp] = rec_child(recs[x], 1);  // right below left: left comes out first
      spar[sp] = (k << 1) | 1;
      ++sp;
      snode[sp] = rec_child(recs[x], 0);
      spar[sp] = k << 1;
      ++sp;
    }
  }
  *count = (uint32_t)cnt;
}

hipError_t launch_patch_tree(const BuildBufs& b, uint64_t n, uint64_t od, int fp32, int64_t* out, uint32_t* count,
                             uint64_t* stack, hipStream_t s) {
  if (!out || !count || !stack) return hipErrorInvalidValue;
  if (fp32) hipLaunchKernelGGL(k_patch_tree<true>, dim3(1), dim3(64), 0, s, b, n, od, out, count, stack);
  else hipLaunchKernelGGL(k_patch_tree<false>, dim3(1), dim3(64), 0, s, b, n, od, out, count, stack);
  return hipGetLastError();
}

// ---- launchers ---------------------------------------------------------------------
template <int KIND, int NV>
static void prep_dispatch(const ElemInput& in, const BuildBufs& b, const ZeroList& zl, hipStream_t s) {
  hipLaunchKernelGGL((k_prep<KIND, NV>), dim3((unsigned)prep_blocks(in.n)), dim3(kPrepBlock), 0, s, in, b, zl);
}

hipError_t launch_prep(const ElemInput& in, const BuildBufs& b, hipStream_t s) {
  (void)in;
  ZeroList zl{};
  zl.count = 0;
  // b.hist holds kHistCopies copies of the Morton histograms (8*256)
  // followed by the depth histogram (256)
  zl.r[zl.count++] = {b.hist, kHistCopies * kHistWords, 0u};
  // look-back words of the Morton sort (8 passes) and of the depth sort
  const uint64_t lb_words = (8 + 1) * sort_tiles(in.n > 0 ? in.n : 1) * 256;
  zl.r[zl.count++] = {b.lookback, lb_words, 0u};
  // the build's status words (error bits, fp32 slab error); nothing in this
  // launch sets them
  zl.r[zl.count++] = {b.status, 4, 0u};
  if (in.kind == kPoly) {
    switch (in.nv) {
      case 1: prep_dispatch<kPoly, 1>(in, b, zl, s); break;
      case 2: prep_dispatch<kPoly, 2>(in, b, zl, s); break;
      case 3: prep_dispatch<kPoly, 3>(in, b, zl, s); break;
      case 4: prep_dispatch<kPoly, 4>(in, b, zl, s); break;
      case 5: prep_dispatch<kPoly, 5>(in, b, zl, s); break;
      case 6: prep_dispatch<kPoly, 6>(in, b, zl, s); break;
      case 7: prep_dispatch<kPoly, 7>(in, b, zl, s); break;
      case 8: prep_dispatch<kPoly, 8>(in, b, zl, s); break;
      default: return hipErrorInvalidValue;
    }
  } else if (in.kind == kSphere) {
    prep_dispatch<kSphere, 1>(in, b, zl, s);
  } else {
    prep_dispatch<kSnapshot, 1>(in, b, zl, s);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_prep_final, dim3(1), dim3(kPrepBlock), 0, s, b, (uint32_t)prep_blocks(in.n));
  return hipGetLastError();
}

hipError_t launch_morton(const ElemInput& in, const BuildBufs& b, int bits, hipStream_t s) {
  if (in.n == 0) return hipSuccess;
  uint64_t blocks = (in.n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  const dim3 g((unsigned)blocks), t(256);
  if (bits == 63) {
    if (in.kind == kPoly) hipLaunchKernelGGL((k_morton<kPoly, 63>), g, t, 0, s, in, b);
    else if (in.kind == kSphere) hipLaunchKernelGGL((k_morton<kSphere, 63>), g, t, 0, s, in, b);
    else hipLaunchKernelGGL((k_morton<kSnapshot, 63>), g, t, 0, s, in, b);
  } else {
    if (in.kind == kPoly) hipLaunchKernelGGL((k_morton<kPoly, 30>), g, t, 0, s, in, b);
    else if (in.kind == kSphere) hipLaunchKernelGGL((k_morton<kSphere, 30>), g, t, 0, s, in, b);
    else hipLaunchKernelGGL((k_morton<kSnapshot, 30>), g, t, 0, s, in, b);
  }
  return hipGetLastError();
}

hipError_t launch_cluster(const BuildBufs& b, uint64_t n, int bits, uint64_t nsplits, hipStream_t s) {
  if (nsplits == 0) return hipSuccess;
  if (nsplits > (uint64_t)kSelectMax) return hipErrorInvalidValue;
  hipError_t e;
  if (n >= 2 && nsplits > 0) {
    const uint64_t m = n - 1;
    uint64_t blocks = (m + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    uint32_t* hist = b.hist + 8 * 256;
    if (bits == 63)
      hipLaunchKernelGGL(k_cluster_depth<uint64_t>, dim3((unsigned)blocks), dim3(256), 0, s,
                         (const uint64_t*)b.codes_a, n, b.depth0, hist);
    else
      hipLaunchKernelGGL(k_cluster_depth<uint32_t>, dim3((unsigned)blocks), dim3(256), 0, s,
                         (const uint32_t*)b.codes_a, n, b.depth0, hist);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint64_t c = nsplits < m ? nsplits : m;
    const uint64_t tiles = (m + kClusterTile - 1) / kClusterTile;
    uint32_t* counts = b.depth_idx;  // 2 words per tile (n words available)
    hipLaunchKernelGGL(k_cluster_count, dim3((unsigned)tiles), dim3(256), 0, s, (const uint32_t*)b.depth0, m,
                       (const uint32_t*)hist, c, counts);
    hipLaunchKernelGGL(k_cluster_scan, dim3(1), dim3(1024), 0, s, counts, (uint32_t)tiles);
    hipLaunchKernelGGL(k_cluster_emit, dim3((unsigned)tiles), dim3(256), 0, s, (const uint32_t*)b.depth0, m,
                       (const uint32_t*)hist, c, (const uint32_t*)counts, nsplits, b.splits);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_cluster_select, dim3(1), dim3(1024), 0, s, (const uint32_t*)b.depth_idx, n, nsplits, b.splits);
  return hipGetLastError();
}

template <int KIND, int NV>
static void tree_dispatch(const ElemInput& in, const BuildBufs& b, uint64_t od, int fp32, hipStream_t s) {
  const unsigned blocks = (unsigned)((in.n + kTreeChunk - 1) / kTreeChunk);
  if (fp32) hipLaunchKernelGGL((k_tree_leaves<KIND, NV, true>), dim3(blocks), dim3(kTreeChunk), 0, s, in, b, od);
  else hipLaunchKernelGGL((k_tree_leaves<KIND, NV, false>), dim3(blocks), dim3(kTreeChunk), 0, s, in, b, od);
}

hipError_t launch_delta(const ElemInput& in, const BuildBufs& b, int bits, uint64_t od, hipStream_t s) {
  const uint64_t n = in.n;
  if (n == 0) return hipSuccess;
  // one block per kDeltaTile positions (the od patch roots are cleared by a
  // grid-stride loop over the same blocks)
  uint64_t blocks = (n + kDeltaTile - 1) / kDeltaTile;
  if (blocks == 0) blocks = 1;
  const uint32_t nsplits = (uint32_t)(od - 1);
  if (bits == 63)
    hipLaunchKernelGGL(k_delta<uint64_t>, dim3((unsigned)blocks), dim3(256), 0, s,
                       (const uint64_t*)b.codes_a, n, (const uint32_t*)b.splits, nsplits, od, b);
  else
    hipLaunchKernelGGL(k_delta<uint32_t>, dim3((unsigned)blocks), dim3(256), 0, s,
                       (const uint32_t*)b.codes_a, n, (const uint32_t*)b.splits, nsplits, od, b);
  return hipGetLastError();
}

hipError_t launch_tree_leaves(const ElemInput& in, const BuildBufs& b, uint64_t od, int fp32, hipStream_t s) {
  const uint64_t n = in.n;
  if (n == 0) return hipSuccess;
  // every buffer the kernels write must exist for this form
  if (!b.patch_k64 || !b.patch_root || !b.counters || !b.front_a || !b.front_b || !b.delta || !b.perm)
    return hipErrorInvalidValue;
  if (fp32 ? (!b.rec32 || !b.rootbox32 || !b.origin) : (!b.rec64 || !b.rootbox64))
    return hipErrorInvalidValue;
  if (in.kind == kPoly) {
    switch (in.nv) {
      case 1: tree_dispatch<kPoly, 1>(in, b, od, fp32, s); break;
      case 2: tree_dispatch<kPoly, 2>(in, b, od, fp32, s); break;
      case 3: tree_dispatch<kPoly, 3>(in, b, od, fp32, s); break;
      case 4: tree_dispatch<kPoly, 4>(in, b, od, fp32, s); break;
      case 5: tree_dispatch<kPoly, 5>(in, b, od, fp32, s); break;
      case 6: tree_dispatch<kPoly, 6>(in, b, od, fp32, s); break;
      case 7: tree_dispatch<kPoly, 7>(in, b, od, fp32, s); break;
      case 8: tree_dispatch<kPoly, 8>(in, b, od, fp32, s); break;
      default: return hipErrorInvalidValue;
    }
  } else if (in.kind == kSphere) {
    tree_dispatch<kSphere, 1>(in, b, od, fp32, s);
  } else {
    tree_dispatch<kSnapshot, 1>(in, b, od, fp32, s);
  }
  return hipGetLastError();
}

hipError_t launch_tree_up(uint64_t n, const BuildBufs& b, int fp32, hipStream_t s) {
  hipError_t e = hipSuccess;
  // upper phases: chunks kTreeChunk times larger until one chunk holds every leaf
  uint64_t chunk = kTreeChunk;
  FrontNode* fin = b.front_a;
  FrontNode* fout = b.front_b;
  while (chunk < n) {
    const uint64_t nsub = (n + chunk - 1) / chunk;
    const uint64_t next = chunk * kTreeChunk;
    const unsigned blocks = (unsigned)((n + next - 1) / next);
    if (fp32) hipLaunchKernelGGL(k_tree_up<true>, dim3(blocks), dim3(2 * kTreeChunk), 0, s, b, n, next, nsub, fin, fout);
    else hipLaunchKernelGGL(k_tree_up<false>, dim3(blocks), dim3(2 * kTreeChunk), 0, s, b, n, next, nsub, fin, fout);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    FrontNode* tmp = fin;
    fin = fout;
    fout = tmp;
    chunk = next;
  }
  return hipSuccess;
}

hipError_t launch_tree(const ElemInput& in, const BuildBufs& b, uint64_t od, int fp32, hipStream_t s) {
  const hipError_t e = launch_tree_leaves(in, b, od, fp32, s);
  return e != hipSuccess ? e : launch_tree_up(in.n, b, fp32, s);
}

hipError_t launch_patch_k64_init(unsigned long long* pk64, uint64_t od, hipStream_t s) {
  const uint64_t count = od * kSlabs;
  if (count == 0) return hipSuccess;
  if (!pk64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_patch_k64_init, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, pk64, count);
  return hipGetLastError();
}

hipError_t launch_sorted_kdops(const Rec64* rec, const double* rootbox, const uint8_t* delta, uint64_t n, double* out,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (!rec || !rootbox || !delta || !out) return hipErrorInvalidValue;
  uint64_t blocks = (n * kSlabs + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_sorted_kdops, dim3((unsigned)blocks), dim3(256), 0, s, rec, rootbox, delta, n, out);
  return hipGetLastError();
}

hipError_t launch_patch_slabs(const ElemInput& in, const uint32_t* perm, const uint32_t* patch_of_pos,
                              unsigned long long* pk64, hipStream_t s) {
  if (in.n == 0) return hipSuccess;
  if (!perm || !patch_of_pos || !pk64 || !in.data) return hipErrorInvalidValue;
  const dim3 g((unsigned)((in.n + 255) / 256)), th(256);
  if (in.kind == kSphere) {
    hipLaunchKernelGGL((k_patch_slabs<kSphere, 1>), g, th, 0, s, in, perm, patch_of_pos, pk64);
  } else if (in.kind == kSnapshot) {
    hipLaunchKernelGGL((k_patch_slabs<kSnapshot, 1>), g, th, 0, s, in, perm, patch_of_pos, pk64);
  } else {
    switch (in.nv) {
      case 1: hipLaunchKernelGGL((k_patch_slabs<kPoly, 1>), g, th, 0, s, in, perm, patch_of_pos, pk64); break;
      case 2: hipLaunchKernelGGL((k_patch_slabs<kPoly, 2>), g, th, 0, s, in, perm, patch_of_pos, pk64); break;
      case 3: hipLaunchKernelGGL((k_patch_slabs<kPoly, 3>), g, th, 0, s, in, perm, patch_of_pos, pk64); break;
      case 4: hipLaunchKernelGGL((k_patch_slabs<kPoly, 4>), g, th, 0, s, in, perm, patch_of_pos, pk64); break;
      case 5: hipLaunchKernelGGL((k_patch_slabs<kPoly, 5>), g, th, 0, s, in, perm, patch_of_pos, pk64); break;
      case 6: hipLaunchKernelGGL((k_patch_slabs<kPoly, 6>), g, th, 0, s, in, perm, patch_of_pos, pk64); break;
      case 7: hipLaunchKernelGGL((k_patch_slabs<kPoly, 7>), g, th, 0, s, in, perm, patch_of_pos, pk64); break;
      default: hipLaunchKernelGGL((k_patch_slabs<kPoly, 8>), g, th, 0, s, in, perm, patch_of_pos, pk64); break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_slab_error(const ElemInput& in, const double* org, uint32_t* out, hipStream_t s) {
  if (in.n == 0) return hipSuccess;
  if (!org || !out || !in.data) return hipErrorInvalidValue;
  uint64_t blocks = (in.n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  const dim3 g((unsigned)blocks), th(256);
  if (in.kind == kSphere) {
    hipLaunchKernelGGL((k_slab_error<kSphere, 1>), g, th, 0, s, in, org, out);
  } else if (in.kind == kSnapshot) {
    hipLaunchKernelGGL((k_slab_error<kSnapshot, 1>), g, th, 0, s, in, org, out);
  } else {
    switch (in.nv) {
      case 1: hipLaunchKernelGGL((k_slab_error<kPoly, 1>), g, th, 0, s, in, org, out); break;
      case 2: hipLaunchKernelGGL((k_slab_error<kPoly, 2>), g, th, 0, s, in, org, out); break;
      case 3: hipLaunchKernelGGL((k_slab_error<kPoly, 3>), g, th, 0, s, in, org, out); break;
      case 4: hipLaunchKernelGGL((k_slab_error<kPoly, 4>), g, th, 0, s, in, org, out); break;
      case 5: hipLaunchKernelGGL((k_slab_error<kPoly, 5>), g, th, 0, s, in, org, out); break;
      case 6: hipLaunchKernelGGL((k_slab_error<kPoly, 6>), g, th, 0, s, in, org, out); break;
      case 7: hipLaunchKernelGGL((k_slab_error<kPoly, 7>), g, th, 0, s, in, org, out); break;
      default: hipLaunchKernelGGL((k_slab_error<kPoly, 8>), g, th, 0, s, in, org, out); break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_sorted_elem_kdops(const ElemInput& in, const uint32_t* perm, double* out, hipStream_t s) {
  if (in.n == 0) return hipSuccess;
  if (!perm || !out || !in.data) return hipErrorInvalidValue;
  uint64_t blocks = (in.n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_sorted_elem_kdops, dim3((unsigned)blocks), dim3(256), 0, s, in, perm, out);
  return hipGetLastError();
}

}  // namespace dbvh
