// split.hip -- the geom_axis / ml_geom_axis element splitters on gfx950.
//
// Replaces split_permutations and split_permutations_ml with split::mean and
// axis::longest (split/split.hpp:121-323, split/mean.hpp:47-81,
// split/axis.hpp:42-49) as set_entity_data_geom_axis / _ml_geom_axis call
// them (collision_object.hpp:133-162): a permutation of the elements (the
// reference order) and 2^depth - 1 split offsets, depth = log2(od).
//
// The reference recursion works range by range; here every range of one
// level is handled at once:
//  * ml_geom_axis: level l has 2^l ranges.  Per range: the union of its
//    element k-DOPs gives the longest axis; the mean of the centroid
//    projections on that axis is the split point; a two-pointer partition
//    (libstdc++ std::partition) moves the elements; fewer than two elements
//    give offset 0, an empty side gives size / 2.
//  * geom_axis: the reference takes the axis and the split point over all
//    elements at every level (its element span is never narrowed), so only
//    level 0 partitions; every deeper range is all on one side of that split
//    point and is split at its middle without moving anything.
//
// Exact split points without a serial sum.  The reference sums the
// projections in element order (std::accumulate; Kokkos::Serial
// parallel_reduce), and that rounded sum decides which side an element whose
// projection is close to the mean goes to.  Here the sum and the sum of
// magnitudes are accumulated in any order, which bounds both that sum and the
// reference's (recursive summation: |error| <= gamma(m) * sum |x|).  Every
// element whose projection lies outside the resulting interval is on a known
// side; only if some projection lies inside it is the sum recomputed in the
// reference's order (exact; the projections are staged in range order and one
// lane per such range adds them, so the serial part is the fp64 add chain).
//
// The two-pointer partition pairs, in order, the misplaced elements of the
// left part (those that fail the test, from the left) with the misplaced
// elements of the right part (those that pass, from the right) and swaps
// each pair.  With P the exclusive scan of the test bits, the k-th
// misplaced element from the left and from the right are found by rank, so
// a level is a scan and two scatter / gather passes.
#include "device.hpp"
#include "kernels.hpp"

namespace dbvh {

__constant__ double c_split_normals[kAxes][3] = {
    {1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.},
    {DBVH_R3, DBVH_R3, DBVH_R3}, {DBVH_R3, -DBVH_R3, DBVH_R3},
    {DBVH_R3, DBVH_R3, -DBVH_R3}, {DBVH_R3, -DBVH_R3, -DBVH_R3},
    {DBVH_R2, DBVH_R2, 0.0}, {DBVH_R2, 0.0, DBVH_R2}, {0.0, DBVH_R2, DBVH_R2},
    {DBVH_R2, -DBVH_R2, 0.0}, {DBVH_R2, 0.0, -DBVH_R2}, {0.0, DBVH_R2, -DBVH_R2}};

// project(v, axis) with a run-time axis: the operations of project<A>
// (x * 1.0 is exactly x, the zero and +-R terms are the same products)
__device__ __forceinline__ double project_rt(const double* c, int a) {
  return c[0] * c_split_normals[a][0] + c[1] * c_split_normals[a][1] + c[2] * c_split_normals[a][2];
}

__device__ __forceinline__ void split_centroid(const SplitInput& si, uint32_t e, double* c) {
  if (si.cplanes) {
    c[0] = si.cplanes[e];
    c[1] = si.cplanes[si.n + e];
    c[2] = si.cplanes[2 * si.n + e];
  } else {
    c[0] = si.caos[3 * (uint64_t)e];
    c[1] = si.caos[3 * (uint64_t)e + 1];
    c[2] = si.caos[3 * (uint64_t)e + 2];
  }
}

// the range of position i: the last range whose begin is <= i (empty ranges
// come before the non-empty one with the same begin)
__device__ __forceinline__ uint32_t find_seg(const uint32_t* begin, uint32_t nseg, uint32_t i) {
  uint32_t lo = 0, hi = nseg;  // answer in [lo, hi)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (begin[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// in-order index of range s of level l in the complete binary recursion of
// `depth` levels (the order in which the reference pushes the splits)
__device__ __forceinline__ uint32_t inorder(uint32_t l, uint32_t s, uint32_t depth) {
  return ((2u * s + 1u) << (depth - 1u - l)) - 1u;
}

__global__ void k_split_start(SplitBufs sb, uint64_t n, uint64_t od) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sb.seg[0] = 0;
    sb.seg[od] = (uint32_t)n;
  }
}

__global__ void k_split_clear(SplitBufs sb, uint32_t nseg) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  for (int i = 0; i < kSlabs; ++i) sb.seg_k64[(uint64_t)s * kSlabs + i] = (i & 1) ? 0ull : ~0ull;
  sb.seg_acc[5 * s] = 0.0;
  sb.seg_acc[5 * s + 1] = 0.0;
  sb.seg_flag[2 * s] = 0;
}

// union of the element k-DOPs per range (ordered-u64 min / max)
__global__ __launch_bounds__(256) void k_split_union(ElemInput in, SplitInput si, SplitBufs sb, uint32_t nseg,
                                                     uint64_t od) {
  const uint64_t n = si.n;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < n;
  const uint32_t s = act ? find_seg(sb.seg, nseg, (uint32_t)i) : kEnd;
  double kd[kSlabs];
  if (act) {
    element_kdop_any(in, sb.rp[i], kd);
  } else {
#pragma unroll
    for (int a = 0; a < kAxes; ++a) {
      kd[2 * a] = __builtin_inf();
      kd[2 * a + 1] = -__builtin_inf();
    }
  }
  const uint64_t m = __ballot(act);
  if (m == 0) return;
  const uint32_t s0 = __builtin_amdgcn_readlane(s, __ffsll((long long)m) - 1);
  unsigned long long* d0 = sb.seg_k64 + (uint64_t)s0 * kSlabs;
  if (__ballot(act && s != s0) == 0) {
#pragma unroll
    for (int a = 0; a < kSlabs; ++a) {
      const unsigned long long v = ordered_u64(kd[a]);
      const unsigned long long r = (a & 1) ? wave_max_u64(act ? v : 0ull) : wave_min_u64(act ? v : ~0ull);
      if (lane_id() == 0) {
        if (a & 1) atomicMax(d0 + a, r);
        else atomicMin(d0 + a, r);
      }
    }
  } else if (act) {
    unsigned long long* d = sb.seg_k64 + (uint64_t)s * kSlabs;
#pragma unroll
    for (int a = 0; a < kSlabs; ++a) {
      if (a & 1) atomicMax(d + a, (unsigned long long)ordered_u64(kd[a]));
      else atomicMin(d + a, (unsigned long long)ordered_u64(kd[a]));
    }
  }
  (void)od;
}

// axis::longest of each range's union: the first maximum slab width
// (kdop.hpp:354-362, std::max_element with length() <); 0 for an empty range
__global__ void k_split_axis(SplitBufs sb, uint32_t nseg, uint64_t od) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  int best = 0;
  if (sb.seg[od + s] > sb.seg[s]) {
    const unsigned long long* k = sb.seg_k64 + (uint64_t)s * kSlabs;
    double bl = from_ordered_u64(k[1]) - from_ordered_u64(k[0]);
    for (int a = 1; a < kAxes; ++a) {
      const double l = from_ordered_u64(k[2 * a + 1]) - from_ordered_u64(k[2 * a]);
      if (bl < l) {
        best = a;
        bl = l;
      }
    }
  }
  sb.seg_axis[s] = best;
}

// sum and sum of magnitudes of the centroid projections per range (any order)
__global__ __launch_bounds__(256) void k_split_sum(SplitInput si, SplitBufs sb, uint32_t nseg) {
  const uint64_t n = si.n;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < n;
  const uint32_t s = act ? find_seg(sb.seg, nseg, (uint32_t)i) : kEnd;
  double x = 0.0;
  if (act) {
    double c[3];
    split_centroid(si, sb.rp[i], c);
    x = project_rt(c, sb.seg_axis[s]);
  }
  const uint64_t m = __ballot(act);
  if (m == 0) return;
  const uint32_t s0 = __builtin_amdgcn_readlane(s, __ffsll((long long)m) - 1);
  if (__ballot(act && s != s0) == 0) {
    const double sum = wave_sum(x), mag = wave_sum(__builtin_fabs(x));
    if (lane_id() == 0) {
      atomicAdd(sb.seg_acc + 5 * s0, sum);
      atomicAdd(sb.seg_acc + 5 * s0 + 1, mag);
    }
  } else if (act) {
    atomicAdd(sb.seg_acc + 5 * s, x);
    atomicAdd(sb.seg_acc + 5 * s + 1, __builtin_fabs(x));
  }
}

// an interval [lo, hi] that contains the reference's split point sum / m,
// whatever order the sum was rounded in
__global__ void k_split_interval(SplitBufs sb, uint32_t nseg, uint64_t od) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const double m = (double)(sb.seg[od + s] - sb.seg[s]);
  double* acc = sb.seg_acc + 5 * s;
  if (m == 0.0) {
    acc[2] = acc[3] = 0.0;
    return;
  }
  const double u = 0x1p-53;
  const double S = acc[0], A = acc[1];
  // gamma(m) = m u / (1 - m u) bounds recursive summation in any order; once
  // for the reference's order, once for the order used here
  const double g = m * u / (1.0 - m * u);
  const double E = 2.0 * g * A;
  const double W = (__builtin_fabs(S) + E) / m;
  const double slack = 8.0 * u * W + 0x1p-1000;  // rounding of these lines and of the division
  acc[2] = S / m - E / m - slack;
  acc[3] = S / m + E / m + slack;
}

// test bit per position: 1 below the interval, 0 above, 2 undecided (and
// the range flagged for the exact split point)
__global__ __launch_bounds__(256) void k_split_pred(SplitInput si, SplitBufs sb, uint32_t nseg) {
  const uint64_t n = si.n;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == n) sb.scan[n] = 0;
  if (i >= n) return;
  const uint32_t s = find_seg(sb.seg, nseg, (uint32_t)i);
  double c[3];
  split_centroid(si, sb.rp[i], c);
  const double x = project_rt(c, sb.seg_axis[s]);
  const double lo = sb.seg_acc[5 * s + 2], hi = sb.seg_acc[5 * s + 3];
  uint32_t p;
  if (x < lo) p = 1;
  else if (x > hi) p = 0;
  else p = 2;  // also NaN x, lo or hi
  if (p == 2) atomicOr(sb.seg_flag + 2 * s, 1u);
  sb.scan[i] = p;
}

// the projections in range order, for the exact split points of flagged
// ranges (contiguous, so the serial sum reads them coalesced)
__global__ __launch_bounds__(256) void k_split_xbuf(SplitInput si, SplitBufs sb, uint32_t nseg) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= si.n) return;
  const uint32_t s = find_seg(sb.seg, nseg, (uint32_t)i);
  if (sb.seg_flag[2 * s] == 0) return;
  double c[3];
  split_centroid(si, sb.rp[i], c);
  sb.xbuf[i] = project_rt(c, sb.seg_axis[s]);
}

// the reference's split point, for flagged ranges: the projections summed in
// the range's current order from 0 (std::accumulate / Kokkos::Serial
// parallel_reduce), then divided by the size.  One wave per range: the wave
// stages 512 projections in LDS with coalesced loads and lane 0 adds them
// in order (the add chain is the serial part).
constexpr uint32_t kExactChunk = 512;

__global__ __launch_bounds__(64) void k_split_exact(SplitBufs sb, uint32_t nseg, uint64_t od) {
  __shared__ double buf[kExactChunk];
  const uint32_t s = blockIdx.x;
  if (s >= nseg || sb.seg_flag[2 * s] == 0) return;
  const uint32_t b = sb.seg[s], e = sb.seg[od + s];
  const uint32_t lane = threadIdx.x;
  double sum = 0.0;
  for (uint32_t base = b; base < e; base += kExactChunk) {
    const uint32_t cnt = (e - base) < kExactChunk ? (e - base) : kExactChunk;
#pragma unroll
    for (uint32_t k = 0; k < kExactChunk / 64; ++k) {
      const uint32_t j = k * 64 + lane;
      if (j < cnt) buf[j] = sb.xbuf[base + j];
    }
    __syncthreads();
    if (lane == 0)
      for (uint32_t j = 0; j < cnt; ++j) sum = sum + buf[j];
    __syncthreads();
  }
  if (lane == 0) sb.seg_acc[5 * s + 4] = sum / (double)(e - b);
}

__global__ __launch_bounds__(256) void k_split_fix(SplitInput si, SplitBufs sb, uint32_t nseg) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= si.n || sb.scan[i] != 2) return;
  const uint32_t s = find_seg(sb.seg, nseg, (uint32_t)i);
  double c[3];
  split_centroid(si, sb.rp[i], c);
  sb.scan[i] = project_rt(c, sb.seg_axis[s]) < sb.seg_acc[5 * s + 4] ? 1u : 0u;
}

// split offset per range (in the reference's push order), the next level's
// ranges, and the partition point L of the ranges that move (else kEnd)
__global__ void k_split_seg(SplitBufs sb, uint32_t nseg, uint64_t od, uint32_t level, uint32_t depth, int ml,
                            uint32_t* splits) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const uint32_t b = sb.seg[s], e = sb.seg[od + s], m = e - b;
  const uint32_t L = sb.scan[e] - sb.scan[b];
  const bool part = !(ml && m < 2);
  uint32_t delta = 0;
  if (part) delta = (L == 0 || L == m) ? m / 2 : L;
  const uint32_t sp = b + delta;
  splits[inorder(level, s, depth)] = sp;
  sb.seg_next[2 * s] = b;
  sb.seg_next[od + 2 * s] = sp;
  sb.seg_next[2 * s + 1] = sp;
  sb.seg_next[od + 2 * s + 1] = e;
  sb.seg_flag[2 * s + 1] = (part && L > 0 && L < m) ? L : kEnd;
}

// the two-pointer swaps by rank: the k-th failing element left of the
// partition point and the k-th passing element right of it (counted from the
// right end) trade places
__global__ __launch_bounds__(256) void k_split_scatter(SplitInput si, SplitBufs sb, uint32_t nseg, uint64_t od) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= si.n) return;
  const uint32_t s = find_seg(sb.seg, nseg, (uint32_t)i);
  const uint32_t L = sb.seg_flag[2 * s + 1];
  if (L == kEnd) return;
  const uint32_t b = sb.seg[s], e = sb.seg[od + s];
  const uint32_t* P = sb.scan;
  const uint32_t pv = P[i + 1] - P[i];
  if (i < b + L && pv == 0) {
    const uint32_t k = (uint32_t)(i - b) - (P[i] - P[b]);
    sb.tmpf[b + k] = sb.rp[i];
  } else if (i >= b + L && pv == 1) {
    const uint32_t k = P[e] - P[i + 1];
    sb.tmpt[b + k] = sb.rp[i];
  }
}

__global__ __launch_bounds__(256) void k_split_gather(SplitInput si, SplitBufs sb, uint32_t nseg, uint64_t od) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= si.n) return;
  const uint32_t s = find_seg(sb.seg, nseg, (uint32_t)i);
  const uint32_t L = sb.seg_flag[2 * s + 1];
  uint32_t v = sb.rp[i];
  if (L != kEnd) {
    const uint32_t b = sb.seg[s], e = sb.seg[od + s];
    const uint32_t* P = sb.scan;
    const uint32_t pv = P[i + 1] - P[i];
    if (i < b + L && pv == 0) v = sb.tmpt[b + (uint32_t)(i - b) - (P[i] - P[b])];
    else if (i >= b + L && pv == 1) v = sb.tmpf[b + P[e] - P[i + 1]];
  }
  sb.rp_b[i] = v;
}

// geom_axis below level 0: every range is split at its middle (one thread;
// at most od - 1 splits)
__global__ void k_split_midpoints(SplitBufs sb, uint64_t od, uint32_t depth, uint32_t* splits) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t* cur = sb.seg;
  uint32_t* nxt = sb.seg_next;
  for (uint32_t l = 1; l < depth; ++l) {
    const uint32_t nseg = 1u << l;
    for (uint32_t s = 0; s < nseg; ++s) {
      const uint32_t b = cur[s], e = cur[od + s];
      const uint32_t sp = b + (e - b) / 2;
      splits[inorder(l, s, depth)] = sp;
      nxt[2 * s] = b;
      nxt[od + 2 * s] = sp;
      nxt[2 * s + 1] = sp;
      nxt[od + 2 * s + 1] = e;
    }
    uint32_t* t = cur;
    cur = nxt;
    nxt = t;
  }
}

// patch of every input element: the number of split offsets <= its position
__global__ __launch_bounds__(256) void k_split_patch_key(SplitBufs sb, uint64_t n, const uint32_t* splits,
                                                         uint32_t nsplits, uint32_t* patch_key) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t lo = 0, hi = nsplits;  // count of splits <= i
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (splits[mid] <= (uint32_t)i) lo = mid + 1;
    else hi = mid;
  }
  patch_key[sb.rp[i]] = lo;
}

// keys[i] = src[idx[i]]
template <typename T>
__global__ __launch_bounds__(256) void k_gather_by(const T* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                                                   T* __restrict__ dst) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

template <typename T>
hipError_t launch_gather_by(const T* src, const uint32_t* idx, uint64_t n, T* dst, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_by<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, idx, n, dst);
  return hipGetLastError();
}
template hipError_t launch_gather_by<uint32_t>(const uint32_t*, const uint32_t*, uint64_t, uint32_t*, hipStream_t);
template hipError_t launch_gather_by<uint64_t>(const uint64_t*, const uint32_t*, uint64_t, uint64_t*, hipStream_t);

// Patches given by the caller: patch k holds input indices [given[k-1],
// given[k]) (given[-1] = 0, given[od-1] = n); the reference order is the
// input order.  given: od - 1 non-decreasing offsets (device).
hipError_t launch_given_split(SplitBufs sb, uint64_t n, uint64_t od, const uint32_t* given, uint32_t* splits,
                              uint32_t* patch_key, uint32_t** result, hipStream_t s) {
  *result = sb.rp;
  if (n == 0) return hipSuccess;
  hipError_t e = launch_iota_u32(sb.rp, n, s);
  if (e != hipSuccess) return e;
  if (od > 1) {
    e = launch_copy<uint32_t>(splits, given, od - 1, s);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_split_patch_key, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sb, n, splits,
                     (uint32_t)(od - 1), patch_key);
  return hipGetLastError();
}

hipError_t launch_split(const ElemInput& in, const SplitInput& si, SplitBufs sb, uint64_t od, int ml,
                        uint32_t* splits, uint32_t* patch_key, uint32_t** result, hipStream_t s) {
  const uint64_t n = si.n;
  *result = sb.rp;
  if (n == 0 || od < 2) {
    if (n) {
      hipError_t e = launch_iota_u32(sb.rp, n, s);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(k_split_patch_key, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sb, n, splits, 0u,
                         patch_key);
    }
    return hipGetLastError();
  }
  if ((od & (od - 1)) != 0 || n >= (1ull << 30)) return hipErrorInvalidValue;
  if (!si.cplanes && !si.caos) return hipErrorInvalidValue;
  uint32_t depth = 0;
  while ((1ull << (depth + 1)) <= od) ++depth;
  hipError_t e = launch_iota_u32(sb.rp, n, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_split_start, dim3(1), dim3(64), 0, s, sb, n, od);
  const unsigned g = (unsigned)((n + 255) / 256), g1 = (unsigned)((n + 1 + 255) / 256);
  const uint32_t levels = ml ? depth : 1u;
  for (uint32_t l = 0; l < levels; ++l) {
    const uint32_t nseg = 1u << l;
    const unsigned gs = (nseg + 63) / 64;
    hipLaunchKernelGGL(k_split_clear, dim3(gs), dim3(64), 0, s, sb, nseg);
    hipLaunchKernelGGL(k_split_union, dim3(g), dim3(256), 0, s, in, si, sb, nseg, od);
    hipLaunchKernelGGL(k_split_axis, dim3(gs), dim3(64), 0, s, sb, nseg, od);
    hipLaunchKernelGGL(k_split_sum, dim3(g), dim3(256), 0, s, si, sb, nseg);
    hipLaunchKernelGGL(k_split_interval, dim3(gs), dim3(64), 0, s, sb, nseg, od);
    hipLaunchKernelGGL(k_split_pred, dim3(g1), dim3(256), 0, s, si, sb, nseg);
    hipLaunchKernelGGL(k_split_xbuf, dim3(g), dim3(256), 0, s, si, sb, nseg);
    hipLaunchKernelGGL(k_split_exact, dim3(nseg), dim3(64), 0, s, sb, nseg, od);
    hipLaunchKernelGGL(k_split_fix, dim3(g), dim3(256), 0, s, si, sb, nseg);
    e = launch_fill_u32((uint32_t*)sb.scan_lb, 2 * sort_tiles(n + 1), 0u, s);
    if (e != hipSuccess) return e;
    e = launch_fill_u32(sb.scan_ctr, 1, 0u, s);
    if (e != hipSuccess) return e;
    e = launch_scan_u32(sb.scan, n + 1, sb.scan_lb, sb.scan_ctr, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_split_seg, dim3(gs), dim3(64), 0, s, sb, nseg, od, l, depth, ml, splits);
    hipLaunchKernelGGL(k_split_scatter, dim3(g), dim3(256), 0, s, si, sb, nseg, od);
    hipLaunchKernelGGL(k_split_gather, dim3(g), dim3(256), 0, s, si, sb, nseg, od);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t* t = sb.rp;
    sb.rp = sb.rp_b;
    sb.rp_b = t;
    t = sb.seg;
    sb.seg = sb.seg_next;
    sb.seg_next = t;
  }
  if (!ml) hipLaunchKernelGGL(k_split_midpoints, dim3(1), dim3(64), 0, s, sb, od, depth, splits);
  hipLaunchKernelGGL(k_split_patch_key, dim3(g), dim3(256), 0, s, sb, n, splits, (uint32_t)(od - 1), patch_key);
  *result = sb.rp;
  return hipGetLastError();
}

}  // namespace dbvh
