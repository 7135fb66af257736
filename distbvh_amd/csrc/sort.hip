// sort.hip -- stable LSD radix sort for gfx950 (8-bit digits, one kernel per
// digit pass, decoupled look-back across tiles).
//
// Replaces radix_sorter<T, I> (util/sort.hpp:75-144), which runs one 1-bit
// split pass (3 dispatches) per key bit.  The result is the same stable
// permutation: keys ascending, ties in input order.
//
// Per pass, a 256-thread workgroup takes a 4096-key tile (tile ids come from
// an atomic counter, so a tile only ever waits on tiles already running):
//  * keys are read coalesced (wave w, iteration j, lane l -> tile position
//    w*1024 + j*64 + l) and ranked per wave with 8 ballots per iteration
//    (the lanes holding the same digit) plus per-wave LDS digit counters;
//  * the tile's 256 digit counts are published as an aggregate, then each
//    thread (= digit) walks back over earlier tiles until it meets an
//    inclusive prefix (flag and count packed in one 32-bit word, agent-scope
//    atomics);
//  * keys and payloads are reordered by digit in LDS and written out in digit
//    runs, so consecutive lanes write consecutive addresses.
// Digit histograms for all passes come from one read of the keys (fused into
// the Morton kernel for the object build).  A pass whose digit is the same
// for every key is skipped on the device.
#include "device.hpp"
#include "kernels.hpp"

namespace dbvh {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;  // 4096
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;
constexpr int kLookWindow = 4;  // predecessors read per look-back round trip

__device__ __forceinline__ uint32_t lb_load(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t sort_tiles_dev(uint64_t n) { return (n + kSortTile - 1) / kSortTile; }

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* s_tmp /*4*/) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_tmp[w] = x;
  __syncthreads();
  uint32_t wp = 0;
  for (int i = 0; i < w; ++i) wp += s_tmp[i];
  __syncthreads();
  return wp + x - v;
}

// Histograms of all digit passes of `keys` (grid-stride).
template <typename K>
__global__ __launch_bounds__(256) void k_sort_histogram(const K* __restrict__ keys, uint64_t n, int npass,
                                                        uint32_t* __restrict__ hist) {
  __shared__ uint32_t sh[8 * 256];
  for (int i = threadIdx.x; i < npass * 256; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    K k = keys[i];
    for (int p = 0; p < npass; ++p) atomicAdd(&sh[p * 256 + (uint32_t)((k >> (8 * p)) & 255)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < npass * 256; i += blockDim.x)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// One block: digit offsets (exclusive scan per pass), pass activity and the
// ping-pong buffer selectors (the last active pass writes buffer A).
__global__ __launch_bounds__(256) void k_sort_plan(const uint32_t* __restrict__ hist, uint64_t n, int npass,
                                                   SortCtl* __restrict__ ctl, int copies, uint32_t copy_stride) {
  __shared__ uint32_t s_wtot[8][4];
  __shared__ uint32_t s_active[8];
  const int t = threadIdx.x;
  // every pass's count of digit t, summed over the copies (all loads issued
  // before any is used)
  // (every load is made, at a clamped address, and the unwanted ones count
  // zero: no branch separates them)
  uint32_t hv[8][8];
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const int pp = p < npass ? p : 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int cc = c < copies ? c : 0;
      hv[p][c] = hist[(uint64_t)cc * copy_stride + pp * 256 + t];
    }
  }
  uint32_t hs[8];
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    hs[p] = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) hs[p] += (p < npass && c < copies) ? hv[p][c] : 0u;
  }
  // the exclusive scans of all passes side by side (wave scans, then the
  // four wave totals of each pass through LDS: two barriers for all passes);
  // a pass is needed unless one bin holds every key
  const int lane = t & 63, w = t >> 6;
  uint32_t xs[8];
#pragma unroll
  for (int p = 0; p < 8; ++p) xs[p] = hs[p];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const uint32_t y = __shfl_up(xs[p], o, 64);
      if (lane >= o) xs[p] += y;
    }
  }
  if (t < 8) s_active[t] = 1u;
  __syncthreads();
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    if (lane == 63) s_wtot[p][w] = xs[p];
    if (p < npass && (uint64_t)hs[p] == n) s_active[p] = 0u;
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    if (p >= npass) break;  // (uniform)
    uint32_t wp = 0;
    for (int i = 0; i < w; ++i) wp += s_wtot[p][i];
    ctl->offsets[p * 256 + t] = wp + xs[p] - hs[p];
  }
  if (t == 0) {
    int nact = 0;
    for (int p = 0; p < npass; ++p) nact += s_active[p];
    if (nact == 0) {
      s_active[0] = 1;
      nact = 1;
    }
    // walk backwards: last active pass -> A (1), previous -> B (2), ...
    uint32_t out_sel[8];
    uint32_t cur = 1;
    for (int p = npass - 1; p >= 0; --p) {
      if (s_active[p]) {
        out_sel[p] = cur;
        cur = (cur == 1) ? 2 : 1;
      } else {
        out_sel[p] = 0;
      }
    }
    uint32_t prev_out = 0;  // input of the first active pass is the caller's buffer (0)
    for (int p = 0; p < npass; ++p) {
      SortPassInfo pi;
      pi.active = s_active[p];
      pi.in_sel = prev_out;
      pi.out_sel = out_sel[p];
      pi.first = (prev_out == 0) ? 1u : 0u;
      ctl->pass[p] = pi;
      if (s_active[p]) prev_out = out_sel[p];
      ctl->tile_counter[p] = 0;
    }
    for (int p = npass; p < 8; ++p) {
      ctl->pass[p].active = 0;
      ctl->tile_counter[p] = 0;
    }
  }
}

template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_sort_pass(SortArgs<K> a, int pass) {
  const SortPassInfo pi = a.ctl->pass[pass];
  if (!pi.active) return;
  const K* __restrict__ kin = pi.in_sel == 0 ? a.k0 : (pi.in_sel == 1 ? a.ka : a.kb);
  K* __restrict__ kout = pi.out_sel == 1 ? a.ka : a.kb;
  const uint32_t* __restrict__ vin = pi.in_sel == 0 ? a.v0 : (pi.in_sel == 1 ? a.va : a.vb);
  uint32_t* __restrict__ vout = pi.out_sel == 1 ? a.va : a.vb;
  const bool has_vals = a.va != nullptr;
  const int shift = 8 * pass;
  const uint64_t n = a.n;

  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wcount[4][256];
  __shared__ uint32_t s_texcl[256];
  __shared__ uint32_t s_gbase[256];
  __shared__ uint32_t s_tmp[4];
  __shared__ uint32_t s_hist[256];
  __shared__ K s_keys[kSortTile];
  __shared__ uint32_t s_vals[kSortTile];

  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  if (tid == 0) s_tile = atomicAdd(&a.ctl->tile_counter[pass], 1u);
  for (int i = tid; i < 4 * 256; i += kSortBlock) (&s_wcount[0][0])[i] = 0;
  s_hist[tid] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  if (tile >= a.ntiles) return;  // uniform per block; no tile waits on it
  const uint64_t base = (uint64_t)tile * kSortTile;

  K key[kSortItems];
  uint32_t val[kSortItems];
  uint32_t rank[kSortItems];
  uint32_t dig[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    uint64_t pos = base + (uint64_t)(w * (kSortItems * 64) + j * 64 + lane);
    bool valid = pos < n;
    key[j] = valid ? kin[pos] : (K)0;
    if (has_vals) val[j] = valid ? (vin ? vin[pos] : (uint32_t)pos) : 0u;
    dig[j] = valid ? (uint32_t)((key[j] >> shift) & 255) : 256u;
  }
  // the tile's digit counts first, published as its aggregate before the
  // ranking, so that later tiles' look-backs find it sooner
#pragma unroll
  for (int j = 0; j < kSortItems; ++j)
    if (dig[j] < 256u) atomicAdd(&s_hist[dig[j]], 1u);
  __syncthreads();
  uint32_t* lb = a.lookback + ((uint64_t)pass * a.ntiles + tile) * 256;
  {
    const uint32_t tot = s_hist[tid];
    if (tile == 0) lb_store(&lb[tid], kFlagInc | tot);
    else lb_store(&lb[tid], kFlagAgg | tot);
  }
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t d = dig[j];
    const bool valid = d < 256u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t before = __popcll(m & lt);
    uint32_t old = 0;
    if (valid) old = s_wcount[w][d];
    rank[j] = old + before;
    __builtin_amdgcn_wave_barrier();
    if (valid && before == 0) s_wcount[w][d] = old + (uint32_t)__popcll(m);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();

  // thread t owns digit t
  const int t = tid;
  uint32_t c0 = s_wcount[0][t], c1 = s_wcount[1][t], c2 = s_wcount[2][t], c3 = s_wcount[3][t];
  const uint32_t total = c0 + c1 + c2 + c3;
  __syncthreads();
  s_wcount[0][t] = 0;
  s_wcount[1][t] = c0;
  s_wcount[2][t] = c0 + c1;
  s_wcount[3][t] = c0 + c1 + c2;

  const uint32_t texcl = block_excl_scan256(total, s_tmp);
  s_texcl[t] = texcl;

  uint32_t excl = 0;
  if (tile > 0) {
    // decoupled look-back over a window of kLookWindow predecessors whose
    // words are loaded together (one round trip per window, not per tile):
    // their aggregates are added in order up to the first inclusive prefix;
    // a word not yet published is polled until it is (tile 0 publishes an
    // inclusive prefix, so the walk ends by it)
    const uint32_t* lb0 = a.lookback + (uint64_t)pass * a.ntiles * 256 + t;
    int64_t prev = (int64_t)tile - 1;
    bool done = false;
    while (!done) {
      uint32_t v[kLookWindow];
#pragma unroll
      for (int k = 0; k < kLookWindow; ++k) v[k] = prev - k >= 0 ? lb_load(lb0 + (uint64_t)(prev - k) * 256) : 0u;
#pragma unroll
      for (int k = 0; k < kLookWindow; ++k) {
        if (done || prev - k < 0) break;
        uint32_t x = v[k];
        while ((x & ~kValMask) == 0) {
          __builtin_amdgcn_s_sleep(1);
          x = lb_load(lb0 + (uint64_t)(prev - k) * 256);
        }
        excl += x & kValMask;
        done = (x & ~kValMask) == kFlagInc;
      }
      prev -= kLookWindow;
    }
    lb_store(&lb[t], kFlagInc | (excl + total));
  }
  s_gbase[t] = a.ctl->offsets[pass * 256 + t] + excl;
  __syncthreads();

#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t d = dig[j];
    if (d < 256u) {
      const uint32_t lp = s_texcl[d] + s_wcount[w][d] + rank[j];
      s_keys[lp] = key[j];
      if (has_vals) s_vals[lp] = val[j];
    }
  }
  __syncthreads();
  const uint32_t cnt = (uint32_t)((n - base) < (uint64_t)kSortTile ? (n - base) : (uint64_t)kSortTile);
  for (uint32_t i = tid; i < cnt; i += kSortBlock) {
    const K k = s_keys[i];
    const uint32_t d = (uint32_t)((k >> shift) & 255);
    const uint32_t op = s_gbase[d] + (i - s_texcl[d]);
    if (op < n) {
      kout[op] = k;
      if (has_vals) vout[op] = s_vals[i];
    }
  }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_iota_u32(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

template <typename T>
__global__ void k_copy(T* __restrict__ dst, const T* __restrict__ src, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---- host launchers ----------------------------------------------------------------
uint64_t sort_tiles(uint64_t n) { return (n + kSortTile - 1) / kSortTile; }

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_iota_u32(uint32_t* p, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_iota_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_copy(T* dst, const T* src, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_copy<T>, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
  return hipGetLastError();
}
template hipError_t launch_copy<uint64_t>(uint64_t*, const uint64_t*, uint64_t, hipStream_t);
template hipError_t launch_copy<uint32_t>(uint32_t*, const uint32_t*, uint64_t, hipStream_t);
template hipError_t launch_copy<double>(double*, const double*, uint64_t, hipStream_t);

template <typename K>
hipError_t launch_sort_histogram(const K* keys, uint64_t n, int npass, uint32_t* hist, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_sort_histogram<K>, dim3((unsigned)blocks), dim3(256), 0, s, keys, n, npass, hist);
  return hipGetLastError();
}
template hipError_t launch_sort_histogram<uint64_t>(const uint64_t*, uint64_t, int, uint32_t*, hipStream_t);
template hipError_t launch_sort_histogram<uint32_t>(const uint32_t*, uint64_t, int, uint32_t*, hipStream_t);

// Sort after the histograms are in `hist` (npass * 256 counters) and the
// look-back words are zero.  Result in a.ka (and a.va).
template <typename K>
hipError_t launch_sort_plan(const SortArgs<K>& a, int npass, const uint32_t* hist, hipStream_t s, int copies,
                            uint32_t copy_stride) {
  if (a.n == 0) return hipSuccess;
  if (npass < 1 || npass > 8 || copies < 1 || copies > 8) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_sort_plan, dim3(1), dim3(256), 0, s, hist, a.n, npass, a.ctl, copies, copy_stride);
  return hipGetLastError();
}

template <typename K>
hipError_t launch_sort_pass(const SortArgs<K>& a, int p, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_pass<K>, dim3((unsigned)sort_tiles(a.n)), dim3(kSortBlock), 0, s, a, p);
  return hipGetLastError();
}

template <typename K>
hipError_t launch_sort_passes(const SortArgs<K>& a, int npass, const uint32_t* hist, hipStream_t s) {
  hipError_t e = launch_sort_plan<K>(a, npass, hist, s);
  for (int p = 0; p < npass && e == hipSuccess; ++p) e = launch_sort_pass<K>(a, p, s);
  return e;
}
template hipError_t launch_sort_plan<uint64_t>(const SortArgs<uint64_t>&, int, const uint32_t*, hipStream_t, int,
                                              uint32_t);
template hipError_t launch_sort_plan<uint32_t>(const SortArgs<uint32_t>&, int, const uint32_t*, hipStream_t, int,
                                              uint32_t);
template hipError_t launch_sort_pass<uint64_t>(const SortArgs<uint64_t>&, int, hipStream_t);
template hipError_t launch_sort_pass<uint32_t>(const SortArgs<uint32_t>&, int, hipStream_t);
template hipError_t launch_sort_passes<uint64_t>(const SortArgs<uint64_t>&, int, const uint32_t*, hipStream_t);
template hipError_t launch_sort_passes<uint32_t>(const SortArgs<uint32_t>&, int, const uint32_t*, hipStream_t);


// ---- exclusive scan (prefix_sum, util/prefix_sum.hpp:41-54) ---------------------------
// Single pass: 256 threads x 16 values per tile, decoupled look-back with the
// flag in the top two bits of a 64-bit word.  Values wrap modulo 2^32 as the
// reference's uint32 scan does.
constexpr uint64_t kScanAgg = 1ull << 62;
constexpr uint64_t kScanInc = 2ull << 62;
constexpr uint64_t kScanMask = (1ull << 62) - 1;

__global__ __launch_bounds__(256) void k_scan_u32(uint32_t* __restrict__ v, uint64_t n, uint64_t* __restrict__ lb,
                                                  uint32_t* __restrict__ counter) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_tmp[4];
  __shared__ uint64_t s_excl;
  const int tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(counter, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  if ((uint64_t)tile >= sort_tiles_dev(n)) return;
  const uint64_t base = (uint64_t)tile * kSortTile + (uint64_t)tid * kSortItems;
  uint32_t x[kSortItems];
  uint32_t sum = 0;
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint64_t i = base + j;
    x[j] = i < n ? v[i] : 0u;
    sum += x[j];
  }
  const uint32_t texcl = block_excl_scan256(sum, s_tmp);
  if (tid == 255) {
    const uint64_t total = (uint64_t)(texcl + sum);
    uint64_t excl = 0;
    if (tile == 0) {
      __hip_atomic_store(&lb[0], kScanInc | (total & kScanMask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(&lb[tile], kScanAgg | (total & kScanMask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int64_t prev = (int64_t)tile - 1;
      for (;;) {
        uint64_t w = __hip_atomic_load(&lb[prev], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while ((w & ~kScanMask) == 0) {
          __builtin_amdgcn_s_sleep(1);
          w = __hip_atomic_load(&lb[prev], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        excl += w & kScanMask;
        if ((w & ~kScanMask) == kScanInc) break;
        --prev;
      }
      __hip_atomic_store(&lb[tile], kScanInc | ((excl + total) & kScanMask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_excl = excl;
  }
  __syncthreads();
  uint32_t run = (uint32_t)s_excl + texcl;
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint64_t i = base + j;
    if (i < n) v[i] = run;
    run += x[j];
  }
}

hipError_t launch_scan_u32(uint32_t* v, uint64_t n, uint64_t* lb, uint32_t* counter, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scan_u32, dim3((unsigned)sort_tiles(n)), dim3(256), 0, s, v, n, lb, counter);
  return hipGetLastError();
}

}  // namespace dbvh
