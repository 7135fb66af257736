// sweep.hip -- sort-and-sweep between the elements of two patches.
//
// Replaces sort_and_sweep_local and max_variant_axis (narrowphase/kokkos.hpp:
// 119-248).  The axis with the largest variance of the two patches'
// centroids; then B's slab minima on that axis, scaled to [0, UINT32_MAX]
// over the union of the two patches' slabs and truncated, are sorted (stable
// LSD radix sort); every element a of A visits the sorted B from the start,
// skips b while b's scaled maximum (fp64) lies below a's truncated minimum,
// stops at the first b whose truncated minimum exceeds a's truncated maximum,
// and reports (a, b) when the two 26-DOPs overlap.
//
// The stop is the first sorted position whose minimum exceeds a's maximum,
// i.e. an upper bound found by binary search; the visited range is then
// walked 64 positions at a time by one wave per a, so each a's pairs come out
// in B's sorted order, as the reference's callback sees them.  A count pass
// and a write pass at per-a offsets (exclusive scan) make the output order
// (a ascending, then sorted b) independent of scheduling.
#include "device.hpp"
#include "kernels.hpp"

namespace dbvh {

namespace {
constexpr int kMomentBlock = 256;
constexpr uint32_t kMomentBlocks = 256;
}  // namespace

// Per-block sums of (x, y, z, x^2, y^2, z^2) over A's centroids then B's
// (the concatenation of kokkos.hpp:127-136), in a fixed order: a thread sums
// its strided elements, the block folds its threads pairwise.
__global__ __launch_bounds__(kMomentBlock) void k_moments(const double* __restrict__ ca, uint64_t na,
                                                          const double* __restrict__ cb, uint64_t nb,
                                                          double* __restrict__ part) {
  __shared__ double s[6][kMomentBlock];
  const uint64_t n = na + nb;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (uint64_t i = (uint64_t)blockIdx.x * kMomentBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kMomentBlock) {
    const double* c = i < na ? ca + 3 * i : cb + 3 * (i - na);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double x = c[j];
      acc[j] = acc[j] + x;
      acc[3 + j] = acc[3 + j] + x * x;
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int h = kMomentBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = s[k][threadIdx.x] + s[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 6) part[(uint64_t)blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// The block sums folded in block order, then var = s2 - s * s / n per axis
// and the first axis of largest variance (kokkos.hpp:144-158).
__global__ void k_variant_axis(const double* __restrict__ part, uint32_t nblocks, uint64_t n,
                               int32_t* __restrict__ axis) {
  if (threadIdx.x != 0) return;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t b = 0; b < nblocks; ++b)
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = s[k] + part[(uint64_t)b * 6 + k];
  double var[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) var[i] = s[3 + i] - s[i] * s[i] / (double)n;
  int ax = 0;
  if (var[1] > var[0]) ax = 1;
  if (var[2] > var[ax]) ax = 2;
  *axis = ax;
}

// B's truncated minima on the axis and the identity payload (kokkos.hpp:
// 183-193); the casts are the reference's double -> uint32 conversions.
__global__ __launch_bounds__(256) void k_sweep_quantize(const double* __restrict__ kb, uint64_t nb, int axis,
                                                        double gmin, double conv, uint32_t* __restrict__ bq,
                                                        uint32_t* __restrict__ bidx) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  bq[i] = (uint32_t)float_to_uint((kb[i * 26 + 2 * axis] - gmin) * conv);
  bidx[i] = (uint32_t)i;
}

// B's k-DOPs and scaled maxima in sorted order, so that the sweep reads
// consecutive positions.
__global__ __launch_bounds__(256) void k_sweep_gather(const double* __restrict__ kb, const uint32_t* __restrict__ bidx,
                                                      uint64_t nb, int axis, double gmin, double conv,
                                                      double* __restrict__ kbs, double* __restrict__ bmax) {
  const uint64_t total = nb * 26;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t j = t / 26;
    const int k = (int)(t - j * 26);
    const double v = kb[(uint64_t)bidx[j] * 26 + k];
    kbs[t] = v;
    if (k == 2 * axis + 1) bmax[j] = (v - gmin) * conv;
  }
}

// The sweep of one element a of A by one wave: COUNT writes a's pair count
// and adds it to the 64-bit total; WRITE writes the pairs at a's exclusive
// offset.
template <bool WRITE>
__device__ __forceinline__ void sweep_one(const SweepArgs& a, uint64_t ai, int lane) {
  const double* ak = a.ka + ai * 26;
  const uint32_t amin = (uint32_t)float_to_uint((ak[2 * a.axis] - a.gmin) * a.conv);
  const uint32_t amax = (uint32_t)float_to_uint((ak[2 * a.axis + 1] - a.gmin) * a.conv);
  // the first sorted position whose minimum exceeds amax: the sweep's stop
  uint64_t lo = 0, hi = a.nb;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a.bq[mid] > amax)
      hi = mid;
    else
      lo = mid + 1;
  }
  const uint64_t ub = lo;
  uint32_t cnt = 0;
  uint64_t base = 0;
  uint64_t gid_a = 0;
  if constexpr (WRITE) {
    base = a.count[ai];
    gid_a = a.ga ? a.ga[ai] : ai;
  }
  for (uint64_t j0 = 0; j0 < ub; j0 += 64) {
    const uint64_t j = j0 + lane;
    bool hit = false;
    if (j < ub) {
      // skipped while b lies below a (bmax < amin, fp64 against the integer)
      if (!(a.bmax[j] < (double)amin)) hit = overlap(ak, a.kbs + j * 26);
    }
    const uint64_t m = __ballot(hit);
    if constexpr (WRITE) {
      if (hit) {
        const uint64_t pos = base + cnt + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        const uint32_t bj = a.bidx[j];
        a.out[2 * pos] = gid_a;
        a.out[2 * pos + 1] = a.gb ? a.gb[bj] : bj;
      }
    }
    cnt += (uint32_t)__popcll(m);
  }
  if constexpr (!WRITE) {
    if (lane == 0) {
      a.count[ai] = cnt;
      atomicAdd(a.total, (unsigned long long)cnt);
    }
  }
}

// One wave per element a of A, striding over A.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_sweep(SweepArgs a) {
  const int lane = threadIdx.x & 63;
  // a is uniform across the wave: every wave leaves the loop together
  for (uint64_t ai = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); ai < a.na; ai += (uint64_t)gridDim.x * 4)
    sweep_one<WRITE>(a, ai, lane);
}

// kdop::from_vertices for a batch of polytopes (kdop.hpp:238-264): element
// e's 26-DOP (n x 26) with the build's own projections, for callers that
// sweep elements outside an object.
template <int NV>
__global__ __launch_bounds__(256) void k_element_kdops(ElemInput in, double* __restrict__ out) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= in.n) return;
  double kd[kSlabs];
  element_kdop<kPoly, NV>(in, (uint32_t)e, kd);
#pragma unroll
  for (int i = 0; i < kSlabs; ++i) out[e * kSlabs + i] = kd[i];
}

hipError_t launch_element_kdops(const ElemInput& in, double* out, hipStream_t s) {
  if (in.n == 0) return hipSuccess;
  const dim3 g((uint32_t)((in.n + 255) / 256)), b(256);
  switch (in.nv) {
    case 1: hipLaunchKernelGGL(k_element_kdops<1>, g, b, 0, s, in, out); break;
    case 2: hipLaunchKernelGGL(k_element_kdops<2>, g, b, 0, s, in, out); break;
    case 3: hipLaunchKernelGGL(k_element_kdops<3>, g, b, 0, s, in, out); break;
    case 4: hipLaunchKernelGGL(k_element_kdops<4>, g, b, 0, s, in, out); break;
    case 5: hipLaunchKernelGGL(k_element_kdops<5>, g, b, 0, s, in, out); break;
    case 6: hipLaunchKernelGGL(k_element_kdops<6>, g, b, 0, s, in, out); break;
    case 7: hipLaunchKernelGGL(k_element_kdops<7>, g, b, 0, s, in, out); break;
    case 8: hipLaunchKernelGGL(k_element_kdops<8>, g, b, 0, s, in, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_variant_axis(const double* ca, uint64_t na, const double* cb, uint64_t nb, double* part,
                               int32_t* axis, hipStream_t s) {
  const uint64_t n = na + nb;
  uint64_t blocks = (n + kMomentBlock - 1) / kMomentBlock;
  if (blocks > kMomentBlocks) blocks = kMomentBlocks;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_moments, dim3((uint32_t)blocks), dim3(kMomentBlock), 0, s, ca, na, cb, nb, part);
  hipLaunchKernelGGL(k_variant_axis, dim3(1), dim3(64), 0, s, part, (uint32_t)blocks, n, axis);
  return hipGetLastError();
}

uint32_t variant_axis_partials() { return kMomentBlocks; }

hipError_t launch_sweep_prepare(const double* kb, uint64_t nb, int axis, double gmin, double conv, uint32_t* bq,
                                uint32_t* bidx, hipStream_t s) {
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sweep_quantize, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, s, kb, nb, axis, gmin, conv,
                     bq, bidx);
  return hipGetLastError();
}

hipError_t launch_sweep_gather(const double* kb, const uint32_t* bidx, uint64_t nb, int axis, double gmin,
                               double conv, double* kbs, double* bmax, hipStream_t s) {
  if (nb == 0) return hipSuccess;
  uint64_t blocks = (nb * 26 + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_sweep_gather, dim3((uint32_t)blocks), dim3(256), 0, s, kb, bidx, nb, axis, gmin,
                     conv, kbs, bmax);
  return hipGetLastError();
}

hipError_t launch_sweep(const SweepArgs& a, bool write, hipStream_t s) {
  if (a.na == 0) return hipSuccess;
  uint64_t blocks = (a.na + 3) / 4;
  if (blocks > 65536) blocks = 65536;
  if (write)
    hipLaunchKernelGGL(k_sweep<true>, dim3((uint32_t)blocks), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_sweep<false>, dim3((uint32_t)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace dbvh
