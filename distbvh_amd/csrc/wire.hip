// wire.hip -- byte layouts of what the reference sends between ranks, packed
// and unpacked on the device, so that a multi-GPU driver can hand them to
// RCCL without a host round trip.
//
// Each layout follows the field order of the reference's serialize() and
// writes every field as its in-memory bytes (DARMA checkpoint copies
// byte-copyable types whole, serialization/bvh_serialize.hpp:66-79):
//  * patch record, 256 B (patch<>::serialize, patch.hpp:93-97):
//      u64 global_id | u64 size | dop_26 extents: 13 x {f64 min, f64 max}
//      (kdop.hpp:57-60) | centroid vec3<double>, whose storage is one chunk
//      of 4 doubles (math/storage/vec_storage.hpp:64-71): x, y, z, 0;
//  * element payload (the patch's element bytes, the `bytes` of
//    narrowphase_patch_collection_type, collision_object/types.hpp:187-201,
//    gathered in patch order as prepare_local_patch_for_sending does,
//    collision_object/impl.hpp:90-124): one fixed-stride record per element,
//      u64 global_id | the element's doubles (a 3-vertex triangle: 9,
//      v0 v1 v2 x y z each; nv vertices: 3 nv; a sphere: x y z r; a
//      snapshot: 26 k-DOP slabs then 3 centroid coordinates);
//  * result message (result_msg::serialize, collision_object/types.hpp:
//    109-138): u64 stride | u64 byte count | the records (32-byte
//    detailed_narrowphase_result each).
// DARMA checkpoint's own framing around a message is not reproduced.
#include "device.hpp"
#include "kernels.hpp"

namespace dbvh {

// od patch records from the exact patch slabs (ordered u64), the split
// offsets and the patch centroids (od x 3)
__global__ __launch_bounds__(64) void k_pack_patches(const unsigned long long* __restrict__ patch_k64,
                                                     const uint32_t* __restrict__ splits, const uint32_t* roots,
                                                     const double* __restrict__ cent, uint64_t n, uint64_t od,
                                                     uint64_t* __restrict__ out) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= od) return;
  const uint64_t b = (k == 0) ? 0 : splits[k - 1];
  const uint64_t e = (k == od - 1) ? n : splits[k];
  uint64_t* r = out + k * (kPatchRecordBytes / 8);
  r[0] = k;
  r[1] = e - b;
  const bool empty = roots[k] == kEnd;
  double* kd = reinterpret_cast<double*>(r + 2);
#pragma unroll
  for (int a = 0; a < kAxes; ++a) {
    // an empty patch keeps the default extent {max(), lowest()} (kdop.hpp:59-60)
    kd[2 * a] = empty ? 1.7976931348623157e308 : from_ordered_u64(patch_k64[k * kSlabs + 2 * a]);
    kd[2 * a + 1] = empty ? -1.7976931348623157e308 : from_ordered_u64(patch_k64[k * kSlabs + 2 * a + 1]);
  }
  double* c = kd + kSlabs;
  c[0] = cent[3 * k];
  c[1] = cent[3 * k + 1];
  c[2] = cent[3 * k + 2];
  c[3] = 0.0;
}

// One element record per thread: the element at input index src[i]
__global__ __launch_bounds__(256) void k_pack_elements(ElemInput in, const uint64_t* __restrict__ gids,
                                                       const uint32_t* __restrict__ src, uint64_t count,
                                                       uint32_t planes, uint64_t* __restrict__ out) {
  const uint64_t stride = 1 + planes;  // in 8-byte words
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t e = src[i];
    uint64_t* r = out + i * stride;
    r[0] = gids ? gids[e] : (uint64_t)e;
    double* d = reinterpret_cast<double*>(r + 1);
    if (in.kind == kSnapshot) {
      for (int j = 0; j < kSlabs; ++j) d[j] = in.data[(uint64_t)e * kSlabs + j];
      for (int c = 0; c < 3; ++c) d[kSlabs + c] = in.centroids[(uint64_t)e * 3 + c];
    } else {
      // plane p of an element: vertex p / 3, coordinate p % 3 (spheres: x y z r)
      for (uint32_t p = 0; p < planes; ++p) d[p] = in.data[(uint64_t)p * in.n + e];
    }
  }
}

// records back into SoA planes (planes x count) and global ids; snapshots
// into count x 26 k-DOPs and count x 3 centroids (planes == 29)
__global__ __launch_bounds__(256) void k_unpack_elements(const uint64_t* __restrict__ in, uint64_t count,
                                                         uint32_t planes, int snapshot, double* __restrict__ data,
                                                         double* __restrict__ cent, uint64_t* __restrict__ gids) {
  const uint64_t stride = 1 + planes;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t* r = in + i * stride;
    if (gids) gids[i] = r[0];
    const double* d = reinterpret_cast<const double*>(r + 1);
    if (snapshot) {
      for (int j = 0; j < kSlabs; ++j) data[i * kSlabs + j] = d[j];
      for (int c = 0; c < 3; ++c) cent[i * 3 + c] = d[kSlabs + c];
    } else {
      for (uint32_t p = 0; p < planes; ++p) data[(uint64_t)p * count + i] = d[p];
    }
  }
}

// positions [begin[k], end[k]) of the listed patches, concatenated: the
// input index of the element at each position of the reference order
__global__ __launch_bounds__(256) void k_patch_positions(const uint32_t* __restrict__ ref_perm,
                                                         const uint64_t* __restrict__ range, uint64_t npatch,
                                                         uint32_t* __restrict__ src) {
  // range: per listed patch {begin, end, output offset}
  const uint64_t k = blockIdx.y;
  if (k >= npatch) return;
  const uint64_t b = range[3 * k], e = range[3 * k + 1], o = range[3 * k + 2];
  for (uint64_t p = b + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < e; p += (uint64_t)gridDim.x * blockDim.x)
    src[o + (p - b)] = ref_perm[p];
}

__global__ void k_result_header(uint64_t* out, uint64_t count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out[0] = kResultRecordBytes;
    out[1] = count * kResultRecordBytes;
  }
}

static unsigned blocks_for(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (unsigned)(b > 8192 ? 8192 : (b ? b : 1));
}

hipError_t launch_pack_patches(const unsigned long long* patch_k64, const uint32_t* splits, const uint32_t* roots,
                               const double* cent, uint64_t n, uint64_t od, void* out, hipStream_t s) {
  if (od == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_patches, dim3((unsigned)((od + 63) / 64)), dim3(64), 0, s, patch_k64, splits, roots, cent,
                     n, od, (uint64_t*)out);
  return hipGetLastError();
}

hipError_t launch_patch_positions(const uint32_t* ref_perm, const uint64_t* range, uint64_t npatch,
                                  uint64_t max_len, uint32_t* src, hipStream_t s) {
  if (npatch == 0 || max_len == 0) return hipSuccess;
  if (npatch > 65535) return hipErrorInvalidValue;
  uint64_t bx = (max_len + 255) / 256;
  if (bx > 1024) bx = 1024;
  hipLaunchKernelGGL(k_patch_positions, dim3((unsigned)bx, (unsigned)npatch), dim3(256), 0, s, ref_perm, range,
                     npatch, src);
  return hipGetLastError();
}

hipError_t launch_pack_elements(const ElemInput& in, const uint64_t* gids, const uint32_t* src, uint64_t count,
                                uint32_t planes, void* out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_elements, dim3(blocks_for(count)), dim3(256), 0, s, in, gids, src, count, planes,
                     (uint64_t*)out);
  return hipGetLastError();
}

hipError_t launch_unpack_elements(const void* in, uint64_t count, uint32_t planes, int snapshot, double* data,
                                  double* cent, uint64_t* gids, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_unpack_elements, dim3(blocks_for(count)), dim3(256), 0, s, (const uint64_t*)in, count, planes,
                     snapshot, data, cent, gids);
  return hipGetLastError();
}

hipError_t launch_result_header(void* out, uint64_t count, hipStream_t s) {
  hipLaunchKernelGGL(k_result_header, dim3(1), dim3(64), 0, s, (uint64_t*)out, count);
  return hipGetLastError();
}

}  // namespace dbvh
