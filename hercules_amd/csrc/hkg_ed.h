// Device helpers shared by the query kernels: the streaming row load, one
// 8-float chunk of the squared Euclidean distance in the reference's order,
// and the (distance, id) candidate key.  Compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hkg {

typedef float nf4 __attribute__((ext_vector_type(4)));

// streaming 16-byte load (the rows are touched once per query)
__device__ __forceinline__ float4 ld_stream(const float *p)
{
  nf4 v = __builtin_nontemporal_load(reinterpret_cast<const nf4 *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

// one 8-float chunk of ts_euclidean_distance_SIMD (code/hercules/src/ts.c:83-109):
// separate sub and mul roundings, then the two-hadd tree
// ((d0+d1)+(d2+d3)) + ((d4+d5)+(d6+d7))
__device__ __forceinline__ float ed_chunk(float4 qa, float4 qb, float4 xa, float4 xb)
{
  float d0 = qa.x - xa.x, d1 = qa.y - xa.y, d2 = qa.z - xa.z, d3 = qa.w - xa.w;
  float d4 = qb.x - xb.x, d5 = qb.y - xb.y, d6 = qb.z - xb.z, d7 = qb.w - xb.w;
  d0 = d0 * d0; d1 = d1 * d1; d2 = d2 * d2; d3 = d3 * d3;
  d4 = d4 * d4; d5 = d5 * d5; d6 = d6 * d6; d7 = d7 * d7;
  return ((d0 + d1) + (d2 + d3)) + ((d4 + d5) + (d6 + d7));
}

// a candidate: FP32 squared distances are non-negative, so the bits order
// like the values and the key orders by (d2, id)
__device__ __forceinline__ uint64_t make_key(float d2, uint32_t id)
{
  return ((uint64_t)__float_as_uint(d2) << 32) | (uint64_t)id;
}

}  // namespace hkg
