// Internal declarations shared by the host side and the HIP kernels of
// libhkg.  Nothing here is part of the C ABI (include/hkg.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/hkg.h"

namespace hkg {

// ---- limits ---------------------------------------------------------------
constexpr int kPaaSegments = 16;    // iSAX words are 16 symbols (sax.c:320-344)
constexpr int kMaxK = 256;          // per-query k
constexpr int kTopKCap = 512;       // LDS merge buffer (list + staged)
constexpr int kBlock = 256;         // threads per block of the query kernels
constexpr uint32_t kMaxLength = 16384;  // longest series the query paths take (64 KB of LDS)
constexpr uint32_t kMaxBuildLength = 2048;  // longest series the device builder takes
constexpr uint32_t kNoId32 = 0xFFFFFFFFu;
constexpr uint32_t kInfBits = 0x7F7FFFFFu;  // FLT_MAX bits: "no bound yet"

// Lower bounds are compared with a certified slack: an FP32 squared
// distance can sit below the exact one by up to ~L*2^-24 relative, and the
// node/iSAX summaries were rounded to FP32 when built, so a candidate is
// dropped only when  lb > kth * (1 + kLbRel) + kLbAbs.
constexpr float kLbRel = 1.0e-4f;
constexpr float kLbAbs = 1.0e-5f;

// ---- host-side tree (parsed root.idx or built) ------------------------------
struct HostNode {
  uint8_t is_leaf = 0;
  uint32_t level = 0;
  uint32_t size = 0;
  int32_t left = -1, right = -1;  // node indices
  uint32_t seg_off = 0;           // into seg_end / ind
  uint16_t nseg = 0;
  uint64_t file_pos = 0;          // leaves only
  int32_t leaf_idx = -1;          // leaves only (DFS order)
  // internal nodes: the split policy (hercules_node_split.h:22-36)
  int16_t split_from = 0, split_to = 0, split_segment = 0;
  int32_t indicator_split_idx = 0;
  float indicator_split_value = 0.0f;
  std::string fname;              // leaves only
};

struct HostTree {
  // header (hercules_index.c:1755-1763)
  uint64_t leaf_count_hdr = 0;
  double buffered_memory_size = 0.0;
  uint32_t L = 0, init_segments = 1;
  int32_t paa_segments = 16, sax_bit_cardinality = 8;
  uint32_t max_leaf_size = 0;
  int32_t serial = 1, sims = 1;
  // pre-order nodes
  std::vector<HostNode> nodes;
  std::vector<int16_t> seg_end;
  std::vector<float> ind;     // 4 per segment {max mean, min mean, max sd, min sd}
  std::vector<int32_t> leaves;  // node index of each leaf, DFS (file) order
  uint64_t n_series = 0;
};

// parse root.idx; returns HKG_OK or HKG_EIO/HKG_EFORMAT with a message
hkg_status parse_root_idx(const std::string &path, HostTree &t, std::string &err);
// serialise root.idx in the same format
hkg_status write_root_idx(const std::string &path, const HostTree &t, std::string &err);

// ---- device-resident index ----------------------------------------------------
struct DevIndex {
  int device = 0;
  uint64_t n = 0;          // series
  uint32_t L = 0;
  uint32_t nleaf = 0;
  uint32_t max_leaf = 0;
  bool has_tree = false;
  // SoA in HBM
  float *X = nullptr;          // n x L, leaf order
  uint8_t *sax = nullptr;      // n x 16
  float2 *lohi = nullptr;      // 256 {lower, upper} breakpoints per iSAX symbol
  uint32_t *ids = nullptr;     // n, public id per row (nullable: row = id)
  uint64_t *leaf_pos = nullptr;   // nleaf
  uint32_t *leaf_size = nullptr;  // nleaf
  uint32_t *leaf_seg_off = nullptr;  // nleaf + 1
  int16_t *seg_end = nullptr;        // total segments
  float4 *seg_ind = nullptr;         // total segments
  // every node (pre-order), for the approximate modes' traversals
  uint32_t nnodes = 0;
  int4 *nd_link = nullptr;     // {left, right, leaf index or -1, is_leaf}
  int4 *nd_split = nullptr;    // internal: {split_from, split_to, indicator (0 mean, 1 sd), value bits}
  uint2 *nd_seg = nullptr;     // {first segment, segments} into nd_seg_end / nd_ind
  int16_t *nd_seg_end = nullptr;
  float4 *nd_ind = nullptr;
  int32_t *leaf_node = nullptr;  // nleaf: node of each leaf
  // BF16 filter data (optional)
  uint16_t *Xb = nullptr;      // n x L bf16 bits (round to nearest even)
  float4 *xst = nullptr;       // n x {|x|^2 down, 2|x| up, 2(|x - x~| + g |x~|) up, g |x|^2 up}, x~ = bf16(x),
                               // g = 2 (L + 1) 2^-24
  float *xn2 = nullptr;        // n x |x|^2 (down), contiguous for the filter's tile loads
  float2 *tile_slack = nullptr;  // per 96-row filter tile: max of xst.y and of xst.z over its rows
  uint32_t ntiles = 0;         // 96-row tiles of the BF16 filter
  uint32_t *tile_leaf0 = nullptr, *tile_leaf1 = nullptr;  // leaves overlapping each tile
  uint64_t device_bytes = 0;
  // host copies kept for writing / info
  HostTree tree;
  std::vector<uint64_t> host_ids;  // optional, full 64-bit
};

// ---- kernels (hkg_kernels.hip) ------------------------------------------------
struct QueryBuffers;  // defined in hkg_engine.cpp

void upload_breakpoints(int device);
void host_breakpoints(float *out255);

}  // namespace hkg

// error helpers
#define HKG_HIP_TRY(expr)                                                       \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      hkg::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));        \
      return HKG_EDEVICE;                                                       \
    }                                                                           \
  } while (0)

namespace hkg {
void set_error(const std::string &msg);
}
