// Exact full-scan k-NN that allocates nothing: hkg_scan_knn_device_noalloc.
//
// The same kernels as the full-scan path of hkg_knn_device (k_scan, then
// k_select), with their scratch -- the per-query k-th distance bounds,
// candidate counts, the candidate buffer and a few status words -- in one
// file-scope device array instead of a buffer set from hipMalloc.  For
// callers that cannot allocate device memory (code between HIP graph
// captures, tight serving loops, hosts that forbid it).
//
// The array is a host-side symbol: its device address comes from
// hipGetSymbolAddress once per device, when an index is opened on that device
// (noalloc_setup, called from Index::init_device), and every later use goes
// through that address.  Each device holds its own copy until the process
// ends.  Calls on one device are serialised by a lock here (each returns only
// after its work is done), so they never share the array at the same time;
// the array keeps the last call's contents, and everything a call reads is
// written by that call first (the bounds and counts by a reset kernel, the
// status words by a fill kernel).
#include <mutex>

#include "hkg_engine.h"
#include "hkg_kernels.h"

namespace hkg {

constexpr size_t kNoAllocBytes = 16ull << 20;  // the array: 16 MiB per device
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

static std::mutex g_noalloc_mu[kNoAllocMaxDevices];
static void *g_noalloc_ptr[kNoAllocMaxDevices];  // device address of scratch, per device (guarded by the lock)

// once per device, with that device current
hkg_status noalloc_setup(int device)
{
  if (device < 0 || device >= kNoAllocMaxDevices) {
    set_error("device ordinal beyond the no-allocation entry point's table");
    return HKG_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_noalloc_mu[device]);
  if (g_noalloc_ptr[device]) return HKG_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return HKG_EDEVICE;
  }
  g_noalloc_ptr[device] = p;
  return HKG_OK;
}

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// bytes of the array a call of nq queries at kmax needs (0: too many)
static size_t noalloc_need(uint32_t nq, uint32_t cap)
{
  return 4 * up256((size_t)nq * 4) + up256(8 * 4) + (size_t)nq * cap * 8;
}

hkg_status scan_knn_noalloc(Index &ix, const float *dQ, uint32_t nq, const uint32_t *dk, uint32_t kmax,
                            uint64_t *dids, float *dd2, hipStream_t st)
{
  DevIndex &dv = ix.dev;
  if (!dQ || !dk || !dids || !dd2) {
    set_error("null argument");
    return HKG_EINVAL;
  }
  if (kmax == 0 || kmax > (uint32_t)kMaxK) {
    set_error("k must be in [1, 256]");
    return HKG_EINVAL;
  }
  if (nq == 0) return HKG_OK;
  // the candidate buffer holds every block's flushed list: it cannot
  // overflow (one list of at most k entries per block of the scan grid).
  // The grid is the default one, never a tuned one, so the limit holds.
  const uint32_t gx = scan_grid_x(nq, dv.L, kScanBlocksPerCU);
  const uint32_t cap = gx * kmax + 256;
  const size_t need = noalloc_need(nq, cap);
  if (need > kNoAllocBytes) {
    set_error("hkg_scan_knn_device_noalloc: " + std::to_string(nq) + " queries at kmax " + std::to_string(kmax) +
              " need " + std::to_string(need) + " bytes of scratch, beyond its " + std::to_string(kNoAllocBytes) +
              " (at most 64 queries per call at any kmax)");
    return HKG_EINVAL;
  }
  if (dv.device < 0 || dv.device >= kNoAllocMaxDevices) {
    set_error("device ordinal beyond the no-allocation entry point's table");
    return HKG_EINVAL;
  }
  HKG_HIP_TRY(hipSetDevice(dv.device));
  if (!st) st = ix.stream;
  std::lock_guard<std::mutex> lock(g_noalloc_mu[dv.device]);
  uint8_t *base = static_cast<uint8_t *>(g_noalloc_ptr[dv.device]);
  if (!base) {
    set_error("no-allocation scratch not set up on this device");
    return HKG_EINTERNAL;
  }
  // layout: kq, bound, cand_cnt, (one spare), status words, candidates
  uint32_t *kq = reinterpret_cast<uint32_t *>(base);
  uint32_t *bound = reinterpret_cast<uint32_t *>(base + up256((size_t)nq * 4));
  uint32_t *cand_cnt = reinterpret_cast<uint32_t *>(base + 2 * up256((size_t)nq * 4));
  uint32_t *status = reinterpret_cast<uint32_t *>(base + 4 * up256((size_t)nq * 4));
  uint64_t *cand = reinterpret_cast<uint64_t *>(base + 4 * up256((size_t)nq * 4) + up256(8 * 4));
  HKG_HIP_TRY(hipMemcpyAsync(kq, dk, (size_t)nq * 4, hipMemcpyDeviceToDevice, st));
  launch_reset(st, nq, bound, cand_cnt, nullptr);
  launch_fill(st, status, 8, 0);
  launch_check_k(st, nq, kmax, kq, status + 3);
  launch_scan(st, gx, dv.X, dv.n, dv.L, dQ, nq, kq, dv.ids, cap, cand, cand_cnt, bound);
  launch_select(st, nq, kmax, kq, cap, cand, cand_cnt, ix.idmap64, dids, dd2, status + 2);
  HKG_HIP_TRY(hipGetLastError());
  uint32_t host_status[8] = {0};
  HKG_HIP_TRY(hipMemcpyAsync(host_status, status, sizeof(host_status), hipMemcpyDeviceToHost, st));
  HKG_HIP_TRY(hipStreamSynchronize(st));
  if (host_status[3]) {
    set_error(std::to_string(host_status[3]) + " per-query k value(s) outside [1, kmax = " + std::to_string(kmax) +
              "]");
    return HKG_EINVAL;
  }
  if (host_status[2]) {
    set_error("hkg_scan_knn_device_noalloc: a candidate buffer overflowed");
    return HKG_EINTERNAL;
  }
  return HKG_OK;
}

}  // namespace hkg

extern "C" hkg_status hkg_scan_knn_device_noalloc(hkg_index *ix, const float *dQ, uint32_t nq, const uint32_t *dk,
                                                  uint32_t kmax, uint64_t *dids, float *dd2, void *stream)
{
  if (!ix) return HKG_EINVAL;
  return hkg::scan_knn_noalloc(*reinterpret_cast<hkg::Index *>(ix), dQ, nq, dk, kmax, dids, dd2,
                               (hipStream_t)stream);
}
