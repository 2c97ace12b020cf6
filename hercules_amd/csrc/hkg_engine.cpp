// Host orchestration of the query kernels and the C ABI entry points for
// opening, querying and closing an index.
#include <errno.h>
#include <float.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>

#include "hkg_engine.h"
#include "hkg_kernels.h"

namespace hkg {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const std::string &last_error() { return g_err; }

// ---------------------------------------------------------------------------
// device memory helpers
// ---------------------------------------------------------------------------
template <class T>
static hkg_status dalloc(T **p, size_t count, uint64_t *acct)
{
  *p = nullptr;
  if (count == 0) return HKG_OK;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error(std::string("hipMalloc(") + std::to_string(count * sizeof(T)) + " B): " + hipGetErrorString(e));
    *p = nullptr;
    return HKG_ENOMEM;
  }
  if (acct) *acct += count * sizeof(T);
  return HKG_OK;
}

template <class T>
static void dfree(T *&p)
{
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// host file -> device, through two pinned staging buffers
static hkg_status upload_file(const std::string &path, void *dst, uint64_t bytes, hipStream_t st)
{
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) {
    set_error(path + ": " + strerror(errno));
    return HKG_EIO;
  }
  fseek(f, 0, SEEK_END);
  uint64_t sz = (uint64_t)ftell(f);
  fseek(f, 0, SEEK_SET);
  if (sz != bytes) {
    fclose(f);
    set_error(path + ": size " + std::to_string(sz) + " != expected " + std::to_string(bytes));
    return HKG_EFORMAT;
  }
  const size_t chunk = 64ull << 20;
  void *pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2];
  if (hipHostMalloc(&pin[0], chunk, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&pin[1], chunk, hipHostMallocDefault) != hipSuccess) {
    fclose(f);
    if (pin[0]) (void)hipHostFree(pin[0]);
    set_error("hipHostMalloc failed");
    return HKG_ENOMEM;
  }
  (void)hipEventCreate(&ev[0]);
  (void)hipEventCreate(&ev[1]);
  hkg_status rc = HKG_OK;
  bool used[2] = {false, false};
  uint64_t off = 0;
  int b = 0;
  while (off < bytes) {
    size_t n = (size_t)std::min<uint64_t>(chunk, bytes - off);
    if (used[b]) (void)hipEventSynchronize(ev[b]);
    if (fread(pin[b], 1, n, f) != n) {
      set_error(path + ": short read");
      rc = HKG_EIO;
      break;
    }
    if (hipMemcpyAsync((char *)dst + off, pin[b], n, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipEventRecord(ev[b], st) != hipSuccess) {
      set_error("hipMemcpyAsync failed while uploading " + path);
      rc = HKG_EDEVICE;
      break;
    }
    used[b] = true;
    off += n;
    b ^= 1;
  }
  (void)hipStreamSynchronize(st);
  fclose(f);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  (void)hipHostFree(pin[0]);
  (void)hipHostFree(pin[1]);
  return rc;
}

// ---------------------------------------------------------------------------
// index lifetime
// ---------------------------------------------------------------------------
Index::~Index()
{
  (void)hipSetDevice(dev.device);
  bufs.clear();
  dfree(dev.X);
  dfree(dev.sax);
  dfree(dev.lohi);
  dfree(dev.ids);
  dfree(dev.leaf_pos);
  dfree(dev.leaf_size);
  dfree(dev.leaf_seg_off);
  dfree(dev.seg_end);
  dfree(dev.seg_ind);
  dfree(dev.nd_link);
  dfree(dev.nd_split);
  dfree(dev.nd_seg);
  dfree(dev.nd_seg_end);
  dfree(dev.nd_ind);
  dfree(dev.leaf_node);
  dfree(dev.Xb);
  dfree(dev.xst);
  dfree(dev.xn2);
  dfree(dev.tile_slack);
  dfree(dev.tile_leaf0);
  dfree(dev.tile_leaf1);
  dfree(idmap64);
  if (stream) (void)hipStreamDestroy(stream);
}

hkg_status Index::init_device(int device)
{
  dev.device = device;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    set_error("no HIP device available");
    return HKG_EDEVICE;
  }
  if (device < 0 || device >= ndev) {
    set_error("device ordinal out of range");
    return HKG_EINVAL;
  }
  HKG_HIP_TRY(hipSetDevice(device));
  HKG_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HKG_HIP_TRY(upload_breakpoints_dev());
  hkg_status rc = upload_isax_tables(dev);
  if (rc != HKG_OK) return rc;
  set_lds_limits();
  set_mode_lds_limits();
  return noalloc_setup(device);
}

// leaf tables (CSR of segments) from a host tree
hkg_status Index::upload_tree()
{
  const HostTree &t = dev.tree;
  const uint32_t nl = (uint32_t)t.leaves.size();
  std::vector<uint64_t> pos;
  std::vector<uint32_t> size, off;
  std::vector<int16_t> se;
  std::vector<float4> ind;
  pos.reserve(nl);
  size.reserve(nl);
  off.reserve(nl + 1);
  uint32_t maxleaf = 0;
  for (uint32_t i = 0; i < nl; ++i) {
    const HostNode &nd = t.nodes[t.leaves[i]];
    pos.push_back(nd.file_pos);
    size.push_back(nd.size);
    maxleaf = std::max(maxleaf, nd.size);
    off.push_back((uint32_t)se.size());
    for (uint32_t s = 0; s < nd.nseg; ++s) {
      se.push_back(t.seg_end[nd.seg_off + s]);
      const float *g = &t.ind[4 * (nd.seg_off + s)];
      ind.push_back(make_float4(g[0], g[1], g[2], g[3]));
    }
  }
  off.push_back((uint32_t)se.size());
  dev.nleaf = nl;
  dev.max_leaf = maxleaf;
  hkg_status rc;
  if ((rc = dalloc(&dev.leaf_pos, nl, &dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&dev.leaf_size, nl, &dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&dev.leaf_seg_off, nl + 1, &dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&dev.seg_end, se.size(), &dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&dev.seg_ind, ind.size(), &dev.device_bytes)) != HKG_OK) return rc;
  HKG_HIP_TRY(hipMemcpy(dev.leaf_pos, pos.data(), nl * 8, hipMemcpyHostToDevice));
  HKG_HIP_TRY(hipMemcpy(dev.leaf_size, size.data(), nl * 4, hipMemcpyHostToDevice));
  HKG_HIP_TRY(hipMemcpy(dev.leaf_seg_off, off.data(), (nl + 1) * 4, hipMemcpyHostToDevice));
  if (!se.empty()) {
    HKG_HIP_TRY(hipMemcpy(dev.seg_end, se.data(), se.size() * 2, hipMemcpyHostToDevice));
    HKG_HIP_TRY(hipMemcpy(dev.seg_ind, ind.data(), ind.size() * sizeof(float4), hipMemcpyHostToDevice));
  }
  // every node, for the approximate modes
  const uint32_t nn = (uint32_t)t.nodes.size();
  std::vector<int4> link(nn), split(nn);
  std::vector<uint2> nseg(nn);
  for (uint32_t i = 0; i < nn; ++i) {
    const HostNode &nd = t.nodes[i];
    link[i] = make_int4(nd.left, nd.right, nd.leaf_idx, nd.is_leaf ? 1 : 0);
    int32_t vbits;
    memcpy(&vbits, &nd.indicator_split_value, 4);
    split[i] = make_int4(nd.split_from, nd.split_to, nd.indicator_split_idx, vbits);
    nseg[i] = make_uint2(nd.seg_off, nd.nseg);
  }
  std::vector<float4> all_ind(t.seg_end.size());
  for (size_t s = 0; s < t.seg_end.size(); ++s)
    all_ind[s] = make_float4(t.ind[4 * s], t.ind[4 * s + 1], t.ind[4 * s + 2], t.ind[4 * s + 3]);
  dev.nnodes = nn;
  if ((rc = dalloc(&dev.nd_link, std::max<uint32_t>(nn, 1), &dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&dev.nd_split, std::max<uint32_t>(nn, 1), &dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&dev.nd_seg, std::max<uint32_t>(nn, 1), &dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&dev.nd_seg_end, std::max<size_t>(t.seg_end.size(), 1), &dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&dev.nd_ind, std::max<size_t>(t.seg_end.size(), 1), &dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&dev.leaf_node, std::max<uint32_t>(nl, 1), &dev.device_bytes)) != HKG_OK) return rc;
  if (nn) {
    HKG_HIP_TRY(hipMemcpy(dev.nd_link, link.data(), nn * sizeof(int4), hipMemcpyHostToDevice));
    HKG_HIP_TRY(hipMemcpy(dev.nd_split, split.data(), nn * sizeof(int4), hipMemcpyHostToDevice));
    HKG_HIP_TRY(hipMemcpy(dev.nd_seg, nseg.data(), nn * sizeof(uint2), hipMemcpyHostToDevice));
  }
  if (!t.seg_end.empty()) {
    HKG_HIP_TRY(hipMemcpy(dev.nd_seg_end, t.seg_end.data(), t.seg_end.size() * 2, hipMemcpyHostToDevice));
    HKG_HIP_TRY(hipMemcpy(dev.nd_ind, all_ind.data(), all_ind.size() * sizeof(float4), hipMemcpyHostToDevice));
  }
  if (nl) HKG_HIP_TRY(hipMemcpy(dev.leaf_node, t.leaves.data(), nl * 4, hipMemcpyHostToDevice));
  dev.has_tree = true;
  return HKG_OK;
}

hkg_status open_dir(const char *dir, const hkg_options *opt, Index **out)
{
  *out = nullptr;
  if (!dir) {
    set_error("null index directory");
    return HKG_EINVAL;
  }
  hkg_options o;
  hkg_default_options(&o);
  if (opt) o = *opt;
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d += '/';
  std::unique_ptr<Index> ix(new Index());
  std::string err;
  hkg_status rc = parse_root_idx(d + "root.idx", ix->dev.tree, err);
  if (rc != HKG_OK) {
    set_error(d + "root.idx: " + err);
    return rc;
  }
  const HostTree &t = ix->dev.tree;
  if (t.L > kMaxLength) {
    set_error("series longer than " + std::to_string(kMaxLength) + " are not supported");
    return HKG_EINVAL;
  }
  if (t.paa_segments != kPaaSegments || t.sax_bit_cardinality != 8) {
    set_error("only 16 iSAX segments of 8-bit cardinality are supported");
    return HKG_EFORMAT;
  }
  if (!t.sims) {
    set_error("index was built without --sims (no leaves_sims.idx)");
    return HKG_EFORMAT;
  }
  if ((rc = ix->init_device(o.device)) != HKG_OK) return rc;
  ix->dev.n = t.n_series;
  ix->dev.L = t.L;
  if ((rc = ix->upload_tree()) != HKG_OK) return rc;
  const uint64_t n = ix->dev.n, L = ix->dev.L;
  if ((rc = dalloc(&ix->dev.X, n * L, &ix->dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&ix->dev.sax, n * kPaaSegments, &ix->dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = upload_file(d + "leaves_raw.idx", ix->dev.X, n * L * 4, ix->stream)) != HKG_OK) return rc;
  if ((rc = upload_file(d + "leaves_sims.idx", ix->dev.sax, n * kPaaSegments, ix->stream)) != HKG_OK) return rc;
  if (o.keep_bf16 && (rc = ix->make_bf16()) != HKG_OK) return rc;
  if (o.verbose)
    fprintf(stderr, "hkg: loaded %s: %llu series x %u, %u leaves, %.2f GB in HBM\n", dir,
            (unsigned long long)n, (unsigned)L, ix->dev.nleaf, ix->dev.device_bytes / 1e9);
  *out = ix.release();
  return HKG_OK;
}

hkg_status open_raw(const float *X, uint64_t n, uint32_t L, const hkg_options *opt, Index **out)
{
  *out = nullptr;
  if (!X || n == 0 || L == 0 || L % 8 != 0 || L > kMaxLength || n >= (1ull << 32)) {
    set_error("open_raw: need n in [1, 2^32) rows and L a positive multiple of 8, at most " +
              std::to_string(kMaxLength));
    return HKG_EINVAL;
  }
  hkg_options o;
  hkg_default_options(&o);
  if (opt) o = *opt;
  std::unique_ptr<Index> ix(new Index());
  hkg_status rc;
  if ((rc = ix->init_device(o.device)) != HKG_OK) return rc;
  ix->dev.n = n;
  ix->dev.L = L;
  if ((rc = dalloc(&ix->dev.X, n * L, &ix->dev.device_bytes)) != HKG_OK) return rc;
  HKG_HIP_TRY(hipMemcpy(ix->dev.X, X, n * L * 4, hipMemcpyHostToDevice));
  if (o.keep_bf16 && (rc = ix->make_bf16()) != HKG_OK) return rc;
  *out = ix.release();
  return HKG_OK;
}

hkg_status Index::set_ids(const uint64_t *ids, uint64_t n)
{
  if (n != dev.n) {
    set_error("id map length differs from the number of series");
    return HKG_EINVAL;
  }
  // the device keys carry 32-bit ids; the 64-bit map translates them back
  std::vector<uint32_t> rank(n);
  // the key's low word must order like the public id: use the rank of the id
  std::vector<uint64_t> sorted(ids, ids + n);
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) {
    set_error("id map has duplicate ids");
    return HKG_EINVAL;
  }
  for (uint64_t i = 0; i < n; ++i)
    rank[i] = (uint32_t)(std::lower_bound(sorted.begin(), sorted.end(), ids[i]) - sorted.begin());
  hkg_status rc;
  (void)hipSetDevice(dev.device);
  dfree(dev.ids);
  dfree(idmap64);
  if ((rc = dalloc(&dev.ids, n, &dev.device_bytes)) != HKG_OK) return rc;
  if ((rc = dalloc(&idmap64, n, &dev.device_bytes)) != HKG_OK) return rc;
  HKG_HIP_TRY(hipMemcpy(dev.ids, rank.data(), n * 4, hipMemcpyHostToDevice));
  HKG_HIP_TRY(hipMemcpy(idmap64, sorted.data(), n * 8, hipMemcpyHostToDevice));
  dev.host_ids.assign(ids, ids + n);
  return HKG_OK;
}

// ---------------------------------------------------------------------------
// per-batch device buffers
// ---------------------------------------------------------------------------
void QueryBuffers::free_bf16()
{
  dfree(bf_cnt);
  dfree(bf_surv);
  dfree(Qb);
  dfree(qst);
  dfree(qperm);
  if (bf_sort_tmp) (void)hipFree(bf_sort_tmp);
  bf_sort_tmp = nullptr;
  dfree(need);
  dfree(tiles);
  dfree(ntiles_g);
  dfree(thr);
  dfree(hist);
  dfree(bf_region);
  dfree(bf_region_cnt);
  dfree(bf_prog);
  dfree(slot_need8);
  dfree(slot_need);
}

QueryBuffers::~QueryBuffers()
{
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  dfree(Q);
  dfree(kq);
  dfree(paa);
  dfree(psum);
  dfree(qn2);
  dfree(lb);
  dfree(best);
  dfree(bound);
  dfree(cand);
  dfree(cand_cnt);
  dfree(work_keys);
  dfree(work_vals);
  if (sort_tmp) (void)hipFree(sort_tmp);
  dfree(seed);
  free_bf16();
  free_isax();
  dfree(misc);
  dfree(counters);
  dfree(out_ids);
  dfree(out_d2);
  dfree(nlb);
  dfree(heap);
  dfree(bfs);
  dfree(info);
  dfree(mstate);
  dfree(vtop);
  dfree(lab);
}

// buffer sets kept alive per index (each a batch's scratch: up to a few GB)
constexpr size_t kMaxBufferSets = 16;

hkg_status Index::get_buffers(uint32_t B, uint32_t kmax, uint32_t cap, QueryBuffers **out)
{
  B = std::max(B, call_settings().min_buffer_batch);
  const auto key = std::make_tuple(B, kmax, cap);
  std::lock_guard<std::mutex> lock(mu_bufs);
  // a cached set of the same shape that no call holds (an overflow re-run
  // of a whole batch asks for the shape of its caller)
  auto range = bufs.equal_range(key);
  for (auto it = range.first; it != range.second; ++it) {
    QueryBuffers &c = *it->second;
    if (!c.busy && c.B == B && c.kmax == kmax && c.cap == cap) {
      c.busy = true;
      c.last_use = ++buf_clock;
      *out = &c;
      return HKG_OK;
    }
  }
  // keep at most kMaxBufferSets sets alive, dropping the least recently
  // used one a call does not hold (a server's workers lease several shapes
  // in turn: freeing them all at once would re-allocate gigabytes per batch)
  auto evict_lru = [&]() {
    auto lru = bufs.end();
    for (auto e = bufs.begin(); e != bufs.end(); ++e)
      if (!e->second->busy && (lru == bufs.end() || e->second->last_use < lru->second->last_use)) lru = e;
    if (lru == bufs.end()) return false;
    bufs.erase(lru);
    ++sets_evicted;
    return true;
  };
  while (bufs.size() >= kMaxBufferSets && evict_lru()) {
  }
  ++sets_allocated;
  hkg_status rc = alloc_buffers(B, kmax, cap, out);
  if (rc == HKG_ENOMEM) {
    // device memory held by idle sets: drop them all and try once more
    while (evict_lru()) {
    }
    rc = alloc_buffers(B, kmax, cap, out);
  }
  return rc;
}

// a new buffer set (mu_bufs held), leased to the caller
hkg_status Index::alloc_buffers(uint32_t B, uint32_t kmax, uint32_t cap, QueryBuffers **out)
{
  const uint32_t L = dev.L, nleaf = dev.nleaf;
  const auto key = std::make_tuple(B, kmax, cap);
  std::unique_ptr<QueryBuffers> b(new QueryBuffers());
  b->B = B;
  b->kmax = kmax;
  b->cap = cap;
  if ((uint64_t)B * std::max<uint32_t>(nleaf, 1) >= 0x7FFFFFFFull) {
    set_error("batch x leaves exceeds 2^31; use a smaller batch");
    return HKG_EINVAL;
  }
  b->work_cap = B * std::max<uint32_t>(nleaf, 1);
  uint64_t acct = 0;
  hkg_status rc;
#define A(ptr, cnt) \
  if ((rc = dalloc(&b->ptr, (cnt), &acct)) != HKG_OK) return rc;
  A(Q, (uint64_t)B * L);
  A(kq, B);
  A(paa, (uint64_t)B * kPaaSegments);
  A(psum, (uint64_t)B * 2 * (L + 1));
  A(qn2, B);
  A(lb, (uint64_t)B * std::max<uint32_t>(nleaf, 1));
  A(best, (uint64_t)B * kBestLeaves);
  A(bound, B);
  A(cand, (uint64_t)B * cap);
  A(cand_cnt, B);
  const uint64_t pairs = dev.has_tree ? (uint64_t)B * nleaf : 1;
  A(work_keys, 2 * pairs);
  A(work_vals, 2 * pairs);
  A(seed, B);
  if (dev.has_tree) {
    b->sort_bytes = sort_work_temp_bytes((uint32_t)pairs);
    if (hipMalloc(&b->sort_tmp, std::max<size_t>(b->sort_bytes, 16)) != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipMalloc of the sort scratch failed");
      return HKG_ENOMEM;
    }
    acct += b->sort_bytes;
  }
  A(misc, 8);
  A(counters, kNumCounters);
  A(out_ids, (uint64_t)B * kmax);
  A(out_d2, (uint64_t)B * kmax);
#undef A
  b->bytes = acct;
  b->busy = true;
  b->last_use = ++buf_clock;
  *out = b.get();
  bufs.emplace(key, std::move(b));
  return HKG_OK;
}

hkg_status Index::prepare_buffers(hkg_path path, uint32_t B, uint32_t kmax, uint32_t count)
{
  if (path == HKG_PATH_AUTO) path = HKG_PATH_BATCHED;
  if (!dev.has_tree) path = HKG_PATH_SCAN;
  hkg_status rc = HKG_OK;
  if (path == HKG_PATH_BF16 && (rc = make_bf16()) != HKG_OK) return rc;
  HKG_HIP_TRY(hipSetDevice(dev.device));
  // the caps the batch paths compute for a set of B queries
  const uint32_t cap = path == HKG_PATH_BF16 ? cand_cap(HKG_PATH_INDEX, B, kmax) : cand_cap(path, B, kmax);
  std::vector<QueryBuffers *> held;
  for (uint32_t i = 0; i < count && rc == HKG_OK; ++i) {
    QueryBuffers *b = nullptr;
    rc = get_buffers(B, kmax, cap, &b);
    if (rc != HKG_OK) break;
    held.push_back(b);
    if (path == HKG_PATH_BF16) rc = ensure_bf16_scratch(*b);
  }
  for (QueryBuffers *b : held) release_buffers(b);
  return rc;
}

void Index::release_buffers(QueryBuffers *b)
{
  std::lock_guard<std::mutex> lock(mu_bufs);
  b->busy = false;
}

uint64_t Index::scratch_per_query(hkg_path path, uint32_t kmax) const
{
  const uint64_t L = dev.L, nl = std::max<uint32_t>(dev.nleaf, 1);
  const uint64_t cap = cand_cap(path, 1024, kmax);
  uint64_t b = L * 4 + 4 + kPaaSegments * 4 + 2 * (L + 1) * 8 + 8 + 4 * kBestLeaves + 8 + cap * 8 + 4 + 4 +
               kmax * 12 + 64;
  if (dev.has_tree) b += nl * 4 /* lb */ + nl * 16 /* work keys, values */ + nl * 8 /* sort scratch */;
  if (path == HKG_PATH_BF16) b += bf16_scratch_per_query(dev.L, nl, dev.ntiles);
  return b;
}

CallSettings &call_settings()
{
  static thread_local CallSettings s;
  return s;
}

// ---------------------------------------------------------------------------
// one batch through a path (device pointers; results left in b->out_*)
// ---------------------------------------------------------------------------
static uint32_t cu_count()
{
  static uint32_t cus = 0;
  if (!cus) {
    int dev = 0, n = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    cus = (uint32_t)std::max(n, 1);
  }
  return cus;
}

// persistent grid of k_filter_refine: blocks per CU from the call, else the
// tuned shape, else kFilterBlocksPerCU (any grid is correct: the blocks
// take work items from one queue)
uint32_t Index::filter_grid() const
{
  uint32_t per = call_settings().fblocks;
  if (!per) per = tuned_fblocks.load(std::memory_order_relaxed);
  if (!per) per = kFilterBlocksPerCU;
  return cu_count() * per;
}

bool Index::leaf_ordered(uint32_t nq) const
{
  if (!dev.has_tree) return false;
  const uint32_t f = call_settings().batched_filter;
  if (f == 1) return false;
  if (f == 2) return true;
  const uint32_t tf = tuned_filter.load(std::memory_order_relaxed);
  if (tf && nq == tuned_batch[HKG_PATH_BATCHED].load(std::memory_order_relaxed)) return tf == 2;
  return isax_batched(nq, dev.nleaf);
}


void Index::enqueue_index_batch(QueryBuffers &b, uint32_t nq, hipStream_t st, Prof *pf, bool seeded)
{
  const uint32_t L = dev.L, nleaf = dev.nleaf;
  uint32_t *nwork = b.misc + 0, *work_next = b.misc + 1, *overflow = b.misc + 2;
  launch_reset(st, nq, b.bound, b.cand_cnt, seeded ? b.seed : nullptr);
  launch_fill(st, b.misc, 8, 0);
  launch_check_k(st, nq, b.kmax, b.kq, b.misc + 3);
  launch_fill(st, reinterpret_cast<uint32_t *>(b.counters), 2 * kNumCounters, 0);
  if (pf) pf->mark(st, -1);
  launch_prep(st, b.Q, nq, L, b.paa, b.psum, b.qn2);
  if (pf) pf->mark(st, HKG_STAGE_PREP);
  launch_eapca(st, nq, nleaf, L, dev.leaf_seg_off, dev.seg_end, dev.seg_ind, b.psum, b.lb);
  if (pf) pf->mark(st, HKG_STAGE_EAPCA);
  launch_best_leaf(st, nq, nleaf, b.lb, b.best);
  if (pf) pf->mark(st, HKG_STAGE_SELECT_LEAF);
  launch_scan_best(st, nq, dev.X, L, b.Q, b.kq, b.best, dev.leaf_pos, dev.leaf_size, dev.ids, b.cap, b.cand,
                   b.cand_cnt, b.bound, b.counters);
  if (pf) pf->mark(st, HKG_STAGE_APPROX);
  const uint32_t pairs = nq * nleaf;
  uint32_t *keys2 = b.work_keys + b.work_cap, *vals2 = b.work_vals + b.work_cap;
  launch_collect(st, nq, nleaf, b.lb, b.best, b.bound, b.work_keys, b.work_vals, nwork);
  (void)launch_sort_work(st, b.sort_tmp, b.sort_bytes, b.work_keys, keys2, b.work_vals, vals2, pairs);
  if (pf) pf->mark(st, HKG_STAGE_COLLECT);
  launch_filter_refine(st, filter_grid(), dev.X, dev.sax, L, b.Q, b.paa, b.kq, dev.leaf_pos, dev.leaf_size,
                       dev.ids, nleaf, b.lb, vals2, nwork, work_next, b.cap, b.cand, b.cand_cnt, b.bound,
                       b.counters);
  if (pf) pf->mark(st, HKG_STAGE_REFINE);
  launch_select(st, nq, b.kmax, b.kq, b.cap, b.cand, b.cand_cnt, idmap64, b.out_ids, b.out_d2, overflow);
  if (pf) pf->mark(st, HKG_STAGE_TOPK);
}

// The approximate modes: exact node bounds, one block per query replaying
// the reference's traversal, then (epsilon search, refined branch) the
// candidate leaves through the work list and the iSAX filter + refinement.
hkg_status Index::enqueue_mode_batch(QueryBuffers &b, uint32_t nq, hipStream_t st, Prof *pf, bool seeded)
{
  const uint32_t L = dev.L, nleaf = dev.nleaf, nn = dev.nnodes;
  hkg_status rc;
  if (!b.nlb && (rc = dalloc(&b.nlb, (uint64_t)b.B * nn, &b.bytes)) != HKG_OK) return rc;
  if (!b.info && (rc = dalloc(&b.info, (uint64_t)b.B * 4, &b.bytes)) != HKG_OK) return rc;
  if (!mode_heap_in_lds(L, nn) && !b.heap && (rc = dalloc(&b.heap, (uint64_t)b.B * (nn + 1), &b.bytes)) != HKG_OK)
    return rc;
  if (!b.bfs && (rc = dalloc(&b.bfs, (uint64_t)b.B * 2 * (nn + 1), &b.bytes)) != HKG_OK) return rc;
  if (!b.mstate && (rc = dalloc(&b.mstate, (uint64_t)b.B * kModeState, &b.bytes)) != HKG_OK) return rc;
  if (!b.vtop && (rc = dalloc(&b.vtop, (uint64_t)b.B * b.kmax, &b.bytes)) != HKG_OK) return rc;
  if (!b.lab && (rc = dalloc(&b.lab, (uint64_t)b.B, &b.bytes)) != HKG_OK) return rc;
  ModeParams mp = call_settings().mp;
  mp.nleaves = nleaf;
  mp.nseries = dev.n;
  uint32_t *nwork = b.misc + 0, *work_next = b.misc + 1, *overflow = b.misc + 2;
  launch_reset(st, nq, b.bound, b.cand_cnt, seeded ? b.seed : nullptr);
  launch_fill(st, b.misc, 8, 0);
  launch_check_k(st, nq, b.kmax, b.kq, b.misc + 3);
  launch_fill(st, reinterpret_cast<uint32_t *>(b.counters), 2 * kNumCounters, 0);
  if (pf) pf->mark(st, -1);
  launch_prep(st, b.Q, nq, L, b.paa, b.psum, b.qn2);
  if (pf) pf->mark(st, HKG_STAGE_PREP);
  launch_eapca(st, nq, nleaf, L, dev.leaf_seg_off, dev.seg_end, dev.seg_ind, b.psum, b.lb);
  launch_node_lb(st, nq, nn, L, b.Q, dev.nd_seg, dev.nd_seg_end, dev.nd_ind, b.nlb);
  if (pf) pf->mark(st, HKG_STAGE_EAPCA);
  const uint32_t pairs = nq * nleaf;
  launch_mode_keys(st, pairs, b.work_keys, b.work_vals);
  launch_mode_search(st, nq, dev.X, dev.sax, L, b.Q, b.paa, b.kq, nn, dev.nd_link, dev.nd_split, b.nlb, nleaf,
                     dev.leaf_pos, dev.leaf_size, dev.leaf_node, dev.ids, b.lb, b.heap, b.bfs, mp, b.cap, b.cand,
                     b.cand_cnt, b.bound, b.work_keys, nwork, b.info, b.counters, b.mstate, b.vtop, b.kmax);
  if (mp.nprobes == 0) {
    launch_fill(st, reinterpret_cast<uint32_t *>(b.lab), 2 * nq, 0);
    launch_mode_count(st, nq, dev.sax, L, b.paa, nn, b.nlb, nleaf, dev.leaf_pos, dev.leaf_size, dev.leaf_node,
                      b.work_keys, b.mstate, b.lab);
    launch_mode_finish(st, nq, dev.X, L, b.Q, b.kq, nn, b.nlb, nleaf, dev.leaf_pos, dev.leaf_size, dev.leaf_node,
                       dev.ids, mp, b.cap, b.cand, b.cand_cnt, b.bound, b.work_keys, nwork, b.info, b.counters,
                       b.mstate, b.vtop, b.kmax, b.lab);
  }
  if (pf) pf->mark(st, HKG_STAGE_APPROX);
  if (mp.nprobes == 0) {
    uint32_t *keys2 = b.work_keys + b.work_cap, *vals2 = b.work_vals + b.work_cap;
    HKG_HIP_TRY(launch_sort_work(st, b.sort_tmp, b.sort_bytes, b.work_keys, keys2, b.work_vals, vals2, pairs));
    if (pf) pf->mark(st, HKG_STAGE_COLLECT);
    launch_filter_refine(st, filter_grid(), dev.X, dev.sax, L, b.Q, b.paa, b.kq, dev.leaf_pos, dev.leaf_size,
                         dev.ids, nleaf, b.lb, vals2, nwork, work_next, b.cap, b.cand, b.cand_cnt, b.bound,
                         b.counters);
    if (pf) pf->mark(st, HKG_STAGE_REFINE);
  }
  launch_select(st, nq, b.kmax, b.kq, b.cap, b.cand, b.cand_cnt, idmap64, b.out_ids, b.out_d2, overflow);
  if (pf) pf->mark(st, HKG_STAGE_TOPK);
  return HKG_OK;
}

// one pass over X per group of kScanQ queries
uint64_t Index::total_scan_bytes(uint32_t nq, uint32_t batch) const
{
  uint64_t passes = 0;
  for (uint32_t q0 = 0; q0 < nq; q0 += batch) {
    uint32_t n = std::min(batch, nq - q0);
    passes += (n + scan_group_size(n, dev.L) - 1) / scan_group_size(n, dev.L);
  }
  return passes * dev.n * dev.L * 4;
}

uint64_t Index::total_segs() const { return dev.tree.seg_end.size(); }

uint32_t scan_grid_x(uint32_t nq, uint32_t L, uint32_t blocks_per_cu)
{
  // a multiple of 8 (the XCD count): block x of every query group runs on
  // the same XCD and reads the same rows
  uint32_t groups = (nq + scan_group_size(nq, L) - 1) / scan_group_size(nq, L);
  uint32_t total = cu_count() * std::max<uint32_t>(blocks_per_cu, 1);
  uint32_t gx = total / std::max<uint32_t>(groups, 1);
  return std::max<uint32_t>(gx / 8 * 8, 64);
}

// persistent grid of the full scan for the running call: blocks per CU from
// the call, else the tuned shape, else kScanBlocksPerCU (any grid gives the
// same answers)
uint32_t Index::scan_grid(uint32_t nq) const
{
  uint32_t per = call_settings().sblocks;
  if (!per) per = tuned_sblocks.load(std::memory_order_relaxed);
  if (!per) per = kScanBlocksPerCU;
  return scan_grid_x(nq, dev.L, per);
}

void Index::enqueue_scan_batch(QueryBuffers &b, uint32_t nq, hipStream_t st, Prof *pf)
{
  uint32_t *overflow = b.misc + 2;
  launch_reset(st, nq, b.bound, b.cand_cnt, nullptr);
  launch_fill(st, b.misc, 8, 0);
  launch_check_k(st, nq, b.kmax, b.kq, b.misc + 3);
  launch_fill(st, reinterpret_cast<uint32_t *>(b.counters), 2 * kNumCounters, 0);
  if (pf) pf->mark(st, -1);
  launch_scan(st, scan_grid(nq), dev.X, dev.n, dev.L, b.Q, nq, b.kq, dev.ids, b.cap, b.cand, b.cand_cnt,
              b.bound);
  if (pf) pf->mark(st, HKG_STAGE_SCAN);
  launch_select(st, nq, b.kmax, b.kq, b.cap, b.cand, b.cand_cnt, idmap64, b.out_ids, b.out_d2, overflow);
  if (pf) pf->mark(st, HKG_STAGE_TOPK);
}

// capacity of the per-query candidate buffer for a path
uint32_t Index::cand_cap(hkg_path path, uint32_t nq, uint32_t kmax) const
{
  if (path == HKG_PATH_SCAN || !dev.has_tree) return scan_grid(nq) * kmax + 256;
  const uint32_t cap_override = call_settings().cap_override;
  if (cap_override) return std::max<uint32_t>(cap_override, kmax);
  // the batched iSAX filter keeps every series within the bound of the
  // approximate stage, without a per-query top-k
  if (path == HKG_PATH_BATCHED && leaf_ordered(nq)) return std::max<uint32_t>(32768, 16 * kmax);
  return std::max<uint32_t>(8192, 16 * kmax);
}

hkg_status Index::run_batch(hkg_path path, const float *dQ, const uint32_t *dk, uint32_t nq, uint32_t kmax,
                            bool use_graph, hipStream_t st, QueryBuffers **used, Counters *ctr, Prof *pf,
                            const uint32_t *seed_host)
{
  const bool scan = (path == HKG_PATH_SCAN) || !dev.has_tree;
  // sized for the buffer set's batch, not this call's: a server's batches of
  // different sizes then lease the same sets
  const uint32_t cap = scan ? cand_cap(HKG_PATH_SCAN, nq, kmax)
                            : cand_cap(path, std::max(nq, call_settings().min_buffer_batch), kmax);
  QueryBuffers *b = nullptr;
  hkg_status rc = get_buffers(nq, kmax, cap, &b);
  if (rc != HKG_OK) return rc;
  BufLease lease{this, b};  // handed to the caller on success
  HKG_HIP_TRY(hipMemcpyAsync(b->Q, dQ, (size_t)nq * dev.L * 4, hipMemcpyDeviceToDevice, st));
  HKG_HIP_TRY(hipMemcpyAsync(b->kq, dk, (size_t)nq * 4, hipMemcpyDeviceToDevice, st));
  if (seed_host) {
    HKG_HIP_TRY(hipMemcpyAsync(b->seed, seed_host, (size_t)nq * 4, hipMemcpyHostToDevice, st));
    HKG_HIP_TRY(hipStreamSynchronize(st));
    use_graph = false;
  }
  if (call_settings().mode && !scan) {
    rc = enqueue_mode_batch(*b, nq, st, pf, seed_host != nullptr);
    if (rc != HKG_OK) return rc;
  } else if (use_graph && !scan) {
    rc = run_captured(*b, 1, nq, 0, st, pf, [&](Prof *gp) {
      enqueue_index_batch(*b, nq, st, gp, false);
      return HKG_OK;
    });
    if (rc != HKG_OK) return rc;
  } else if (scan) {
    enqueue_scan_batch(*b, nq, st, pf);
  } else {
    enqueue_index_batch(*b, nq, st, pf, seed_host != nullptr);
  }
  HKG_HIP_TRY(hipGetLastError());
  if (ctr) {
    HKG_HIP_TRY(hipMemcpyAsync(ctr->v, b->counters, sizeof(ctr->v), hipMemcpyDeviceToHost, st));
    HKG_HIP_TRY(hipMemcpyAsync(&ctr->overflow, b->misc + 2, 4, hipMemcpyDeviceToHost, st));
    HKG_HIP_TRY(hipMemcpyAsync(&ctr->nwork, b->misc + 0, 4, hipMemcpyDeviceToHost, st));
    HKG_HIP_TRY(hipMemcpyAsync(&ctr->bad_k, b->misc + 3, 4, hipMemcpyDeviceToHost, st));
  }
  *used = b;
  lease.b = nullptr;
  return HKG_OK;
}

hkg_status Index::run_captured(QueryBuffers &b, uint32_t kind, uint32_t nq, uint32_t flags, hipStream_t st, Prof *pf,
                               const std::function<hkg_status(Prof *)> &enqueue)
{
  const bool prof = pf != nullptr;
  const uint32_t fgrid = filter_grid();
  if (!b.graph_exec || b.graph_stream != st || b.graph_kind != kind || b.graph_nq != nq || b.graph_flags != flags ||
      b.graph_fgrid != fgrid || b.graph_profiled != prof) {
    if (b.graph_exec) (void)hipGraphExecDestroy(b.graph_exec);
    b.graph_exec = nullptr;
    Prof *gp = nullptr;
    if (prof) {
      if (!b.graph_prof) {
        b.graph_prof.reset(new Prof());
        if (!b.graph_prof->init()) {
          b.graph_prof.reset();
          set_error("stage timing set-up failed (wall clock rate or stamp buffer)");
          return HKG_EDEVICE;
        }
      }
      gp = b.graph_prof.get();
      gp->n = 0;
    }
    HKG_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const hkg_status rc = enqueue(gp);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != HKG_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (rc != HKG_OK) return rc;
      HKG_HIP_TRY(e);
    }
    e = hipGraphInstantiate(&b.graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HKG_HIP_TRY(e);
    b.graph_stream = st;
    b.graph_kind = kind;
    b.graph_nq = nq;
    b.graph_flags = flags;
    b.graph_fgrid = fgrid;
    b.graph_profiled = prof;
  }
  HKG_HIP_TRY(hipGraphLaunch(b.graph_exec, st));
  if (pf) {
    pf->n = 0;
    pf->src = b.graph_prof.get();
  }
  return HKG_OK;
}

// Overflowed queries (their candidate buffer filled up) run again on the
// index path, seeded with the k-th distance of the candidates that were
// kept -- a valid upper bound, usually tight, so the re-run keeps few
// candidates.  A query still overflowing after a few rounds is answered by
// the full scan, whose buffer is sized so that it cannot overflow.
hkg_status Index::fix_overflow(QueryBuffers &b, uint32_t nq, uint32_t kmax, hipStream_t st, uint32_t *retries)
{
  const uint32_t L = dev.L;
  std::vector<uint32_t> cnt(nq), kv(nq);
  std::vector<float> d2((size_t)nq * kmax);
  HKG_HIP_TRY(hipMemcpyAsync(cnt.data(), b.cand_cnt, nq * 4, hipMemcpyDeviceToHost, st));
  HKG_HIP_TRY(hipMemcpyAsync(kv.data(), b.kq, nq * 4, hipMemcpyDeviceToHost, st));
  HKG_HIP_TRY(hipMemcpyAsync(d2.data(), b.out_d2, d2.size() * 4, hipMemcpyDeviceToHost, st));
  HKG_HIP_TRY(hipStreamSynchronize(st));
  std::vector<uint32_t> todo;
  for (uint32_t q = 0; q < nq; ++q)
    if (cnt[q] > b.cap) todo.push_back(q);
  if (todo.empty()) return HKG_OK;
  uint32_t m = (uint32_t)todo.size();
  float *tq = nullptr;
  uint32_t *tk = nullptr;
  hkg_status rc = dalloc(&tq, (uint64_t)m * L, nullptr);
  if (rc != HKG_OK) return rc;
  if ((rc = dalloc(&tk, m, nullptr)) != HKG_OK) {
    dfree(tq);
    return rc;
  }
  std::vector<uint32_t> seed(m);
  for (uint32_t i = 0; i < m; ++i) {
    uint32_t q = todo[i];
    (void)hipMemcpyAsync(tq + (uint64_t)i * L, b.Q + (uint64_t)q * L, L * 4, hipMemcpyDeviceToDevice, st);
    (void)hipMemcpyAsync(tk + i, b.kq + q, 4, hipMemcpyDeviceToDevice, st);
    float kth = d2[(size_t)q * kmax + kv[q] - 1];
    uint32_t bits;
    memcpy(&bits, &kth, 4);
    seed[i] = kth < FLT_MAX ? bits : kInfBits;
  }
  // an approximate mode answers over the rows its traversal examines, so
  // its last round is the same search with a buffer that cannot overflow
  // (every flush holds at most k entries: one per query plus one per
  // candidate leaf) rather than the full scan
  const bool mode = call_settings().mode;
  const uint32_t saved_cap = call_settings().cap_override;
  for (int round = 0; round < 4 && rc == HKG_OK; ++round) {
    ++*retries;
    QueryBuffers *sb = nullptr;
    Counters c;
    const bool last = round == 3;
    if (mode && last) call_settings().cap_override = kmax * (dev.nleaf + 2);
    rc = run_batch(last && !mode ? HKG_PATH_SCAN : HKG_PATH_BATCHED, tq, tk, m, kmax, false, st, &sb, &c, nullptr,
                   last ? nullptr : seed.data());
    call_settings().cap_override = saved_cap;
    if (rc != HKG_OK) break;
    BufLease lease{this, sb};
    std::vector<uint32_t> c2(m);
    std::vector<float> o2((size_t)m * kmax);
    if (hipMemcpyAsync(c2.data(), sb->cand_cnt, m * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(o2.data(), sb->out_d2, o2.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      set_error("overflow re-run readback failed");
      rc = HKG_EDEVICE;
      break;
    }
    bool again = false;
    for (uint32_t i = 0; i < m; ++i) {
      if (c2[i] > sb->cap) {
        again = true;
        float kth = o2[(size_t)i * kmax + kv[todo[i]] - 1];
        uint32_t bits;
        memcpy(&bits, &kth, 4);
        if (kth < FLT_MAX && bits < seed[i]) seed[i] = bits;
        continue;
      }
      (void)hipMemcpyAsync(b.out_ids + (uint64_t)todo[i] * kmax, sb->out_ids + (uint64_t)i * kmax, kmax * 8,
                           hipMemcpyDeviceToDevice, st);
      (void)hipMemcpyAsync(b.out_d2 + (uint64_t)todo[i] * kmax, sb->out_d2 + (uint64_t)i * kmax, kmax * 4,
                           hipMemcpyDeviceToDevice, st);
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
      rc = HKG_EDEVICE;
      break;
    }
    if (!again) break;
    // keep only the queries that overflowed again
    std::vector<uint32_t> keep;
    for (uint32_t i = 0; i < m; ++i)
      if (c2[i] > sb->cap) keep.push_back(i);
    for (uint32_t j = 0; j < keep.size(); ++j) {
      uint32_t i = keep[j];
      (void)hipMemcpyAsync(tq + (uint64_t)j * L, tq + (uint64_t)i * L, L * 4, hipMemcpyDeviceToDevice, st);
      (void)hipMemcpyAsync(tk + j, tk + i, 4, hipMemcpyDeviceToDevice, st);
      seed[j] = seed[i];
      todo[j] = todo[i];
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
      rc = HKG_EDEVICE;
      break;
    }
    m = (uint32_t)keep.size();
  }
  dfree(tq);
  dfree(tk);
  return rc;
}

static bool debug_enabled() { return getenv("HKG_DEBUG") != nullptr; }

// HKG_DEBUG: how far the filter's first threshold was from the final k-th
// distance, and how many survivors each query had
static void print_bf16_summary(const QueryBuffers *b, uint32_t n, uint32_t kmax, const std::vector<uint32_t> &cnt,
                               const std::vector<uint32_t> &thr_bits)
{
  std::vector<uint32_t> kv(n);
  std::vector<float> d2((size_t)n * kmax);
  if (hipMemcpy(kv.data(), b->kq, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(d2.data(), b->out_d2, d2.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return;
  std::vector<double> gap, ratio, surv;
  uint32_t dropped = 0;
  for (uint32_t q = 0; q < n; ++q) {
    float thr;
    memcpy(&thr, &thr_bits[q], 4);
    const float kth = d2[(size_t)q * kmax + kv[q] - 1];
    if (thr < FLT_MAX && kth < FLT_MAX) {
      gap.push_back((double)thr - kth);
      ratio.push_back(kth > 0 ? (double)thr / kth : 1.0);
    }
    if (cnt[q] == 0xFFFFFFFFu) ++dropped;
    else surv.push_back(cnt[q]);
  }
  auto qs = [](std::vector<double> v, const char *name) {
    if (v.empty()) return;
    std::sort(v.begin(), v.end());
    auto at = [&](double f) { return v[std::min<size_t>(v.size() - 1, (size_t)(f * v.size()))]; };
    fprintf(stderr, "hkg:   %-22s p10 %.4g p50 %.4g p90 %.4g max %.4g\n", name, at(0.1), at(0.5), at(0.9), v.back());
  };
  fprintf(stderr, "hkg: bf16 batch of %u: %u queries dropped survivors\n", n, dropped);
  qs(gap, "threshold - kth");
  qs(ratio, "threshold / kth");
  qs(surv, "survivors (kept lists)");
}

hkg_status Index::knn_device(const float *dQ, uint32_t nq, const uint32_t *dk, uint32_t kmax,
                             const hkg_query_opts *qo, uint64_t *dids, float *dd2, hipStream_t st,
                             hkg_query_stats *stats)
{
  hkg_query_opts o;
  hkg_default_query_opts(&o);
  if (qo) o = *qo;
  if (kmax == 0 || kmax > (uint32_t)kMaxK) {
    set_error("k must be in [1, 256]");
    return HKG_EINVAL;
  }
  if (nq == 0) return HKG_OK;
  if ((uint32_t)o.path > (uint32_t)HKG_PATH_BF16) {
    set_error("unknown query path " + std::to_string((uint32_t)o.path));
    return HKG_EINVAL;
  }
  if (o.batched_filter > 2) {
    set_error("batched_filter must be 0 (by size), 1 (per item) or 2 (leaf-ordered)");
    return HKG_EINVAL;
  }
  const bool mode = o.nprobes > 0 || o.epsilon != 0.0f;
  if (!(o.epsilon >= 0.0f) || std::isinf(o.epsilon) || std::isnan(o.eapca_threshold) ||
      std::isnan(o.sax_threshold)) {
    set_error("epsilon must be finite and >= 0; thresholds must be numbers");
    return HKG_EINVAL;
  }
  if (mode && !dev.has_tree) {
    set_error("the epsilon and nprobes modes need an index tree (hkg_open_dir or a built index)");
    return HKG_EINVAL;
  }
  hkg_path path = o.path;
  if (path == HKG_PATH_AUTO) path = HKG_PATH_BATCHED;
  if (mode) path = HKG_PATH_BATCHED;  // the modes run their own stage sequence
  if (path == HKG_PATH_BF16) {
    hkg_status rc = make_bf16();  // once per index
    if (rc != HKG_OK) return rc;
    if (o.bf16_query_tiles && !bf16_query_tiles_ok(dev.L, o.bf16_query_tiles)) {
      set_error("bf16_query_tiles must be 2, 3 or, for series of length 128 or less, 4");
      return HKG_EINVAL;
    }
  }
  if (!dev.has_tree && path != HKG_PATH_SCAN) path = HKG_PATH_SCAN;
  uint32_t batch = o.batch;
  if (batch == 0) batch = tuned_batch[path].load(std::memory_order_relaxed);
  if (batch == 0) batch = path == HKG_PATH_INDEX ? 1 : (path == HKG_PATH_SCAN ? 64 : 1024);
  if (path == HKG_PATH_INDEX) batch = 1;
  if (batch > 16384) batch = 16384;  // grid.y of the per-query launches
  // per-call settings; a nested call (the last-resort re-run of the BF16
  // path) restores its caller's on return
  struct Restore {
    CallSettings saved;
    ~Restore() { call_settings() = saved; }
  } restore{call_settings()};
  call_settings().cap_override = o.cand_cap;
  call_settings().bound_pass = o.bf16_bound_pass;
  call_settings().batched_filter = o.batched_filter;
  call_settings().bf16_qt = o.bf16_query_tiles ? o.bf16_query_tiles : bf16_qt();
  call_settings().mode = mode;
  call_settings().mp = ModeParams{o.epsilon, o.eapca_threshold, o.sax_threshold, o.approx_leaves, o.nprobes, 0, 0};
  HKG_HIP_TRY(hipSetDevice(dev.device));
  if (!st) st = stream;
  hipEvent_t e0, e1;
  HKG_HIP_TRY(hipEventCreate(&e0));
  HKG_HIP_TRY(hipEventCreate(&e1));
  HKG_HIP_TRY(hipEventRecord(e0, st));
  Counters total;
  uint32_t retries = 0, launches = 0;
  std::unique_ptr<Prof> pf;
  double stage_ms[HKG_NUM_STAGES] = {0};
  uint64_t stage_launches[HKG_NUM_STAGES] = {0};
  if (o.profile) {
    pf.reset(new Prof());
    if (!pf->init()) {
      set_error("stage timing set-up failed (wall clock rate or stamp buffer)");
      return HKG_EDEVICE;
    }
  }
  for (uint32_t q0 = 0; q0 < nq; q0 += batch) {
    uint32_t n = std::min(batch, nq - q0);
    QueryBuffers *b = nullptr;
    Counters c;
    hkg_status rc;
    if (pf) {
      pf->n = 0;
      pf->src = nullptr;
    }
    const bool isax = path == HKG_PATH_BATCHED && !mode && leaf_ordered(n);
    if (path == HKG_PATH_BF16)
      rc = run_bf16_batch(dQ + (uint64_t)q0 * dev.L, dk + q0, n, kmax, st, &b, &c, pf.get(), nullptr,
                          o.bf16_bound_pass == 1, n == batch);
    else if (isax)
      rc = run_isax_batch(dQ + (uint64_t)q0 * dev.L, dk + q0, n, kmax, st, &b, &c, pf.get(), nullptr, n == batch);
    else
      rc = run_batch(path, dQ + (uint64_t)q0 * dev.L, dk + q0, n, kmax, path == HKG_PATH_BATCHED && n == batch, st,
                     &b, &c, pf.get());
    if (rc != HKG_OK) return rc;
    BufLease lease{this, b};
    HKG_HIP_TRY(hipStreamSynchronize(st));
    if (c.bad_k) {
      set_error(std::to_string(c.bad_k) + " per-query k value(s) outside [1, kmax = " + std::to_string(kmax) + "]");
      return HKG_EINVAL;
    }
    ++launches;
    if (pf) pf->accumulate(stage_ms, stage_launches);
    std::vector<uint32_t> dbg_cnt, dbg_thr;
    if (path == HKG_PATH_BF16 && debug_enabled()) {
      dbg_cnt.resize(n);
      dbg_thr.resize(n);
      HKG_HIP_TRY(hipMemcpy(dbg_cnt.data(), b->bf_cnt, n * 4, hipMemcpyDeviceToHost));
      HKG_HIP_TRY(hipMemcpy(dbg_thr.data(), b->thr, n * 4, hipMemcpyDeviceToHost));
    }
    c.isax = isax;
    if (c.overflow) {
      // the batched iSAX path's candidate buffers overflow like the
      // per-item filter's, and re-run through it (fix_overflow)
      rc = path == HKG_PATH_BF16 ? fix_bf16_overflow(*b, n, kmax, st, &retries)
                                 : fix_overflow(*b, n, kmax, st, &retries);
      if (rc != HKG_OK) return rc;
    }
    if (!dbg_cnt.empty()) {
      HKG_HIP_TRY(hipStreamSynchronize(st));
      print_bf16_summary(b, n, kmax, dbg_cnt, dbg_thr);
      fprintf(stderr, "hkg:   approximate stage: %llu series in its leaves, %llu ranked\n",
              (unsigned long long)c.v[kCtrApprox], (unsigned long long)c.v[kCtrApproxRanked]);
    }
    HKG_HIP_TRY(hipMemcpyAsync(dids + (uint64_t)q0 * kmax, b->out_ids, (size_t)n * kmax * 8,
                               hipMemcpyDeviceToDevice, st));
    HKG_HIP_TRY(hipMemcpyAsync(dd2 + (uint64_t)q0 * kmax, b->out_d2, (size_t)n * kmax * 4,
                               hipMemcpyDeviceToDevice, st));
    HKG_HIP_TRY(hipStreamSynchronize(st));  // before the set is released
    for (int i = 0; i < kNumCounters; ++i) total.v[i] += c.v[i];
    total.nwork += c.nwork;
    total.isax |= c.isax;
    total.bf16_rows += c.bf16_rows;
    total.bf16_rows_bound += c.bf16_rows_bound;
  }
  HKG_HIP_TRY(hipEventRecord(e1, st));
  HKG_HIP_TRY(hipStreamSynchronize(st));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->queries = nq;
    const bool scan = path == HKG_PATH_SCAN;
    stats->leaves_total = (uint64_t)nq * dev.nleaf;
    // pairs within the k-th bound after the approximate stage, the leaves
    // that stage scanned included
    stats->leaves_after_eapca =
        scan ? 0
             : (path == HKG_PATH_BF16 ? total.v[kCtrLeafPairs] : total.nwork + total.v[kCtrApproxLeaves]);
    stats->series_total = (uint64_t)nq * dev.n;
    // BF16: the pairs of the row tiles the filter walked, less those of the
    // tiles a wave skipped (no query of it needs them)
    const uint64_t bf16_seen = total.v[kCtrBf16] - std::min(total.v[kCtrBf16], total.v[kCtrBf16Skip]);
    stats->series_lb_checked = scan ? 0 : total.v[kCtrSaxChecked] + bf16_seen;
    stats->bf16_pairs = bf16_seen;
    stats->series_refined = scan ? (uint64_t)nq * dev.n : total.v[kCtrRefined];
    stats->device_ms = ms;
    stats->overflow_retries = retries;
    stats->launches = launches;
    for (int i = 0; i < HKG_NUM_STAGES; ++i) {
      stats->stage_ms[i] = stage_ms[i];
      stats->stage_launches[i] = stage_launches[i];
    }
    const uint64_t L4 = (uint64_t)dev.L * 4;
    stats->stage_bytes[HKG_STAGE_APPROX] = total.v[kCtrApprox] * L4;
    // the leaf-ordered iSAX filter reads each candidate leaf's words once per
    // wave chunk (up to 64 queries), and refines its survivors (at most 4 L
    // bytes each: they are abandoned early)
    stats->stage_bytes[HKG_STAGE_REFINE] =
        path == HKG_PATH_BF16 ? total.bf16_rows * ((uint64_t)dev.L * 2 + 16)
        : total.isax          ? total.v[kCtrSaxWords] * 16
                              : total.v[kCtrSaxChecked] * 16 + total.v[kCtrRefined] * L4;
    if (total.isax) stats->stage_bytes[HKG_STAGE_REFINE] += total.v[kCtrRefined] * L4;
    if (path == HKG_PATH_BF16) {
      stats->stage_bytes[HKG_STAGE_TOPK] = total.v[kCtrRefined] * L4;
      stats->stage_bytes[HKG_STAGE_BOUND] = total.bf16_rows_bound * ((uint64_t)dev.L * 2 + 16);
    }
    stats->stage_bytes[HKG_STAGE_SCAN] = scan ? (uint64_t)launches * 0 + total_scan_bytes(nq, batch) : 0;
    stats->stage_bytes[HKG_STAGE_EAPCA] = scan ? 0 : (uint64_t)nq * dev.nleaf * 8 + total_segs() * 18ull * nq;
    stats->mode_leaves_first = total.v[kCtrModeFirst];
    stats->mode_candidates = total.v[kCtrModeCand];
    stats->mode_candidates_scanned = total.v[kCtrModeScanned];
    stats->mode_sequential = total.v[kCtrModeSeq];
    stats->bytes_algorithmic = 0;
    for (int i = 0; i < HKG_NUM_STAGES; ++i) stats->bytes_algorithmic += stats->stage_bytes[i];
  }
  return HKG_OK;
}

hkg_status Index::knn_host(const float *Q, uint32_t nq, const uint32_t *k, uint32_t kmax, const hkg_query_opts *qo,
                           uint64_t *ids, float *d2, hkg_query_stats *stats)
{
  if (!Q || !ids || !d2) {
    set_error("null argument");
    return HKG_EINVAL;
  }
  if (nq == 0) return HKG_OK;
  HKG_HIP_TRY(hipSetDevice(dev.device));
  std::vector<uint32_t> kv(nq, kmax);
  if (k) {
    for (uint32_t i = 0; i < nq; ++i) {
      if (k[i] == 0 || k[i] > kmax) {
        set_error("per-query k must be in [1, kmax]");
        return HKG_EINVAL;
      }
      kv[i] = k[i];
    }
  }
  float *dQ = nullptr, *dd2 = nullptr;
  uint32_t *dk = nullptr;
  uint64_t *dids = nullptr;
  hkg_status rc;
  if ((rc = dalloc(&dQ, (uint64_t)nq * dev.L, nullptr)) != HKG_OK) return rc;
  if ((rc = dalloc(&dk, nq, nullptr)) != HKG_OK) {
    dfree(dQ);
    return rc;
  }
  if ((rc = dalloc(&dids, (uint64_t)nq * kmax, nullptr)) != HKG_OK || (rc = dalloc(&dd2, (uint64_t)nq * kmax, nullptr)) != HKG_OK) {
    dfree(dQ);
    dfree(dk);
    dfree(dids);
    return rc;
  }
  hipError_t e1 = hipMemcpy(dQ, Q, (size_t)nq * dev.L * 4, hipMemcpyHostToDevice);
  hipError_t e2 = hipMemcpy(dk, kv.data(), (size_t)nq * 4, hipMemcpyHostToDevice);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    set_error("query upload failed");
    rc = HKG_EDEVICE;
  } else {
    rc = knn_device(dQ, nq, dk, kmax, qo, dids, dd2, stream, stats);
  }
  if (rc == HKG_OK) {
    if (hipMemcpy(ids, dids, (size_t)nq * kmax * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(d2, dd2, (size_t)nq * kmax * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      set_error("result download failed");
      rc = HKG_EDEVICE;
    }
  }
  dfree(dQ);
  dfree(dk);
  dfree(dids);
  dfree(dd2);
  return rc;
}

// Launch-shape tuner: each candidate (queries per batch x filter grid, or
// on the BF16 path x query tiles per filter wave) runs the whole batches of
// the sample three times and keeps its fastest device time per query; the default
// shape runs first, after one untimed warm-up call (BF16 copy, buffer sets,
// code objects).  Among shapes within kTuneMargin of the fastest, the one
// with the largest batch is kept: repeated runs of one shape spread by about
// that much, and a larger batch spreads its per-batch work over more
// queries.  Answers must not depend on the shape: any difference is an
// engine fault, reported instead of tuning.
constexpr double kTuneMargin = 0.01;
hkg_status Index::autotune(const float *Q, uint32_t nq, const uint32_t *k, uint32_t kmax, uint32_t path_in,
                           hkg_tuning *out)
{
  if (!Q || !out || nq == 0 || kmax == 0 || path_in > HKG_PATH_BF16) {
    set_error("autotune: null argument, no queries, kmax 0 or unknown path");
    return HKG_EINVAL;
  }
  hkg_path path = path_in == HKG_PATH_AUTO ? HKG_PATH_BATCHED : (hkg_path)path_in;
  if (!dev.has_tree) path = HKG_PATH_SCAN;
  // a shape: queries per batch, the main kernel's grid (blocks per CU) and,
  // on the batched path, its filter (1 per item, 2 leaf-ordered; the grid
  // only shapes the per-item filter); on the BF16 path g is the filter's
  // query tiles per wave, with batches of whole block groups of that shape
  struct Shape {
    uint32_t g, f;
  };
  std::vector<uint32_t> batches;
  std::vector<Shape> shapes;
  std::vector<std::pair<uint32_t, Shape>> extra;  // (batch, shape) pairs beyond batches x shapes
  uint32_t def_batch = 1024, def_grid = 0;
  switch (path) {
    case HKG_PATH_INDEX:
      batches = {1};
      shapes = {{2, 0}, {4, 0}, {8, 0}};
      def_batch = 1;
      def_grid = kFilterBlocksPerCU;
      break;
    case HKG_PATH_BATCHED:
      batches = {256, 512, 1024, 2048};
      shapes = {{2, 1}, {4, 1}, {8, 1}, {0, 2}};
      def_grid = kFilterBlocksPerCU;
      break;
    case HKG_PATH_BF16:
      batches = {512, 1024, 2048, 4096};
      shapes = {{kBf16QueryTiles, 0}};
      def_grid = kBf16QueryTiles;
      // three tiles per wave: 384 queries per block group
      for (uint32_t b : {768u, 1536u, 3072u}) extra.push_back({b, {3, 0}});
      if (bf16_query_tiles_ok(dev.L, 4))
        for (uint32_t b : {1024u, 2048u, 4096u}) extra.push_back({b, {4, 0}});
      break;
    default:
      batches = {1, 4, 16, 64, 256};
      shapes = {{3, 0}, {6, 0}, {8, 0}};  // full-scan blocks per CU
      def_batch = 64;
      def_grid = kScanBlocksPerCU;
      break;
  }
  // batches beyond the sample time the same as the sample-sized one
  std::vector<uint32_t> bs;
  for (uint32_t b : batches)
    if (b <= nq || bs.empty()) bs.push_back(b);
  std::vector<std::pair<uint32_t, Shape>> cands;
  for (uint32_t b : bs)
    for (const Shape &sh : shapes) cands.push_back({b, sh});
  for (const auto &c : extra)
    if (c.first <= nq) cands.push_back(c);
  struct Restore {
    CallSettings saved;
    ~Restore() { call_settings() = saved; }
  } restore{call_settings()};
  const bool scan_path = path == HKG_PATH_SCAN;
  const bool batched = path == HKG_PATH_BATCHED;
  const bool bf16 = path == HKG_PATH_BF16;
  // the default shape is the untuned one
  tuned_batch[path].store(0, std::memory_order_relaxed);
  if (batched) tuned_filter.store(0, std::memory_order_relaxed);
  if (bf16) tuned_qt.store(0, std::memory_order_relaxed);
  hkg_query_opts qo;
  hkg_default_query_opts(&qo);
  qo.path = path;
  const size_t m = (size_t)nq * kmax;
  std::vector<uint64_t> ids0(m), ids(m);
  std::vector<float> d20(m), d2(m);
  // a shape is timed on the whole batches the sample holds (a partial last
  // batch would charge the shape for a launch it would not make at scale):
  // device time per query over the first used(batch) queries
  auto used = [&](uint32_t batch) { return batch <= nq ? nq / batch * batch : nq; };
  auto run = [&](uint32_t batch, uint32_t grid, uint32_t filt, uint32_t n, uint64_t *I, float *D,
                 double *ms) -> hkg_status {
    qo.batch = batch;
    qo.batched_filter = filt;
    if (bf16)
      qo.bf16_query_tiles = grid;
    else if (scan_path)
      call_settings().sblocks = grid;
    else
      call_settings().fblocks = grid;
    *ms = 1e300;
    for (int rep = 0; rep < 3; ++rep) {
      hkg_query_stats st;
      hkg_status rc = knn_host(Q, n, k, kmax, &qo, I, D, &st);
      if (rc != HKG_OK) return rc;
      *ms = std::min(*ms, st.device_ms / n);
    }
    return HKG_OK;
  };
  // the default shape: default batch and grid, the filter its size picks
  const uint32_t def_filt = batched ? (isax_batched(def_batch, dev.nleaf) ? 2u : 1u) : 0u;
  double ms_def = 0.0, ms;
  // warm-up and the answers of every sample query, then the default's time
  hkg_status rc = run(def_batch, def_grid, def_filt, nq, ids0.data(), d20.data(), &ms);
  if (rc != HKG_OK) return rc;
  if ((rc = run(def_batch, def_grid, def_filt, used(def_batch), ids.data(), d2.data(), &ms_def)) != HKG_OK) return rc;
  struct Timed {
    uint32_t b, g, f;
    double ms;
  };
  std::vector<Timed> timed{{def_batch, def_filt == 2 ? 0u : def_grid, def_filt, ms_def}};
  uint32_t ncand = 1;
  for (const auto &c : cands) {
    const uint32_t b = c.first;
    const Shape &sh = c.second;
    if (b == def_batch && sh.f == def_filt && (sh.f == 2 || sh.g == def_grid)) continue;
    const uint32_t n = used(b);
    if ((rc = run(b, sh.g, sh.f, n, ids.data(), d2.data(), &ms)) != HKG_OK) return rc;
    ++ncand;
    const size_t mu = (size_t)n * kmax;
    if (memcmp(ids.data(), ids0.data(), mu * 8) != 0 || memcmp(d2.data(), d20.data(), mu * 4) != 0) {
      set_error("autotune: a launch shape changed the answers (batch " + std::to_string(b) + ", grid " +
                std::to_string(sh.g) + (bf16 ? " query tiles per wave" : " blocks per CU") + ", filter " +
                std::to_string(sh.f) + ")");
      return HKG_EINTERNAL;
    }
    timed.push_back({b, sh.g, sh.f, ms});
  }
  double fastest = ms_def;
  for (const Timed &t : timed) fastest = std::min(fastest, t.ms);
  // the largest batch within the margin of the fastest, then its fastest shape
  const Timed *best = nullptr;
  for (const Timed &t : timed) {
    if (t.ms > fastest * (1.0 + kTuneMargin)) continue;
    if (!best || t.b > best->b || (t.b == best->b && t.ms < best->ms)) best = &t;
  }
  tuned_batch[path].store(best->b, std::memory_order_relaxed);
  if (bf16)
    tuned_qt.store(best->g, std::memory_order_relaxed);
  else if (best->g)
    (scan_path ? tuned_sblocks : tuned_fblocks).store(best->g, std::memory_order_relaxed);
  if (batched) tuned_filter.store(best->f, std::memory_order_relaxed);
  memset(out, 0, sizeof(*out));
  out->path = path;
  out->length = dev.L;
  out->batch = best->b;
  out->filter_blocks_per_cu = scan_path || bf16 ? 0 : best->g;
  out->scan_blocks_per_cu = scan_path ? best->g : 0;
  out->batched_filter = best->f;
  out->bf16_query_tiles = bf16 ? best->g : 0;
  out->candidates = ncand;
  out->ms_per_query = best->ms;
  out->ms_per_query_default = ms_def;
  return HKG_OK;
}

}  // namespace hkg

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using hkg::Index;

struct hkg_index {
  Index impl;
};

extern "C" {

void hkg_default_options(hkg_options *opt)
{
  if (!opt) return;
  memset(opt, 0, sizeof(*opt));
}

void hkg_default_query_opts(hkg_query_opts *q)
{
  if (!q) return;
  memset(q, 0, sizeof(*q));
  q->path = HKG_PATH_AUTO;
  q->delta = 1.0f;             // hercules.c:49
  q->approx_leaves = 80;       // hercules.c:70
  q->eapca_threshold = 0.25f;  // hercules.c:71
  q->sax_threshold = 0.9f;     // hercules.c:72
}

const char *hkg_status_str(hkg_status s)
{
  switch (s) {
  case HKG_OK: return "ok";
  case HKG_EINVAL: return "invalid argument";
  case HKG_ENOMEM: return "out of memory";
  case HKG_EIO: return "i/o error";
  case HKG_EFORMAT: return "malformed index";
  case HKG_EDEVICE: return "device error";
  case HKG_EBUSY: return "busy";
  case HKG_ENOTREADY: return "not ready";
  case HKG_EINTERNAL: return "internal check failed";
  }
  return "unknown status";
}

const char *hkg_last_error(void) { return hkg::last_error().c_str(); }

void hkg_breakpoints256(float *out255) { hkg::host_breakpoints(out255); }

hkg_status hkg_open_dir(const char *index_dir, const hkg_options *opt, hkg_index **out)
{
  if (!out) return HKG_EINVAL;
  Index *ix = nullptr;
  hkg_status rc = hkg::open_dir(index_dir, opt, &ix);
  *out = reinterpret_cast<hkg_index *>(ix);
  return rc;
}

hkg_status hkg_open_raw(const float *X, uint64_t n, uint32_t L, const hkg_options *opt, hkg_index **out)
{
  if (!out) return HKG_EINVAL;
  Index *ix = nullptr;
  hkg_status rc = hkg::open_raw(X, n, L, opt, &ix);
  *out = reinterpret_cast<hkg_index *>(ix);
  return rc;
}

hkg_status hkg_set_ids(hkg_index *ix, const uint64_t *ids, uint64_t n)
{
  if (!ix || !ids) return HKG_EINVAL;
  return reinterpret_cast<Index *>(ix)->set_ids(ids, n);
}

hkg_status hkg_inspect_dir(const char *index_dir, hkg_index_info *info)
{
  if (!index_dir || !info) return HKG_EINVAL;
  std::string d(index_dir);
  if (!d.empty() && d.back() != '/') d += '/';
  hkg::HostTree t;
  std::string err;
  hkg_status rc = hkg::parse_root_idx(d + "root.idx", t, err);
  if (rc != HKG_OK) {
    hkg::set_error(d + "root.idx: " + err);
    return rc;
  }
  struct stat sr, ss;
  if (stat((d + "leaves_raw.idx").c_str(), &sr) != 0 || stat((d + "leaves_sims.idx").c_str(), &ss) != 0) {
    hkg::set_error(d + ": leaves_raw.idx or leaves_sims.idx missing");
    return HKG_EIO;
  }
  if ((uint64_t)sr.st_size != t.n_series * t.L * 4 ||
      (uint64_t)ss.st_size != t.n_series * (uint64_t)t.paa_segments) {
    hkg::set_error(d + ": leaf file sizes do not match root.idx");
    return HKG_EFORMAT;
  }
  memset(info, 0, sizeof(*info));
  info->n_series = t.n_series;
  info->length = t.L;
  info->paa_segments = (uint32_t)t.paa_segments;
  info->sax_bit_cardinality = (uint32_t)t.sax_bit_cardinality;
  info->max_leaf_size = t.max_leaf_size;
  info->n_leaves = t.leaves.size();
  info->n_nodes = t.nodes.size();
  info->has_tree = 1;
  return HKG_OK;
}

hkg_status hkg_info(const hkg_index *ix, hkg_index_info *info)
{
  if (!ix || !info) return HKG_EINVAL;
  const Index *p = reinterpret_cast<const Index *>(ix);
  memset(info, 0, sizeof(*info));
  info->n_series = p->dev.n;
  info->length = p->dev.L;
  info->paa_segments = hkg::kPaaSegments;
  info->sax_bit_cardinality = 8;
  info->max_leaf_size = p->dev.tree.max_leaf_size;
  info->n_leaves = p->dev.nleaf;
  info->n_nodes = p->dev.tree.nodes.size();
  info->device_bytes = p->dev.device_bytes;
  info->has_tree = p->dev.has_tree;
  info->has_ids = p->dev.ids != nullptr;
  return HKG_OK;
}

void hkg_close(hkg_index *ix)
{
  delete reinterpret_cast<Index *>(ix);
}

hkg_status hkg_knn(hkg_index *ix, const float *Q, uint32_t nq, const uint32_t *k, uint32_t kmax,
                   const hkg_query_opts *qo, uint64_t *ids, float *d2, hkg_query_stats *stats)
{
  if (!ix) return HKG_EINVAL;
  return reinterpret_cast<Index *>(ix)->knn_host(Q, nq, k, kmax, qo, ids, d2, stats);
}

hkg_status hkg_autotune(hkg_index *ix, const float *Q, uint32_t nq, const uint32_t *k, uint32_t kmax,
                        uint32_t path, hkg_tuning *out)
{
  if (!ix) return HKG_EINVAL;
  return reinterpret_cast<Index *>(ix)->autotune(Q, nq, k, kmax, path, out);
}

hkg_status hkg_knn_device(hkg_index *ix, const float *dQ, uint32_t nq, const uint32_t *dk, uint32_t kmax,
                          const hkg_query_opts *qo, uint64_t *dids, float *dd2, void *stream,
                          hkg_query_stats *stats)
{
  if (!ix || !dQ || !dk || !dids || !dd2) return HKG_EINVAL;
  return reinterpret_cast<Index *>(ix)->knn_device(dQ, nq, dk, kmax, qo, dids, dd2, (hipStream_t)stream, stats);
}

}  // extern "C"
