// Index files <-> device-resident SoA.
//
// root.idx format (written by hercules_index_write, read by
// hercules_index_read / hercules_node_read,
// code/hercules/src/hercules_index.c:1694-1905, 2079-2340):
//   header: u64 leaf_count, f64 buffered_memory_size, u32 timeseries_size,
//           u32 init_segments, i32 paa_segments, i32 sax_bit_cardinality,
//           u32 max_leaf_size, i32 serial, i32 sims                 (44 B)
//   node:   u8 is_leaf, u32 level, then
//     internal: <left> <right> u32 node_size, i16 split_segment,
//               48 B struct node_split_policy, i16 S, i16 seg_end[S],
//               S x {i32 n, f32 ind[n]}
//     leaf:     u32 node_size, i32 fname_len, and if fname_len > 0:
//               char fname[fname_len], i16 S, i16 seg_end[S],
//               S x {i32 n, f32 ind[n]}, u64 file_pos (serial indexes)
// leaves_raw.idx is f32[N][L] and leaves_sims.idx u8[N][paa_segments],
// both in left-first DFS leaf order (hercules_file_buffer.c:631-651).
#include <errno.h>
#include <stdio.h>
#include <string.h>

#include <memory>

#include "hkg_internal.h"

namespace hkg {

namespace {

struct Reader {
  const uint8_t *p;
  size_t n, off = 0;
  bool ok = true;
  template <class T>
  T get()
  {
    T v{};
    if (off + sizeof(T) > n) {
      ok = false;
      return v;
    }
    memcpy(&v, p + off, sizeof(T));
    off += sizeof(T);
    return v;
  }
  bool bytes(void *dst, size_t len)
  {
    if (off + len > n) {
      ok = false;
      return false;
    }
    memcpy(dst, p + off, len);
    off += len;
    return true;
  }
};

// the 48-byte node_split_policy (hercules_node_split.h:22-36, x86-64 layout)
constexpr size_t kPolicyBytes = 48;

bool read_sketches(Reader &r, HostTree &t, HostNode &nd, std::string &err)
{
  int16_t S = r.get<int16_t>();
  if (!r.ok || S <= 0 || (uint32_t)S > t.L) {
    err = "bad segment count";
    return false;
  }
  nd.nseg = (uint16_t)S;
  nd.seg_off = (uint32_t)t.seg_end.size();
  int prev = 0;
  for (int i = 0; i < S; ++i) {
    int16_t e = r.get<int16_t>();
    if (!r.ok || e <= prev || (uint32_t)e > t.L) {
      err = "bad segment end";
      return false;
    }
    prev = e;
    t.seg_end.push_back(e);
  }
  if ((uint32_t)prev != t.L) {
    err = "segments do not cover the series";
    return false;
  }
  for (int i = 0; i < S; ++i) {
    int32_t n = r.get<int32_t>();
    if (!r.ok || n != 4) {
      err = "bad indicator count";
      return false;
    }
    for (int j = 0; j < 4; ++j) t.ind.push_back(r.get<float>());
    if (!r.ok) {
      err = "truncated sketch";
      return false;
    }
  }
  return true;
}

int32_t read_node(Reader &r, HostTree &t, int depth, std::string &err)
{
  if (depth > 4096) {
    err = "tree too deep";
    return -1;
  }
  HostNode nd;
  nd.is_leaf = r.get<uint8_t>();
  nd.level = r.get<uint32_t>();
  if (!r.ok) {
    err = "truncated node";
    return -1;
  }
  int32_t me = (int32_t)t.nodes.size();
  t.nodes.push_back(nd);
  if (!nd.is_leaf) {
    int32_t l = read_node(r, t, depth + 1, err);
    if (l < 0) return -1;
    int32_t rr = read_node(r, t, depth + 1, err);
    if (rr < 0) return -1;
    HostNode &m = t.nodes[me];
    m.left = l;
    m.right = rr;
    m.size = r.get<uint32_t>();
    m.split_segment = r.get<int16_t>();
    uint8_t pol[kPolicyBytes];
    if (!r.bytes(pol, kPolicyBytes)) {
      err = "truncated split policy";
      return -1;
    }
    memcpy(&m.split_from, pol + 0, 2);
    memcpy(&m.split_to, pol + 2, 2);
    memcpy(&m.indicator_split_idx, pol + 12, 4);
    memcpy(&m.indicator_split_value, pol + 16, 4);
    HostNode tmp = m;
    if (!read_sketches(r, t, tmp, err)) return -1;
    t.nodes[me] = tmp;
    return me;
  }
  HostNode &m = t.nodes[me];
  m.size = r.get<uint32_t>();
  int32_t flen = r.get<int32_t>();
  if (!r.ok || flen < 0 || flen > 1 << 20) {
    err = "bad leaf name length";
    return -1;
  }
  if (flen == 0) {  // empty leaf (hercules_index.c:2331-2335)
    m.size = 0;
    return me;
  }
  std::string name((size_t)flen, '\0');
  if (!r.bytes(&name[0], (size_t)flen)) {
    err = "truncated leaf name";
    return -1;
  }
  HostNode tmp = m;
  tmp.fname = name;
  if (!read_sketches(r, t, tmp, err)) return -1;
  if (t.serial) {
    tmp.file_pos = r.get<uint64_t>();
    if (!r.ok) {
      err = "truncated file position";
      return -1;
    }
  }
  tmp.leaf_idx = (int32_t)t.leaves.size();
  t.leaves.push_back(me);
  t.nodes[me] = tmp;
  return me;
}

bool read_file(const std::string &path, std::vector<uint8_t> &buf, std::string &err)
{
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) {
    err = path + ": " + strerror(errno);
    return false;
  }
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  buf.resize(sz > 0 ? (size_t)sz : 0);
  size_t got = sz > 0 ? fread(buf.data(), 1, (size_t)sz, f) : 0;
  fclose(f);
  if ((long)got != sz) {
    err = path + ": short read";
    return false;
  }
  return true;
}

}  // namespace

hkg_status parse_root_idx(const std::string &path, HostTree &t, std::string &err)
{
  std::vector<uint8_t> buf;
  if (!read_file(path, buf, err)) return HKG_EIO;
  Reader r{buf.data(), buf.size()};
  t = HostTree();
  t.leaf_count_hdr = r.get<uint64_t>();
  t.buffered_memory_size = r.get<double>();
  t.L = r.get<uint32_t>();
  t.init_segments = r.get<uint32_t>();
  t.paa_segments = r.get<int32_t>();
  t.sax_bit_cardinality = r.get<int32_t>();
  t.max_leaf_size = r.get<uint32_t>();
  t.serial = r.get<int32_t>();
  t.sims = r.get<int32_t>();
  if (!r.ok) {
    err = "truncated header";
    return HKG_EFORMAT;
  }
  if (t.L == 0 || t.L % 8 != 0 || t.L > 32767) {
    err = "series length must be a positive multiple of 8";
    return HKG_EFORMAT;
  }
  if (!t.serial) {
    err = "index was built without --serial 1 (no leaf file positions)";
    return HKG_EFORMAT;
  }
  if (read_node(r, t, 0, err) < 0) return HKG_EFORMAT;
  if (r.off != buf.size()) {
    err = "trailing bytes after the tree";
    return HKG_EFORMAT;
  }
  // leaves must tile [0, N) in DFS order
  uint64_t pos = 0;
  for (int32_t li : t.leaves) {
    const HostNode &nd = t.nodes[li];
    if (nd.file_pos != pos) {
      err = "leaf file positions are not contiguous";
      return HKG_EFORMAT;
    }
    pos += nd.size;
  }
  t.n_series = pos;
  return HKG_OK;
}

namespace {
void put_bytes(std::vector<uint8_t> &o, const void *p, size_t n)
{
  const uint8_t *b = (const uint8_t *)p;
  o.insert(o.end(), b, b + n);
}
template <class T>
void put(std::vector<uint8_t> &o, T v)
{
  put_bytes(o, &v, sizeof(T));
}
void put_sketches(std::vector<uint8_t> &o, const HostTree &t, const HostNode &nd)
{
  put<int16_t>(o, (int16_t)nd.nseg);
  for (uint32_t i = 0; i < nd.nseg; ++i) put<int16_t>(o, t.seg_end[nd.seg_off + i]);
  for (uint32_t i = 0; i < nd.nseg; ++i) {
    put<int32_t>(o, 4);
    for (int j = 0; j < 4; ++j) put<float>(o, t.ind[4 * (nd.seg_off + i) + j]);
  }
}
void put_node(std::vector<uint8_t> &o, const HostTree &t, int32_t i)
{
  const HostNode &nd = t.nodes[i];
  put<uint8_t>(o, nd.is_leaf);
  put<uint32_t>(o, nd.level);
  if (!nd.is_leaf) {
    put_node(o, t, nd.left);
    put_node(o, t, nd.right);
    put<uint32_t>(o, nd.size);
    put<int16_t>(o, nd.split_segment);
    uint8_t pol[kPolicyBytes];
    memset(pol, 0, sizeof(pol));
    memcpy(pol + 0, &nd.split_from, 2);
    memcpy(pol + 2, &nd.split_to, 2);
    memcpy(pol + 4, &nd.indicator_split_value, 4);
    memcpy(pol + 8, &nd.indicator_split_idx, 4);
    memcpy(pol + 12, &nd.indicator_split_idx, 4);
    memcpy(pol + 16, &nd.indicator_split_value, 4);
    int32_t four = 4;
    memcpy(pol + 32, &four, 4);
    put_bytes(o, pol, sizeof(pol));
    put_sketches(o, t, nd);
    return;
  }
  put<uint32_t>(o, nd.size);
  if (nd.size == 0) {
    put<int32_t>(o, 0);
    return;
  }
  put<int32_t>(o, (int32_t)nd.fname.size());
  put_bytes(o, nd.fname.data(), nd.fname.size());
  put_sketches(o, t, nd);
  put<uint64_t>(o, nd.file_pos);
}
}  // namespace

hkg_status write_root_idx(const std::string &path, const HostTree &t, std::string &err)
{
  std::vector<uint8_t> o;
  put<uint64_t>(o, (uint64_t)t.leaves.size());
  put<double>(o, t.buffered_memory_size);
  put<uint32_t>(o, t.L);
  put<uint32_t>(o, t.init_segments);
  put<int32_t>(o, t.paa_segments);
  put<int32_t>(o, t.sax_bit_cardinality);
  put<uint32_t>(o, t.max_leaf_size);
  put<int32_t>(o, 1);
  put<int32_t>(o, 1);
  if (!t.nodes.empty()) put_node(o, t, 0);
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) {
    err = path + ": " + strerror(errno);
    return HKG_EIO;
  }
  size_t w = fwrite(o.data(), 1, o.size(), f);
  if (fclose(f) != 0 || w != o.size()) {
    err = path + ": write failed";
    return HKG_EIO;
  }
  return HKG_OK;
}

}  // namespace hkg
