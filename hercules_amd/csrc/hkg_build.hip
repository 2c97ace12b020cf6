// GPU bulk loader of an EAPCA tree, written in the reference's on-disk
// format.
//
// The reference builds by inserting series one at a time and splitting a
// leaf when it overflows (hercules_index_build, hercules_index.c:2354-2929;
// split policies hercules_node_split.c:883-1041, 1043-1346).  Here the tree
// is built top-down, level by level, over all series at once:
//   1. stats: every series of a frontier node contributes its per-segment
//      mean and sd to the node's sketch ({max mean, min mean, max sd,
//      min sd}) over the node's segmentation and over the refined
//      segmentation that splits every segment of length >= 2 at its middle
//      (node_init_segments with min_length 1, hercules_node.c:141-168);
//   2. policy (host): the reference's split choice -- for each segment and
//      each policy (mean, then sd) the benefit QoS(seg) - mean(QoS(children))
//      with QoS = len * (mean_width^2 + max_sd^2) (range_calc), children
//      halving the indicator range; a refined segment must beat twice the
//      best plain one; a node with at most leaf_size series is a leaf;
//   3. route: a series goes left when its statistic on the split segment is
//      below the split value (node_split_policy_route_to_left);
//   4. partition: a deterministic stable partition of the permutation (an
//      exclusive scan of the left flags) keeps every node's series
//      contiguous, left child first, so leaves come out in left-first DFS
//      order -- the order of leaves_raw.idx.
// Member statistics are accumulated in FP64 and the sketches rounded
// outward to FP32, so every member's exact statistics lie inside its
// node's intervals.  Finally rows are gathered into leaf order and each
// row's iSAX word is computed with the reference's PAA and symbol rules
// (sax.c:86-135).
#include <float.h>
#include <math.h>
#include <string.h>

#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <memory>
#include <vector>

#include "hkg_engine.h"
#include "hkg_kernels.h"

namespace hkg {

namespace {

constexpr int kMaxFine = 512;    // refined segments per node (LDS sketch)
constexpr uint32_t kNone = 0xFFFFFFFFu;

__host__ __device__ inline uint32_t fenc(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float fdec(uint32_t e)
{
  uint32_t u = (e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
__device__ inline float round_up(double v)
{
  float f = (float)v;
  if ((double)f < v) f = nextafterf(f, FLT_MAX);
  return f;
}
__device__ inline float round_down(double v)
{
  float f = (float)v;
  if ((double)f > v) f = nextafterf(f, -FLT_MAX);
  return f;
}

// per-level frontier tables (device)
struct Level {
  uint32_t nf;
  const uint32_t *fine_off;    // nf+1
  const int16_t *fine_end;     // refined segment ends
  const uint16_t *fine_coarse; // coarse segment (node-local) of each refined segment
  const uint32_t *coarse_off;  // nf+1
  uint32_t *fine_sk;           // 4 per refined segment (encoded)
  uint32_t *coarse_sk;         // 4 per coarse segment (encoded)
};

// works on LDS and global sketches alike
__device__ inline void sk_update(uint32_t *sk, float mu_hi, float mu_lo, float sd_hi, float sd_lo)
{
  atomicMax(&sk[0], fenc(mu_hi));
  atomicMin(&sk[1], fenc(mu_lo));
  atomicMax(&sk[2], fenc(sd_hi));
  atomicMin(&sk[3], fenc(sd_lo));
}

__device__ inline float tree8(const float *v)
{
  return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

// a segment's mean and sd exactly as the reference computes them in FP32
// (calc_mean_stdev_SIMD, calc_utils.c:48-112): 8-chunks from the segment
// start, the last zero-padded; sd is NaN when the FP32 variance goes
// negative, and such a value is ignored like fmaxf/fminf ignore it
__device__ inline void ref_seg_stats(const float *x, int s, int e, float &mu, float &sd)
{
  const int n = e - s;
  float sx = 0.0f, sq = 0.0f;
  for (int b = s; b < e; b += 8) {
    float c[8], c2[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      c[j] = b + j < e ? x[b + j] : 0.0f;
      c2[j] = c[j] * c[j];
    }
    sx = sx + tree8(c);
    sq = sq + tree8(c2);
  }
  mu = sx / (float)n;
  sq = sq - (sx * sx) / (float)n;
  sd = (float)sqrt((double)(sq / (float)n));
}

// sketch update covering both the exact (FP64) statistics, rounded
// outward, and the reference's FP32 statistics of the same segment
__device__ inline void sk_update_both(uint32_t *sk, double mu, double sd, float mu32, float sd32)
{
  float mh = round_up(mu), ml = round_down(mu), sh = round_up(sd), sl = round_down(sd);
  mh = fmaxf(mh, mu32);
  ml = fminf(ml, mu32);
  if (sd32 == sd32) {
    sh = fmaxf(sh, sd32);
    sl = fminf(sl, sd32);
  }
  sk_update(sk, mh, ml, sh, sl);
}

// one block per 1024 positions; when the whole range belongs to one node
// the sketch is reduced in LDS first
__global__ void __launch_bounds__(256) k_level_stats(uint64_t n, uint32_t L, const float *__restrict__ X,
                                                     const uint32_t *__restrict__ perm,
                                                     const uint32_t *__restrict__ pnode, Level lv)
{
  __shared__ uint32_t s_fine[kMaxFine * 4];
  __shared__ uint32_t s_coarse[kMaxFine * 4];
  __shared__ uint32_t s_node;
  __shared__ int s_uniform;
  const uint64_t p0 = (uint64_t)blockIdx.x * 1024;
  if (p0 >= n) return;
  const uint64_t p1 = p0 + 1024 < n ? p0 + 1024 : n;
  if (threadIdx.x == 0) {
    uint32_t a = pnode[p0], b = pnode[p1 - 1];
    s_node = a;
    s_uniform = (a == b && a != kNone);
  }
  __syncthreads();
  const bool uni = s_uniform;
  const uint32_t un = s_node;
  uint32_t nfine = 0, ncoarse = 0;
  if (uni) {
    nfine = lv.fine_off[un + 1] - lv.fine_off[un];
    ncoarse = lv.coarse_off[un + 1] - lv.coarse_off[un];
    for (uint32_t i = threadIdx.x; i < nfine; i += 256) {
      s_fine[4 * i + 0] = fenc(-FLT_MAX);
      s_fine[4 * i + 1] = fenc(FLT_MAX);
      s_fine[4 * i + 2] = fenc(-FLT_MAX);
      s_fine[4 * i + 3] = fenc(FLT_MAX);
    }
    for (uint32_t i = threadIdx.x; i < ncoarse; i += 256) {
      s_coarse[4 * i + 0] = fenc(-FLT_MAX);
      s_coarse[4 * i + 1] = fenc(FLT_MAX);
      s_coarse[4 * i + 2] = fenc(-FLT_MAX);
      s_coarse[4 * i + 3] = fenc(FLT_MAX);
    }
  }
  __syncthreads();
  for (uint64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    const uint32_t node = pnode[p];
    if (node == kNone) continue;
    const float *x = X + (uint64_t)perm[p] * L;
    const uint32_t fb = lv.fine_off[node], fe = lv.fine_off[node + 1];
    const uint32_t cb = lv.coarse_off[node];
    int start = 0;
    double cs = 0.0, cs2 = 0.0;
    int cstart = 0;
    for (uint32_t f = fb; f < fe; ++f) {
      const int end = lv.fine_end[f];
      double s = 0.0, s2 = 0.0;
      for (int t = start; t < end; ++t) {
        double v = (double)x[t];
        s += v;
        s2 += v * v;
      }
      const double len = (double)(end - start);
      double mu = s / len, var = s2 / len - mu * mu;
      double sd = var > 0.0 ? sqrt(var) : 0.0;
      uint32_t *fsk = uni ? &s_fine[4 * (f - fb)] : &lv.fine_sk[4 * f];
      float mu32, sd32;
      ref_seg_stats(x, start, end, mu32, sd32);
      sk_update_both(fsk, mu, sd, mu32, sd32);
      cs += s;
      cs2 += s2;
      const uint32_t c = lv.fine_coarse[f];
      const bool last = (f + 1 == fe) || lv.fine_coarse[f + 1] != c;
      if (last) {
        const double cl = (double)(end - cstart);
        double cm = cs / cl, cv = cs2 / cl - cm * cm;
        double csd = cv > 0.0 ? sqrt(cv) : 0.0;
        uint32_t *csk = uni ? &s_coarse[4 * c] : &lv.coarse_sk[4 * (cb + c)];
        float cm32, csd32;
        ref_seg_stats(x, cstart, end, cm32, csd32);
        sk_update_both(csk, cm, csd, cm32, csd32);
        cs = cs2 = 0.0;
        cstart = end;
      }
      start = end;
    }
  }
  __syncthreads();
  if (uni) {
    const uint32_t fb = lv.fine_off[un], cb = lv.coarse_off[un];
    for (uint32_t i = threadIdx.x; i < nfine; i += 256) {
      atomicMax(&lv.fine_sk[4 * (fb + i) + 0], s_fine[4 * i + 0]);
      atomicMin(&lv.fine_sk[4 * (fb + i) + 1], s_fine[4 * i + 1]);
      atomicMax(&lv.fine_sk[4 * (fb + i) + 2], s_fine[4 * i + 2]);
      atomicMin(&lv.fine_sk[4 * (fb + i) + 3], s_fine[4 * i + 3]);
    }
    for (uint32_t i = threadIdx.x; i < ncoarse; i += 256) {
      atomicMax(&lv.coarse_sk[4 * (cb + i) + 0], s_coarse[4 * i + 0]);
      atomicMin(&lv.coarse_sk[4 * (cb + i) + 1], s_coarse[4 * i + 1]);
      atomicMax(&lv.coarse_sk[4 * (cb + i) + 2], s_coarse[4 * i + 2]);
      atomicMin(&lv.coarse_sk[4 * (cb + i) + 3], s_coarse[4 * i + 3]);
    }
  }
}

struct Split {
  int32_t from, to;   // segment [from, to)
  int32_t idx;        // 0 mean, 1 sd; -1: not split (leaf)
  float value;
  uint32_t left, right;  // next-level node ids
};

// left flag of every position of a splitting node; leaves get pnode = none
__global__ void k_route(uint64_t n, uint32_t L, const float *__restrict__ X, const uint32_t *__restrict__ perm,
                        uint32_t *__restrict__ pnode, const Split *__restrict__ sp, uint32_t *__restrict__ flag)
{
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t node = pnode[p];
  uint32_t f = 0;
  if (node != kNone) {
    Split s = sp[node];
    if (s.idx < 0) {
      pnode[p] = kNone;
    } else {
      const float *x = X + (uint64_t)perm[p] * L;
      double a = 0.0, a2 = 0.0;
      for (int t = s.from; t < s.to; ++t) {
        double v = (double)x[t];
        a += v;
        a2 += v * v;
      }
      double len = (double)(s.to - s.from);
      double mu = a / len;
      double stat = mu;
      if (s.idx == 1) {
        double var = a2 / len - mu * mu;
        stat = var > 0.0 ? sqrt(var) : 0.0;
      }
      f = stat < (double)s.value ? 1u : 0u;
    }
  }
  flag[p] = f;
}

// stable partition: left series first, then right, inside every node
__global__ void k_scatter(uint64_t n, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ pnode,
                          const uint32_t *__restrict__ flag, const uint32_t *__restrict__ ex,
                          const uint64_t *__restrict__ nbeg, const uint32_t *__restrict__ nleft,
                          const Split *__restrict__ sp, uint32_t *__restrict__ perm2,
                          uint32_t *__restrict__ pnode2)
{
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t node = pnode[p];
  if (node == kNone) {
    perm2[p] = perm[p];
    pnode2[p] = kNone;
    return;
  }
  uint64_t b = nbeg[node];
  uint32_t lr = ex[p] - ex[b];  // lefts before p inside the node
  uint64_t dst;
  uint32_t child;
  if (flag[p]) {
    dst = b + lr;
    child = sp[node].left;
  } else {
    dst = b + nleft[node] + (uint32_t)(p - b - lr);
    child = sp[node].right;
  }
  perm2[dst] = perm[p];
  pnode2[dst] = child;
}

__global__ void k_count_left(uint32_t nf, const uint64_t *__restrict__ nbeg, const uint64_t *__restrict__ nend,
                             const uint32_t *__restrict__ flag, const uint32_t *__restrict__ ex,
                             const Split *__restrict__ sp, uint32_t *__restrict__ nleft)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  if (sp[i].idx < 0) {
    nleft[i] = 0;
    return;
  }
  uint64_t b = nbeg[i], e = nend[i];
  nleft[i] = ex[e - 1] + flag[e - 1] - ex[b];
}

// grid-stride: n * L / 4 vectors can exceed the 2^32 work-items one
// dispatch can address
__global__ void k_gather_rows(uint64_t n, uint32_t L, const float *__restrict__ X, const uint32_t *__restrict__ perm,
                              float *__restrict__ out)
{
  const uint32_t q = L / 4;
  const uint64_t nv = n * q, stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
    const uint64_t p = i / q, c = i - p * q;
    reinterpret_cast<float4 *>(out)[i] = reinterpret_cast<const float4 *>(X)[(uint64_t)perm[p] * q + c];
  }
}

// iSAX word of every row: PAA (sax.c:46-56, sequential FP32 sum then
// division) and the symbol = number of breakpoints strictly below it
struct Breakpoints {
  float v[256];
};

__global__ void k_sax_words(uint64_t n, uint32_t L, const float *__restrict__ X, uint8_t *__restrict__ sax,
                            Breakpoints bpv)
{
  __shared__ float bp[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) bp[i] = bpv.v[i];
  __syncthreads();
  uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const float *x = X + r * L;
  const int per = (int)(L / kPaaSegments);
  uint32_t w[4] = {0, 0, 0, 0};
  for (int s = 0; s < kPaaSegments; ++s) {
    float a = 0.0f;
    for (int i = 0; i < per; ++i) a = a + x[s * per + i];
    float pv = a / (float)per;
    int lo = 0, hi = 255;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (bp[mid] < pv) lo = mid + 1;
      else hi = mid;
    }
    w[s >> 2] |= (uint32_t)lo << ((s & 3) * 8);
  }
  reinterpret_cast<uint4 *>(sax)[r] = make_uint4(w[0], w[1], w[2], w[3]);
}

// ---- host side -------------------------------------------------------------
struct BNode {
  uint64_t begin, end;
  std::vector<int16_t> pts;  // coarse segment ends
  std::vector<float> sk;     // 4 per coarse segment (decoded)
  uint32_t level;
  int32_t left = -1, right = -1;  // BNode indices
  int32_t from = 0, to = 0, idx = -1;
  float value = 0.0f;
  bool force_leaf = false;  // a split of the parent separated nothing
};

double range_calc(const float *sk, int len)
{
  // len * (mean_width^2 + max_sd^2), in FP32 like the reference
  float w = sk[0] - sk[1];
  float u = sk[2];
  return (double)((float)len * (w * w + u * u));
}

// split choice of hercules_index_split_node for one node; returns false when
// no split separates the members (all statistics equal)
bool choose_split(const std::vector<int16_t> &pts, const std::vector<float> &csk, const std::vector<int16_t> &fpts,
                  const std::vector<float> &fsk, BNode &nd, int16_t &new_point)
{
  float best = -FLT_MAX;
  bool found = false;
  new_point = -1;
  auto consider = [&](const float *sk, int s, int e, bool fine) {
    const int len = e - s;
    const float q = (float)range_calc(sk, len);
    for (int pol = 0; pol < 2; ++pol) {
      float child[2][4];
      float value;
      for (int c = 0; c < 2; ++c)
        for (int j = 0; j < 4; ++j) child[c][j] = sk[j];
      if (pol == 0) {
        value = (sk[0] + sk[1]) / 2;
        child[0][1] = value;
        child[1][0] = value;
      } else {
        value = (sk[2] + sk[3]) / 2;
        child[0][2] = value;
        child[1][3] = value;
      }
      float r0 = (float)range_calc(child[0], len), r1 = (float)range_calc(child[1], len);
      float diff = q - (r0 + r1) / 2;
      const float lo = pol == 0 ? sk[1] : sk[3], hi = pol == 0 ? sk[0] : sk[2];
      if (diff > best && hi > lo) {
        best = diff;
        nd.from = s;
        nd.to = e;
        nd.idx = pol;
        nd.value = value;
        found = true;
        if (fine) {
          bool in_coarse = std::binary_search(pts.begin(), pts.end(), (int16_t)e);
          new_point = in_coarse ? (int16_t)s : (int16_t)e;
        } else {
          new_point = -1;
        }
      }
    }
  };
  int s = 0;
  for (size_t i = 0; i < pts.size(); ++i) {
    consider(&csk[4 * i], s, pts[i], false);
    s = pts[i];
  }
  best = best * 2;  // trade-off for refining the segmentation
  s = 0;
  for (size_t i = 0; i < fpts.size(); ++i) {
    // a refined segment equal to a coarse one was already considered
    bool is_coarse = std::binary_search(pts.begin(), pts.end(), fpts[i]) &&
                     (s == 0 || std::binary_search(pts.begin(), pts.end(), (int16_t)s));
    if (!is_coarse) consider(&fsk[4 * i], s, fpts[i], true);
    s = fpts[i];
  }
  return found;
}

void refined_points(const std::vector<int16_t> &pts, std::vector<int16_t> &f, std::vector<uint16_t> &coarse_of)
{
  f.clear();
  coarse_of.clear();
  int s = 0;
  for (size_t i = 0; i < pts.size(); ++i) {
    int len = pts[i] - s;
    if (len >= 2) {
      f.push_back((int16_t)(s + len / 2));
      coarse_of.push_back((uint16_t)i);
    }
    f.push_back(pts[i]);
    coarse_of.push_back((uint16_t)i);
    s = pts[i];
  }
}

#define BTRY(expr)                                                        \
  do {                                                                    \
    hipError_t _e = (expr);                                               \
    if (_e != hipSuccess) {                                               \
      set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
      return HKG_EDEVICE;                                                 \
    }                                                                     \
  } while (0)

template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  ~DBuf()
  {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t count)
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = count;
    return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
  }
  hipError_t up(const std::vector<T> &v)
  {
    if (v.size() > n) {
      hipError_t e = alloc(v.size());
      if (e != hipSuccess) return e;
    }
    return v.empty() ? hipSuccess : hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
};

}  // namespace

hkg_status build_device(float *dX, uint64_t n, uint32_t L, const hkg_build_opts &bo, Index **out)
{
  *out = nullptr;
  std::unique_ptr<Index> ix(new Index());
  hkg_status rc = ix->init_device(bo.device);
  if (rc != HKG_OK) {
    (void)hipFree(dX);
    return rc;
  }
  DevIndex &dv = ix->dev;
  dv.n = n;
  dv.L = L;
  const uint32_t leaf = std::max<uint32_t>(bo.leaf_size, 1);
  hipStream_t st = 0;
  DBuf<uint32_t> perm, perm2, pnode, pnode2, flag, ex;
  BTRY(perm.alloc(n));
  BTRY(perm2.alloc(n));
  BTRY(pnode.alloc(n));
  BTRY(pnode2.alloc(n));
  BTRY(flag.alloc(n));
  BTRY(ex.alloc(n));
  {
    std::vector<uint32_t> iota(n);
    for (uint64_t i = 0; i < n; ++i) iota[i] = (uint32_t)i;
    BTRY(hipMemcpy(perm.p, iota.data(), n * 4, hipMemcpyHostToDevice));
    std::fill(iota.begin(), iota.end(), 0u);
    BTRY(hipMemcpy(pnode.p, iota.data(), n * 4, hipMemcpyHostToDevice));
  }
  size_t scan_tmp_bytes = 0;
  BTRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp_bytes, flag.p, ex.p, (int)std::min<uint64_t>(n, INT32_MAX)));
  DBuf<uint8_t> scan_tmp;
  BTRY(scan_tmp.alloc(scan_tmp_bytes));

  std::vector<BNode> nodes;
  BNode root;
  root.begin = 0;
  root.end = n;
  root.level = 0;
  {
    const uint32_t segs = std::max<uint32_t>(bo.init_segments, 1);
    for (uint32_t i = 0; i < segs; ++i) root.pts.push_back((int16_t)((i + 1) * (L / segs)));
    root.pts.back() = (int16_t)L;  // calc_split_points
  }
  nodes.push_back(root);
  std::vector<uint32_t> frontier = {0};
  DBuf<uint32_t> d_fine_off, d_coarse_off, d_fine_sk, d_coarse_sk, d_nleft;
  DBuf<int16_t> d_fine_end;
  DBuf<uint16_t> d_fine_coarse;
  DBuf<Split> d_split;
  DBuf<uint64_t> d_nbeg, d_nend;
  while (!frontier.empty()) {
    const uint32_t nf = (uint32_t)frontier.size();
    std::vector<uint32_t> fine_off(1, 0), coarse_off(1, 0);
    std::vector<int16_t> fine_end;
    std::vector<uint16_t> fine_coarse;
    std::vector<std::vector<int16_t>> fpts(nf);
    for (uint32_t i = 0; i < nf; ++i) {
      const BNode &nd = nodes[frontier[i]];
      std::vector<uint16_t> co;
      refined_points(nd.pts, fpts[i], co);
      if (fpts[i].size() > (size_t)kMaxFine) {
        set_error("builder: too many segments in a node");
        return HKG_EINVAL;
      }
      fine_end.insert(fine_end.end(), fpts[i].begin(), fpts[i].end());
      fine_coarse.insert(fine_coarse.end(), co.begin(), co.end());
      fine_off.push_back((uint32_t)fine_end.size());
      coarse_off.push_back(coarse_off.back() + (uint32_t)nd.pts.size());
    }
    std::vector<uint32_t> init_f(4 * fine_end.size()), init_c(4 * coarse_off.back());
    for (size_t i = 0; i < init_f.size(); i += 4) {
      init_f[i] = fenc(-FLT_MAX);
      init_f[i + 1] = fenc(FLT_MAX);
      init_f[i + 2] = fenc(-FLT_MAX);
      init_f[i + 3] = fenc(FLT_MAX);
    }
    for (size_t i = 0; i < init_c.size(); i += 4) {
      init_c[i] = fenc(-FLT_MAX);
      init_c[i + 1] = fenc(FLT_MAX);
      init_c[i + 2] = fenc(-FLT_MAX);
      init_c[i + 3] = fenc(FLT_MAX);
    }
    BTRY(d_fine_off.up(fine_off));
    BTRY(d_coarse_off.up(coarse_off));
    BTRY(d_fine_end.up(fine_end));
    BTRY(d_fine_coarse.up(fine_coarse));
    BTRY(d_fine_sk.up(init_f));
    BTRY(d_coarse_sk.up(init_c));
    Level lv{nf, d_fine_off.p, d_fine_end.p, d_fine_coarse.p, d_coarse_off.p, d_fine_sk.p, d_coarse_sk.p};
    hipLaunchKernelGGL(k_level_stats, dim3((uint32_t)((n + 1023) / 1024)), dim3(256), 0, st, n, L, dX, perm.p,
                       pnode.p, lv);
    BTRY(hipGetLastError());
    BTRY(hipMemcpy(init_f.data(), d_fine_sk.p, init_f.size() * 4, hipMemcpyDeviceToHost));
    BTRY(hipMemcpy(init_c.data(), d_coarse_sk.p, init_c.size() * 4, hipMemcpyDeviceToHost));
    // decide
    std::vector<Split> splits(nf);
    std::vector<uint32_t> next;
    std::vector<uint64_t> nbeg(nf), nend(nf);
    for (uint32_t i = 0; i < nf; ++i) {
      BNode &nd = nodes[frontier[i]];
      nbeg[i] = nd.begin;
      nend[i] = nd.end;
      nd.sk.resize(4 * nd.pts.size());
      for (size_t j = 0; j < nd.sk.size(); ++j) nd.sk[j] = fdec(init_c[4 * coarse_off[i] + j]);
      std::vector<float> fsk(4 * fpts[i].size());
      for (size_t j = 0; j < fsk.size(); ++j) fsk[j] = fdec(init_f[4 * fine_off[i] + j]);
      Split s{0, 0, -1, 0.0f, kNone, kNone};
      int16_t np = -1;
      if (!nd.force_leaf && nd.end - nd.begin > leaf && choose_split(nd.pts, nd.sk, fpts[i], fsk, nd, np)) {
        s.from = nd.from;
        s.to = nd.to;
        s.idx = nd.idx;
        s.value = nd.value;
        std::vector<int16_t> cp = nd.pts;
        if (np >= 0) {
          cp.push_back(np);
          std::sort(cp.begin(), cp.end());
        }
        BNode l, r;
        l.pts = cp;
        r.pts = cp;
        l.level = r.level = nd.level + 1;
        nd.left = (int32_t)nodes.size();
        nd.right = nd.left + 1;
        s.left = (uint32_t)next.size();
        s.right = s.left + 1;
        next.push_back((uint32_t)nd.left);
        next.push_back((uint32_t)nd.right);
        nodes.push_back(l);
        nodes.push_back(r);
      } else {
        nodes[frontier[i]].idx = -1;
      }
      splits[i] = s;
    }
    BTRY(d_split.up(splits));
    BTRY(d_nbeg.up(nbeg));
    BTRY(d_nend.up(nend));
    if (d_nleft.n < nf) BTRY(d_nleft.alloc(nf));
    const uint32_t bs = 256;
    const uint32_t gridn = (uint32_t)((n + bs - 1) / bs);
    hipLaunchKernelGGL(k_route, dim3(gridn), dim3(bs), 0, st, n, L, dX, perm.p, pnode.p, d_split.p, flag.p);
    BTRY(hipGetLastError());
    BTRY(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, scan_tmp_bytes, flag.p, ex.p, (int)n, st));
    hipLaunchKernelGGL(k_count_left, dim3((nf + 255) / 256), dim3(256), 0, st, nf, d_nbeg.p, d_nend.p, flag.p,
                       ex.p, d_split.p, d_nleft.p);
    hipLaunchKernelGGL(k_scatter, dim3(gridn), dim3(bs), 0, st, n, perm.p, pnode.p, flag.p, ex.p, d_nbeg.p,
                       d_nleft.p, d_split.p, perm2.p, pnode2.p);
    BTRY(hipGetLastError());
    std::vector<uint32_t> nleft(nf);
    BTRY(hipMemcpy(nleft.data(), d_nleft.p, nf * 4, hipMemcpyDeviceToHost));
    std::swap(perm.p, perm2.p);
    std::swap(pnode.p, pnode2.p);
    for (uint32_t i = 0; i < nf; ++i) {
      BNode &nd = nodes[frontier[i]];
      if (nd.left < 0) continue;
      BNode &l = nodes[nd.left];
      BNode &r = nodes[nd.right];
      l.begin = nd.begin;
      l.end = nd.begin + nleft[i];
      r.begin = l.end;
      r.end = nd.end;
      // a split that separated nothing: the full child keeps the members
      // as an oversized leaf, the other is an empty leaf
      if (l.end == l.begin || r.end == r.begin) {
        l.force_leaf = true;
        r.force_leaf = true;
      }
    }
    frontier.swap(next);
  }
  // host tree in pre-order
  HostTree &t = dv.tree;
  t.L = L;
  t.init_segments = std::max<uint32_t>(bo.init_segments, 1);
  t.paa_segments = kPaaSegments;
  t.sax_bit_cardinality = 8;
  t.max_leaf_size = leaf;
  t.serial = 1;
  t.sims = 1;
  t.n_series = n;
  std::vector<int32_t> stack = {0};
  std::vector<int32_t> pre_of(nodes.size(), -1);
  // iterative pre-order; children indices patched after
  while (!stack.empty()) {
    int32_t b = stack.back();
    stack.pop_back();
    const BNode &nd = nodes[b];
    HostNode h;
    h.is_leaf = nd.left < 0;
    h.level = nd.level;
    h.size = (uint32_t)(nd.end - nd.begin);
    h.nseg = (uint16_t)nd.pts.size();
    h.seg_off = (uint32_t)t.seg_end.size();
    for (size_t i = 0; i < nd.pts.size(); ++i) {
      t.seg_end.push_back(nd.pts[i]);
      for (int j = 0; j < 4; ++j) t.ind.push_back(nd.sk[4 * i + j]);
    }
    if (h.is_leaf && h.size == 0) {
      // empty leaf: written with no name and no sketch (hercules_index.c:2174-2180)
      h.nseg = 0;
      t.seg_end.resize(h.seg_off);
      t.ind.resize(4 * (size_t)h.seg_off);
    } else if (h.is_leaf) {
      h.file_pos = nd.begin;
      h.leaf_idx = (int32_t)t.leaves.size();
      h.fname = "L" + std::to_string(h.leaf_idx);
      t.leaves.push_back((int32_t)t.nodes.size());
    } else {
      h.split_from = (int16_t)nd.from;
      h.split_to = (int16_t)nd.to;
      h.indicator_split_idx = nd.idx;
      h.indicator_split_value = nd.value;
      for (size_t i = 0; i < nd.pts.size(); ++i)
        if (nd.pts[i] == nd.to) h.split_segment = (int16_t)i;
      stack.push_back(nd.right);
      stack.push_back(nd.left);
    }
    pre_of[b] = (int32_t)t.nodes.size();
    t.nodes.push_back(h);
  }
  for (size_t b = 0; b < nodes.size(); ++b) {
    if (nodes[b].left >= 0) {
      t.nodes[pre_of[b]].left = pre_of[nodes[b].left];
      t.nodes[pre_of[b]].right = pre_of[nodes[b].right];
    }
  }
  t.leaf_count_hdr = t.leaves.size();
  // rows in leaf order, iSAX words, ids
  float *Xl = nullptr;
  {
    hipError_t e = hipMalloc((void **)&Xl, n * L * 4);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(dX);
      set_error("builder: no memory for the leaf-ordered copy");
      return HKG_ENOMEM;
    }
  }
  const uint64_t nv = n * (L / 4);
  hipLaunchKernelGGL(k_gather_rows, dim3((uint32_t)std::min<uint64_t>((nv + 255) / 256, 1u << 20)), dim3(256), 0, st,
                     n, L, dX, perm.p, Xl);
  BTRY(hipDeviceSynchronize());
  (void)hipFree(dX);
  dv.X = Xl;
  dv.device_bytes += n * L * 4;
  BTRY(hipMalloc((void **)&dv.sax, n * kPaaSegments));
  dv.device_bytes += n * kPaaSegments;
  Breakpoints bpv;
  host_breakpoints(bpv.v);
  bpv.v[255] = 0.0f;
  hipLaunchKernelGGL(k_sax_words, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, n, L, dv.X, dv.sax, bpv);
  dv.ids = perm.p;  // original ordinal of every leaf-order row
  perm.p = nullptr;
  dv.device_bytes += n * 4;
  BTRY(hipDeviceSynchronize());
  rc = ix->upload_tree();
  if (rc != HKG_OK) return rc;
  *out = ix.release();
  return HKG_OK;
}

hkg_status write_dir(const Index &ix, const char *dir)
{
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d += '/';
  std::string err;
  hkg_status rc = write_root_idx(d + "root.idx", ix.dev.tree, err);
  if (rc != HKG_OK) {
    set_error(err);
    return rc;
  }
  const uint64_t n = ix.dev.n, L = ix.dev.L;
  const uint64_t rows = 1u << 16;
  std::vector<float> buf(rows * L);
  FILE *f = fopen((d + "leaves_raw.idx").c_str(), "wb");
  if (!f) {
    set_error(d + "leaves_raw.idx: cannot create");
    return HKG_EIO;
  }
  for (uint64_t r = 0; r < n; r += rows) {
    uint64_t m = std::min(rows, n - r);
    if (hipMemcpy(buf.data(), ix.dev.X + r * L, m * L * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        fwrite(buf.data(), 4, m * L, f) != m * L) {
      fclose(f);
      set_error("writing leaves_raw.idx failed");
      return HKG_EIO;
    }
  }
  fclose(f);
  std::vector<uint8_t> sx(n * kPaaSegments);
  if (hipMemcpy(sx.data(), ix.dev.sax, sx.size(), hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("reading iSAX words failed");
    return HKG_EDEVICE;
  }
  f = fopen((d + "leaves_sims.idx").c_str(), "wb");
  if (!f || fwrite(sx.data(), 1, sx.size(), f) != sx.size()) {
    if (f) fclose(f);
    set_error("writing leaves_sims.idx failed");
    return HKG_EIO;
  }
  fclose(f);
  if (ix.dev.ids) {
    std::vector<uint32_t> ids(n);
    std::vector<uint64_t> ids64(n);
    if (hipMemcpy(ids.data(), ix.dev.ids, n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      set_error("reading ids failed");
      return HKG_EDEVICE;
    }
    if (!ix.dev.host_ids.empty()) ids64 = ix.dev.host_ids;
    else
      for (uint64_t i = 0; i < n; ++i) ids64[i] = ids[i];
    f = fopen((d + "ids.u64").c_str(), "wb");
    if (!f || fwrite(ids64.data(), 8, n, f) != n) {
      if (f) fclose(f);
      set_error("writing ids.u64 failed");
      return HKG_EIO;
    }
    fclose(f);
  }
  return HKG_OK;
}

}  // namespace hkg

extern "C" {

void hkg_default_build_opts(hkg_build_opts *bo)
{
  if (!bo) return;
  bo->leaf_size = 10000;
  bo->init_segments = 1;
  bo->paa_segments = 16;
  bo->device = 0;
}

hkg_status hkg_build_device(float *dX, uint64_t n, uint32_t L, const hkg_build_opts *bo, hkg_index **out)
{
  if (!out || !dX) return HKG_EINVAL;
  hkg_build_opts o;
  hkg_default_build_opts(&o);
  if (bo) o = *bo;
  if (n == 0 || n >= (1ull << 31) || L == 0 || L % 16 != 0 || L > hkg::kMaxBuildLength || o.paa_segments != 16) {
    hkg::set_error("build: need 0 < n < 2^31, L a multiple of 16 up to 2048, 16 iSAX segments");
    return HKG_EINVAL;
  }
  hkg::Index *ix = nullptr;
  hkg_status rc = hkg::build_device(dX, n, L, o, &ix);
  *out = reinterpret_cast<hkg_index *>(ix);
  return rc;
}

hkg_status hkg_build(const float *X, uint64_t n, uint32_t L, const hkg_build_opts *bo, hkg_index **out)
{
  if (!out || !X) return HKG_EINVAL;
  hkg_build_opts o;
  hkg_default_build_opts(&o);
  if (bo) o = *bo;
  if (hipSetDevice(o.device) != hipSuccess) {
    hkg::set_error("hipSetDevice failed");
    return HKG_EDEVICE;
  }
  float *dX = nullptr;
  if (hipMalloc((void **)&dX, n * L * 4) != hipSuccess) {
    (void)hipGetLastError();
    hkg::set_error("build: no device memory for the series");
    return HKG_ENOMEM;
  }
  if (hipMemcpy(dX, X, n * L * 4, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(dX);
    hkg::set_error("build: upload failed");
    return HKG_EDEVICE;
  }
  return hkg_build_device(dX, n, L, &o, out);
}

hkg_status hkg_write_dir(const hkg_index *ix, const char *index_dir)
{
  if (!ix || !index_dir) return HKG_EINVAL;
  return hkg::write_dir(*reinterpret_cast<const hkg::Index *>(ix), index_dir);
}

}  // extern "C"
