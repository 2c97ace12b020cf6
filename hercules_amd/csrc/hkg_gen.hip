// Workload generator on the device.
//
// Counter-based, so any series can be produced from (seed, id) alone and in
// any order: draw(seed, stream, j) = mix64(mix64(seed ^ G) ^ mix64(stream<<20 | j))
// with mix64 the SplitMix64 finaliser; uniforms ((bits >> 11) + 0.5) 2^-53;
// normal pairs by Box-Muller with log/sin/cos written with basic IEEE
// operations only (no libm), so the host specification and this kernel give
// bit-identical series.  Compiled with -ffp-contract=off.
//
// kind 0: random walk (sequential FP64 prefix sum of N(0,1) steps),
//         z-normalised (population sd) and rounded to FP32;
// kind 1: N(0,1) vector, L2-normalised, then z-normalised.
#include <float.h>

#include <vector>

#include "hkg_internal.h"

namespace hkg {

__host__ __device__ inline uint64_t mix64(uint64_t x)
{
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}

__host__ __device__ inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t j)
{
  uint64_t key = mix64(seed ^ 0x9E3779B97F4A7C15ULL);
  uint64_t c = (stream << 20) | (j & 0xFFFFFULL);
  return mix64(key ^ mix64(c));
}

__device__ inline double gen_uniform(uint64_t bits) { return ((double)(bits >> 11) + 0.5) * 0x1.0p-53; }

__device__ inline double gen_log(double u)
{
  uint64_t b = (uint64_t)__double_as_longlong(u);
  int e = (int)((b >> 52) & 0x7FF) - 1022;
  uint64_t mb = (b & 0x000FFFFFFFFFFFFFULL) | (1022ULL << 52);
  double m = __longlong_as_double((long long)mb);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double t = 1.0 / 23.0;
  t = t * s2 + 1.0 / 21.0;
  t = t * s2 + 1.0 / 19.0;
  t = t * s2 + 1.0 / 17.0;
  t = t * s2 + 1.0 / 15.0;
  t = t * s2 + 1.0 / 13.0;
  t = t * s2 + 1.0 / 11.0;
  t = t * s2 + 1.0 / 9.0;
  t = t * s2 + 1.0 / 7.0;
  t = t * s2 + 1.0 / 5.0;
  t = t * s2 + 1.0 / 3.0;
  t = t * s2 + 1.0;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  return (double)e * ln2_hi + ((2.0 * s) * t + (double)e * ln2_lo);
}

// correctly rounded square root: the hardware result, then the neighbour
// whose square is closest to x
__device__ inline double gen_sqrt(double x)
{
  double y = sqrt(x);
  double yp = nextafter(y, DBL_MAX), ym = nextafter(y, 0.0);
  double r = fabs(fma(-y, y, x)), rp = fabs(fma(-yp, yp, x)), rm = fabs(fma(-ym, ym, x));
  if (rp < r) {
    y = yp;
    r = rp;
  }
  if (rm < r) y = ym;
  return y;
}

__device__ inline void gen_normal_pair(uint64_t seed, uint64_t stream, uint64_t pair, double &z0, double &z1)
{
  double u1 = gen_uniform(draw(seed, stream, 2 * pair));
  double u2 = gen_uniform(draw(seed, stream, 2 * pair + 1));
  double r = gen_sqrt(-2.0 * gen_log(u1));
  double t4 = u2 * 4.0;
  int q = (int)t4;
  double a = (t4 - (double)q) * 1.57079632679489661923;
  double a2 = a * a;
  double s = 1.0, c = 1.0;
  for (int k = 13; k >= 1; --k) {
    s = 1.0 - (a2 / (double)((2 * k) * (2 * k + 1))) * s;
    c = 1.0 - (a2 / (double)((2 * k - 1) * (2 * k))) * c;
  }
  double sn = a * s, cs = c;
  double cc, ss;
  switch (q & 3) {
  case 0: cc = cs; ss = sn; break;
  case 1: cc = -sn; ss = cs; break;
  case 2: cc = -cs; ss = -sn; break;
  default: cc = sn; ss = -cs; break;
  }
  z0 = r * cc;
  z1 = r * ss;
}

// raw value t of a series before z-normalisation, produced sequentially
struct RawSeries {
  int kind;
  uint64_t seed, id;
  double walk = 0.0, z1 = 0.0, inv = 1.0;  // inv: 1/||v|| for kind 1 (applied as v/nrm)
  double nrm = 1.0;
  __device__ double next(uint32_t t)
  {
    double z;
    if ((t & 1) == 0) {
      double z0;
      gen_normal_pair(seed, id, t >> 1, z0, z1);
      z = z0;
    } else {
      z = z1;
    }
    if (kind == 0) {
      walk = walk + z;
      return walk;
    }
    return z / nrm;
  }
};

__device__ void series_stats(int kind, uint64_t seed, uint64_t id, uint32_t L, double &nrm, double &mu, double &sd)
{
  nrm = 1.0;
  if (kind == 1) {
    RawSeries r{1, seed, id};
    double n2 = 0.0;
    for (uint32_t t = 0; t < L; ++t) {
      double v = r.next(t);
      n2 = n2 + v * v;
    }
    nrm = gen_sqrt(n2);
  }
  RawSeries a{kind, seed, id};
  a.nrm = nrm;
  double sum = 0.0;
  for (uint32_t t = 0; t < L; ++t) sum = sum + a.next(t);
  mu = sum / (double)L;
  RawSeries b{kind, seed, id};
  b.nrm = nrm;
  double ss = 0.0;
  for (uint32_t t = 0; t < L; ++t) {
    double d = b.next(t) - mu;
    ss = ss + d * d;
  }
  sd = gen_sqrt(ss / (double)L);
}

__global__ void k_gen_series(int kind, uint64_t seed, uint64_t first, uint64_t n, uint32_t L, float *__restrict__ out)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t id = first + i;
  double nrm, mu, sd;
  series_stats(kind, seed, id, L, nrm, mu, sd);
  RawSeries c{kind, seed, id};
  c.nrm = nrm;
  float *o = out + i * L;
  for (uint32_t t = 0; t < L; ++t) o[t] = (float)((c.next(t) - mu) / sd);
}

// q_i = znorm(X[src_i] + sigma_i * N(0,1)) with the noise from stream i of
// query_seed; X[src] is regenerated (FP32-rounded) on the fly
__global__ void k_gen_noisy(int kind, uint64_t data_seed, uint64_t query_seed, uint32_t nq, uint32_t L,
                            const float *__restrict__ sigmas, const uint64_t *__restrict__ src,
                            float *__restrict__ out)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  double nrm, mu, sd;
  series_stats(kind, data_seed, src[i], L, nrm, mu, sd);
  const double sg = (double)sigmas[i];
  // pass 1: sum, pass 2: sum of squared deviations, pass 3: write
  double sum = 0.0, m = 0.0, ss = 0.0;
  for (int pass = 0; pass < 3; ++pass) {
    RawSeries x{kind, data_seed, src[i]};
    x.nrm = nrm;
    double z1 = 0.0;
    for (uint32_t t = 0; t < L; ++t) {
      double xv = (double)(float)((x.next(t) - mu) / sd);
      double z;
      if ((t & 1) == 0) {
        double z0;
        gen_normal_pair(query_seed, i, t >> 1, z0, z1);
        z = z0;
      } else {
        z = z1;
      }
      double v = xv + sg * z;
      if (pass == 0) {
        sum = sum + v;
      } else if (pass == 1) {
        double d = v - m;
        ss = ss + d * d;
      } else {
        out[(uint64_t)i * L + t] = (float)((v - m) / gen_sqrt(ss / (double)L));
      }
    }
    if (pass == 0) m = sum / (double)L;
  }
}

uint64_t noisy_source(uint64_t query_seed, uint32_t i, uint64_t N) { return draw(query_seed, i, 0xFFFFF) % N; }

}  // namespace hkg

extern "C" {

hkg_status hkg_gen_series_device(int kind, uint64_t seed, uint64_t first, uint64_t n, uint32_t L, float *dX,
                                 void *stream)
{
  if (!dX || L == 0 || (L & 1) || (kind != 0 && kind != 1)) return HKG_EINVAL;
  if (n == 0) return HKG_OK;
  hipStream_t st = (hipStream_t)stream;
  const uint32_t bs = 256;
  // bound one launch to 2^24 series so a grid never gets too large
  for (uint64_t b = 0; b < n; b += (1ull << 24)) {
    uint64_t m = n - b < (1ull << 24) ? n - b : (1ull << 24);
    hipLaunchKernelGGL(hkg::k_gen_series, dim3((uint32_t)((m + bs - 1) / bs)), dim3(bs), 0, st, kind, seed,
                       first + b, m, L, dX + b * L);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    hkg::set_error(std::string("k_gen_series: ") + hipGetErrorString(e));
    return HKG_EDEVICE;
  }
  return HKG_OK;
}

hkg_status hkg_gen_noisy_queries_device(int kind, uint64_t data_seed, uint64_t N, uint64_t query_seed, uint32_t nq,
                                        uint32_t L, const float *sigmas, float *dQ, uint64_t *src_ids)
{
  if (!dQ || !sigmas || N == 0 || L == 0 || (L & 1)) return HKG_EINVAL;
  if (nq == 0) return HKG_OK;
  std::vector<uint64_t> src(nq);
  for (uint32_t i = 0; i < nq; ++i) src[i] = hkg::noisy_source(query_seed, i, N);
  if (src_ids)
    for (uint32_t i = 0; i < nq; ++i) src_ids[i] = src[i];
  float *dsig = nullptr;
  uint64_t *dsrc = nullptr;
  if (hipMalloc((void **)&dsig, nq * 4) != hipSuccess || hipMalloc((void **)&dsrc, nq * 8) != hipSuccess) {
    if (dsig) (void)hipFree(dsig);
    hkg::set_error("hipMalloc failed");
    return HKG_ENOMEM;
  }
  hkg_status rc = HKG_OK;
  if (hipMemcpy(dsig, sigmas, nq * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dsrc, src.data(), nq * 8, hipMemcpyHostToDevice) != hipSuccess) {
    rc = HKG_EDEVICE;
  } else {
    hipLaunchKernelGGL(hkg::k_gen_noisy, dim3((nq + 63) / 64), dim3(64), 0, 0, kind, data_seed, query_seed, nq, L,
                       dsig, dsrc, dQ);
    if (hipDeviceSynchronize() != hipSuccess) rc = HKG_EDEVICE;
  }
  (void)hipFree(dsig);
  (void)hipFree(dsrc);
  if (rc != HKG_OK) hkg::set_error("noisy query generation failed");
  return rc;
}

}  // extern "C"
