// Shared device definitions of the gfx950 progressive UMAP engine.
//
// Neighbour pools: every row owns W = 32 slots (the reference's MAX_NEIGHBORS,
// cpp/src/kernels/buffers.cuh:17) of packed 64-bit keys
//     key = float_bits(dist) << 32 | new_flag << 31 | id
// Distances are never negative, so unsigned key order is (dist, flag, id)
// order -- the reference's (dist, id) ResultItem order
// (cpp/src/kernels/update_kernel.cuh:65-69) with the new/old flag
// (InternalID_t, buffers.cuh:139-164) between them.  An empty slot holds
// ~0ull, above every real key.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>

namespace gm {

constexpr int W = 32;        // pool width
constexpr int CAND = 256;    // insertion candidate cap (reference MAX_CANDIDATES)
constexpr uint64_t EMPTY = ~0ull;
constexpr uint64_t NEW_BIT = 0x80000000ull;
constexpr uint32_t ID_MASK = 0x7fffffffu;
constexpr uint32_t NO_ID = 0xffffffffu;
constexpr uint32_t TAG_INIT_NBR = 1, TAG_EMB_INIT = 2, TAG_NOISE = 3,
                   TAG_NEG = 4;

__host__ __device__ inline uint64_t make_key(float d, uint32_t id,
                                             bool is_new) {
  return ((uint64_t)__float_as_uint(d) << 32) | (is_new ? NEW_BIT : 0ull) |
         (uint64_t)id;
}
__host__ __device__ inline uint32_t key_id(uint64_t k) {
  return (uint32_t)k & ID_MASK;
}
__host__ __device__ inline float key_dist(uint64_t k) {
  return __uint_as_float((uint32_t)(k >> 32));
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), constants as published.
__host__ __device__ inline void philox(const uint32_t in[4],
                                       const uint32_t key[2],
                                       uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = key[0], k1 = key[1];
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Counter-mode stream: counter = (a, b, block, tag), key = seed.
struct Rng {
  uint32_t key[2];
  uint32_t ctr[4];
  uint32_t buf[4];
  int pos;
  __device__ Rng(uint64_t seed, uint32_t a, uint32_t b, uint32_t tag) {
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
    ctr[0] = a;
    ctr[1] = b;
    ctr[2] = 0;
    ctr[3] = tag;
    pos = 4;
  }
  __device__ uint32_t next() {
    if (pos == 4) {
      philox(ctr, key, buf);
      ctr[2]++;
      pos = 0;
    }
    return buf[pos++];
  }
  __device__ float normal() { // Box-Muller
    const float u1 = ((float)next() + 1.0f) * 2.3283064365386963e-10f;
    const float u2 = (float)next() * 2.3283064365386963e-10f;
    return sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
  }
};

// IEEE round-to-nearest square root and quotient of floats.  The hardware
// estimate is within one ulp; the candidate is corrected with exact double
// arithmetic (products of 24/25-bit significands fit in 53 bits, and a float
// square root or quotient never lands exactly on a rounding midpoint).
__device__ inline float sqrt_rn(float x) {
  if (!(x > 0.0f) || isinf(x))
    return sqrtf(x);
  float y = sqrtf(x);
  const double xd = (double)x;
  const float up = nextafterf(y, INFINITY), dn = nextafterf(y, 0.0f);
  const double mu = 0.5 * ((double)y + (double)up);
  const double md = 0.5 * ((double)y + (double)dn);
  if (xd >= mu * mu)
    y = up;
  else if (xd < md * md)
    y = dn;
  return y;
}
__device__ inline float div_rn(float x, float d) {
  float q = x / d;
  if (!isfinite(q) || q == 0.0f || !isfinite(d) || d == 0.0f)
    return q;
  const double xd = (double)x, dd = (double)d;
  const float up = nextafterf(q, INFINITY), dn = nextafterf(q, -INFINITY);
  // pick the neighbour whose product is nearest x (products are exact)
  const double e0 = fabs(xd - (double)q * dd);
  const double eu = fabs(xd - (double)up * dd);
  const double ed = fabs(xd - (double)dn * dd);
  if (eu < e0 && eu <= ed)
    return up;
  if (ed < e0)
    return dn;
  return q;
}

// Canonical squared distance: one FMA chain over the features in index
// order.  Every kernel that produces a distance destined for a pool uses this
// exact operation order, so a pair's distance is bit-identical wherever it is
// computed (and identical to the CPU restatement).
__device__ inline float canon_dist_g(const float *__restrict__ a,
                                     const float *__restrict__ b, int dp) {
  float acc = 0.0f;
  for (int i = 0; i < dp; i += 4) {
    const float4 x = *reinterpret_cast<const float4 *>(a + i);
    const float4 y = *reinterpret_cast<const float4 *>(b + i);
    float t = x.x - y.x;
    acc = __builtin_fmaf(t, t, acc);
    t = x.y - y.y;
    acc = __builtin_fmaf(t, t, acc);
    t = x.z - y.z;
    acc = __builtin_fmaf(t, t, acc);
    t = x.w - y.w;
    acc = __builtin_fmaf(t, t, acc);
  }
  return acc;
}

__device__ inline uint64_t ld_relaxed(const uint64_t *p) {
  return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool cas64(uint64_t *p, uint64_t expect, uint64_t val) {
  return __hip_atomic_compare_exchange_strong(p, &expect, val,
                                              __ATOMIC_RELAXED,
                                              __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}

// Ascending sort of one value per lane over the full 64-lane wave.
__device__ inline uint32_t wave_sort64(uint32_t x) {
  const int l = __lane_id();
#pragma unroll
  for (int size = 2; size <= 64; size <<= 1)
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const uint32_t y = __shfl_xor(x, stride);
      const bool up = (size == 64) || ((l & size) == 0);
      const bool lower = (l & stride) == 0;
      x = (lower == up) ? min(x, y) : max(x, y);
    }
  return x;
}

// Ascending sort of one value per lane within each 32-lane half.
template <typename T> __device__ inline T half_sort32(T x) {
  const int l = __lane_id() & 31;
#pragma unroll
  for (int size = 2; size <= 32; size <<= 1)
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const T y = __shfl_xor(x, stride, 32);
      const bool up = (size == 32) || ((l & size) == 0);
      const bool lower = (l & stride) == 0;
      x = (lower == up) ? (x < y ? x : y) : (x < y ? y : x);
    }
  return x;
}

// Lock-free bounded top-W insert, executed by one 32-lane half of a wave
// (lane = slot).  Contract (identical to the CPU restatement): reject a
// duplicate id or a key not below the row maximum; otherwise replace the
// maximum (lowest slot among equal maxima) by a 64-bit CAS and retry on
// failure.  Every successful CAS replaces the row's current maximum, so a
// successful CAS saw the row exactly as it was, the final row is the W
// smallest distinct keys offered regardless of order, and `rank` (keys below
// the new one) is exact.
// Returns 0 rejected, 1 inserted at rank >= k, 2 inserted at rank < k.
__device__ inline int pool_offer_half(uint64_t *__restrict__ row,
                                      uint64_t key, int k, int hl, int half) {
  const uint32_t id = key_id(key);
  const int base = half * 32;
  for (;;) {
    const uint64_t v = ld_relaxed(row + hl);
    const bool dup = (v != EMPTY) && key_id(v) == id;
    const uint64_t dmask = (__ballot(dup) >> base) & 0xffffffffull;
    if (dmask)
      return 0;
    uint64_t mx = v;
    int ml = hl;
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) {
      const uint64_t ov = __shfl_xor(mx, off, 32);
      const int ol = __shfl_xor(ml, off, 32);
      if (ov > mx || (ov == mx && ol < ml)) {
        mx = ov;
        ml = ol;
      }
    }
    if (key >= mx)
      return 0;
    const uint64_t lmask = (__ballot(v < key) >> base) & 0xffffffffull;
    const int rank = __popcll(lmask);
    int ok = 0;
    if (hl == ml)
      ok = cas64(row + ml, mx, key) ? 1 : 0;
    ok = __shfl(ok, ml, 32);
    if (ok) {
      return rank < k ? 2 : 1;
    }
  }
}

// ---------------------------------------------------------- rho / sigma
// smooth_knn_dist of one row in FP64 (reference smooth_knn_dist_kernel,
// embedding_kernel.cuh:135-207) by a 32-lane half-wave: lane c < k holds the
// row's embed-time distance c (0 for the self column, the sqrt of the stored
// squared distance for euclidean rows, 0 for an invalid entry).  rho = first
// non-zero distance; sigma by 64-step bisection of
// sum_{j=2..k-1} exp(-max(0, d_j - rho)/sigma) = log2(k-1), floored at 1e-3
// times the row mean.  A row without a non-zero distance has rho = 0 and
// sigma = -1: its sigma is 1e-3 times the global mean distance, which the
// reader resolves (sigma_of), so the value stored here depends on the row's
// public kNN list only.  Sums are xor-shuffle trees; the bisection's
// decisions are uniform over the half-wave.  All 32 lanes must call this.

// One lane's double from the lane `CTRL` names (a DPP row pattern).
template <int CTRL> __device__ inline double dpp_f64(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(
      0, (int)(unsigned)u, CTRL, 0xf, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(
      0, (int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(
      (long long)(((unsigned long long)hi << 32) | (unsigned long long)lo));
}

// Sum of a double over each 16-lane row, every lane of the row receiving
// it: rotations within the row by 8, 4, 2 and 1 lanes.  After the rotation
// by 8 the values repeat every 8 lanes (then every 4, 2), so each rotation
// hands a lane the value its xor partner holds, and the sums are those of
// the xor-shuffle tree over the row, operand for operand.
__device__ inline double row16_sum(double v) {
  v += dpp_f64<0x128>(v); // row_ror:8
  v += dpp_f64<0x124>(v); // row_ror:4
  v += dpp_f64<0x122>(v); // row_ror:2
  v += dpp_f64<0x121>(v); // row_ror:1
  return v;
}

__device__ inline void row_sigma_rho(double di, int c, int k, int half,
                                     double &sigma_out, double &rho_out) {
  double sum = di;
#pragma unroll
  for (int off = 16; off > 0; off >>= 1)
    sum += __shfl_xor(sum, off, 32);
  const uint32_t nzm =
      (uint32_t)((__ballot(c < k && di > 0.0) >> (half * 32)) & 0xffffffffull);
  if (nzm == 0) {
    rho_out = 0.0;
    sigma_out = -1.0;
    return;
  }
  const double r = __shfl(di, __ffs(nzm) - 1, 32);
  const double row_mean = sum / (double)k;
  const double target = log2((double)(k - 1));
  const bool term = c >= 2 && c < k;
  const double x = di - r;
  // At k <= 16 every term sits in lanes 2 .. 15, the tree's first step (xor
  // 16) adds exact zeros, and what follows is a sum over 16 lanes: lanes c
  // and c + 16 both take entry c, and each 16-lane row sums by DPP
  // (row16_sum), with the same result as the 32-lane tree.
  const bool narrow = k <= 16;
  const double xn = narrow ? __shfl(x, c & 15, 32) : x;
  const bool tn = narrow ? ((c & 15) >= 2 && (c & 15) < k) : term;
  double lo = 0.0, hi = INFINITY, mid = 1.0;
  for (int it = 0; it < 64; ++it) {
    double psum = tn ? ((xn > 0.0) ? exp(-xn / mid) : 1.0) : 0.0;
    if (narrow) {
      psum = row16_sum(psum);
    } else {
#pragma unroll
      for (int off = 16; off > 0; off >>= 1)
        psum += __shfl_xor(psum, off, 32);
    }
    if (fabs(psum - target) < 1e-5)
      break;
    if (psum > target) {
      hi = mid;
      mid = (lo + hi) * 0.5;
    } else {
      lo = mid;
      mid = isinf(hi) ? mid * 2.0 : (lo + hi) * 0.5;
    }
  }
  if (!isfinite(mid) || mid <= 0.0)
    mid = 1.0;
  const double min_scale = 1e-3 * row_mean; // r > 0 here
  rho_out = r;
  sigma_out = mid < min_scale ? min_scale : mid;
}

// a row's sigma with the global-mean floor resolved
__device__ inline double sigma_of(const double *sigma, uint64_t i,
                                  double mean_dist) {
  const double s = sigma[i];
  return s < 0.0 ? 1e-3 * mean_dist : s;
}

// an embed-time distance from a stored squared distance (the reference's
// sqrt at embed time, embedding_kernel.cuh:627-629): invalid entries
// (sentinels) stay as they are
__device__ inline bool valid_embed_d(double d) {
  return isfinite(d) && d >= 0.0 && d < 3.402823466e38 / 2.0;
}
__device__ inline double embed_d(float v, bool euclid) {
  if (!valid_embed_d((double)v))
    return (double)v;
  return euclid ? (double)sqrt_rn(fmaxf(v, 0.0f)) : (double)v;
}

// The same bounded insert by a single lane: read the W slots (relaxed loads,
// possibly at different times), reject a duplicate id or a key not below the
// maximum, else CAS the maximum slot (lowest index among ties) and retry on
// failure.  Slots only ever decrease, so the values read are >= the row's
// current ones; a CAS that succeeds found its slot still holding the
// maximum read, which is then the row's current maximum -- the row ends as
// the W smallest distinct keys offered, in any order of the offers.
// Returns true when the key was inserted.
__device__ inline bool pool_offer_lane(uint64_t *__restrict__ row,
                                       uint64_t key) {
  const uint32_t id = key_id(key);
  for (;;) {
    uint64_t mx = 0;
    int ml = 0;
    bool dup = false;
#pragma unroll
    for (int s = 0; s < W; ++s) {
      const uint64_t v = ld_relaxed(row + s);
      dup |= (v != EMPTY) && key_id(v) == id;
      if (v > mx) {
        mx = v;
        ml = s;
      }
    }
    if (dup || key >= mx)
      return false;
    if (cas64(row + ml, mx, key))
      return true;
  }
}

// OCP FP8 E4M3 (the gfx950 matrix cores' E4M3 operands): sign, 4 exponent
// bits (bias 7), 3 mantissa bits; normals 2^-6 .. 448, subnormals k 2^-9,
// code 0x7f is NaN.  e4m3_rn rounds to nearest, ties to an even mantissa,
// and saturates at +-448 (|x| above 464, which would round beyond the range,
// also gives 448); the screens only convert values with |x| <= 448.  The
// rounding error is at most 2^-4 |x| (normal range) or 2^-10 (subnormal).
__host__ __device__ inline float e4m3_dec(uint8_t v) {
  const int e = (v >> 3) & 15, m = v & 7;
  const float r = e == 0 ? (float)m * 0.001953125f // m 2^-9
                         : ldexpf(1.0f + (float)m * 0.125f, e - 7);
  return (v & 0x80) ? -r : r;
}

__host__ __device__ inline uint8_t e4m3_rn(float x) {
  uint32_t b;
  memcpy(&b, &x, 4);
  const uint8_t sign = (uint8_t)((b >> 24) & 0x80u);
  b &= 0x7fffffffu;
  float a;
  memcpy(&a, &b, 4);
  if (!(a < 464.0f))
    return (uint8_t)(sign | 0x7e); // 448 (NaN also lands here)
  if (a < 0.015625f) {             // below 2^-6: multiples of 2^-9
    const float q = rintf(a * 512.0f); // exact scaling, then nearest-even
    return (uint8_t)(sign | (uint8_t)q); // q = 8 is 2^-6, code 0x08
  }
  // round the 23-bit mantissa to 3 bits, nearest-even (a carry moves to
  // the exponent); 2^-6 <= a < 464 keeps the exponent within 1 .. 15
  const uint32_t r = (b + 0x7ffffu + ((b >> 20) & 1u)) & ~0xfffffu;
  const int e = (int)(r >> 23) - 127 + 7;
  const uint32_t m = (r >> 20) & 7u;
  if (e > 15 || (e == 15 && m == 7))
    return (uint8_t)(sign | 0x7e);
  return (uint8_t)(sign | (uint8_t)(e << 3) | (uint8_t)m);
}

} // namespace gm
