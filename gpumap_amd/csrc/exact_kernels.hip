// Exact insertion by distance tiles (MI355X path of north-star subsystem (a)).
//
// The reference builds a new row's neighbours by NN-Descent from random
// starting points (insertion_kernel.cuh:926-1095).  On MI355X the brute-force
// alternative is cheap -- one insertion of B rows into n costs B * n * D
// multiply-adds, a few TFLOP for a whole 70,000 x 784 session -- and makes the
// graph exact: every inserted row gets the exact 32 nearest rows inserted so
// far, and every earlier row is offered every new row.  By induction every
// pool is the exact 32-NN set, so the NN-Descent update has nothing left to
// improve.
//
// FP32 path: every pair's canonical distance (one FMA chain over the features,
// bit-identical to the CPU restatement) in 64 x 64 LDS-staged tiles.
// Forward: each block keeps, per batch row, the 32 smallest keys of its column
// split in LDS (threshold, append, bitonic compaction); k_exact_merge takes the
// 32 smallest over all splits.  Reverse: pairs (batch row u, earlier row j)
// whose distance does not exceed j's current pool maximum are offered to j's
// pool with the lock-free insert, whose result does not depend on the order.
#include "common.hpp"
#include "kernels.hpp"

namespace gm {

namespace {

template <typename T>
__device__ void block_bitonic_e(T *a, int len, int seg) {
  for (int size = 2; size <= seg; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < len / 2; t += blockDim.x) {
        const int i = 2 * stride * (t / stride) + (t % stride);
        const int j = i + stride;
        const bool up = (size == seg) || ((i & size) == 0);
        const T x = a[i], y = a[j];
        if ((x > y) == up) {
          a[i] = y;
          a[j] = x;
        }
      }
      __syncthreads();
    }
}

constexpr int ER = 64;   // batch rows per block
constexpr int EC = 64;   // columns per tile
constexpr int EB = 64;   // per-row buffer: W kept + 32 appended per half tile
constexpr int EQ = 2048; // reverse-offer queue per half tile

} // namespace

__global__ __launch_bounds__(256) void k_exact_tiles(
    const float *__restrict__ X, uint64_t dp, uint64_t *__restrict__ pool,
    const float *__restrict__ pmax, uint64_t b0, uint64_t nb, uint64_t n,
    int k, int nsplit, uint64_t *__restrict__ part, uint8_t *__restrict__ flag,
    int new_flag, const unsigned *__restrict__ guard) {
  // guard: the BF16 screening's queue state; run only when it overflowed
  if (guard && guard[1] == 0)
    return;
  __shared__ float s_vec[128 * 33];
  __shared__ uint64_t s_buf[ER * EB];
  __shared__ uint64_t s_th[ER];
  __shared__ int s_cnt[ER];
  __shared__ float s_pm[EC];
  __shared__ uint64_t s_qkey[EQ];
  __shared__ uint8_t s_qtgt[EQ];
  __shared__ int s_qn, s_need;
  const int tid = threadIdx.x;
  const int rt = tid >> 4, ct = tid & 15;
  const int hw = tid >> 5, hl = tid & 31, half = (tid >> 5) & 1;
  const uint64_t r0 = (uint64_t)blockIdx.x * ER;
  const int rows = (int)min((uint64_t)ER, nb - r0);
  const int split = blockIdx.y;
  const uint64_t ntile = (n + EC - 1) / EC;

  if (tid < ER) {
    s_cnt[tid] = 0;
    s_th[tid] = EMPTY;
  }
  __syncthreads();

  for (uint64_t t = split; t < ntile; t += nsplit) {
    const uint64_t c0 = t * EC;
    const int cols = (int)min((uint64_t)EC, n - c0);
    if (tid < EC) {
      const uint64_t j = c0 + tid;
      s_pm[tid] = (tid < cols && j < b0) ? pmax[j] : -1.0f;
    }
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = 0.0f;
    for (uint64_t base = 0; base < dp; base += 32) {
      const int fw = (int)min((uint64_t)32, dp - base);
      for (int e = tid; e < 128 * 8; e += blockDim.x) {
        const int m = e >> 3, f4 = (e & 7) * 4;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (f4 < fw) {
          if (m < ER) {
            if (m < rows)
              v = *reinterpret_cast<const float4 *>(
                  X + (b0 + r0 + m) * dp + base + f4);
          } else if (m - ER < cols) {
            v = *reinterpret_cast<const float4 *>(
                X + (c0 + (m - ER)) * dp + base + f4);
          }
        }
        float *dst = s_vec + m * 33 + f4;
        dst[0] = v.x;
        dst[1] = v.y;
        dst[2] = v.z;
        dst[3] = v.w;
      }
      __syncthreads();
      for (int f = 0; f < fw; ++f) {
        float rv[4], cv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
          rv[i] = s_vec[(rt + 16 * i) * 33 + f];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          cv[j] = s_vec[(ER + ct + 16 * j) * 33 + f];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float d = rv[i] - cv[j];
            acc[i][j] = __builtin_fmaf(d, d, acc[i][j]);
          }
      }
      __syncthreads();
    }

    // two half tiles of 32 columns: forward appends, compaction, reverse
    for (int hc = 0; hc < 2; ++hc) {
      if (tid == 0) {
        s_need = 0;
        s_qn = 0;
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = rt + 16 * i;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int cj = ct + 16 * (2 * hc + jj);
          const uint64_t u = b0 + r0 + r, j = c0 + cj;
          if (r >= rows || cj >= cols || j == u)
            continue;
          const float d = acc[i][2 * hc + jj];
          const uint64_t key = make_key(d, (uint32_t)j, new_flag != 0);
          if (key < s_th[r]) {
            const int slot = atomicAdd(&s_cnt[r], 1);
            s_buf[r * EB + slot] = key;
            if (slot + 1 > W)
              s_need = 1;
          }
          if (d <= s_pm[cj]) {
            const int at = atomicAdd(&s_qn, 1);
            s_qkey[at] = make_key(d, (uint32_t)u, new_flag != 0);
            s_qtgt[at] = (uint8_t)cj;
          }
        }
      }
      __syncthreads();
      if (s_need) {
        for (int e = tid; e < ER * EB; e += blockDim.x)
          if ((e % EB) >= s_cnt[e / EB])
            s_buf[e] = EMPTY;
        __syncthreads();
        block_bitonic_e(s_buf, ER * EB, EB);
        if (tid < ER && s_cnt[tid] >= W) {
          s_cnt[tid] = W;
          s_th[tid] = s_buf[tid * EB + W - 1];
        }
        __syncthreads();
      }
      const int qn = s_qn;
      for (int e = hw; e < qn; e += 8) {
        const uint64_t j = c0 + s_qtgt[e];
        if (pool_offer_half(pool + j * W, s_qkey[e], k, hl, half) && hl == 0)
          flag[j] = 1;
      }
      __syncthreads();
    }
  }

  // final per-row top W of this split, sorted
  for (int e = tid; e < ER * EB; e += blockDim.x)
    if ((e % EB) >= s_cnt[e / EB])
      s_buf[e] = EMPTY;
  __syncthreads();
  block_bitonic_e(s_buf, ER * EB, EB);
  for (int e = tid; e < rows * W; e += blockDim.x) {
    const int r = e / W, s = e % W;
    part[((r0 + r) * (uint64_t)nsplit + split) * W + s] = s_buf[r * EB + s];
  }
}

// Top W over all splits for one batch row -> its pool.
__global__ __launch_bounds__(256) void k_exact_merge(
    const uint64_t *__restrict__ part, int nsplit, uint64_t *__restrict__ pool,
    uint64_t b0, uint64_t nb, uint8_t *__restrict__ flag,
    const unsigned *__restrict__ guard) {
  __shared__ uint64_t s[4096];
  const uint64_t bi = blockIdx.x;
  if (bi >= nb || (guard && guard[1] == 0))
    return;
  const int len = nsplit * W; // power of two, <= 4096
  for (int e = threadIdx.x; e < len; e += blockDim.x)
    s[e] = part[bi * (uint64_t)len + e];
  __syncthreads();
  block_bitonic_e(s, len, len);
  const uint64_t u = b0 + bi;
  if (threadIdx.x < W)
    pool[u * W + threadIdx.x] = s[threadIdx.x];
  if (threadIdx.x == 0)
    flag[u] = 1;
}

} // namespace gm
