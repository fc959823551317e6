// Exact insertion screened by BF16 MFMA distance tiles.
//
// Same result as the FP32 tiles of exact_kernels.hip -- every inserted row
// gets the exact 32 nearest rows by the canonical FP32 distance, every earlier
// row is offered every new row -- with the distance work moved to the matrix
// cores:
//   * each row x is stored as two BF16 vectors hi = bf16(x), lo = bf16(x - hi)
//     (x - hi - lo is below 2^-16 |x|) plus its FP32 squared norm;
//   * a 64 x 64 tile of approximate squared distances
//       d~ = |u|^2 + |j|^2 - 2 (hi_u.hi_j + hi_u.lo_j + lo_u.hi_j)
//     takes three v_mfma_f32_32x32x16_bf16 per 16 features (FP32
//     accumulation);
//   * |d~ - d| <= bound(u, j) = coef * (|u|^2 + |j|^2) where d is the
//     canonical distance and coef covers the BF16 split, any FP32 summation
//     order of the 3D products, the norm and d~ roundings and the canonical
//     distance's own rounding (derived on the host, engine.hip);
//   * pass 1 keeps, per batch row and column split, the 32 smallest upper
//     bounds d~ + bound; their 32nd smallest over all splits, tau, is at least
//     the row's 32nd smallest exact distance;
//   * pass 2 recomputes the tiles and re-ranks in exact FP32 every pair whose
//     lower bound d~ - bound is <= tau (forward candidates) or <= the earlier
//     row's pool maximum (reverse candidates).  No pair that can enter a pool
//     is screened out, so the pools equal the FP32 path's bit for bit.
// Rows are read straight from HBM/L2 into the MFMA operand registers (16 B
// per lane per 16 features); the FP32 rows are read only for the re-ranked
// pairs.
#include "common.hpp"
#include "kernels.hpp"

namespace gm {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

template <typename T>
__device__ void block_bitonic_m(T *a, int len, int seg) {
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

constexpr int MR = 64;    // batch rows per block
constexpr int MC = 64;    // columns per tile
constexpr int MB = 64;    // per-row key buffer: W kept + 32 appended per phase
constexpr int MS = 4096;  // survivor queue (a full tile)
constexpr int MQ = 1024;  // reverse offers per drain round

// per-row top-W buffers: append a key if below the row's threshold
__device__ inline void row_append(uint64_t *buf, int *cnt, const uint64_t *th,
                                  int *need, int r, uint64_t key) {
  if (key < th[r]) {
    const int slot = atomicAdd(&cnt[r], 1);
    buf[r * MB + slot] = key;
    if (slot + 1 > W)
      *need = 1;
  }
}

// sort every row's buffer and keep the W smallest (all threads)
__device__ inline void row_compact(uint64_t *buf, int *cnt, uint64_t *th) {
  for (int e = threadIdx.x; e < MR * MB; e += blockDim.x)
    if ((e % MB) >= cnt[e / MB])
      buf[e] = EMPTY;
  __syncthreads();
  block_bitonic_m(buf, MR * MB, MB);
  if (threadIdx.x < MR && cnt[threadIdx.x] >= W) {
    cnt[threadIdx.x] = W;
    th[threadIdx.x] = buf[threadIdx.x * MB + W - 1];
  }
  __syncthreads();
}

// after a round of row_append (all threads): compact the rows if any buffer
// went past W.  The flag is read by every thread before anyone can set it
// again.
__device__ inline void settle_rows(uint64_t *buf, int *cnt, uint64_t *th,
                                   int *need) {
  __syncthreads();
  const int nd = *need;
  __syncthreads();
  if (nd) {
    row_compact(buf, cnt, th);
    if (threadIdx.x == 0)
      *need = 0;
    __syncthreads();
  }
}

} // namespace

// hi/lo BF16 split and FP32 squared norm of the batch rows.  One 64-thread
// block per row; the norm is one sequential FMA chain.
__global__ __launch_bounds__(64) void k_bf16_prep(
    const float *__restrict__ X, uint64_t dp, uint64_t db,
    __bf16 *__restrict__ Xh, __bf16 *__restrict__ Xl, float *__restrict__ nrm,
    uint64_t b0, uint64_t nb) {
  if (blockIdx.x >= nb)
    return;
  const uint64_t r = b0 + blockIdx.x;
  const float *x = X + r * dp;
  for (uint64_t i = threadIdx.x; i < db; i += blockDim.x) {
    const float v = i < dp ? x[i] : 0.0f;
    const __bf16 h = (__bf16)v;
    Xh[r * db + i] = h;
    Xl[r * db + i] = (__bf16)(v - (float)h);
  }
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (uint64_t i = 0; i < dp; ++i)
      s = __builtin_fmaf(x[i], x[i], s);
    nrm[r] = s;
  }
}

template <int PASS>
__global__ __launch_bounds__(256) void k_bf16_tiles(
    const float *__restrict__ X, uint64_t dp, const __bf16 *__restrict__ Xh,
    const __bf16 *__restrict__ Xl, uint64_t db, const float *__restrict__ nrm,
    uint64_t *__restrict__ pool, const float *__restrict__ pmax,
    const float *__restrict__ tau, uint64_t b0, uint64_t nb, uint64_t n, int k,
    int nsplit, uint64_t *__restrict__ part, uint8_t *__restrict__ flag,
    float coef, float sfac) {
  __shared__ uint64_t s_buf[MR * MB];
  __shared__ uint64_t s_th[MR];
  __shared__ int s_cnt[MR];
  __shared__ float s_nu[MR], s_tau[MR], s_nj[MC], s_pm[MC];
  __shared__ uint32_t s_sq[MS];
  __shared__ uint64_t s_qkey[MQ];
  __shared__ uint8_t s_qtgt[MQ];
  __shared__ int s_sn, s_qn, s_need;

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = wave & 1, wn = wave >> 1;
  const int hw = tid >> 5, hl = tid & 31, half = (tid >> 5) & 1;
  const uint64_t r0 = (uint64_t)blockIdx.x * MR;
  const int rows = (int)min((uint64_t)MR, nb - r0);
  const int split = blockIdx.y;
  const uint64_t ntile = (n + MC - 1) / MC;

  if (tid < MR) {
    s_cnt[tid] = 0;
    s_th[tid] = EMPTY;
    s_nu[tid] = tid < rows ? nrm[b0 + r0 + tid] : 0.0f;
    s_tau[tid] = (PASS == 2 && tid < rows) ? tau[r0 + tid] : 0.0f;
  }
  // this lane's MFMA operand rows
  const int arow = wm * 32 + (lane & 31);
  const bool a_ok = arow < rows;
  const uint64_t ga = b0 + r0 + (a_ok ? arow : 0);
  const __bf16 *pah = Xh + ga * db + (lane >> 5) * 8;
  const __bf16 *pal = Xl + ga * db + (lane >> 5) * 8;
  const bf16x8 zero8 = {};
  __syncthreads();

  for (uint64_t t = split; t < ntile; t += nsplit) {
    const uint64_t c0 = t * MC;
    const int cols = (int)min((uint64_t)MC, n - c0);
    if (tid < MC) {
      const uint64_t j = c0 + tid;
      s_nj[tid] = tid < cols ? nrm[j] : 0.0f;
      s_pm[tid] = (PASS == 2 && tid < cols && j < b0) ? pmax[j] : -1.0f;
    }
    if (tid == 0) {
      s_sn = 0;
      s_need = 0;
    }
    // ---- approximate distances of the wave's 32 x 32 sub-tile
    const int bcol = wn * 32 + (lane & 31);
    const bool b_ok = bcol < cols;
    const uint64_t gb = c0 + (b_ok ? bcol : 0);
    const __bf16 *pbh = Xh + gb * db + (lane >> 5) * 8;
    const __bf16 *pbl = Xl + gb * db + (lane >> 5) * 8;
    floatx16 acc = {};
    for (uint64_t k0 = 0; k0 < db; k0 += 16) {
      const bf16x8 ah = a_ok ? *reinterpret_cast<const bf16x8 *>(pah + k0) : zero8;
      const bf16x8 al = a_ok ? *reinterpret_cast<const bf16x8 *>(pal + k0) : zero8;
      const bf16x8 bh = b_ok ? *reinterpret_cast<const bf16x8 *>(pbh + k0) : zero8;
      const bf16x8 bl = b_ok ? *reinterpret_cast<const bf16x8 *>(pbl + k0) : zero8;
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
    }
    __syncthreads();

    const int lc = wn * 32 + (lane & 31);
    const uint64_t j = c0 + lc;
    if (PASS == 1) {
      // two phases (one per column half) keep each row's appends <= 32
      for (int ph = 0; ph < 2; ++ph) {
        if (wn == ph) {
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const int lr = wm * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
            const uint64_t u = b0 + r0 + lr;
            if (lr >= rows || lc >= cols || j == u)
              continue;
            const float s = s_nu[lr] + s_nj[lc];
            const float dt = __builtin_fmaf(-2.0f, acc[i], s);
            const float ub = fmaxf(dt + coef * (s * sfac) + 1e-30f, 0.0f);
            row_append(s_buf, s_cnt, s_th, &s_need, lr,
                       ((uint64_t)__float_as_uint(ub) << 32) | (uint64_t)j);
          }
        }
        settle_rows(s_buf, s_cnt, s_th, &s_need);
      }
    } else {
      // ---- survivors: lower bound within the row's tau (forward) or the
      // earlier row's pool maximum (reverse)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int lr = wm * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        const uint64_t u = b0 + r0 + lr;
        if (lr >= rows || lc >= cols || j == u)
          continue;
        const float s = s_nu[lr] + s_nj[lc];
        const float dt = __builtin_fmaf(-2.0f, acc[i], s);
        const float lb = dt - coef * (s * sfac) - 1e-30f;
        const bool fwd = lb <= s_tau[lr];
        const bool rev = lb <= s_pm[lc];
        if (fwd || rev) {
          const int at = atomicAdd(&s_sn, 1);
          s_sq[at] = ((uint32_t)lr << 16) | ((uint32_t)lc << 2) |
                     (fwd ? 1u : 0u) | (rev ? 2u : 0u);
        }
      }
      __syncthreads();
      const int sn = s_sn;
      // exact re-rank in rounds of MQ survivors: forward keys into the row
      // buffers (two column phases), reverse offers through the CAS insert
      for (int base = 0; base < sn; base += MQ) {
        const int cnt = min(MQ, sn - base);
        if (tid == 0)
          s_qn = 0;
        __syncthreads();
        for (int ph = 0; ph < 2; ++ph) {
          for (int q = tid; q < cnt; q += blockDim.x) {
            const uint32_t e = s_sq[base + q];
            const int lr = (int)(e >> 16), lcq = (int)((e >> 2) & 0x3fff);
            if ((lcq >> 5) != ph)
              continue;
            const uint64_t u = b0 + r0 + lr, jj = c0 + lcq;
            const float dc = canon_dist_g(X + u * dp, X + jj * dp, (int)dp);
            if (e & 1u)
              row_append(s_buf, s_cnt, s_th, &s_need, lr,
                         make_key(dc, (uint32_t)jj, false));
            if ((e & 2u) && dc <= s_pm[lcq]) {
              const int at = atomicAdd(&s_qn, 1);
              s_qkey[at] = make_key(dc, (uint32_t)u, false);
              s_qtgt[at] = (uint8_t)lcq;
            }
          }
          settle_rows(s_buf, s_cnt, s_th, &s_need);
        }
        const int qn = s_qn;
        for (int e = hw; e < qn; e += 8) {
          const uint64_t jj = c0 + s_qtgt[e];
          if (pool_offer_half(pool + jj * W, s_qkey[e], k, hl, half) && hl == 0)
            flag[jj] = 1;
        }
        __syncthreads();
      }
    }
  }

  // this split's W smallest keys per row, sorted
  row_compact(s_buf, s_cnt, s_th);
  for (int e = tid; e < rows * W; e += blockDim.x) {
    const int r = e / W, s = e % W;
    part[((r0 + r) * (uint64_t)nsplit + split) * W + s] = s_buf[r * MB + s];
  }
}

template __global__ void k_bf16_tiles<1>(
    const float *, uint64_t, const __bf16 *, const __bf16 *, uint64_t,
    const float *, uint64_t *, const float *, const float *, uint64_t,
    uint64_t, uint64_t, int, int, uint64_t *, uint8_t *, float, float);
template __global__ void k_bf16_tiles<2>(
    const float *, uint64_t, const __bf16 *, const __bf16 *, uint64_t,
    const float *, uint64_t *, const float *, const float *, uint64_t,
    uint64_t, uint64_t, int, int, uint64_t *, uint8_t *, float, float);

// tau of one batch row: the W-th smallest upper bound over all splits
// (+inf when the row has fewer than W other rows).
__global__ __launch_bounds__(256) void k_tau_merge(
    const uint64_t *__restrict__ part, int nsplit, float *__restrict__ tau,
    uint64_t nb) {
  __shared__ uint64_t s[4096];
  const uint64_t bi = blockIdx.x;
  if (bi >= nb)
    return;
  const int len = nsplit * W;
  for (int e = threadIdx.x; e < len; e += blockDim.x)
    s[e] = part[bi * (uint64_t)len + e];
  __syncthreads();
  block_bitonic_m(s, len, len);
  if (threadIdx.x == 0) {
    const uint64_t kw = s[W - 1];
    tau[bi] = kw == EMPTY ? INFINITY : __uint_as_float((uint32_t)(kw >> 32));
  }
}

} // namespace gm
