// NN-Descent insertion and graph update on gfx950 (north-star subsystem (a)).
//
// Reference: cpp/src/kernels/insertion_kernel.cuh (insertion) and
// cpp/src/kernels/update_kernel.cuh (update).  Where the reference moves the
// graph to the host after every step (insertion_kernel.cuh:1023-1089,
// update_kernel.cuh:940-1051), every structure here stays in HBM: the pools,
// the sampled lists, the reverse lists and the owner list are device arrays
// and the host only receives the ascending list of changed rows.
#include "common.hpp"
#include "kernels.hpp"

namespace gm {

namespace {

// Bitonic sort of `len` (power of two) LDS elements in independent ascending
// segments of `seg` elements; every thread of the block must call it.
template <typename T>
__device__ void block_bitonic(T *a, int len, int seg) {
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

} // namespace

// ----------------------------------------------------------------- rows
// Cosine rows are normalised in place (reference preprocess_l2_norms_kernel,
// insertion_kernel.cuh:31-81); the norm is one sequential FMA chain so the
// result matches the CPU restatement bit for bit.
__global__ __launch_bounds__(64) void k_prep_rows(float *X, uint64_t dp, uint64_t d, uint64_t b0,
                            uint64_t nb) {
  const uint64_t r = b0 + blockIdx.x;
  if (blockIdx.x >= nb)
    return;
  float *x = X + r * dp;
  __shared__ float s_norm;
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (uint64_t i = 0; i < dp; ++i)
      s = __builtin_fmaf(x[i], x[i], s);
    s_norm = sqrt_rn(s);
  }
  __syncthreads();
  const float nrm = s_norm;
  if (nrm > 0.0f)
    for (uint64_t i = threadIdx.x; i < d; i += blockDim.x)
      x[i] = div_rn(x[i], nrm);
}

// Random initial neighbours (reference initialize_random_neighbors_kernel,
// insertion_kernel.cuh:170-248): half from the recent window, the rest from
// all rows, distinct, no self.  One thread per batch row writes ids; the
// distances follow in k_init_dist.
__global__ __launch_bounds__(64) void k_random_ids(uint32_t *ids, int *cnt, uint64_t b0,
                             uint64_t nb, uint64_t n, uint32_t window_start,
                             uint32_t seed) {
  const uint64_t bi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (bi >= nb)
    return;
  const uint32_t u = (uint32_t)(b0 + bi);
  uint32_t *out = ids + bi * W;
  if (n <= 1) {
    cnt[bi] = 0;
    return;
  }
  Rng rng(seed, u, 0, TAG_INIT_NBR);
  const int max_fill = (int)min((uint64_t)W, n - 1);
  const int window_size = (int)n - (int)window_start;
  const int quota = min(W / 2, max(0, window_size - 1));
  int filled = 0, attempts = 0;
  while (filled < quota && attempts < quota * 16) {
    ++attempts;
    const uint32_t c = window_start + rng.next() % (uint32_t)window_size;
    bool dup = c == u;
    for (int q = 0; q < filled && !dup; ++q)
      dup = out[q] == c;
    if (!dup)
      out[filled++] = c;
  }
  while (filled < max_fill) {
    const uint32_t c = rng.next() % (uint32_t)n;
    bool dup = c == u;
    for (int q = 0; q < filled && !dup; ++q)
      dup = out[q] == c;
    if (!dup)
      out[filled++] = c;
  }
  cnt[bi] = filled;
}

// One thread per (batch row, slot): canonical distance and pool key.
__global__ void k_init_dist(const float *__restrict__ X, uint64_t dp,
                            uint64_t *__restrict__ pool,
                            const uint32_t *__restrict__ ids,
                            const int *__restrict__ cnt, uint64_t b0,
                            uint64_t nb) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * W)
    return;
  const uint64_t bi = t / W;
  const int s = (int)(t % W);
  const uint64_t u = b0 + bi;
  uint64_t key = EMPTY;
  if (s < cnt[bi]) {
    const uint32_t v = ids[t];
    key = make_key(canon_dist_g(X + u * dp, X + (uint64_t)v * dp, (int)dp), v,
                   true);
  }
  pool[u * W + s] = key;
}

// Neighbours-of-neighbours candidates of one batch row (reference
// sample_forward_candidates_kernel, insertion_kernel.cuh:250-321): slot-major
// over the row's k nearest neighbours' sorted pools, skipping self, current
// pool members and repeats, capped at CAND in that order; then the canonical
// distance of each candidate.  256 threads per batch row.
__global__ __launch_bounds__(256) void k_ins_candidates(
    const float *__restrict__ X, uint64_t dp, const uint64_t *__restrict__ pool,
    uint64_t b0, uint64_t nb, int k, uint32_t *__restrict__ cand,
    float *__restrict__ cdist, int *__restrict__ ccount) {
  __shared__ uint64_t s_nb[W * W];    // neighbours' pools, sorted per row
  __shared__ uint64_t s_sort[1024];   // (id << 10 | position)
  __shared__ uint32_t s_ids[1024];    // id at each position
  __shared__ int s_surv[1024];
  __shared__ uint64_t s_own[W];
  __shared__ uint32_t s_top[W];
  __shared__ int s_nbc[W];
  __shared__ int s_kk, s_total;
  const uint64_t bi = blockIdx.x;
  if (bi >= nb)
    return;
  const uint32_t u = (uint32_t)(b0 + bi);
  const int tid = threadIdx.x;

  if (tid < W)
    s_own[tid] = pool[(uint64_t)u * W + tid];
  __syncthreads();
  block_bitonic(s_own, W, W);
  if (tid == 0) {
    int cu = 0;
    while (cu < W && s_own[cu] != EMPTY)
      ++cu;
    s_kk = min(k, cu);
  }
  __syncthreads();
  const int kk = s_kk;
  if (tid < kk)
    s_top[tid] = key_id(s_own[tid]);
  __syncthreads();
  for (int t = tid; t < W * W; t += blockDim.x) {
    const int j = t / W, s = t % W;
    s_nb[t] = j < kk ? pool[(uint64_t)s_top[j] * W + s] : EMPTY;
  }
  __syncthreads();
  block_bitonic(s_nb, W * W, W);
  if (tid < W) {
    int c = 0;
    while (c < W && s_nb[tid * W + c] != EMPTY)
      ++c;
    s_nbc[tid] = c;
  }
  __syncthreads();

  const int P = k * kk;
  int P2 = 1;
  while (P2 < P)
    P2 <<= 1;
  for (int p = tid; p < P2; p += blockDim.x) {
    uint64_t e = EMPTY;
    uint32_t id = NO_ID;
    if (p < P) {
      const int s = p / kk, j = p % kk;
      if (s < s_nbc[j]) {
        id = key_id(s_nb[j * W + s]);
        bool skip = id == u;
        for (int q = 0; q < W && !skip; ++q)
          skip = s_own[q] != EMPTY && key_id(s_own[q]) == id;
        if (!skip)
          e = ((uint64_t)id << 10) | (uint64_t)p;
      }
    }
    s_sort[p] = e;
    if (p < 1024) {
      s_ids[p] = id;
      s_surv[p] = 0;
    }
  }
  __syncthreads();
  if (P2 > 1)
    block_bitonic(s_sort, P2, P2);
  for (int i = tid; i < P2; i += blockDim.x) {
    const uint64_t e = s_sort[i];
    if (e == EMPTY)
      continue;
    if (i == 0 || (s_sort[i - 1] >> 10) != (e >> 10))
      s_surv[e & 1023] = 1;
  }
  __syncthreads();
  // exclusive scan of the survivor flags in position order (P <= 1024)
  if (tid == 0) {
    int run = 0;
    for (int p = 0; p < P; ++p) {
      const int f = s_surv[p];
      s_surv[p] = f ? run : -1;
      run += f;
    }
    s_total = min(run, CAND);
    ccount[bi] = s_total;
  }
  __syncthreads();
  for (int p = tid; p < P; p += blockDim.x) {
    const int r = s_surv[p];
    if (r >= 0 && r < CAND)
      cand[bi * CAND + r] = s_ids[p];
  }
  __syncthreads();
  const int total = s_total;
  for (int q = tid; q < total; q += blockDim.x) {
    const uint32_t c = cand[bi * CAND + q];
    cdist[bi * CAND + q] =
        canon_dist_g(X + (uint64_t)u * dp, X + (uint64_t)c * dp, (int)dp);
  }
}

// Forward and reverse inserts of the candidates (reference
// update_forward_neighbors_kernel :486-534 and the reverse scatter/update
// :548-640); each of the 8 half-waves of the block handles every 8th
// candidate.  flag[row] = 1 when the row's pool changed.
__global__ __launch_bounds__(256) void k_ins_offers(
    uint64_t *__restrict__ pool, uint8_t *__restrict__ flag,
    const uint32_t *__restrict__ cand, const float *__restrict__ cdist,
    const int *__restrict__ ccount, uint64_t b0, uint64_t nb, int k) {
  const uint64_t bi = blockIdx.x;
  if (bi >= nb)
    return;
  const uint32_t u = (uint32_t)(b0 + bi);
  const int hw = threadIdx.x >> 5, hl = threadIdx.x & 31;
  const int half = (threadIdx.x >> 5) & 1;
  const int cc = ccount[bi];
  for (int q = hw; q < cc; q += 8) {
    const uint32_t c = cand[bi * CAND + q];
    const float d = cdist[bi * CAND + q];
    int r = pool_offer_half(pool + (uint64_t)u * W, make_key(d, c, true), k,
                            hl, half);
    if (r && hl == 0)
      flag[u] = 1;
    r = pool_offer_half(pool + (uint64_t)c * W, make_key(d, u, true), k, hl,
                        half);
    if (r && hl == 0)
      flag[c] = 1;
  }
}

// ---------------------------------------------------------------- update
// Sampling (reference sample_graph, update_kernel.cuh:750-807 with the
// changes of the CPU restatement): every new entry of the row's pool goes to
// its forward-new list and is marked old; every old entry goes to its
// forward-old list.  One 32-lane half per row.
__global__ void k_sample(uint64_t *__restrict__ pool,
                         const uint32_t *__restrict__ rows, uint64_t nrows,
                         uint32_t *__restrict__ fn, uint32_t *__restrict__ fo,
                         int *__restrict__ fnc, int *__restrict__ foc,
                         uint32_t *__restrict__ newmask,
                         uint32_t *__restrict__ owner_mark) {
  const uint64_t g = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
  const int hl = threadIdx.x & 31;
  const int base = (threadIdx.x & 32);
  if (g >= nrows)
    return;
  const uint32_t u = rows[g];
  uint64_t *row = pool + (uint64_t)u * W;
  const uint64_t v = row[hl];
  const bool isnew = v != EMPTY && (v & NEW_BIT);
  const bool isold = v != EMPTY && !(v & NEW_BIT);
  const uint32_t nm = (uint32_t)((__ballot(isnew) >> base) & 0xffffffffull);
  const uint32_t om = (uint32_t)((__ballot(isold) >> base) & 0xffffffffull);
  const uint32_t below = (1u << hl) - 1u;
  if (isnew) {
    fn[(uint64_t)u * W + __popc(nm & below)] = key_id(v);
    row[hl] = v & ~NEW_BIT;
  }
  if (isold)
    fo[(uint64_t)u * W + __popc(om & below)] = key_id(v);
  if (hl == 0) {
    fnc[u] = __popc(nm);
    foc[u] = __popc(om);
    newmask[u] = nm;
    owner_mark[u] = 1;
  }
}

// Reverse lists: each sampled entry (u -> v, d) offers (d, u) to v's reverse
// new/old list, a bounded top-W keeping the nearest reverse neighbours
// (reference add_rev_edges_kernel :122-146 keeps whichever won an atomicAdd).
// A row that receives its first reverse entry joins the owner list.
__global__ void k_reverse(const uint64_t *__restrict__ pool,
                          const uint32_t *__restrict__ rows, uint64_t nrows,
                          const uint32_t *__restrict__ newmask,
                          uint64_t *__restrict__ rev_new,
                          uint64_t *__restrict__ rev_old,
                          uint32_t *__restrict__ owner_mark,
                          uint32_t *__restrict__ owners,
                          uint32_t *__restrict__ owner_count) {
  const uint64_t g = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
  const int hl = threadIdx.x & 31;
  const int half = (threadIdx.x >> 5) & 1;
  if (g >= nrows)
    return;
  const uint32_t u = rows[g];
  const uint64_t v = pool[(uint64_t)u * W + hl];
  const uint32_t nm = newmask[u];
  const uint64_t live = __ballot(v != EMPTY) >> (half * 32);
  for (int j = 0; j < W; ++j) {
    if (!((live >> j) & 1ull))
      continue;
    const uint64_t vj = __shfl(v, j, 32);
    const uint32_t t = key_id(vj);
    uint64_t *dst = ((nm >> j) & 1u) ? rev_new : rev_old;
    pool_offer_half(dst + (uint64_t)t * W, make_key(key_dist(vj), u, false), 0,
                    hl, half);
    if (hl == 0) {
      const uint32_t was = __hip_atomic_fetch_or(
          owner_mark + t, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (!was) {
        const uint32_t at = __hip_atomic_fetch_add(
            owner_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        owners[at] = t;
      }
    }
  }
}

// Local join (reference local_join_kernel, update_kernel.cuh:377-687).  One
// block of 256 threads per owner v:
//  1. new = fwd_new(v) U rev_new(v), old = (fwd_old(v) U rev_old(v)) \ new,
//     de-duplicated by a wave bitonic sort (<= 64 each);
//  2. the members' pools are loaded once (sorted ids + maximum key) to filter
//     offers that cannot change a pool;
//  3. new x (new U old) distances in 32-feature LDS chunks; thread (rt, ct)
//     owns rows rt + 16 i and columns ct + 16 j (conflict-free LDS reads) and
//     one FMA chain per pair -- the canonical distance;
//  4. pairs new x new (c > r) and new x old offer both directions; the
//     surviving offers are queued in LDS and inserted by the 8 half-waves.
// flag[row] = 1 when an insert lands in the row's top k.
constexpr int JQ = 4096;
__global__ __launch_bounds__(256) void k_join(
    const float *__restrict__ X, uint64_t dp, uint64_t *__restrict__ pool,
    const uint32_t *__restrict__ fn, const uint32_t *__restrict__ fo,
    const int *__restrict__ fnc, const int *__restrict__ foc,
    const uint64_t *__restrict__ rev_new, const uint64_t *__restrict__ rev_old,
    const uint32_t *__restrict__ owners,
    const uint32_t *__restrict__ owner_count, uint8_t *__restrict__ flag,
    int k, unsigned long long *__restrict__ stats) {
  __shared__ uint32_t s_ids[128];
  __shared__ uint32_t s_mid[128 * W];  // members' pool ids, sorted
  __shared__ uint64_t s_mmax[128];
  __shared__ float s_vec[128 * 33];
  __shared__ uint64_t s_qkey[JQ];
  __shared__ uint8_t s_qtgt[JQ];
  __shared__ int s_a, s_o, s_qn;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int hw = tid >> 5, hl = tid & 31, half = (tid >> 5) & 1;
  const int rt = tid >> 4, ct = tid & 15;
  const uint32_t n_owner = *owner_count;
  // work counters of this block: list members, pairs, queued offers
  unsigned long long c_members = 0, c_pairs = 0, c_offers = 0;

  for (uint32_t oi = blockIdx.x; oi < n_owner; oi += gridDim.x) {
    const uint32_t v = owners[oi];
    // ---- 1. lists
    if (wave == 0) {
      const int nf = fnc[v];
      uint32_t x = NO_ID;
      if (lane < 32) {
        if (lane < nf)
          x = fn[(uint64_t)v * W + lane];
      } else {
        const uint64_t rk = rev_new[(uint64_t)v * W + lane - 32];
        if (rk != EMPTY)
          x = key_id(rk);
      }
      x = wave_sort64(x);
      const uint32_t prev = __shfl_up(x, 1);
      const bool keep = x != NO_ID && (lane == 0 || prev != x);
      const uint64_t km = __ballot(keep);
      if (keep)
        s_ids[__popcll(km & ((1ull << lane) - 1ull))] = x;
      if (lane == 0)
        s_a = __popcll(km);
    } else if (wave == 1) {
      const int nf = foc[v];
      uint32_t x = NO_ID;
      if (lane < 32) {
        if (lane < nf)
          x = fo[(uint64_t)v * W + lane];
      } else {
        const uint64_t rk = rev_old[(uint64_t)v * W + lane - 32];
        if (rk != EMPTY)
          x = key_id(rk);
      }
      x = wave_sort64(x);
      const uint32_t prev = __shfl_up(x, 1);
      const bool keep = x != NO_ID && (lane == 0 || prev != x);
      const uint64_t km = __ballot(keep);
      if (keep)
        s_mid[__popcll(km & ((1ull << lane) - 1ull))] = x;
      if (lane == 0)
        s_o = __popcll(km);
    }
    __syncthreads();
    const int a = s_a;
    if (a == 0) {
      __syncthreads();
      continue;
    }
    if (wave == 1) { // drop old entries that are also new
      const int o = s_o;
      uint32_t x = lane < o ? s_mid[lane] : NO_ID;
      bool keep = x != NO_ID;
      if (keep) {
        int lo = 0, hi = a;
        while (lo < hi) {
          const int m = (lo + hi) >> 1;
          const uint32_t y = s_ids[m];
          if (y == x) {
            keep = false;
            break;
          }
          if (y < x)
            lo = m + 1;
          else
            hi = m;
        }
      }
      const uint64_t km = __ballot(keep);
      if (keep)
        s_ids[a + __popcll(km & ((1ull << lane) - 1ull))] = x;
      if (lane == 0)
        s_o = __popcll(km);
    }
    __syncthreads();
    const int o = s_o;
    const int m_all = a + o;
    c_members += (unsigned long long)m_all;
    c_pairs += (unsigned long long)(a * (a - 1) / 2 + a * o);
    // ---- 2. members' pools
    for (int t = tid; t < 128 * W; t += blockDim.x) {
      const int m = t / W, s = t % W;
      uint32_t id = NO_ID;
      if (m < m_all) {
        const uint64_t key = pool[(uint64_t)s_ids[m] * W + s];
        id = key == EMPTY ? NO_ID : key_id(key);
      }
      s_mid[t] = id;
    }
    if (tid < 128) {
      uint64_t mx = 0;
      if (tid < m_all) {
        const uint64_t *r = pool + (uint64_t)s_ids[tid] * W;
        for (int s = 0; s < W; ++s) {
          const uint64_t key = r[s];
          mx = key > mx ? key : mx;
        }
      }
      s_mmax[tid] = mx;
    }
    __syncthreads();
    block_bitonic(s_mid, 128 * W, W);

    // ---- 3. distances
    float acc[4][8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j)
        acc[i][j] = 0.0f;
    for (uint64_t base = 0; base < dp; base += 32) {
      const int fw = (int)min((uint64_t)32, dp - base);
      for (int e = tid; e < m_all * 8; e += blockDim.x) {
        const int m = e >> 3, f4 = (e & 7) * 4;
        float4 val = make_float4(0.f, 0.f, 0.f, 0.f);
        if (f4 < fw)
          val = *reinterpret_cast<const float4 *>(
              X + (uint64_t)s_ids[m] * dp + base + f4);
        float *dst = s_vec + m * 33 + f4;
        dst[0] = val.x;
        dst[1] = val.y;
        dst[2] = val.z;
        dst[3] = val.w;
      }
      __syncthreads();
      if (rt < a && ct < m_all) {
        for (int f = 0; f < fw; ++f) {
          float rv[4], cv[8];
#pragma unroll
          for (int i = 0; i < 4; ++i)
            rv[i] = s_vec[(rt + 16 * i) * 33 + f];
#pragma unroll
          for (int j = 0; j < 8; ++j)
            cv[j] = s_vec[(ct + 16 * j) * 33 + f];
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              const float t = rv[i] - cv[j];
              acc[i][j] = __builtin_fmaf(t, t, acc[i][j]);
            }
        }
      }
      __syncthreads();
    }

    // ---- 4. offers, one row-group of the thread's tile per round
    for (int i = 0; i < 4; ++i) {
      if (tid == 0)
        s_qn = 0;
      __syncthreads();
      const int r = rt + 16 * i;
      if (r < a) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int c = ct + 16 * j;
          if (c >= m_all || (c < a && c <= r))
            continue;
          const float d = acc[i][j];
          const uint32_t idr = s_ids[r], idc = s_ids[c];
          // r <- c
          uint64_t key = make_key(d, idc, true);
          if (key < s_mmax[r]) {
            const uint32_t *ids = s_mid + r * W;
            int lo = 0, hi = W;
            while (lo < hi) {
              const int m = (lo + hi) >> 1;
              if (ids[m] < idc)
                lo = m + 1;
              else
                hi = m;
            }
            if (lo == W || ids[lo] != idc) {
              const int at = atomicAdd(&s_qn, 1);
              s_qkey[at] = key;
              s_qtgt[at] = (uint8_t)r;
            }
          }
          // c <- r
          key = make_key(d, idr, true);
          if (key < s_mmax[c]) {
            const uint32_t *ids = s_mid + c * W;
            int lo = 0, hi = W;
            while (lo < hi) {
              const int m = (lo + hi) >> 1;
              if (ids[m] < idr)
                lo = m + 1;
              else
                hi = m;
            }
            if (lo == W || ids[lo] != idr) {
              const int at = atomicAdd(&s_qn, 1);
              s_qkey[at] = key;
              s_qtgt[at] = (uint8_t)c;
            }
          }
        }
      }
      __syncthreads();
      const int qn = s_qn;
      c_offers += (unsigned long long)qn;
      for (int e = hw; e < qn; e += 8) {
        const uint32_t tgt = s_ids[s_qtgt[e]];
        const int res = pool_offer_half(pool + (uint64_t)tgt * W, s_qkey[e], k,
                                        hl, half);
        if (res == 2 && hl == 0)
          flag[tgt] = 1;
      }
      __syncthreads();
    }
  }
  if (tid == 0 && stats && c_members) {
    atomicAdd(stats + 0, c_members);
    atomicAdd(stats + 1, c_pairs);
    atomicAdd(stats + 2, c_offers);
  }
}

// Reset the per-pass state of the owners (reverse lists, marks, counts).
__global__ void k_join_reset(const uint32_t *__restrict__ owners,
                             const uint32_t *__restrict__ owner_count,
                             uint64_t *__restrict__ rev_new,
                             uint64_t *__restrict__ rev_old,
                             uint32_t *__restrict__ owner_mark,
                             int *__restrict__ fnc, int *__restrict__ foc) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n_owner = *owner_count;
  if (t >= n_owner * W)
    return;
  const uint32_t v = owners[t / W];
  const int s = (int)(t % W);
  rev_new[(uint64_t)v * W + s] = EMPTY;
  rev_old[(uint64_t)v * W + s] = EMPTY;
  if (s == 0) {
    owner_mark[v] = 0;
    fnc[v] = 0;
    foc[v] = 0;
  }
}

// ------------------------------------------------------ public kNN rows
// Sorted top-k of each listed row's pool into the public (n x k) arrays
// (reference stage_and_scatter_knn_rows, matrix_utils.cuh:127-174, without
// the host staging).  Counts the ids that were not in the previous row and
// records the pool's maximum distance.
// Public top-k rows of the listed rows from their pools (one half-wave per
// row), the fresh-edge count, the pool maximum (infinite while the pool has
// an empty slot, the reverse filter of exact insertion), and the row's
// sigma / rho for the fuzzy set, which depend on the public row only.
__global__ __launch_bounds__(256) void k_sync_rows(
    const uint64_t *__restrict__ pool, const uint32_t *__restrict__ rows,
    uint64_t nrows, const uint32_t *__restrict__ nrows_dev,
    uint32_t *__restrict__ pidx, float *__restrict__ pdist, int k,
    unsigned long long *__restrict__ fresh, float *__restrict__ pmax,
    int euclid, double *__restrict__ sigma, double *__restrict__ rho) {
  __shared__ unsigned long long s_fresh;
  if (threadIdx.x == 0)
    s_fresh = 0;
  __syncthreads();
  // the row count is nrows, or the device count when one is given (then
  // nrows only bounds it); half-waves stride over the rows
  const uint64_t cnt =
      nrows_dev ? min(nrows, (uint64_t)*nrows_dev) : nrows;
  const uint64_t stride = ((uint64_t)gridDim.x * blockDim.x) >> 5;
  const int hl = threadIdx.x & 31;
  const int half = (threadIdx.x >> 5) & 1;
  for (uint64_t g = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
       g < cnt; g += stride) { // uniform over the half-wave
    const uint32_t u = rows[g];
    const uint64_t key = half_sort32(pool[(uint64_t)u * W + hl]);
    const uint32_t old = hl < k ? pidx[(uint64_t)u * k + hl] : NO_ID;
    const uint32_t id = key == EMPTY ? NO_ID : key_id(key);
    bool seen = false;
    for (int q = 0; q < k; ++q)
      seen |= (__shfl(old, q, 32) == id);
    const bool is_fresh = hl < k && id != NO_ID && !seen;
    const uint64_t fm =
        (__ballot(is_fresh) >> (threadIdx.x & 32)) & 0xffffffffull;
    const float dv = key == EMPTY ? 3.402823466e38f : key_dist(key);
    if (hl < k) {
      pidx[(uint64_t)u * k + hl] = id;
      pdist[(uint64_t)u * k + hl] = dv;
    }
    if (fresh && hl == 0 && fm)
      atomicAdd(&s_fresh, (unsigned long long)__popcll(fm));
    if (hl == 31)
      pmax[u] = key == EMPTY ? INFINITY : key_dist(key);
    // embed-time entry c: the self column (0), then public slot c - 1
    const float vs = __shfl(dv, hl == 0 ? 0 : hl - 1, 32);
    double di = 0.0;
    if (hl > 0 && hl < k) {
      const double d = embed_d(vs, euclid != 0);
      di = valid_embed_d(d) ? d : 0.0;
    }
    double sg, rh;
    row_sigma_rho(di, hl, k, half, sg, rh);
    if (hl == 0) {
      sigma[u] = sg;
      rho[u] = rh;
    }
  }
  __syncthreads();
  if (fresh && threadIdx.x == 0 && s_fresh)
    atomicAdd(fresh, s_fresh);
}

// ------------------------------------------------------------ compaction
// Ascending list of rows in [lo, hi) with flag != 0 or in [f0, f1) (flags are
// cleared).  Pass 1 counts per 1024-row block, pass 2 scans the block counts
// in one block, pass 3 writes.
__global__ void k_flag_count(const uint8_t *__restrict__ flag, uint64_t lo,
                             uint64_t hi, uint32_t *__restrict__ bcount,
                             uint64_t f0, uint64_t f1) {
  __shared__ uint32_t s_c;
  if (threadIdx.x == 0)
    s_c = 0;
  __syncthreads();
  uint32_t c = 0;
  const uint64_t r0 = lo + (uint64_t)blockIdx.x * 1024;
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) {
    const uint64_t r = r0 + i;
    c += (r < hi && (flag[r] || (r >= f0 && r < f1))) ? 1u : 0u;
  }
  atomicAdd(&s_c, c);
  __syncthreads();
  if (threadIdx.x == 0)
    bcount[blockIdx.x] = s_c;
}

__global__ void k_scan_small(uint32_t *__restrict__ v, uint64_t n,
                             uint32_t *__restrict__ total) {
  // exclusive scan of n values with one 1024-thread block
  __shared__ uint32_t s[1024];
  uint32_t carry = 0;
  for (uint64_t b = 0; b < n; b += 1024) {
    const uint64_t i = b + threadIdx.x;
    const uint32_t x = i < n ? v[i] : 0u;
    s[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const uint32_t y = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
      __syncthreads();
      s[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < n)
      v[i] = carry + s[threadIdx.x] - x;
    const uint32_t blk = s[1023];
    __syncthreads();
    carry += blk;
  }
  if (threadIdx.x == 0)
    *total = carry;
}

__global__ void k_flag_write(uint8_t *__restrict__ flag, uint64_t lo,
                             uint64_t hi, const uint32_t *__restrict__ boff,
                             uint32_t *__restrict__ out, uint64_t f0,
                             uint64_t f1) {
  // a 1024-thread block: wave w lists rows [64 w, 64 w + 64) of the
  // block's 1024, after the block's earlier waves' rows
  __shared__ uint32_t s_cnt[16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t r = lo + (uint64_t)blockIdx.x * 1024 + threadIdx.x;
  const bool f = r < hi && (flag[r] || (r >= f0 && r < f1));
  const uint64_t m = __ballot(f);
  if (lane == 0)
    s_cnt[wv] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t at = boff[blockIdx.x];
  for (int v = 0; v < wv; ++v)
    at += s_cnt[v];
  if (f) {
    out[at + __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)r;
    flag[r] = 0;
  }
}

// --------------------------------------------------------- embedding init
// Batch rows: N(0,1) per component, replaced by the exp(-(d - d_min))
// weighted mean of the pre-batch neighbours' embeddings when there are any
// (reference initialize_embeddings, insertion_kernel.cuh:790-878).
__global__ __launch_bounds__(64) void k_init_embedding(float *__restrict__ emb, int c,
                                 const uint32_t *__restrict__ pidx,
                                 const float *__restrict__ pdist, int k,
                                 uint64_t b0, uint64_t nb) {
  const uint64_t bi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (bi >= nb)
    return;
  const uint64_t u = b0 + bi;
  Rng rng(42, (uint32_t)u, 0, TAG_EMB_INIT);
  float e[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int d = 0; d < 4; ++d)
    if (d < c)
      e[d] = rng.normal();
  // the row's k entries, read sixteen at a time as independent loads
  // (entries past k read entry 0 and are skipped; at k <= 16 the first
  // pass's loads serve the second); pre-batch neighbours only
  constexpr int KC = 16;
  uint32_t ids[KC];
  float dv[KC];
  auto load = [&](int j0) {
#pragma unroll
    for (int t = 0; t < KC; ++t) {
      const uint64_t at = u * k + (j0 + t < k ? j0 + t : 0);
      ids[t] = pidx[at];
      dv[t] = pdist[at];
    }
  };
  float dmin = 3.402823466e38f;
  for (int j0 = 0; j0 < k; j0 += KC) {
    load(j0);
#pragma unroll
    for (int t = 0; t < KC; ++t)
      if (j0 + t < k && ids[t] < b0)
        dmin = fminf(dmin, dv[t]);
  }
  if (dmin != 3.402823466e38f) {
    float wsum = 0.0f, acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int j0 = 0; j0 < k; j0 += KC) {
      if (k > KC)
        load(j0);
      float ev[KC][4];
#pragma unroll
      for (int t = 0; t < KC; ++t) {
        const uint64_t r = ids[t] < b0 ? ids[t] : 0;
#pragma unroll
        for (int d = 0; d < 4; ++d)
          ev[t][d] = d < c ? emb[r * c + d] : 0.0f;
      }
      // the weighted sums in entry order
#pragma unroll
      for (int t = 0; t < KC; ++t) {
        if (j0 + t >= k || ids[t] >= b0)
          continue;
        const float wj = expf(-(dv[t] - dmin));
        wsum += wj;
#pragma unroll
        for (int d = 0; d < 4; ++d)
          if (d < c)
            acc[d] += wj * ev[t][d];
      }
    }
    if (wsum > 0.0f) {
#pragma unroll
      for (int d = 0; d < 4; ++d)
        if (d < c)
          e[d] = acc[d] / wsum;
    }
  }
#pragma unroll
  for (int d = 0; d < 4; ++d)
    if (d < c)
      emb[u * c + d] = e[d];
}

} // namespace gm
