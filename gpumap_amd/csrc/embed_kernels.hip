// Fuzzy simplicial set and SGD layout optimiser on gfx950 (north-star
// subsystems (b) and (c)).
//
// Reference: cpp/src/kernels/embedding_kernel.cuh.  The embed-time kNN (self
// at column 0, the first k-1 stored neighbours after it, Euclidean distances
// square-rooted; reference prepend_self_knn_kernel :599-631) is not
// materialised: every kernel reads it through embed_entry() from the public
// (n x k) arrays.
#include "common.hpp"

#include <type_traits>
#include "kernels.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>

namespace gm {

namespace {
constexpr float FLT_MAX_F = 3.402823466e38f;
constexpr float FLT_MIN_F = 1.175494351e-38f;

__device__ inline bool valid_d(double d) {
  return isfinite(d) && d >= 0.0 && d < (double)FLT_MAX_F / 2.0;
}

__device__ inline void embed_entry(const uint32_t *pidx, const float *pdist,
                                   uint64_t i, int c, int k, bool euclid,
                                   uint32_t *id, double *d) {
  if (c == 0) {
    *id = (uint32_t)i;
    *d = 0.0;
    return;
  }
  *id = pidx[i * k + c - 1];
  const float v = pdist[i * k + c - 1];
  if (!valid_d((double)v))
    *d = (double)v;
  else
    *d = euclid ? (double)sqrt_rn(fmaxf(v, 0.0f)) : (double)v;
}

__device__ inline double membership(double d, double rho, double sigma) {
  if (d - rho <= 0.0 || sigma == 0.0)
    return 1.0;
  const double v = exp(-((d - rho) / sigma));
  return v < (double)FLT_MIN_F ? (double)FLT_MIN_F : v;
}

// order-preserving map of floats to unsigned ints (for atomic min/max)
__device__ inline uint32_t f2o(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float o2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
} // namespace

// ------------------------------------------------------------ mean dist
// Global mean of the valid embed-time distances (reference
// compute_mean_dist_kernel :82-128), in FP64 with a fixed reduction order.
__global__ void k_mean_partial(const uint32_t *__restrict__ pidx,
                               const float *__restrict__ pdist, NRows nr,
                               int k, int euclid, double *__restrict__ psum,
                               unsigned long long *__restrict__ pcnt) {
  __shared__ double s_s[256];
  __shared__ unsigned long long s_c[256];
  const uint64_t total = nr.get() * (uint64_t)k;
  double s = 0.0;
  unsigned long long c = 0;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t id;
    double d;
    embed_entry(pidx, pdist, t / k, (int)(t % k), k, euclid != 0, &id, &d);
    if (valid_d(d)) {
      s += d;
      ++c;
    }
  }
  s_s[threadIdx.x] = s;
  s_c[threadIdx.x] = c;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) {
      s_s[threadIdx.x] += s_s[threadIdx.x + off];
      s_c[threadIdx.x] += s_c[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    psum[blockIdx.x] = s_s[0];
    pcnt[blockIdx.x] = s_c[0];
  }
}

// Also marks the batch rows in `mark` (when given) for k_membership and
// k_reposition; k_reposition clears the marks.
__global__ __launch_bounds__(256) void k_mean_final(
    const double *__restrict__ psum, const unsigned long long *__restrict__ pcnt,
    int nparts, double *__restrict__ mean, uint32_t *__restrict__ reset3,
    const uint32_t *__restrict__ batch, uint64_t nb,
    uint8_t *__restrict__ mark) {
  // nparts <= 256 partials, summed as a fixed tree
  __shared__ double s_s[256];
  __shared__ unsigned long long s_c[256];
  const int t = threadIdx.x;
  if (mark)
    for (uint64_t r = t; r < nb; r += blockDim.x)
      mark[batch[r]] = 1;
  s_s[t] = t < nparts ? psum[t] : 0.0;
  s_c[t] = t < nparts ? pcnt[t] : 0ull;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) {
      s_s[t] += s_s[t + off];
      s_c[t] += s_c[t + off];
    }
    __syncthreads();
  }
  if (t == 0) {
    *mean = (s_c[0] > 0 && s_s[0] > 0.0) ? s_s[0] / (double)s_c[0] : 1.0;
    // the accumulators of the kernels that follow: maximum weight, then the
    // minimum and maximum of embedding column 0 (ordered-uint encodings)
    if (reset3) {
      reset3[0] = 0u;
      reset3[1] = 0xffffffffu;
      reset3[2] = 0u;
    }
  }
}

// -------------------------------------------------------- rho / sigma
// Each row's rho / sigma are computed when its public kNN row is written
// (k_sync_rows, row_sigma_rho in common.hpp); a sigma < 0 stands for the
// global-mean floor and is resolved here with sigma_of.

// ------------------------------------------------------------ membership
// Directed weights of the batch rows and the fuzzy union w = p + q - p q
// (reference compute_membership_strength_kernel :214-250 and
// symmetrize_and_expand_kernel :266-320, without the self-loop).  Also folds
// the maximum weight into *maxw_bits.  flags bit 0: live edge (i, j).  The
// union holds (j, i) too when i is absent from j's list (the reverse edge,
// same weight) or when j is a batch row with i in its list (j's own entry
// for i, whose weight p + q - p q is the same double expression with the
// operands swapped, so the same float).  The COO places (j, i) in the slot
// after (i, j): bit 1 says that slot holds (j, i); in the second case the
// entry of the larger row is not placed on its own (bit 2) since its twin
// holds it.  `inb` marks the batch rows (null: every row is one).
__global__ void k_membership(const uint32_t *__restrict__ pidx,
                             const float *__restrict__ pdist, NRows nr,
                             int k, int euclid, const double *__restrict__ sigma,
                             const double *__restrict__ rho,
                             const double *__restrict__ mean_p,
                             const uint32_t *__restrict__ batch, uint64_t nb,
                             const uint8_t *__restrict__ inb,
                             float *__restrict__ w, uint8_t *__restrict__ flags,
                             uint32_t *__restrict__ maxw_bits) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n = nr.get();
  const double mean_dist = *mean_p;
  float wf = 0.0f;
  if (t < nb * (uint64_t)k) {
    const uint64_t i = batch[t / k];
    const int c = (int)(t % k);
    uint32_t j;
    double d;
    embed_entry(pidx, pdist, i, c, k, euclid != 0, &j, &d);
    uint8_t f = 0;
    if (j < n && j != i && valid_d(d)) {
      const double p = membership(d, rho[i], sigma_of(sigma, i, mean_dist));
      double q = 0.0;
      bool found = false;
      // i's first entry in j's list (j's self entry aside), its ids read
      // four at a time as independent loads
      int hit = -1;
      for (int e0 = 0; e0 < k - 1 && hit < 0; e0 += 4) {
        uint32_t ids[4];
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4)
          ids[t4] = pidx[(uint64_t)j * k + (e0 + t4 < k - 1 ? e0 + t4 : 0)];
#pragma unroll
        for (int t4 = 3; t4 >= 0; --t4)
          if (e0 + t4 < k - 1 && ids[t4] == (uint32_t)i)
            hit = e0 + t4;
      }
      if (hit >= 0) {
        uint32_t jj;
        double dj;
        embed_entry(pidx, pdist, j, hit + 1, k, euclid != 0, &jj, &dj);
        if (valid_d(dj)) {
          q = membership(dj, rho[j], sigma_of(sigma, j, mean_dist));
          found = true;
        }
      }
      wf = (float)(p + q - p * q);
      const bool twin = found && (!inb || inb[j]);
      f = 1 | (!found || (twin && i < j) ? 2 : 0) | (twin && i > j ? 4 : 0);
    }
    w[t] = wf;
    flags[t] = f;
  }
  // wave max -> one atomic per wave
  for (int off = 32; off > 0; off >>= 1)
    wf = fmaxf(wf, __shfl_xor(wf, off));
  if ((threadIdx.x & 63) == 0 && wf > 0.0f)
    atomicMax(maxw_bits, __float_as_uint(wf));
}

// ------------------------------------------------------ COO compaction
// Keep the live edges with w >= max_w / n_epochs (reference trim_graph
// :348-400) in (batch, slot) order, each placed directed edge (i, j) as a
// twin pair of slots 2p, 2p + 1: (i, j), then (j, i) where k_membership
// found the union to hold it (the same weight, so kept with (i, j)), else an
// empty slot (head = tail = -1, w = 0).  The SGD kernels skip empty slots
// and use the pairing to add both edges' contributions to a row with one
// atomic.  A deterministic COO with the reference's edge set.
__device__ inline bool coo_keep(const float *w, const uint8_t *flags,
                                uint64_t s, float thr) {
  return (flags[s] & 5) == 1 && w[s] >= thr;
}

__device__ inline float coo_thr(const uint32_t *maxw_bits, int n_epochs) {
  const float mw = __uint_as_float(*maxw_bits);
  return mw > 0.0f ? div_rn(mw, (float)n_epochs) : 3.402823466e38f;
}

__global__ void k_coo_count(const float *__restrict__ w,
                            const uint8_t *__restrict__ flags, uint64_t nd,
                            const uint32_t *__restrict__ maxw_bits,
                            int n_epochs, uint32_t *__restrict__ bcount) {
  __shared__ uint32_t s_c;
  if (threadIdx.x == 0)
    s_c = 0;
  __syncthreads();
  const float thr = coo_thr(maxw_bits, n_epochs);
  const uint64_t t0 = (uint64_t)blockIdx.x * 1024;
  uint32_t c = 0;
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) {
    const uint64_t t = t0 + i;
    c += (t < nd && coo_keep(w, flags, t, thr)) ? 2u : 0u;
  }
  atomicAdd(&s_c, c);
  __syncthreads();
  if (threadIdx.x == 0)
    bcount[blockIdx.x] = s_c;
}

__global__ void k_coo_write(const uint32_t *__restrict__ pidx, int k,
                            const uint32_t *__restrict__ batch,
                            const float *__restrict__ w,
                            const uint8_t *__restrict__ flags, uint64_t nd,
                            const uint32_t *__restrict__ maxw_bits,
                            int n_epochs, const uint32_t *__restrict__ boff,
                            int32_t *__restrict__ head,
                            int32_t *__restrict__ tail,
                            float *__restrict__ wout) {
  // a 1024-thread block: wave w places entries [64 w, 64 w + 64) of the
  // block's 1024, after the block's earlier waves' entries
  __shared__ uint32_t s_cnt[16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const float thr = coo_thr(maxw_bits, n_epochs);
  const uint64_t s = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
  const bool f = s < nd && coo_keep(w, flags, s, thr);
  const uint64_t m = __ballot(f);
  if (lane == 0)
    s_cnt[wv] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t at = boff[blockIdx.x];
  for (int v = 0; v < wv; ++v)
    at += 2u * s_cnt[v];
  if (f) {
    const int32_t hi = (int32_t)batch[s / k];
    const int c = (int)(s % k);
    const int32_t hj = c == 0 ? hi : (int32_t)pidx[(uint64_t)hi * k + c - 1];
    const bool rev = flags[s] & 2;
    const float ws = w[s];
    const uint32_t pos = at + 2u * __popcll(m & ((1ull << lane) - 1ull));
    head[pos] = hi;
    tail[pos] = hj;
    wout[pos] = ws;
    head[pos + 1] = rev ? hj : -1;
    tail[pos + 1] = rev ? hi : -1;
    wout[pos + 1] = rev ? ws : 0.0f;
  }
}

// ------------------------------------------------------------ SGD setup
// epochs_per_sample and epoch state (reference make_epochs_per_sample_kernel
// :469-477, init_epoch_state_kernel :485-495), and head degrees for the
// gradient rounding factor (:405-460).
__global__ void k_sgd_init(const float *__restrict__ w,
                           const int32_t *__restrict__ head,
                           const uint32_t *__restrict__ nnz_p,
                           const uint32_t *__restrict__ maxw_bits,
                           int n_epochs, float nsr_inv,
                           float *__restrict__ eps, float *__restrict__ nxt,
                           float *__restrict__ nneg, int *__restrict__ deg,
                           int *__restrict__ maxdeg, unsigned *__restrict__ bar,
                           unsigned nbar) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // block 0 clears what the next kernels accumulate into: the maximum
  // degree (k_max_int) and the cooperative SGD's barrier words
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0)
      *maxdeg = 0;
    for (unsigned i = threadIdx.x; i < nbar; i += blockDim.x)
      bar[i] = 0u;
  }
  if (e >= *nnz_p)
    return;
  const float mw = __uint_as_float(*maxw_bits);
  const float ns = (float)n_epochs * div_rn(w[e], mw);
  const float ep = ns > 0.0f ? div_rn((float)n_epochs, ns) : -1.0f;
  eps[e] = ep;
  nxt[e] = ep;
  nneg[e] = ep * nsr_inv;
  const int32_t h = head[e];
  if (h >= 0) // not an empty slot
    atomicAdd(deg + h, 1);
}

__device__ inline float rounding_of(int md, float alpha) {
  if (md == 0)
    return 1.0f;
  const float max_abs = (float)md * 4.0f * fabsf(alpha);
  const float delta =
      div_rn(max_abs, 1.0f - 2.0f * (float)md * 1.1920928955078125e-7f);
  int e;
  frexpf(delta, &e);
  return ldexpf(1.0f, e);
}

// Maximum head degree, and from it the gradient rounding grid (reference
// :405-460): the last block to finish reads the maximum and writes the grid.
// `done` counts finished blocks and is left at 0 for the next launch.
__global__ __launch_bounds__(256) void k_max_int(const int *__restrict__ v,
                                                 NRows nr,
                                                 int *__restrict__ out,
                                                 float alpha,
                                                 float *__restrict__ rounding,
                                                 unsigned *__restrict__ done) {
  __shared__ int s_m[4];
  __shared__ bool s_last;
  const uint64_t n = nr.get();
  int m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    m = max(m, v[i]);
  for (int off = 32; off > 0; off >>= 1)
    m = max(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0)
    s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) { // one atomic per block
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      m = max(m, s_m[w]);
    if (m > 0)
      atomicMax(out, m);
    __threadfence();
    s_last = atomicAdd(done, 1u) == gridDim.x - 1;
    if (s_last) {
      __threadfence();
      const int md = __hip_atomic_load(out, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
      *rounding = rounding_of(md, alpha);
      *done = 0u;
    }
  }
}

// -------------------------------------------------------------- noise
__global__ __launch_bounds__(256) void k_minmax_col0(
    const float *__restrict__ emb, NRows nr, int c, uint32_t *__restrict__ mn,
    uint32_t *__restrict__ mx, int *__restrict__ deg_zero) {
  __shared__ uint32_t s_lo[4], s_hi[4];
  const uint64_t n = nr.get();
  uint32_t lo = 0xffffffffu, hi = 0u;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    // the per-row head degrees of the SGD set-up start from zero
    if (deg_zero)
      deg_zero[i] = 0;
    const uint32_t o = f2o(emb[i * c]);
    lo = min(lo, o);
    hi = max(hi, o);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off));
    hi = max(hi, __shfl_xor(hi, off));
  }
  if ((threadIdx.x & 63) == 0) {
    s_lo[threadIdx.x >> 6] = lo;
    s_hi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) { // one atomic pair per block
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      lo = min(lo, s_lo[w]);
      hi = max(hi, s_hi[w]);
    }
    atomicMin(mn, lo);
    atomicMax(mx, hi);
  }
}

// N(0, 0.001 * x_range) on the selected rows (reference add_embedding_noise
// :953-1007).
__global__ __launch_bounds__(256) void k_noise(float *__restrict__ emb, int c,
                        const uint32_t *__restrict__ batch, uint64_t nb,
                        const uint32_t *__restrict__ mn,
                        const uint32_t *__restrict__ mx, float noise_scale,
                        uint64_t seed) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * (uint64_t)c)
    return;
  const float scale = noise_scale * (o2f(*mx) - o2f(*mn));
  Rng rng(seed, (uint32_t)t, (uint32_t)(t >> 32), TAG_NOISE);
  const uint32_t row = batch[t / c];
  emb[(uint64_t)row * c + t % c] += rng.normal() * scale;
}

// --------------------------------------------------------------- SGD
// x^y for x > 0 through the hardware base-2 exponential and logarithm
// (about 2 ulp; the SGD is compared statistically)
__device__ inline float pow_pos(float x, float y) {
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

__device__ inline float fclip(float v) {
  if (!isfinite(v))
    return 0.0f;
  return fmaxf(fminf(v, 4.0f), -4.0f);
}

// The first 8 raw draws of edge e's negative-sample stream for one epoch
// (two Philox4x32-10 blocks; the stream continues at block 2 when an edge
// needs more than 8 samples).  A sample's index is the high word of draw * n.
__device__ inline void neg_draws(uint64_t seed, uint64_t e, uint32_t out[8]) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t ctr[4] = {(uint32_t)e, (uint32_t)(e >> 32), 0u, TAG_NEG};
  philox(ctr, key, out);
  ctr[2] = 1u;
  philox(ctr, key, out + 4);
}

// One positive sample of edge e and its negative samples (reference
// optimize_batch_kernel :509-589) for any n_components <= 4, in two halves so
// that a caller can overlap independent work with the memory round trip:
// sgd_gather issues the loads of the edge's two endpoints and of its first 8
// negative samples (positions taken before the sequential updates of the
// head: the values a sequential read would see in a Hogwild run);
// sgd_update does the arithmetic and returns the edge's contributions to its
// head (gh) and tail (gt), each truncated to the rounding grid as the
// reference does before its atomic adds (the caller adds them, seg_add).
// The gradient coefficients use the hardware reciprocal (1 ulp); the epoch
// schedule keeps IEEE division.
template <int NC> struct EdgeLoads {
  float cur[NC], oth[NC], neg[8][NC];
  uint32_t r[8];
  int n_neg;
};

template <int NC>
__device__ inline void sgd_gather(const float *__restrict__ emb, int32_t j,
                                  int32_t kk, uint32_t n, float ep_neg,
                                  float nn, int epoch, const uint32_t draw[8],
                                  EdgeLoads<NC> &L) {
  L.n_neg = (int)(((float)epoch - nn) / ep_neg);
  const int cnt = min(8, L.n_neg);
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (q < cnt) {
      L.r[q] = __umulhi(draw[q], n);
      const float *np_ = emb + (uint64_t)L.r[q] * NC;
#pragma unroll
      for (int d = 0; d < NC; ++d)
        L.neg[q][d] = np_[d];
    }
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    L.cur[d] = emb[(uint64_t)j * NC + d];
    L.oth[d] = emb[(uint64_t)kk * NC + d];
  }
}

template <int NC>
__device__ inline void neg_step(float cur[NC], float g[NC], const float neg[NC],
                                float alpha, float gamma, float a, float b) {
  float d2n = 0.0f;
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    const float t = cur[d] - neg[d];
    d2n += t * t;
  }
  float rep = 0.0f;
  if (d2n > 0.0f)
    rep = (2.0f * gamma * b) *
          __builtin_amdgcn_rcpf((0.001f + d2n) * (a * pow_pos(d2n, b) + 1.0f));
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    const float gg =
        rep > 0.0f ? fclip(rep * (cur[d] - neg[d])) * alpha : 4.0f * alpha;
    cur[d] += gg;
    g[d] += gg;
  }
}

// Where the SGD reads rows: the float embedding, or its fixed-point copy
// (FixRows, below); load() gives a row's NC coordinates as floats.
// load_raw / decode split a load from its use: the raw words stay in
// registers (the load in flight) until decode, so that a caller can issue
// rows long before it needs them.
template <int NC> struct EmbRows {
  const float *__restrict__ emb;
  static constexpr int RAW = NC;
  __device__ void load(uint64_t row, float out[NC]) const {
#pragma unroll
    for (int d = 0; d < NC; ++d)
      out[d] = emb[row * NC + d];
  }
  __device__ void load_raw(uint64_t row, int32_t out[RAW]) const {
#pragma unroll
    for (int d = 0; d < NC; ++d)
      out[d] = __float_as_int(emb[row * NC + d]);
  }
  __device__ void decode(const int32_t in[RAW], float out[NC]) const {
#pragma unroll
    for (int d = 0; d < NC; ++d)
      out[d] = __int_as_float(in[d]);
  }
};

// the arithmetic of sgd_update; the caller advances the edge's schedule
template <int NC, typename Rows>
__device__ inline void sgd_apply_rows(const Rows &rw, int32_t j, uint64_t e,
                                      uint32_t n, float ep_neg, float alpha,
                                      float gamma, float a, float b,
                                      uint64_t seed, float rounding,
                                      EdgeLoads<NC> &L, float gh[NC],
                                      float gt[NC],
                                      unsigned long long &c_pos,
                                      unsigned long long &c_neg) {
  ++c_pos;
  float g[NC];
  float d2 = 0.0f;
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    const float t = L.cur[d] - L.oth[d];
    d2 += t * t;
  }
  float attr = 0.0f;
  if (d2 > 0.0f) {
    const float pb = pow_pos(d2, b);
    attr = (-2.0f * a * b * pb) *
           __builtin_amdgcn_rcpf(d2 * (a * pb + 1.0f));
  }
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    g[d] = fclip(attr * (L.cur[d] - L.oth[d])) * alpha;
    L.cur[d] += g[d];
    gt[d] = (rounding + -g[d]) - rounding;
  }
  const int n_neg = L.n_neg;
  if (n_neg > 0) {
    const int cnt = min(8, n_neg);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < cnt && (int32_t)L.r[q] != j) {
        ++c_neg;
        neg_step<NC>(L.cur, g, L.neg[q], alpha, gamma, a, b);
      }
    if (n_neg > 8) {
      // more than 8 negative samples (an edge skipped many epochs): the
      // stream continues at Philox block 2
      Rng rng(seed, (uint32_t)e, (uint32_t)(e >> 32), TAG_NEG);
      rng.ctr[2] = 2u;
      for (int p = 8; p < n_neg; ++p) {
        const uint32_t r = __umulhi(rng.next(), n);
        if ((int32_t)r == j)
          continue;
        ++c_neg;
        float neg[NC];
        rw.load(r, neg);
        neg_step<NC>(L.cur, g, neg, alpha, gamma, a, b);
      }
    }
  }
#pragma unroll
  for (int d = 0; d < NC; ++d)
    gh[d] = (rounding + g[d]) - rounding;
}

template <int NC>
__device__ inline void sgd_apply(const float *__restrict__ emb, int32_t j,
                                 uint64_t e, uint32_t n, float ep_neg,
                                 float alpha, float gamma, float a, float b,
                                 uint64_t seed, float rounding,
                                 EdgeLoads<NC> &L, float gh[NC], float gt[NC],
                                 unsigned long long &c_pos,
                                 unsigned long long &c_neg) {
  sgd_apply_rows<NC>(EmbRows<NC>{emb}, j, e, n, ep_neg, alpha, gamma, a, b,
                     seed, rounding, L, gh, gt, c_pos, c_neg);
}

// the edge's schedule after a sample with L.n_neg negative samples
__device__ inline void advance_schedule(float ep, float ep_neg, int n_neg,
                                        float &nx, float &nn) {
  nx += ep;
  if (n_neg > 0)
    nn += (float)n_neg * ep_neg;
}

template <int NC>
__device__ inline void sgd_update(const float *__restrict__ emb, int32_t j,
                                  uint64_t e, uint32_t n, float ep, float &nx,
                                  float &nn, float ep_neg, float alpha,
                                  float gamma, float a, float b, uint64_t seed,
                                  float rounding, EdgeLoads<NC> &L,
                                  float gh[NC], float gt[NC],
                                  unsigned long long &c_pos,
                                  unsigned long long &c_neg) {
  sgd_apply<NC>(emb, j, e, n, ep_neg, alpha, gamma, a, b, seed, rounding, L,
                gh, gt, c_pos, c_neg);
  advance_schedule(ep, ep_neg, L.n_neg, nx, nn);
}

// sgd_gather without branches: every lane issues all ten loads (endpoints
// and 8 negative samples); a lane without a sample this epoch, and samples
// past its count, read row 0 (one cache line per wave).  With a fixed number
// of loads the compiler's waits count exactly the operations issued after
// them.
template <int NC, typename Rows>
__device__ inline void sgd_gather_rows(const Rows &rw, bool act, int32_t j,
                                       int32_t kk, uint32_t n, float ep_neg,
                                       float nn, int epoch,
                                       const uint32_t draw[8],
                                       EdgeLoads<NC> &L) {
  L.n_neg = act ? (int)(((float)epoch - nn) / ep_neg) : 0;
  const int cnt = min(8, L.n_neg);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    L.r[q] = __umulhi(draw[q], n);
    const uint32_t rq = q < cnt ? L.r[q] : 0u;
    rw.load(rq, L.neg[q]);
  }
  const uint64_t jj = act ? (uint64_t)j : 0, kq = act ? (uint64_t)kk : 0;
  rw.load(jj, L.cur);
  rw.load(kq, L.oth);
}

template <int NC>
__device__ inline void sgd_gather_all(const float *__restrict__ emb, bool act,
                                      int32_t j, int32_t kk, uint32_t n,
                                      float ep_neg, float nn, int epoch,
                                      const uint32_t draw[8],
                                      EdgeLoads<NC> &L) {
  sgd_gather_rows<NC>(EmbRows<NC>{emb}, act, j, kk, n, ep_neg, nn, epoch,
                      draw, L);
}

// Runs of equal row ids over the lanes of a wave (edges in COO order, so a
// row's edges as head, or as tail, are mostly adjacent), cut at the 16-lane
// DPP rows.  m bit s (s < 4): lane + 2^s is in the same run and row; start:
// first lane of its run within its row.  Lanes without an edge pass a key of
// their own (< 0).
struct Seg {
  uint32_t m;
  bool start;
};

__device__ inline Seg seg_of(int32_t key) {
  const int lane = threadIdx.x & 63;
  const int32_t prev = __shfl_up(key, 1, 64);
  const bool start = (lane & 15) == 0 || prev != key;
  const uint64_t sm = __ballot(start);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const int run = __popcll(sm & upto);
  uint32_t m = 0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int ro = __shfl_down(run, 1 << s, 64);
    if ((lane & 15) + (1 << s) < 16 && ro == run)
      m |= 1u << s;
  }
  return Seg{m, start};
}

// v of lane ^ 1 (a DPP quad permutation, with no LDS round trip)
__device__ inline float lane_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xf,
                                                 0xf, false));
}

// v from lane + 2^S of the same 16-lane row (0 past the row's end)
template <int S> __device__ inline float dpp_from_above(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(
      0, __float_as_int(v), 0x100 | (1 << S), 0xf, 0xf, true));
}

// Sum each run's contributions (within a 16-lane row) into its first lane
// with four DPP row-shift steps and add them to the row there: one atomic
// per run and component instead of one per edge.  The contributions are
// multiples of the rounding grid and a run holds at most max-degree of them,
// so their sum is exact in any order.  All lanes of the wave must call this
// together.
template <int NC>
__device__ inline void seg_add(float *__restrict__ emb, int32_t row, Seg sg,
                               float v[NC]) {
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    float o = dpp_from_above<0>(v[d]);
    if (sg.m & 1u)
      v[d] += o;
    o = dpp_from_above<1>(v[d]);
    if (sg.m & 2u)
      v[d] += o;
    o = dpp_from_above<2>(v[d]);
    if (sg.m & 4u)
      v[d] += o;
    o = dpp_from_above<3>(v[d]);
    if (sg.m & 8u)
      v[d] += o;
  }
  bool any = false;
#pragma unroll
  for (int d = 0; d < NC; ++d)
    any |= v[d] != 0.0f;
  if (sg.start && any) {
#pragma unroll
    for (int d = 0; d < NC; ++d)
      unsafeAtomicAdd(emb + (uint64_t)row * NC + d, v[d]);
  }
}

// Runs of equal keys over the even lanes of a wave (stride 2, cut at the
// 16-lane DPP rows); odd lanes are runs of their own.  m bit s (s = 1..3):
// lane + 2^s is in the same run.
__device__ inline Seg seg2_of(int32_t key) {
  const int lane = threadIdx.x & 63;
  const bool even = !(lane & 1);
  const int32_t prev = __shfl_up(key, 2, 64);
  const bool start = !even || (lane & 15) == 0 || prev != key;
  const uint64_t sm = __ballot(even && start);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const int run = __popcll(sm & upto);
  uint32_t m = 0;
#pragma unroll
  for (int s = 1; s < 4; ++s) {
    const int ro = __shfl_down(run, 1 << s, 64);
    if (even && (lane & 15) + (1 << s) < 16 && ro == run)
      m |= 1u << s;
  }
  return Seg{m, start};
}

// seg_add over seg2_of's runs: DPP row shifts of 2, 4 and 8 lanes.
template <int NC>
__device__ inline void seg2_add(float *__restrict__ emb, int32_t row, Seg sg,
                                float v[NC]) {
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    float o = dpp_from_above<1>(v[d]);
    if (sg.m & 2u)
      v[d] += o;
    o = dpp_from_above<2>(v[d]);
    if (sg.m & 4u)
      v[d] += o;
    o = dpp_from_above<3>(v[d]);
    if (sg.m & 8u)
      v[d] += o;
  }
  bool any = false;
#pragma unroll
  for (int d = 0; d < NC; ++d)
    any |= v[d] != 0.0f;
  if (sg.start && any) {
#pragma unroll
    for (int d = 0; d < NC; ++d)
      unsafeAtomicAdd(emb + (uint64_t)row * NC + d, v[d]);
  }
}

// How a wave adds its edges' contributions.  The embedding's COO places
// each edge (i, j) and its reverse (j, i) (or an empty slot) in lanes 2p,
// 2p + 1: then each lane adds its own head's contribution and its twin's
// tail's (the same row), lane 2p for i and lane 2p + 1 for j, and runs of
// equal i over the even lanes share an atomic: about half the atomics of
// adding each edge's head and tail on their own.  Any other wave (a caller's
// COO) adds heads and tails by their own runs.  Either way every
// contribution is added once, exactly (all are multiples of the rounding
// grid).  Lanes without an edge carry negative row ids and zero
// contributions.
struct Agg {
  bool paired;  // wave-uniform
  Seg sh, st;   // unpaired: runs of heads and of tails
  Seg sp;       // paired: runs of kp
  int32_t kp;   // paired: the lane's row (its head, else its twin's tail)
};

__device__ inline Agg agg_of(int32_t hh, int32_t tt) {
  const int32_t ph = __shfl_xor(hh, 1, 64), pt = __shfl_xor(tt, 1, 64);
  const bool twins = hh == pt && tt == ph;
  const bool ok = twins || hh < 0 || ph < 0;
  Agg a;
  // uniform in a scalar register, so that its branches are scalar branches
  a.paired = __builtin_amdgcn_readfirstlane(__ballot(!ok) == 0 ? 1 : 0) != 0;
  a.kp = hh >= 0 ? hh : pt;
  if (a.paired) {
    a.sp = seg2_of(a.kp);
    a.sh = a.st = Seg{0u, true};
  } else {
    a.sh = seg_of(hh);
    a.st = seg_of(tt);
    a.sp = Seg{0u, true};
  }
  return a;
}

template <int NC>
__device__ inline void agg_add(float *__restrict__ emb, const Agg &a,
                               int32_t hh, int32_t tt, float gh[NC],
                               float gt[NC]) {
  if (a.paired) {
    float tot[NC];
#pragma unroll
    for (int d = 0; d < NC; ++d)
      tot[d] = gh[d] + __shfl_xor(gt[d], 1, 64);
    seg2_add<NC>(emb, a.kp, a.sp, tot);
  } else {
    seg_add<NC>(emb, tt, a.st, gt);
    seg_add<NC>(emb, hh, a.sh, gh);
  }
}

// ----------------------------------------------- branch-free row updates
// The per-cycle SGD kernel issues the next epoch's loads before this epoch's
// atomics, so that waiting for the loads does not wait for the atomics (the
// vector memory counter completes in issue order).  That works only if the
// operations issued after the loads do not depend on a branch (the compiler
// then has to assume the fewest of them and waits for the rest): the row
// updates go through a raw buffer over the embedding, and a lane that adds
// nothing passes an offset past the buffer's end, which the hardware drops.
constexpr uint32_t BUF_SKIP = 0x80000000u;

template <int NC>
__device__ inline __amdgpu_buffer_rsrc_t emb_rsrc(float *emb, uint32_t n) {
  // num_records in bytes; rows stay below 2^31 / 16
  return __builtin_amdgcn_make_buffer_rsrc(emb, 0, (int)(n * (NC * 4u)),
                                           0x00020000);
}

// run sum into each run's first lane (seg_add's DPP steps from shift 2^S0)
template <int NC, int S0>
__device__ inline bool seg_sum(Seg sg, float v[NC]) {
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    if constexpr (S0 == 0) {
      const float o = dpp_from_above<0>(v[d]);
      if (sg.m & 1u)
        v[d] += o;
    }
    float o = dpp_from_above<1>(v[d]);
    if (sg.m & 2u)
      v[d] += o;
    o = dpp_from_above<2>(v[d]);
    if (sg.m & 4u)
      v[d] += o;
    o = dpp_from_above<3>(v[d]);
    if (sg.m & 8u)
      v[d] += o;
  }
  bool any = false;
#pragma unroll
  for (int d = 0; d < NC; ++d)
    any |= v[d] != 0.0f;
  return sg.start && any;
}

template <int NC>
__device__ inline void buf_row_add(__amdgpu_buffer_rsrc_t r, int32_t row,
                                   bool wr, const float v[NC]) {
  const uint32_t off =
      (wr && row >= 0) ? (uint32_t)row * (NC * 4u) : BUF_SKIP;
#pragma unroll
  for (int d = 0; d < NC; ++d)
    __builtin_amdgcn_raw_ptr_buffer_atomic_fadd_f32(v[d], r,
                                                    (int)(off + 4u * d), 0, 0);
}

// agg_add without branches: always two sets of NC row updates (a paired
// wave's second set is dropped), selected by the uniform `paired`.  seg2_of's
// runs have no bit 0, so seg_sum's first step adds nothing for them.
template <int NC>
__device__ inline void agg_add_buf(__amdgpu_buffer_rsrc_t r, const Agg &a,
                                   int32_t hh, int32_t tt, float gh[NC],
                                   float gt[NC]) {
  float v1[NC], v2[NC];
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    const float tot = gh[d] + lane_xor1(gt[d]);
    v1[d] = a.paired ? tot : gt[d];
    v2[d] = a.paired ? 0.0f : gh[d];
  }
  // field by field: a select of whole structs goes through scratch memory
  const Seg s1{a.paired ? a.sp.m : a.st.m, a.paired ? a.sp.start : a.st.start};
  const int32_t k1 = a.paired ? a.kp : tt;
  const bool w1 = seg_sum<NC, 0>(s1, v1);
  const bool w2 = seg_sum<NC, 0>(a.sh, v2);
  buf_row_add<NC>(r, k1, w1, v1);
  buf_row_add<NC>(r, hh, w2 && !a.paired, v2);
}

// ------------------------------------------------ fixed-point row updates
// The cooperative SGD kernels keep each row as a record (FixFrame): its
// coordinates at the launch's start p0 (float) and the sum D of the
// launch's updates as integers in units of q, the rounding grid's finest
// spacing (rounding factor x 2^-24: every contribution is an exact integer
// multiple of it).  A row reads as p0 + D q, the same value the float
// embedding holds under float atomics whenever those adds are exact (the
// grid's purpose); coordinates 2i and 2i + 1 share one 64-bit delta word
// M = D_2i + 2^32 D_2i+1 (as integers, modulo 2^64), so one 64-bit integer
// atomic adds to both.  A device-scope atomic costs per active lane: a 2-D
// or 4-D row update takes half the lanes of one float atomic per
// coordinate, and integer sums are exact in any order.  k_fix_out adds each
// row's D q back to the float embedding (rows with D = 0 keep their bits).
// Record layout: fix_stride<NC> float words, then as many delta words; the
// low word of a pair's M is D_2i, its high word D_2i+1 - 1 when D_2i < 0
// (the borrow), which fix_decode adds back.
template <int NC> constexpr int fix_stride() { return NC == 3 ? 4 : NC; }

template <int NC>
__device__ inline void fix_decode(const int32_t *__restrict__ dw,
                                  int32_t D[NC]) {
  int32_t x[fix_stride<NC>()];
#pragma unroll
  for (int d = 0; d < fix_stride<NC>(); ++d)
    x[d] = dw[d];
#pragma unroll
  for (int d = 0; d < NC; ++d)
    D[d] = (d & 1) ? x[d] + (int32_t)((uint32_t)x[d - 1] >> 31) : x[d];
}

// a 64-bit word holding lo + 2^32 hi (modulo 2^64)
__device__ inline unsigned long long fix_pack(int32_t lo, int32_t hi) {
  return (unsigned long long)(int64_t)lo +
         ((unsigned long long)(uint32_t)hi << 32);
}

template <int NC> struct FixRows {
  const int32_t *__restrict__ rec;
  float q;
  static constexpr int RAW = 2 * fix_stride<NC>();
  __device__ void load(uint64_t row, float out[NC]) const {
    constexpr int F = fix_stride<NC>();
    const int32_t *r = rec + row * (2 * F);
    int32_t D[NC];
    fix_decode<NC>(r + F, D);
#pragma unroll
    for (int d = 0; d < NC; ++d)
      out[d] = __int_as_float(r[d]) + (float)D[d] * q;
  }
  __device__ void load_raw(uint64_t row, int32_t out[RAW]) const {
    const int32_t *r = rec + row * RAW;
#pragma unroll
    for (int d = 0; d < RAW; ++d)
      out[d] = r[d];
  }
  __device__ void decode(const int32_t in[RAW], float out[NC]) const {
    constexpr int F = fix_stride<NC>();
    int32_t D[NC];
    fix_decode<NC>(in + F, D);
#pragma unroll
    for (int d = 0; d < NC; ++d)
      out[d] = __int_as_float(in[d]) + (float)D[d] * q;
  }
};

// One row's contributions v (multiples of q) into its record: one 64-bit
// atomic per coordinate pair, a 32-bit one for a lone third coordinate (or
// the only one).
template <int NC>
__device__ inline void fix_row_add(int32_t *__restrict__ rec, float qinv,
                                   int32_t row, bool wr, const float v[NC]) {
  if (!(wr && row >= 0))
    return;
  constexpr int F = fix_stride<NC>();
  int32_t *p = rec + (uint64_t)row * (2 * F) + F;
  int32_t dx[NC];
#pragma unroll
  for (int d = 0; d < NC; ++d)
    dx[d] = (int32_t)rintf(v[d] * qinv);
#pragma unroll
  for (int d = 0; d + 1 < NC; d += 2) {
    const unsigned long long pk = fix_pack(dx[d], dx[d + 1]);
    if (pk != 0ull)
      __hip_atomic_fetch_add(reinterpret_cast<unsigned long long *>(p + d),
                             pk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if constexpr (NC & 1) {
    if (dx[NC - 1] != 0)
      __hip_atomic_fetch_add(reinterpret_cast<unsigned *>(p + NC - 1),
                             (unsigned)dx[NC - 1], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

// agg_add_buf's row updates on the records
template <int NC>
__device__ inline void agg_add_fix(int32_t *__restrict__ rec, float qinv,
                                   const Agg &a, int32_t hh, int32_t tt,
                                   float gh[NC], float gt[NC]) {
  float v1[NC], v2[NC];
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    const float tot = gh[d] + lane_xor1(gt[d]);
    v1[d] = a.paired ? tot : gt[d];
    v2[d] = a.paired ? 0.0f : gh[d];
  }
  const Seg s1{a.paired ? a.sp.m : a.st.m, a.paired ? a.sp.start : a.st.start};
  const int32_t k1 = a.paired ? a.kp : tt;
  const bool w1 = seg_sum<NC, 0>(s1, v1);
  const bool w2 = seg_sum<NC, 0>(a.sh, v2);
  fix_row_add<NC>(rec, qinv, k1, w1, v1);
  fix_row_add<NC>(rec, qinv, hh, w2 && !a.paired, v2);
}

// fix_row_add without branches, for kernels that keep a fixed set of memory
// operations in flight: buffer atomics on the records, a lane that adds
// nothing passing an offset past the buffer's end (dropped by the
// hardware).  Integer buffer atomics have no clang builtin; these are the
// LLVM intrinsics (no-return forms when the result is unused).
__device__ long long gm_buf_atomic_add_i64(long long v,
                                           __amdgpu_buffer_rsrc_t r, int off,
                                           int soff, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.atomic.add.i64");
__device__ int gm_buf_atomic_add_i32(int v, __amdgpu_buffer_rsrc_t r, int off,
                                     int soff, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.atomic.add.i32");

template <int NC>
__device__ inline __amdgpu_buffer_rsrc_t fix_rsrc(int32_t *rec, uint32_t n) {
  // num_records in bytes; records stay below 2^31 bytes
  return __builtin_amdgcn_make_buffer_rsrc(
      rec, 0, (int)(n * (2u * fix_stride<NC>() * 4u)), 0x00020000);
}

template <int NC>
__device__ inline void fix_row_add_buf(__amdgpu_buffer_rsrc_t r, float qinv,
                                       int32_t row, bool wr,
                                       const float v[NC]) {
  constexpr uint32_t F = fix_stride<NC>();
  const uint32_t off =
      (wr && row >= 0) ? ((uint32_t)row * (2u * F) + F) * 4u : BUF_SKIP;
  int32_t dx[NC];
#pragma unroll
  for (int d = 0; d < NC; ++d)
    dx[d] = (int32_t)rintf(v[d] * qinv);
#pragma unroll
  for (int d = 0; d + 1 < NC; d += 2)
    gm_buf_atomic_add_i64((long long)fix_pack(dx[d], dx[d + 1]), r,
                          (int)(off + 4u * d), 0, 0);
  if constexpr (NC & 1)
    gm_buf_atomic_add_i32(dx[NC - 1], r, (int)(off + 4u * (NC - 1)), 0, 0);
}

// agg_add_fix with fix_row_add_buf: always the same atomics per call
template <int NC>
__device__ inline void agg_add_fix_buf(__amdgpu_buffer_rsrc_t r, float qinv,
                                       const Agg &a, int32_t hh, int32_t tt,
                                       float gh[NC], float gt[NC]) {
  float v1[NC], v2[NC];
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    const float tot = gh[d] + lane_xor1(gt[d]);
    v1[d] = a.paired ? tot : gt[d];
    v2[d] = a.paired ? 0.0f : gh[d];
  }
  const Seg s1{a.paired ? a.sp.m : a.st.m, a.paired ? a.sp.start : a.st.start};
  const int32_t k1 = a.paired ? a.kp : tt;
  const bool w1 = seg_sum<NC, 0>(s1, v1);
  const bool w2 = seg_sum<NC, 0>(a.sh, v2);
  fix_row_add_buf<NC>(r, qinv, k1, w1, v1);
  fix_row_add_buf<NC>(r, qinv, hh, w2 && !a.paired, v2);
}

// The rows and row updates of one SGD kernel: the float embedding (buffer
// float atomics) or the records.
template <int NC> struct FloatStore {
  EmbRows<NC> rows;
  __amdgpu_buffer_rsrc_t er;
  __device__ void add(const Agg &a, int32_t hh, int32_t tt, float gh[NC],
                      float gt[NC]) const {
    agg_add_buf<NC>(er, a, hh, tt, gh, gt);
  }
};
template <int NC> struct FixStore {
  FixRows<NC> rows;
  int32_t *rec;
  float qinv;
  __device__ void add(const Agg &a, int32_t hh, int32_t tt, float gh[NC],
                      float gt[NC]) const {
    agg_add_fix<NC>(rec, qinv, a, hh, tt, gh, gt);
  }
};
template <int NC>
__device__ inline FixStore<NC> fix_store(int32_t *rec, float rounding) {
  const float q = ldexpf(rounding, -24); // a power of two, as is 1 / q
  return FixStore<NC>{FixRows<NC>{rec, q}, rec, 1.0f / q};
}
// FixStore with branch-free row updates (agg_add_fix_buf)
template <int NC> struct FixStoreBuf {
  FixRows<NC> rows;
  __amdgpu_buffer_rsrc_t rr;
  float qinv;
  __device__ void add(const Agg &a, int32_t hh, int32_t tt, float gh[NC],
                      float gt[NC]) const {
    agg_add_fix_buf<NC>(rr, qinv, a, hh, tt, gh, gt);
  }
};
// whether n rows' records fit one buffer descriptor (below 2^31 bytes)
template <int NC> __host__ __device__ inline bool fix_buf_fits(uint64_t n) {
  return n * (2u * fix_stride<NC>() * 4u) < 0x80000000ull;
}
template <int NC>
__device__ inline FixStoreBuf<NC> fix_store_buf(int32_t *rec, float rounding,
                                                uint32_t n) {
  const float q = ldexpf(rounding, -24);
  return FixStoreBuf<NC>{FixRows<NC>{rec, q}, fix_rsrc<NC>(rec, n), 1.0f / q};
}

// The batched kernel's store for the float embedding: row updates by the
// branching seg_add / seg2_add (agg_add), which issue only the atomics a
// wave needs.
template <int NC> struct FloatStoreBranchy {
  EmbRows<NC> rows;
  float *emb;
  __device__ void add(const Agg &a, int32_t hh, int32_t tt, float gh[NC],
                      float gt[NC]) const {
    agg_add<NC>(emb, a, hh, tt, gh, gt);
  }
};

// sgd_gather on any row store: only the samples' loads, in a branch
template <int NC, typename Rows>
__device__ inline void sgd_gather_r(const Rows &rw, int32_t j, int32_t kk,
                                    uint32_t n, float ep_neg, float nn,
                                    int epoch, const uint32_t draw[8],
                                    EdgeLoads<NC> &L) {
  L.n_neg = (int)(((float)epoch - nn) / ep_neg);
  const int cnt = min(8, L.n_neg);
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (q < cnt) {
      L.r[q] = __umulhi(draw[q], n);
      rw.load(L.r[q], L.neg[q]);
    }
  rw.load((uint64_t)j, L.cur);
  rw.load((uint64_t)kk, L.oth);
}

// every row's record: its coordinates, no updates yet
template <int NC>
__global__ __launch_bounds__(256) void k_fix_in(const float *__restrict__ emb,
                                                uint64_t n, FixFrame fx) {
  constexpr int F = fix_stride<NC>();
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    int32_t w[2 * F];
#pragma unroll
    for (int d = 0; d < 2 * F; ++d)
      w[d] = d < NC ? __float_as_int(emb[r * NC + d]) : 0;
    int32_t *p = fx.rec + r * (2 * F);
#pragma unroll
    for (int d = 0; d < 2 * F; ++d)
      p[d] = w[d];
  }
}

// each row's updates back into the float embedding: p0 + D q, one rounding
template <int NC>
__global__ __launch_bounds__(256) void k_fix_out(float *__restrict__ emb,
                                                 uint64_t n, FixFrame fx,
                                                 const float *__restrict__ rounding_p) {
  constexpr int F = fix_stride<NC>();
  const float q = ldexpf(*rounding_p, -24);
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const int32_t *p = fx.rec + r * (2 * F);
    int32_t D[NC];
    fix_decode<NC>(p + F, D);
#pragma unroll
    for (int d = 0; d < NC; ++d)
      if (D[d] != 0)
        emb[r * NC + d] = __int_as_float(p[d]) + (float)D[d] * q;
  }
}

__device__ inline void flush_sgd_stats(unsigned long long *stats,
                                       unsigned long long c_pos,
                                       unsigned long long c_neg) {
  if (!stats)
    return;
  for (int off = 32; off > 0; off >>= 1) {
    c_pos += __shfl_xor(c_pos, off);
    c_neg += __shfl_xor(c_neg, off);
  }
  if ((threadIdx.x & 63) == 0 && (c_pos | c_neg)) {
    atomicAdd(stats + 3, c_pos);
    atomicAdd(stats + 4, c_neg);
  }
}

// One edge of a wave-wide step with the schedule in memory (k_sgd_epoch,
// k_sgd_stream): sample it if due, then add the wave's contributions.
template <int NC>
__device__ inline void sgd_step(float *__restrict__ emb,
                                const int32_t *__restrict__ head,
                                const int32_t *__restrict__ tail, uint64_t e,
                                uint64_t nnz, uint32_t n,
                                const float *__restrict__ eps,
                                float *__restrict__ nxt,
                                float *__restrict__ nneg, float alpha,
                                int epoch, float gamma, float a, float b,
                                float nsr_inv, uint64_t seed, float rounding,
                                unsigned long long &c_pos,
                                unsigned long long &c_neg) {
  const int32_t h0 = e < nnz ? head[e] : -1;
  const bool valid = h0 >= 0; // an empty slot of the COO has head -1
  const int32_t hv = valid ? h0 : -1 - (int32_t)(threadIdx.x & 63);
  const int32_t tv = valid ? tail[e] : -1 - (int32_t)(threadIdx.x & 63);
  float gh[NC], gt[NC];
#pragma unroll
  for (int d = 0; d < NC; ++d)
    gh[d] = gt[d] = 0.0f;
  if (valid) {
    float nx = nxt[e];
    if (!(nx > (float)epoch)) {
      float nn = nneg[e];
      const float ep = eps[e];
      const float ep_neg = ep * nsr_inv;
      uint32_t dr[8];
      neg_draws(seed, e, dr);
      EdgeLoads<NC> L;
      sgd_gather<NC>(emb, hv, tv, n, ep_neg, nn, epoch, dr, L);
      sgd_update<NC>(emb, hv, e, n, ep, nx, nn, ep_neg, alpha, gamma, a, b,
                     seed, rounding, L, gh, gt, c_pos, c_neg);
      nxt[e] = nx;
      nneg[e] = nn;
    }
  }
  agg_add<NC>(emb, agg_of(hv, tv), hv, tv, gh, gt);
}

// One epoch of the Hogwild edge sampler, one launch per epoch (used when the
// edge set is too large for the register-resident kernel below).
template <int NC>
__global__ __launch_bounds__(256) void k_sgd_epoch(
    float *__restrict__ emb, const int32_t *__restrict__ head,
    const int32_t *__restrict__ tail, const uint32_t *__restrict__ nnz_p,
    uint32_t n, const float *__restrict__ eps, float *__restrict__ nxt,
    float *__restrict__ nneg, float alpha, int epoch, float gamma, float a,
    float b, float nsr_inv, uint64_t seed, const float *__restrict__ rounding_p,
    unsigned long long *__restrict__ stats) {
  const uint64_t nnz = *nnz_p;
  const float rounding = *rounding_p;
  unsigned long long c_pos = 0, c_neg = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // whole waves iterate together (seg_add is a wave operation)
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t base = t0 - (threadIdx.x & 63); base < nnz; base += stride) {
    const uint64_t e = base + (threadIdx.x & 63);
    sgd_step<NC>(emb, head, tail, e, nnz, n, eps, nxt, nneg, alpha, epoch,
                 gamma, a, b, nsr_inv, seed, rounding, c_pos, c_neg);
  }
  flush_sgd_stats(stats, c_pos, c_neg);
}

// Epoch boundary of the fused SGD kernels: one monotonically increasing
// arrival counter, incremented once per block per epoch.  A block finishing
// epoch e may start epoch e + 1 once every block has finished epoch e - lag
// (counter >= (e - lag + 1) * nblocks); lag 0 is a full grid barrier.  The
// block barriers are plain s_barrier (no release fence) and the arrival is a
// relaxed atomic: the Hogwild updates need no ordering between epochs, so the
// epoch's atomic updates stay in flight across the boundary.  Every thread
// reads the counter at the start of the epoch (`peeked`, one request per
// wave) and uses it only at the boundary, so the read's latency overlaps the
// epoch's own memory accesses and the boundary usually waits for nothing.
// The wait is bounded: a boundary that times out sets bar[2] (the host
// reports it), the later boundaries then stop waiting within 1024 polls, and
// the kernel still runs to completion.
__device__ inline unsigned bar_peek(unsigned *bar) {
  return __hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void epoch_end(unsigned *bar, unsigned nblocks, int epoch,
                                 int lag, unsigned peeked, unsigned &seen) {
  __builtin_amdgcn_s_barrier();
  if (threadIdx.x == 0) {
    seen = max(seen, peeked);
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (epoch >= lag) {
      const unsigned target = (unsigned)(epoch - lag + 1) * nblocks;
      unsigned spins = 0;
      while (seen < target) {
        seen = bar_peek(bar);
        if (seen >= target)
          break;
        __builtin_amdgcn_s_sleep(1);
        if ((++spins & 1023u) == 0 &&
            (spins > (1u << 22) || bar_peek(bar + 2))) {
          __hip_atomic_store(bar + 2, 1u, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
  }
  __builtin_amdgcn_s_barrier();
}

// All epochs in one cooperative launch: each thread keeps EPT edges' ids and
// schedule in registers for the whole run, and epoch_end separates the epochs
// (the reference launches one kernel per epoch, :785-793).
template <int NC, int EPT>
__global__ __launch_bounds__(256) void k_sgd_fused(
    float *__restrict__ emb, const int32_t *__restrict__ head,
    const int32_t *__restrict__ tail, const uint32_t *__restrict__ nnz_p,
    uint32_t n, const float *__restrict__ eps, int n_epochs, float alpha0,
    float gamma, float a, float b, float nsr_inv, uint64_t seed,
    const float *__restrict__ rounding_p, unsigned *__restrict__ bar,
    unsigned long long *__restrict__ stats, int lag) {
  const uint64_t nnz = *nnz_p;
  const float rounding = *rounding_p;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int32_t hh[EPT], tt[EPT];
  float ep[EPT], nx[EPT], nn[EPT];
  uint32_t dr[EPT][8]; // this epoch's first 8 negative-sample draws
  Agg ag[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const uint64_t e = t0 + (uint64_t)q * stride;
    if (e < nnz && head[e] >= 0) {
      hh[q] = head[e];
      tt[q] = tail[e];
      ep[q] = eps[e];
      nx[q] = ep[q];
      nn[q] = ep[q] * nsr_inv;
    } else {
      hh[q] = tt[q] = -1 - (int32_t)(threadIdx.x & 63);
      ep[q] = -1.0f;
      nx[q] = INFINITY;
      nn[q] = 0.0f;
    }
    ag[q] = agg_of(hh[q], tt[q]);
    neg_draws(seed, e, dr[q]);
  }
  unsigned long long c_pos = 0, c_neg = 0;
  unsigned seen = 0;
  float alpha = alpha0;
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const unsigned peeked = bar_peek(bar);
    const uint64_t es = seed + (uint64_t)epoch;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const uint64_t e = t0 + (uint64_t)q * stride;
      const bool act = !(nx[q] > (float)epoch);
      const float ep_neg = ep[q] * nsr_inv;
      EdgeLoads<NC> L;
      if (act)
        sgd_gather<NC>(emb, hh[q], tt[q], n, ep_neg, nn[q], epoch, dr[q], L);
      // next epoch's draws, by every lane while the loads are in flight
      neg_draws(es + 1, e, dr[q]);
      float gh[NC], gt[NC];
#pragma unroll
      for (int d = 0; d < NC; ++d)
        gh[d] = gt[d] = 0.0f;
      if (act)
        sgd_update<NC>(emb, hh[q], e, n, ep[q], nx[q], nn[q], ep_neg, alpha,
                       gamma, a, b, es, rounding, L, gh, gt, c_pos, c_neg);
      agg_add<NC>(emb, ag[q], hh[q], tt[q], gh, gt);
    }
    alpha = alpha0 * (1.0f - (float)epoch / (float)n_epochs);
    epoch_end(bar, gridDim.x, epoch, lag, peeked, seen);
  }
  flush_sgd_stats(stats, c_pos, c_neg);
}

// The wave's arrival slots, read without a branch: lanes past nw read the
// last slot, and slot_min masks them out where the values are used (so that
// the loads stay in flight until then).
__device__ inline void poll_slots(const unsigned *__restrict__ slots,
                                  unsigned nw, int lane, unsigned sv[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const unsigned idx = (unsigned)lane + 64u * i;
    sv[i] = __hip_atomic_load(slots + min(idx, nw - 1u), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
  }
}

// the smallest slot value over the wave
__device__ inline unsigned slot_min(const unsigned sv[8], unsigned nw,
                                    int lane) {
  unsigned m = ~0u;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const unsigned idx = (unsigned)lane + 64u * i;
    m = min(m, sv[i] | (0u - (unsigned)(idx >= nw)));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    m = min(m, (unsigned)__shfl_xor((int)m, off, 64));
  return m;
}

// All epochs in one cooperative launch for edge sets of at most 512 waves
// (one edge per thread, one wave per block): the per-cycle embedding
// batches.  With a few hundred waves on 1024 SIMDs an epoch costs about one
// wave's instruction stream plus what its memory operations cannot hide.
//
// Per epoch a wave issues, in this order: the next epoch's gathers (two
// epochs in flight, so that the loads overlap the arithmetic), then the
// previous epoch's row updates (device-scope float atomics, whose cost is
// their device-wide throughput rather than their place in the stream), then
// the epoch's arithmetic.  The buffers of the two epochs in flight swap
// roles from one epoch to the next (the loop is unrolled), so that no
// register copy waits for loads in flight.
//
// The positions an epoch reads were gathered before the previous two
// epochs' updates of this wave were issued: each lane adds its own pair's
// share of those two epochs' updates (what its edge and its twin added to
// the edge's two rows), so that its edge steps from where its own previous
// steps left it.  Updates of the same rows by other lanes are up to two
// epochs old when read, a Hogwild read.  On a conflict-free edge set all of
// this is deterministic (the wave's own issue order decides it).
//
// Epoch boundary: wave w stores the count of epochs it has finished into
// its arrival slot every SGD_POLL epochs, and every SGD_POLL epochs reads
// all slots (coherent stores and reads on a few shared cache lines slow the
// whole device when every wave makes them every epoch).  Epoch e + 1 starts
// once the slots last read show e - lag + 1, lag being the counter kernels'
// lag plus SGD_POLL; a boundary that the last read does not clear reads the
// slots again until it does.  A stored count trails its wave by at most
// SGD_POLL - 1 epochs, so the slowest wave never waits and every wave is at
// most lag + 2 (SGD_POLL - 1) epochs ahead of it.  The wait is bounded
// (*timeout is set, as epoch_end's bar[2]).
// the smallest of eight values per lane over the wave: DPP row shifts within
// each 16-lane row, then the four rows' minima (no LDS round trips)
__device__ inline unsigned wave_min_dpp(const unsigned v[8]) {
  unsigned m = v[0];
#pragma unroll
  for (int i = 1; i < 8; ++i)
    m = min(m, v[i]);
  // row_shr:1, 2, 4, 8 (a lane without a source keeps its own value)
  m = min(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x111, 0xf,
                                                   0xf, false));
  m = min(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x112, 0xf,
                                                   0xf, false));
  m = min(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x114, 0xf,
                                                   0xf, false));
  m = min(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x118, 0xf,
                                                   0xf, false));
  // lane 15 of each row holds its row's minimum
  return min(min(__builtin_amdgcn_readlane(m, 15),
                 __builtin_amdgcn_readlane(m, 31)),
             min(__builtin_amdgcn_readlane(m, 47),
                 __builtin_amdgcn_readlane(m, 63)));
}

constexpr int SGD_POLL = 4;

template <int NC, typename Store>
__device__ __attribute__((always_inline)) inline void sgd_waves_body(
    const Store &st, const int32_t *__restrict__ head,
    const int32_t *__restrict__ tail, uint64_t nnz, uint32_t n,
    const float *__restrict__ eps, int n_epochs, float alpha0, float gamma,
    float a, float b, float nsr_inv, uint64_t seed, float rounding,
    unsigned *__restrict__ slots, unsigned nw, unsigned w,
    unsigned *__restrict__ timeout, unsigned long long *__restrict__ stats,
    int lag) {
  const int lane = threadIdx.x & 63;
  const uint64_t e = (uint64_t)w * 64 + lane;
  lag += SGD_POLL;
  int32_t hh, tt;
  float ep, nx, nn;
  if (e < nnz && head[e] >= 0) {
    hh = head[e];
    tt = tail[e];
    ep = eps[e];
    nx = ep;
    nn = ep * nsr_inv;
  } else {
    hh = tt = -1 - lane;
    ep = -1.0f;
    nx = INFINITY;
    nn = 0.0f;
  }
  const Agg ag = agg_of(hh, tt);
  const float ep_neg = ep * nsr_inv;
  unsigned long long c_pos = 0, c_neg = 0;
  float alpha = alpha0;
  uint32_t dr[8];
  neg_draws(seed, e, dr);
  EdgeLoads<NC> L0, L1;
  // the own pair's share of the last two epochs' updates (1: the previous
  // epoch, 2: the one before), and the previous epoch's contributions, whose
  // row updates are issued in this epoch
  float own1_h[NC], own1_t[NC], own2_h[NC], own2_t[NC], pgh[NC], pgt[NC];
#pragma unroll
  for (int d = 0; d < NC; ++d)
    own1_h[d] = own1_t[d] = own2_h[d] = own2_t[d] = pgh[d] = pgt[d] = 0.0f;
  bool act0 = !(nx > 0.0f), act1 = false;
  sgd_gather_rows<NC>(st.rows, act0, hh, tt, n, ep_neg, nn, 0, dr, L0);
  if (act0)
    advance_schedule(ep, ep_neg, L0.n_neg, nx, nn);
  neg_draws(seed + 1, e, dr);
  // what a step leaves in flight after its gathers, so that the loop is
  // entered with the same pending operations as it repeats with: the row
  // updates (zero here, dropped) and the arrival
  st.add(ag, hh, tt, pgh, pgt);
  __hip_atomic_store(slots + w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned smin = 0; // smallest slot at the last read taken in
  unsigned sv[8];    // a read of the slots in flight
  unsigned spins = 0;
  bool gave_up = false;
  // One epoch: Lc holds its positions, Ln receives the next epoch's.  A
  // `read` step issues a read of the slots, which the next step (`take`)
  // takes in and where the wave stores its arrival: the read was issued
  // after that step's row updates, and in order, so waiting for it at once
  // would wait for them.  The loop runs SGD_POLL epochs per pass with the
  // buffers swapping roles, so that no register copy between them waits for
  // loads in flight.
  auto step = [&](int epoch, EdgeLoads<NC> &Lc, bool &ac, EdgeLoads<NC> &Ln,
                  bool &an, auto read, auto take) {
    const uint64_t es = seed + (uint64_t)epoch;
    // the next epoch's gathers (the last epoch issues them for nothing:
    // every lane reads row 0)
    an = epoch + 1 < n_epochs && !(nx > (float)(epoch + 1));
    sgd_gather_rows<NC>(st.rows, an, hh, tt, n, ep_neg, nn, epoch + 1, dr,
                        Ln);
    if (an)
      advance_schedule(ep, ep_neg, Ln.n_neg, nx, nn);
    // the previous epoch's row updates (zero contributions at epoch 0)
    st.add(ag, hh, tt, pgh, pgt);
    if constexpr (decltype(read)::value)
      poll_slots(slots, nw, lane, sv);
    neg_draws(es + 2, e, dr);
    float gh[NC], gt[NC];
#pragma unroll
    for (int d = 0; d < NC; ++d) {
      gh[d] = gt[d] = 0.0f;
      Lc.cur[d] += own1_h[d] + own2_h[d];
      Lc.oth[d] += own1_t[d] + own2_t[d];
    }
    // The epoch's gathers have landed, on every path (also when no lane
    // samples, where the compiler would otherwise consume them only inside
    // the branch below): a load left pending here would make the next write
    // to its registers wait for every memory operation in flight.
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int d = 0; d < NC; ++d)
        __asm__ volatile("" ::"v"(Lc.neg[q][d]));
#pragma unroll
    for (int d = 0; d < NC; ++d)
      __asm__ volatile("" ::"v"(Lc.cur[d]), "v"(Lc.oth[d]));
    if (ac)
      sgd_apply_rows<NC>(st.rows, hh, e, n, ep_neg, alpha, gamma, a, b, es,
                         rounding, Lc, gh, gt, c_pos, c_neg);
    // this epoch's share of the pair's rows (paired: the head row gets the
    // edge's head and the twin's tail contributions, the tail row the
    // twin's total)
#pragma unroll
    for (int d = 0; d < NC; ++d) {
      const float twin_t = lane_xor1(gt[d]);
      const float pair_h = gh[d] + twin_t;
      own2_h[d] = own1_h[d];
      own2_t[d] = own1_t[d];
      own1_h[d] = ag.paired ? pair_h : gh[d];
      own1_t[d] = ag.paired ? lane_xor1(pair_h) : gt[d];
      pgh[d] = gh[d];
      pgt[d] = gt[d];
    }
    alpha = alpha0 * (1.0f - (float)epoch / (float)n_epochs);
    // boundary: epoch + 1 starts once every slot shows epoch - lag + 1
    if constexpr (decltype(take)::value)
      smin = slot_min(sv, nw, lane);
    if (epoch >= lag && epoch + 1 < n_epochs) {
      const unsigned target = (unsigned)(epoch - lag + 1);
      while (!gave_up && smin < target) {
        // bounded: a timed-out wait anywhere in the launch ends the others
        // within 1024 polls
        if ((++spins & 1023u) == 0 &&
            (spins > (1u << 20) ||
             __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                 timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))) {
          if (lane == 0)
            __hip_atomic_store(timeout, 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
          gave_up = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        unsigned sw[8];
        poll_slots(slots, nw, lane, sw);
        smin = slot_min(sw, nw, lane);
      }
    }
    // arrival, every SGD_POLL epochs (every lane stores the same value)
    if constexpr (decltype(take)::value)
      __hip_atomic_store(slots + w, (unsigned)epoch + 1u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  };
  using P1 = std::integral_constant<bool, true>;
  using P0 = std::integral_constant<bool, false>;
  static_assert(SGD_POLL == 4, "the loop below runs four epochs per pass");
  int epoch = 0;
  for (; epoch + 3 < n_epochs; epoch += 4) {
    step(epoch, L0, act0, L1, act1, P1{}, P0{});
    step(epoch + 1, L1, act1, L0, act0, P0{}, P1{});
    step(epoch + 2, L0, act0, L1, act1, P0{}, P0{});
    step(epoch + 3, L1, act1, L0, act0, P0{}, P0{});
  }
  // the last one to three epochs
  if (epoch < n_epochs)
    step(epoch, L0, act0, L1, act1, P1{}, P0{});
  if (epoch + 1 < n_epochs)
    step(epoch + 1, L1, act1, L0, act0, P0{}, P1{});
  if (epoch + 2 < n_epochs)
    step(epoch + 2, L0, act0, L1, act1, P0{}, P0{});
  // the last epoch's row updates and arrival
  st.add(ag, hh, tt, pgh, pgt);
  __hip_atomic_store(slots + w, (unsigned)n_epochs, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  flush_sgd_stats(stats, c_pos, c_neg);
}

// sgd_waves_body for the batched kernel (several sessions' edge sets in one
// launch, thousands of waves): there the kernel is bound by the memory
// pipeline's throughput rather than by one wave's latency, and the
// latency-hiding form above costs more than it hides (every lane's ten
// gathers and the dropped row updates are issued whether or not the lane
// samples), so each epoch issues only the memory operations it needs:
// the slot reads and the sampling lanes' gathers at the start of the epoch,
// the row updates (global atomics, seg_add / seg2_add) at its end.
template <int NC, typename Store>
__device__ __attribute__((always_inline)) inline void sgd_waves_body_batch(
    const Store &st, const int32_t *__restrict__ head,
    const int32_t *__restrict__ tail, uint64_t nnz, uint32_t n,
    const float *__restrict__ eps, int n_epochs, float alpha0, float gamma,
    float a, float b, float nsr_inv, uint64_t seed, float rounding,
    unsigned *__restrict__ slots, unsigned nw, unsigned w,
    unsigned *__restrict__ timeout, unsigned long long *__restrict__ stats,
    int lag) {
  const int lane = threadIdx.x & 63;
  const uint64_t e = (uint64_t)w * 64 + lane;
  int32_t hh, tt;
  float ep, nx, nn;
  if (e < nnz && head[e] >= 0) {
    hh = head[e];
    tt = tail[e];
    ep = eps[e];
    nx = ep;
    nn = ep * nsr_inv;
  } else {
    hh = tt = -1 - lane;
    ep = -1.0f;
    nx = INFINITY;
    nn = 0.0f;
  }
  const Agg ag = agg_of(hh, tt);
  uint32_t dr[8];
  neg_draws(seed, e, dr);
  unsigned long long c_pos = 0, c_neg = 0;
  float alpha = alpha0;
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    unsigned sv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const unsigned idx = (unsigned)lane + 64u * i;
      sv[i] = idx < nw ? __hip_atomic_load(slots + idx, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT)
                       : ~0u;
    }
    const uint64_t es = seed + (uint64_t)epoch;
    const bool act = !(nx > (float)epoch);
    const float ep_neg = ep * nsr_inv;
    EdgeLoads<NC> L;
    if (act)
      sgd_gather_r<NC>(st.rows, hh, tt, n, ep_neg, nn, epoch, dr, L);
    neg_draws(es + 1, e, dr);
    float gh[NC], gt[NC];
#pragma unroll
    for (int d = 0; d < NC; ++d)
      gh[d] = gt[d] = 0.0f;
    if (act) {
      sgd_apply_rows<NC>(st.rows, hh, e, n, ep_neg, alpha, gamma, a, b, es,
                         rounding, L, gh, gt, c_pos, c_neg);
      advance_schedule(ep, ep_neg, L.n_neg, nx, nn);
    }
    st.add(ag, hh, tt, gh, gt);
    alpha = alpha0 * (1.0f - (float)epoch / (float)n_epochs);
    // boundary
    if (lane == 0)
      __hip_atomic_store(slots + w, (unsigned)epoch + 1u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    if (epoch >= lag && epoch + 1 < n_epochs) {
      const unsigned target = (unsigned)(epoch - lag + 1);
      unsigned spins = 0;
      for (;;) {
        if (wave_min_dpp(sv) >= target)
          break;
        // bounded: a timed-out wait anywhere in the launch ends the others
        // within 1024 polls
        if ((++spins & 1023u) == 0 &&
            (spins > (1u << 20) ||
             __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                 timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))) {
          if (lane == 0)
            __hip_atomic_store(timeout, 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const unsigned idx = (unsigned)lane + 64u * i;
          sv[i] = idx < nw ? __hip_atomic_load(slots + idx, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)
                           : ~0u;
        }
      }
    }
  }
  flush_sgd_stats(stats, c_pos, c_neg);
}


// On the records when the frame has them (k_fix_in made them), else on the
// float embedding.
template <int NC>
__global__ __launch_bounds__(64) void k_sgd_waves(
    float *__restrict__ emb, const int32_t *__restrict__ head,
    const int32_t *__restrict__ tail, const uint32_t *__restrict__ nnz_p,
    uint32_t n, const float *__restrict__ eps, int n_epochs, float alpha0,
    float gamma, float a, float b, float nsr_inv, uint64_t seed,
    const float *__restrict__ rounding_p, unsigned *__restrict__ bar,
    unsigned long long *__restrict__ stats, int lag, FixFrame fx) {
  if (fx.rec) {
    const FixStore<NC> st = fix_store<NC>(fx.rec, *rounding_p);
    sgd_waves_body<NC>(st, head, tail, *nnz_p, n, eps, n_epochs, alpha0,
                       gamma, a, b, nsr_inv, seed, *rounding_p, bar + 4,
                       gridDim.x, blockIdx.x, bar + 2, stats, lag);
  } else {
    const FloatStore<NC> st{EmbRows<NC>{emb}, emb_rsrc<NC>(emb, n)};
    sgd_waves_body<NC>(st, head, tail, *nnz_p, n, eps, n_epochs, alpha0,
                       gamma, a, b, nsr_inv, seed, *rounding_p, bar + 4,
                       gridDim.x, blockIdx.x, bar + 2, stats, lag);
  }
}

// ------------------------------------------------------- pipelined SGD
// The per-cycle kernel with each group of 64 edges on four waves of one
// block (the four SIMDs of one CU) that hand work to each other through LDS
// slots and counters, with no block barrier:
//   chain wave  the epochs' arithmetic only (each edge's positive step and
//               its sequential negative steps), inputs read from LDS;
//   row wave    gathers the edge's own two rows three epochs ahead, hands
//               them to the chain wave, adds the chain wave's
//               contributions to the rows, and keeps the launch's blocks
//               within the drift bound;
//   draw waves  (two, for the even and the odd epochs) compute the
//               negative-sample draws (Philox's 32-bit multiplies run at a
//               quarter of the vector rate: one epoch's draws cost about as
//               much as the chain wave's epoch), gather the sample rows and
//               hand them to the chain wave.
// With one wave per SIMD an epoch of the single-wave kernel costs that
// wave's whole instruction stream plus the memory latency it cannot hide;
// split, the chain wave's stream is about a third of it, and the scattered
// row loads of the other waves overlap it.  A wave waits for another only
// when it has run ahead of it.
//
// Row wave, tick k (k = -3 .. Ep; Ep = n_epochs rounded up to a multiple of
// 4, the epochs from n_epochs on sample nothing):
//   (b) once the chain wave has finished epoch k - 3, epoch k + 1's two
//       rows (loads issued at tick k - 2) into chain slot (k + 1) mod 4;
//   (c) epoch k + 3's two row loads issued (row 0 for a lane that does not
//       sample), so that (b)'s wait counts exactly the loads after its own;
//   (d) once the chain wave has finished epoch k - 1, its contributions
//       (ring slot (k - 1) mod 4) added to the rows;
//   (e) every fourth tick, the arrival: row updates of epochs < k issued;
//   (f) every fourth tick (k = 0 mod 4) a read of the other blocks' arrival
//       slots is issued, taken in at tick k + 3 (its wait then covers only
//       operations issued at least two ticks earlier), and before (c) the
//       drift bound: epoch k + 3's rows are read once every other block's
//       slot shows k + 3 - lag (read again, at once, until it does).
//       A slot read stays in flight for three ticks: waiting for it at
//       once would wait for every operation issued before it.
// Epoch g's own rows are loaded after the row updates of epoch g - 5 and
// before those of epoch g - 4 (one wave, issue order), so the chain wave
// adds its own pair's share of epochs g - 4 .. g - 1 to what they read and,
// on a conflict-free edge set, follows the sequential trajectory exactly.
// A draw wave loads epoch g's sample rows after the row wave's tick g - 3,
// so after its drift bound for epoch g.  Other blocks' updates are at most
// lag + 1 epochs old when read (a Hogwild read).  A block whose row
// wave is at tick k' has stored at least k' - 4, so with a lag >= 7 the
// row wave furthest behind never waits for another block and the launch
// always progresses.  With four chain slots the row and draw waves hand
// over an epoch while the chain wave is still two epochs before it, so the
// chain wave waits only when they fall behind.  Every wait is bounded: a
// time-out sets *timeout, and the wave's later waits are skipped.
template <int NC> struct PipeSlot {
  float cur[NC][64], oth[NC][64], neg[8][NC][64];
  // n_neg when the lane samples (bits 0..15) and the usable negative
  // samples (bits 16..23: q < min(8, n_neg) and not the head), else -1
  int meta[64];
  int qmax; // 1 + the highest usable sample slot over the wave (0: none)
  int many; // some lane has more than 8 negative samples
};

// fclip with one clamp instruction (v_med3: equal to min then max for every
// finite value; a value that is not finite is 0, as in fclip)
__device__ inline float fclip_med3(float v) {
  return isfinite(v) ? __builtin_amdgcn_fmed3f(v, -4.0f, 4.0f) : 0.0f;
}

// The chain wave's epoch: sgd_apply_rows without branches per lane.  Every
// lane runs the positive step and the wave's negative-sample steps q <
// qmax, and a lane keeps only the steps it takes (act, usable sample q):
// the same operations as sgd_apply_rows on the lanes that take them.  At
// d^2 = 0 the coefficient is computed and discarded (the select keeps 0,
// as sgd_apply_rows' branch does).
template <int NC, typename Rows>
__device__ inline void pipe_apply(const Rows &rw, int32_t j, uint64_t e,
                                  uint32_t n, float ep_neg, float alpha,
                                  float gamma, float a, float b,
                                  uint64_t seed, float rounding, bool act,
                                  int n_neg, int vm, int qmax, bool many,
                                  EdgeLoads<NC> &L, float gh[NC],
                                  float gt[NC], unsigned &c_pos,
                                  unsigned &c_neg) {
  c_pos += act ? 1u : 0u;
  c_neg += (unsigned)__popc((unsigned)vm);
  float g[NC];
  float d2 = 0.0f;
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    const float t = L.cur[d] - L.oth[d];
    d2 += t * t;
  }
  const float pb = pow_pos(d2, b);
  const float attr_c =
      (-2.0f * a * b * pb) * __builtin_amdgcn_rcpf(d2 * (a * pb + 1.0f));
  const float attr = d2 > 0.0f ? attr_c : 0.0f;
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    const float gp = fclip_med3(attr * (L.cur[d] - L.oth[d])) * alpha;
    g[d] = act ? gp : 0.0f;
    L.cur[d] += g[d];
    gt[d] = (rounding + -g[d]) - rounding;
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if (q >= qmax)
      break;
    const bool take = (vm >> q) & 1;
    float d2n = 0.0f;
#pragma unroll
    for (int d = 0; d < NC; ++d) {
      const float t = L.cur[d] - L.neg[q][d];
      d2n += t * t;
    }
    const float rep_c =
        (2.0f * gamma * b) *
        __builtin_amdgcn_rcpf((0.001f + d2n) * (a * pow_pos(d2n, b) + 1.0f));
    const float rep = d2n > 0.0f ? rep_c : 0.0f;
#pragma unroll
    for (int d = 0; d < NC; ++d) {
      const float gg = rep > 0.0f
                           ? fclip_med3(rep * (L.cur[d] - L.neg[q][d])) * alpha
                           : 4.0f * alpha;
      const float gk = take ? gg : 0.0f;
      L.cur[d] += gk;
      g[d] += gk;
    }
  }
  if (many && act && n_neg > 8) {
    // more than 8 negative samples (an edge skipped many epochs): the
    // stream continues at Philox block 2
    Rng rng(seed, (uint32_t)e, (uint32_t)(e >> 32), TAG_NEG);
    rng.ctr[2] = 2u;
    for (int p = 8; p < n_neg; ++p) {
      const uint32_t r = __umulhi(rng.next(), n);
      if ((int32_t)r == j)
        continue;
      ++c_neg;
      float neg[NC];
      rw.load(r, neg);
      neg_step<NC>(L.cur, g, neg, alpha, gamma, a, b);
    }
  }
#pragma unroll
  for (int d = 0; d < NC; ++d)
    gh[d] = (rounding + g[d]) - rounding;
}
struct PipeSync {
  int c_done;    // epochs the chain wave has finished
  int m_hand;    // epochs whose own rows are in their chain slot
  int m_done;    // row-wave ticks finished, + 4 (tick k done: k + 4)
  int w_hand[2]; // draw wave h: its last epoch in its chain slot, + 1
};
// the block's LDS state (one allocation per kernel, whichever row store)
template <int NC> struct PipeShared {
  PipeSlot<NC> ps[4];       // chain slots, epoch g in slot g mod 4
  float ctr[4][2 * NC][64]; // the chain wave's contributions, g mod 4
  PipeSync sy;
};
// the drift bound's default lag (GPUMAP_AMD_PIPE_LAG overrides it, at least
// PIPE_LAG_MIN).  In steady state a slot read is up to six ticks old when
// used and each slot trails its block by up to four ticks, so a small lag
// repeats reads at once; a large one lets blocks drift apart, which was
// slower too, and slots that move in steps of four make the lag's residue
// mod 4 matter (70,000 x 784 per-cycle sets, ms per launch: lag 7 0.44,
// 8 0.37, 9 0.335, 10 0.355, 11 0.42, 12 0.39, 16 0.41)
constexpr int PIPE_LAG = 9;
constexpr int PIPE_LAG_MIN = 7;
constexpr unsigned PIPE_SPIN_MAX = 1u << 22;

// wait until *p >= target, a counter that another wave of the block raises
__device__ inline void pipe_wait(const int *p, int target, bool &gave_up,
                                 unsigned *__restrict__ timeout) {
  if (gave_up)
    return;
  unsigned spins = 0;
  while (__builtin_amdgcn_readfirstlane(__hip_atomic_load(
             p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) < target) {
    if (++spins > PIPE_SPIN_MAX) {
      if ((threadIdx.x & 63) == 0)
        __hip_atomic_store(timeout, 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      gave_up = true;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  // the slot reads below stay after the counter read (compiler order; the
  // wave has waited for the counter's value)
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// both counters, read together (one LDS round trip per poll)
__device__ inline void pipe_wait2(const int *p1, int t1, const int *p2, int t2,
                                  bool &gave_up,
                                  unsigned *__restrict__ timeout) {
  if (gave_up)
    return;
  unsigned spins = 0;
  for (;;) {
    const int v1 = __hip_atomic_load(p1, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
    const int v2 = __hip_atomic_load(p2, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
    if (__builtin_amdgcn_readfirstlane(v1 >= t1 && v2 >= t2))
      break;
    if (++spins > PIPE_SPIN_MAX) {
      if ((threadIdx.x & 63) == 0)
        __hip_atomic_store(timeout, 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      gave_up = true;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// The slot writes before it become visible no later than the counter: one
// wave's LDS operations are performed in issue order, so only the
// compiler's order needs holding (no wait for the writes).
__device__ inline void pipe_post(int *p, int v) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// the smallest of a read of the slots over the wave, slot w left out
__device__ inline unsigned pipe_slots_min(const unsigned sv[8], unsigned nw,
                                          unsigned w, int lane) {
  unsigned v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const unsigned idx = (unsigned)lane + 64u * i;
    v[i] = (idx >= nw || idx == w) ? ~0u : sv[i];
  }
  return wave_min_dpp(v);
}

// the smallest arrival slot of the other blocks (slot w left out), read now
__device__ inline unsigned pipe_slot_min(const unsigned *__restrict__ slots,
                                         unsigned nw, unsigned w, int lane) {
  unsigned sv[8];
  poll_slots(slots, nw, lane, sv);
  return pipe_slots_min(sv, nw, w, lane);
}

template <int NC, typename Store>
__device__ __attribute__((always_inline)) inline void sgd_pipe_body(
    const Store &st, const int32_t *__restrict__ head,
    const int32_t *__restrict__ tail, uint64_t nnz, uint32_t n,
    const float *__restrict__ eps, int n_epochs, float alpha0, float gamma,
    float a, float b, float nsr_inv, uint64_t seed, float rounding,
    unsigned *__restrict__ slots, unsigned nw, unsigned w,
    unsigned *__restrict__ timeout, unsigned long long *__restrict__ stats,
    int lag, PipeShared<NC> &sh) {
  PipeSlot<NC>(&ps)[4] = sh.ps;
  float(&ctr)[4][2 * NC][64] = sh.ctr;
  PipeSync &sy = sh.sy;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t e = (uint64_t)w * 64 + lane;
  int32_t hh, tt;
  float ep;
  if (e < nnz && head[e] >= 0) {
    hh = head[e];
    tt = tail[e];
    ep = eps[e];
  } else {
    hh = tt = -1 - lane;
    ep = -1.0f;
  }
  const float ep_neg = ep * nsr_inv;
  const int E = n_epochs;
  const int Ep = (E + 3) & ~3;
  // every edge's schedule, run by the row wave and both draw waves alike:
  // whether epoch g samples, and its negative-sample count
  float nx = ep > 0.0f ? ep : INFINITY, nn = ep * nsr_inv;
  auto schedule = [&](int g, int &n_neg) {
    const bool act = g < E && ep > 0.0f && !(nx > (float)g);
    n_neg = act ? (int)(((float)g - nn) / ep_neg) : 0;
    if (act)
      advance_schedule(ep, ep_neg, n_neg, nx, nn);
    return act;
  };
  if (threadIdx.x == 0)
    sy = PipeSync{0, 0, 0, {0, 0}};
  if (wv == 1) {
    // the row wave's adds of ticks -3 .. 0 (epochs -4 .. -1) add zeros
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int d = 0; d < 2 * NC; ++d)
        ctr[c][d][lane] = 0.0f;
    __hip_atomic_store(slots + w, 0u, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  bool gave_up = false;
  if (wv == 0) {
    // ---------------------------------------------------------- chain
    const Agg ag = agg_of(hh, tt);
    unsigned c_pos = 0, c_neg = 0;
    float own_h[4][NC], own_t[4][NC];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int d = 0; d < NC; ++d)
        own_h[k][d] = own_t[k][d] = 0.0f;
    float alpha = alpha0;
    auto step = [&](int g, auto ring) {
      constexpr int R = decltype(ring)::value; // g mod 4
      pipe_wait2(&sy.m_hand, g + 1, &sy.w_hand[g & 1], g + 1, gave_up,
                 timeout);
      const PipeSlot<NC> &sl = ps[g & 3];
      const int meta = sl.meta[lane];
      const int qmax = __builtin_amdgcn_readfirstlane(sl.qmax);
      const bool many = __builtin_amdgcn_readfirstlane(sl.many) != 0;
      const bool act = meta >= 0;
      EdgeLoads<NC> L;
#pragma unroll
      for (int d = 0; d < NC; ++d) {
        L.cur[d] = sl.cur[d][lane] + own_h[0][d] + own_h[1][d] + own_h[2][d] +
                   own_h[3][d];
        L.oth[d] = sl.oth[d][lane] + own_t[0][d] + own_t[1][d] + own_t[2][d] +
                   own_t[3][d];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int d = 0; d < NC; ++d)
          L.neg[q][d] = sl.neg[q][d][lane];
      // every slot value read now, with one wait: left to itself the
      // compiler moves each sample's read into its step (behind the step's
      // uniform exit) and waits an LDS round trip per step
#pragma unroll
      for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int d = 0; d < NC; ++d)
          __asm__ volatile("" ::"v"(L.neg[q][d]));
      float gh[NC], gt[NC];
      pipe_apply<NC>(st.rows, hh, e, n, ep_neg, alpha, gamma, a, b,
                     seed + (uint64_t)g, rounding, act, act ? meta & 0xffff : 0,
                     act ? (meta >> 16) & 0xff : 0, qmax, many, L, gh, gt,
                     c_pos, c_neg);
#pragma unroll
      for (int d = 0; d < NC; ++d) {
        ctr[R][d][lane] = gh[d];
        ctr[R][NC + d][lane] = gt[d];
        // this epoch's share of the pair's rows (paired: the head row gets
        // the edge's head and the twin's tail contributions, the tail row
        // the twin's total)
        const float twin_t = lane_xor1(gt[d]);
        const float pair_h = gh[d] + twin_t;
        own_h[R][d] = ag.paired ? pair_h : gh[d];
        own_t[R][d] = ag.paired ? lane_xor1(pair_h) : gt[d];
      }
      alpha = alpha0 * (1.0f - (float)g / (float)E);
      pipe_post(&sy.c_done, g + 1);
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    for (int g = 0; g < Ep; g += 4) {
      step(g, I0{});
      step(g + 1, I1{});
      step(g + 2, I2{});
      step(g + 3, I3{});
    }
    flush_sgd_stats(stats, c_pos, c_neg);
  } else if (wv == 1) {
    // ------------------------------------------------------------ rows
    const Agg ag = agg_of(hh, tt);
    using Rows = std::remove_cv_t<std::remove_reference_t<decltype(st.rows)>>;
    constexpr int RW = Rows::RAW;
    // two epochs' own rows (cur, oth) as raw words, in flight from (c) to
    // (b)
    int32_t G0[2][RW], G1[2][RW];
    int next_g = 0; // the next epoch whose schedule is to run
    unsigned sv[8];    // a read of the slots in flight
    unsigned smin = 0; // the smallest other slot at the last read taken in
    unsigned spins = 0;
    bool slots_gave_up = false;
    // tick phase k mod 4 = Q: (e) at Q = 3, the slot read issued at Q = 0
    // and taken in at Q = 3; before the loop Q = -1 (no slot work) and -2
    // (the slot's starting value stored again, so that the loop is entered
    // with the memory operations in flight that it repeats with)
    auto tick = [&](int k, auto par, auto hb, auto hc, auto hd, auto phase) {
      constexpr int P = decltype(par)::value;   // k mod 2
      constexpr int Q = decltype(phase)::value; // k mod 4, or -1, -2
      int32_t(&G)[2][RW] = P ? G1 : G0;
      // (b)
      if constexpr (decltype(hb)::value) {
        if (k + 1 < Ep) {
          pipe_wait(&sy.c_done, k - 2, gave_up, timeout);
          PipeSlot<NC> &sl = ps[(k + 1) & 3];
          float v[NC];
          st.rows.decode(G[0], v);
#pragma unroll
          for (int d = 0; d < NC; ++d)
            sl.cur[d][lane] = v[d];
          st.rows.decode(G[1], v);
#pragma unroll
          for (int d = 0; d < NC; ++d)
            sl.oth[d][lane] = v[d];
          pipe_post(&sy.m_hand, k + 2);
        }
      }
      // (f): the read issued at tick k - 3 taken in, and the drift bound
      if constexpr (Q == 3)
        smin = pipe_slots_min(sv, nw, w, lane);
      if constexpr (decltype(hc)::value) {
        if (k + 3 > lag) {
          const unsigned target = (unsigned)(k + 3 - lag);
          while (!slots_gave_up && smin < target) {
            if ((++spins & 1023u) == 0 &&
                (spins > (1u << 20) ||
                 __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                     timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))) {
              if (lane == 0)
                __hip_atomic_store(timeout, 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
              slots_gave_up = true;
              break;
            }
            smin = pipe_slot_min(slots, nw, w, lane);
            if (smin < target)
              __builtin_amdgcn_s_sleep(1);
          }
        }
        // (c): the schedule through epoch k + 3
        bool act = false;
        for (; next_g <= k + 3; ++next_g) {
          int n_neg;
          act = schedule(next_g, n_neg);
        }
        st.rows.load_raw(act ? (uint64_t)hh : 0, G[0]);
        st.rows.load_raw(act ? (uint64_t)tt : 0, G[1]);
      }
      // (d)
      if constexpr (decltype(hd)::value) {
        pipe_wait(&sy.c_done, k, gave_up, timeout);
        float gh[NC], gt[NC];
        const int c = (k - 1) & 3;
#pragma unroll
        for (int d = 0; d < NC; ++d) {
          gh[d] = ctr[c][d][lane];
          gt[d] = ctr[c][NC + d][lane];
        }
        st.add(ag, hh, tt, gh, gt);
      }
      // (e)
      if constexpr (Q == 3 || Q == -2)
        __hip_atomic_store(slots + w, Q == 3 ? (unsigned)k : 0u,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // (f): the slot read, issued after this tick's loads and updates
      if constexpr (Q == 0)
        poll_slots(slots, nw, lane, sv);
      pipe_post(&sy.m_done, k + 4);
    };
    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;
    using T = std::integral_constant<bool, true>;
    using F = std::integral_constant<bool, false>;
    using QN = std::integral_constant<int, -1>;
    using QS = std::integral_constant<int, -2>;
    using Q0 = std::integral_constant<int, 0>;
    using Q1 = std::integral_constant<int, 1>;
    using Q2 = std::integral_constant<int, 2>;
    using Q3 = std::integral_constant<int, 3>;
    tick(-3, B1{}, F{}, T{}, T{}, QN{});
    tick(-2, B0{}, F{}, T{}, T{}, QN{});
    tick(-1, B1{}, T{}, T{}, T{}, QS{});
    for (int k = 0; k < Ep; k += 4) {
      tick(k, B0{}, T{}, T{}, T{}, Q0{});
      tick(k + 1, B1{}, T{}, T{}, T{}, Q1{});
      tick(k + 2, B0{}, T{}, T{}, T{}, Q2{});
      tick(k + 3, B1{}, T{}, T{}, T{}, Q3{});
    }
    tick(Ep, B0{}, F{}, F{}, T{}, QN{});
    __hip_atomic_store(slots + w, (unsigned)Ep + 8u, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  } else {
    // ----------------------------------------------------------- draws
    // draw wave h takes the epochs g = h mod 2 and runs the schedule of
    // every epoch
    const int h = wv - 2;
    using Rows = std::remove_cv_t<std::remove_reference_t<decltype(st.rows)>>;
    constexpr int RW = Rows::RAW;
    for (int g = 0; g < Ep; ++g) {
      int n_neg;
      const bool act = schedule(g, n_neg);
      if ((g & 1) != h)
        continue;
      uint32_t dr[8];
      neg_draws(seed + (uint64_t)g, e, dr);
      const int cnt = min(8, n_neg);
      int vm = 0;
      uint32_t r[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint32_t rq = __umulhi(dr[q], n);
        r[q] = q < cnt ? rq : 0u;
        vm |= (q < cnt && (int32_t)rq != hh) ? (1 << q) : 0;
      }
      // the sample rows are read after the row wave's drift bound for
      // epoch g (its tick g - 3)
      pipe_wait(&sy.m_done, g + 1, gave_up, timeout);
      int32_t Gn[8][RW];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        st.rows.load_raw(r[q], Gn[q]);
      // chain slot g mod 4 is free once the chain wave has finished g - 4
      pipe_wait(&sy.c_done, g - 3, gave_up, timeout);
      PipeSlot<NC> &sl = ps[g & 3];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        float v[NC];
        st.rows.decode(Gn[q], v);
#pragma unroll
        for (int d = 0; d < NC; ++d)
          sl.neg[q][d][lane] = v[d];
      }
      sl.meta[lane] = act ? (n_neg | (vm << 16)) : -1;
      // wave-wide values, computed by every lane
      int qmax = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        qmax = __ballot((vm >> q) & 1) ? q + 1 : qmax;
      const int many = __ballot(act && n_neg > 8) ? 1 : 0;
      if (lane == 0) {
        sl.qmax = qmax;
        sl.many = many;
      }
      pipe_post(&sy.w_hand[h], g + 1);
    }
  }
}

template <int NC>
__global__ __launch_bounds__(256) void k_sgd_pipe(
    float *__restrict__ emb, const int32_t *__restrict__ head,
    const int32_t *__restrict__ tail, const uint32_t *__restrict__ nnz_p,
    uint32_t n, const float *__restrict__ eps, int n_epochs, float alpha0,
    float gamma, float a, float b, float nsr_inv, uint64_t seed,
    const float *__restrict__ rounding_p, unsigned *__restrict__ bar,
    unsigned long long *__restrict__ stats, int lag, FixFrame fx) {
  __shared__ PipeShared<NC> sh;
  if (fx.rec) {
    const FixStoreBuf<NC> st = fix_store_buf<NC>(fx.rec, *rounding_p, n);
    sgd_pipe_body<NC>(st, head, tail, *nnz_p, n, eps, n_epochs, alpha0, gamma,
                      a, b, nsr_inv, seed, *rounding_p, bar + 4, gridDim.x,
                      blockIdx.x, bar + 2, stats, lag, sh);
  } else {
    const FloatStore<NC> st{EmbRows<NC>{emb}, emb_rsrc<NC>(emb, n)};
    sgd_pipe_body<NC>(st, head, tail, *nnz_p, n, eps, n_epochs, alpha0, gamma,
                      a, b, nsr_inv, seed, *rounding_p, bar + 4, gridDim.x,
                      blockIdx.x, bar + 2, stats, lag, sh);
  }
}

// Block b runs wave b - job.wave0 of the job whose wave range holds b; the
// slots of job j are bar[4 + job.wave0 ...], the timeout flag bar[2].
template <int NC>
__global__ __launch_bounds__(64) void k_sgd_waves_multi(SgdBatch batch,
                                                        unsigned *__restrict__ bar,
                                                        int lag) {
  const unsigned b = blockIdx.x;
  int j = 0;
  for (int q = 1; q < batch.njobs; ++q)
    if (b >= batch.job[q].wave0)
      j = q;
  const SgdJob &J = batch.job[j];
  if (b >= J.wave0 + J.nwaves)
    return; // not part of any job (never launched so)
  if (J.rec) {
    const FixStore<NC> st = fix_store<NC>(J.rec, *J.rounding_p);
    sgd_waves_body_batch<NC>(st, J.head, J.tail, *J.nnz_p, J.n, J.eps,
                             J.n_epochs, J.alpha0, J.gamma, J.a, J.b,
                             J.nsr_inv, J.seed, *J.rounding_p,
                             bar + 4 + J.wave0, J.nwaves, b - J.wave0,
                             bar + 2, J.stats, lag);
  } else {
    const FloatStoreBranchy<NC> st{EmbRows<NC>{J.emb}, J.emb};
    sgd_waves_body_batch<NC>(st, J.head, J.tail, *J.nnz_p, J.n, J.eps,
                             J.n_epochs, J.alpha0, J.gamma, J.a, J.b,
                             J.nsr_inv, J.seed, *J.rounding_p,
                             bar + 4 + J.wave0, J.nwaves, b - J.wave0,
                             bar + 2, J.stats, lag);
  }
}

// A slot's schedule as the streaming kernel reads it: all five fields in one
// round trip (an index past the edge set reads the last slot and is marked
// empty, head -1).
struct SlotSched {
  int32_t h, t;
  float nx, nn, ep;
};

__device__ inline SlotSched slot_load(const int32_t *__restrict__ head,
                                      const int32_t *__restrict__ tail,
                                      const float *__restrict__ nxt,
                                      const float *__restrict__ nneg,
                                      const float *__restrict__ eps,
                                      uint64_t e, uint64_t nnz) {
  const uint64_t q = e < nnz ? e : nnz - 1;
  SlotSched r;
  const int32_t h = head[q];
  r.t = tail[q];
  r.nx = nxt[q];
  r.nn = nneg[q];
  r.ep = eps[q];
  r.h = e < nnz ? h : -1;
  return r;
}

// All epochs in one cooperative launch for edge sets larger than the
// register-resident grid: the schedule stays in memory (eps, nxt, nneg as in
// k_sgd_epoch) and each wave walks a grid-strided range of slots per epoch.
// The walk is pipelined: the next slot's schedule is loaded before this
// slot's work, and everything issued after it -- the sample's ten gathers
// (sgd_gather_all), the schedule's write-back and the row updates (buffer
// operations whose inactive lanes pass an offset past the buffer) -- is a
// fixed set of operations, so that waiting for a slot's schedule waits for
// nothing issued later.  Each slot then costs one memory round trip (its
// gathers) instead of four dependent ones.
template <int NC, typename Store>
__device__ __attribute__((always_inline)) inline void sgd_stream_body(
    const Store &st, const int32_t *__restrict__ head,
    const int32_t *__restrict__ tail, const uint32_t *__restrict__ nnz_p,
    uint32_t n, const float *__restrict__ eps, float *__restrict__ nxt,
    float *__restrict__ nneg, int n_epochs, float alpha0, float gamma, float a,
    float b, float nsr_inv, uint64_t seed, const float *__restrict__ rounding_p,
    unsigned *__restrict__ bar, unsigned long long *__restrict__ stats,
    int lag) {
  const uint64_t nnz = *nnz_p;
  const float rounding = *rounding_p;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  // whole waves iterate together (the row updates are wave operations); the
  // wave's first slot in a scalar register, so that the per-slot buffers
  // below are uniform
  const uint64_t w0 = (uint64_t)blockIdx.x * blockDim.x +
                      (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x - lane);
  unsigned long long c_pos = 0, c_neg = 0;
  unsigned seen = 0;
  float alpha = alpha0;
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const unsigned peeked = bar_peek(bar);
    const uint64_t es = seed + (uint64_t)epoch;
    if (w0 < nnz) {
      SlotSched cur = slot_load(head, tail, nxt, nneg, eps, w0 + lane, nnz);
      // the first schedule complete before the loop: otherwise the loop
      // head, reached with these loads in flight, would wait for every
      // operation still in flight on each iteration
      __builtin_amdgcn_s_waitcnt(0);
      for (uint64_t base = w0; base < nnz; base += stride) {
        const uint64_t e = base + lane;
        const SlotSched nxs =
            slot_load(head, tail, nxt, nneg, eps, e + stride, nnz);
        const bool valid = cur.h >= 0; // an empty slot of the COO has head -1
        const int32_t hv = valid ? cur.h : -1 - lane;
        const int32_t tv = valid ? cur.t : -1 - lane;
        const bool act = valid && !(cur.nx > (float)epoch);
        const float ep_neg = cur.ep * nsr_inv;
        uint32_t dr[8];
        neg_draws(es, e, dr);
        EdgeLoads<NC> L;
        sgd_gather_rows<NC>(st.rows, act, hv, tv, n, ep_neg, cur.nn, epoch, dr,
                            L);
        float gh[NC], gt[NC];
#pragma unroll
        for (int d = 0; d < NC; ++d)
          gh[d] = gt[d] = 0.0f;
        float nx = cur.nx, nn = cur.nn;
        if (act) {
          sgd_apply_rows<NC>(st.rows, hv, e, n, ep_neg, alpha, gamma, a, b,
                             es, rounding, L, gh, gt, c_pos, c_neg);
          advance_schedule(cur.ep, ep_neg, L.n_neg, nx, nn);
        }
        // the sampled slots' schedules back, the others' stores dropped
        const uint32_t so = act ? (uint32_t)lane * 4u : BUF_SKIP;
        __builtin_amdgcn_raw_buffer_store_b32(
            __float_as_uint(nx),
            __builtin_amdgcn_make_buffer_rsrc(nxt + base, 0, 256, 0x00020000),
            (int)so, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b32(
            __float_as_uint(nn),
            __builtin_amdgcn_make_buffer_rsrc(nneg + base, 0, 256, 0x00020000),
            (int)so, 0, 0);
        st.add(agg_of(hv, tv), hv, tv, gh, gt);
        cur = nxs;
      }
    }
    alpha = alpha0 * (1.0f - (float)epoch / (float)n_epochs);
    epoch_end(bar, gridDim.x, epoch, lag, peeked, seen);
  }
  flush_sgd_stats(stats, c_pos, c_neg);
}

template <int NC>
__global__ __launch_bounds__(1024) void k_sgd_stream(
    float *__restrict__ emb, const int32_t *__restrict__ head,
    const int32_t *__restrict__ tail, const uint32_t *__restrict__ nnz_p,
    uint32_t n, const float *__restrict__ eps, float *__restrict__ nxt,
    float *__restrict__ nneg, int n_epochs, float alpha0, float gamma, float a,
    float b, float nsr_inv, uint64_t seed, const float *__restrict__ rounding_p,
    unsigned *__restrict__ bar, unsigned long long *__restrict__ stats,
    int lag, FixFrame fx) {
  if (fx.rec) {
    const FixStore<NC> st = fix_store<NC>(fx.rec, *rounding_p);
    sgd_stream_body<NC>(st, head, tail, nnz_p, n, eps, nxt, nneg, n_epochs,
                        alpha0, gamma, a, b, nsr_inv, seed, rounding_p, bar,
                        stats, lag);
  } else {
    const FloatStore<NC> st{EmbRows<NC>{emb}, emb_rsrc<NC>(emb, n)};
    sgd_stream_body<NC>(st, head, tail, nnz_p, n, eps, nxt, nneg, n_epochs,
                        alpha0, gamma, a, b, nsr_inv, seed, rounding_p, bar,
                        stats, lag);
  }
}

// --------------------------------------------------- row-owner SGD
// For edge sets beyond the resident grid (the full sweeps).  Every COO slot
// (h, t) is listed twice, as an entry of row h (side 0) and of row t (side
// 1), the entries grouped by row, each entry with its own copy of the
// slot's schedule.  An epoch reads every row from the positions the
// previous epoch left (a Jacobi step: no row moves while an epoch reads),
// each entry computes its row's share of its slot -- on side 0 the head's
// positive step and negative samples (the reference's head update,
// embedding_kernel.cuh:509-589), on side 1 the tail's positive step, from
// the same positions the head side reads -- and each run of a row's
// entries within a wave adds its sum with one atomic.  The rows are read
// from one of two record buffers, the records (B0) and PullFrame::d1 (B1, a
// copy of the records' p0 with its own delta words), one load per row.
// Epoch t reads B(t mod 2) and
// adds into B(t + 1 mod 2) the run's sums of epochs t - 1 and t, so that
// B(t + 1 mod 2) = D(t + 1), the sum of all updates of epochs 0 .. t (both
// buffers start at zero).  The sums are integers in units of q, exact in
// any order: the result does not depend on the order of the entries or of
// the atomics.  A full grid barrier with agent-scope release and acquire
// separates the epochs.

// row-major counting of each slot under both its rows
__global__ void k_pull_count(const int32_t *__restrict__ head,
                             const int32_t *__restrict__ tail,
                             const uint32_t *__restrict__ nnz_p,
                             uint32_t *__restrict__ cnt) {
  const uint64_t nnz = *nnz_p;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const int32_t h = head[e];
    if (h < 0)
      continue; // an empty slot of the twin-paired COO
    atomicAdd(&cnt[h], 1u);
    atomicAdd(&cnt[tail[e]], 1u);
  }
}

// sum of cnt over each block of 1024 rows
__global__ __launch_bounds__(1024) void k_pull_bsum(
    const uint32_t *__restrict__ cnt, uint64_t n, uint32_t *__restrict__ bsum) {
  __shared__ uint32_t s_w[16];
  const uint64_t r = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
  uint32_t v = r < n ? cnt[r] : 0u;
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0)
    s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < 16; ++w)
      t += s_w[w];
    bsum[blockIdx.x] = t;
  }
}

// off[r] = bsum[block] (exclusive, scanned) + exclusive scan within the block
__global__ __launch_bounds__(1024) void k_pull_offsets(
    const uint32_t *__restrict__ cnt, uint64_t n,
    const uint32_t *__restrict__ bsum, uint32_t *__restrict__ off) {
  __shared__ uint32_t s[1024];
  const uint64_t r = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
  const uint32_t x = r < n ? cnt[r] : 0u;
  s[threadIdx.x] = x;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const uint32_t y = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += y;
    __syncthreads();
  }
  if (r < n)
    off[r] = bsum[blockIdx.x] + s[threadIdx.x] - x;
}

// the entries: slot e under its head (side 0) and its tail (side 1), placed
// by each row's cursor (cnt, zeroed again); the order within a row does not
// change the result
__global__ void k_pull_fill(const int32_t *__restrict__ head,
                            const int32_t *__restrict__ tail,
                            const uint32_t *__restrict__ nnz_p,
                            const float *__restrict__ eps,
                            const float *__restrict__ nxt,
                            const float *__restrict__ nneg, PullFrame pf) {
  const uint64_t nnz = *nnz_p;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const int32_t h = head[e];
    if (h < 0)
      continue;
    const int32_t t = tail[e];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int32_t row = side ? t : h;
      const uint32_t at = pf.off[row] + atomicAdd(&pf.cnt[row], 1u);
      pf.erc[at] = make_int2(row, side ? h : t);
      pf.esch[at] = make_float4(
          __uint_as_float((uint32_t)e | (side ? 0x80000000u : 0u)), eps[e],
          nxt[e], nneg[e]);
    }
  }
}

__global__ void k_fill_i32(int32_t *__restrict__ p, uint64_t n, int32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// d1 = the records' p0 with zero deltas (before the launch), or the
// records' deltas = d1's (after an odd number of epochs, when the final
// sums are in d1)
template <int NC>
__global__ void k_pull_copy(int32_t *__restrict__ d1,
                            int32_t *__restrict__ rec, uint64_t n, int after) {
  constexpr int F = fix_stride<NC>();
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int d = 0; d < F; ++d) {
      if (after)
        rec[r * (2 * F) + F + d] = d1[r * (2 * F) + F + d];
      else {
        d1[r * (2 * F) + d] = rec[r * (2 * F) + d];
        d1[r * (2 * F) + F + d] = 0;
      }
    }
}

// full grid barrier between epochs: every wave's row updates complete
// before the block arrives (each wave waits for its own memory operations;
// the arriving thread's release fence then covers the CU's caches), every
// row read after the block leaves (the leaving thread's acquire fence
// invalidates the CU's and the XCD's cached lines before any wave reads)
__device__ inline void pull_barrier(unsigned *bar, unsigned nblocks,
                                    int epoch, unsigned &seen) {
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_s_barrier();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = (unsigned)(epoch + 1) * nblocks;
    unsigned spins = 0;
    while (seen < target) {
      seen = bar_peek(bar);
      if (seen >= target)
        break;
      __builtin_amdgcn_s_sleep(1);
      // bounded: a timed-out wait anywhere in the launch ends the others
      if ((++spins & 1023u) == 0 &&
          (spins > (1u << 22) || bar_peek(bar + 2))) {
        __hip_atomic_store(bar + 2, 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __builtin_amdgcn_s_barrier();
}

template <int NC>
__global__ __launch_bounds__(1024) void k_sgd_pull(
    PullFrame pf, int32_t *__restrict__ rec, const uint32_t *__restrict__ nent_p,
    uint32_t n, int n_epochs, float alpha0, float gamma, float a, float b,
    float nsr_inv, uint64_t seed, const float *__restrict__ rounding_p,
    unsigned *__restrict__ bar, unsigned long long *__restrict__ stats) {
  constexpr int F = fix_stride<NC>();
  const uint64_t nent = *nent_p;
  const float rounding = *rounding_p;
  const float q = ldexpf(rounding, -24); // a power of two, as is 1 / q
  const float qinv = 1.0f / q;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  // whole waves iterate together (the runs are wave operations)
  const uint64_t w0 = (uint64_t)blockIdx.x * blockDim.x +
                      (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x - lane);
  unsigned long long c_pos = 0, c_neg = 0;
  unsigned seen = 0;
  float alpha = alpha0;
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const bool odd = epoch & 1;
    const FixRows<NC> rw{odd ? pf.d1 : rec, q};
    int32_t *const wb = (odd ? rec : pf.d1) + F; // delta words, stride 2F
    const uint64_t es = seed + (uint64_t)epoch;
    // the walk is pipelined: the next entry's words (and its run's sum of
    // the previous epoch) load while this entry works
    uint64_t base = w0;
    int2 rc_c = make_int2(0, 0), rc_n;
    float4 sc_c = make_float4(0.f, 0.f, 0.f, 0.f), sc_n;
    int32_t pc_c[NC], pc_n[NC];
    auto load_entry = [&](uint64_t e0, int2 &rc, float4 &sc,
                          int32_t (&pc)[NC]) {
      const uint64_t ei = e0 < nent ? e0 : nent - 1;
      rc = pf.erc[ei];
      sc = pf.esch[ei];
#pragma unroll
      for (int d = 0; d < NC; ++d)
        pc[d] = pf.epc[ei * NC + d];
    };
    if (base < nent)
      load_entry(base + lane, rc_c, sc_c, pc_c);
    for (; base < nent; base += stride) {
      load_entry(base + stride + lane, rc_n, sc_n, pc_n);
      const uint64_t e = base + lane;
      const bool valid = e < nent;
      const int32_t row = rc_c.x, oth = rc_c.y;
      const uint32_t sl = __float_as_uint(sc_c.x);
      const float ep = sc_c.y;
      float nx = sc_c.z, nn = sc_c.w;
      const bool tail_side = (sl >> 31) != 0;
      const uint64_t slot = sl & 0x7fffffffu;
      const bool act = valid && !(nx > (float)epoch);
      const bool negs = act && !tail_side;
      // the slot's head and tail, as the reference's edge reads them
      const int32_t hh = tail_side ? oth : row, tt = tail_side ? row : oth;
      const float ep_neg = ep * nsr_inv;
      float gh[NC], gt[NC];
#pragma unroll
      for (int d = 0; d < NC; ++d)
        gh[d] = gt[d] = 0.0f;
      if (act) {
        // only the sampled entries load: both rows, and on the head side
        // the draws and rows of the first 8 negative samples
        EdgeLoads<NC> L;
        L.n_neg = negs ? (int)(((float)epoch - nn) / ep_neg) : 0;
        rw.load((uint64_t)hh, L.cur);
        rw.load((uint64_t)tt, L.oth);
        if (L.n_neg > 0) {
          uint32_t dr[8];
          neg_draws(es, slot, dr);
          const int cnt = min(8, L.n_neg);
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            L.r[k] = __umulhi(dr[k], n);
            if (k < cnt)
              rw.load(L.r[k], L.neg[k]);
          }
        }
        unsigned long long cp = 0, cn = 0;
        sgd_apply_rows<NC>(rw, hh, slot, n, ep_neg, alpha, gamma, a, b, es,
                           rounding, L, gh, gt, cp, cn);
        if (!tail_side) {
          c_pos += cp;
          c_neg += cn;
        }
        nx += ep;
        if (L.n_neg > 0)
          nn += (float)L.n_neg * ep_neg;
        // the schedule back (slot and eps unchanged)
        reinterpret_cast<float2 *>(pf.esch + e)[1] = make_float2(nx, nn);
      }
      // this entry's share of its row, in units of q
      int32_t dx[NC];
#pragma unroll
      for (int d = 0; d < NC; ++d)
        dx[d] = act ? (int32_t)rintf((tail_side ? gt[d] : gh[d]) * qinv) : 0;
      // runs of equal rows over the wave (entries grouped by row): each
      // run's sum into its first lane
      const int32_t key = valid ? row : -1 - lane;
      const int32_t prev = __shfl_up(key, 1, 64);
      const bool start = lane == 0 || prev != key;
      const uint64_t sm = __ballot(start);
      const uint64_t after = lane == 63 ? 0ull : (sm >> (lane + 1));
      const int nstart = after ? lane + 1 + __builtin_ctzll(after) : 64;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
        for (int d = 0; d < NC; ++d) {
          const int32_t t = __shfl_down(dx[d], o, 64);
          if (lane + o < nstart)
            dx[d] += t;
        }
      }
      if (start && valid) {
        // the run's sums of the previous epoch and this one
        int32_t tot[NC];
#pragma unroll
        for (int d = 0; d < NC; ++d) {
          tot[d] = pc_c[d] + dx[d];
          pf.epc[e * NC + d] = dx[d];
        }
        int32_t *p = wb + (uint64_t)row * (2 * F);
#pragma unroll
        for (int d = 0; d + 1 < NC; d += 2) {
          const unsigned long long pk = fix_pack(tot[d], tot[d + 1]);
          if (pk != 0ull)
            __hip_atomic_fetch_add(reinterpret_cast<unsigned long long *>(p + d),
                                   pk, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if constexpr (NC & 1) {
          if (tot[NC - 1] != 0)
            __hip_atomic_fetch_add(reinterpret_cast<unsigned *>(p + NC - 1),
                                   (unsigned)tot[NC - 1], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      rc_c = rc_n;
      sc_c = sc_n;
#pragma unroll
      for (int d = 0; d < NC; ++d)
        pc_c[d] = pc_n[d];
    }
    alpha = alpha0 * (1.0f - (float)epoch / (float)n_epochs);
    pull_barrier(bar, gridDim.x, epoch, seen);
  }
  flush_sgd_stats(stats, c_pos, c_neg);
}

template <int NC>
static hipError_t launch_pull_nc(const PullFrame &pf, FixFrame fx,
                                 const int32_t *head, const int32_t *tail,
                                 const uint32_t *nnz_p, uint64_t cap,
                                 uint32_t n, const float *eps,
                                 const float *nxt, const float *nneg,
                                 int n_epochs, float alpha0, float gamma,
                                 float a, float b, float nsr_inv,
                                 uint64_t seed, const float *rounding_p,
                                 unsigned *bar, unsigned long long *stats,
                                 int num_cu, hipStream_t s) {
  static int per_cu = 0;
  if (per_cu <= 0)
    per_cu = blocks_per_cu(&k_sgd_pull<NC>, 1024);
  if (per_cu <= 0 || !fx.rec || cap == 0 || cap > pf.slots || n > pf.rows)
    return hipErrorInvalidConfiguration;
  const unsigned g = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((cap + 255) / 256, 16 * (uint64_t)num_cu));
  const unsigned nb = (unsigned)((n + 1023) / 1024);
  k_fill_i32<<<g, 256, 0, s>>>(reinterpret_cast<int32_t *>(pf.cnt), n + 1, 0);
  k_pull_count<<<g, 256, 0, s>>>(head, tail, nnz_p, pf.cnt);
  k_pull_bsum<<<nb, 1024, 0, s>>>(pf.cnt, n, pf.bsum);
  k_scan_small<<<1, 1024, 0, s>>>(pf.bsum, nb, pf.off + n);
  k_pull_offsets<<<nb, 1024, 0, s>>>(pf.cnt, n, pf.bsum, pf.off);
  k_fill_i32<<<g, 256, 0, s>>>(reinterpret_cast<int32_t *>(pf.cnt), n + 1, 0);
  k_pull_fill<<<g, 256, 0, s>>>(head, tail, nnz_p, eps, nxt, nneg, pf);
  k_fill_i32<<<g, 256, 0, s>>>(pf.epc, 2 * cap * NC, 0);
  k_pull_copy<NC><<<g, 256, 0, s>>>(pf.d1, fx.rec, n, 0);
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess)
    return le;
  // the co-resident grid, no more blocks than the most entries need
  const uint64_t blocks = std::min<uint64_t>((uint64_t)per_cu * num_cu,
                                             (2 * cap + 1023) / 1024);
  PullFrame pfa = pf;
  int32_t *reca = fx.rec;
  const uint32_t *nent_p = pf.off + n;
  stats = stats ? stats + 8 : nullptr; // the streaming kernels' slots
  void *args[] = {&pfa,  &reca,    &nent_p, &n,    &n_epochs,
                  &alpha0, &gamma, &a,      &b,    &nsr_inv,
                  &seed, &rounding_p, &bar, &stats};
  const hipError_t e = hipLaunchCooperativeKernel(
      reinterpret_cast<const void *>(&k_sgd_pull<NC>), dim3((unsigned)blocks),
      dim3(1024), args, 0, s);
  if (e != hipSuccess)
    return e;
  if (n_epochs & 1)
    k_pull_copy<NC><<<g, 256, 0, s>>>(pf.d1, fx.rec, n, 1);
  return hipGetLastError();
}

hipError_t launch_sgd_pull(int nc, const PullFrame &pf, FixFrame fx,
                           const int32_t *head, const int32_t *tail,
                           const uint32_t *nnz_p, uint64_t cap, uint32_t n,
                           const float *eps, const float *nxt,
                           const float *nneg, int n_epochs, float alpha0,
                           float gamma, float a, float b, float nsr_inv,
                           uint64_t seed, const float *rounding_p,
                           unsigned *bar, unsigned long long *stats,
                           int num_cu, hipStream_t s) {
#define GM_PULL(NC)                                                            \
  return launch_pull_nc<NC>(pf, fx, head, tail, nnz_p, cap, n, eps, nxt, nneg, \
                            n_epochs, alpha0, gamma, a, b, nsr_inv, seed,     \
                            rounding_p, bar, stats, num_cu, s)
  switch (nc) {
  case 1:
    GM_PULL(1);
  case 2:
    GM_PULL(2);
  case 3:
    GM_PULL(3);
  default:
    GM_PULL(4);
  }
#undef GM_PULL
}

// Host launchers of the cooperative SGD kernels, in the kernels' own
// translation unit: the register-resident kernel with 1 or 2 edges per
// thread when its co-resident grid holds every edge, else the streaming
// kernel on the co-resident grid.  Returns hipSuccess, or
// the launch error (the caller then launches one kernel per epoch).
// GPUMAP_AMD_SGD_PIPE=0 keeps the per-cycle SGD on k_sgd_waves (read at
// each launch)
static bool sgd_pipe_enabled() {
  const char *v = std::getenv("GPUMAP_AMD_SGD_PIPE");
  return !(v && v[0] == '0');
}

template <typename K> static int blocks_per_cu(K kernel, int threads) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void *>(kernel), threads, 0) !=
      hipSuccess)
    return 0;
  return per_cu;
}

// Launch of a per-cycle grid-wide SGD kernel (k_sgd_pipe, k_sgd_waves,
// k_sgd_waves_multi), whose blocks wait for one another.  Its callers keep
// the grid within the occupancy the device holds at once and hold the device
// exclusively while it runs (no other kernel of this process is in flight),
// so every block is resident without the cooperative launch's guarantee; and
// every wait in these kernels is bounded, reporting a timeout rather than
// hanging.  The plain launch leaves the device idle for less time around the
// kernel: 0.320 against 0.334 ms per k_sgd_pipe launch at 70,000 rows (HIP
// events around the SGD's launch sequence, MI355X).  GPUMAP_AMD_SGD_COOP=1
// restores the cooperative launch.
static hipError_t grid_launch(const void *fn, unsigned blocks,
                              unsigned threads, void **args, hipStream_t s) {
  const char *v = std::getenv("GPUMAP_AMD_SGD_COOP");
  if (v && v[0] == '1')
    return hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(threads), args,
                                      0, s);
  return hipLaunchKernel(fn, dim3(blocks), dim3(threads), args, 0, s);
}

template <int NC, int EPT>
static hipError_t try_fused(float *emb, const int32_t *head,
                            const int32_t *tail, const uint32_t *nnz_p,
                            uint32_t n, const float *eps, int n_epochs,
                            float alpha0, float gamma, float a, float b,
                            float nsr_inv, uint64_t seed,
                            const float *rounding_p, unsigned *bar,
                            unsigned long long *stats, int lag, uint64_t nnz,
                            int num_cu, hipStream_t s) {
  static int per_cu = 0; // cached once the occupancy query succeeds
  if (per_cu <= 0)
    per_cu = blocks_per_cu(&k_sgd_fused<NC, EPT>, 256);
  // 256-thread blocks: every block arrives on one counter per epoch, so
  // fewer, larger blocks keep that serialised atomic traffic small
  const uint64_t need = (nnz + 256ull * EPT - 1) / (256ull * EPT);
  if (need == 0 || need > (uint64_t)per_cu * (uint64_t)num_cu)
    return hipErrorInvalidConfiguration;
  void *args[] = {&emb,     &head,   &tail,  &nnz_p, &n,       &eps,
                  &n_epochs, &alpha0, &gamma, &a,     &b,       &nsr_inv,
                  &seed,    &rounding_p, &bar, &stats, &lag};
  return hipLaunchCooperativeKernel(
      reinterpret_cast<const void *>(&k_sgd_fused<NC, EPT>),
      dim3((unsigned)need), dim3(256), args, 0, s);
}

template <int NC>
static hipError_t try_waves(float *emb, const int32_t *head,
                            const int32_t *tail, const uint32_t *nnz_p,
                            uint32_t n, const float *eps, int n_epochs,
                            float alpha0, float gamma, float a, float b,
                            float nsr_inv, uint64_t seed,
                            const float *rounding_p, unsigned *bar,
                            unsigned long long *stats, int lag, uint64_t nnz,
                            int num_cu, hipStream_t s, FixFrame fx) {
  static int per_cu = 0;
  if (per_cu <= 0)
    per_cu = blocks_per_cu(&k_sgd_waves<NC>, 64);
  const uint64_t need = (nnz + 63) / 64;
  if (need == 0 || need > SGD_MAX_WAVES ||
      need > (uint64_t)per_cu * (uint64_t)num_cu)
    return hipErrorInvalidConfiguration;
  void *args[] = {&emb,     &head,   &tail,  &nnz_p, &n,       &eps,
                  &n_epochs, &alpha0, &gamma, &a,     &b,       &nsr_inv,
                  &seed,    &rounding_p, &bar, &stats, &lag,    &fx};
  return grid_launch(reinterpret_cast<const void *>(&k_sgd_waves<NC>),
                     (unsigned)need, 64, args, s);
}

template <int NC>
static hipError_t try_pipe(float *emb, const int32_t *head,
                           const int32_t *tail, const uint32_t *nnz_p,
                           uint32_t n, const float *eps, int n_epochs,
                           float alpha0, float gamma, float a, float b,
                           float nsr_inv, uint64_t seed,
                           const float *rounding_p, unsigned *bar,
                           unsigned long long *stats, uint64_t nnz,
                           int num_cu, hipStream_t s, FixFrame fx) {
  static int per_cu = 0;
  if (per_cu <= 0)
    per_cu = blocks_per_cu(&k_sgd_pipe<NC>, 256);
  const uint64_t need = (nnz + 63) / 64;
  if (need == 0 || need > SGD_MAX_WAVES ||
      need > (uint64_t)per_cu * (uint64_t)num_cu)
    return hipErrorInvalidConfiguration;
  // the row updates address the records through a buffer of at most 2^31
  // bytes (larger sessions take k_sgd_waves, whose atomics use 64-bit
  // addresses)
  if (fx.rec && !fix_buf_fits<NC>(n))
    return hipErrorInvalidConfiguration;
  int lag = PIPE_LAG;
  if (const char *v = std::getenv("GPUMAP_AMD_PIPE_LAG"))
    lag = std::max(PIPE_LAG_MIN, std::atoi(v));
  void *args[] = {&emb,     &head,   &tail,  &nnz_p, &n,       &eps,
                  &n_epochs, &alpha0, &gamma, &a,     &b,       &nsr_inv,
                  &seed,    &rounding_p, &bar, &stats, &lag,    &fx};
  return grid_launch(reinterpret_cast<const void *>(&k_sgd_pipe<NC>),
                     (unsigned)need, 256, args, s);
}

template <int NC>
static hipError_t try_stream(float *emb, const int32_t *head,
                             const int32_t *tail, const uint32_t *nnz_p,
                             uint32_t n, const float *eps, float *nxt,
                             float *nneg, int n_epochs, float alpha0,
                             float gamma, float a, float b, float nsr_inv,
                             uint64_t seed, const float *rounding_p,
                             unsigned *bar, unsigned long long *stats, int lag,
                             uint64_t nnz, int num_cu, hipStream_t s,
                             FixFrame fx) {
  static int per_cu = 0;
  if (per_cu <= 0)
    per_cu = blocks_per_cu(&k_sgd_stream<NC>, 1024);
  if (per_cu <= 0 || nnz == 0)
    return hipErrorInvalidConfiguration;
  // 1024-thread blocks: one arrival per block and epoch on the counter
  const uint64_t blocks = std::min<uint64_t>((uint64_t)per_cu * num_cu,
                                             (nnz + 1023) / 1024);
  // its work counters go to their own slots (stats + 8)
  stats = stats ? stats + 8 : nullptr;
  void *args[] = {&emb,   &head,    &tail,    &nnz_p, &n,  &eps,
                  &nxt,   &nneg,    &n_epochs, &alpha0, &gamma, &a,
                  &b,     &nsr_inv, &seed,    &rounding_p, &bar, &stats,
                  &lag,   &fx};
  return hipLaunchCooperativeKernel(
      reinterpret_cast<const void *>(&k_sgd_stream<NC>),
      dim3((unsigned)blocks), dim3(1024), args, 0, s);
}

template <int NC>
static hipError_t try_fused_any(float *emb, const int32_t *head,
                                const int32_t *tail, const uint32_t *nnz_p,
                                uint32_t n, const float *eps, float *nxt,
                                float *nneg, int n_epochs, float alpha0,
                                float gamma, float a, float b, float nsr_inv,
                                uint64_t seed, const float *rounding_p,
                                unsigned *bar, unsigned long long *stats,
                                int lag, uint64_t nnz, int num_cu,
                                hipStream_t s, int *variant,
                                int stream_mode, FixFrame fx) {
  *variant = 1;
  if (stream_mode == SGD_STREAM_FORCE)
    return try_stream<NC>(emb, head, tail, nnz_p, n, eps, nxt, nneg, n_epochs,
                          alpha0, gamma, a, b, nsr_inv, seed, rounding_p, bar,
                          stats, lag, nnz, num_cu, s, fx);
  if (sgd_pipe_enabled()) {
    // small edge sets: chain, row and draw waves (variant 4)
    *variant = 4;
    const hipError_t e =
        try_pipe<NC>(emb, head, tail, nnz_p, n, eps, n_epochs, alpha0, gamma,
                     a, b, nsr_inv, seed, rounding_p, bar, stats, nnz, num_cu,
                     s, fx);
    if (e != hipErrorInvalidConfiguration)
      return e;
  }
  *variant = 2;
  {
    // small edge sets: per-wave arrival slots, twice the counter kernels' lag
    const hipError_t e =
        try_waves<NC>(emb, head, tail, nnz_p, n, eps, n_epochs, alpha0, gamma,
                      a, b, nsr_inv, seed, rounding_p, bar, stats, 2 * lag,
                      nnz, num_cu, s, fx);
    if (e != hipErrorInvalidConfiguration)
      return e;
  }
  *variant = 0;
#define GM_TRY(EPT)                                                            \
  {                                                                            \
    const hipError_t e = try_fused<NC, EPT>(                                   \
        emb, head, tail, nnz_p, n, eps, n_epochs, alpha0, gamma, a, b,         \
        nsr_inv, seed, rounding_p, bar, stats, lag, nnz, num_cu, s);           \
    if (e != hipErrorInvalidConfiguration)                                     \
      return e;                                                                \
  }
  GM_TRY(1)
  GM_TRY(2)
#undef GM_TRY
  if (stream_mode == SGD_STREAM_NONE)
    return hipErrorInvalidConfiguration;
  *variant = 1;
  return try_stream<NC>(emb, head, tail, nnz_p, n, eps, nxt, nneg, n_epochs,
                        alpha0, gamma, a, b, nsr_inv, seed, rounding_p, bar,
                        stats, lag, nnz, num_cu, s, fx);
}

template <int NC>
static hipError_t try_batch(const SgdBatch &batch, unsigned waves,
                            unsigned *bar, int lag, int num_cu,
                            hipStream_t s) {
  static int per_cu = 0;
  if (per_cu <= 0)
    per_cu = blocks_per_cu(&k_sgd_waves_multi<NC>, 64);
  if (waves == 0 || waves > (unsigned)per_cu * (unsigned)num_cu)
    return hipErrorInvalidConfiguration;
  SgdBatch bb = batch;
  void *args[] = {&bb, &bar, &lag};
  return grid_launch(reinterpret_cast<const void *>(&k_sgd_waves_multi<NC>),
                     waves, 64, args, s);
}

hipError_t launch_sgd_batch(int nc, const SgdBatch &batch, unsigned waves,
                            unsigned *bar, int lag, int num_cu,
                            hipStream_t s) {
  switch (nc) {
  case 1:
    return try_batch<1>(batch, waves, bar, lag, num_cu, s);
  case 2:
    return try_batch<2>(batch, waves, bar, lag, num_cu, s);
  case 3:
    return try_batch<3>(batch, waves, bar, lag, num_cu, s);
  default:
    return try_batch<4>(batch, waves, bar, lag, num_cu, s);
  }
}

hipError_t launch_sgd_fused(int nc, float *emb, const int32_t *head,
                            const int32_t *tail, const uint32_t *nnz_p,
                            uint32_t n, const float *eps, float *nxt,
                            float *nneg, int n_epochs, float alpha0,
                            float gamma, float a, float b, float nsr_inv,
                            uint64_t seed, const float *rounding_p,
                            unsigned *bar, unsigned long long *stats, int lag,
                            uint64_t nnz, int num_cu, hipStream_t s,
                            int *variant, int stream_mode, FixFrame fx) {
  switch (nc) {
  case 1:
    return try_fused_any<1>(emb, head, tail, nnz_p, n, eps, nxt, nneg,
                            n_epochs, alpha0, gamma, a, b, nsr_inv, seed,
                            rounding_p, bar, stats, lag, nnz, num_cu, s, variant,
                            stream_mode, fx);
  case 2:
    return try_fused_any<2>(emb, head, tail, nnz_p, n, eps, nxt, nneg,
                            n_epochs, alpha0, gamma, a, b, nsr_inv, seed,
                            rounding_p, bar, stats, lag, nnz, num_cu, s, variant,
                            stream_mode, fx);
  case 3:
    return try_fused_any<3>(emb, head, tail, nnz_p, n, eps, nxt, nneg,
                            n_epochs, alpha0, gamma, a, b, nsr_inv, seed,
                            rounding_p, bar, stats, lag, nnz, num_cu, s, variant,
                            stream_mode, fx);
  default:
    return try_fused_any<4>(emb, head, tail, nnz_p, n, eps, nxt, nneg,
                            n_epochs, alpha0, gamma, a, b, nsr_inv, seed,
                            rounding_p, bar, stats, lag, nnz, num_cu, s, variant,
                            stream_mode, fx);
  }
}

// The records around a cooperative SGD launch: phase 0 (before) makes
// them, phase 1 (after a kernel that used them) adds the updates back into
// the float embedding.
template <int NC>
static void fix_phase(int phase, float *emb, uint64_t n, FixFrame fx,
                      const float *rounding_p, int num_cu, hipStream_t s) {
  const unsigned grid = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((n + 255) / 256, 8 * (uint64_t)num_cu));
  if (phase == 0)
    k_fix_in<NC><<<grid, 256, 0, s>>>(emb, n, fx);
  else
    k_fix_out<NC><<<grid, 256, 0, s>>>(emb, n, fx, rounding_p);
}

hipError_t launch_fix(int nc, int phase, float *emb, uint64_t n, FixFrame fx,
                      const float *rounding_p, int num_cu, hipStream_t s) {
  if (!fx.rec || n == 0)
    return hipSuccess;
  switch (nc) {
  case 1:
    fix_phase<1>(phase, emb, n, fx, rounding_p, num_cu, s);
    break;
  case 2:
    fix_phase<2>(phase, emb, n, fx, rounding_p, num_cu, s);
    break;
  case 3:
    fix_phase<3>(phase, emb, n, fx, rounding_p, num_cu, s);
    break;
  default:
    fix_phase<4>(phase, emb, n, fx, rounding_p, num_cu, s);
  }
  return hipGetLastError();
}

#define GM_SGD_FUSED(NC, EPT)                                                  \
  template __global__ void k_sgd_fused<NC, EPT>(                               \
      float *, const int32_t *, const int32_t *, const uint32_t *, uint32_t,   \
      const float *, int, float, float, float, float, float, uint64_t,         \
      const float *, unsigned *, unsigned long long *, int);
GM_SGD_FUSED(1, 1)
GM_SGD_FUSED(1, 2)
GM_SGD_FUSED(2, 1)
GM_SGD_FUSED(2, 2)
GM_SGD_FUSED(3, 1)
GM_SGD_FUSED(3, 2)
GM_SGD_FUSED(4, 1)
GM_SGD_FUSED(4, 2)

template __global__ void k_sgd_epoch<2>(float *, const int32_t *,
                                        const int32_t *, const uint32_t *,
                                        uint32_t, const float *, float *,
                                        float *, float, int, float, float,
                                        float, float, uint64_t, const float *,
                                        unsigned long long *);
template __global__ void k_sgd_epoch<3>(float *, const int32_t *,
                                        const int32_t *, const uint32_t *,
                                        uint32_t, const float *, float *,
                                        float *, float, int, float, float,
                                        float, float, uint64_t, const float *,
                                        unsigned long long *);
template __global__ void k_sgd_epoch<1>(float *, const int32_t *,
                                        const int32_t *, const uint32_t *,
                                        uint32_t, const float *, float *,
                                        float *, float, int, float, float,
                                        float, float, uint64_t, const float *,
                                        unsigned long long *);
template __global__ void k_sgd_epoch<4>(float *, const int32_t *,
                                        const int32_t *, const uint32_t *,
                                        uint32_t, const float *, float *,
                                        float *, float, int, float, float,
                                        float, float, uint64_t, const float *,
                                        unsigned long long *);

// ----------------------------------------------------------- reposition
// Unselected rows move to the membership-weighted mean of their embed-time
// neighbours, read from a snapshot (reference heuristic_neighbor_avg_kernel
// :811-894), any n_components <= 4.  A thread reads its k - 1 neighbour
// entries four at a time, each group's entries and snapshot rows as
// independent loads (entries past k - 1 read entry 0 and are skipped), and
// sums them in neighbour order.
__global__ void k_reposition(const uint32_t *__restrict__ pidx,
                             const float *__restrict__ pdist, NRows nr,
                             int k, int euclid, const double *__restrict__ sigma,
                             const double *__restrict__ rho,
                             const double *__restrict__ mean_p,
                             uint8_t *__restrict__ in_batch,
                             const float *__restrict__ snap,
                             float *__restrict__ emb, int c) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n = nr.get();
  if (i >= n)
    return;
  if (in_batch[i]) { // a batch row: keeps its place; clear its mark
    in_batch[i] = 0;
    return;
  }
  const double sig_i = sigma_of(sigma, i, *mean_p);
  const double rho_i = rho[i];
  double wsum = 0.0, acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int q0 = 0; q0 < k - 1; q0 += 4) {
    uint32_t js[4];
    float vs[4], sv[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint64_t at = i * k + (q0 + t < k - 1 ? q0 + t : 0);
      js[t] = pidx[at];
      vs[t] = pdist[at];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint64_t jr = js[t] < n ? js[t] : 0;
#pragma unroll
      for (int dd = 0; dd < 4; ++dd)
        sv[t][dd] = dd < c ? snap[jr * c + dd] : 0.0f;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint32_t j = js[t];
      const float v = vs[t];
      const double d = !valid_d((double)v) ? (double)v
                       : euclid ? (double)sqrt_rn(fmaxf(v, 0.0f))
                                : (double)v;
      if (q0 + t < k - 1 && j != i && j < n && valid_d(d)) {
        const double wq = membership(d, rho_i, sig_i);
        wsum += wq;
#pragma unroll
        for (int dd = 0; dd < 4; ++dd)
          if (dd < c)
            acc[dd] += wq * sv[t][dd];
      }
    }
  }
  if (wsum > 0.0) {
#pragma unroll
    for (int dd = 0; dd < 4; ++dd)
      if (dd < c)
        emb[i * c + dd] = (float)(acc[dd] / wsum);
  }
}

// dst[0, n c) = src[0, n c): the embedding's snapshot for k_reposition
__global__ void k_copy_rows(const float *__restrict__ src,
                            float *__restrict__ dst, NRows nr, int c) {
  const uint64_t total = nr.get() * (uint64_t)c;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ctr += nnz * n_epochs: the edge-epochs an SGD launch scans, counted on the
// device so that measuring needs no read-back of the edge count
__global__ void k_add_edge_epochs(const uint32_t *__restrict__ nnz_p,
                                  int n_epochs,
                                  unsigned long long *__restrict__ ctr) {
  if (threadIdx.x == 0)
    atomicAdd(ctr, (unsigned long long)*nnz_p * (unsigned long long)n_epochs);
}

} // namespace gm
