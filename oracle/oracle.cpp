// CPU oracle of the progressive UMAP hot path -- TEST INFRASTRUCTURE ONLY.
// Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
// never by the product.  Declarations and the list of what pins each function
// are in oracle.h.
//
// Method restated (reference files, path:line):
//   engine / queues / batch selection   cpp/src/impl.cuh:146-500, queue.cuh
//   insertion                            cpp/src/kernels/insertion_kernel.cuh:926-1095
//   graph update (NN-Descent join)       cpp/src/kernels/update_kernel.cuh:377-687,913-1054
//   fuzzy simplicial set                 cpp/src/kernels/embedding_kernel.cuh:135-320,649-716
//   SGD                                  cpp/src/kernels/embedding_kernel.cuh:44-75,325-589,726-798
//   noise / reposition                   cpp/src/kernels/embedding_kernel.cuh:811-894,953-1007
//
// Deviations from the reference, all deliberate (the reference's behaviour at
// each cited line is a defect of that code, not part of the method):
//  D1 Forward insertion takes effect.  The reference sorts its -1 "invalid
//     candidate" sentinels to the front and then stops at the first one
//     (insertion_kernel.cuh:407-444,462-464), so for k<=16 new rows keep their
//     random FLT_MAX neighbours.  Here valid candidates are inserted.
//  D2 Random initial neighbours get their real distance (reference: FLT_MAX,
//     insertion_kernel.cuh:240-243), so a later offer of the same id is a
//     true duplicate rather than a rejected improvement (update_kernel.cuh:341).
//  D3 One 32-wide pool per row (reference: 32-wide GPU pool + k-wide host list,
//     buffers.cuh:17-22, update_kernel.cuh:809-848) whose entries carry the
//     new/old flag; sampling takes every new and every old entry of the pool.
//     Classic NN-Descent never joins old x old, so no Bloom filter is needed
//     (reference: buffers.cuh:35-137, update_kernel.cuh:787-797).
//  D4 Reverse lists keep the 32 nearest reverse neighbours (ties by id)
//     instead of whichever 32 won an atomicAdd race (update_kernel.cuh:122-146).
//  D5 No self-loops in the fuzzy union.  The reference maps the self column to
//     (i,i) and then finds p_ii=1 in the transpose scan, giving every batch row a
//     w=1 self edge (embedding_kernel.cuh:229-247,286-305).
//  D6 Empty kNN slots stay sentinel (FLT_MAX) through the embed-time sqrt
//     (reference: sqrt(FLT_MAX) becomes a "valid" 1.8e19, :627-629).
//  D7 rho/sigma bisection and membership weights in FP64.
//  D8 SGD is general in n_components (reference hard-codes 2, :532-533,570).
//  D9 Update convergence threshold max(1, 1e-4 n) rows: the reference's
//     (size_t)(1e-4 n) is 0 below n = 10,000 and `size < 0` never holds
//     (impl.cuh:261-263,320), so small sessions would never complete.
//  D10 A negative sample's index is the high word of r * n for a 32-bit
//     Philox draw r (reference: curand() % n, embedding_kernel.cuh:567): the
//     same near-uniform choice without an integer division on the device.
//     The reference's cuRAND streams are not reproducible here anyway.
//
// knn_mode 1 (exact insertion, the MI355X path's default): each inserted row's
// pool is the exact top-32 (canonical distance, ties by id) over all rows
// inserted so far, and every earlier row is offered every new row; entries
// enter as old, so the NN-Descent update has nothing to join and only drains
// its queue.  By induction every pool is the exact 32-NN set of the rows
// inserted so far.
// Everything else (constants, seeds, queue semantics, completion rules,
// trim/eps/rounding/alpha schedule, noise, reposition) follows the reference.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <numeric>
#include <queue>
#include <random>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int W = 32;            // pool width (reference MAX_NEIGHBORS)
constexpr int C = 256;           // candidate cap (reference MAX_CANDIDATES)
constexpr int WINDOW_N_BATCHES = 3;
constexpr uint64_t EMPTY = ~0ull;
constexpr uint64_t NEW_BIT = 0x80000000ull;
constexpr uint32_t ID_MASK = 0x7fffffffu;
constexpr uint32_t TAG_INIT_NBR = 1, TAG_EMB_INIT = 2, TAG_NOISE = 3,
                   TAG_NEG = 4;

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): multipliers and Weyl
// constants as published there.
void philox(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = key[0], k1 = key[1];
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
  int pos = 4;
  Rng(uint64_t seed, uint32_t a, uint32_t b, uint32_t tag) {
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
    ctr[0] = a;
    ctr[1] = b;
    ctr[2] = 0;
    ctr[3] = tag;
  }
  uint32_t next() {
    if (pos == 4) {
      philox(ctr, key, buf);
      ctr[2]++;
      pos = 0;
    }
    return buf[pos++];
  }
  // Box-Muller standard normal.
  float normal() {
    const float u1 = ((float)next() + 1.0f) * 2.3283064365386963e-10f;
    const float u2 = (float)next() * 2.3283064365386963e-10f;
    return sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
  }
};

// -------------------------------------------------------- distance / keys
// Canonical squared distance: sequential FMA over features in index order.
inline float canon_dist(const float *a, const float *b, int d) {
  float acc = 0.0f;
  for (int i = 0; i < d; ++i) {
    const float t = a[i] - b[i];
    acc = std::fmaf(t, t, acc);
  }
  return acc;
}

inline uint64_t make_key(float d, uint32_t id, bool is_new) {
  uint32_t db;
  std::memcpy(&db, &d, 4);
  return ((uint64_t)db << 32) | (is_new ? NEW_BIT : 0ull) | (uint64_t)id;
}
inline uint32_t key_id(uint64_t k) { return (uint32_t)k & ID_MASK; }
inline float key_dist(uint64_t k) {
  const uint32_t db = (uint32_t)(k >> 32);
  float d;
  std::memcpy(&d, &db, 4);
  return d;
}

// Bounded top-W insert of one key; the same contract as the device CAS
// insert: reject a duplicate id or a key not below the current maximum,
// otherwise replace the maximum (lowest slot index among equal maxima).
// Returns 0 rejected, 1 inserted at rank >= k, 2 inserted at rank < k.
inline int offer(uint64_t *row, int k, uint64_t key) {
  const uint32_t id = key_id(key);
  int m = 0;
  uint64_t mx = row[0];
  for (int j = 0; j < W; ++j) {
    const uint64_t v = row[j];
    if (v != EMPTY && key_id(v) == id)
      return 0;
    if (v > mx) {
      mx = v;
      m = j;
    }
  }
  if (key >= mx)
    return 0;
  int rank = 0;
  for (int j = 0; j < W; ++j)
    rank += (row[j] < key);
  row[m] = key;
  return rank < k ? 2 : 1;
}

inline int sorted_row(const uint64_t *row, uint64_t *out) {
  int cnt = 0;
  for (int j = 0; j < W; ++j)
    if (row[j] != EMPTY)
      out[cnt++] = row[j];
  std::sort(out, out + cnt);
  return cnt;
}

// --------------------------------------------------------------- queues
struct DedupQueue { // reference UpdateQueue (queue.cuh:46-80)
  std::queue<uint32_t> q;
  std::unordered_set<uint32_t> in;
  size_t size() const { return q.size(); }
  void push(uint32_t v) {
    if (in.insert(v).second)
      q.push(v);
  }
  std::vector<uint32_t> pop_n(size_t n) {
    std::vector<uint32_t> out;
    n = std::min(n, q.size());
    for (size_t i = 0; i < n; ++i) {
      out.push_back(q.front());
      in.erase(q.front());
      q.pop();
    }
    return out;
  }
};
struct Fifo { // reference Queue (queue.cuh:10-40)
  std::deque<uint32_t> q;
  size_t size() const { return q.size(); }
  std::vector<uint32_t> pop_n(size_t n) {
    n = std::min(n, q.size());
    std::vector<uint32_t> out(q.begin(), q.begin() + n);
    q.erase(q.begin(), q.begin() + n);
    return out;
  }
};

// ----------------------------------------------------------- fuzzy set
// smooth_knn_dist for one row of the embed-time kNN (self at column 0),
// FP64 (reference embedding_kernel.cuh:135-207 restated in double).
void smooth_row(const double *d, int k, double mean_dist, double *sigma,
                double *rho) {
  const double target = std::log2((double)(k - 1));
  int start_nz = -1;
  double sum = 0.0;
  for (int i = 0; i < k; ++i) {
    double di = d[i];
    if (!std::isfinite(di) || di < 0.0 || di >= FLT_MAX / 2.0)
      di = 0.0;
    sum += di;
    if (di > 0.0 && start_nz == -1)
      start_nz = i;
  }
  if (start_nz == -1) {
    *rho = 0.0;
    *sigma = 1e-3 * mean_dist;
    return;
  }
  const double row_mean = sum / (double)k;
  const double r = d[start_nz];
  double lo = 0.0, hi = INFINITY, mid = 1.0;
  for (int it = 0; it < 64; ++it) {
    double psum = 0.0;
    for (int j = 2; j < k; ++j) {
      double dj = d[j];
      if (!std::isfinite(dj) || dj < 0.0 || dj >= FLT_MAX / 2.0)
        dj = 0.0;
      const double x = dj - r;
      psum += (x > 0.0) ? std::exp(-x / mid) : 1.0;
    }
    if (std::fabs(psum - target) < 1e-5)
      break;
    if (psum > target) {
      hi = mid;
      mid = (lo + hi) * 0.5;
    } else {
      lo = mid;
      mid = std::isinf(hi) ? mid * 2.0 : (lo + hi) * 0.5;
    }
  }
  if (!std::isfinite(mid) || mid <= 0.0)
    mid = 1.0;
  const double min_scale = 1e-3 * (r > 0.0 ? row_mean : mean_dist);
  *rho = r;
  *sigma = mid < min_scale ? min_scale : mid;
}

inline bool valid_dist(double d) {
  return std::isfinite(d) && d >= 0.0 && d < FLT_MAX / 2.0;
}
inline double membership(double d, double rho, double sigma) {
  if (d - rho <= 0.0 || sigma == 0.0)
    return 1.0;
  const double v = std::exp(-((d - rho) / sigma));
  return v < FLT_MIN ? (double)FLT_MIN : v;
}

// Embed-time kNN row: column 0 = self (d=0), columns 1..k-1 = the first k-1
// stored neighbours; Euclidean squared distances are square-rooted
// (reference prepend_self_knn_kernel, embedding_kernel.cuh:599-631).
inline void embed_row(const uint32_t *idx, const float *dist, uint64_t i,
                      int k, bool euclid, uint32_t *ri, double *rd) {
  ri[0] = (uint32_t)i;
  rd[0] = 0.0;
  for (int c = 1; c < k; ++c) {
    ri[c] = idx[i * k + c - 1];
    const float d = dist[i * k + c - 1];
    if (!valid_dist(d))
      rd[c] = d;
    else
      rd[c] = euclid ? (double)sqrtf(fmaxf(d, 0.0f)) : (double)d;
  }
}

struct Fuzzy {
  std::vector<double> sigma, rho;
  std::vector<int32_t> head, tail;
  std::vector<float> w;
};

// Fuzzy simplicial set for the batch rows with the union restricted to them
// (reference compute_fuzzy_graph, embedding_kernel.cuh:649-716) followed by
// the trim of run_optimize_layout (:749-765).  Deterministic order: directed
// entries in (batch, slot) order, then the reverse entries in the same order.
void fuzzy_set(const uint32_t *idx, const float *dist, uint64_t n, int k,
               bool euclid, const uint32_t *batch, uint64_t nb, int n_epochs,
               Fuzzy &out) {
  std::vector<uint32_t> ei((size_t)n * k);
  std::vector<double> ed((size_t)n * k);
  for (uint64_t i = 0; i < n; ++i)
    embed_row(idx, dist, i, k, euclid, &ei[i * k], &ed[i * k]);
  double s = 0.0;
  uint64_t cnt = 0;
  for (size_t t = 0; t < ed.size(); ++t)
    if (valid_dist(ed[t])) {
      s += ed[t];
      ++cnt;
    }
  const double mean = (cnt > 0 && s > 0.0) ? s / (double)cnt : 1.0;
  out.sigma.assign(n, 0.0);
  out.rho.assign(n, 0.0);
  for (uint64_t i = 0; i < n; ++i)
    smooth_row(&ed[i * k], k, mean, &out.sigma[i], &out.rho[i]);

  const size_t nd = (size_t)nb * k;
  std::vector<double> wv(nd, 0.0);
  std::vector<uint8_t> rev(nd, 0), live(nd, 0);
  std::vector<int32_t> hs(nd), ts(nd);
  for (uint64_t bi = 0; bi < nb; ++bi) {
    const uint64_t i = batch[bi];
    for (int c = 0; c < k; ++c) {
      const size_t t = bi * k + c;
      const uint32_t j = ei[i * k + c];
      const double d = ed[i * k + c];
      hs[t] = (int32_t)i;
      ts[t] = (int32_t)j;
      if (j >= n || j == i || !valid_dist(d))
        continue;
      const double p = membership(d, out.rho[i], out.sigma[i]);
      double q = 0.0;
      bool found = false;
      for (int cj = 0; cj < k; ++cj)
        if (ei[(size_t)j * k + cj] == i) {
          const double dj = ed[(size_t)j * k + cj];
          if (valid_dist(dj)) {
            q = membership(dj, out.rho[j], out.sigma[j]);
            found = true;
          }
          break;
        }
      wv[t] = p + q - p * q;
      live[t] = 1;
      rev[t] = !found;
    }
  }
  float max_w = 0.0f;
  for (size_t t = 0; t < nd; ++t)
    if (live[t])
      max_w = std::max(max_w, (float)wv[t]);
  out.head.clear();
  out.tail.clear();
  out.w.clear();
  if (!(max_w > 0.0f))
    return;
  const float thr = max_w / (float)n_epochs;
  for (int pass = 0; pass < 2; ++pass)
    for (size_t t = 0; t < nd; ++t) {
      if (!live[t] || (pass == 1 && !rev[t]))
        continue;
      const float wf = (float)wv[t];
      if (!(wf >= thr))
        continue;
      out.head.push_back(pass == 0 ? hs[t] : ts[t]);
      out.tail.push_back(pass == 0 ? ts[t] : hs[t]);
      out.w.push_back(wf);
    }
}

// ----------------------------------------------------------------- SGD
inline float fclip(float v) {
  if (!std::isfinite(v))
    return 0.0f;
  return std::fmax(std::fmin(v, 4.0f), -4.0f);
}

float rounding_factor(const std::vector<int32_t> &head, uint64_t n,
                      float alpha) {
  if (head.empty())
    return 1.0f;
  std::vector<int> deg(n, 0);
  for (int32_t h : head)
    deg[h]++;
  int max_deg = 0;
  for (int v : deg)
    max_deg = std::max(max_deg, v);
  if (max_deg == 0)
    return 1.0f;
  const float max_abs = (float)max_deg * 4.0f * std::fabs(alpha);
  const float delta = max_abs / (1.0f - 2.0f * (float)max_deg * FLT_EPSILON);
  int e;
  std::frexp(delta, &e);
  return std::ldexp(1.0f, e);
}

// Epoch loop of run_optimize_layout / optimize_batch_kernel
// (embedding_kernel.cuh:509-589,766-793) for any n_components, on the
// already-trimmed COO with weights w.  n_threads > 1 runs the edges
// Hogwild-style with relaxed atomic adds like the device kernel.  With
// `trace` (serial runs only) every gradient step is recorded as it is
// applied, for the tests that check each one.
void sgd_epochs(float *emb, uint64_t n, int c, const int32_t *head,
                const int32_t *tail, const float *w, int64_t nnz,
                int n_epochs, int nsr, float a, float b, float alpha0,
                float gamma, uint64_t seed, int n_threads,
                std::vector<or_sgd_event> *trace = nullptr,
                float *rounding_out = nullptr) {
  if (nnz <= 0 || n == 0 || n_epochs <= 0)
    return;
  float max_w = 0.0f;
  for (int64_t e = 0; e < nnz; ++e)
    max_w = std::max(max_w, w[e]);
  if (!(max_w > 0.0f))
    return;
  std::vector<float> eps(nnz), nxt(nnz), nneg(nnz);
  const float nsr_inv = 1.0f / (float)nsr;
  for (int64_t e = 0; e < nnz; ++e) {
    const float ns = (float)n_epochs * (w[e] / max_w);
    eps[e] = ns > 0.0f ? (float)n_epochs / ns : -1.0f;
    nxt[e] = eps[e];
    nneg[e] = eps[e] * nsr_inv;
  }
  std::vector<int32_t> hv(head, head + nnz);
  const float rounding = rounding_factor(hv, n, alpha0);
  if (rounding_out)
    *rounding_out = rounding;
  if (trace)
    n_threads = 1;
  auto add = [&](float *p, float v) {
    std::atomic_ref<float>(*p).fetch_add(v, std::memory_order_relaxed);
  };
  auto record = [&](int epoch, int64_t e, int kind, int32_t other, float alpha,
                    float d2, float coef, const float *y, const float *o,
                    const float *g) {
    or_sgd_event ev{};
    ev.epoch = epoch;
    ev.edge = (int32_t)e;
    ev.kind = kind;
    ev.other = other;
    ev.alpha = alpha;
    ev.d2 = d2;
    ev.coef = coef;
    for (int d = 0; d < c; ++d) {
      ev.y[d] = y[d];
      ev.o[d] = o ? o[d] : 0.0f;
      ev.g[d] = g ? g[d] : 0.0f;
    }
    trace->push_back(ev);
  };
  float alpha = alpha0;
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const uint64_t es = seed + (uint64_t)epoch;
#pragma omp parallel for schedule(static) num_threads(n_threads) if (n_threads > 1)
    for (int64_t e = 0; e < nnz; ++e) {
      const float ns = nxt[e];
      if (ns > (float)epoch)
        continue;
      const float ep = eps[e];
      const float ep_neg = ep * nsr_inv;
      const int32_t j = head[e], kk = tail[e];
      float *cur_p = emb + (size_t)j * c;
      float *oth_p = emb + (size_t)kk * c;
      float cur[4], oth[4], grads[4], step[4];
      for (int d = 0; d < c; ++d) {
        cur[d] = std::atomic_ref<float>(cur_p[d]).load(std::memory_order_relaxed);
        oth[d] = std::atomic_ref<float>(oth_p[d]).load(std::memory_order_relaxed);
      }
      float d2 = 0.0f;
      for (int d = 0; d < c; ++d) {
        const float t = cur[d] - oth[d];
        d2 += t * t;
      }
      float attr = 0.0f;
      if (d2 > 0.0f)
        attr = (-2.0f * a * b * powf(d2, b - 1.0f)) / (a * powf(d2, b) + 1.0f);
      float y0[4];
      for (int d = 0; d < c; ++d) {
        y0[d] = cur[d];
        const float g = fclip(attr * (cur[d] - oth[d])) * alpha;
        cur[d] += g;
        grads[d] = g;
      }
      if (trace)
        record(epoch, e, 0, kk, alpha, d2, attr, y0, oth, grads);
      for (int d = 0; d < c; ++d)
        add(&oth_p[d], (rounding + -grads[d]) - rounding);
      nxt[e] = ns + ep;
      const float nn = nneg[e];
      const int n_neg = (int)(((float)epoch - nn) / ep_neg);
      if (n_neg > 0) {
        Rng rng(es, (uint32_t)e, (uint32_t)((uint64_t)e >> 32), TAG_NEG);
        for (int p = 0; p < n_neg; ++p) {
          // index in [0, n): high word of the 32 x 32-bit product
          const uint32_t r =
              (uint32_t)(((uint64_t)rng.next() * (uint64_t)n) >> 32);
          if ((int32_t)r == j) {
            if (trace)
              record(epoch, e, 2, (int32_t)r, alpha, 0.0f, 0.0f, cur, nullptr,
                     nullptr);
            continue;
          }
          const float *neg_p = emb + (size_t)r * c;
          float neg[4];
          for (int d = 0; d < c; ++d)
            neg[d] = std::atomic_ref<float>(const_cast<float &>(neg_p[d]))
                         .load(std::memory_order_relaxed);
          float d2n = 0.0f;
          for (int d = 0; d < c; ++d) {
            const float t = cur[d] - neg[d];
            d2n += t * t;
          }
          float rep = 0.0f;
          if (d2n > 0.0f)
            rep = (2.0f * gamma * b) / ((0.001f + d2n) * (a * powf(d2n, b) + 1.0f));
          for (int d = 0; d < c; ++d) {
            y0[d] = cur[d];
            const float g =
                rep > 0.0f ? fclip(rep * (cur[d] - neg[d])) * alpha : 4.0f * alpha;
            cur[d] += g;
            grads[d] += g;
            step[d] = g;
          }
          if (trace)
            record(epoch, e, 1, (int32_t)r, alpha, d2n, rep, y0, neg, step);
        }
        nneg[e] = nn + (float)n_neg * ep_neg;
      }
      for (int d = 0; d < c; ++d)
        add(&cur_p[d], (rounding + grads[d]) - rounding);
    }
    alpha = alpha0 * (1.0f - (float)epoch / (float)n_epochs);
  }
}

// --------------------------------------------------------------- engine
struct Oracle {
  or_params p;
  uint64_t N, D, Dp;
  int k, c;
  bool euclid;
  uint64_t n = 0; // inserted rows
  std::vector<float> X;
  std::vector<uint64_t> pool;
  std::vector<uint32_t> pidx;
  std::vector<float> pdist;
  std::vector<float> emb;
  Fifo insertion_queue, embedding_new_queue;
  DedupQueue update_queue;
  std::deque<uint32_t> batch_start_history;
  size_t embedding_cursor = SIZE_MAX;
  size_t embedding_sweeps_done = 0;
  bool insertion_completed = false, update_completed = false,
       embedding_completed = false;
  std::mt19937 rng{42};
  int nt;

  explicit Oracle(const or_params &pp) : p(pp) {
    N = p.n_instances;
    D = p.n_features;
    Dp = (D + 3) / 4 * 4;
    k = (int)p.n_neighbors;
    c = (int)p.n_components;
    euclid = p.metric == 0;
    nt = p.n_threads > 0 ? p.n_threads : 1;
    X.assign(N * Dp, 0.0f);
    pool.assign(N * W, EMPTY);
    pidx.assign(N * k, 0xffffffffu);
    pdist.assign(N * k, FLT_MAX);
    emb.assign(N * c, 0.0f);
  }

  float dist(uint32_t a, uint32_t b) const {
    return canon_dist(&X[a * Dp], &X[b * Dp], (int)Dp);
  }

  int append(const float *x, uint64_t rows) {
    if (rows == 0)
      return 0;
    if (n + rows > N)
      return -1;
    for (uint64_t r = 0; r < rows; ++r) {
      std::memcpy(&X[(n + r) * Dp], x + r * D, D * sizeof(float));
      insertion_queue.q.push_back((uint32_t)(n + r));
      embedding_new_queue.q.push_back((uint32_t)(n + r));
    }
    n += rows;
    return 0;
  }

  // Sorted top-k of the pool -> public row; returns #ids not in the old row.
  size_t sync_row(uint64_t i) {
    uint64_t s[W];
    const int cnt = sorted_row(&pool[i * W], s);
    uint32_t old[64];
    std::memcpy(old, &pidx[i * k], sizeof(uint32_t) * k);
    size_t fresh = 0;
    for (int j = 0; j < k; ++j) {
      if (j < cnt) {
        const uint32_t id = key_id(s[j]);
        pidx[i * k + j] = id;
        pdist[i * k + j] = key_dist(s[j]);
        bool seen = false;
        for (int q = 0; q < k; ++q)
          seen |= (old[q] == id);
        fresh += !seen;
      } else {
        pidx[i * k + j] = 0xffffffffu;
        pdist[i * k + j] = FLT_MAX;
      }
    }
    return fresh;
  }

  // ---------------------------------------------------------- insertion
  // reference Impl::insert (impl.cuh:220-253) + run_insertion
  // (insertion_kernel.cuh:926-1095), with deviations D1-D3.
  uint32_t insert(uint32_t n_ops) {
    if (insertion_queue.size() == 0 || n_ops == 0)
      return 0;
    std::vector<uint32_t> batch = insertion_queue.pop_n(n_ops);
    batch_start_history.push_back(batch.front());
    if ((int)batch_start_history.size() > WINDOW_N_BATCHES)
      batch_start_history.pop_front();
    const uint32_t window_start = batch_start_history.front();
    const uint32_t seed = n_ops * 9973u + (uint32_t)n * 7919u;
    const uint64_t b0 = batch.front(), B = batch.size();

    // norms / cosine normalisation (preprocess_l2_norms_kernel :31-81)
    for (uint64_t r = b0; r < b0 + B; ++r) {
      float *x = &X[r * Dp];
      if (!euclid) {
        float s = 0.0f;
        for (uint64_t i = 0; i < Dp; ++i)
          s = std::fmaf(x[i], x[i], s);
        const float nrm = sqrtf(s);
        if (nrm > 0.0f)
          for (uint64_t i = 0; i < D; ++i)
            x[i] = x[i] / nrm;
      }
    }

    std::vector<uint8_t> dirty(n, 0);
    if (p.knn_mode == 1) {
      exact_insert(b0, B, dirty);
    } else {
      nnd_insert(b0, B, window_start, seed, dirty);
    }

    // affected rows (CPU diff :1034-1089): batch rows + changed existing rows
    std::vector<uint32_t> affected(batch.begin(), batch.end());
    for (uint64_t r = 0; r < b0; ++r)
      if (dirty[r])
        affected.push_back((uint32_t)r);
    for (uint32_t r : affected)
      sync_row(r);
    init_embedding(b0, B);
    for (uint32_t r : affected)
      update_queue.push(r);
    return (uint32_t)B;
  }

  // exact insertion: forward top-W over [0, n) and reverse offers to [0, b0)
  void exact_insert(uint64_t b0, uint64_t B, std::vector<uint8_t> &dirty) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (uint64_t bi = 0; bi < B; ++bi) {
      const uint32_t u = (uint32_t)(b0 + bi);
      uint64_t *row = &pool[(size_t)u * W];
      for (int j = 0; j < W; ++j)
        row[j] = EMPTY;
      for (uint64_t j = 0; j < n; ++j) {
        if (j == u)
          continue;
        offer(row, k, make_key(dist(u, (uint32_t)j), (uint32_t)j, false));
      }
    }
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (uint64_t j = 0; j < b0; ++j) {
      for (uint64_t bi = 0; bi < B; ++bi) {
        const uint32_t u = (uint32_t)(b0 + bi);
        if (offer(&pool[j * W], k, make_key(dist(u, (uint32_t)j), u, false)))
          dirty[j] = 1;
      }
    }
  }

  void nnd_insert(uint64_t b0, uint64_t B, uint32_t window_start,
                  uint32_t seed, std::vector<uint8_t> &dirty) {
    // random initial neighbours (initialize_random_neighbors_kernel :170-248)
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (uint64_t bi = 0; bi < B; ++bi) {
      const uint32_t u = (uint32_t)(b0 + bi);
      uint64_t *row = &pool[(size_t)u * W];
      for (int j = 0; j < W; ++j)
        row[j] = EMPTY;
      if (n <= 1)
        continue;
      Rng rng(seed, u, 0, TAG_INIT_NBR);
      const int max_fill = (int)std::min<uint64_t>(W, n - 1);
      const int window_size = (int)n - (int)window_start;
      const int quota = std::min(W / 2, std::max(0, window_size - 1));
      uint32_t ids[W];
      int filled = 0, attempts = 0;
      auto has = [&](uint32_t v) {
        for (int q = 0; q < filled; ++q)
          if (ids[q] == v)
            return true;
        return false;
      };
      while (filled < quota && attempts < quota * 16) {
        ++attempts;
        const uint32_t cand = window_start + rng.next() % (uint32_t)window_size;
        if (cand == u || has(cand))
          continue;
        ids[filled++] = cand;
      }
      while (filled < max_fill) {
        const uint32_t cand = rng.next() % (uint32_t)n;
        if (cand == u || has(cand))
          continue;
        ids[filled++] = cand;
      }
      for (int q = 0; q < filled; ++q)
        row[q] = make_key(dist(u, ids[q]), ids[q], true);
    }

    std::vector<uint32_t> cand((size_t)B * C);
    std::vector<float> cdist((size_t)B * C);
    std::vector<int> ccount(B);
    const int max_iter = p.max_iter > 0 ? p.max_iter : 10;
    for (int it = 0; it < max_iter; ++it) {
      // neighbours-of-neighbours candidates (sample_forward_candidates_kernel
      // :250-321) + canonical distances (compute_candidate_distances :337-405)
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
      for (uint64_t bi = 0; bi < B; ++bi) {
        const uint32_t u = (uint32_t)(b0 + bi);
        uint64_t su[W];
        const int cu = sorted_row(&pool[(size_t)u * W], su);
        const int kk = std::min(k, cu);
        uint32_t top[64];
        uint64_t nb[64][W];
        int nbc[64];
        for (int j = 0; j < kk; ++j) {
          top[j] = key_id(su[j]);
          nbc[j] = sorted_row(&pool[(size_t)top[j] * W], nb[j]);
        }
        uint32_t *cl = &cand[bi * C];
        int cc = 0;
        for (int s = 0; s < k && cc < C; ++s)
          for (int j = 0; j < kk && cc < C; ++j) {
            if (s >= nbc[j])
              continue;
            const uint32_t cid = key_id(nb[j][s]);
            if (cid == u)
              continue;
            bool skip = false;
            for (int q = 0; q < W && !skip; ++q)
              skip = pool[(size_t)u * W + q] != EMPTY &&
                     key_id(pool[(size_t)u * W + q]) == cid;
            for (int q = 0; q < cc && !skip; ++q)
              skip = cl[q] == cid;
            if (!skip)
              cl[cc++] = cid;
          }
        ccount[bi] = cc;
        for (int q = 0; q < cc; ++q)
          cdist[bi * C + q] = dist(u, cl[q]);
      }
      // forward + reverse inserts (update_forward_neighbors_kernel :486-534,
      // scatter/update_reverse :548-640); order-independent top-W updates
      for (uint64_t bi = 0; bi < B; ++bi) {
        const uint32_t u = (uint32_t)(b0 + bi);
        for (int q = 0; q < ccount[bi]; ++q) {
          const uint32_t cid = cand[bi * C + q];
          const float d = cdist[bi * C + q];
          if (offer(&pool[(size_t)u * W], k, make_key(d, cid, true)))
            dirty[u] = 1;
          if (offer(&pool[(size_t)cid * W], k, make_key(d, u, true)))
            dirty[cid] = 1;
        }
      }
    }

  }

  // embedding init (initialize_embeddings :790-878)
  void init_embedding(uint64_t b0, uint64_t B) {
    for (uint64_t bi = 0; bi < B; ++bi) {
      const uint64_t u = b0 + bi;
      Rng rng(42, (uint32_t)u, 0, TAG_EMB_INIT);
      for (int d = 0; d < c; ++d)
        emb[u * c + d] = rng.normal();
      float dmin = FLT_MAX;
      for (int j = 0; j < k; ++j) {
        const uint32_t id = pidx[u * k + j];
        if (id < b0)
          dmin = std::min(dmin, pdist[u * k + j]);
      }
      if (dmin == FLT_MAX)
        continue;
      float wsum = 0.0f, acc[4] = {0, 0, 0, 0};
      for (int j = 0; j < k; ++j) {
        const uint32_t id = pidx[u * k + j];
        if (id >= b0)
          continue;
        const float wj = expf(-(pdist[u * k + j] - dmin));
        wsum += wj;
        for (int d = 0; d < c; ++d)
          acc[d] += wj * emb[(size_t)id * c + d];
      }
      if (wsum > 0.0f)
        for (int d = 0; d < c; ++d)
          emb[u * c + d] = acc[d] / wsum;
    }
  }

  // ------------------------------------------------------------- update
  // One NN-Descent pass over `rows` (UpdateAlgo::run, update_kernel.cuh
  // :913-1054) with deviations D3/D4.  Returns #fresh top-k edges.
  size_t nnd_pass(std::vector<uint32_t> rows,
                  std::vector<uint32_t> &affected) {
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    std::vector<uint32_t> fn((size_t)n * W, 0xffffffffu),
        fo((size_t)n * W, 0xffffffffu);
    std::vector<int> fnc(n, 0), foc(n, 0);
    std::vector<uint64_t> rn((size_t)n * W, EMPTY), ro((size_t)n * W, EMPTY);
    std::vector<uint8_t> owner(n, 0), dirty(n, 0);
    // sampling (sample_graph :750-807): every new entry goes to the new list
    // and is marked old; every old entry goes to the old list
    for (uint32_t u : rows) {
      uint64_t *row = &pool[(size_t)u * W];
      for (int j = 0; j < W; ++j) {
        if (row[j] == EMPTY)
          continue;
        if (row[j] & NEW_BIT) {
          fn[(size_t)u * W + fnc[u]++] = key_id(row[j]);
          row[j] &= ~NEW_BIT;
        } else {
          fo[(size_t)u * W + foc[u]++] = key_id(row[j]);
        }
      }
    }
    // reverse lists (add_rev_edges_kernel :122-146): 32 nearest by (d, id)
    for (uint32_t u : rows) {
      const uint64_t *row = &pool[(size_t)u * W];
      for (int j = 0; j < W; ++j) {
        if (row[j] == EMPTY)
          continue;
        const uint32_t v = key_id(row[j]);
        const float d = key_dist(row[j]);
        bool is_new = false;
        for (int q = 0; q < fnc[u]; ++q)
          is_new |= fn[(size_t)u * W + q] == v;
        offer(is_new ? &rn[(size_t)v * W] : &ro[(size_t)v * W], 0,
              make_key(d, u, false));
      }
    }
    std::vector<uint32_t> owners;
    for (uint64_t v = 0; v < n; ++v) {
      bool has_new = fnc[v] > 0;
      for (int j = 0; j < W && !has_new; ++j)
        has_new = rn[v * W + j] != EMPTY;
      if (has_new)
        owners.push_back((uint32_t)v);
    }
    // local join (local_join_kernel :377-687): offers computed in parallel,
    // applied in owner order (the result does not depend on the order)
    const size_t chunk = 256;
    for (size_t o0 = 0; o0 < owners.size(); o0 += chunk) {
      const size_t o1 = std::min(owners.size(), o0 + chunk);
      std::vector<std::vector<std::pair<uint32_t, uint64_t>>> offs(o1 - o0);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
      for (size_t oi = o0; oi < o1; ++oi) {
        const uint32_t v = owners[oi];
        std::vector<uint32_t> nw, od;
        for (int q = 0; q < fnc[v]; ++q)
          nw.push_back(fn[(size_t)v * W + q]);
        for (int j = 0; j < W; ++j)
          if (rn[(size_t)v * W + j] != EMPTY)
            nw.push_back(key_id(rn[(size_t)v * W + j]));
        for (int q = 0; q < foc[v]; ++q)
          od.push_back(fo[(size_t)v * W + q]);
        for (int j = 0; j < W; ++j)
          if (ro[(size_t)v * W + j] != EMPTY)
            od.push_back(key_id(ro[(size_t)v * W + j]));
        std::sort(nw.begin(), nw.end());
        nw.erase(std::unique(nw.begin(), nw.end()), nw.end());
        std::sort(od.begin(), od.end());
        od.erase(std::unique(od.begin(), od.end()), od.end());
        std::vector<uint32_t> od2;
        for (uint32_t x : od)
          if (!std::binary_search(nw.begin(), nw.end(), x))
            od2.push_back(x);
        auto &o = offs[oi - o0];
        const size_t a = nw.size();
        for (size_t r = 0; r < a; ++r) {
          for (size_t q = r + 1; q < a; ++q) {
            const float d = dist(nw[r], nw[q]);
            o.push_back({nw[r], make_key(d, nw[q], true)});
            o.push_back({nw[q], make_key(d, nw[r], true)});
          }
          for (uint32_t y : od2) {
            const float d = dist(nw[r], y);
            o.push_back({nw[r], make_key(d, y, true)});
            o.push_back({y, make_key(d, nw[r], true)});
          }
        }
      }
      for (auto &o : offs)
        for (auto &pr : o)
          if (offer(&pool[(size_t)pr.first * W], k, pr.second) == 2)
            dirty[pr.first] = 1;
    }
    size_t fresh = 0;
    for (uint64_t v = 0; v < n; ++v)
      if (dirty[v]) {
        fresh += sync_row(v);
        affected.push_back((uint32_t)v);
      }
    return fresh;
  }

  // reference Impl::update (impl.cuh:255-333)
  uint32_t update(uint32_t n_ops, size_t &changed_out) {
    changed_out = 0;
    if (n_ops == 0 || update_completed || n == 0)
      return 0;
    // converged once the queue holds fewer than 1e-4 n rows; at least the
    // empty queue counts (reference impl.cuh:261-263 uses a threshold of 0
    // below n = 10,000, which no queue size satisfies)
    const size_t thr = std::max<size_t>(1, (size_t)(0.0001f * (float)n));
    uint32_t left = n_ops;
    while (left > 0) {
      std::vector<uint32_t> b;
      if ((size_t)left >= n) {
        update_queue.pop_n(update_queue.size());
        b.resize(n);
        std::iota(b.begin(), b.end(), 0u);
      } else if ((size_t)left >= update_queue.size()) {
        b = update_queue.pop_n(update_queue.size());
        std::unordered_set<uint32_t> seen(b.begin(), b.end());
        std::uniform_int_distribution<uint32_t> dd(0, (uint32_t)(n - 1));
        uint32_t off = dd(rng);
        while (b.size() < (size_t)left) {
          const uint32_t cnd = (off++) % (uint32_t)n;
          if (seen.insert(cnd).second)
            b.push_back(cnd);
        }
      } else {
        b = update_queue.pop_n(left);
      }
      if (b.empty())
        break;
      std::vector<uint32_t> affected;
      changed_out += nnd_pass(b, affected);
      left -= (uint32_t)b.size();
      for (uint32_t r : affected)
        update_queue.push(r);
      if (update_queue.size() < thr) {
        update_completed = (n == N);
        break;
      }
    }
    return n_ops - left;
  }

  // -------------------------------------------------------------- embed
  // reference Impl::embed (impl.cuh:335-500)
  uint32_t embed(uint32_t n_ops) {
    if (n_ops == 0 || n == 0)
      return 0;
    std::vector<uint32_t> batch;
    if ((size_t)n_ops >= n) {
      embedding_new_queue.pop_n(embedding_new_queue.size());
      batch.resize(n);
      std::iota(batch.begin(), batch.end(), 0u);
      embedding_cursor = n - 1;
      if (embedding_cursor == N - 1 && update_completed)
        ++embedding_sweeps_done;
    } else {
      if (embedding_new_queue.size() > 0) {
        auto it = embedding_new_queue.pop_n(embedding_new_queue.size());
        batch.insert(batch.end(), it.begin(), it.end());
      }
      size_t left = batch.size() >= n_ops ? 0 : n_ops - batch.size();
      if (left > 0) {
        std::unordered_set<uint32_t> seen(batch.begin(), batch.end());
        size_t cur = embedding_cursor == SIZE_MAX ? 0 : (embedding_cursor + 1) % n;
        for (size_t chk = 0; left > 0 && chk < n; ++chk) {
          if (seen.insert((uint32_t)cur).second) {
            batch.push_back((uint32_t)cur);
            embedding_cursor = cur;
            if (embedding_cursor == N - 1 && update_completed)
              ++embedding_sweeps_done;
            --left;
          }
          cur = (cur + 1) % n;
        }
      }
    }
    std::vector<uint8_t> in_batch(n, 0);
    for (uint32_t id : batch)
      in_batch[id] = 1;
    std::vector<uint32_t> unsel;
    for (uint64_t i = 0; i < n; ++i)
      if (!in_batch[i])
        unsel.push_back((uint32_t)i);

    Fuzzy fz;
    fuzzy_set(pidx.data(), pdist.data(), n, k, euclid, batch.data(),
              batch.size(), (int)p.max_epoch, fz);
    if (embedding_sweeps_done > 2)
      embedding_completed = update_completed;

    // noise (add_embedding_noise :953-1007)
    {
      float mn = FLT_MAX, mx = -FLT_MAX;
      for (uint64_t i = 0; i < n; ++i) {
        mn = std::min(mn, emb[i * c]);
        mx = std::max(mx, emb[i * c]);
      }
      const float scale = 0.001f * (mx - mn);
      const uint64_t nseed = 42ull ^ (uint64_t)embedding_sweeps_done;
      for (size_t t = 0; t < batch.size() * (size_t)c; ++t) {
        Rng rng(nseed, (uint32_t)t, (uint32_t)(t >> 32), TAG_NOISE);
        emb[(size_t)batch[t / c] * c + t % c] += rng.normal() * scale;
      }
    }
    sgd_epochs(emb.data(), n, c, fz.head.data(), fz.tail.data(), fz.w.data(),
               (int64_t)fz.w.size(), (int)p.max_epoch,
               (int)p.negative_sample_rate, p.a, p.b, p.alpha, 1.0f, 42ull, nt);

    // reposition unselected rows (heuristic_neighbor_avg_kernel :811-894)
    if (!unsel.empty()) {
      std::vector<float> snap(emb.begin(), emb.begin() + n * c);
      std::vector<uint32_t> ri(k);
      std::vector<double> rd(k);
      for (uint32_t i : unsel) {
        embed_row(pidx.data(), pdist.data(), i, k, euclid, ri.data(), rd.data());
        double wsum = 0.0, acc[4] = {0, 0, 0, 0};
        for (int q = 0; q < k; ++q) {
          if (ri[q] == i || ri[q] >= n || !valid_dist(rd[q]))
            continue;
          const double wq = membership(rd[q], fz.rho[i], fz.sigma[i]);
          wsum += wq;
          for (int d = 0; d < c; ++d)
            acc[d] += wq * snap[(size_t)ri[q] * c + d];
        }
        if (wsum > 0.0)
          for (int d = 0; d < c; ++d)
            emb[(size_t)i * c + d] = (float)(acc[d] / wsum);
      }
    }
    return (uint32_t)batch.size();
  }

  // reference Impl::run (impl.cuh:146-218)
  void run(uint32_t ni, uint32_t nu, uint32_t ne, or_result *r) {
    using clk = std::chrono::steady_clock;
    std::memset(r, 0, sizeof(*r));
    if (!insertion_completed) {
      auto t0 = clk::now();
      r->n_insertion_ops = insert(ni);
      r->insertion_time = std::chrono::duration<float>(clk::now() - t0).count();
      insertion_completed = insertion_queue.size() == 0 && n == N;
    }
    if (!(nu == 0 || update_completed)) {
      size_t ch = 0;
      auto t0 = clk::now();
      r->n_update_ops = update(nu, ch);
      r->update_time = std::chrono::duration<float>(clk::now() - t0).count();
      r->n_update_changed_edges = ch;
    }
    r->embedding_pending_edges = n;
    if (!(ne == 0 || n == 0 || embedding_completed)) {
      auto t0 = clk::now();
      r->n_embedding_ops = embed(ne);
      r->embedding_time = std::chrono::duration<float>(clk::now() - t0).count();
    }
    r->n_inserted_instances = n;
    r->insertion_queue_size = insertion_queue.size();
    r->update_queue_size = update_queue.size();
    r->embedding_queue_size = r->embedding_pending_edges;
    r->insertion_completed = insertion_completed;
    r->update_completed = update_completed;
    r->embedding_completed = embedding_completed;
  }
};

} // namespace

extern "C" {

void *or_create(const or_params *p) {
  if (!p || p->n_instances == 0 || p->n_features == 0 || p->n_neighbors < 2 ||
      p->n_neighbors > 32 || p->n_components < 1 || p->n_components > 4)
    return nullptr;
  return new Oracle(*p);
}
void or_destroy(void *h) { delete static_cast<Oracle *>(h); }
int or_append(void *h, const float *x, uint64_t rows, uint64_t cols) {
  auto *o = static_cast<Oracle *>(h);
  if (cols != o->D)
    return -2;
  return o->append(x, rows);
}
int or_run(void *h, uint32_t a, uint32_t b, uint32_t c, or_result *out) {
  static_cast<Oracle *>(h)->run(a, b, c, out);
  return 0;
}
void or_get_knn(void *h, uint32_t *idx, float *dist) {
  auto *o = static_cast<Oracle *>(h);
  std::memcpy(idx, o->pidx.data(), o->n * o->k * sizeof(uint32_t));
  std::memcpy(dist, o->pdist.data(), o->n * o->k * sizeof(float));
}
void or_get_embedding(void *h, float *e) {
  auto *o = static_cast<Oracle *>(h);
  std::memcpy(e, o->emb.data(), o->n * o->c * sizeof(float));
}
void or_set_embedding(void *h, const float *e) {
  auto *o = static_cast<Oracle *>(h);
  std::memcpy(o->emb.data(), e, o->n * o->c * sizeof(float));
}
void or_get_pool(void *h, uint64_t *keys) {
  auto *o = static_cast<Oracle *>(h);
  std::memcpy(keys, o->pool.data(), o->n * W * sizeof(uint64_t));
}
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2],
                   uint32_t out[4]) {
  philox(ctr, key, out);
}
float or_canon_dist(const float *a, const float *b, int d) {
  return canon_dist(a, b, d);
}
int64_t or_fuzzy(const uint32_t *idx, const float *dist, uint64_t n, int k,
                 int euclidean, const uint32_t *batch, uint64_t n_batch,
                 int n_epochs, int32_t *head, int32_t *tail, float *w,
                 double *sigmas, double *rhos) {
  Fuzzy fz;
  fuzzy_set(idx, dist, n, k, euclidean != 0, batch, n_batch, n_epochs, fz);
  const size_t nnz = fz.w.size();
  if (head)
    std::memcpy(head, fz.head.data(), nnz * 4);
  if (tail)
    std::memcpy(tail, fz.tail.data(), nnz * 4);
  if (w)
    std::memcpy(w, fz.w.data(), nnz * 4);
  if (sigmas)
    std::memcpy(sigmas, fz.sigma.data(), n * 8);
  if (rhos)
    std::memcpy(rhos, fz.rho.data(), n * 8);
  return (int64_t)nnz;
}
int64_t or_sgd_trace(float *emb, uint64_t n, int c, const int32_t *head,
                     const int32_t *tail, const float *w, int64_t nnz,
                     int n_epochs, int nsr, float a, float b, float alpha0,
                     float gamma, uint64_t seed, or_sgd_event *events,
                     int64_t cap, float *rounding) {
  std::vector<or_sgd_event> tr;
  float r = 0.0f;
  sgd_epochs(emb, n, c, head, tail, w, nnz, n_epochs, nsr, a, b, alpha0,
             gamma, seed, 1, &tr, &r);
  if (rounding)
    *rounding = r;
  const int64_t m = std::min<int64_t>(cap, (int64_t)tr.size());
  for (int64_t i = 0; i < m; ++i)
    events[i] = tr[i];
  return (int64_t)tr.size();
}

float or_rounding_factor(const int32_t *head, int64_t nnz, uint64_t n,
                         float alpha) {
  std::vector<int32_t> hv(head, head + nnz);
  return rounding_factor(hv, n, alpha);
}

void or_sgd(float *emb, uint64_t n, int c, const int32_t *head,
            const int32_t *tail, const float *w, int64_t nnz, int n_epochs,
            int nsr, float a, float b, float alpha0, float gamma,
            uint64_t seed, int n_threads) {
  sgd_epochs(emb, n, c, head, tail, w, nnz, n_epochs, nsr, a, b, alpha0, gamma,
             seed, n_threads);
}

} // extern "C"
