// NN-Descent local join screened by BF16 MFMA distance tiles
// (knn_mode nndescent_bf16).
//
// Same result as k_join (nnd_kernels.hip, the reference's local_join_kernel,
// update_kernel.cuh:377-687): every pair of an owner's lists whose canonical
// FP32 distance beats an endpoint's pool maximum is offered to that pool, so
// the pools equal the FP32 join's bit for bit.  The distance work moves to
// the matrix cores:
//   * every inserted row is stored as hi = bf16(x), lo = bf16(x - hi) in
//     32-feature chunks [hi 32 | lo 32] (128 bytes, k_join_prep) beside its
//     FP32 squared norm, summed as k_bf16_prep sums it;
//   * per owner, the members' chunks are staged in LDS and 32 x 32 blocks of
//       d~ = |r|^2 + |c|^2 - 2 (hi_c.hi_r + hi_c.lo_r + lo_c.hi_r)
//     accumulate on v_mfma_f32_32x32x16_bf16, BF16_GROUP_STEPS feature steps
//     per group on the matrix core and the groups added in round-to-nearest
//     FP32, which is the accumulation the BF16 insertion screen's bound
//     covers (mfma_kernels.hip; coef and cfac from the host);
//   * a pair survives when its lower bound (d~ - coef s)(1 - cfac) is at or
//     below either endpoint's pool-maximum distance as loaded for the owner.
//     A pool maximum only falls while the pass runs, so a screened-out pair
//     would have been refused by both pools;
//   * the survivors' canonical distances (one FMA chain over the features,
//     the FP32 rows staged in LDS chunk by chunk) go through k_join's offer
//     filter and the lock-free bounded insert, whose final rows do not depend
//     on the order of the offers.
//
// One block of 256 threads per owner (grid-strided), two blocks per CU.
#include "common.hpp"
#include "kernels.hpp"

namespace gm {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

constexpr int JM = 128;        // list members at most (64 new + 64 old)
constexpr int JKC = 32;        // features per staged chunk
constexpr int JRS = 144;       // LDS bytes per member row and chunk (128 + pad)
constexpr int JBUF = JM * JRS; // one staging buffer
constexpr int JSV = 64 * JM;   // survivors at most (new rows x members)
constexpr int JRR = 2048;      // survivors re-ranked per round (8 per thread)
constexpr int JQ = 512;        // offers per sub-round (2 per thread)
// survivor entry: new row << 7 | member, and the directions that may offer
constexpr int SV_FWD = 0x2000; // the member to the new row's pool
constexpr int SV_REV = 0x4000; // the new row to the member's pool

// is `id` among the 32 ascending ids at `ids`
__device__ inline bool in_sorted32(const uint32_t *ids, uint32_t id) {
  int lo = 0, hi = W;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (ids[m] < id)
      lo = m + 1;
    else
      hi = m;
  }
  return lo < W && ids[lo] == id;
}

} // namespace

// hi/lo BF16 split of rows [b0, b0 + nb) into the join's operand store (row
// u at Xj + u * dj * 2: dj / 32 chunks of 32 hi then 32 lo values) and their
// FP32 squared norms.  One 64-thread block per row; each lane sums its own
// 8-feature groups in order, then a 6-level xor tree, the summation the
// screening bound's norm term assumes.
__global__ __launch_bounds__(64) void k_join_prep(
    const float *__restrict__ X, uint64_t dp, uint64_t dj,
    __bf16 *__restrict__ Xj, float *__restrict__ nrm, uint64_t b0,
    uint64_t nb) {
  const uint64_t r = blockIdx.x;
  if (r >= nb)
    return;
  const uint64_t u = b0 + r;
  const float *x = X + u * dp;
  __bf16 *row = Xj + u * dj * 2;
  const int lane = threadIdx.x;
  float s = 0.0f;
  for (uint64_t t = lane; t < dj / 8; t += 64) {
    const uint64_t f0 = t * 8;
    bf16x8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float v = f0 + e < dp ? x[f0 + e] : 0.0f;
      s = __builtin_fmaf(v, v, s);
      const __bf16 h = (__bf16)v;
      hi[e] = h;
      lo[e] = (__bf16)(v - (float)h);
    }
    __bf16 *chunk = row + (f0 / JKC) * (2 * JKC) + (f0 % JKC);
    *reinterpret_cast<bf16x8 *>(chunk) = hi;
    *reinterpret_cast<bf16x8 *>(chunk + JKC) = lo;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    s += __shfl_xor(s, off, 64);
  if (lane == 0)
    nrm[u] = s;
}

__global__ __launch_bounds__(256, 2) void k_join_bf16(
    const float *__restrict__ X, uint64_t dp, const __bf16 *__restrict__ Xj,
    uint64_t dj, const float *__restrict__ nrm, uint64_t *__restrict__ pool,
    const uint32_t *__restrict__ fn, const uint32_t *__restrict__ fo,
    const int *__restrict__ fnc, const int *__restrict__ foc,
    const uint64_t *__restrict__ rev_new, const uint64_t *__restrict__ rev_old,
    const uint32_t *__restrict__ owners,
    const uint32_t *__restrict__ owner_count, uint8_t *__restrict__ flag,
    int k, float coef, float cfac, unsigned long long *__restrict__ stats) {
  __shared__ uint32_t s_ids[JM];
  __shared__ uint64_t s_mmax[JM];
  __shared__ float s_nrm[JM], s_thr[JM];
  // staging ring (BF16 chunks, then FP32 chunks); before them the scratch
  // of the old list, between them the members' sorted pool ids
  __shared__ __attribute__((aligned(16))) unsigned char s_ring[2 * JBUF];
  __shared__ uint16_t s_sv[JSV]; // survivors (SV_FWD, SV_REV)
  __shared__ float s_svd[JRR];   // their canonical distances (one round)
  __shared__ uint64_t s_qkey[JQ];
  __shared__ uint8_t s_qtgt[JQ];
  __shared__ int s_a, s_o, s_nsv, s_qn;
  uint32_t *const s_tmp = reinterpret_cast<uint32_t *>(s_ring);
  uint32_t *const s_mid = reinterpret_cast<uint32_t *>(s_ring);
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int hw = tid >> 5, hl = tid & 31, half = (tid >> 5) & 1;
  const uint32_t n_owner = *owner_count;
  const int nch = (int)(dj / JKC);          // BF16 chunks
  const int nsteps = 2 * nch;               // 16-feature MFMA steps
  const int nchf = (int)((dp + JKC - 1) / JKC); // FP32 chunks
  unsigned long long c_members = 0, c_pairs = 0, c_offers = 0, c_surv = 0;
  const floatx16 z = {};

  for (uint32_t oi = blockIdx.x; oi < n_owner; oi += gridDim.x) {
    const uint32_t v = owners[oi];
    // ---- 1. lists: new = fwd_new U rev_new, old = (fwd_old U rev_old) \ new
    if (tid == 0)
      s_nsv = 0;
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
        s_tmp[__popcll(km & ((1ull << lane) - 1ull))] = x;
      if (lane == 0)
        s_o = __popcll(km);
    }
    __syncthreads();
    const int a = s_a;
    if (a == 0) {
      __syncthreads();
      continue;
    }
    if (wave == 1) {
      const int o0 = s_o;
      const uint32_t x = lane < o0 ? s_tmp[lane] : NO_ID;
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

    // ---- 2. members' norms and pool maxima (key and distance threshold)
    if (tid < JM) {
      uint64_t mx = 0;
      float nv = 0.0f;
      if (tid < m_all) {
        const uint32_t id = s_ids[tid];
        const uint64_t *r = pool + (uint64_t)id * W;
        for (int s = 0; s < W; ++s) {
          const uint64_t key = r[s];
          mx = key > mx ? key : mx;
        }
        nv = nrm[id];
      }
      s_mmax[tid] = mx;
      s_nrm[tid] = nv;
      s_thr[tid] = mx == EMPTY ? INFINITY : key_dist(mx);
    }

    // ---- 3. screening tiles: 32 new rows x 32 members per MFMA block; wave
    // w takes blocks w and w + 4 of the (new block, member block) grid, when
    // a block holds a pair the join computes (member above the row, or old)
    const int RB = (a + 31) >> 5, CB = (m_all + 31) >> 5;
    bool need[2];
    int trb[2], tcb[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int t = wave + 4 * q;
      trb[q] = t / CB;
      tcb[q] = t % CB;
      const int max_c = min(tcb[q] * 32 + 31, m_all - 1);
      need[q] = t < RB * CB && (max_c >= a || max_c > trb[q] * 32);
    }
    floatx16 acc[2], grp[2];
#pragma unroll
    for (int q = 0; q < 2; ++q)
      acc[q] = grp[q] = z;
    // staging: 8 threads per member row, 16 bytes each; members past the
    // list stage zeros.  Two register sets in flight (chunks ch + 1, ch + 2).
    uint4 ra[4], rb[4];
    auto bload = [&](int ch, uint4 (&r)[4]) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int seg = tid + 256 * j, m = seg >> 3, pc = seg & 7;
        r[j] = m < m_all ? *reinterpret_cast<const uint4 *>(
                               Xj + (uint64_t)s_ids[m] * dj * 2 +
                               (uint64_t)ch * (2 * JKC) + pc * 8)
                         : make_uint4(0u, 0u, 0u, 0u);
      }
    };
    auto bstore = [&](int buf, const uint4 (&r)[4]) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int seg = tid + 256 * j, m = seg >> 3, pc = seg & 7;
        *reinterpret_cast<uint4 *>(s_ring + buf * JBUF + m * JRS + pc * 16) =
            r[j];
      }
    };
    int st = 0; // MFMA step
    // one chunk's two feature steps from LDS buffer `buf`
    auto bstep = [&](int buf) {
      const unsigned char *rbuf = s_ring + buf * JBUF;
#pragma unroll
      for (int h = 0; h < 2; ++h, ++st) {
        const bool fresh = (st % BF16_GROUP_STEPS) == 0;
        const bool close =
            (st % BF16_GROUP_STEPS) == BF16_GROUP_STEPS - 1 || st == nsteps - 1;
        const int off = h * 32 + (lane >> 5) * 16;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (!need[q])
            continue;
          // A: the member block (columns of the distance block), B: the new
          // rows; C register i of lane l is member (i / 4) * 8 + (l / 32) * 4
          // + i % 4, new row l % 32 of the block
          const unsigned char *pa = rbuf + (tcb[q] * 32 + (lane & 31)) * JRS;
          const unsigned char *pb = rbuf + (trb[q] * 32 + (lane & 31)) * JRS;
          const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(pa + off);
          const bf16x8 al = *reinterpret_cast<const bf16x8 *>(pa + 64 + off);
          const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(pb + off);
          const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(pb + 64 + off);
          grp[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh,
                                                           fresh ? z : grp[q],
                                                           0, 0, 0);
          grp[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, grp[q], 0, 0,
                                                           0);
          grp[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, grp[q], 0, 0,
                                                           0);
          if (close)
            acc[q] += grp[q];
        }
      }
    };
    bload(0, ra);
    if (nch > 1)
      bload(1, rb);
    bstore(0, ra);
    __syncthreads();
    // two chunks per pass so that the register sets keep their roles
    for (int ch = 0; ch < nch; ch += 2) {
      if (ch + 2 < nch)
        bload(ch + 2, ra);
      bstep(0);
      if (ch + 1 < nch)
        bstore(1, rb);
      __syncthreads();
      if (ch + 1 >= nch)
        break;
      if (ch + 3 < nch)
        bload(ch + 3, rb);
      bstep(1);
      if (ch + 2 < nch)
        bstore(0, ra);
      __syncthreads();
    }

    // ---- 4. survivors: a direction survives when the lower bound is at or
    // below the receiving pool's maximum and the other member is not in that
    // pool (the join's offer filter, with the pools' ids sorted per member;
    // filtering an offer never changes a pool: an id in a pool is refused,
    // and once evicted it lies above the pool's new maximum)
    for (int m = hw; m < m_all; m += 8) {
      const uint64_t key = pool[(uint64_t)s_ids[m] * W + hl];
      s_mid[m * W + hl] = half_sort32(key == EMPTY ? NO_ID : key_id(key));
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (!need[q])
        continue;
      const int r = trb[q] * 32 + (lane & 31);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int c = tcb[q] * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        if (r >= a || c >= m_all || (c < a && c <= r))
          continue;
        const float s = s_nrm[r] + s_nrm[c];
        const float dt = __builtin_fmaf(-2.0f, acc[q][i], s);
        const float bnd = __builtin_fmaf(coef, s, 1e-30f);
        const float lb = (dt - bnd) * (1.0f - cfac);
        // (a NaN bound survives: the exact distance decides)
        const bool fwd =
            !(lb > s_thr[r]) && !in_sorted32(s_mid + r * W, s_ids[c]);
        const bool rev =
            !(lb > s_thr[c]) && !in_sorted32(s_mid + c * W, s_ids[r]);
        if (fwd || rev) {
          const int at = atomicAdd(&s_nsv, 1);
          s_sv[at] = (uint16_t)((r << 7) | c | (fwd ? SV_FWD : 0) |
                                (rev ? SV_REV : 0));
        }
      }
    }
    __syncthreads();
    const int nsv = s_nsv;
    c_surv += (unsigned long long)nsv;

    // ---- 5. canonical distances of the survivors and their offers
    for (int base = 0; base < nsv; base += JRR) {
      const int cnt = min(JRR, nsv - base);
      const int jn = (cnt + 255) >> 8; // survivors per thread, uniform
      float dacc[8];
      int pr[8], pm[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        dacc[j] = 0.0f;
        const int li = tid + 256 * j;
        const int e = li < cnt ? (int)s_sv[base + li] : 0;
        pr[j] = (e >> 7) & 63;
        pm[j] = e & 127;
      }
      float4 fa[4], fb[4];
      auto fload = [&](int ch, float4 (&r)[4]) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int seg = tid + 256 * j, m = seg >> 3, pc = seg & 7;
          const uint64_t f = (uint64_t)ch * JKC + pc * 4;
          r[j] = (m < m_all && f < dp)
                     ? *reinterpret_cast<const float4 *>(
                           X + (uint64_t)s_ids[m] * dp + f)
                     : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      };
      auto fstore = [&](int buf, const float4 (&r)[4]) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int seg = tid + 256 * j, m = seg >> 3, pc = seg & 7;
          *reinterpret_cast<float4 *>(s_ring + buf * JBUF + m * JRS +
                                      pc * 16) = r[j];
        }
      };
      // one chunk of every chain: features in order, one FMA each (zero
      // features past dp leave a chain unchanged)
      auto fstep = [&](int buf) {
        const unsigned char *fbuf = s_ring + buf * JBUF;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if (j >= jn || tid + 256 * j >= cnt)
            continue;
          const float4 *xr =
              reinterpret_cast<const float4 *>(fbuf + pr[j] * JRS);
          const float4 *xc =
              reinterpret_cast<const float4 *>(fbuf + pm[j] * JRS);
          float d = dacc[j];
#pragma unroll
          for (int f4 = 0; f4 < JKC / 4; ++f4) {
            const float4 p = xr[f4], q = xc[f4];
            float t = p.x - q.x;
            d = __builtin_fmaf(t, t, d);
            t = p.y - q.y;
            d = __builtin_fmaf(t, t, d);
            t = p.z - q.z;
            d = __builtin_fmaf(t, t, d);
            t = p.w - q.w;
            d = __builtin_fmaf(t, t, d);
          }
          dacc[j] = d;
        }
      };
      fload(0, fa);
      if (nchf > 1)
        fload(1, fb);
      fstore(0, fa);
      __syncthreads();
      for (int ch = 0; ch < nchf; ch += 2) {
        if (ch + 2 < nchf)
          fload(ch + 2, fa);
        fstep(0);
        if (ch + 1 < nchf)
          fstore(1, fb);
        __syncthreads();
        if (ch + 1 >= nchf)
          break;
        if (ch + 3 < nchf)
          fload(ch + 3, fb);
        fstep(1);
        if (ch + 2 < nchf)
          fstore(0, fa);
        __syncthreads();
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (tid + 256 * j < cnt)
          s_svd[tid + 256 * j] = dacc[j];
      __syncthreads();
      for (int j = 0; j < jn; ++j) {
        if (tid == 0)
          s_qn = 0;
        __syncthreads();
        const int li = tid + 256 * j;
        if (li < cnt) {
          const int e = s_sv[base + li];
          const int r = (e >> 7) & 63, c = e & 127;
          const float d = s_svd[li];
          const uint32_t idr = s_ids[r], idc = s_ids[c];
          // r <- c
          uint64_t key = make_key(d, idc, true);
          if ((e & SV_FWD) && key < s_mmax[r]) {
            const int at = atomicAdd(&s_qn, 1);
            s_qkey[at] = key;
            s_qtgt[at] = (uint8_t)r;
          }
          // c <- r
          key = make_key(d, idr, true);
          if ((e & SV_REV) && key < s_mmax[c]) {
            const int at = atomicAdd(&s_qn, 1);
            s_qkey[at] = key;
            s_qtgt[at] = (uint8_t)c;
          }
        }
        __syncthreads();
        const int qn = s_qn;
        c_offers += (unsigned long long)qn;
        for (int q = hw; q < qn; q += 8) {
          const uint32_t tgt = s_ids[s_qtgt[q]];
          const int res = pool_offer_half(pool + (uint64_t)tgt * W, s_qkey[q],
                                          k, hl, half);
          if (res == 2 && hl == 0)
            flag[tgt] = 1;
        }
        __syncthreads();
      }
    }
  }
  if (tid == 0 && stats && c_members) {
    atomicAdd(stats + 0, c_members);
    atomicAdd(stats + 1, c_pairs);
    atomicAdd(stats + 2, c_offers);
    atomicAdd(stats + 7, c_surv);
  }
}

} // namespace gm
