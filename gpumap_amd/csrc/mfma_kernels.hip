// Exact insertion screened by BF16 MFMA distance tiles.
//
// Same result as the FP32 tiles of exact_kernels.hip -- every inserted row
// gets the exact 32 nearest rows by the canonical FP32 distance, every earlier
// row is offered every new row that beats its pool maximum -- with the
// distance work moved to the matrix cores:
//   * each row x is stored as two BF16 vectors hi = bf16(x), lo = bf16(x - hi)
//     (x - hi - lo is below 2^-16 |x|) plus its FP32 squared norm, in an
//     MFMA-ready layout (below);
//   * a 32 x 32 block of approximate squared distances
//       d~ = |u|^2 + |j|^2 - 2 (hi_j.hi_u + hi_j.lo_u + lo_j.hi_u)
//     takes three v_mfma_f32_32x32x16_bf16 per 16 features (FP32
//     accumulation);
//   * d~ - coef (|u|^2 + |j|^2) <= d / (1 - cfac) and d / (1 + cfac) <=
//     d~ + coef (|u|^2 + |j|^2), where d is the canonical distance: coef
//     covers the BF16 split, the summation of each group of
//     BF16_GROUP_STEPS feature steps' products on the matrix core and the
//     round-to-nearest adds across groups, the norm
//     and d~ roundings; cfac covers the canonical distance's own rounding
//     (derived on the host, engine.hip);
//   * pass 1 keeps, per lane and batch row, the TOPL smallest upper bounds
//     d~ + bound it sees (dropping larger ones can only raise what follows);
//     the 32nd smallest of all of them, tau, is an upper bound of the row's
//     32nd smallest exact distance (k_tau_select); it also stores every
//     pair's lower bound d~ - bound;
//   * the survivor scan (k_bf16_lbscan, or a second pass over the tiles when
//     the store does not fit) lists, per batch row, every column whose lower
//     bound is <= tau (forward) or <= the column row's pool maximum
//     (reverse); k_bf16_rerank computes the listed pairs'
//     canonical distances, writes each batch row's 32 smallest forward keys to
//     its pool and offers the reverse ones with the lock-free bounded insert,
//     whose result does not depend on the order.  No pair that can enter a
//     pool is screened out, so the pools equal the FP32 path's bit for bit;
//   * a full list sets an overflow flag; the FP32 tiles then redo the
//     insertion (their kernels take the flag as a guard).
//
// Operand layout.  Rows are grouped by 32; for group g and feature step s (16
// features) the 32 x 16 hi block and the 32 x 16 lo block are stored as
// 2 x 1 KB in MFMA lane order -- lane l holds row l % 32, features
// (l / 32) * 8 .. +7 -- so one 16-byte load per lane (1 KB per wave, fully
// coalesced) fetches a whole operand of v_mfma_f32_32x32x16_bf16.
#include "common.hpp"
#include "kernels.hpp"

namespace gm {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

using screen::LFWD;
using screen::LID;
using screen::LREV;
using screen::PARTW;
using screen::TB;
using screen::TC;
using screen::TOPL;
// a buffer offset past every buffer's end: the access is dropped
constexpr uint32_t BUF_OFF = 0x80000000u;
constexpr int RC = 32;                 // features per re-rank chunk
constexpr int RS = RC + 1;             // padded LDS row of a chunk


// element offset of (row, feature) in the MFMA-ready layout
__device__ inline uint64_t swz(uint64_t row, uint64_t f, uint64_t nsteps,
                               int part) {
  return (((row >> 5) * nsteps + (f >> 4)) * 2 + part) * 512 +
         ((((f & 15) >> 3) << 5) + (row & 31)) * 8 + (f & 7);
}

__device__ inline bf16x8 ld8(const __bf16 *p) {
  return *reinterpret_cast<const bf16x8 *>(p);
}

// stages of the screening kernel's operand ring (16 KB each)
constexpr int NSTG = 4;
// s_waitcnt for the ring: vector memory count (NSTG - 2) * 4 (a wave's loads
// of the NSTG - 2 stages issued after the awaited one), no other wait
constexpr int RING_WAIT = ((NSTG - 2) * 4) | (7 << 4) | (15 << 8);

// A step's eight operands of one wave from the ring: four blocks from the
// column half at a, four from the batch-row half at b (LDS byte addresses of
// this lane's 16 bytes in the first block).  Written as one asm block that
// waits for its own reads: the caller has already waited for the stage's
// direct-to-LDS loads (its own, and every wave's at the barrier), which the
// compiler cannot tell apart from the later stages' loads still in flight
// and would otherwise wait for as well.
__device__ inline void ring_read(uint32_t a, uint32_t b, bf16x8 &ah0,
                                 bf16x8 &al0, bf16x8 &ah1, bf16x8 &al1,
                                 bf16x8 &bh0, bf16x8 &bl0, bf16x8 &bh1,
                                 bf16x8 &bl1) {
  asm volatile("ds_read_b128 %0, %8\n\t"
               "ds_read_b128 %1, %8 offset:1024\n\t"
               "ds_read_b128 %2, %8 offset:2048\n\t"
               "ds_read_b128 %3, %8 offset:3072\n\t"
               "ds_read_b128 %4, %9\n\t"
               "ds_read_b128 %5, %9 offset:1024\n\t"
               "ds_read_b128 %6, %9 offset:2048\n\t"
               "ds_read_b128 %7, %9 offset:3072\n\t"
               "s_waitcnt lgkmcnt(0)"
               : "=&v"(ah0), "=&v"(al0), "=&v"(ah1), "=&v"(al1), "=&v"(bh0),
                 "=&v"(bl0), "=&v"(bh1), "=&v"(bl1)
               : "v"(a), "v"(b)
               : "memory");
}

// block-wide barrier for LDS data only: completes this wave's LDS accesses,
// not its global loads (__syncthreads would also wait for the operand ring's
// loads in flight)
__device__ inline void lds_barrier() {
  __builtin_amdgcn_s_waitcnt((15 | (3 << 14)) | (7 << 4)); // lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
}

} // namespace

// hi/lo BF16 split of the batch rows into the column operand store Xs (global
// row index) and the batch operand store Xb (batch-local index), their FP32
// squared norms, a set flag and an empty survivor list per batch row, and a
// cleared overflow flag.  One 64-thread block per batch row.
__global__ __launch_bounds__(64) void k_bf16_prep(
    const float *__restrict__ X, uint64_t dp, uint64_t db,
    __bf16 *__restrict__ Xs, __bf16 *__restrict__ Xb, float *__restrict__ nrm,
    uint8_t *__restrict__ flag, unsigned *__restrict__ rcnt,
    unsigned *__restrict__ qctl, uint64_t b0, uint64_t nb) {
  const uint64_t r = blockIdx.x;
  if (r >= nb)
    return;
  const uint64_t u = b0 + r;
  const float *x = X + u * dp;
  const uint64_t nsteps = db / 16;
  const int lane = threadIdx.x;
  float s = 0.0f;
  for (uint64_t t = lane; t < db / 8; t += 64) {
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
    *reinterpret_cast<bf16x8 *>(Xs + swz(u, f0, nsteps, 0)) = hi;
    *reinterpret_cast<bf16x8 *>(Xs + swz(u, f0, nsteps, 1)) = lo;
    *reinterpret_cast<bf16x8 *>(Xb + swz(r, f0, nsteps, 0)) = hi;
    *reinterpret_cast<bf16x8 *>(Xb + swz(r, f0, nsteps, 1)) = lo;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    s += __shfl_xor(s, off, 64);
  if (lane == 0) {
    nrm[u] = s;
    flag[u] = 1;
    rcnt[r] = 0;
    rcnt[nb + r] = 0; // the row's forward keys within tau (k_bf16_rerank)
  }
  if (r == 0 && lane == 1)
    qctl[1] = 0;
}

// One pass of the screening tiles.  Grid: rtiles * nsplit blocks of 256
// threads; block (rt, split) covers batch rows [128 rt, 128 rt + 128) and the
// column tiles split, split + nsplit, ...  Blocks of one split are placed on
// one XCD so that their column reads share its L2.  Wave w computes the
// 64 x 64 sub-tile (columns (w & 1) * 64, batch rows (w >> 1) * 64) as 2 x 2
// MFMA blocks; C register i of lane l is column (i / 4) * 8 + (l / 32) * 4 +
// i % 4 and batch row l % 32 of its block.
// Two blocks per CU (at most 256 registers per lane): the second wave on each
// SIMD runs its matrix work while the first waits for operands.
template <int PASS>
__global__ __launch_bounds__(256, 2) void k_bf16_screen(
    const __bf16 *__restrict__ Xs, const __bf16 *__restrict__ Xb, int nsteps,
    const float *__restrict__ nrm, const float *__restrict__ pmax,
    const float *__restrict__ tau, uint64_t b0, uint64_t nb, uint64_t n,
    int nsplit, int rtiles, float *__restrict__ part,
    uint32_t *__restrict__ rlist, unsigned *__restrict__ rcnt, unsigned rcap,
    unsigned *__restrict__ qctl, float coef, float cfac,
    float *__restrict__ lbm) {
  __shared__ float s_nj[TC], s_pm[TC];
  __shared__ __attribute__((aligned(16))) __bf16 ring[NSTG][16 * 512];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wa = wave & 1, wb = wave >> 1;
  int rt, split;
  {
    const int b = blockIdx.x;
    if ((nsplit & 7) == 0) {
      const int q = b >> 3;
      split = (q / rtiles) * 8 + (b & 7);
      rt = q % rtiles;
    } else {
      rt = b % rtiles;
      split = b / rtiles;
    }
  }
  const uint64_t r0 = (uint64_t)rt * TB;
  const uint64_t ntile = (n + TC - 1) / TC;

  // this lane's two batch rows
  float nu[2], tr[2];
  bool rok[2];
#pragma unroll
  for (int sb = 0; sb < 2; ++sb) {
    const uint64_t rl = r0 + wb * 64 + sb * 32 + (lane & 31);
    rok[sb] = rl < nb;
    nu[sb] = rok[sb] ? nrm[b0 + rl] : 0.0f;
    tr[sb] = (PASS == 2 && rok[sb]) ? tau[rl] : 0.0f;
  }
  // pass 1: the TOPL smallest upper bounds per batch row, ascending, as
  // bit patterns (the bounds are >= 0, so their bits order as their values)
  uint32_t top[2][TOPL];
#pragma unroll
  for (int sb = 0; sb < 2; ++sb)
#pragma unroll
    for (int q = 0; q < TOPL; ++q)
      top[sb][q] = 0x7f800000u; // +inf

  const uint64_t stepw = 1024; // elements per (group, step): hi + lo blocks
  const uint64_t gstride = (uint64_t)nsteps * stepw; // one group's steps
  // Operand ring in LDS: NSTG stages of one feature step each, 16 blocks of
  // 1 KB per stage (blocks 0-7: the tile's 4 column groups, hi then lo;
  // blocks 8-15: the 4 batch-row groups).  Wave w loads blocks 4w .. 4w + 3
  // of each stage with direct-to-LDS loads (16 bytes a lane, MFMA lane
  // order), NSTG - 1 steps ahead of the step on the matrix cores; the steps
  // run over this block's tiles in sequence, so the next tile's first steps
  // load during this tile's last.  Each operand block is fetched once per
  // workgroup (the waves that share it read it from LDS).
  const bool aw = wave < 2; // this wave loads column blocks, else batch rows
  const __bf16 *const bsrc = // waves 2, 3: batch-row groups 0-1, 2-3
      Xb + ((r0 / 32 + (uint64_t)(wave & 1) * 2) * gstride) + lane * 8;
  const uint32_t my_tiles =
      split < ntile ? (uint32_t)((ntile - 1 - (uint64_t)split) / nsplit + 1)
                    : 0u;
  uint32_t pti = 0, pst = 0; // the next step to load: tile index, step
  // issue the next step's 4 blocks of this wave into stage `stg` (past the
  // last step a repeat of it, so that every step issues the same loads)
#define GM_RING_ISSUE(stg)                                                      \
  {                                                                            \
    const bool live = pti < my_tiles;                                          \
    const uint32_t qt = live ? pti : my_tiles - 1;                             \
    const uint32_t qs = live ? pst : (uint32_t)nsteps - 1;                     \
    const __bf16 *src =                                                        \
        (aw ? Xs + (((split + (uint64_t)qt * nsplit) * TC / 32 +               \
                     (uint64_t)wave * 2) * gstride) + lane * 8                 \
            : bsrc) +                                                          \
        (uint64_t)qs * stepw;                                                  \
    __bf16 *dst = &ring[(stg)][wave * 4 * 512];                                \
    __builtin_amdgcn_global_load_lds(                                          \
        src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);         \
    __builtin_amdgcn_global_load_lds(                                          \
        src + 512, (__attribute__((address_space(3))) void *)(dst + 512), 16,  \
        0, 0);                                                                 \
    __builtin_amdgcn_global_load_lds(                                          \
        src + gstride, (__attribute__((address_space(3))) void *)(dst + 1024), \
        16, 0, 0);                                                             \
    __builtin_amdgcn_global_load_lds(                                          \
        src + gstride + 512,                                                   \
        (__attribute__((address_space(3))) void *)(dst + 1536), 16, 0, 0);     \
    if (++pst == (uint32_t)nsteps) {                                           \
      pst = 0;                                                                 \
      ++pti;                                                                   \
    }                                                                          \
  }
  // this lane's LDS byte address in stage 0 of the ring
  const uint32_t ring_base =
      (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const char *)(
          (const char *)&ring[0][0]) +
      (uint32_t)lane * 16u;
  uint32_t gs = 0; // the step on the matrix cores, over the block's tiles
  if (my_tiles > 0) {
#pragma unroll
    for (int q = 0; q < NSTG - 1; ++q)
      GM_RING_ISSUE(q)
  }

  for (uint32_t ti = 0; ti < my_tiles; ++ti) {
    const uint64_t c0 = (split + (uint64_t)ti * nsplit) * TC;
    lds_barrier(); // previous tile's readers of s_nj / s_pm are done
    if (tid < TC) {
      const uint64_t j = c0 + tid;
      s_nj[tid] = j < n ? nrm[j] : 0.0f;
      s_pm[tid] = (PASS == 2 && j < b0) ? pmax[j] : -INFINITY;
    }
    // acc: the running dot products; grp: the current group's products,
    // summed on the matrix core
    floatx16 acc[2][2], grp[2][2];
#pragma unroll
    for (int sa = 0; sa < 2; ++sa)
#pragma unroll
      for (int sb = 0; sb < 2; ++sb)
        acc[sa][sb] = floatx16{};
    const floatx16 z = {};
    // One feature step: wait until this step's stage has landed (this
    // wave's loads of it, then every wave's at the barrier, which also
    // frees the stage read by the previous step), load the step NSTG - 1
    // ahead into that stage, then this step's 48 products per pair into the
    // group's accumulators (C = 0 on the group's first step).  Every
    // BF16_GROUP_STEPS steps the group is added to the running dot product
    // in round-to-nearest FP32, so the summation error grows with the
    // group, not with D.
#define GM_SCREEN_STEP(FRESH)                                                  \
  {                                                                            \
    __builtin_amdgcn_s_waitcnt(RING_WAIT);                                     \
    __builtin_amdgcn_s_barrier();                                              \
    GM_RING_ISSUE((gs + NSTG - 1) % NSTG)                                      \
    bf16x8 ah0, al0, ah1, al1, bh0, bl0, bh1, bl1;                            \
    ring_read(ring_base + (uint32_t)(gs % NSTG) * 16384u + (uint32_t)wa * 4096u, \
              ring_base + (uint32_t)(gs % NSTG) * 16384u + 8192u +             \
                  (uint32_t)wb * 4096u,                                        \
              ah0, al0, ah1, al1, bh0, bl0, bh1, bl1);                         \
    grp[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(                       \
        ah0, bh0, FRESH ? z : grp[0][0], 0, 0, 0);                             \
    grp[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(                       \
        ah0, bh1, FRESH ? z : grp[0][1], 0, 0, 0);                             \
    grp[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(                       \
        ah1, bh0, FRESH ? z : grp[1][0], 0, 0, 0);                             \
    grp[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(                       \
        ah1, bh1, FRESH ? z : grp[1][1], 0, 0, 0);                             \
    grp[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0, bl0, grp[0][0],   \
                                                        0, 0, 0);              \
    grp[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0, bl1, grp[0][1],   \
                                                        0, 0, 0);              \
    grp[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1, bl0, grp[1][0],   \
                                                        0, 0, 0);              \
    grp[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1, bl1, grp[1][1],   \
                                                        0, 0, 0);              \
    grp[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al0, bh0, grp[0][0],   \
                                                        0, 0, 0);              \
    grp[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al0, bh1, grp[0][1],   \
                                                        0, 0, 0);              \
    grp[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al1, bh0, grp[1][0],   \
                                                        0, 0, 0);              \
    grp[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al1, bh1, grp[1][1],   \
                                                        0, 0, 0);              \
    ++gs;                                                                      \
  }
    for (int st = 0; st < nsteps;) {
      const int gend = min(st + BF16_GROUP_STEPS, nsteps);
      GM_SCREEN_STEP(true)
      for (++st; st < gend; ++st)
        GM_SCREEN_STEP(false)
#pragma unroll
      for (int sa = 0; sa < 2; ++sa)
#pragma unroll
        for (int sb = 0; sb < 2; ++sb)
          acc[sa][sb] += grp[sa][sb];
    }
#undef GM_SCREEN_STEP
    lds_barrier(); // s_nj / s_pm written

    // pass 1 stores every pair's lower bound, column-major, for
    // k_bf16_lbscan: through a buffer over this tile's columns, so that a
    // lane past the batch (offset BUF_OFF) or a column past n (beyond
    // num_records) is dropped by the hardware, without a branch per store
    // (num_records 0 when there is no store; the host passes lbm only for
    // batches whose 128-column region stays below 2^31 bytes)
    const uint64_t cols = min((uint64_t)TC, n - c0);
    const __amdgpu_buffer_rsrc_t lr = __builtin_amdgcn_make_buffer_rsrc(
        lbm ? lbm + c0 * nb : nullptr, 0,
        lbm ? (int)(cols * nb * sizeof(float)) : 0, 0x00020000);
    // a column's stride, taken as a new value per tile so that the 64
    // per-pair offsets below are formed here and not hoisted out of the tile
    // loop (where they would not fit in registers)
    uint32_t cstride = (uint32_t)nb * 4u;
    asm volatile("" : "+s"(cstride));
    uint32_t lbase[2];
#pragma unroll
    for (int sb = 0; sb < 2; ++sb)
      lbase[sb] = (rok[sb] ? (uint32_t)(r0 + wb * 64 + sb * 32 + (lane & 31)) * 4u
                           : BUF_OFF) +
                  (uint32_t)(wa * 64 + (lane >> 5) * 4) * cstride;
#pragma unroll
    for (int sa = 0; sa < 2; ++sa)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int jl = wa * 64 + sa * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        const uint64_t j = c0 + jl;
        const float nj = s_nj[jl];
        const float pm = s_pm[jl];
#pragma unroll
        for (int sb = 0; sb < 2; ++sb) {
          const uint64_t rl = r0 + wb * 64 + sb * 32 + (lane & 31);
          const bool ok = rok[sb] && j < n && j != b0 + rl;
          const float s = nu[sb] + nj;
          const float dt = __builtin_fmaf(-2.0f, acc[sa][sb][i], s);
          const float bnd = __builtin_fmaf(coef, s, 1e-30f);
          if (PASS == 1) {
            __builtin_amdgcn_raw_buffer_store_b32(
                __float_as_uint((dt - bnd) * (1.0f - cfac)), lr,
                (int)(lbase[sb] +
                      (uint32_t)(sa * 32 + (i >> 2) * 8 + (i & 3)) * cstride),
                0, 0);
            const float ub = fmaxf((dt + bnd) * (1.0f + cfac), 0.0f);
            // branch-free insertion: the value is carried down the sorted
            // list, each slot keeps the smaller, the largest drops out; a
            // pair that may not enter carries ~0u (above every bound)
            uint32_t x = ok ? __float_as_uint(ub) & 0x7fffffffu : 0xffffffffu;
#pragma unroll
            for (int q = 0; q < TOPL; ++q) {
              const uint32_t lo = min(top[sb][q], x);
              x = max(top[sb][q], x);
              top[sb][q] = lo;
            }
          } else {
            const float lb = (dt - bnd) * (1.0f - cfac);
            const bool fwd = lb <= tr[sb];
            const bool rev = lb <= pm; // pm = -inf unless j < b0
            if (ok && (fwd || rev)) {
              const unsigned at = atomicAdd(&rcnt[rl], 1u);
              if (at < rcap)
                rlist[rl * rcap + at] = (uint32_t)j | (fwd ? LFWD : 0u) |
                                        (rev ? LREV : 0u);
              else
                atomicOr(&qctl[1], 1u);
            }
          }
        }
      }
  }

#undef GM_RING_ISSUE

  if (PASS == 1) {
    // PARTW values per (batch row, split): 4 lanes x TOPL, each list sorted
#pragma unroll
    for (int sb = 0; sb < 2; ++sb) {
      const uint64_t rl = r0 + wb * 64 + sb * 32 + (lane & 31);
      if (!rok[sb])
        continue;
      float *dst = part + (rl * (uint64_t)nsplit + split) * PARTW +
                   (wa * 2 + (lane >> 5)) * TOPL;
#pragma unroll
      for (int q = 0; q < TOPL; ++q)
        dst[q] = __uint_as_float(top[sb][q]);
    }
  }
}

template __global__ void k_bf16_screen<1>(
    const __bf16 *, const __bf16 *, int, const float *, const float *,
    const float *, uint64_t, uint64_t, uint64_t, int, int, float *,
    uint32_t *, unsigned *, unsigned, unsigned *, float, float, float *);
template __global__ void k_bf16_screen<2>(
    const __bf16 *, const __bf16 *, int, const float *, const float *,
    const float *, uint64_t, uint64_t, uint64_t, int, int, float *,
    uint32_t *, unsigned *, unsigned, unsigned *, float, float, float *);

// The survivor lists from the lower bounds stored by the first pass (in
// place of a second pass over the tiles): the pair (batch row rl, column j)
// is listed as forward when lb <= tau[rl] and as reverse when lb <=
// pmax[j] for an earlier row j, exactly as k_bf16_screen<2> lists it (the
// order within a list differs, which nothing downstream depends on).  A
// wave takes 64 batch rows (one per lane, tau in a register) and a range of
// columns, eight columns per step: each step's eight loads of
// lbm[j * nb + rl] are 256-byte rows, in flight together.
__global__ __launch_bounds__(256) void k_bf16_lbscan(
    const float *__restrict__ lbm, uint64_t nb, uint64_t n, uint64_t b0,
    const float *__restrict__ tau, const float *__restrict__ pmax,
    uint32_t *__restrict__ rlist, unsigned *__restrict__ rcnt, unsigned rcap,
    unsigned *__restrict__ qctl, uint64_t cols_per_wave) {
  const int lane = threadIdx.x & 63;
  const uint64_t wid = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint64_t ngroups = (nb + 63) / 64;
  const uint64_t g = wid % ngroups;
  const uint64_t j0 = (wid / ngroups) * cols_per_wave;
  if (j0 >= n)
    return;
  const uint64_t j1 = min(n, j0 + cols_per_wave);
  const uint64_t rl = g * 64 + lane;
  const bool rok = rl < nb;
  const float tr = rok ? tau[rl] : -INFINITY;
  const uint64_t u = b0 + rl;
  // a lane's entries are buffered and appended four at a time: one atomic
  // (and its round trip) per four entries instead of one per entry
  uint32_t buf0 = 0, buf1 = 0, buf2 = 0, buf3 = 0;
  unsigned nbuf = 0;
  auto flush = [&]() {
    const unsigned at = atomicAdd(&rcnt[rl], nbuf);
    const uint32_t e4[4] = {buf0, buf1, buf2, buf3};
#pragma unroll
    for (unsigned q = 0; q < 4; ++q)
      if (q < nbuf) {
        if (at + q < rcap)
          rlist[rl * rcap + at + q] = e4[q];
        else
          atomicOr(&qctl[1], 1u);
      }
    nbuf = 0;
  };
  auto emit = [&](uint64_t j, float lb, float pm) {
    const bool fwd = lb <= tr;
    const bool rev = lb <= pm;
    if (rok && j != u && (fwd || rev)) {
      const uint32_t ent = (uint32_t)j | (fwd ? LFWD : 0u) | (rev ? LREV : 0u);
      buf3 = nbuf == 3 ? ent : buf3;
      buf2 = nbuf == 2 ? ent : buf2;
      buf1 = nbuf == 1 ? ent : buf1;
      buf0 = nbuf == 0 ? ent : buf0;
      if (++nbuf == 4)
        flush();
    }
  };
  const uint64_t rc = rok ? rl : 0; // an in-bounds address for idle lanes
  uint64_t j = j0;
  for (; j + 8 <= j1; j += 8) {
    float lb[8], pm[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      lb[q] = lbm[(j + q) * nb + rc];
      pm[q] = j + q < b0 ? pmax[j + q] : -INFINITY;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
      emit(j + q, lb[q], pm[q]);
  }
  for (; j < j1; ++j)
    emit(j, lbm[j * nb + rc], j < b0 ? pmax[j] : -INFINITY);
  if (nbuf)
    flush();
}

// tau of one batch row: the 32nd smallest of its nsplit * PARTW upper bounds
// (+inf when it has fewer than 32 other rows), found by bisection over the
// float bit patterns (all values are >= 0) with block-wide counts; the
// values stay in registers and no bin is contended.  A wave counts with
// ballots (one per register, popcounts on the scalar unit), and the four
// waves' counts meet in LDS behind one barrier per step (two alternating
// count slots: a wave writes a slot again only after the next step's
// barrier, which every wave reaches after reading it).
__global__ __launch_bounds__(256) void k_tau_select(
    const float *__restrict__ part, int nsplit, float *__restrict__ tau,
    uint64_t nb) {
  __shared__ unsigned s_cnt[2][4];
  const uint64_t r = blockIdx.x;
  if (r >= nb)
    return;
  constexpr int PER = 128 * PARTW / 256; // values per thread for nsplit 128
  const int len = nsplit * PARTW;
  const unsigned *v =
      reinterpret_cast<const unsigned *>(part) + r * (uint64_t)len;
  unsigned x[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = (int)threadIdx.x + 256 * i;
    x[i] = e < len ? v[e] : 0xffffffffu;
  }
  // the smallest t with #{x <= t} >= W, by bisection over the 32-bit
  // patterns (the values are compared as unsigned patterns, which orders
  // the non-negative floats); counts are exact, so t is the W-th smallest
  unsigned lo = 0u, hi = 0xffffffffu;
  for (int it = 0; it < 32 && lo < hi; ++it) {
    const unsigned mid = lo + ((hi - lo) >> 1);
    unsigned c = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i)
      c += (unsigned)__popcll(__ballot(x[i] <= mid));
    unsigned *sc = s_cnt[it & 1];
    if ((threadIdx.x & 63) == 0)
      sc[threadIdx.x >> 6] = c;
    __syncthreads();
    const unsigned tot = sc[0] + sc[1] + sc[2] + sc[3];
    if (tot >= (unsigned)W)
      hi = mid;
    else
      lo = mid + 1u;
  }
  if (threadIdx.x == 0)
    tau[r] = __uint_as_float(lo);
}

// Canonical distances of the listed pairs and the reverse offers.  Block
// (r, g) is one wave and takes entries [64 g, 64 g + 64) of batch row
// u = b0 + r's list, then every (gridDim.y)-th group after it (blocks past
// the list end return at once), so the work spreads evenly however uneven
// the lists are.  The wave walks the
// features in chunks of 32: the 64 column rows' chunk is staged in LDS from
// coalesced 128-byte row reads (the next chunk's reads are in flight
// meanwhile), then every lane continues its own pair's FMA chain over the
// chunk in feature order (the canonical distance; u's features are
// wave-uniform loads).  Forward keys within the row's tau are appended to
// the row's key list in fkeys for k_bf16_select; every lane offers its
// reverse pair, when within the column row's pool maximum, with the
// single-lane lock-free insert.  Does nothing when a list overflowed.
__global__ __launch_bounds__(64) void k_bf16_rerank(
    const float *__restrict__ X, uint64_t dp,
    const uint32_t *__restrict__ rlist, unsigned *__restrict__ rcnt,
    unsigned rcap, const unsigned *__restrict__ qctl,
    uint64_t *__restrict__ fkeys, uint64_t *__restrict__ pool,
    const float *__restrict__ pmax, uint8_t *__restrict__ flag, uint64_t b0,
    const float *__restrict__ tau, uint64_t nb,
    unsigned long long *__restrict__ stats) {
  __shared__ float sl[64 * RS];
  const uint64_t r = blockIdx.x;
  const unsigned cnt = min(rcnt[r], rcap);
  if (qctl[1] || blockIdx.y * 64u >= cnt)
    return;
  const int lane = threadIdx.x;
  if (stats && blockIdx.y == 0 && lane == 0)
    atomicAdd(&stats[5], (unsigned long long)cnt);
  const uint64_t u = b0 + r;
  const float *__restrict__ xu = X + u * dp;
  const int c4 = (lane & 7) * 4; // this lane's float4 of a staged row chunk
  // groups of 64 entries, grid-strided over the block rows of the grid's y
  for (unsigned b = blockIdx.y * 64u; b < cnt; b += gridDim.y * 64u) {
    const bool valid = b + lane < cnt;
    const uint32_t ent = valid ? rlist[r * (uint64_t)rcap + b + lane] : 0u;
    const uint64_t j = ent & LID;
    // staged rows lane / 8 + 8 q: loads are branch-free (a row past the list
    // reads row 0 and a float4 past dp re-reads the row's last one; both are
    // then replaced by zeros)
    uint32_t jr[8];
    uint32_t live = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int row = (lane >> 3) + 8 * q;
      const uint32_t jq = (uint32_t)__shfl((uint32_t)j, row, 64);
      const bool ok = b + row < cnt;
      jr[q] = ok ? jq : 0u;
      live |= ok ? 1u << q : 0u;
    }
    // (the loads are issued here and zeroed only when they are written to
    // LDS, two chunks later: zeroing them here would wait for them at once)
    auto stage = [&](uint64_t f, float4 out[8]) {
      const bool fin = f + c4 < dp;
      const uint64_t col = fin ? f + c4 : dp - 4;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        out[q] =
            *reinterpret_cast<const float4 *>(X + (uint64_t)jr[q] * dp + col);
    };
    float acc = 0.0f;
    // One chunk of RC features: its staged rows (nv, loaded two chunks
    // earlier) into LDS, the loads of the chunk two ahead into nv, then this
    // chunk's FMA chain.  Two register buffers alternate, so that two
    // chunks' row loads are in flight during each chain; the chain runs over
    // the features in order, one FMA each (the canonical distance).
    auto chunk = [&](uint64_t f0, float4 nv[8]) {
      // zeros for a row past the list and a float4 past dp (the loads
      // re-read row 0 and the row's last float4; a bit mask rather than a
      // select, which would let the compiler move the loads under a branch)
      const bool fin = f0 + c4 < dp;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint32_t m = (fin && ((live >> q) & 1u)) ? ~0u : 0u;
        float *d = sl + ((lane >> 3) + 8 * q) * RS + c4;
        d[0] = __uint_as_float(__float_as_uint(nv[q].x) & m);
        d[1] = __uint_as_float(__float_as_uint(nv[q].y) & m);
        d[2] = __uint_as_float(__float_as_uint(nv[q].z) & m);
        d[3] = __uint_as_float(__float_as_uint(nv[q].w) & m);
      }
      __syncthreads();
      stage(f0 + 2 * RC, nv);
      __builtin_amdgcn_sched_barrier(0);
      // u's chunk (wave-uniform); features past dp are zero on both sides and
      // fma(0, 0, acc) == acc
      float xv[RC];
      if (f0 + RC <= dp) {
#pragma unroll
        for (int i = 0; i < RC; ++i)
          xv[i] = xu[f0 + i];
      } else {
#pragma unroll
        for (int i = 0; i < RC; ++i)
          xv[i] = f0 + i < dp ? xu[f0 + i] : 0.0f;
      }
#pragma unroll
      for (int i = 0; i < RC; ++i) {
        const float t = xv[i] - sl[lane * RS + i];
        acc = __builtin_fmaf(t, t, acc);
      }
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();
    };
    float4 nva[8], nvb[8];
    stage(0, nva);
    stage(RC, nvb);
    for (uint64_t f0 = 0; f0 < dp; f0 += 2 * RC) {
      chunk(f0, nva);
      if (f0 + RC < dp)
        chunk(f0 + RC, nvb);
    }
    // forward keys within tau, appended to the row's key list (rcnt[nb + r]
    // counts them): the row's 32 nearest are among them, since the 32 columns
    // with the smallest upper bounds all have exact distances <= tau
    const bool fw = valid && (ent & LFWD) && acc <= tau[r];
    const uint64_t fm = __ballot(fw);
    unsigned base = 0;
    if (lane == 0 && fm)
      base = atomicAdd(&rcnt[nb + r], (unsigned)__popcll(fm));
    base = (unsigned)__shfl((int)base, 0, 64);
    if (fw)
      fkeys[r * (uint64_t)rcap + base +
            __popcll(fm & ((1ull << lane) - 1ull))] =
          make_key(acc, (uint32_t)j, false);
    if (valid && (ent & LREV) && acc <= pmax[j] &&
        pool_offer_lane(pool + j * W, make_key(acc, (uint32_t)u, false)))
      flag[j] = 1;
  }
}

// Batch row u = b0 + r's pool: the 32 smallest of its forward keys (those
// within tau, fcnt[r] of them), one wave per row.  Up to 64 keys are sorted in registers (bitonic network over
// the lanes); longer lists are sorted in LDS (bitonic over the list length
// rounded up to a power of two).
__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
  return ((uint64_t)hi << 32) | lo;
}

__global__ __launch_bounds__(64) void k_bf16_select(
    const uint64_t *__restrict__ fkeys, const unsigned *__restrict__ fcnt,
    unsigned rcap, const unsigned *__restrict__ qctl,
    uint64_t *__restrict__ pool, uint64_t b0,
    unsigned long long *__restrict__ stats) {
  extern __shared__ uint64_t s_keys[]; // rcap keys
  const uint64_t r = blockIdx.x;
  const int lane = threadIdx.x;
  if (qctl[1]) {
    if (stats && r == 0 && lane == 0)
      atomicAdd(&stats[6], 1ull);
    return;
  }
  const unsigned cnt = min(fcnt[r], rcap);
  const uint64_t *src = fkeys + r * (uint64_t)rcap;
  if (cnt <= 64) {
    uint64_t key = (unsigned)lane < cnt ? src[lane] : EMPTY;
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1)
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        const uint64_t o = shfl_xor_u64(key, stride);
        // ascending runs where lane & size == 0; the lower lane of a pair
        // keeps the smaller key in an ascending run
        const bool keep_min = ((lane & stride) == 0) == ((lane & size) == 0);
        key = keep_min ? (o < key ? o : key) : (o > key ? o : key);
      }
    if (lane < W)
      pool[(b0 + r) * W + lane] = key;
    return;
  }
  int len = 128;
  while ((unsigned)len < cnt)
    len <<= 1;
  for (int e = lane; e < len; e += 64)
    s_keys[e] = (unsigned)e < cnt ? src[e] : EMPTY;
  __syncthreads();
  for (int size = 2; size <= len; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = lane; t < len / 2; t += 64) {
        const int i = 2 * stride * (t / stride) + (t % stride);
        const int jj = i + stride;
        const bool up = (i & size) == 0;
        const uint64_t x = s_keys[i], y = s_keys[jj];
        if ((x > y) == up) {
          s_keys[i] = y;
          s_keys[jj] = x;
        }
      }
      __syncthreads();
    }
  if (lane < W)
    pool[(b0 + r) * W + lane] = s_keys[lane];
}

} // namespace gm
