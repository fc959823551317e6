// Exact insertion screened by FP8 (OCP E4M3) MFMA distance tiles with
// per-row scales.
//
// The same screen-and-rerank scheme as the BF16 tiles of mfma_kernels.hip
// (pass 1: upper bounds, tau, lower bounds; the survivor scan and the exact
// FP32 rerank are shared), with the operands in FP8:
//   * row x gets the power-of-two scale s = 2^floor(log2(448 / max|x_f|))
//     (1 for a zero row), so that a = s x is exact and max|a| lies in
//     (224, 448] (a row below 2^-117 everywhere, whose 1 / s would not be
//     a normal float, is not screened: all its pairs go to the rerank);
//     a is stored as two E4M3 vectors hi = e4m3(a),
//     lo = e4m3(a - hi) (a - hi is exact in FP32), with 1 / s and the FP32
//     squared norm of x;
//   * the dot product of a batch row u and a column row j is
//       P = (hi_u.hi_j + hi_u.lo_j + lo_u.hi_j) / (s_u s_j),
//     three v_mfma_scale_f32_32x32x64_f8f6f4 per 64 features (unit block
//     scales, FP32 accumulation; every E4M3 product is exact in FP32), and
//     d~ = |u|^2 + |j|^2 - 2 P;
//   * |d~ - d*| <= coef (|u|^2 + |j|^2), with coef derived on the host
//     (engine.hip) from the E4M3 rounding (2^-4 relative, 2^-10 absolute
//     below 2^-6), the three-term split's residual, the summation on the
//     matrix core and the norm roundings.  The bound is about 300 times the
//     BF16 split's, so more pairs survive to the exact rerank; the pools are
//     the FP32 path's bit for bit all the same.
//
// Operand layout.  Rows are grouped by 32; for group g and feature step t
// (64 features) the 32 x 64 hi block and the 32 x 64 lo block are stored as
// 2 x 2 KB in MFMA lane order -- lane l holds row l % 32, features
// (l / 32) * 32 .. +31, 32 bytes -- so two 16-byte loads per lane fetch an
// operand of v_mfma_scale_f32_32x32x64_f8f6f4.  The order of the features
// within an operand does not change a dot product as long as both operands
// use the same one, which they do (one layout for columns and batch rows).
#include "common.hpp"
#include "kernels.hpp"

namespace gm {

namespace {

typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;
typedef __attribute__((__vector_size__(8 * sizeof(int)))) int intx8;
typedef __attribute__((__vector_size__(4 * sizeof(int)))) int intx4;

using screen::LFWD;
using screen::LREV;
using screen::PARTW;
using screen::TB;
using screen::TC;
using screen::TOPL;

// a buffer offset past every buffer's end: the access is dropped
constexpr uint32_t BUF_OFF = 0x80000000u;

// byte offset of (row, feature, part) in the MFMA-ready layout
__device__ inline uint64_t swz8(uint64_t row, uint64_t f, uint64_t nsteps,
                                int part) {
  return (((row >> 5) * nsteps + (f >> 6)) * 2 + part) * 2048 +
         ((((f & 63) >> 5) << 5) + (row & 31)) * 32 + (f & 31);
}

// operand ring: NSTG stages of one feature step (16 blocks of 2 KB)
constexpr int NSTG = 2;
constexpr int STAGE = 16 * 2048;

// A step's eight operands of one wave from the ring (two 16-byte reads per
// operand), waiting for its own reads; see ring_read in mfma_kernels.hip.
__device__ inline void ring_read8(uint32_t a, uint32_t b, intx8 &ah0,
                                  intx8 &al0, intx8 &ah1, intx8 &al1,
                                  intx8 &bh0, intx8 &bl0, intx8 &bh1,
                                  intx8 &bl1) {
  intx4 r[16];
  asm volatile("ds_read_b128 %0, %16\n\t"
               "ds_read_b128 %1, %16 offset:16\n\t"
               "ds_read_b128 %2, %16 offset:2048\n\t"
               "ds_read_b128 %3, %16 offset:2064\n\t"
               "ds_read_b128 %4, %16 offset:4096\n\t"
               "ds_read_b128 %5, %16 offset:4112\n\t"
               "ds_read_b128 %6, %16 offset:6144\n\t"
               "ds_read_b128 %7, %16 offset:6160\n\t"
               "ds_read_b128 %8, %17\n\t"
               "ds_read_b128 %9, %17 offset:16\n\t"
               "ds_read_b128 %10, %17 offset:2048\n\t"
               "ds_read_b128 %11, %17 offset:2064\n\t"
               "ds_read_b128 %12, %17 offset:4096\n\t"
               "ds_read_b128 %13, %17 offset:4112\n\t"
               "ds_read_b128 %14, %17 offset:6144\n\t"
               "ds_read_b128 %15, %17 offset:6160\n\t"
               "s_waitcnt lgkmcnt(0)"
               : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]),
                 "=&v"(r[4]), "=&v"(r[5]), "=&v"(r[6]), "=&v"(r[7]),
                 "=&v"(r[8]), "=&v"(r[9]), "=&v"(r[10]), "=&v"(r[11]),
                 "=&v"(r[12]), "=&v"(r[13]), "=&v"(r[14]), "=&v"(r[15])
               : "v"(a), "v"(b)
               : "memory");
  auto cat = [](intx4 x, intx4 y) {
    return intx8{x[0], x[1], x[2], x[3], y[0], y[1], y[2], y[3]};
  };
  ah0 = cat(r[0], r[1]);
  al0 = cat(r[2], r[3]);
  ah1 = cat(r[4], r[5]);
  al1 = cat(r[6], r[7]);
  bh0 = cat(r[8], r[9]);
  bl0 = cat(r[10], r[11]);
  bh1 = cat(r[12], r[13]);
  bl1 = cat(r[14], r[15]);
}

__device__ inline void lds_barrier8() {
  __builtin_amdgcn_s_waitcnt((15 | (3 << 14)) | (7 << 4)); // lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
}

__device__ inline floatx16 mfma8(intx8 a, intx8 b, floatx16 c) {
  // FP8 E4M3 operands (format 0), unit E8M0 block scales (127)
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, 127,
                                                         0, 127);
}

} // namespace

// Per-row scale and hi/lo E4M3 split of the batch rows into the column
// operand store Xs (global row index) and the batch operand store Xb
// (batch-local index), 1 / s, the FP32 squared norms (the same canonical sum
// as k_bf16_prep), a set flag and an empty survivor list per batch row, and
// a cleared overflow flag.  One 64-thread block per batch row.
__global__ __launch_bounds__(64) void k_fp8_prep(
    const float *__restrict__ X, uint64_t dp, uint64_t db,
    uint8_t *__restrict__ Xs, uint8_t *__restrict__ Xb,
    float *__restrict__ nrm, float *__restrict__ isc,
    uint8_t *__restrict__ flag, unsigned *__restrict__ rcnt,
    unsigned *__restrict__ qctl, uint64_t b0, uint64_t nb) {
  const uint64_t r = blockIdx.x;
  if (r >= nb)
    return;
  const uint64_t u = b0 + r;
  const float *x = X + u * dp;
  const uint64_t nsteps = db / 64;
  const int lane = threadIdx.x;
  // the row's largest magnitude, then its scale 2^floor(log2(448 / max))
  float mx = 0.0f;
  for (uint64_t f = lane; f < dp; f += 64)
    mx = fmaxf(mx, fabsf(x[f]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  int ex = 0;
  bool tiny = false;
  if (mx > 0.0f) {
    // largest e with mx * 2^e <= 448: frexp gives mx = m 2^q, m in [0.5, 1)
    int q = 0;
    (void)frexpf(mx, &q);
    ex = 9 - q; // mx 2^(9 - q) < 512: one step down when above 448
    if (ldexpf(mx, ex) > 448.0f)
      --ex;
    // 1 / s must stay a normal float: a row below 2^-117 everywhere is
    // left unscreened (its pairs always go to the exact rerank)
    if (ex > 126) {
      ex = 126;
      tiny = true;
    }
  }
  // the norm as k_bf16_prep forms it: each lane's terms of its 8-feature
  // chunks in sequence, then the butterfly
  float s = 0.0f;
  for (uint64_t t = lane; t < db / 8; t += 64) {
    const uint64_t f0 = t * 8;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float v = f0 + e < dp ? x[f0 + e] : 0.0f;
      s = __builtin_fmaf(v, v, s);
    }
  }
  // hi/lo bytes, 32 consecutive features per lane and pass
  for (uint64_t t = lane; t < db / 32; t += 64) {
    const uint64_t f0 = t * 32;
    uint32_t hw[8], lw[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      uint32_t h4 = 0, l4 = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint64_t f = f0 + w * 4 + e;
        const float a = f < dp ? ldexpf(x[f], ex) : 0.0f;
        const uint8_t h = e4m3_rn(a);
        const uint8_t l = e4m3_rn(a - e4m3_dec(h));
        h4 |= (uint32_t)h << (8 * e);
        l4 |= (uint32_t)l << (8 * e);
      }
      hw[w] = h4;
      lw[w] = l4;
    }
    const intx8 hv = {(int)hw[0], (int)hw[1], (int)hw[2], (int)hw[3],
                      (int)hw[4], (int)hw[5], (int)hw[6], (int)hw[7]};
    const intx8 lv = {(int)lw[0], (int)lw[1], (int)lw[2], (int)lw[3],
                      (int)lw[4], (int)lw[5], (int)lw[6], (int)lw[7]};
    *reinterpret_cast<intx8 *>(Xs + swz8(u, f0, nsteps, 0)) = hv;
    *reinterpret_cast<intx8 *>(Xs + swz8(u, f0, nsteps, 1)) = lv;
    *reinterpret_cast<intx8 *>(Xb + swz8(r, f0, nsteps, 0)) = hv;
    *reinterpret_cast<intx8 *>(Xb + swz8(r, f0, nsteps, 1)) = lv;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    s += __shfl_xor(s, off, 64);
  if (lane == 0) {
    nrm[u] = s;
    isc[u] = tiny ? -1.0f : ldexpf(1.0f, -ex);
    flag[u] = 1;
    rcnt[r] = 0;
    rcnt[nb + r] = 0;
  }
  if (r == 0 && lane == 1)
    qctl[1] = 0;
}

// One pass of the FP8 screening tiles: the grid, the block and wave tiling,
// the passes' outputs and the bound arithmetic are k_bf16_screen's (see
// there); a feature step is 64 features, three MFMAs per operand pair, and
// the group's sums are scaled by 1 / (s_u s_j) when the tile is done.  Two
// ring stages (32 KB each) keep two blocks per CU.
template <int PASS>
__global__ __launch_bounds__(256, 2) void k_fp8_screen(
    const uint8_t *__restrict__ Xs, const uint8_t *__restrict__ Xb,
    int nsteps, const float *__restrict__ nrm, const float *__restrict__ isc,
    const float *__restrict__ pmax, const float *__restrict__ tau,
    uint64_t b0, uint64_t nb, uint64_t n, int nsplit, int rtiles,
    float *__restrict__ part, uint32_t *__restrict__ rlist,
    unsigned *__restrict__ rcnt, unsigned rcap, unsigned *__restrict__ qctl,
    float coef, float cfac, float *__restrict__ lbm) {
  __shared__ float s_nj[TC], s_pm[TC], s_ij[TC];
  __shared__ __attribute__((aligned(16))) uint8_t ring[NSTG][STAGE];
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

  float nu[2], tr[2], iu[2];
  bool rok[2];
#pragma unroll
  for (int sb = 0; sb < 2; ++sb) {
    const uint64_t rl = r0 + wb * 64 + sb * 32 + (lane & 31);
    rok[sb] = rl < nb;
    nu[sb] = rok[sb] ? nrm[b0 + rl] : 0.0f;
    iu[sb] = rok[sb] ? isc[b0 + rl] : 0.0f;
    tr[sb] = (PASS == 2 && rok[sb]) ? tau[rl] : 0.0f;
  }
  uint32_t top[2][TOPL];
#pragma unroll
  for (int sb = 0; sb < 2; ++sb)
#pragma unroll
    for (int q = 0; q < TOPL; ++q)
      top[sb][q] = 0x7f800000u; // +inf

  const uint64_t stepw = 4096; // bytes per (group, step): hi + lo blocks
  const uint64_t gstride = (uint64_t)nsteps * stepw; // one group's steps
  // the operand ring as in k_bf16_screen: wave w loads blocks 4w .. 4w + 3
  // of each stage (2 KB each, two 16-byte direct-to-LDS loads per lane and
  // block), one step ahead of the step on the matrix cores
  const bool aw = wave < 2;
  const uint8_t *const bsrc =
      Xb + ((r0 / 32 + (uint64_t)(wave & 1) * 2) * gstride) + lane * 16;
  const uint32_t my_tiles =
      split < ntile ? (uint32_t)((ntile - 1 - (uint64_t)split) / nsplit + 1)
                    : 0u;
  uint32_t pti = 0, pst = 0;
#define GM_RING8_ISSUE(stg)                                                    \
  {                                                                            \
    const bool live = pti < my_tiles;                                          \
    const uint32_t qt = live ? pti : my_tiles - 1;                             \
    const uint32_t qs = live ? pst : (uint32_t)nsteps - 1;                     \
    const uint8_t *src =                                                       \
        (aw ? Xs + (((split + (uint64_t)qt * nsplit) * TC / 32 +               \
                     (uint64_t)wave * 2) * gstride) + lane * 16                \
            : bsrc) +                                                          \
        (uint64_t)qs * stepw;                                                  \
    uint8_t *dst = &ring[(stg)][wave * 4 * 2048];                              \
    _Pragma("unroll") for (int bk = 0; bk < 4; ++bk) {                         \
      /* block bk: group (bk >> 1) of this wave's pair, part bk & 1 */         \
      const uint8_t *s0 = src + (uint64_t)(bk >> 1) * gstride +               \
                          (uint64_t)(bk & 1) * 2048;                           \
      __builtin_amdgcn_global_load_lds(                                        \
          s0, (__attribute__((address_space(3))) void *)(dst + bk * 2048), 16, \
          0, 0);                                                               \
      __builtin_amdgcn_global_load_lds(                                        \
          s0 + 1024,                                                           \
          (__attribute__((address_space(3))) void *)(dst + bk * 2048 + 1024),  \
          16, 0, 0);                                                           \
    }                                                                          \
    if (++pst == (uint32_t)nsteps) {                                           \
      pst = 0;                                                                 \
      ++pti;                                                                   \
    }                                                                          \
  }
  const uint32_t ring_base =
      (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const char *)(
          (const char *)&ring[0][0]) +
      (uint32_t)lane * 32u;
  uint32_t gs = 0;
  if (my_tiles > 0)
    GM_RING8_ISSUE(0)

  for (uint32_t ti = 0; ti < my_tiles; ++ti) {
    const uint64_t c0 = (split + (uint64_t)ti * nsplit) * TC;
    lds_barrier8();
    if (tid < TC) {
      const uint64_t j = c0 + tid;
      s_nj[tid] = j < n ? nrm[j] : 0.0f;
      s_ij[tid] = j < n ? isc[j] : 0.0f;
      s_pm[tid] = (PASS == 2 && j < b0) ? pmax[j] : -INFINITY;
    }
    floatx16 acc[2][2], grp[2][2];
#pragma unroll
    for (int sa = 0; sa < 2; ++sa)
#pragma unroll
      for (int sb = 0; sb < 2; ++sb)
        acc[sa][sb] = floatx16{};
    const floatx16 z = {};
    // one feature step: wait for this step's stage (this wave's loads, then
    // every wave's at the barrier, which also frees the other stage), load
    // the next step into the other stage, then 12 MFMAs
#define GM_SCREEN8_STEP(FRESH)                                                 \
  {                                                                            \
    __builtin_amdgcn_s_waitcnt((7 << 4) | (15 << 8)); /* vmcnt(0) */           \
    __builtin_amdgcn_s_barrier();                                              \
    GM_RING8_ISSUE((gs + 1) % NSTG)                                            \
    intx8 ah0, al0, ah1, al1, bh0, bl0, bh1, bl1;                              \
    ring_read8(ring_base + (uint32_t)(gs % NSTG) * (uint32_t)STAGE +           \
                   (uint32_t)wa * 8192u,                                       \
               ring_base + (uint32_t)(gs % NSTG) * (uint32_t)STAGE + 16384u +  \
                   (uint32_t)wb * 8192u,                                       \
               ah0, al0, ah1, al1, bh0, bl0, bh1, bl1);                        \
    grp[0][0] = mfma8(ah0, bh0, FRESH ? z : grp[0][0]);                        \
    grp[0][1] = mfma8(ah0, bh1, FRESH ? z : grp[0][1]);                        \
    grp[1][0] = mfma8(ah1, bh0, FRESH ? z : grp[1][0]);                        \
    grp[1][1] = mfma8(ah1, bh1, FRESH ? z : grp[1][1]);                        \
    grp[0][0] = mfma8(ah0, bl0, grp[0][0]);                                    \
    grp[0][1] = mfma8(ah0, bl1, grp[0][1]);                                    \
    grp[1][0] = mfma8(ah1, bl0, grp[1][0]);                                    \
    grp[1][1] = mfma8(ah1, bl1, grp[1][1]);                                    \
    grp[0][0] = mfma8(al0, bh0, grp[0][0]);                                    \
    grp[0][1] = mfma8(al0, bh1, grp[0][1]);                                    \
    grp[1][0] = mfma8(al1, bh0, grp[1][0]);                                    \
    grp[1][1] = mfma8(al1, bh1, grp[1][1]);                                    \
    ++gs;                                                                      \
  }
    for (int st = 0; st < nsteps;) {
      const int gend = min(st + FP8_GROUP_STEPS, nsteps);
      GM_SCREEN8_STEP(true)
      for (++st; st < gend; ++st)
        GM_SCREEN8_STEP(false)
#pragma unroll
      for (int sa = 0; sa < 2; ++sa)
#pragma unroll
        for (int sb = 0; sb < 2; ++sb)
          acc[sa][sb] += grp[sa][sb];
    }
#undef GM_SCREEN8_STEP
    lds_barrier8(); // s_nj / s_ij / s_pm written

    const uint64_t cols = min((uint64_t)TC, n - c0);
    const __amdgpu_buffer_rsrc_t lr = __builtin_amdgcn_make_buffer_rsrc(
        lbm ? lbm + c0 * nb : nullptr, 0,
        lbm ? (int)(cols * nb * sizeof(float)) : 0, 0x00020000);
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
        const float ij = s_ij[jl];
        const float pm = s_pm[jl];
#pragma unroll
        for (int sb = 0; sb < 2; ++sb) {
          const uint64_t rl = r0 + wb * 64 + sb * 32 + (lane & 31);
          const bool ok = rok[sb] && j < n && j != b0 + rl;
          const float s = nu[sb] + nj;
          // the scaled vectors' dot back to the rows' (powers of two)
          const float dot = (acc[sa][sb][i] * iu[sb]) * ij;
          const float dt = __builtin_fmaf(-2.0f, dot, s);
          // an unscreened row (negative inverse scale) bounds nothing
          const float bnd = (iu[sb] < 0.0f || ij < 0.0f)
                                ? INFINITY
                                : __builtin_fmaf(coef, s, 1e-30f);
          if (PASS == 1) {
            __builtin_amdgcn_raw_buffer_store_b32(
                __float_as_uint((dt - bnd) * (1.0f - cfac)), lr,
                (int)(lbase[sb] +
                      (uint32_t)(sa * 32 + (i >> 2) * 8 + (i & 3)) * cstride),
                0, 0);
            const float ub = fmaxf((dt + bnd) * (1.0f + cfac), 0.0f);
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
            const bool rev = lb <= pm;
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

#undef GM_RING8_ISSUE

  if (PASS == 1) {
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

template __global__ void k_fp8_screen<1>(
    const uint8_t *, const uint8_t *, int, const float *, const float *,
    const float *, const float *, uint64_t, uint64_t, uint64_t, int, int,
    float *, uint32_t *, unsigned *, unsigned, unsigned *, float, float,
    float *);
template __global__ void k_fp8_screen<2>(
    const uint8_t *, const uint8_t *, int, const float *, const float *,
    const float *, const float *, uint64_t, uint64_t, uint64_t, int, int,
    float *, uint32_t *, unsigned *, unsigned, unsigned *, float, float,
    float *);

} // namespace gm
