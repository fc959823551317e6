// gm_sgd_noalloc: the SGD of gm_debug_sgd (all epochs of the engine's
// kernels on a caller-given embedding and COO) for callers that cannot
// allocate device memory -- code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc.  Inputs and outputs are the
// caller's device arrays; every buffer the kernels need beyond them (the
// per-edge schedule, head degrees, the rounding grid, the cooperative
// kernels' barrier words and wave slots, their row records, counts and
// flags) lives in the file-scope device array `scratch`, of which each
// device holds one copy for the life of the process.  Nothing here calls hipMalloc, hipFree or an
// allocating library routine.
#include "../../include/gpumap_c.h"
#include "kernels.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

namespace gm {
namespace {

// scratch layout (bytes): eps, nxt, nneg (4 B per edge each), deg (4 B per
// row), the SGD's row records (32 B per row), then the barrier words and
// wave slots, then the control words
constexpr uint64_t NA_EDGES = GM_SGD_NOALLOC_MAX_EDGES;
constexpr uint64_t NA_ROWS = GM_SGD_NOALLOC_MAX_ROWS;
constexpr uint64_t OFF_EPS = 0;
constexpr uint64_t OFF_NXT = OFF_EPS + 4 * NA_EDGES;
constexpr uint64_t OFF_NNEG = OFF_NXT + 4 * NA_EDGES;
constexpr uint64_t OFF_DEG = OFF_NNEG + 4 * NA_EDGES;
constexpr uint64_t OFF_FXP = OFF_DEG + 4 * NA_ROWS;
constexpr uint64_t OFF_BAR = OFF_FXP + 32 * NA_ROWS;
constexpr uint64_t BAR_WORDS = 4 + SGD_MAX_WAVES;
constexpr uint64_t OFF_CTL = (OFF_BAR + 4 * BAR_WORDS + 255) / 256 * 256;
// control words: [0] edge count, [1] maximum weight (float bits), [2] an
// index out of range, [3] rounding grid (float), [4] maximum head degree
constexpr uint64_t CTL_WORDS = 8;
constexpr uint64_t SCRATCH_BYTES = OFF_CTL + 4 * CTL_WORDS;

__attribute__((used)) static __device__ __align__(256)
    unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_na_mu;
unsigned char *g_na_ptr[MAX_DEVICES] = {};
int g_na_cu[MAX_DEVICES] = {};

struct NaFail : std::runtime_error {
  int code;
  NaFail(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void na_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw NaFail(GM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// The device address of this device's scratch (looked up once per device,
// with the device current), and its CU count.
unsigned char *scratch_of(int dev, int *num_cu) {
  if (dev < 0 || dev >= MAX_DEVICES)
    throw NaFail(GM_ERR_INVALID, "device id out of range");
  std::lock_guard<std::mutex> l(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    na_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
             "hipGetSymbolAddress(scratch)");
    int cu = 0;
    na_check(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount,
                                   dev),
             "hipDeviceGetAttribute");
    g_na_ptr[dev] = static_cast<unsigned char *>(p);
    g_na_cu[dev] = cu;
  }
  *num_cu = g_na_cu[dev];
  return g_na_ptr[dev];
}

// Per call: the edge count, cleared maximum weight and range flag, zeroed
// head degrees (what the previous call left there is not reused).
__global__ void k_na_reset(uint32_t *ctl, int *deg, uint64_t n, uint32_t nnz) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    ctl[0] = nnz;
    ctl[1] = 0u;
    ctl[2] = 0u;
  }
  for (uint64_t r = i; r < n; r += (uint64_t)gridDim.x * blockDim.x)
    deg[r] = 0;
}

// Every slot's indices in range (an empty slot has head = tail = -1) and the
// largest positive weight over the live slots (non-negative floats order as
// their bit patterns).
__global__ void k_na_check(const int32_t *head, const int32_t *tail,
                           const float *w, uint64_t nnz, uint64_t n,
                           uint32_t *ctl) {
  uint32_t mw = 0u, bad = 0u;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const int32_t h = head[e], t = tail[e];
    if (h == -1 && t == -1)
      continue;
    if (h < 0 || (uint64_t)h >= n || t < 0 || (uint64_t)t >= n) {
      bad = 1u;
      continue;
    }
    const float v = w[e];
    if (v > 0.0f)
      mw = max(mw, __float_as_uint(v));
  }
  for (int off = 32; off > 0; off >>= 1) {
    mw = max(mw, (uint32_t)__shfl_xor((int)mw, off, 64));
    bad |= (uint32_t)__shfl_xor((int)bad, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (mw)
      atomicMax(ctl + 1, mw);
    if (bad)
      atomicOr(ctl + 2, 1u);
  }
}

inline unsigned grid_for(uint64_t n, unsigned tpb, unsigned cap) {
  return (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((n + tpb - 1) / tpb, cap));
}

} // namespace
} // namespace gm

extern "C" int gm_sgd_noalloc(float *emb, uint64_t n, int32_t c,
                              const int32_t *head, const int32_t *tail,
                              const float *w, int64_t nnz, int32_t n_epochs,
                              int32_t nsr, float a, float b, float alpha0,
                              float gamma, uint64_t seed, int32_t per_epoch) {
  using namespace gm;
  try {
    if (n == 0 || c < 1 || c > 4 || nnz < 0 || nsr < 1 || per_epoch < 0 ||
        per_epoch > 2 || n_epochs < 0)
      throw NaFail(GM_ERR_INVALID, "invalid SGD arguments");
    if (n > NA_ROWS || (uint64_t)nnz > NA_EDGES)
      throw NaFail(GM_ERR_INVALID,
                   "gm_sgd_noalloc supports at most " +
                       std::to_string(NA_ROWS) + " rows and " +
                       std::to_string(NA_EDGES) + " COO slots (got " +
                       std::to_string(n) + " rows, " + std::to_string(nnz) +
                       " slots)");
    if (nnz > 0 && (!emb || !head || !tail || !w))
      throw NaFail(GM_ERR_INVALID, "null array");
    int dev = 0;
    na_check(hipGetDevice(&dev), "hipGetDevice");
    int num_cu = 0;
    unsigned char *base = scratch_of(dev, &num_cu);
    float *eps = reinterpret_cast<float *>(base + OFF_EPS);
    float *nxt = reinterpret_cast<float *>(base + OFF_NXT);
    float *nneg = reinterpret_cast<float *>(base + OFF_NNEG);
    int *deg = reinterpret_cast<int *>(base + OFF_DEG);
    unsigned *bar = reinterpret_cast<unsigned *>(base + OFF_BAR);
    uint32_t *ctl = reinterpret_cast<uint32_t *>(base + OFF_CTL);
    float *rounding = reinterpret_cast<float *>(ctl + 3);
    int *maxdeg = reinterpret_cast<int *>(ctl + 4);
    const FixFrame fx{reinterpret_cast<int32_t *>(base + OFF_FXP)};
    hipStream_t s = nullptr; // the null stream
    // no session's kernels (or another call's) run on the device meanwhile:
    // the cooperative kernels need every CU, and the scratch is shared
    ExclusiveDevice excl;
    k_na_reset<<<grid_for(n, 256, 1024), 256, 0, s>>>(ctl, deg, n,
                                                      (uint32_t)nnz);
    na_check(hipGetLastError(), "reset kernel");
    if (nnz > 0) {
      k_na_check<<<grid_for((uint64_t)nnz, 256, 1024), 256, 0, s>>>(
          head, tail, w, (uint64_t)nnz, n, ctl);
      na_check(hipGetLastError(), "check kernel");
    }
    uint32_t got[3] = {0, 0, 0};
    na_check(hipMemcpyAsync(got, ctl, sizeof(got), hipMemcpyDeviceToHost, s),
             "read-back of the checks");
    na_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (got[2])
      throw NaFail(GM_ERR_INVALID, "COO index out of range");
    float mw = 0.0f;
    std::memcpy(&mw, &got[1], 4);
    if (!(mw > 0.0f) || nnz == 0 || n_epochs == 0) {
      set_last_error(nullptr);
      return GM_OK; // nothing to optimise (gm_debug_sgd's rule)
    }
    const float nsr_inv = 1.0f / (float)nsr;
    const uint64_t cap = (uint64_t)nnz;
    k_sgd_init<<<grid_for(cap, 256, 1u << 30), 256, 0, s>>>(
        w, head, ctl, ctl + 1, n_epochs, nsr_inv, eps, nxt, nneg, deg, maxdeg,
        bar, (unsigned)BAR_WORDS);
    na_check(hipGetLastError(), "k_sgd_init");
    k_max_int<<<grid_for(n, 256, 64), 256, 0, s>>>(
        deg, NRows{nullptr, n}, maxdeg, alpha0, rounding, bar + 3);
    na_check(hipGetLastError(), "k_max_int");
    int variant = 0;
    const bool coop = sgd_launches(
        c, emb, head, tail, ctl, (uint32_t)n, eps, nxt, nneg, n_epochs, alpha0,
        gamma, a, b, nsr_inv, seed, rounding, bar, nullptr, cap, num_cu, s,
        per_epoch == 1   ? SGD_PER_EPOCH
        : per_epoch == 2 ? SGD_STREAM
                         : SGD_AUTO,
        &variant, false, fx);
    uint32_t timeout = 0;
    if (coop)
      na_check(hipMemcpyAsync(&timeout, bar + 2, 4, hipMemcpyDeviceToHost, s),
               "read-back of the barrier flag");
    na_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (timeout)
      throw NaFail(GM_ERR_HIP, "SGD device-wide barrier timed out");
    set_last_error(nullptr);
    return GM_OK;
  } catch (const NaFail &x) {
    set_last_error(x.what());
    return x.code;
  } catch (const std::exception &x) {
    set_last_error(x.what());
    return GM_ERR_HIP;
  }
}
