// Host engine of the gfx950 progressive UMAP path and its C ABI
// (include/gpumap_c.h).
//
// Engine restates the reference's GPUMAP::Impl (cpp/src/impl.cuh:29-534):
// the same queues (queue.cuh), batch selection, seeds, convergence and
// completion rules, run() timing and result fields.  Unlike the reference,
// all graph state lives in HBM for the whole session: the host keeps only the
// queues and receives, per insertion/update pass, the ascending list of rows
// whose neighbour lists changed.  One HIP stream per session, used in order
// (the reference's three streams are used strictly serially,
// impl.cuh:64-68,161,181,201).
#include "../../include/gpumap_c.h"
#include "common.hpp"
#include "kernels.hpp"
#include "dlpack_abi.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <numeric>
#include <queue>
#include <random>
#include <stdexcept>
#include <string>
#include <map>
#include <tuple>
#include <unordered_set>
#include <functional>
#include <mutex>
#include <shared_mutex>
#include <condition_variable>
#include <vector>

namespace gm {
namespace {

thread_local std::string g_err;

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what, int line) {
  if (e == hipSuccess)
    return;
  const int code = (e == hipErrorOutOfMemory) ? GM_ERR_OOM : GM_ERR_HIP;
  throw Fail(code, std::string(what) + " failed at engine.hip:" +
                       std::to_string(line) + ": " + hipGetErrorString(e) +
                       " (hipError " + std::to_string((int)e) + ")");
}
#define HIPC(x) ::gm::hip_check((x), #x, __LINE__)
#define LAUNCH_CHECK() HIPC(hipGetLastError())

template <typename T> T *dalloc(size_t n) {
  void *p = nullptr;
  if (n == 0)
    n = 1;
  HIPC(hipMalloc(&p, n * sizeof(T)));
  return static_cast<T *>(p);
}

inline unsigned blocks_for(uint64_t n, unsigned tpb) {
  return (unsigned)((n + tpb - 1) / tpb);
}

// The end of a phase's device work: the stream is polled for up to 2 ms (a
// cycle's phases end within that), then waited for by blocking.
inline void stream_wait(hipStream_t s) {
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t e = hipStreamQuery(s);
    if (e == hipSuccess)
      return;
    if (e != hipErrorNotReady)
      HIPC(e);
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2))
      break;
  }
  HIPC(hipStreamSynchronize(s));
}

// reference UpdateQueue (queue.cuh:46-80): FIFO without duplicates; the
// membership test is a byte per row (rows are < the session's capacity)
struct DedupQueue {
  std::deque<uint32_t> q;
  std::vector<uint8_t> in;
  size_t size() const { return q.size(); }
  void push(uint32_t v) {
    if (v >= in.size())
      in.resize(std::max<size_t>((size_t)v + 1, 2 * in.size()), 0);
    if (!in[v]) {
      in[v] = 1;
      q.push_back(v);
    }
  }
  std::vector<uint32_t> pop_n(size_t n) {
    n = std::min(n, q.size());
    std::vector<uint32_t> out(q.begin(), q.begin() + n);
    q.erase(q.begin(), q.begin() + n);
    for (uint32_t v : out)
      in[v] = 0;
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

constexpr int NPART = 256;
constexpr int WINDOW_N_BATCHES = 3;

// device scalars, one array
enum Scalar {
  S_TOTAL = 0,
  S_OWNERS,
  S_MAXW,
  S_MIN,
  S_MAX,
  S_NNZ,
  S_NROWS, // the session's row count, for kernels replayed from a HIP graph
  S_COUNT
};

} // namespace

// Held from the start of a cooperative SGD launch until it has completed, so
// that no two such launches (of any sessions in this process) are in flight
// on the device at once: each relies on all its blocks being resident.
std::mutex g_coop_mu;

// The sessions' device work in this process: an insertion or update phase of
// run() holds it shared (the phase ends with its stream synchronised); an
// embedding phase of run() (a grid-wide SGD kernel) and a batched cycle hold
// it exclusively, so that a grid-wide kernel never shares the device with
// other sessions' kernels, whose CUs its resident blocks would wait for.
// Writers are preferred: a waiting exclusive holder blocks new shared ones.
class DevLock {
  std::mutex m_;
  std::condition_variable cv_;
  int readers_ = 0, waiting_ = 0;
  bool writer_ = false;

public:
  void lock_shared() {
    std::unique_lock<std::mutex> l(m_);
    cv_.wait(l, [&] { return !writer_ && waiting_ == 0; });
    ++readers_;
  }
  void unlock_shared() {
    std::lock_guard<std::mutex> l(m_);
    if (--readers_ == 0)
      cv_.notify_all();
  }
  void lock() {
    std::unique_lock<std::mutex> l(m_);
    ++waiting_;
    cv_.wait(l, [&] { return !writer_ && readers_ == 0; });
    --waiting_;
    writer_ = true;
  }
  void unlock() {
    std::lock_guard<std::mutex> l(m_);
    writer_ = false;
    cv_.notify_all();
  }
};
DevLock g_dev_mu;

ExclusiveDevice::ExclusiveDevice() {
  g_dev_mu.lock();
  g_coop_mu.lock();
}
ExclusiveDevice::~ExclusiveDevice() {
  g_coop_mu.unlock();
  g_dev_mu.unlock();
}

void set_last_error(const char *msg) {
  if (msg && *msg)
    g_err = msg;
  else
    g_err.clear();
}

// The SGD's launches on a prepared schedule (eps, nxt, nneg, the rounding
// grid and zeroed barrier words, sgd_setup's work): all epochs in one
// cooperative launch of the kernel that fits the edge set (SGD_AUTO: the
// register-resident kernels, else the row-owner kernel when `pf` is given,
// else the streaming kernel), the streaming cooperative kernel (SGD_STREAM),
// the row-owner kernel (SGD_PULL), or one launch per epoch (SGD_PER_EPOCH,
// and whenever no cooperative launch is possible).  Returns true when a
// cooperative kernel ran (its timeout flag is gbar[2]); *variant names it
// (4 pipelined chain/row/draw waves, 2 per-wave slots, 0 resident counter grid, 1
// streaming, 3 row-owner).
// Epoch lag of the cooperative SGD: a block starts epoch e + 1 once every
// block has finished epoch e - SGD_LAG (0 would be a full grid barrier per
// epoch, the reference's one launch per epoch).  The Hogwild updates are
// unordered within an epoch already; the lag lets the boundary's counter read
// overlap the next epoch's work.
bool sgd_launches(int c, float *e, const int32_t *head, const int32_t *tail,
                  const uint32_t *nnz_p, uint32_t nrows, const float *eps,
                  float *nxt, float *nneg, int n_epochs, float alpha0,
                  float gamma, float a, float b, float nsr_inv, uint64_t seed,
                  const float *rounding, unsigned *gbar,
                  unsigned long long *st, uint64_t cap, int num_cu,
                  hipStream_t s, int sgd_mode, int *variant, bool verbose,
                  FixFrame fx, const PullFrame *pf) {
  if (sgd_mode != SGD_PER_EPOCH) {
    // the per-wave, streaming and row-owner kernels run on the fixed-point
    // copy (made here, folded back below when one of them ran)
    if (fx.rec)
      HIPC(launch_fix(c, 0, e, nrows, fx, rounding, num_cu, s));
    const bool pull = pf && fx.rec && sgd_mode != SGD_STREAM;
    hipError_t fe = hipErrorInvalidConfiguration;
    if (sgd_mode != SGD_PULL)
      fe = launch_sgd_fused(
          c, e, head, tail, nnz_p, nrows, eps, nxt, nneg, n_epochs, alpha0,
          gamma, a, b, nsr_inv, seed, rounding, gbar, st, SGD_LAG, cap,
          num_cu, s, variant,
          sgd_mode == SGD_STREAM ? SGD_STREAM_FORCE
          : pull                 ? SGD_STREAM_NONE
                                 : SGD_STREAM_FALLBACK,
          fx);
    if (sgd_mode == SGD_PULL && !pull)
      throw Fail(GM_ERR_HIP, "the row-owner SGD kernel needs its scratch and "
                             "the row records");
    if (fe == hipErrorInvalidConfiguration && pull) {
      (void)hipGetLastError();
      *variant = 3;
      fe = launch_sgd_pull(c, *pf, fx, head, tail, nnz_p, cap, nrows, eps,
                           nxt, nneg, n_epochs, alpha0, gamma, a, b, nsr_inv,
                           seed, rounding, gbar, st, num_cu, s);
      if (fe != hipSuccess && sgd_mode == SGD_PULL)
        throw Fail(GM_ERR_HIP, "the row-owner SGD kernel could not be "
                               "launched");
      if (fe == hipErrorInvalidConfiguration) {
        (void)hipGetLastError();
        fe = launch_sgd_fused(c, e, head, tail, nnz_p, nrows, eps, nxt, nneg,
                              n_epochs, alpha0, gamma, a, b, nsr_inv, seed,
                              rounding, gbar, st, SGD_LAG, cap, num_cu, s,
                              variant, SGD_STREAM_FORCE, fx);
      }
    }
    if (fe == hipSuccess) {
      if (fx.rec && *variant != 0)
        HIPC(launch_fix(c, 1, e, nrows, fx, rounding, num_cu, s));
      return true;
    }
    // too many edges for the resident grid, or cooperative launch not
    // available in this context: one launch per epoch instead
    (void)hipGetLastError();
    if (fe != hipErrorInvalidConfiguration && verbose) {
      fprintf(stderr, "[gpumap_amd] cooperative SGD launch unavailable; one "
                      "launch per epoch\n");
      fflush(stderr);
    }
  }
  const unsigned grid = std::min<unsigned>(blocks_for(cap, 256), 16384);
  float alpha = alpha0;
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const uint64_t es = seed + (uint64_t)epoch;
    switch (c) {
    case 1:
      k_sgd_epoch<1><<<grid, 256, 0, s>>>(e, head, tail, nnz_p, nrows, eps,
                                          nxt, nneg, alpha, epoch, gamma, a, b,
                                          nsr_inv, es, rounding, st);
      break;
    case 2:
      k_sgd_epoch<2><<<grid, 256, 0, s>>>(e, head, tail, nnz_p, nrows, eps,
                                          nxt, nneg, alpha, epoch, gamma, a, b,
                                          nsr_inv, es, rounding, st);
      break;
    case 3:
      k_sgd_epoch<3><<<grid, 256, 0, s>>>(e, head, tail, nnz_p, nrows, eps,
                                          nxt, nneg, alpha, epoch, gamma, a, b,
                                          nsr_inv, es, rounding, st);
      break;
    default:
      k_sgd_epoch<4><<<grid, 256, 0, s>>>(e, head, tail, nnz_p, nrows, eps,
                                          nxt, nneg, alpha, epoch, gamma, a, b,
                                          nsr_inv, es, rounding, st);
    }
    LAUNCH_CHECK();
    alpha = alpha0 * (1.0f - (float)epoch / (float)n_epochs);
  }
  return false;
}

struct Engine;
// Every live session of the process, for gm_shutdown.
// ------------------------------------------------- screen shape tuner
// The BF16 screen's launch shape -- the column splits per 128-row tile
// (bsplit, the grid's second factor) and how the survivors are listed (a
// scan of the stored lower bounds, or a second pass over the tiles) -- is
// chosen per (device, screen, padded feature count, row count within a
// factor of 4,
// batch tiles) by timing the candidates around the size heuristic on the
// first insertions that meet the key: each candidate once, on the session
// stream between two events, then the fastest for every later insertion of
// any session in the process.  Results do not depend on the shape (the
// pools are exact); only the time does.  GPUMAP_AMD_TUNE=0 keeps the
// heuristic.
struct ScreenShape {
  int bsplit;
  bool lbscan;
};
struct TuneKey {
  int dev;
  uint64_t mode; // knn_mode: the BF16 and FP8 screens are tuned apart
  uint64_t db;
  int nlog4;
  uint64_t brt;
  bool operator<(const TuneKey &o) const {
    return std::tie(dev, mode, db, nlog4, brt) <
           std::tie(o.dev, o.mode, o.db, o.nlog4, o.brt);
  }
};
struct TuneEntry {
  std::vector<ScreenShape> cand;
  std::vector<float> ms; // < 0: not timed yet
  int issued = 0;        // candidates handed out for timing
  int best = -1;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tune;

bool tune_enabled() {
  static const bool on = [] {
    const char *v = std::getenv("GPUMAP_AMD_TUNE");
    return !(v && v[0] == '0');
  }();
  return on;
}

// The shape for an insertion under `key`; *trial is the candidate index to
// time, or -1 when the shape is the tuned (or the only) one.
ScreenShape tune_choose(const TuneKey &key, int heuristic, int max_split,
                        bool lbm_ok, int *trial) {
  *trial = -1;
  if (!tune_enabled())
    return ScreenShape{heuristic, lbm_ok};
  std::lock_guard<std::mutex> l(g_tune_mu);
  TuneEntry &t = g_tune[key];
  if (t.cand.empty()) {
    for (int f : {1, 2, 4}) {
      const int b = std::max(1, std::min(heuristic * f / 2, max_split));
      for (bool scan : {true, false}) {
        if (scan && !lbm_ok)
          continue;
        bool dup = false;
        for (auto &c : t.cand)
          dup |= c.bsplit == b && c.lbscan == scan;
        if (!dup)
          t.cand.push_back(ScreenShape{b, scan});
      }
    }
    t.ms.assign(t.cand.size(), -1.0f);
  }
  if (t.best >= 0)
    return t.cand[t.best];
  if (t.cand.size() == 1) {
    t.best = 0;
    return t.cand[0];
  }
  if (t.issued < (int)t.cand.size()) {
    *trial = t.issued++;
    return t.cand[*trial];
  }
  // every candidate handed out, some not timed yet: the heuristic meanwhile
  return ScreenShape{heuristic, lbm_ok};
}

void tune_report(const TuneKey &key, int trial, float ms) {
  std::lock_guard<std::mutex> l(g_tune_mu);
  auto it = g_tune.find(key);
  if (it == g_tune.end() || trial < 0 || trial >= (int)it->second.ms.size())
    return;
  TuneEntry &t = it->second;
  t.ms[trial] = ms;
  int best = -1;
  for (int i = 0; i < (int)t.ms.size(); ++i) {
    if (t.ms[i] < 0.0f)
      return; // not all timed yet
    if (best < 0 || t.ms[i] < t.ms[best])
      best = i;
  }
  t.best = best;
}

std::mutex g_live_mu;
std::vector<Engine *> g_live;

struct Engine {
  gm_params p;
  uint64_t N, D, Dp;
  int k, c;
  bool euclid;
  int device = 0;
  hipStream_t s = nullptr;
  std::atomic<int> refs{1};
  // gm_release: the device memory is freed as soon as no exported tensor
  // holds the session (at the call, or when the last export is deleted);
  // rel_mu orders that decision against the exports' deleters
  std::mutex rel_mu;
  bool release_pending = false;
  std::atomic<bool> released{false};

  // device state
  float *X = nullptr;
  uint64_t *pool = nullptr;
  uint32_t *pidx = nullptr;
  float *pdist = nullptr;
  float *emb = nullptr;
  // the cooperative SGD kernels' row records (FixFrame): N records of
  // 2 (c, or 4 for c = 3) words; GPUMAP_AMD_SGD_FIX=0 keeps the kernels on
  // the float embedding
  int32_t *fxp = nullptr;
  bool sgd_fix = true;
  // With GPUMAP_AMD_SGD_PULL=1, embedding ops whose edge sets exceed the
  // register-resident kernels (PULL_MIN_SLOTS slots) run on the row-owner
  // kernel: reproducible bit for bit from run to run (every epoch reads the
  // positions the previous one left, integer sums), measured 10% faster
  // than the streaming kernel on the 70,000 x 784 full sweep but 1.5x and
  // 2.6x slower on the 200,000 x 300 (k = 30) and 2,000,000 x 128 (3-D)
  // ones, so off by default.  Its scratch is allocated on first use, sized
  // for a full sweep.
  PullFrame pullf{};
  bool sgd_pull = false;
  static constexpr uint64_t PULL_MIN_SLOTS = 1u << 17;
  // With GPUMAP_AMD_GRAPH=1 the embedding op's kernel sequences before and
  // after its SGD launch run as HIP graphs: captured once (stream capture)
  // and relaunched while the host values they depend on stay the same, the
  // row count read from a device scalar; re-captured and patched into the
  // executable graph (hipGraphExecUpdate) when they change.  The cooperative
  // SGD launch stays outside the graphs, between them on the stream, with
  // its timing events.  Off by default: on the 70,000 x 784 op plan the
  // graph launches measured 3% slower per session than the direct launches.
  bool use_graph = false;
  hipGraphExec_t gx_pre = nullptr, gx_post = nullptr;
  uint64_t graph_replays = 0, graph_instantiations = 0;
  // the host values a captured sequence depends on (its kernels read the
  // row count from S_NROWS and the batch from dbatch): an op with the same
  // key relaunches the executable graph as it is
  struct GraphKey {
    uint64_t nb = 0, sweeps = 0;
    bool partial = false, setup = false, valid = false;
    bool operator==(const GraphKey &o) const {
      return nb == o.nb && sweeps == o.sweeps && partial == o.partial &&
             setup == o.setup && valid == o.valid;
    }
  };
  GraphKey gk_pre, gk_post;
  uint8_t *flag = nullptr;
  uint32_t *list = nullptr;
  float *pmax = nullptr;      // per-row pool maximum distance (inf if not full)
  uint64_t *part = nullptr;   // exact insertion: per (row, split) top W
  uint64_t part_cap = 0;
  // BF16 / FP8 screening (knn_mode exact_bf16, exact_fp8): hi/lo split
  // rows (Db features, 4 or 2 bytes each), FP32 norms, the FP8 rows' inverse
  // scales, per-batch-row tau
  uint64_t Db = 0;
  __bf16 *Xs = nullptr, *Xb = nullptr; // column and batch operand stores
  float *nrm = nullptr, *tau = nullptr, *isc = nullptr;
  uint32_t *rlist = nullptr; // survivors per batch row: column | flags
  uint64_t *fkeys = nullptr; // forward keys of the survivors
  unsigned *rcnt = nullptr;  // survivors per batch row
  unsigned *qctl = nullptr;  // [1]: a survivor list overflowed
  // pass-1 lower bounds of every (column, batch row) pair, column-major, so
  // that the survivor lists need no second pass over the tiles; absent when
  // N x batch floats exceed lbm_max_bytes (8 GiB, or the environment's
  // GPUMAP_AMD_LBM_MAX_BYTES at creation)
  float *lbm = nullptr;
  uint64_t lbm_max_bytes = 8ull << 30;
  uint64_t xb_cap = 0;
  // survivors per batch row: 2048, or up to 16384 for sessions above 2^18
  // rows, where a new row can lie within many earlier rows' pool maxima (in
  // concentrated clusters nearly every row of its cluster); bounded by the
  // select kernel's LDS (8 bytes per key)
  unsigned rcap = 2048;
  uint64_t ins_pairs = 0; // (batch row, column) pairs screened or computed
  uint32_t *bcount = nullptr;
  uint32_t *scal = nullptr;
  unsigned long long *fresh = nullptr;
  unsigned long long *kstats = nullptr; // device work counters (16)
  // profiling: HIP events around the timed kernels on this stream
  bool prof = false;
  struct Timed {
    int which;
    hipEvent_t a, b;
    uint64_t launches;
  };
  std::vector<Timed> pending;
  std::vector<hipEvent_t> ev_free; // recorded and folded, ready for reuse
  gm_kernel_stats totals[GM_KSTAT_COUNT] = {};
  // device counters kstats[13], kstats[14]: edge-epochs scanned by the
  // resident and by the streaming cooperative SGD kernels
  static constexpr int KS_SGD_EE = 13, KS_STREAM_EE = 14;
  uint64_t sgd_waves_launches = 0;
  // last cooperative SGD launch: 0 counter-paced resident grid, 1 stream,
  // 2 per-wave arrival slots
  int sgd_variant = 0;
  // insertion scratch
  uint64_t ins_cap = 0;
  uint32_t *ins_ids = nullptr, *cand = nullptr;
  int *ins_cnt = nullptr, *ccount = nullptr;
  float *cdist = nullptr;
  // update
  uint32_t *fn = nullptr, *fo = nullptr, *newmask = nullptr,
           *owner_mark = nullptr, *owners = nullptr, *rows = nullptr;
  int *fnc = nullptr, *foc = nullptr;
  uint64_t *rev_new = nullptr, *rev_old = nullptr;
  // embed
  double *sigma = nullptr, *rho = nullptr, *psum = nullptr, *mean = nullptr;
  unsigned long long *pcnt = nullptr;
  float *w = nullptr, *wout = nullptr, *eps = nullptr, *nxt = nullptr,
        *nneg = nullptr, *snap = nullptr, *rounding = nullptr;
  uint8_t *wflags = nullptr;
  int32_t *head = nullptr, *tail = nullptr;
  int *deg = nullptr, *maxdeg = nullptr;
  unsigned *gbar = nullptr; // fused SGD: counter, -, timeout, -, wave slots
  int num_cu = 0;
  bool fused_used = false;
  uint32_t *dbatch = nullptr;
  uint8_t *bmark = nullptr; // batch rows of the current embedding op
  // pinned host staging
  float *h_stage = nullptr;
  uint64_t h_stage_rows = 0;
  uint32_t *h_list = nullptr;
  uint32_t *h_scal = nullptr;
  // rows flagged by the last insertion (ascending), copied back behind the
  // kernels: a prefix of flag_prefix rows and the count; the host reads them
  // once the stream has caught up (insert_finish)
  uint32_t *h_flagged = nullptr;
  uint64_t flag_prefix = 1024;
  struct PendingInsert {
    bool on = false;
    uint64_t b0 = 0, B = 0, pre = 0;
  } pins;
  bool screened() const {
    return p.knn_mode == GM_KNN_EXACT_BF16 || p.knn_mode == GM_KNN_EXACT_FP8;
  }
  bool nndescent() const {
    return p.knn_mode == GM_KNN_NNDESCENT ||
           p.knn_mode == GM_KNN_NNDESCENT_BF16;
  }
  // a screen shape being timed for the tuner (read in insert_finish)
  struct TunePending {
    bool on = false;
    TuneKey key{};
    int trial = -1;
    hipEvent_t a = nullptr, b = nullptr;
  } tune_pend;

  // host state (reference Impl members, impl.cuh:33-55)
  uint64_t n = 0;
  Fifo insertion_queue, embedding_new_queue;
  DedupQueue update_queue;
  std::deque<uint32_t> batch_start_history;
  size_t embedding_cursor = SIZE_MAX;
  size_t embedding_sweeps_done = 0;
  bool insertion_completed = false, update_completed = false,
       embedding_completed = false;
  bool closed = false;
  std::mt19937 rng{42};

  explicit Engine(const gm_params &pp) : p(pp) {
    N = p.n_instances;
    D = p.n_features;
    Dp = (D + 3) / 4 * 4;
    k = (int)p.n_neighbors;
    c = (int)p.n_components;
    euclid = p.metric == GM_EUCLIDEAN;
    HIPC(hipGetDevice(&device));
    update_queue.in.assign(N, 0);
    if (const char *v = std::getenv("GPUMAP_AMD_LBM_MAX_BYTES"))
      lbm_max_bytes = std::strtoull(v, nullptr, 10);
    if (const char *v = std::getenv("GPUMAP_AMD_SGD_FIX"))
      sgd_fix = v[0] != '0';
    if (const char *v = std::getenv("GPUMAP_AMD_GRAPH"))
      use_graph = v[0] != '0';
    if (const char *v = std::getenv("GPUMAP_AMD_SGD_PULL"))
      sgd_pull = v[0] != '0';
    try {
      allocate();
    } catch (...) {
      release();
      throw;
    }
    std::lock_guard<std::mutex> l(g_live_mu);
    g_live.push_back(this);
  }
  ~Engine() {
    {
      std::lock_guard<std::mutex> l(g_live_mu);
      g_live.erase(std::remove(g_live.begin(), g_live.end(), this),
                   g_live.end());
    }
    release();
  }

  void allocate() {
    HIPC(hipStreamCreate(&s));
    const uint64_t nk = N * (uint64_t)k;
    X = dalloc<float>(N * Dp);
    HIPC(hipMemsetAsync(X, 0, N * Dp * sizeof(float), s));
    pool = dalloc<uint64_t>(N * W);
    HIPC(hipMemsetAsync(pool, 0xff, N * W * sizeof(uint64_t), s));
    pidx = dalloc<uint32_t>(nk);
    HIPC(hipMemsetAsync(pidx, 0xff, nk * sizeof(uint32_t), s));
    pdist = dalloc<float>(nk);
    HIPC(hipMemsetD32Async((hipDeviceptr_t)pdist, 0x7f7fffff, nk, s));
    emb = dalloc<float>(N * c);
    fxp = dalloc<int32_t>(N * 2 * (c == 3 ? 4 : c));
    HIPC(hipMemsetAsync(emb, 0, N * c * sizeof(float), s));
    flag = dalloc<uint8_t>(N);
    HIPC(hipMemsetAsync(flag, 0, N, s));
    list = dalloc<uint32_t>(N);
    pmax = dalloc<float>(N);
    HIPC(hipMemsetD32Async((hipDeviceptr_t)pmax, 0x7f800000, N, s));
    if (screened()) {
      const bool f8 = p.knn_mode == GM_KNN_EXACT_FP8;
      Db = f8 ? (D + 63) / 64 * 64 : (D + 15) / 16 * 16;
      const uint64_t rows_ = (N + 127) / 128 * 128;
      // BF16: hi + lo, 2 bytes each; FP8: 1 byte each
      const uint64_t xs_elems = f8 ? rows_ * Db : rows_ * Db * 2;
      Xs = dalloc<__bf16>(xs_elems);
      HIPC(hipMemsetAsync(Xs, 0, xs_elems * sizeof(__bf16), s));
      nrm = dalloc<float>(rows_);
      HIPC(hipMemsetAsync(nrm, 0, rows_ * sizeof(float), s));
      if (f8) {
        isc = dalloc<float>(rows_);
        HIPC(hipMemsetAsync(isc, 0, rows_ * sizeof(float), s));
      }
      // survivor list per batch row: the FP8 screen's looser bound keeps
      // about five times the BF16 screen's survivors
      rcap = f8 ? 8192 : 2048;
      if (N > (1u << 18) || f8) {
        int smem = 0;
        HIPC(hipDeviceGetAttribute(
            &smem, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        rcap = 16384;
        while (rcap > 2048 && (uint64_t)rcap * 8 > (uint64_t)smem)
          rcap /= 2;
      }
      qctl = dalloc<unsigned>(4);
      HIPC(hipMemsetAsync(qctl, 0, 4 * sizeof(unsigned), s));
    }
    if (p.knn_mode == GM_KNN_NNDESCENT_BF16) {
      // the local join's hi/lo operand store (32-feature chunks) and norms
      Db = (D + 31) / 32 * 32;
      Xs = dalloc<__bf16>(N * Db * 2);
      HIPC(hipMemsetAsync(Xs, 0, N * Db * 2 * sizeof(__bf16), s));
      nrm = dalloc<float>(N);
      HIPC(hipMemsetAsync(nrm, 0, N * sizeof(float), s));
    }
    bcount = dalloc<uint32_t>(std::max<uint64_t>(2 * nk, N) / 1024 + 2);
    scal = dalloc<uint32_t>(S_COUNT);
    fresh = dalloc<unsigned long long>(1);
    kstats = dalloc<unsigned long long>(16);
    HIPC(hipMemsetAsync(kstats, 0, 16 * sizeof(unsigned long long), s));
    fn = dalloc<uint32_t>(N * W);
    fo = dalloc<uint32_t>(N * W);
    newmask = dalloc<uint32_t>(N);
    owner_mark = dalloc<uint32_t>(N);
    HIPC(hipMemsetAsync(owner_mark, 0, N * sizeof(uint32_t), s));
    owners = dalloc<uint32_t>(N);
    rows = dalloc<uint32_t>(N);
    fnc = dalloc<int>(N);
    foc = dalloc<int>(N);
    HIPC(hipMemsetAsync(fnc, 0, N * sizeof(int), s));
    HIPC(hipMemsetAsync(foc, 0, N * sizeof(int), s));
    rev_new = dalloc<uint64_t>(N * W);
    rev_old = dalloc<uint64_t>(N * W);
    HIPC(hipMemsetAsync(rev_new, 0xff, N * W * sizeof(uint64_t), s));
    HIPC(hipMemsetAsync(rev_old, 0xff, N * W * sizeof(uint64_t), s));
    sigma = dalloc<double>(N);
    rho = dalloc<double>(N);
    psum = dalloc<double>(NPART);
    pcnt = dalloc<unsigned long long>(NPART);
    mean = dalloc<double>(1);
    w = dalloc<float>(nk);
    wflags = dalloc<uint8_t>(nk);
    head = dalloc<int32_t>(2 * nk);
    tail = dalloc<int32_t>(2 * nk);
    wout = dalloc<float>(2 * nk);
    eps = dalloc<float>(2 * nk);
    nxt = dalloc<float>(2 * nk);
    nneg = dalloc<float>(2 * nk);
    snap = dalloc<float>(N * c);
    rounding = dalloc<float>(1);
    deg = dalloc<int>(N);
    maxdeg = dalloc<int>(1);
    gbar = dalloc<unsigned>(4 + SGD_MAX_WAVES);
    HIPC(hipMemsetAsync(gbar, 0, (4 + SGD_MAX_WAVES) * sizeof(unsigned), s));
    HIPC(hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount,
                               device));
    dbatch = dalloc<uint32_t>(N);
    bmark = dalloc<uint8_t>(N);
    HIPC(hipMemsetAsync(bmark, 0, N, s));
    HIPC(hipHostMalloc((void **)&h_list, N * sizeof(uint32_t)));
    HIPC(hipHostMalloc((void **)&h_scal, S_COUNT * sizeof(uint32_t) + 64));
    HIPC(hipHostMalloc((void **)&h_flagged, (N + 1) * sizeof(uint32_t)));
    HIPC(hipStreamSynchronize(s));
  }

  void ensure_ins(uint64_t B) {
    if (B <= ins_cap)
      return;
    (void)hipFree(ins_ids);
    (void)hipFree(cand);
    (void)hipFree(ins_cnt);
    (void)hipFree(ccount);
    (void)hipFree(cdist);
    ins_ids = cand = nullptr;
    ins_cnt = ccount = nullptr;
    cdist = nullptr;
    ins_cap = 0;
    ins_ids = dalloc<uint32_t>(B * W);
    ins_cnt = dalloc<int>(B);
    cand = dalloc<uint32_t>(B * CAND);
    cdist = dalloc<float>(B * CAND);
    ccount = dalloc<int>(B);
    ins_cap = B;
  }

  void ensure_stage(uint64_t rows_) {
    if (rows_ <= h_stage_rows)
      return;
    if (h_stage)
      (void)hipHostFree(h_stage);
    h_stage = nullptr;
    h_stage_rows = 0;
    HIPC(hipHostMalloc((void **)&h_stage, rows_ * D * sizeof(float)));
    h_stage_rows = rows_;
  }

  void release() {
    if (s)
      (void)hipStreamSynchronize(s);
    for (auto &t : pending) {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
    }
    pending.clear();
    for (hipEvent_t ev : ev_free)
      (void)hipEventDestroy(ev);
    ev_free.clear();
    for (hipGraphExec_t *gx : {&gx_pre, &gx_post}) {
      if (*gx)
        (void)hipGraphExecDestroy(*gx);
      *gx = nullptr;
    }
    // every device pointer is freed once and cleared, so that release() may
    // run again (from the destructor after an early gm_release)
    void **ptrs[] = {
        (void **)&kstats, (void **)&gbar,    (void **)&Xs,      (void **)&Xb,
        (void **)&nrm,    (void **)&tau,     (void **)&rlist,   (void **)&rcnt,
        (void **)&qctl,   (void **)&fkeys,   (void **)&lbm,     (void **)&pmax,
        (void **)&part,   (void **)&X,       (void **)&pool,    (void **)&pidx,
        (void **)&pdist,  (void **)&emb,     (void **)&flag,    (void **)&list,
        (void **)&bcount, (void **)&scal,    (void **)&fresh,   (void **)&ins_ids,
        (void **)&cand,   (void **)&ins_cnt, (void **)&ccount,  (void **)&cdist,
        (void **)&fn,     (void **)&fo,      (void **)&newmask, (void **)&owner_mark,
        (void **)&owners, (void **)&rows,    (void **)&fnc,     (void **)&foc,
        (void **)&rev_new, (void **)&rev_old, (void **)&sigma,  (void **)&rho,
        (void **)&psum,   (void **)&pcnt,    (void **)&mean,    (void **)&w,
        (void **)&wflags, (void **)&head,    (void **)&tail,    (void **)&wout,
        (void **)&eps,    (void **)&nxt,     (void **)&nneg,    (void **)&snap,
        (void **)&rounding, (void **)&deg,   (void **)&maxdeg,  (void **)&dbatch,
        (void **)&bmark,  (void **)&isc,     (void **)&fxp,
        (void **)&pullf.cnt, (void **)&pullf.off, (void **)&pullf.bsum,
        (void **)&pullf.erc, (void **)&pullf.esch, (void **)&pullf.epc,
        (void **)&pullf.d1};
    for (void **q : ptrs) {
      if (*q)
        (void)hipFree(*q);
      *q = nullptr;
    }
    part_cap = ins_cap = xb_cap = 0;
    pullf.slots = pullf.rows = 0;
    if (h_stage)
      (void)hipHostFree(h_stage);
    if (h_list)
      (void)hipHostFree(h_list);
    if (h_scal)
      (void)hipHostFree(h_scal);
    if (h_flagged)
      (void)hipHostFree(h_flagged);
    h_stage = nullptr;
    h_stage_rows = 0;
    h_list = nullptr;
    h_scal = nullptr;
    h_flagged = nullptr;
    if (s)
      (void)hipStreamDestroy(s);
    s = nullptr;
  }

  hipEvent_t ev_get() {
    if (!ev_free.empty()) {
      hipEvent_t ev = ev_free.back();
      ev_free.pop_back();
      return ev;
    }
    hipEvent_t ev = nullptr;
    HIPC(hipEventCreate(&ev));
    return ev;
  }
  hipEvent_t ev_start(int which) {
    if (!prof)
      return nullptr;
    Timed t{which, ev_get(), ev_get(), 0};
    pending.push_back(t);
    HIPC(hipEventRecord(t.a, s));
    return t.a;
  }
  void ev_stop(uint64_t launches) {
    if (!prof)
      return;
    Timed &t = pending.back();
    t.launches = launches;
    HIPC(hipEventRecord(t.b, s));
  }
  // fold recorded event pairs into the totals (after run()'s own stream
  // synchronisation they have completed, so this waits for nothing)
  void fold_events() {
    for (auto &t : pending) {
      HIPC(hipEventSynchronize(t.b));
      float ms = 0.0f;
      HIPC(hipEventElapsedTime(&ms, t.a, t.b));
      totals[t.which].launches += t.launches;
      totals[t.which].device_ms += ms;
      ev_free.push_back(t.a);
      ev_free.push_back(t.b);
    }
    pending.clear();
  }
  // the device work counters into the totals (one read-back, on request)
  void read_counters() {
    HIPC(hipStreamSynchronize(s));
    fold_events();
    unsigned long long c[16];
    HIPC(hipMemcpy(c, kstats, sizeof(c), hipMemcpyDeviceToHost));
    totals[GM_KSTAT_JOIN].work[0] = c[0];
    totals[GM_KSTAT_JOIN].work[1] = c[1];
    totals[GM_KSTAT_JOIN].work[2] = c[2];
    totals[GM_KSTAT_JOIN].work[3] = c[7];
    totals[GM_KSTAT_SGD].work[0] = c[3];
    totals[GM_KSTAT_SGD].work[1] = c[4];
    totals[GM_KSTAT_SGD].work[2] = c[KS_SGD_EE];
    totals[GM_KSTAT_SGD].work[3] = sgd_waves_launches;
    totals[GM_KSTAT_SGD_STREAM].work[0] = c[11];
    totals[GM_KSTAT_SGD_STREAM].work[1] = c[12];
    totals[GM_KSTAT_SGD_STREAM].work[2] = c[KS_STREAM_EE];
    totals[GM_KSTAT_INSERT].work[0] = ins_pairs;
    totals[GM_KSTAT_INSERT].work[1] = c[5];
    totals[GM_KSTAT_INSERT].work[2] = c[6];
    totals[GM_KSTAT_GRAPH].launches = graph_replays;
    totals[GM_KSTAT_GRAPH].work[0] = graph_instantiations;
  }

  void log(const char *msg) const {
    if (p.verbose) {
      fprintf(stderr, "[gpumap_amd] %s\n", msg);
      fflush(stderr);
    }
  }

  // ------------------------------------------------------------- append
  void append(const float *x, uint64_t nr) {
    if (nr == 0)
      return;
    if (n + nr > N)
      throw Fail(GM_ERR_CAPACITY, "Exceeding maximum number of rows in storage");
    ensure_stage(nr);
    std::memcpy(h_stage, x, nr * D * sizeof(float));
    HIPC(hipMemcpy2DAsync(X + n * Dp, Dp * sizeof(float), h_stage,
                          D * sizeof(float), D * sizeof(float), nr,
                          hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
    for (uint64_t r = 0; r < nr; ++r) {
      insertion_queue.q.push_back((uint32_t)(n + r));
      embedding_new_queue.q.push_back((uint32_t)(n + r));
    }
    n += nr;
  }

  void append_device(const float *x, uint64_t nr) {
    if (nr == 0)
      return;
    if (n + nr > N)
      throw Fail(GM_ERR_CAPACITY, "Exceeding maximum number of rows in storage");
    HIPC(hipMemcpy2DAsync(X + n * Dp, Dp * sizeof(float), x, D * sizeof(float),
                          D * sizeof(float), nr, hipMemcpyDeviceToDevice, s));
    HIPC(hipStreamSynchronize(s));
    for (uint64_t r = 0; r < nr; ++r) {
      insertion_queue.q.push_back((uint32_t)(n + r));
      embedding_new_queue.q.push_back((uint32_t)(n + r));
    }
    n += nr;
  }

  // ascending rows in [0, hi) with flag set or in [f0, f1) into `list`, their
  // count into scal[S_TOTAL]; clears the flags
  void compact_flags(uint64_t hi, uint64_t f0 = 0, uint64_t f1 = 0) {
    const unsigned nblk = blocks_for(hi, 1024);
    k_flag_count<<<nblk, 256, 0, s>>>(flag, 0, hi, bcount, f0, f1);
    LAUNCH_CHECK();
    k_scan_small<<<1, 1024, 0, s>>>(bcount, nblk, scal + S_TOTAL);
    LAUNCH_CHECK();
    k_flag_write<<<nblk, 1024, 0, s>>>(flag, 0, hi, bcount, list, f0, f1);
    LAUNCH_CHECK();
  }

  // compact_flags with the list copied to the host
  uint64_t collect_flags(uint64_t hi, std::vector<uint32_t> &out) {
    out.clear();
    if (hi == 0)
      return 0;
    compact_flags(hi);
    HIPC(hipMemcpyAsync(h_scal, scal + S_TOTAL, sizeof(uint32_t),
                        hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    const uint64_t cnt = h_scal[0];
    if (cnt) {
      HIPC(hipMemcpyAsync(h_list, list, cnt * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
      HIPC(hipStreamSynchronize(s));
      out.assign(h_list, h_list + cnt);
    }
    return cnt;
  }

  // public top-k rows for `rows_` (host list); returns fresh edge count
  uint64_t sync_rows(const std::vector<uint32_t> &rows_) {
    if (rows_.empty())
      return 0;
    std::memcpy(h_list, rows_.data(), rows_.size() * sizeof(uint32_t));
    HIPC(hipMemcpyAsync(rows, h_list, rows_.size() * sizeof(uint32_t),
                        hipMemcpyHostToDevice, s));
    HIPC(hipMemsetAsync(fresh, 0, sizeof(unsigned long long), s));
    k_sync_rows<<<blocks_for(rows_.size() * 32, 256), 256, 0, s>>>(
        pool, rows, rows_.size(), nullptr, pidx, pdist, k, fresh, pmax,
        euclid ? 1 : 0, sigma, rho);
    LAUNCH_CHECK();
    unsigned long long f = 0;
    HIPC(hipMemcpyAsync(h_scal, fresh, sizeof(f), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    std::memcpy(&f, h_scal, sizeof(f));
    return f;
  }

  // ---------------------------------------------------------- insertion
  // reference Impl::insert (impl.cuh:220-253) + run_insertion
  // (insertion_kernel.cuh:926-1095), graph kept on the device.  The rows to
  // publish (the batch and every earlier row whose pool changed) are found
  // and synced on the device; the host learns them for its update queue
  // from a copy made behind the kernels (insert_finish).
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
    if (!euclid) {
      k_prep_rows<<<(unsigned)B, 64, 0, s>>>(X, Dp, D, b0, B);
      LAUNCH_CHECK();
    }
    if (p.knn_mode == GM_KNN_NNDESCENT_BF16) {
      k_join_prep<<<(unsigned)B, 64, 0, s>>>(X, Dp, Db, Xs, nrm, b0, B);
      LAUNCH_CHECK();
    }
    if (nndescent())
      insert_nndescent(b0, B, window_start, seed);
    else
      insert_exact(b0, B);
    // the batch rows are published whether or not their pools were flagged:
    // the compacted list is the earlier flagged rows (ascending), then
    // b0 .. b0 + B - 1 (the batch rows are the newest rows, nothing after
    // them has a pool yet)
    compact_flags(n, b0, b0 + B);
    const unsigned sgrid = (unsigned)std::min<uint64_t>(
        blocks_for(n * 32, 256), 8 * (uint64_t)num_cu);
    k_sync_rows<<<sgrid, 256, 0, s>>>(pool, list, n, scal + S_TOTAL, pidx,
                                      pdist, k, nullptr, pmax, euclid ? 1 : 0,
                                      sigma, rho);
    LAUNCH_CHECK();
    k_init_embedding<<<blocks_for(B, 64), 64, 0, s>>>(emb, c, pidx, pdist, k,
                                                      b0, B);
    LAUNCH_CHECK();
    const uint64_t pre = std::min<uint64_t>(flag_prefix, n);
    HIPC(hipMemcpyAsync(h_flagged + N, scal + S_TOTAL, sizeof(uint32_t),
                        hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(h_flagged, list, pre * sizeof(uint32_t),
                        hipMemcpyDeviceToHost, s));
    pins = PendingInsert{true, b0, B, pre};
    return (uint32_t)B;
  }

  // Host side of the last insertion once the stream has caught up: the
  // batch rows, then the earlier rows whose pools changed (ascending), go to
  // the update queue, as the reference's insert pushes them.
  void insert_finish() {
    if (tune_pend.on) {
      tune_pend.on = false;
      float ms = 0.0f;
      HIPC(hipEventSynchronize(tune_pend.b));
      HIPC(hipEventElapsedTime(&ms, tune_pend.a, tune_pend.b));
      tune_report(tune_pend.key, tune_pend.trial, ms);
      ev_free.push_back(tune_pend.a);
      ev_free.push_back(tune_pend.b);
    }
    if (!pins.on)
      return;
    pins.on = false;
    const uint64_t cnt = h_flagged[N];
    if (cnt < pins.B)
      throw Fail(GM_ERR_HIP, "insertion published fewer rows than its batch");
    const uint64_t nold = cnt - pins.B;
    if (nold > pins.pre) {
      HIPC(hipMemcpyAsync(h_flagged + pins.pre, list + pins.pre,
                          (nold - pins.pre) * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
      HIPC(hipStreamSynchronize(s));
    }
    // next prefix: room for half again as many rows as this one needed
    flag_prefix = std::max<uint64_t>(1024, nold + nold / 2);
    for (uint64_t r = 0; r < pins.B; ++r)
      update_queue.push((uint32_t)(pins.b0 + r));
    for (uint64_t i = 0; i < nold; ++i)
      update_queue.push(h_flagged[i]);
  }

  // NN-Descent insertion (the reference's method)
  void insert_nndescent(uint64_t b0, uint64_t B, uint32_t window_start,
                        uint32_t seed) {
    ensure_ins(B);
    k_random_ids<<<blocks_for(B, 64), 64, 0, s>>>(ins_ids, ins_cnt, b0, B, n,
                                                  window_start, seed);
    LAUNCH_CHECK();
    k_init_dist<<<blocks_for(B * W, 256), 256, 0, s>>>(X, Dp, pool, ins_ids,
                                                       ins_cnt, b0, B);
    LAUNCH_CHECK();
    ev_start(GM_KSTAT_INSERT);
    for (int it = 0; it < 10; ++it) {
      k_ins_candidates<<<(unsigned)B, 256, 0, s>>>(X, Dp, pool, b0, B, k, cand,
                                                   cdist, ccount);
      LAUNCH_CHECK();
      k_ins_offers<<<(unsigned)B, 256, 0, s>>>(pool, flag, cand, cdist, ccount,
                                               b0, B, k);
      LAUNCH_CHECK();
    }
    ev_stop(20);
  }

  // A of the BF16 screens' bound |d~ - d*| <= A s (derivation below, in
  // insert_exact) for db padded features: norms, s and the fma, the hi/lo
  // split, and the summation of BF16_GROUP_STEPS-step groups on the matrix
  // core plus the round-to-nearest group adds
  static double bf16_screen_A(uint64_t db) {
    const double u24 = std::ldexp(1.0, -24);
    const double nsteps_d = (double)db / 16.0;
    const double gnorm = (8.0 * (double)((db + 511) / 512) + 6.0) * u24;
    const double G = (double)BF16_GROUP_STEPS;
    return gnorm + 3.02 * u24 + 3.01 * std::ldexp(1.0, -16) +
           1.01 * (96.0 * G + std::ceil(nsteps_d / G) + 1.0) * u24;
  }

  // exact insertion by distance tiles (exact_kernels.hip, mfma_kernels.hip)
  void insert_exact(uint64_t b0, uint64_t B) {
    const uint64_t row_tiles = (B + 63) / 64;
    const uint64_t col_tiles = (n + 63) / 64;
    int nsplit = 1;
    while (nsplit < 128 && row_tiles * (uint64_t)nsplit < 1024 &&
           (uint64_t)nsplit * 2 <= col_tiles)
      nsplit *= 2;
    // BF16 screening: 128 x 128 block tiles; the heuristic takes about two
    // blocks per CU, the tuner times its neighbours (tune_choose)
    const uint64_t brt = (B + 127) / 128, bct = (n + 127) / 128;
    int bsplit = 1;
    ScreenShape shape{1, false};
    TuneKey tkey{};
    int trial = -1;
    if (screened()) {
      ensure_bf16_batch(B);
      while (bsplit < 128 && brt * (uint64_t)bsplit * 2 <= 2 * (uint64_t)num_cu &&
             (uint64_t)bsplit * 2 <= bct)
        bsplit *= 2;
      int max_split = 1;
      while (max_split < 128 && (uint64_t)max_split * 2 <= bct)
        max_split *= 2;
      int nlog4 = 0;
      while (nlog4 < 31 && (4ull << (2 * nlog4)) <= n)
        ++nlog4;
      tkey = TuneKey{device, (uint64_t)p.knn_mode, Db, nlog4, brt};
      shape = tune_choose(tkey, bsplit, max_split, lbm != nullptr, &trial);
      bsplit = shape.bsplit;
    }
    const uint64_t need = B * (uint64_t)std::max(nsplit, bsplit) * W;
    if (need > part_cap) {
      if (part)
        (void)hipFree(part);
      part = nullptr;
      part_cap = 0;
      part = dalloc<uint64_t>(need);
      part_cap = need;
    }
    ins_pairs += B * n;
    ev_start(GM_KSTAT_INSERT);
    const dim3 grid((unsigned)row_tiles, (unsigned)nsplit);
    const unsigned *guard = nullptr;
    if (screened()) {
      const bool f8 = p.knn_mode == GM_KNN_EXACT_FP8;
      // Error of d~ = s - 2 P against the exact squared distance d*, with
      // u = 2^-24, s = fl(n_u + n_j), P the screened dot product:
      //   norms n_u, n_j: each lane's 8 ceil(Db / 512) terms in sequence,
      //     then a 6-level tree: gamma_n = (8 ceil(Db / 512) + 6) u relative;
      //   s: u;  d~ (one fma): 2.02 u (|s - 2P| <= 2.02 s);
      //   hi/lo split with round-to-nearest BF16: 3.01 2^-16 of sum |u_i j_i|
      //     (<= s / 2) in the dot, so 3.01 2^-16 s in d~;
      //   summation: each group of G = BF16_GROUP_STEPS feature steps' 48 G
      //     exact products summed on the matrix core (at most 2u per
      //     addition), the ceil(nsteps / G) groups and one final partial group
      //     added in round-to-nearest FP32: (96 G + ceil(nsteps / G) + 1) u of
      //     sum |products|;
      // so |d~ - d*| <= A s.  The canonical distance d (one FMA chain over D
      // rounded differences) is within (D + 2) u relative of d*.  coef = 1.25
      // A and cfac = 1.25 (D + 2) u; the factor 1.25 covers second-order
      // terms and the roundings of the bound arithmetic itself.
      //
      // FP8 (E4M3 hi/lo of a = s x, s a power of two with max|a| in
      // (224, 448]): per feature |a - hi| <= 2^-4 |a| + 2^-10 and the
      // residual e = a - hi - lo has |e| <= g |a| + h (g = 2^-8,
      // h = 1.0625 2^-10), |lo| <= al |a| + be (al = 1.0625 2^-4,
      // be = 2.0625 2^-10).  a.b - P' = lo_a lo_b + (hi_a + lo_a) e_b +
      // e_a (hi_b + lo_b) + e_a e_b, so summed over the features
      //   |a.b - P'| <= C1 sum|a||b| + C2 (sum|a| + sum|b|) + C3 D,
      // C1 = al^2 + 2 (1 + g) g + g^2, C2 = al be + (1 + g) h + 2 g h,
      // C3 = be^2 + 3 h^2; with sum|a| <= sqrt(D) |a|, 1 / s <= |x| / 224
      // and sum|a||b| <= |a||b| the dot's split error is at most
      // (C1 + 2 C2 sqrt(D) / 224 + C3 D / 224^2) |u||j| <= that times s / 2.
      // Summation: each group of G = FP8_GROUP_STEPS steps' 192 G exact
      // products (their magnitudes sum to at most 1.28 |u||j|) at 2u per
      // addition, then the group adds: 1.28 (384 G + ceil(nsteps / G) + 1) u
      // of |u||j|.  The scalings by 1 / s are exact.  The split and the
      // summation terms replace the BF16 ones in A.
      const double u24 = std::ldexp(1.0, -24);
      const double step = f8 ? 64.0 : 16.0;
      const double nsteps_d = (double)Db / step;
      const double gnorm = (8.0 * (double)((Db + 511) / 512) + 6.0) * u24;
      double A;
      if (!f8) {
        A = bf16_screen_A(Db);
      } else {
        const double G = (double)FP8_GROUP_STEPS;
        const double g = std::ldexp(1.0, -8), h = 1.0625 * std::ldexp(1.0, -10);
        const double al = 1.0625 * std::ldexp(1.0, -4);
        const double be = 2.0625 * std::ldexp(1.0, -10);
        const double C1 = al * al + 2.0 * (1.0 + g) * g + g * g;
        const double C2 = al * be + (1.0 + g) * h + 2.0 * g * h;
        const double C3 = be * be + 3.0 * h * h;
        const double Dd = (double)Db;
        const double split =
            C1 + 2.0 * C2 * std::sqrt(Dd) / 224.0 + C3 * Dd / (224.0 * 224.0);
        A = gnorm + 3.02 * u24 + split +
            1.01 * 1.28 * (384.0 * G + std::ceil(nsteps_d / G) + 1.0) * u24;
      }
      const float coef = (float)(1.25 * A);
      const float cfac = (float)(1.25 * (double)(D + 2) * u24);
      const int nsteps = (int)(Db / (uint64_t)step);
      const unsigned blocks = (unsigned)(brt * bsplit);
      float *part_f = reinterpret_cast<float *>(part);
      // the lower-bound store only when the survivors come from its scan
      float *const lbs = shape.lbscan ? lbm : nullptr;
      uint8_t *const xs8 = reinterpret_cast<uint8_t *>(Xs);
      uint8_t *const xb8 = reinterpret_cast<uint8_t *>(Xb);
      if (f8)
        k_fp8_prep<<<(unsigned)B, 64, 0, s>>>(X, Dp, Db, xs8, xb8, nrm, isc,
                                              flag, rcnt, qctl, b0, B);
      else
        k_bf16_prep<<<(unsigned)B, 64, 0, s>>>(X, Dp, Db, Xs, Xb, nrm, flag,
                                               rcnt, qctl, b0, B);
      LAUNCH_CHECK();
      if (trial >= 0) {
        tune_pend = TunePending{true, tkey, trial, ev_get(), ev_get()};
        HIPC(hipEventRecord(tune_pend.a, s));
      }
      if (f8)
        k_fp8_screen<1><<<blocks, 256, 0, s>>>(
            xs8, xb8, nsteps, nrm, isc, pmax, tau, b0, B, n, bsplit, (int)brt,
            part_f, rlist, rcnt, rcap, qctl, coef, cfac, lbs);
      else
        k_bf16_screen<1><<<blocks, 256, 0, s>>>(
            Xs, Xb, nsteps, nrm, pmax, tau, b0, B, n, bsplit, (int)brt, part_f,
            rlist, rcnt, rcap, qctl, coef, cfac, lbs);
      LAUNCH_CHECK();
      k_tau_select<<<(unsigned)B, 256, 0, s>>>(part_f, bsplit, tau, B);
      LAUNCH_CHECK();
      if (lbs) {
        // waves = row groups x column ranges, about 16 waves per CU
        const uint64_t groups = (B + 63) / 64;
        const uint64_t ranges =
            std::max<uint64_t>(1, 16 * (uint64_t)num_cu / groups);
        const uint64_t cpw = std::max<uint64_t>(8, (n + ranges - 1) / ranges);
        const uint64_t waves = groups * ((n + cpw - 1) / cpw);
        k_bf16_lbscan<<<(unsigned)((waves + 3) / 4), 256, 0, s>>>(
            lbs, B, n, b0, tau, pmax, rlist, rcnt, rcap, qctl, cpw);
      } else if (f8) {
        k_fp8_screen<2><<<blocks, 256, 0, s>>>(
            xs8, xb8, nsteps, nrm, isc, pmax, tau, b0, B, n, bsplit, (int)brt,
            part_f, rlist, rcnt, rcap, qctl, coef, cfac, nullptr);
      } else {
        k_bf16_screen<2><<<blocks, 256, 0, s>>>(
            Xs, Xb, nsteps, nrm, pmax, tau, b0, B, n, bsplit, (int)brt,
            part_f, rlist, rcnt, rcap, qctl, coef, cfac, nullptr);
      }
      LAUNCH_CHECK();
      k_bf16_rerank<<<dim3((unsigned)B, std::min(rcap, 2048u) / 64), 64, 0,
                      s>>>(
          X, Dp, rlist, rcnt, rcap, qctl, fkeys, pool, pmax, flag, b0, tau, B,
          prof ? kstats : nullptr);
      LAUNCH_CHECK();
      k_bf16_select<<<(unsigned)B, 64, rcap * sizeof(uint64_t), s>>>(
          fkeys, rcnt + B, rcap, qctl, pool, b0, prof ? kstats : nullptr);
      LAUNCH_CHECK();
      if (trial >= 0)
        HIPC(hipEventRecord(tune_pend.b, s));
      // the FP32 tiles below run only if the survivor queue overflowed
      guard = qctl;
    }
    k_exact_tiles<<<grid, 256, 0, s>>>(X, Dp, pool, pmax, b0, B, n, k, nsplit,
                                       part, flag, 0, guard);
    LAUNCH_CHECK();
    k_exact_merge<<<(unsigned)B, 256, 0, s>>>(part, nsplit, pool, b0, B, flag,
                                              guard);
    LAUNCH_CHECK();
    ev_stop(1);
  }

  // the row-owner SGD's scratch for COOs of up to `slots` slots over the
  // session's N rows
  void ensure_pull(uint64_t slots) {
    if (pullf.slots >= slots && pullf.rows >= N)
      return;
    for (void **q :
         {(void **)&pullf.cnt, (void **)&pullf.off, (void **)&pullf.bsum,
          (void **)&pullf.erc, (void **)&pullf.esch, (void **)&pullf.epc,
          (void **)&pullf.d1}) {
      if (*q)
        (void)hipFree(*q);
      *q = nullptr;
    }
    pullf.slots = pullf.rows = 0;
    const uint64_t ent = 2 * slots;
    pullf.cnt = dalloc<uint32_t>(N + 1);
    pullf.off = dalloc<uint32_t>(N + 1);
    pullf.bsum = dalloc<uint32_t>((N + 1023) / 1024 + 1);
    pullf.erc = dalloc<int2>(ent);
    pullf.esch = dalloc<float4>(ent);
    pullf.epc = dalloc<int32_t>(ent * (uint64_t)c);
    pullf.d1 = dalloc<int32_t>(N * 2 * (uint64_t)(c == 3 ? 4 : c));
    pullf.slots = slots;
    pullf.rows = N;
  }

  // batch operand store, tau and survivor queue of the BF16 screening
  void ensure_bf16_batch(uint64_t B) {
    if (B <= xb_cap)
      return;
    for (void *q : {(void *)Xb, (void *)tau, (void *)rlist, (void *)rcnt,
                    (void *)fkeys, (void *)lbm})
      if (q)
        (void)hipFree(q);
    Xb = nullptr;
    tau = nullptr;
    rlist = nullptr;
    rcnt = nullptr;
    fkeys = nullptr;
    lbm = nullptr;
    xb_cap = 0;
    const uint64_t rows_ = (B + 127) / 128 * 128;
    Xb = dalloc<__bf16>(rows_ * Db * 2);
    HIPC(hipMemsetAsync(Xb, 0, rows_ * Db * 2 * sizeof(__bf16), s));
    tau = dalloc<float>(B);
    rlist = dalloc<uint32_t>(B * (uint64_t)rcap);
    rcnt = dalloc<unsigned>(2 * B); // survivors, then forward keys in tau
    fkeys = dalloc<uint64_t>(B * (uint64_t)rcap);
    // the screening kernel addresses each 128-column region of the store
    // with 32-bit offsets: B < 2^22 rows keeps it below 2^31 bytes
    if (N * B * sizeof(float) <= lbm_max_bytes && B < (1ull << 22)) {
      // optional: without it the survivors come from a second tile pass
      if (hipMalloc((void **)&lbm, N * B * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        lbm = nullptr;
      }
    }
    xb_cap = B;
  }

  // ------------------------------------------------------------- update
  // One NN-Descent pass over `b` (reference UpdateAlgo::run,
  // update_kernel.cuh:913-1054); returns the fresh top-k edge count.
  uint64_t nnd_pass(std::vector<uint32_t> b, std::vector<uint32_t> &affected) {
    affected.clear();
    if (!nndescent())
      return 0; // exact pools hold only old entries: nothing to join
    std::sort(b.begin(), b.end());
    b.erase(std::unique(b.begin(), b.end()), b.end());
    const uint64_t B = b.size();
    std::memcpy(h_list, b.data(), B * sizeof(uint32_t));
    HIPC(hipMemcpyAsync(rows, h_list, B * sizeof(uint32_t),
                        hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(owners, rows, B * sizeof(uint32_t),
                        hipMemcpyDeviceToDevice, s));
    h_scal[8] = (uint32_t)B;
    HIPC(hipMemcpyAsync(scal + S_OWNERS, h_scal + 8, sizeof(uint32_t),
                        hipMemcpyHostToDevice, s));
    k_sample<<<blocks_for(B * 32, 256), 256, 0, s>>>(pool, rows, B, fn, fo, fnc,
                                                     foc, newmask, owner_mark);
    LAUNCH_CHECK();
    k_reverse<<<blocks_for(B * 32, 256), 256, 0, s>>>(
        pool, rows, B, newmask, rev_new, rev_old, owner_mark, owners,
        scal + S_OWNERS);
    LAUNCH_CHECK();
    const unsigned jgrid = (unsigned)std::min<uint64_t>(n, 2048);
    ev_start(GM_KSTAT_JOIN);
    if (p.knn_mode == GM_KNN_NNDESCENT_BF16) {
      // the BF16 screen's bound for Db features; cfac covers the canonical
      // distance's own rounding, (D + 2) u relative
      const float coef = (float)(1.25 * bf16_screen_A(Db));
      const float cfac =
          (float)(1.25 * (double)(D + 2) * std::ldexp(1.0, -24));
      k_join_bf16<<<jgrid, 256, 0, s>>>(
          X, Dp, Xs, Db, nrm, pool, fn, fo, fnc, foc, rev_new, rev_old, owners,
          scal + S_OWNERS, flag, k, coef, cfac, kstats);
    } else {
      k_join<<<jgrid, 256, 0, s>>>(X, Dp, pool, fn, fo, fnc, foc, rev_new,
                                   rev_old, owners, scal + S_OWNERS, flag, k,
                                   kstats);
    }
    LAUNCH_CHECK();
    ev_stop(1);
    k_join_reset<<<blocks_for(n * W, 256), 256, 0, s>>>(
        owners, scal + S_OWNERS, rev_new, rev_old, owner_mark, fnc, foc);
    LAUNCH_CHECK();
    collect_flags(n, affected);
    return sync_rows(affected);
  }

  // reference Impl::update (impl.cuh:255-333)
  uint32_t update(uint32_t n_ops, uint64_t &changed_out) {
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
      const uint32_t bsz = (uint32_t)b.size();
      changed_out += nnd_pass(std::move(b), affected);
      left -= bsz;
      for (uint32_t r : affected)
        update_queue.push(r);
      if (update_queue.size() < thr) {
        update_completed = (n == N);
        break;
      }
    }
    return n_ops - left;
  }

  // --------------------------------------------------------- fuzzy set
  // Builds the trimmed COO of the batch rows in dbatch[0, nb); leaves nnz in
  // scal[S_NNZ] and the maximum weight in scal[S_MAXW].  A batch of fewer
  // than n rows is marked in bmark (k_reposition, or the caller, clears it).
  // nr: the row count as the kernels read it (a value, or S_NROWS when the
  // launches are captured into a graph)
  void build_fuzzy(uint64_t nb, NRows nr) {
    uint8_t *mark = nb < n ? bmark : nullptr;
    k_mean_partial<<<NPART, 256, 0, s>>>(pidx, pdist, nr, k, euclid ? 1 : 0,
                                         psum, pcnt);
    LAUNCH_CHECK();
    static_assert(S_MIN == S_MAXW + 1 && S_MAX == S_MAXW + 2,
                  "k_mean_final resets three consecutive scalars");
    k_mean_final<<<1, 256, 0, s>>>(psum, pcnt, NPART, mean, scal + S_MAXW,
                                   dbatch, nb, mark);
    LAUNCH_CHECK();
    const uint64_t nd = nb * (uint64_t)k;
    k_membership<<<blocks_for(nd, 256), 256, 0, s>>>(
        pidx, pdist, nr, k, euclid ? 1 : 0, sigma, rho, mean, dbatch, nb, mark,
        w, wflags, scal + S_MAXW);
    LAUNCH_CHECK();
    const unsigned nblk = blocks_for(nd, 1024);
    k_coo_count<<<nblk, 256, 0, s>>>(w, wflags, nd, scal + S_MAXW,
                                     (int)p.max_epoch, bcount);
    LAUNCH_CHECK();
    k_scan_small<<<1, 1024, 0, s>>>(bcount, nblk, scal + S_NNZ);
    LAUNCH_CHECK();
    k_coo_write<<<nblk, 1024, 0, s>>>(pidx, k, dbatch, w, wflags, nd,
                                    scal + S_MAXW, (int)p.max_epoch, bcount,
                                    head, tail, wout);
    LAUNCH_CHECK();
  }

  // SGD on the COO in head/tail/wout with nnz in scal[S_NNZ] and the
  // maximum weight in scal[S_MAXW] (reference run_optimize_layout,
  // embedding_kernel.cuh:726-798).  `cap` bounds nnz for the launch grids.
  // The SGD's schedule and rounding grid: epochs_per_sample and epoch state
  // per edge, head degrees, the grid from the maximum degree.  deg is cleared
  // by the caller's k_minmax_col0 on the embedding path; k_sgd_init clears
  // maxdeg and the barrier words, and k_max_int counts its finished blocks in
  // gbar[3] (unused by the SGD kernels).
  // dev_rows: the row count read from S_NROWS (a captured launch; the grid
  // then covers the capacity)
  void sgd_setup(uint64_t nrows, uint64_t cap, int n_epochs, float nsr_inv,
                 float alpha0, bool deg_zeroed, bool dev_rows = false) {
    if (!deg_zeroed)
      HIPC(hipMemsetAsync(deg, 0, nrows * sizeof(int), s));
    k_sgd_init<<<blocks_for(cap, 256), 256, 0, s>>>(
        wout, head, scal + S_NNZ, scal + S_MAXW, n_epochs, nsr_inv, eps, nxt,
        nneg, deg, maxdeg, gbar, 4 + SGD_MAX_WAVES);
    LAUNCH_CHECK();
    k_max_int<<<std::min<unsigned>(blocks_for(dev_rows ? N : nrows, 256), 64),
                256, 0, s>>>(deg,
                             dev_rows ? NRows{scal + S_NROWS, 0}
                                      : NRows{nullptr, nrows},
                             maxdeg, alpha0, rounding, gbar + 3);
    LAUNCH_CHECK();
  }

  // sgd_mode: SGD_AUTO (the kernel that fits), SGD_PER_EPOCH (one launch per
  // epoch) or SGD_STREAM (the streaming cooperative kernel at any size)
  // set_up: the schedule was prepared already (sgd_setup)
  void run_sgd(float *e, uint64_t nrows, uint64_t cap, int n_epochs, int nsr,
               float a, float b, float alpha0, float gamma, uint64_t seed,
               int sgd_mode = 0, bool deg_zeroed = false,
               bool set_up = false) {
    if (cap == 0 || n_epochs <= 0)
      return;
    const float nsr_inv = 1.0f / (float)nsr;
    if (!set_up)
      sgd_setup(nrows, cap, n_epochs, nsr_inv, alpha0, deg_zeroed);
    unsigned long long *st = prof ? kstats : nullptr;
    const bool want_pull =
        sgd_fix && (sgd_mode == SGD_PULL ||
                    (sgd_pull && sgd_mode == SGD_AUTO && cap >= PULL_MIN_SLOTS));
    if (want_pull)
      ensure_pull(std::max<uint64_t>(cap, 2 * N * (uint64_t)k));
    ev_start(GM_KSTAT_SGD);
    const bool fused = sgd_launches(
        c, e, head, tail, scal + S_NNZ, (uint32_t)nrows, eps, nxt, nneg,
        n_epochs, alpha0, gamma, a, b, nsr_inv, seed, rounding, gbar, st, cap,
        num_cu, s, sgd_mode, &sgd_variant, p.verbose != 0,
        FixFrame{sgd_fix ? fxp : nullptr}, want_pull ? &pullf : nullptr);
    // the streaming and row-owner kernels count as the large-set SGD
    const bool large = fused && (sgd_variant == 1 || sgd_variant == 3);
    if (fused) {
      fused_used = true;
      if (sgd_variant == 2 || sgd_variant == 4)
        ++sgd_waves_launches;
      if (prof && large)
        pending.back().which = GM_KSTAT_SGD_STREAM;
    }
    ev_stop(fused ? 1 : (uint64_t)n_epochs);
    if (prof) {
      k_add_edge_epochs<<<1, 64, 0, s>>>(
          scal + S_NNZ, n_epochs, kstats + (large ? KS_STREAM_EE : KS_SGD_EE));
      LAUNCH_CHECK();
    }
  }

  // Enqueue `body`'s launches on the stream: directly, or as the executable
  // graph `gx`.  While the host values the launches depend on (`key`) stay
  // those `gx` was made with, `gx` is relaunched as it is; otherwise `body`
  // is captured and patched into `gx` (hipGraphExecUpdate), or instantiated
  // anew when its shape changed.
  template <typename F>
  void as_graph(hipGraphExec_t &gx, GraphKey &last, const GraphKey &key,
                F &&body) {
    if (!use_graph) {
      body();
      return;
    }
    if (gx && last == key) {
      HIPC(hipGraphLaunch(gx, s));
      ++graph_replays;
      return;
    }
    last = GraphKey{};
    hipGraph_t g = nullptr;
    HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      body();
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);
      if (g)
        (void)hipGraphDestroy(g);
      throw;
    }
    HIPC(hipStreamEndCapture(s, &g));
    bool ok = false;
    if (gx) {
      hipGraphNode_t err_node = nullptr;
      hipGraphExecUpdateResult res = hipGraphExecUpdateError;
      ok = hipGraphExecUpdate(gx, g, &err_node, &res) == hipSuccess &&
           res == hipGraphExecUpdateSuccess;
      if (!ok) {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(gx);
        gx = nullptr;
      }
    }
    if (!gx) {
      const hipError_t e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPC(e);
      ++graph_instantiations;
    } else {
      (void)hipGraphDestroy(g);
    }
    last = key;
    HIPC(hipGraphLaunch(gx, s));
    ++graph_replays;
  }

  // -------------------------------------------------------------- embed
  // reference Impl::embed (impl.cuh:335-500)
  uint32_t embed(uint32_t n_ops) {
    const uint32_t nb = embed_begin(n_ops, nullptr);
    embed_end();
    return nb;
  }

  // The embedding op up to its SGD.  With `job` null the SGD runs here on
  // the session's stream; otherwise, when the op's edge set fits the
  // per-wave kernel, its schedule is set up and `job` describes its SGD for a
  // batched launch (job->nwaves = 0 when the SGD ran here after all).  The
  // reposition follows in embed_end, once the SGD is done.
  uint64_t pend_nb = 0;
  // `before_own_sgd` (may be empty) runs right before an SGD launched on the
  // session's own stream, the one place it is known that one will be
  uint32_t embed_begin(uint32_t n_ops, SgdJob *job,
                       const std::function<void()> &before_own_sgd = {}) {
    if (job)
      job->nwaves = 0;
    pend_nb = 0;
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
        size_t cur =
            embedding_cursor == SIZE_MAX ? 0 : (embedding_cursor + 1) % n;
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
    const uint64_t nb = batch.size();
    std::memcpy(h_list, batch.data(), nb * sizeof(uint32_t));
    if (embedding_sweeps_done > 2)
      embedding_completed = update_completed;
    // the launch grids are sized for the most edges the batch can give (the
    // union keeps at most both directions of each of its nb k entries); the
    // kernels read the edge count itself from the device
    const uint64_t cap = 2 * nb * (uint64_t)k;
    const uint64_t waves = (cap + 63) / 64;
    const bool batched =
        job && cap > 0 && p.max_epoch > 0 && waves <= SGD_MAX_WAVES;
    const bool own_sgd = !batched && cap > 0 && p.max_epoch > 0;
    const float nsr_inv = 1.0f / (float)p.negative_sample_rate;
    // with graphs the kernels read the row count from S_NROWS (copied from
    // pinned memory by the sequence's first node) and cover the capacity
    const NRows nr = use_graph ? NRows{scal + S_NROWS, 0} : NRows{nullptr, n};
    const uint64_t gn = use_graph ? N : n;
    h_scal[16] = (uint32_t)n;
    const GraphKey key{nb, embedding_sweeps_done, nb < n, batched || own_sgd,
                       true};
    as_graph(gx_pre, gk_pre, key, [&] {
      if (use_graph)
        HIPC(hipMemcpyAsync(scal + S_NROWS, h_scal + 16, sizeof(uint32_t),
                            hipMemcpyHostToDevice, s));
      HIPC(hipMemcpyAsync(dbatch, h_list, nb * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
      build_fuzzy(nb, nr);
      // noise (reference add_embedding_noise, embedding_kernel.cuh:953-1007);
      // the column-0 range accumulators were reset by k_mean_final, and the
      // range pass also clears the SGD's head degrees
      k_minmax_col0<<<std::min<unsigned>(blocks_for(gn, 256), 64), 256, 0,
                      s>>>(emb, nr, c, scal + S_MIN, scal + S_MAX, deg);
      LAUNCH_CHECK();
      k_noise<<<blocks_for(nb * c, 256), 256, 0, s>>>(
          emb, c, dbatch, nb, scal + S_MIN, scal + S_MAX, 0.001f,
          42ull ^ (uint64_t)embedding_sweeps_done);
      LAUNCH_CHECK();
      if (batched || own_sgd)
        sgd_setup(n, cap, (int)p.max_epoch, nsr_inv, p.alpha, true,
                  use_graph);
    });
    pend_nb = nb;
    if (batched) {
      *job = SgdJob{emb,     head,         tail,
                    scal + S_NNZ, eps,     rounding,
                    prof ? kstats : nullptr, 42ull, (uint32_t)n,
                    (int)p.max_epoch, p.alpha, 1.0f, p.a, p.b, nsr_inv, 0u,
                    (unsigned)waves, sgd_fix ? fxp : nullptr};
      return (uint32_t)nb;
    }
    if (before_own_sgd)
      before_own_sgd();
    run_sgd(emb, n, cap, (int)p.max_epoch, (int)p.negative_sample_rate, p.a,
            p.b, p.alpha, 1.0f, 42ull, SGD_AUTO, true, true);
    if (fused_used)
      HIPC(hipMemcpyAsync(h_scal + 12, gbar + 2, 4, hipMemcpyDeviceToHost, s));
    return (uint32_t)nb;
  }

  // reposition of the rows outside the op's batch (reference
  // heuristic_neighbor_avg_kernel :811-894, impl.cuh:422-449), from a
  // snapshot so that no row reads a neighbour's new value; the batch rows
  // were marked by build_fuzzy and k_reposition clears the marks
  void embed_end() {
    const uint64_t nb = pend_nb;
    pend_nb = 0;
    if (nb > 0 && nb < n) {
      // the row count from S_NROWS, set by the op's first sequence
      const NRows nr =
          use_graph ? NRows{scal + S_NROWS, 0} : NRows{nullptr, n};
      const uint64_t gn = use_graph ? N : n;
      as_graph(gx_post, gk_post, GraphKey{0, 0, true, false, true}, [&] {
        k_copy_rows<<<(unsigned)std::min<uint64_t>(blocks_for(gn * c, 256),
                                                   8 * (uint64_t)num_cu),
                      256, 0, s>>>(emb, snap, nr, c);
        LAUNCH_CHECK();
        k_reposition<<<blocks_for(gn, 128), 128, 0, s>>>(
            pidx, pdist, nr, k, euclid ? 1 : 0, sigma, rho, mean, bmark, snap,
            emb, c);
        LAUNCH_CHECK();
      });
    }
  }

  // reference Impl::run (impl.cuh:146-218)
  void run(uint32_t ni, uint32_t nu, uint32_t ne, gm_result *r) {
    using clk = std::chrono::steady_clock;
    std::memset(r, 0, sizeof(*r));
    if (!insertion_completed) {
      std::shared_lock<DevLock> dl(g_dev_mu);
      check_live();
      auto t0 = clk::now();
      r->n_insertion_ops = insert(ni);
      stream_wait(s);
      insert_finish();
      r->insertion_time = std::chrono::duration<float>(clk::now() - t0).count();
      insertion_completed = insertion_queue.size() == 0 && n == N;
    }
    if (!(nu == 0 || update_completed)) {
      std::shared_lock<DevLock> dl(g_dev_mu);
      check_live();
      uint64_t ch = 0;
      auto t0 = clk::now();
      r->n_update_ops = update(nu, ch);
      stream_wait(s);
      r->update_time = std::chrono::duration<float>(clk::now() - t0).count();
      r->n_update_changed_edges = ch;
    }
    r->embedding_pending_edges = n;
    if (embedding_due(ne)) {
      std::unique_lock<DevLock> dl(g_dev_mu);
      std::lock_guard<std::mutex> lk(g_coop_mu);
      check_live();
      auto t0 = clk::now();
      r->n_embedding_ops = embed(ne);
      stream_wait(s);
      r->embedding_time = std::chrono::duration<float>(clk::now() - t0).count();
    }
    finish_result(r);
  }

  // a session freed by gm_shutdown (or gm_release) between two phases of a
  // call in flight on another thread takes no further device work
  void check_live() const {
    if (release_pending)
      throw Fail(GM_ERR_STATE, "the session was released");
  }

  bool embedding_due(uint32_t ne) const {
    return !(ne == 0 || n == 0 || embedding_completed);
  }

  // the result's state fields once the cycle's work has completed
  void finish_result(gm_result *r) {
    r->n_inserted_instances = n;
    r->insertion_queue_size = insertion_queue.size();
    r->update_queue_size = update_queue.size();
    r->embedding_queue_size = r->embedding_pending_edges;
    r->insertion_completed = insertion_completed;
    r->update_completed = update_completed;
    r->embedding_completed = embedding_completed;
    if (fused_used) {
      // the timeout flag was copied back behind the SGD launch, before the
      // embedding phase's synchronisation
      fused_used = false;
      if (h_scal[12])
        throw Fail(GM_ERR_HIP, "SGD device-wide barrier timed out");
    }
    fold_events();
  }

  void check_barrier() {
    if (!fused_used)
      return;
    fused_used = false;
    unsigned flagv = 0;
    HIPC(hipMemcpyAsync(h_scal + 12, gbar + 2, 4, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    flagv = h_scal[12];
    if (flagv)
      throw Fail(GM_ERR_HIP, "SGD device-wide barrier timed out");
  }
};

// ---------------------------------------------------------------- DLPack
struct Export {
  DLManagedTensor mt;
  int64_t shape[2];
  int64_t strides[2];
};

// Frees a released session's device memory once only the handle holds it.
// The caller holds e->rel_mu.
void release_if_unshared(Engine *e) {
  if (e->release_pending && !e->released.load() && e->refs.load() == 1) {
    (void)hipSetDevice(e->device);
    e->release();
    e->released.store(true);
  }
}

void engine_unref(Engine *e) {
  int left;
  {
    std::lock_guard<std::mutex> l(e->rel_mu);
    left = e->refs.fetch_sub(1) - 1;
    if (left == 1)
      release_if_unshared(e);
  }
  if (left == 0)
    delete e;
}

void dl_deleter(DLManagedTensor *self) {
  Export *x = reinterpret_cast<Export *>(self);
  engine_unref(static_cast<Engine *>(self->manager_ctx));
  delete x;
}

// ------------------------------------------------------- batched cycles
// One cycle of several sessions at once (gm_run_batch): every session's
// insertion kernels are enqueued before any phase waits, so they overlap on
// the device; the per-cycle SGDs of all sessions whose edge sets fit the
// per-wave kernel run as one cooperative launch of k_sgd_waves_multi, each
// paced only by its own waves; every session's results are those of gm_run
// on its own (the phases' times are the batch's).  The batched launch uses a
// stream, arrival slots and a pinned timeout flag shared per device.
struct BatchCtx {
  int device = -1;
  hipStream_t s = nullptr;
  unsigned *bar = nullptr;       // 4 words, then one slot per wave (the
                                 // launches run in order on s and reuse it)
  unsigned *h_flag = nullptr;    // pinned: each chunk's timeout flag
  int num_cu = 0;
};
static constexpr unsigned BATCH_MAX_WAVES = 2048; // per launch
static constexpr int BATCH_MAX_CHUNKS = 64;
std::mutex g_batch_mu;
std::vector<BatchCtx> g_batch;

BatchCtx &batch_ctx(int device) {
  for (auto &b : g_batch)
    if (b.device == device)
      return b;
  BatchCtx b;
  b.device = device;
  HIPC(hipStreamCreate(&b.s));
  HIPC(hipMalloc((void **)&b.bar, (4 + (size_t)BATCH_MAX_WAVES) *
                                      sizeof(unsigned)));
  HIPC(hipHostMalloc((void **)&b.h_flag, BATCH_MAX_CHUNKS * sizeof(unsigned)));
  HIPC(hipDeviceGetAttribute(&b.num_cu, hipDeviceAttributeMultiprocessorCount,
                             device));
  g_batch.push_back(b);
  return g_batch.back();
}

void run_batch(Engine *const *es, int cnt, const uint32_t *ops,
               gm_result *rs) {
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point t0) {
    return std::chrono::duration<float>(clk::now() - t0).count();
  };
  std::lock_guard<std::mutex> lk(g_batch_mu);
  std::unique_lock<DevLock> dl(g_dev_mu);
  for (int i = 0; i < cnt; ++i)
    es[i]->check_live();
  BatchCtx &bc = batch_ctx(es[0]->device);
  for (int i = 0; i < cnt; ++i)
    std::memset(&rs[i], 0, sizeof(gm_result));

  // Every session's insertion kernels are enqueued first.  The sessions are
  // then finished in order: wait for its insertion, its host bookkeeping and
  // update, then its embedding op up to the SGD, so that the host prepares
  // session i's embedding while the device still runs the later sessions'
  // insertions.  A session's phase times run from the batch start to the
  // end of its insertion, over its update, and from the start of its
  // embedding op to the end of the batch.  An SGD too large for the batched
  // kernel runs in embed_begin and completes before the next cooperative
  // launch.
  std::vector<char> did(cnt, 0);
  const auto tb = clk::now();
  for (int i = 0; i < cnt; ++i)
    if (!es[i]->insertion_completed) {
      did[i] = 1;
      rs[i].n_insertion_ops = es[i]->insert(ops[3 * i]);
    }
  std::lock_guard<std::mutex> coop(g_coop_mu);
  std::vector<SgdJob> jobs(cnt);
  std::vector<char> emb(cnt, 0);
  std::vector<clk::time_point> te(cnt);
  int nchunks = 0;
  unsigned *const bar = bc.bar;
  // completion of the last batched launch so far: an SGD that a session runs
  // on its own stream (too large for the batched kernel) waits for it, since
  // two grid-wide SGD kernels on the device at once could each hold CUs the
  // other's waiting blocks need
  hipEvent_t last_done = nullptr;
  // the launches' completion events go back to their pools when the cycle
  // ends (also when it throws), so none is re-recorded while waited on
  struct DoneEvents {
    std::vector<std::pair<Engine *, hipEvent_t>> v;
    ~DoneEvents() {
      for (auto &p : v)
        p.first->ev_free.push_back(p.second);
    }
  } dones;
  // One batched SGD launch over `members` (one component count, at most
  // SGD_MAX_JOBS jobs and BATCH_MAX_WAVES waves), after their set-up; their
  // streams continue after it.  Launches run in order on the batch stream.
  // A batched launch is a grid-wide kernel whose jobs' waves wait for each
  // other at the epoch boundaries: it starts only once everything the cycle
  // has enqueued so far, on every session's stream, has finished (a launch
  // closed mid-loop would otherwise share the device with the insertions
  // of the sessions not yet reached, whose kernels hold CUs that the
  // launch's waiting waves need).
  auto launch_group = [&](const std::vector<int> &members) {
    if (members.empty())
      return;
    if (nchunks >= BATCH_MAX_CHUNKS)
      throw Fail(GM_ERR_INVALID, "too many sessions in one batched cycle");
    const int c = es[members[0]]->c;
    SgdBatch b{};
    unsigned waves = 0;
    for (int i : members) {
      b.job[b.njobs] = jobs[i];
      b.job[b.njobs].wave0 = waves;
      waves += jobs[i].nwaves;
      ++b.njobs;
    }
    for (int i = 0; i < cnt; ++i) {
      hipEvent_t ev = es[i]->ev_get();
      HIPC(hipEventRecord(ev, es[i]->s));
      HIPC(hipStreamWaitEvent(bc.s, ev, 0));
      dones.v.emplace_back(es[i], ev);
    }
    HIPC(hipMemsetAsync(bar, 0, (4 + (size_t)waves) * sizeof(unsigned), bc.s));
    // the members' row records (the kernel updates those, not the float
    // embeddings), folded back after the launch
    for (int i : members)
      if (jobs[i].rec)
        HIPC(launch_fix(c, 0, jobs[i].emb, jobs[i].n,
                        FixFrame{jobs[i].rec}, jobs[i].rounding_p, bc.num_cu,
                        bc.s));
    std::vector<Engine::Timed> tm;
    for (int i : members)
      if (es[i]->prof) {
        Engine::Timed x{GM_KSTAT_SGD, es[i]->ev_get(), es[i]->ev_get(), 1};
        HIPC(hipEventRecord(x.a, bc.s));
        tm.push_back(x);
      }
    const hipError_t err = launch_sgd_batch(c, b, waves, bar,
                                            2 * SGD_LAG, bc.num_cu,
                                            bc.s);
    if (err != hipSuccess) {
      (void)hipGetLastError();
      throw Fail(GM_ERR_HIP, std::string("batched SGD launch failed: ") +
                                 hipGetErrorString(err));
    }
    size_t ti = 0;
    for (int i : members)
      if (es[i]->prof) {
        HIPC(hipEventRecord(tm[ti].b, bc.s));
        es[i]->pending.push_back(tm[ti++]);
      }
    for (int i : members)
      if (jobs[i].rec)
        HIPC(launch_fix(c, 1, jobs[i].emb, jobs[i].n,
                        FixFrame{jobs[i].rec}, jobs[i].rounding_p, bc.num_cu,
                        bc.s));
    HIPC(hipMemcpyAsync(bc.h_flag + nchunks, bar + 2, sizeof(unsigned),
                        hipMemcpyDeviceToHost, bc.s));
    hipEvent_t done = es[members[0]]->ev_get();
    HIPC(hipEventRecord(done, bc.s));
    last_done = done;
    for (int i : members) {
      HIPC(hipStreamWaitEvent(es[i]->s, done, 0));
      ++es[i]->sgd_waves_launches;
      if (es[i]->prof) {
        k_add_edge_epochs<<<1, 64, 0, es[i]->s>>>(
            jobs[i].nnz_p, jobs[i].n_epochs,
            es[i]->kstats + Engine::KS_SGD_EE);
        LAUNCH_CHECK();
      }
    }
    dones.v.emplace_back(es[members[0]], done);
    ++nchunks;
  };
  // One launch for all prepared sessions of a component count (a launch is
  // closed early only at SGD_MAX_JOBS jobs or BATCH_MAX_WAVES waves):
  // splitting the sessions into launches that overlap the later sessions'
  // insertions measured slower, as the two contend for the CUs.
  const unsigned gmax = (unsigned)SGD_MAX_JOBS;
  std::vector<std::vector<int>> open; // one open group per component count
  std::vector<unsigned> open_waves;
  auto group_of = [&](int c) -> size_t {
    for (size_t g = 0; g < open.size(); ++g)
      if (!open[g].empty() && es[open[g][0]]->c == c)
        return g;
    open.emplace_back();
    open_waves.push_back(0);
    return open.size() - 1;
  };
  for (int i = 0; i < cnt; ++i) {
    Engine &E = *es[i];
    if (did[i]) {
      stream_wait(E.s);
      E.insert_finish();
      rs[i].insertion_time = secs(tb);
      E.insertion_completed = E.insertion_queue.size() == 0 && E.n == E.N;
    }
    if (!(ops[3 * i + 1] == 0 || E.update_completed)) {
      const auto t0 = clk::now();
      uint64_t ch = 0;
      rs[i].n_update_ops = E.update(ops[3 * i + 1], ch);
      rs[i].n_update_changed_edges = ch;
      stream_wait(E.s);
      rs[i].update_time = secs(t0);
    }
    rs[i].embedding_pending_edges = E.n;
    emb[i] = E.embedding_due(ops[3 * i + 2]);
    if (!emb[i])
      continue;
    te[i] = clk::now();
    // An SGD on the session's own stream is a grid-wide kernel that needs
    // every CU: it starts once everything the cycle has enqueued so far, on
    // every session's stream and the batch stream, has finished.  Other work
    // still running would hold CUs while its resident blocks wait at the
    // epoch boundaries for the rest (and a batched SGD could wait for them
    // in turn).
    auto quiesce = [&]() {
      if (last_done)
        HIPC(hipStreamWaitEvent(E.s, last_done, 0));
      for (int j = 0; j < cnt; ++j) {
        if (j == i)
          continue;
        hipEvent_t ev = es[j]->ev_get();
        HIPC(hipEventRecord(ev, es[j]->s));
        HIPC(hipStreamWaitEvent(E.s, ev, 0));
        dones.v.emplace_back(es[j], ev);
      }
    };
    rs[i].n_embedding_ops = E.embed_begin(ops[3 * i + 2], &jobs[i], quiesce);
    if (!jobs[i].nwaves) {
      stream_wait(E.s);
      continue;
    }
    size_t g = group_of(E.c);
    if (open_waves[g] + jobs[i].nwaves > BATCH_MAX_WAVES ||
        open[g].size() >= (size_t)SGD_MAX_JOBS) {
      launch_group(open[g]);
      open[g].clear();
      open_waves[g] = 0;
      g = group_of(E.c);
    }
    open[g].push_back(i);
    open_waves[g] += jobs[i].nwaves;
    if (open[g].size() >= gmax) {
      launch_group(open[g]);
      open[g].clear();
      open_waves[g] = 0;
    }
  }
  for (auto &m : open)
    launch_group(m);
  for (int i = 0; i < cnt; ++i)
    if (emb[i])
      es[i]->embed_end();
  for (int i = 0; i < cnt; ++i)
    if (emb[i])
      stream_wait(es[i]->s);
  stream_wait(bc.s);
  for (int i = 0; i < cnt; ++i)
    if (emb[i])
      rs[i].embedding_time = secs(te[i]);
  for (int q = 0; q < nchunks; ++q)
    if (bc.h_flag[q])
      throw Fail(GM_ERR_HIP, "batched SGD device-wide barrier timed out");
  for (int i = 0; i < cnt; ++i)
    es[i]->finish_result(&rs[i]);
}

__global__ void k_e4m3_probe(const float *in, uint8_t *out, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    out[i] = e4m3_rn(in[i]);
}

__global__ void k_philox_probe(const uint32_t *in, uint32_t *out) {
  uint32_t c[4] = {in[0], in[1], in[2], in[3]};
  uint32_t kk[2] = {in[4], in[5]};
  uint32_t o[4];
  philox(c, kk, o);
  for (int i = 0; i < 4; ++i)
    out[i] = o[i];
}

} // namespace gm

struct gm_handle {
  gm::Engine *e;
};

using gm::Fail;

namespace {
template <typename F> int guarded(F &&f) {
  try {
    f();
    gm::g_err.clear();
    return GM_OK;
  } catch (const Fail &x) {
    gm::g_err = x.what();
    return x.code;
  } catch (const std::bad_alloc &) {
    gm::g_err = "host allocation failed";
    return GM_ERR_OOM;
  } catch (const std::exception &x) {
    gm::g_err = x.what();
    return GM_ERR_HIP;
  }
}

void check_handle(gm_handle *h) {
  if (!h || !h->e)
    throw Fail(GM_ERR_INVALID, "null or destroyed gm_handle");
  if (h->e->release_pending)
    throw Fail(GM_ERR_STATE, "the session was released (gm_release)");
  HIPC(hipSetDevice(h->e->device));
}

void view_of(gm::Engine *e, int32_t which, gm_view *v) {
  std::memset(v, 0, sizeof(*v));
  v->rows = e->n;
  v->device_id = e->device;
  switch (which) {
  case GM_EMBEDDING:
    v->data = e->emb;
    v->cols = (uint64_t)e->c;
    v->row_stride = (uint64_t)e->c;
    v->dtype_code = 2;
    v->dtype_bits = 32;
    break;
  case GM_NEIGHBOR_INDICES:
    v->data = e->pidx;
    v->cols = (uint64_t)e->k;
    v->row_stride = (uint64_t)e->k;
    v->dtype_code = 1;
    v->dtype_bits = 32;
    break;
  case GM_NEIGHBOR_DISTS:
    v->data = e->pdist;
    v->cols = (uint64_t)e->k;
    v->row_stride = (uint64_t)e->k;
    v->dtype_code = 2;
    v->dtype_bits = 32;
    break;
  default:
    throw Fail(GM_ERR_INVALID, "unknown array id " + std::to_string(which));
  }
}
} // namespace

extern "C" {

void gm_default_params(gm_params *o) {
  if (!o)
    return;
  std::memset(o, 0, sizeof(*o));
  o->n_neighbors = 15;
  o->n_components = 2;
  o->min_dist = 0.1f;
  o->negative_sample_rate = 5;
  o->max_epoch = 200;
  o->a = 1.0f;
  o->b = 1.0f;
  o->alpha = 1.0f;
  o->verbose = 0;
  o->metric = GM_EUCLIDEAN;
  o->knn_mode = GM_KNN_EXACT_BF16;
}

int gm_create(const gm_params *params, gm_handle **out) {
  return guarded([&] {
    if (!params || !out)
      throw Fail(GM_ERR_INVALID, "null argument");
    *out = nullptr;
    const gm_params &q = *params;
    if (q.n_neighbors < 2 || q.n_neighbors > (uint64_t)gm::W)
      throw Fail(GM_ERR_INVALID, "n_neighbors must be in [2, 32]");
    if (q.n_components < 1 || q.n_components > 4)
      throw Fail(GM_ERR_INVALID, "n_components must be in [1, 4]");
    if (q.n_instances == 0 || q.n_instances >= 0x7fffffffull)
      throw Fail(GM_ERR_INVALID, "n_instances must be in [1, 2^31-1)");
    // the SGD kernels update the embedding through a buffer addressed with
    // 32-bit byte offsets below 2^31
    if (q.n_instances * q.n_components * sizeof(float) >= 0x80000000ull)
      throw Fail(GM_ERR_INVALID,
                 "n_instances * n_components must be below 2^29");
    if (q.n_features == 0)
      throw Fail(GM_ERR_INVALID, "n_features must be positive");
    if (q.metric != GM_EUCLIDEAN && q.metric != GM_COSINE)
      throw Fail(GM_ERR_INVALID, "metric must be 0 (euclidean) or 1 (cosine)");
    if (q.knn_mode < GM_KNN_NNDESCENT || q.knn_mode > GM_KNN_NNDESCENT_BF16)
      throw Fail(GM_ERR_INVALID, "knn_mode must be 0 (nndescent), 1 "
                                 "(exact_fp32), 2 (exact_bf16), 3 "
                                 "(exact_fp8) or 4 (nndescent_bf16)");
    if ((q.knn_mode == GM_KNN_EXACT_BF16 || q.knn_mode == GM_KNN_EXACT_FP8) &&
        q.n_instances >= (1ull << 30))
      throw Fail(GM_ERR_INVALID,
                 "exact_bf16 and exact_fp8 support n_instances < 2^30");
    if (q.negative_sample_rate < 1)
      throw Fail(GM_ERR_INVALID, "negative_sample_rate must be >= 1");
    if (q.max_epoch > 1000000)
      throw Fail(GM_ERR_INVALID, "max_epoch too large");
    auto *h = new gm_handle{nullptr};
    try {
      h->e = new gm::Engine(q);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int gm_initialize(gm_handle *h, uint64_t n_instances, uint64_t n_features) {
  return guarded([&] {
    check_handle(h);
    if (n_instances != h->e->N || n_features != h->e->D)
      throw Fail(GM_ERR_INVALID,
                 "initialize(): sizes differ from the params the handle was "
                 "created with");
  });
}

int gm_append_data(gm_handle *h, const float *x, uint64_t rows,
                   uint64_t cols) {
  return guarded([&] {
    check_handle(h);
    if (cols != h->e->D)
      throw Fail(GM_ERR_INVALID,
                 "n_features mismatch with numpy array second dimension.");
    if (rows == 0)
      return;
    if (!x)
      throw Fail(GM_ERR_INVALID, "null data pointer");
    h->e->append(x, rows);
  });
}

int gm_run(gm_handle *h, uint32_t ni, uint32_t nu, uint32_t ne,
           gm_result *out) {
  return guarded([&] {
    check_handle(h);
    if (!out)
      throw Fail(GM_ERR_INVALID, "null result pointer");
    h->e->run(ni, nu, ne, out);
  });
}

int gm_close(gm_handle *h) {
  return guarded([&] {
    if (h && h->e && h->e->release_pending)
      return; // released sessions are closed already
    check_handle(h);
    HIPC(hipStreamSynchronize(h->e->s));
    h->e->closed = true;
  });
}

int gm_release(gm_handle *h) {
  return guarded([&] {
    if (!h || !h->e)
      throw Fail(GM_ERR_INVALID, "null or destroyed gm_handle");
    gm::Engine *e = h->e;
    std::lock_guard<std::mutex> l(e->rel_mu);
    if (e->release_pending)
      return;
    HIPC(hipSetDevice(e->device));
    HIPC(hipStreamSynchronize(e->s));
    e->closed = true;
    e->release_pending = true;
    gm::release_if_unshared(e);
  });
}

void gm_shutdown(void) {
  // no device work of any session is in flight once both locks are held
  std::lock_guard<std::mutex> bl(gm::g_batch_mu);
  std::unique_lock<gm::DevLock> dl(gm::g_dev_mu);
  std::vector<gm::Engine *> live;
  {
    std::lock_guard<std::mutex> l(gm::g_live_mu);
    live = gm::g_live;
  }
  for (gm::Engine *e : live) {
    std::lock_guard<std::mutex> l(e->rel_mu);
    (void)hipSetDevice(e->device);
    e->closed = true;
    e->release_pending = true;
    e->release();
    e->released.store(true);
  }
  for (auto &b : gm::g_batch) {
    (void)hipSetDevice(b.device);
    (void)hipStreamSynchronize(b.s);
    (void)hipStreamDestroy(b.s);
    (void)hipFree(b.bar);
    (void)hipHostFree(b.h_flag);
  }
  gm::g_batch.clear();
}

int gm_is_released(gm_handle *h, int32_t *out) {
  return guarded([&] {
    if (!h || !h->e || !out)
      throw Fail(GM_ERR_INVALID, "null argument");
    *out = h->e->released.load() ? 1 : 0;
  });
}

void gm_destroy(gm_handle *h) {
  if (!h)
    return;
  if (h->e) {
    // a session freed by gm_shutdown makes no further runtime call (its
    // handle may be destroyed after the runtime has finalized)
    if (!h->e->released.load())
      (void)hipSetDevice(h->e->device);
    gm::engine_unref(h->e);
  }
  h->e = nullptr;
  delete h;
}

int gm_get_view(gm_handle *h, int32_t which, gm_view *out) {
  return guarded([&] {
    check_handle(h);
    if (!out)
      throw Fail(GM_ERR_INVALID, "null view pointer");
    view_of(h->e, which, out);
  });
}

int gm_export_dlpack(gm_handle *h, int32_t which, void **dl) {
  return guarded([&] {
    check_handle(h);
    if (!dl)
      throw Fail(GM_ERR_INVALID, "null output pointer");
    gm_view v;
    view_of(h->e, which, &v);
    HIPC(hipStreamSynchronize(h->e->s));
    auto *x = new gm::Export();
    x->shape[0] = (int64_t)v.rows;
    x->shape[1] = (int64_t)v.cols;
    x->strides[0] = (int64_t)v.row_stride;
    x->strides[1] = 1;
    x->mt.dl_tensor.data = v.data;
    x->mt.dl_tensor.device = {10 /* kDLROCM */, v.device_id};
    x->mt.dl_tensor.ndim = 2;
    x->mt.dl_tensor.dtype = {(uint8_t)v.dtype_code, (uint8_t)v.dtype_bits, 1};
    x->mt.dl_tensor.shape = x->shape;
    x->mt.dl_tensor.strides = x->strides;
    x->mt.dl_tensor.byte_offset = 0;
    x->mt.manager_ctx = h->e;
    x->mt.deleter = gm::dl_deleter;
    h->e->refs.fetch_add(1);
    *dl = &x->mt;
  });
}

int gm_append_data_device(gm_handle *h, const float *device_rows,
                          uint64_t rows, uint64_t cols) {
  return guarded([&] {
    check_handle(h);
    if (cols != h->e->D)
      throw Fail(GM_ERR_INVALID,
                 "n_features mismatch with numpy array second dimension.");
    if (rows == 0)
      return;
    if (!device_rows)
      throw Fail(GM_ERR_INVALID, "null data pointer");
    h->e->append_device(device_rows, rows);
  });
}

int64_t gm_screen_shapes(char *buf, uint64_t cap) {
  try {
    std::string out = "[";
    {
      std::lock_guard<std::mutex> l(gm::g_tune_mu);
      bool first = true;
      for (const auto &kv : gm::g_tune) {
        const gm::TuneKey &k = kv.first;
        const gm::TuneEntry &t = kv.second;
        char head[200];
        std::snprintf(head, sizeof head,
                      "%s{\"device\": %d, \"knn_mode\": %llu, \"db\": %llu, "
                      "\"rows_log4\": %d, \"batch_tiles\": %llu, \"best\": %d, "
                      "\"candidates\": [",
                      first ? "" : ", ", k.dev, (unsigned long long)k.mode,
                      (unsigned long long)k.db, k.nlog4,
                      (unsigned long long)k.brt, t.best);
        out += head;
        for (size_t c = 0; c < t.cand.size(); ++c) {
          char item[96];
          std::snprintf(item, sizeof item,
                        "%s{\"bsplit\": %d, \"lbscan\": %s, \"ms\": %.6g}",
                        c ? ", " : "", t.cand[c].bsplit,
                        t.cand[c].lbscan ? "true" : "false", (double)t.ms[c]);
          out += item;
        }
        out += "]}";
        first = false;
      }
    }
    out += "]";
    if (buf && cap > out.size())
      std::memcpy(buf, out.c_str(), out.size() + 1);
    gm::set_last_error(nullptr);
    return (int64_t)out.size();
  } catch (const std::exception &x) {
    gm::set_last_error(x.what());
    return GM_ERR_HIP;
  }
}

int gm_run_batch(gm_handle *const *hs, int32_t count, const uint32_t *ops,
                 gm_result *results) {
  return guarded([&] {
    if (!hs || !ops || !results || count < 1)
      throw Fail(GM_ERR_INVALID, "bad batched run request");
    std::vector<gm::Engine *> es((size_t)count);
    for (int32_t i = 0; i < count; ++i) {
      check_handle(hs[i]);
      es[i] = hs[i]->e;
      if (es[i]->device != es[0]->device)
        throw Fail(GM_ERR_INVALID, "batched sessions must share one device");
      for (int32_t j = 0; j < i; ++j)
        if (es[j] == es[i])
          throw Fail(GM_ERR_INVALID, "a session appears twice in one batch");
    }
    HIPC(hipSetDevice(es[0]->device));
    gm::run_batch(es.data(), count, ops, results);
  });
}

int gm_set_profiling(gm_handle *h, int32_t enable) {
  return guarded([&] {
    check_handle(h);
    h->e->prof = enable != 0;
  });
}

int gm_get_kernel_stats(gm_handle *h, int32_t which, gm_kernel_stats *out) {
  return guarded([&] {
    check_handle(h);
    if (which < 0 || which >= GM_KSTAT_COUNT || !out)
      throw Fail(GM_ERR_INVALID, "bad kernel-stats request");
    h->e->read_counters();
    *out = h->e->totals[which];
  });
}

int gm_reset_kernel_stats(gm_handle *h) {
  return guarded([&] {
    check_handle(h);
    gm::Engine *e = h->e;
    HIPC(hipStreamSynchronize(e->s));
    for (auto &t : e->totals)
      t = gm_kernel_stats{};
    e->fold_events();
    e->sgd_waves_launches = 0;
    e->ins_pairs = 0;
    e->graph_replays = e->graph_instantiations = 0;
    HIPC(hipMemsetAsync(e->kstats, 0, 16 * sizeof(unsigned long long), e->s));
    HIPC(hipStreamSynchronize(e->s));
  });
}

int gm_copy_knn(gm_handle *h, uint32_t *idx, float *dist) {
  return guarded([&] {
    check_handle(h);
    gm::Engine *e = h->e;
    const size_t nk = e->n * (size_t)e->k;
    if (idx)
      HIPC(hipMemcpyAsync(idx, e->pidx, nk * 4, hipMemcpyDeviceToHost, e->s));
    if (dist)
      HIPC(hipMemcpyAsync(dist, e->pdist, nk * 4, hipMemcpyDeviceToHost, e->s));
    HIPC(hipStreamSynchronize(e->s));
  });
}

int gm_copy_embedding(gm_handle *h, float *emb) {
  return guarded([&] {
    check_handle(h);
    gm::Engine *e = h->e;
    HIPC(hipMemcpyAsync(emb, e->emb, e->n * e->c * 4, hipMemcpyDeviceToHost,
                        e->s));
    HIPC(hipStreamSynchronize(e->s));
  });
}

int gm_set_embedding(gm_handle *h, const float *emb) {
  return guarded([&] {
    check_handle(h);
    gm::Engine *e = h->e;
    HIPC(hipMemcpyAsync(e->emb, emb, e->n * e->c * 4, hipMemcpyHostToDevice,
                        e->s));
    HIPC(hipStreamSynchronize(e->s));
  });
}

const char *gm_last_error(void) { return gm::g_err.c_str(); }

int gm_debug_pool(gm_handle *h, uint64_t *keys) {
  return guarded([&] {
    check_handle(h);
    gm::Engine *e = h->e;
    HIPC(hipMemcpyAsync(keys, e->pool, e->n * gm::W * 8, hipMemcpyDeviceToHost,
                        e->s));
    HIPC(hipStreamSynchronize(e->s));
  });
}

int gm_debug_fuzzy(gm_handle *h, const uint32_t *batch, uint64_t nb,
                   int32_t *head, int32_t *tail, float *w, double *sigma,
                   double *rho, int64_t *nnz) {
  return guarded([&] {
    check_handle(h);
    gm::Engine *e = h->e;
    if (nb > e->n || nb == 0)
      throw Fail(GM_ERR_INVALID, "batch size out of range");
    std::vector<bool> seen(e->n, false);
    for (uint64_t i = 0; i < nb; ++i) {
      if (batch[i] >= e->n)
        throw Fail(GM_ERR_INVALID, "batch row out of range");
      if (seen[batch[i]])
        throw Fail(GM_ERR_INVALID, "duplicate batch row");
      seen[batch[i]] = true;
    }
    std::shared_lock<gm::DevLock> dl(gm::g_dev_mu);
    HIPC(hipMemcpyAsync(e->dbatch, batch, nb * 4, hipMemcpyHostToDevice, e->s));
    e->build_fuzzy(nb, gm::NRows{nullptr, e->n});
    if (nb < e->n) // no reposition follows to clear the batch marks
      HIPC(hipMemsetAsync(e->bmark, 0, e->n, e->s));
    uint32_t cnt = 0;
    HIPC(hipMemcpyAsync(&cnt, e->scal + gm::S_NNZ, 4, hipMemcpyDeviceToHost,
                        e->s));
    HIPC(hipStreamSynchronize(e->s));
    // the device COO holds twin pairs with empty slots (head -1); the caller
    // gets the edges alone, in their order
    std::vector<int32_t> hv(cnt), tv(cnt);
    std::vector<float> wv(cnt);
    HIPC(hipMemcpyAsync(hv.data(), e->head, cnt * 4ull, hipMemcpyDeviceToHost,
                        e->s));
    HIPC(hipMemcpyAsync(tv.data(), e->tail, cnt * 4ull, hipMemcpyDeviceToHost,
                        e->s));
    HIPC(hipMemcpyAsync(wv.data(), e->wout, cnt * 4ull, hipMemcpyDeviceToHost,
                        e->s));
    if (sigma)
      HIPC(hipMemcpyAsync(sigma, e->sigma, e->n * 8, hipMemcpyDeviceToHost, e->s));
    if (rho)
      HIPC(hipMemcpyAsync(rho, e->rho, e->n * 8, hipMemcpyDeviceToHost, e->s));
    double mean = 0.0;
    HIPC(hipMemcpyAsync(&mean, e->mean, 8, hipMemcpyDeviceToHost, e->s));
    HIPC(hipStreamSynchronize(e->s));
    // a stored sigma < 0 stands for the global-mean floor
    if (sigma)
      for (uint64_t i = 0; i < e->n; ++i)
        if (sigma[i] < 0.0)
          sigma[i] = 1e-3 * mean;
    int64_t m = 0;
    for (uint32_t i = 0; i < cnt; ++i) {
      if (hv[i] < 0)
        continue;
      if (head)
        head[m] = hv[i];
      if (tail)
        tail[m] = tv[i];
      if (w)
        w[m] = wv[i];
      ++m;
    }
    *nnz = m;
  });
}

int gm_debug_sgd(float *emb, uint64_t n, int32_t c, const int32_t *head,
                 const int32_t *tail, const float *w, int64_t nnz,
                 int32_t n_epochs, int32_t nsr, float a, float b, float alpha0,
                 float gamma, uint64_t seed, int32_t per_epoch) {
  gm_handle *h = nullptr;
  int rc = guarded([&] {
    if (n == 0 || c < 1 || c > 4 || nnz < 0 || nsr < 1 || per_epoch < 0 ||
        per_epoch > 3)
      throw Fail(GM_ERR_INVALID, "invalid SGD arguments");
    // the SGD kernels address the embedding with 32-bit byte offsets (the
    // limit gm_create applies to n_instances * n_components)
    if (n * (uint64_t)c * sizeof(float) >= 0x80000000ull)
      throw Fail(GM_ERR_INVALID, "n * c must be below 2^29");
    for (int64_t i = 0; i < nnz; ++i) {
      if (head[i] == -1 && tail[i] == -1)
        continue; // an empty slot
      if (head[i] < 0 || (uint64_t)head[i] >= n || tail[i] < 0 ||
          (uint64_t)tail[i] >= n)
        throw Fail(GM_ERR_INVALID, "COO index out of range");
    }
    gm_params q;
    gm_default_params(&q);
    q.n_instances = std::max<uint64_t>(n, (uint64_t)(nnz + 1) / 2 + 1);
    q.n_features = 1;
    q.n_neighbors = 2;
    q.n_components = (uint64_t)c;
    q.knn_mode = GM_KNN_NNDESCENT;
    h = new gm_handle{new gm::Engine(q)};
    gm::Engine *e = h->e;
    e->n = n;
    float mw = 0.0f;
    for (int64_t i = 0; i < nnz; ++i)
      if (head[i] >= 0)
        mw = std::max(mw, w[i]);
    uint32_t sc[2] = {(uint32_t)nnz, 0};
    std::memcpy(&sc[1], &mw, 4);
    HIPC(hipMemcpyAsync(e->emb, emb, n * c * 4, hipMemcpyHostToDevice, e->s));
    HIPC(hipMemcpyAsync(e->head, head, nnz * 4, hipMemcpyHostToDevice, e->s));
    HIPC(hipMemcpyAsync(e->tail, tail, nnz * 4, hipMemcpyHostToDevice, e->s));
    HIPC(hipMemcpyAsync(e->wout, w, nnz * 4, hipMemcpyHostToDevice, e->s));
    HIPC(hipMemcpyAsync(e->scal + gm::S_NNZ, &sc[0], 4, hipMemcpyHostToDevice,
                        e->s));
    HIPC(hipMemcpyAsync(e->scal + gm::S_MAXW, &sc[1], 4, hipMemcpyHostToDevice,
                        e->s));
    std::unique_lock<gm::DevLock> dl(gm::g_dev_mu);
    std::lock_guard<std::mutex> lk(gm::g_coop_mu);
    if (mw > 0.0f && nnz > 0)
      e->run_sgd(e->emb, n, (uint64_t)nnz, n_epochs, nsr, a, b, alpha0, gamma,
                 seed,
                 per_epoch == 1   ? gm::SGD_PER_EPOCH
                 : per_epoch == 2 ? gm::SGD_STREAM
                 : per_epoch == 3 ? gm::SGD_PULL
                                  : gm::SGD_AUTO);
    HIPC(hipMemcpyAsync(emb, e->emb, n * c * 4, hipMemcpyDeviceToHost, e->s));
    HIPC(hipStreamSynchronize(e->s));
    e->check_barrier();
  });
  gm_destroy(h);
  return rc;
}

int gm_debug_e4m3(const float *in, uint8_t *out, uint64_t n,
                  int32_t on_device) {
  if (n > 0 && (!in || !out))
    return GM_ERR_INVALID;
  if (!on_device) {
    for (uint64_t i = 0; i < n; ++i)
      out[i] = gm::e4m3_rn(in[i]);
    gm::set_last_error(nullptr);
    return GM_OK;
  }
  float *d = nullptr;
  int rc = guarded([&] {
    if (n == 0)
      return;
    HIPC(hipMalloc(&d, n * 5));
    HIPC(hipMemcpy(d, in, n * 4, hipMemcpyHostToDevice));
    uint8_t *o = reinterpret_cast<uint8_t *>(d + n);
    gm::k_e4m3_probe<<<(unsigned)((n + 255) / 256), 256>>>(d, o, n);
    HIPC(hipGetLastError());
    HIPC(hipMemcpy(out, o, n, hipMemcpyDeviceToHost));
  });
  if (d)
    (void)hipFree(d);
  return rc;
}

int gm_debug_philox(const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4]) {
  uint32_t *d = nullptr;
  int rc = guarded([&] {
    HIPC(hipMalloc(&d, 16 * sizeof(uint32_t)));
    uint32_t in[6] = {ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]};
    HIPC(hipMemcpy(d, in, sizeof(in), hipMemcpyHostToDevice));
    gm::k_philox_probe<<<1, 1>>>(d, d + 8);
    HIPC(hipGetLastError());
    HIPC(hipMemcpy(out, d + 8, 16, hipMemcpyDeviceToHost));
  });
  if (d)
    (void)hipFree(d);
  return rc;
}

} // extern "C"
