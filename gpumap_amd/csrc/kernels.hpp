// Kernel declarations shared by the engine and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gm {

// A row count that the kernels read at run time: a device scalar (kernels
// captured into a HIP graph replay while the session grows) or a value.
struct NRows {
  const uint32_t *p;
  uint64_t v;
  __device__ uint64_t get() const { return p ? (uint64_t)*p : v; }
};

// nnd_kernels.hip
__global__ void k_prep_rows(float *X, uint64_t dp, uint64_t d, uint64_t b0,
                            uint64_t nb);
__global__ void k_random_ids(uint32_t *ids, int *cnt, uint64_t b0,
                             uint64_t nb, uint64_t n, uint32_t window_start,
                             uint32_t seed);
__global__ void k_init_dist(const float *X, uint64_t dp, uint64_t *pool,
                            const uint32_t *ids, const int *cnt, uint64_t b0,
                            uint64_t nb);
__global__ void k_ins_candidates(const float *X, uint64_t dp,
                                 const uint64_t *pool, uint64_t b0,
                                 uint64_t nb, int k, uint32_t *cand,
                                 float *cdist, int *ccount);
__global__ void k_ins_offers(uint64_t *pool, uint8_t *flag,
                             const uint32_t *cand, const float *cdist,
                             const int *ccount, uint64_t b0, uint64_t nb,
                             int k);
__global__ void k_sample(uint64_t *pool, const uint32_t *rows, uint64_t nrows,
                         uint32_t *fn, uint32_t *fo, int *fnc, int *foc,
                         uint32_t *newmask, uint32_t *owner_mark);
__global__ void k_reverse(const uint64_t *pool, const uint32_t *rows,
                          uint64_t nrows, const uint32_t *newmask,
                          uint64_t *rev_new, uint64_t *rev_old,
                          uint32_t *owner_mark, uint32_t *owners,
                          uint32_t *owner_count);
__global__ void k_join(const float *X, uint64_t dp, uint64_t *pool,
                       const uint32_t *fn, const uint32_t *fo, const int *fnc,
                       const int *foc, const uint64_t *rev_new,
                       const uint64_t *rev_old, const uint32_t *owners,
                       const uint32_t *owner_count, uint8_t *flag, int k,
                       unsigned long long *stats);
__global__ void k_join_reset(const uint32_t *owners,
                             const uint32_t *owner_count, uint64_t *rev_new,
                             uint64_t *rev_old, uint32_t *owner_mark, int *fnc,
                             int *foc);
__global__ void k_sync_rows(const uint64_t *pool, const uint32_t *rows,
                            uint64_t nrows, const uint32_t *nrows_dev,
                            uint32_t *pidx, float *pdist, int k,
                            unsigned long long *fresh, float *pmax, int euclid,
                            double *sigma, double *rho);
__global__ void k_flag_count(const uint8_t *flag, uint64_t lo, uint64_t hi,
                             uint32_t *bcount, uint64_t f0, uint64_t f1);
__global__ void k_scan_small(uint32_t *v, uint64_t n, uint32_t *total);
__global__ void k_flag_write(uint8_t *flag, uint64_t lo, uint64_t hi,
                             const uint32_t *boff, uint32_t *out, uint64_t f0,
                             uint64_t f1);
__global__ void k_init_embedding(float *emb, int c, const uint32_t *pidx,
                                 const float *pdist, int k, uint64_t b0,
                                 uint64_t nb);

// exact_kernels.hip
__global__ void k_exact_tiles(const float *X, uint64_t dp, uint64_t *pool,
                              const float *pmax, uint64_t b0, uint64_t nb,
                              uint64_t n, int k, int nsplit, uint64_t *part,
                              uint8_t *flag, int new_flag,
                              const unsigned *guard);
__global__ void k_exact_merge(const uint64_t *part, int nsplit, uint64_t *pool,
                              uint64_t b0, uint64_t nb, uint8_t *flag,
                              const unsigned *guard);

// mfma_kernels.hip, fp8_kernels.hip
// Shared by the BF16 and FP8 screens and the survivor kernels: block tiles
// of TB batch rows x TC columns; TOPL upper bounds kept per lane and batch
// row in pass 1 (a lane sees a quarter of a block's columns for a row, so a
// list of 8 is cut short only when more than 8 of the row's 32 nearest fall
// among them, and tau is then larger, still valid); PARTW pass-1 values per
// (batch row, split); survivor list entries: the column row and its flags.
namespace screen {
constexpr int TB = 128;
constexpr int TC = 128;
constexpr int TOPL = 8;
constexpr int PARTW = 4 * TOPL;
constexpr uint32_t LFWD = 0x80000000u; // forward pair
constexpr uint32_t LREV = 0x40000000u; // reverse pair
constexpr uint32_t LID = 0x3fffffffu;  // column row
} // namespace screen
// feature steps (16 features each) whose products k_bf16_screen sums on the
// matrix core before one round-to-nearest FP32 add into the running dot
// product; the screening bound's summation term depends on it
constexpr int BF16_GROUP_STEPS = 4;
__global__ void k_bf16_prep(const float *X, uint64_t dp, uint64_t db,
                            __bf16 *Xs, __bf16 *Xb, float *nrm, uint8_t *flag,
                            unsigned *rcnt, unsigned *qctl, uint64_t b0,
                            uint64_t nb);
template <int PASS>
__global__ void k_bf16_screen(const __bf16 *Xs, const __bf16 *Xb, int nsteps,
                              const float *nrm, const float *pmax,
                              const float *tau, uint64_t b0, uint64_t nb,
                              uint64_t n, int nsplit, int rtiles, float *part,
                              uint32_t *rlist, unsigned *rcnt, unsigned rcap,
                              unsigned *qctl, float coef, float cfac,
                              float *lbm);
// fp8_kernels.hip
// feature steps (64 features each) of k_fp8_screen's groups (as
// BF16_GROUP_STEPS)
constexpr int FP8_GROUP_STEPS = 2;
__global__ void k_fp8_prep(const float *X, uint64_t dp, uint64_t db,
                           uint8_t *Xs, uint8_t *Xb, float *nrm, float *isc,
                           uint8_t *flag, unsigned *rcnt, unsigned *qctl,
                           uint64_t b0, uint64_t nb);
template <int PASS>
__global__ void k_fp8_screen(const uint8_t *Xs, const uint8_t *Xb, int nsteps,
                             const float *nrm, const float *isc,
                             const float *pmax, const float *tau, uint64_t b0,
                             uint64_t nb, uint64_t n, int nsplit, int rtiles,
                             float *part, uint32_t *rlist, unsigned *rcnt,
                             unsigned rcap, unsigned *qctl, float coef,
                             float cfac, float *lbm);
__global__ void k_bf16_lbscan(const float *lbm, uint64_t nb, uint64_t n,
                              uint64_t b0, const float *tau, const float *pmax,
                              uint32_t *rlist, unsigned *rcnt, unsigned rcap,
                              unsigned *qctl, uint64_t cols_per_wave);
__global__ void k_tau_select(const float *part, int nsplit, float *tau,
                             uint64_t nb);
__global__ void k_bf16_rerank(const float *X, uint64_t dp,
                              const uint32_t *rlist, unsigned *rcnt,
                              unsigned rcap, const unsigned *qctl,
                              uint64_t *fkeys, uint64_t *pool,
                              const float *pmax, uint8_t *flag, uint64_t b0,
                              const float *tau, uint64_t nb,
                              unsigned long long *stats);
__global__ void k_bf16_select(const uint64_t *fkeys, const unsigned *fcnt,
                              unsigned rcap, const unsigned *qctl,
                              uint64_t *pool, uint64_t b0,
                              unsigned long long *stats);

// join_kernels.hip
__global__ void k_join_prep(const float *X, uint64_t dp, uint64_t dj,
                            __bf16 *Xj, float *nrm, uint64_t b0, uint64_t nb);
__global__ void k_join_bf16(const float *X, uint64_t dp, const __bf16 *Xj,
                            uint64_t dj, const float *nrm, uint64_t *pool,
                            const uint32_t *fn, const uint32_t *fo,
                            const int *fnc, const int *foc,
                            const uint64_t *rev_new, const uint64_t *rev_old,
                            const uint32_t *owners,
                            const uint32_t *owner_count, uint8_t *flag, int k,
                            float coef, float cfac, unsigned long long *stats);

// embed_kernels.hip
__global__ void k_mean_partial(const uint32_t *pidx, const float *pdist,
                               NRows n, int k, int euclid, double *psum,
                               unsigned long long *pcnt);
__global__ void k_mean_final(const double *psum, const unsigned long long *pcnt,
                             int nparts, double *mean, uint32_t *reset3,
                             const uint32_t *batch, uint64_t nb, uint8_t *mark);
__global__ void k_membership(const uint32_t *pidx, const float *pdist,
                             NRows n, int k, int euclid,
                             const double *sigma, const double *rho,
                             const double *mean, const uint32_t *batch,
                             uint64_t nb, const uint8_t *inb, float *w,
                             uint8_t *flags, uint32_t *maxw_bits);
__global__ void k_coo_count(const float *w, const uint8_t *flags, uint64_t nd,
                            const uint32_t *maxw_bits, int n_epochs,
                            uint32_t *bcount);
__global__ void k_coo_write(const uint32_t *pidx, int k, const uint32_t *batch,
                            const float *w, const uint8_t *flags, uint64_t nd,
                            const uint32_t *maxw_bits, int n_epochs,
                            const uint32_t *boff, int32_t *head, int32_t *tail,
                            float *wout);
__global__ void k_sgd_init(const float *w, const int32_t *head,
                           const uint32_t *nnz_p, const uint32_t *maxw_bits,
                           int n_epochs, float nsr_inv, float *eps, float *nxt,
                           float *nneg, int *deg, int *maxdeg, unsigned *bar,
                           unsigned nbar);
__global__ void k_max_int(const int *v, NRows n, int *out, float alpha,
                          float *rounding, unsigned *done);
__global__ void k_minmax_col0(const float *emb, NRows n, int c,
                              uint32_t *mn, uint32_t *mx, int *deg_zero);
__global__ void k_noise(float *emb, int c, const uint32_t *batch, uint64_t nb,
                        const uint32_t *mn, const uint32_t *mx,
                        float noise_scale, uint64_t seed);
template <int NC>
__global__ void k_sgd_epoch(float *emb, const int32_t *head,
                            const int32_t *tail, const uint32_t *nnz_p,
                            uint32_t n, const float *eps, float *nxt,
                            float *nneg, float alpha, int epoch, float gamma,
                            float a, float b, float nsr_inv, uint64_t seed,
                            const float *rounding_p, unsigned long long *stats);
template <int NC, int EPT>
__global__ void k_sgd_fused(float *emb, const int32_t *head,
                            const int32_t *tail, const uint32_t *nnz_p,
                            uint32_t n, const float *eps, int n_epochs,
                            float alpha0, float gamma, float a, float b,
                            float nsr_inv, uint64_t seed,
                            const float *rounding_p, unsigned *bar,
                            unsigned long long *stats, int lag);
// arrival slots of k_sgd_waves after the 4 barrier words (bar + 4)
constexpr unsigned SGD_MAX_WAVES = 512;

// One session's per-cycle SGD in a batched launch (k_sgd_waves_multi): its
// edge set, schedule and parameters, and its range of waves in the grid.
struct SgdJob {
  float *emb;
  const int32_t *head, *tail;
  const uint32_t *nnz_p;
  const float *eps, *rounding_p;
  unsigned long long *stats;
  uint64_t seed;
  uint32_t n;
  int n_epochs;
  float alpha0, gamma, a, b, nsr_inv;
  unsigned wave0, nwaves;
  int32_t *rec; // the session's row records (FixFrame), or nullptr
};
constexpr int SGD_MAX_JOBS = 16;
struct SgdBatch { // passed by value as a kernel argument
  SgdJob job[SGD_MAX_JOBS];
  int njobs;
};
// one cooperative launch of k_sgd_waves_multi over the jobs' waves (job q
// starts at wave job[q].wave0, ascending, each <= SGD_MAX_WAVES waves); bar
// holds 4 words and then one slot per wave, zeroed by the caller; returns
// hipErrorInvalidConfiguration when the waves cannot all be resident
hipError_t launch_sgd_batch(int nc, const SgdBatch &batch, unsigned waves,
                            unsigned *bar, int lag, int num_cu,
                            hipStream_t s);
// The cooperative SGD kernels' row records (embed_kernels.hip): n records
// of 2 F words (F = NC, or 4 for NC = 3): the row's coordinates at the
// launch's start, then its updates as integer multiples of the rounding
// grid.  rec == nullptr: the kernels update the float embedding.
struct FixFrame {
  int32_t *rec;
};
// Scratch of the row-owner SGD kernel (k_sgd_pull, embed_kernels.hip) for
// edge sets beyond the resident grid: every COO slot listed under both its
// rows (an entry), the entries grouped by row, each with its own copy of the
// slot's schedule; sized for up to `slots` COO slots and `rows` rows.
struct PullFrame {
  uint32_t *cnt;   // rows + 1: entries per row, then the fill cursors
  uint32_t *off;   // rows + 1: first entry of each row, off[n] = entries
  uint32_t *bsum;  // per 1024 rows: block sums of cnt
  int2 *erc;       // entry's row, the slot's other row
  // slot | side << 31 (side 1: the row is the slot's tail; as float bits),
  // then the slot's schedule: eps, next sample, next negative sample (used
  // on the head side)
  float4 *esch;
  int32_t *epc;    // per entry n_components ints: a run's previous sum
  int32_t *d1;     // second record buffer: rows x 2 fix_stride words
  uint64_t slots, rows;
};
// cooperative launch of k_sgd_waves (*variant 2), k_sgd_fused (0) or, for
// edge sets larger than its resident grid or when forced, k_sgd_stream (1,
// counting its work at stats + 8) (embed_kernels.hip);
// waves and stream run on fx's copy when it has one; returns the launch
// error when none can be launched
hipError_t launch_sgd_fused(int nc, float *emb, const int32_t *head,
                            const int32_t *tail, const uint32_t *nnz_p,
                            uint32_t n, const float *eps, float *nxt,
                            float *nneg, int n_epochs, float alpha0,
                            float gamma, float a, float b, float nsr_inv,
                            uint64_t seed, const float *rounding_p,
                            unsigned *bar, unsigned long long *stats, int lag,
                            uint64_t nnz, int num_cu, hipStream_t s,
                            int *variant, int stream_mode, FixFrame fx);
// stream_mode of launch_sgd_fused: the streaming kernel as the fallback
// (SGD_STREAM_FALLBACK), always (SGD_STREAM_FORCE) or never
// (SGD_STREAM_NONE: hipErrorInvalidConfiguration instead)
constexpr int SGD_STREAM_FALLBACK = 0, SGD_STREAM_FORCE = 1,
              SGD_STREAM_NONE = 2;
// the row-owner SGD (k_sgd_pull) on a COO of up to `cap` slots over n rows:
// builds pf's entries, then one cooperative launch of all epochs on fx's
// records; hipErrorInvalidConfiguration when pf or fx cannot hold it
hipError_t launch_sgd_pull(int nc, const PullFrame &pf, FixFrame fx,
                           const int32_t *head, const int32_t *tail,
                           const uint32_t *nnz_p, uint64_t cap, uint32_t n,
                           const float *eps, const float *nxt,
                           const float *nneg, int n_epochs, float alpha0,
                           float gamma, float a, float b, float nsr_inv,
                           uint64_t seed, const float *rounding_p,
                           unsigned *bar, unsigned long long *stats,
                           int num_cu, hipStream_t s);
// the fixed-point copy's kernels around a cooperative SGD launch (phase 0
// before it, phase 1 after a kernel that used the copy)
hipError_t launch_fix(int nc, int phase, float *emb, uint64_t n, FixFrame fx,
                      const float *rounding_p, int num_cu, hipStream_t s);
// engine.hip: holds the process's device lock exclusively (no session's
// device work is in flight meanwhile) and the cooperative-launch mutex
struct ExclusiveDevice {
  ExclusiveDevice();
  ~ExclusiveDevice();
  ExclusiveDevice(const ExclusiveDevice &) = delete;
  ExclusiveDevice &operator=(const ExclusiveDevice &) = delete;
};
// engine.hip: the message gm_last_error() returns on this thread
void set_last_error(const char *msg);

// engine.hip: the SGD's launches on a prepared schedule (engine and the
// allocation-free entry point share them)
constexpr int SGD_AUTO = 0, SGD_PER_EPOCH = 1, SGD_STREAM = 2, SGD_PULL = 3;
constexpr int SGD_LAG = 1;
bool sgd_launches(int c, float *e, const int32_t *head, const int32_t *tail,
                  const uint32_t *nnz_p, uint32_t nrows, const float *eps,
                  float *nxt, float *nneg, int n_epochs, float alpha0,
                  float gamma, float a, float b, float nsr_inv, uint64_t seed,
                  const float *rounding, unsigned *gbar,
                  unsigned long long *st, uint64_t cap, int num_cu,
                  hipStream_t s, int sgd_mode, int *variant, bool verbose,
                  FixFrame fx, const PullFrame *pf = nullptr);
__global__ void k_reposition(const uint32_t *pidx, const float *pdist,
                             NRows n, int k, int euclid, const double *sigma,
                             const double *rho, const double *mean,
                             uint8_t *in_batch, const float *snap,
                             float *emb, int c);
__global__ void k_copy_rows(const float *src, float *dst, NRows n, int c);
__global__ void k_add_edge_epochs(const uint32_t *nnz_p, int n_epochs,
                                  unsigned long long *ctr);

} // namespace gm
