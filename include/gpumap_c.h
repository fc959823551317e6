/*
 * gpumap_c.h -- C ABI of the MI355X (gfx950) progressive UMAP engine.
 *
 * This is the drop-in boundary: the reference's public C++ class
 * gpumap::GPUMAP (cpp/include/gpumap/gpumap.cuh:60-96, PImpl over
 * cpp/src/impl.cuh) and its pybind11 module _core (cpp/bindings/binding.cu:
 * 45-116) become plain C entry points with POD structs, status codes instead
 * of C++ exceptions, and device arrays exported through DLPack (kDLROCM).
 *
 * Conventions: every function returns GM_OK (0) or a negative gm_status; the
 * message of the last failure on the calling thread is gm_last_error().  No
 * C++ exception crosses this boundary.  One handle is not thread-safe; calls
 * on one handle are serialised by the caller (the reference's contract,
 * binding.cu:109-111).
 */
#ifndef GPUMAP_C_H
#define GPUMAP_C_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gm_status {
  GM_OK = 0,
  GM_ERR_INVALID = -1,  /* bad argument: Python ValueError / RuntimeError */
  GM_ERR_CAPACITY = -2, /* append beyond n_instances (impl.cuh:81-83) */
  GM_ERR_OOM = -3,      /* device allocation failed: Python MemoryError */
  GM_ERR_HIP = -4,      /* any other HIP runtime error */
  GM_ERR_STATE = -5     /* call not valid in the handle's state */
} gm_status;

typedef enum gm_metric { GM_EUCLIDEAN = 0, GM_COSINE = 1 } gm_metric;

/* How insertion builds a new row's neighbours.
 *   GM_KNN_NNDESCENT   the reference's method: random start + 10 rounds of
 *                      neighbours-of-neighbours candidates
 *                      (insertion_kernel.cuh:926-1095), refined by the
 *                      NN-Descent update
 *   GM_KNN_EXACT_FP32  exact: FP32 distance tiles against every inserted row
 *   GM_KNN_EXACT_BF16  exact: BF16 MFMA tiles screen with a proven error
 *                      bound, survivors re-ranked in FP32 (same graph as
 *                      GM_KNN_EXACT_FP32)
 *   GM_KNN_EXACT_FP8   exact: FP8 E4M3 MFMA tiles (per-row power-of-two
 *                      scales, hi/lo split) screen with a proven, looser
 *                      bound; survivors re-ranked in FP32 (same graph)
 *   GM_KNN_NNDESCENT_BF16  the reference's method with the local join's
 *                      distance blocks on BF16 MFMA tiles, screened with
 *                      the proven bound and re-ranked in FP32 (same graph
 *                      as GM_KNN_NNDESCENT) */
typedef enum gm_knn_mode {
  GM_KNN_NNDESCENT = 0,
  GM_KNN_EXACT_FP32 = 1,
  GM_KNN_EXACT_BF16 = 2,
  GM_KNN_EXACT_FP8 = 3,
  GM_KNN_NNDESCENT_BF16 = 4
} gm_knn_mode;

/* Mirrors gpumap::GPUMAPParams (cpp/include/gpumap/gpumap.cuh:15-36) with
 * fixed-width fields; defaults from gm_default_params(). */
typedef struct gm_params {
  uint64_t n_neighbors;          /* k, 2..32 (default 15) */
  uint64_t n_components;         /* 1..4 (default 2) */
  float min_dist;                /* default 0.1 */
  uint64_t negative_sample_rate; /* default 5 */
  uint64_t max_epoch;            /* default 200 */
  float a, b;                    /* curve parameters (default 1, 1) */
  float alpha;                   /* initial learning rate (default 1) */
  int32_t verbose;
  int32_t metric;                /* gm_metric */
  uint64_t n_instances;          /* capacity N */
  uint64_t n_features;           /* D */
  int32_t knn_mode;              /* gm_knn_mode (no counterpart in the
                                    reference; default GM_KNN_EXACT_BF16) */
} gm_params;

/* Mirrors gpumap::GPUMAPResult (cpp/include/gpumap/gpumap.cuh:38-56). */
typedef struct gm_result {
  uint32_t n_insertion_ops;
  uint32_t n_update_ops;
  uint32_t n_embedding_ops;
  float insertion_time;
  float update_time;
  float embedding_time;
  int32_t insertion_completed;
  int32_t update_completed;
  int32_t embedding_completed;
  uint64_t n_inserted_instances;
  uint64_t n_update_changed_edges;
  uint64_t embedding_pending_edges;
  uint64_t insertion_queue_size;
  uint64_t update_queue_size;
  uint64_t embedding_queue_size;
} gm_result;

typedef enum gm_array {
  GM_EMBEDDING = 0,        /* float32 (n_inserted, n_components) */
  GM_NEIGHBOR_INDICES = 1, /* uint32  (n_inserted, n_neighbors), ascending */
  GM_NEIGHBOR_DISTS = 2    /* float32 (n_inserted, n_neighbors), squared L2
                              (Euclidean) or 2 - 2 cos (cosine) */
} gm_array;

/* A non-owning description of a device array (row-major). */
typedef struct gm_view {
  void *data;
  uint64_t rows, cols, row_stride; /* strides in elements */
  int32_t dtype_code;              /* DLPack code: 1 uint, 2 float */
  int32_t dtype_bits;
  int32_t device_id;
} gm_view;

typedef struct gm_handle gm_handle;

void gm_default_params(gm_params *out);

/* GPUMAP(const GPUMAPParams&) -- gpumap.cuh:73, impl.cu:5, impl.cuh:57-62.
 * Allocates every device array for n_instances rows up front.  Refuses
 * (GM_ERR_INVALID) n_instances >= 2^31 - 1, n_instances * n_components >=
 * 2^29 (the embedding is addressed with 32-bit byte offsets) and, in
 * exact_bf16 mode, n_instances >= 2^30. */
int gm_create(const gm_params *params, gm_handle **out);

/* GPUMAP::initialize -- gpumap.cuh:83 (a no-op in the reference,
 * impl.cu:12-15); here it checks that the sizes match the params. */
int gm_initialize(gm_handle *h, uint64_t n_instances, uint64_t n_features);

/* GPUMAP::append_data -- gpumap.cuh:84, impl.cuh:77-101,
 * gpumap_wrapper.cuh:29-40.  Copies `rows` x `cols` row-major float32 host
 * rows before returning; 0 rows is a no-op; cols must equal n_features. */
int gm_append_data(gm_handle *h, const float *host_rows, uint64_t rows,
                   uint64_t cols);

/* GPUMAP::run -- gpumap.cuh:85-86, impl.cuh:146-218: insertion, update and
 * embedding with the given op counts, timed each, blocking. */
int gm_run(gm_handle *h, uint32_t n_insertion_ops, uint32_t n_update_ops,
           uint32_t n_embedding_ops, gm_result *out);

/* One cycle of several sessions together (no counterpart in the reference,
 * whose sessions run one at a time; the serving form of its run()):
 * session i runs ops[3i], ops[3i+1], ops[3i+2] insertion, update and
 * embedding ops exactly as gm_run would, with the same results, except that
 * the three phase times are the batch's.  The sessions' insertion kernels
 * overlap on the device and their per-cycle SGDs run as one grid-wide
 * launch, each paced only by its own edges.  All sessions must be on one
 * device and distinct; calls on one device are serialised. */
int gm_run_batch(gm_handle *const *handles, int32_t count,
                 const uint32_t *ops, gm_result *results);

/* GPUMAP::close -- gpumap.cuh:87, impl.cuh:502-507: synchronises the device.
 * Arrays stay readable until gm_destroy. */
int gm_close(gm_handle *h);

/* ~GPUMAP.  Device memory is released when the last exported DLPack tensor
 * of the handle is deleted. */
void gm_destroy(gm_handle *h);

/* Close the session and free its device memory as soon as no exported DLPack
 * tensor holds it: at this call, or when the last such tensor is deleted.
 * Every later call on the handle except gm_close, gm_is_released and
 * gm_destroy fails with GM_ERR_STATE.  No counterpart in the reference,
 * whose storage lives until the object is destroyed (impl.cuh:502-507); a
 * server uses it to return a closed session's HBM before admitting others. */
int gm_release(gm_handle *h);

/* *out = 1 once the device memory of a released session has been freed. */
int gm_is_released(gm_handle *h, int32_t *out);

/* Process teardown: waits for the device work in flight, then frees the
 * device memory of every live session (exported tensors included) and the
 * batched cycles' streams, while the HIP runtime is still up.  Afterwards
 * every session call fails with GM_ERR_STATE.  Meant for an exit hook (the
 * Python package registers one); no counterpart in the reference. */
void gm_shutdown(void);

/* get_embedding / get_neighbor_indices / get_neighbor_dists --
 * gpumap.cuh:89-91, impl.cuh:509-533.  The reference returns the full
 * capacity; these describe the n_inserted rows (the slicing the Python facade
 * does at gpumap.py:148-162). */
int gm_get_view(gm_handle *h, int32_t which, gm_view *out);

/* Same array as a DLManagedTensor* (DLPack, device type kDLROCM = 10) whose
 * deleter keeps the handle's memory alive (cai_wrapper.cuh:39-62 exports a
 * non-owning __cuda_array_interface__ instead). */
int gm_export_dlpack(gm_handle *h, int32_t which, void **dl_managed_tensor);

/* append_data from device memory (rows x cols float32, row-major): the same
 * contract as gm_append_data for callers whose chunks are already in HBM. */
int gm_append_data_device(gm_handle *h, const float *device_rows,
                          uint64_t rows, uint64_t cols);

/* Host copies of the public arrays (n_inserted rows). */
int gm_copy_knn(gm_handle *h, uint32_t *idx, float *dist);
int gm_copy_embedding(gm_handle *h, float *emb);
int gm_set_embedding(gm_handle *h, const float *emb);

/* Message of the last failure on this thread ("" if none). */
const char *gm_last_error(void);

/* ---- stage entry points (no counterpart in the reference; they expose the
 * engine's stages on caller-given inputs for parity tests and profiling) */

/* Raw 32-slot neighbour pools of the inserted rows, packed keys. */
int gm_debug_pool(gm_handle *h, uint64_t *keys);

/* Fuzzy set + trim for the given batch rows (distinct) on the handle's
 * current graph (what one embed call builds before the SGD).  head/tail/w
 * need room for 2 * n_batch * n_neighbors entries; sigma/rho (may be NULL)
 * n_inserted. */
int gm_debug_fuzzy(gm_handle *h, const uint32_t *batch, uint64_t n_batch,
                   int32_t *head, int32_t *tail, float *w, double *sigma,
                   double *rho, int64_t *nnz);

/* SGD epochs of the engine on a caller-given embedding (n x c, host) and COO
 * (host), with the engine's kernels: per_epoch 0 runs all epochs in one
 * grid-wide launch of the kernel that fits the edge set (one launch per
 * epoch when none can be launched), 1 one launch per epoch, 2 the streaming
 * grid-wide kernel, 3 the row-owner kernel the engine uses for edge sets
 * beyond the resident grid, at any size.  A slot with head = tail = -1 is
 * empty (the
 * engine's own COO keeps each edge's reverse in the slot after it, or an
 * empty slot). */
int gm_debug_sgd(float *emb, uint64_t n, int32_t c, const int32_t *head,
                 const int32_t *tail, const float *w, int64_t nnz,
                 int32_t n_epochs, int32_t nsr, float a, float b, float alpha0,
                 float gamma, uint64_t seed, int32_t per_epoch);

/* gm_debug_sgd without any device allocation: the same SGD with the same
 * kernels and launch shapes, on the caller's device arrays (emb n x c float32,
 * updated in place; head, tail int32 and w float32 of nnz COO slots, all in
 * device memory), with the schedule, head degrees, rounding grid, the
 * grid-wide kernels' barrier words and row records in one static device
 * array per device (about 48 MiB, held until the process ends).  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Refuses (GM_ERR_INVALID) n above
 * GM_SGD_NOALLOC_MAX_ROWS, nnz above GM_SGD_NOALLOC_MAX_EDGES and COO
 * indices out of range (checked on the device).  Blocking; calls in one
 * process are serialised with the sessions' device work by the engine's
 * device lock, so they do not race on the shared array, but each call
 * overwrites what the previous one left in it. */
#define GM_SGD_NOALLOC_MAX_ROWS (1u << 20)
#define GM_SGD_NOALLOC_MAX_EDGES (1u << 20)
int gm_sgd_noalloc(float *emb, uint64_t n, int32_t c, const int32_t *head,
                   const int32_t *tail, const float *w, int64_t nnz,
                   int32_t n_epochs, int32_t nsr, float a, float b,
                   float alpha0, float gamma, uint64_t seed,
                   int32_t per_epoch);

/* Per-kernel measurement (no counterpart in the reference, which only times
 * whole operations, impl.cuh:159-203).  With profiling on, HIP events on the
 * session's stream bracket the timed kernels and device counters record their
 * algorithmic work:
 *   GM_KSTAT_JOIN        local join; work = {list members, distance pairs,
 *                        queued pool offers, pairs re-ranked in FP32 after
 *                        the BF16 screen (GM_KNN_NNDESCENT_BF16)}
 *   GM_KSTAT_SGD         SGD with a register-resident grid-wide kernel
 *                        (one launch per embedding op: for up to 512 groups
 *                        of 64 edges k_sgd_pipe, or k_sgd_waves with
 *                        GPUMAP_AMD_SGD_PIPE=0, else k_sgd_fused) or, when
 *                        a grid-wide launch is unavailable, one launch per
 *                        epoch; work = {positive samples, negative samples,
 *                        COO slot-epochs scanned (edges and the empty slots
 *                        of the twin-paired COO), launches that ran
 *                        k_sgd_pipe or k_sgd_waves}
 *   GM_KSTAT_INSERT      insertions (all kernels of one insertion); work =
 *                        {(batch row, row) pairs, BF16 screening survivors
 *                        re-ranked in FP32, insertions redone by the FP32
 *                        tiles after a survivor-list overflow}
 *   GM_KSTAT_SGD_STREAM  SGD of edge sets larger than the register-resident
 *                        grid: the row-owner kernel (or, without its
 *                        scratch, the streaming kernel); work as
 *                        GM_KSTAT_SGD
 *   GM_KSTAT_GRAPH       the embedding op's kernel sequences run as HIP
 *                        graphs (counted with or without profiling);
 *                        launches = graph launches, work[0] = executable
 *                        graphs instantiated (the others were updated in
 *                        place), device_ms unused */
typedef enum gm_kstat {
  GM_KSTAT_JOIN = 0,
  GM_KSTAT_SGD = 1,
  GM_KSTAT_INSERT = 2,
  GM_KSTAT_SGD_STREAM = 3,
  GM_KSTAT_GRAPH = 4,
  GM_KSTAT_COUNT = 5
} gm_kstat;

typedef struct gm_kernel_stats {
  uint64_t launches;
  double device_ms; /* summed event time over the launches */
  uint64_t work[4];
} gm_kernel_stats;

/* Profiling adds no host synchronisation to gm_run: event pairs are folded
 * after run()'s own per-phase synchronisation and the device counters are
 * read back by gm_get_kernel_stats (which synchronises the session). */
int gm_set_profiling(gm_handle *h, int32_t enable);
int gm_get_kernel_stats(gm_handle *h, int32_t which, gm_kernel_stats *out);

/* The BF16 screen's launch shapes chosen so far by the process's shape
 * tuner.  No counterpart in the reference, whose launch shapes are constants
 * (insertion_kernel.cuh:961-969, update_kernel.cuh:386-396).  Per key
 * (device, screen "knn_mode", padded feature count "db", row count as
 * floor(log4 n) "rows_log4",
 * 128-row batch tiles "batch_tiles"): the candidates timed so far
 * ({"bsplit", "lbscan", "ms"}, ms < 0 when not timed yet) and the chosen
 * one's index ("best", -1 while tuning), as a NUL-terminated JSON array
 * written to buf when buf holds it.  Returns the text's length without the
 * NUL, whether or not it fitted, or a negative gm_status. */
int64_t gm_screen_shapes(char *buf, uint64_t cap);
int gm_reset_kernel_stats(gm_handle *h);

/* Philox4x32-10 of the device code, for known-answer tests. */
int gm_debug_philox(const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4]);

/* The FP8 screen's float -> E4M3 rounding (round to nearest, ties to even,
 * saturating at 448) of n values, on the host (on_device 0, no GPU needed)
 * or in a kernel (on_device 1), for known-answer tests. */
int gm_debug_e4m3(const float *in, uint8_t *out, uint64_t n,
                  int32_t on_device);

#ifdef __cplusplus
}
#endif
#endif
