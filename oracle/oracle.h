/*
 * CPU oracle of the progressive UMAP hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker for the HIP engine in gpumap_amd/csrc.  Only the
 * test suite, the smoke() entry and the bench's cpu_baseline leg may load it;
 * the product never links it.
 *
 * It restates, in plain C++ (OpenMP for the timed baseline), the method of the
 * reference's engine (cpp/src/impl.cuh), its NN-Descent insertion
 * (cpp/src/kernels/insertion_kernel.cuh) and update
 * (cpp/src/kernels/update_kernel.cuh), the fuzzy simplicial set and the SGD
 * optimiser (cpp/src/kernels/embedding_kernel.cuh), with the documented
 * deviations listed in oracle.cpp's header comment.
 *
 * Pinning (what each function is checked against, tests/test_oracle_*.py):
 *   or_philox4x32          Random123 published known-answer vectors
 *   or_canon_dist          numpy FP64 squared L2 (1e-5 rel) + exact bit
 *                          reproduction of the sequential-FMA definition
 *   kNN (or_run/or_get_knn) FP64 brute-force kNN: recall@k, invariants
 *                          (sorted, unique, no self, distances == canonical)
 *   or_fuzzy               numpy FP64 restatement of smooth_knn_dist +
 *                          membership + union (1e-5 rel), bisection invariant
 *                          sum == log2(k-1) or floor active
 *   or_sgd_epochs          hand-derived single-edge gradient (closed form)
 *   embedding              trustworthiness@15 vs sklearn on the input
 */
#ifndef GPUMAP_ORACLE_H
#define GPUMAP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_params {
  uint32_t n_neighbors;
  uint32_t n_components;
  uint32_t negative_sample_rate;
  uint32_t max_epoch;
  float min_dist;
  float a;
  float b;
  float alpha;
  int32_t metric; /* 0 euclidean, 1 cosine */
  int32_t verbose;
  uint64_t n_instances;
  uint64_t n_features;
  int32_t n_threads; /* OpenMP threads for the timed baseline; 1 = serial */
  int32_t max_iter;  /* insertion NN-D rounds (reference: 10) */
  int32_t knn_mode;  /* 0 NN-Descent insertion (reference method),
                        1 exact insertion (brute force over all rows) */
} or_params;

typedef struct or_result {
  uint32_t n_insertion_ops, n_update_ops, n_embedding_ops;
  float insertion_time, update_time, embedding_time;
  int32_t insertion_completed, update_completed, embedding_completed;
  uint64_t n_inserted_instances, n_update_changed_edges,
      embedding_pending_edges, insertion_queue_size, update_queue_size,
      embedding_queue_size;
} or_result;

void *or_create(const or_params *p);
void or_destroy(void *h);
int or_append(void *h, const float *x, uint64_t rows, uint64_t cols);
int or_run(void *h, uint32_t n_ins, uint32_t n_upd, uint32_t n_emb,
           or_result *out);
/* public sorted top-k rows [n_inserted x k] */
void or_get_knn(void *h, uint32_t *idx, float *dist);
void or_get_embedding(void *h, float *emb);
void or_set_embedding(void *h, const float *emb);
/* raw 32-wide pools [n_inserted x 32] packed keys */
void or_get_pool(void *h, uint64_t *keys);

/* stage-level restatements used to pin the oracle and the HIP kernels */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2],
                   uint32_t out[4]);
float or_canon_dist(const float *a, const float *b, int d);
/* fuzzy set on a given public graph; returns nnz, writes COO (head, tail, w)
 * after trim; sigmas/rhos (may be NULL) receive all n rows */
int64_t or_fuzzy(const uint32_t *idx, const float *dist, uint64_t n, int k,
                 int euclidean, const uint32_t *batch, uint64_t n_batch,
                 int n_epochs, int32_t *head, int32_t *tail, float *w,
                 double *sigmas, double *rhos);
/* One gradient step of the SGD, as applied: kind 0 the positive sample
 * (other = the tail), 1 a negative sample (other = the drawn row), 2 a
 * negative draw equal to the head, skipped.  y is the head's local position
 * before the step, o the other row's position read, g the step added to the
 * head (alpha and clipping included); d2 and coef are the squared distance
 * and the gradient coefficient of the step. */
typedef struct or_sgd_event {
  int32_t epoch, edge, kind, other;
  float alpha, d2, coef;
  float y[4], o[4], g[4];
} or_sgd_event;

/* or_sgd run serially with every step recorded (up to cap events); returns
 * the number of steps; *rounding (may be NULL) receives the rounding grid */
int64_t or_sgd_trace(float *emb, uint64_t n, int c, const int32_t *head,
                     const int32_t *tail, const float *w, int64_t nnz,
                     int n_epochs, int nsr, float a, float b, float alpha0,
                     float gamma, uint64_t seed, or_sgd_event *events,
                     int64_t cap, float *rounding);
/* the SGD's rounding grid for a COO's head degrees (embedding_kernel.cuh:
 * 405-460) */
float or_rounding_factor(const int32_t *head, int64_t nnz, uint64_t n,
                         float alpha);
/* SGD epochs on a given COO (weights w, untrimmed allowed) */
void or_sgd(float *emb, uint64_t n, int c, const int32_t *head,
            const int32_t *tail, const float *w, int64_t nnz, int n_epochs,
            int nsr, float a, float b, float alpha0, float gamma,
            uint64_t seed, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
