/*
 * oracle.h -- CPU restatement of Replay's Monte Carlo SEIR step loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * in replay_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product never links, loads or calls it.
 *
 * Every function restates one piece of the reference (paths relative to the
 * reference repository root):
 *   tool/temporal_contact_matrix.c      ("tcm")  parser, node map, pair packing
 *   tool/ncls/ncls/src/intervaldb.c     ("idb")  nested containment list
 *   tool/gpu_cpu_temporal_sim.cu        ("cu")   window arithmetic, CSR build,
 *                                                CPU twin kernels, CSV writer
 *   tool/Eigen/src/SparseCore/SparseDenseProduct.h:34-70  exposure order
 *
 * Deliberate, documented divergences from the reference CPU path:
 *   - uniforms come from Philox4x32-10 keyed by a seed (or an injected
 *     stream), not srand(time(NULL)) + rand() (cu:365-367, 468-475);
 *   - exp() in the infection probability is the project exponential
 *     or_expf (double-evaluated, rounded to float); or_params.use_libm_exp
 *     switches to glibc expf for the reference-faithful CPU timing leg;
 *   - edge endpoints that would receive an index >= N (undefined behaviour
 *     in the reference) are rejected with OR_E_NODE_RANGE.
 */
#ifndef REPLAY_ORACLE_H
#define REPLAY_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field layout as the reference IntervalMap (intervaldb.h:11-18). */
typedef struct {
  int64_t start;
  int64_t end;
  int64_t target_id; /* (min_index << 32) | max_index, tcm:80-85 */
  int64_t sublist;
} or_interval;

/* Same field layout as the reference SublistHeader (intervaldb.h:25-28). */
typedef struct {
  int start;
  int len;
} or_sublist;

typedef struct {
  int n_nodes;         /* NODE_LIST count (tcm:242) */
  int n_edges;         /* valid contact lines */
  int n_indexed;       /* distinct endpoint indices handed out */
  or_interval *iv;     /* intervals; after or_build_ncls: NCLS order */
  int ncls_built;
  int n_top;           /* top-level list length (dbw.n) */
  int nlists;          /* sublist count as returned by the modified build */
  or_sublist *sub;     /* nlists headers (or one dummy) */
  int64_t *index_to_id;/* node index -> id from the file */
} or_graph;

typedef struct {
  int32_t M;
  float step_size;
  int32_t iterations;
  float initial_infected;
  float infect_prob;
  int32_t exposed_duration;
  int32_t infectious_duration;
  int32_t resistant_duration;
  int64_t start_time;
  int64_t static_network_duration;
  uint64_t seed;
  int32_t threads;      /* >1 enables OpenMP rows when nnz > 20000 */
  int32_t use_libm_exp; /* 1: glibc expf (reference-faithful timing) */
  const float *uniforms;/* optional injected stream, see or_uniform() */
  uint32_t stream;      /* Philox counter word 3 */
} or_params;

enum {
  OR_OK = 0,
  OR_E_IO = -1,
  OR_E_HEADER = -2,
  OR_E_NODE_ID = -3,
  OR_E_DUP_ID = -4,
  OR_E_NO_NODES = -5,
  OR_E_NO_EDGES = -6,
  OR_E_COUNT = -7,
  OR_E_NODE_RANGE = -8,
  OR_E_NOMEM = -9,
  OR_E_PARAM = -10,
  OR_E_OVERFLOW = -11
};

/* ---- ingestion (tcm) ---- */
int64_t or_parse_iso8601(const char *s);
int64_t or_pack_pair(int a, int b);
int or_parse_file(const char *path, or_graph **out);
int or_graph_from_contacts(int n_list, const int64_t *node_list, int n_contacts,
                           const int32_t *src_id, const int32_t *tgt_id,
                           const int64_t *start, const int64_t *end,
                           or_graph **out);
void or_graph_free(or_graph *g);
int or_graph_n_nodes(const or_graph *g);
int or_graph_n_edges(const or_graph *g);
const or_interval *or_graph_intervals(const or_graph *g);

/* ---- NCLS (idb) ---- */
int or_build_ncls(or_graph *g);
int or_graph_n_top(const or_graph *g);
int or_graph_nlists(const or_graph *g);
int or_find_intervals(const or_graph *g, int64_t ws, int64_t we, int nbuf,
                      or_interval *buf, int *n_out);
/* The hit query the step loop makes: cap = nlists, buffer sized as cu:973-974. */
int or_step_hits(const or_graph *g, int64_t ws, int64_t we, or_interval *buf,
                 int bufcap, int *n_out);

/* ---- window arithmetic (cu:860-959) ---- */
typedef struct {
  int valid;
  int64_t cached_start;
  int cached_steps;
} or_window_cache;

int or_window(const or_params *p, or_window_cache *cache, int iter,
              int64_t *current_time, int64_t *ws, int64_t *we, float *scale);

/* ---- CSR build (cu:238-350) ---- */
int or_build_csr(const or_interval *hits, int n_hits, int n_nodes, int64_t ws,
                 int64_t we, float weight_scale, int32_t *rowptr,
                 int32_t *colind, float *val, int64_t *nnz);
/* CSR of iteration `iter` (NCLS query + build), arrays sized by caller:
   rowptr n_nodes+1, colind/val at least 2*max(nlists,1).  */
int or_step_csr(const or_graph *g, const or_params *p, int iter, int32_t *rowptr,
                int32_t *colind, float *val, int64_t *nnz, int64_t *ws,
                int64_t *we, float *scale);

/* ---- step pieces (cu:358-500, Eigen SparseDenseProduct.h:34-70) ---- */
float or_expf(float x);
uint32_t or_philox_word(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2,
                        uint32_t c3);
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]);
float or_uniform(const or_params *p, int slot, int node, int replica, int n_nodes);
void or_init_countdown(const or_params *p, int n_nodes, int32_t *c);
void or_infectious_vector(const or_params *p, int n_nodes, const int32_t *c,
                          float *inf);
void or_spmm(const int32_t *rowptr, const int32_t *colind, const float *val,
             int64_t nnz, int n, const float *B, float *C, int m, int threads);
void or_spmm_f64(const int32_t *rowptr, const int32_t *colind, const float *val,
                 int n, const float *B, double *C, int m);
void or_infection_probability(const or_params *p, int n_nodes, const int32_t *c,
                              float *E);
void or_monte_carlo(const or_params *p, int slot, int n_nodes, const float *P,
                    int32_t *c);
void or_update_countdown(const or_params *p, int n_nodes, int32_t *c);
/* cu:505-568 classification; counts per sim: exposed, infectious, resistant,
   susceptible (the CSV column order). */
void or_classify_counts(const or_params *p, int n_nodes, const int32_t *c,
                        int32_t *counts4);
int or_state_class(const or_params *p, int32_t c); /* 0 S, 1 E, 2 I, 3 R */

/* ---- the whole loop (cu:727-1150) ---- */
/* counts: (iterations+1) * M * 4 int32, block 0 = initial state.
   final_state (optional): N*M int32 column-major (idx = node + sim*N).
   summary/node (optional): CSV streams written in the cu:505-568 format. */
int or_simulate(const or_graph *g, const or_params *p, int32_t *counts,
                int32_t *final_state, FILE *summary, FILE *node,
                double *loop_seconds);
int or_simulate_csv(const or_graph *g, const or_params *p,
                    const char *summary_path, const char *node_path,
                    int32_t *counts, double *loop_seconds);

#ifdef __cplusplus
}
#endif
#endif
