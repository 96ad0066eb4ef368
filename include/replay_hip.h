/*
 * replay_hip.h -- C ABI of libreplay_hip, the MI355X (gfx950) engine for
 * Replay's Monte Carlo SIR/SEIR step loop.
 *
 * The reference exposes no library API: its boundary is the `replay`
 * executable (tool/gpu_cpu_temporal_sim.cu:570-1185) plus the C functions it
 * shares through extern "C" headers (tool/temporal_contact_matrix.h:59-94,
 * tool/ncls/ncls/src/intervaldb.h:76-122).  Each entry point below names the
 * part of that code it replaces.  Conventions:
 *   - plain C types and caller-allocated buffers, so ctypes / cgo / JNI can
 *     bind it; no torch or HIP types cross the boundary;
 *   - every call returns REPLAY_OK (0) or a negative REPLAY_E_* code and never
 *     exits the process (the reference calls exit() inside, e.g.
 *     gpu_cpu_temporal_sim.cu:59-67, temporal_contact_matrix.c:308-312);
 *   - one replay_ctx owns its device buffers; contexts share no globals, so
 *     different host threads may drive different contexts;
 *   - state layout at the boundary is the reference's: N x M int32 countdowns,
 *     column-major, idx = node + sim * N (gpu_cpu_temporal_sim.cu:754).
 */
#ifndef REPLAY_HIP_H
#define REPLAY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REPLAY_HIP_ABI_VERSION 3

enum {
  REPLAY_OK = 0,
  REPLAY_E_ARG = -1,        /* bad argument / parameter */
  REPLAY_E_IO = -2,         /* file cannot be opened or read */
  REPLAY_E_FORMAT = -3,     /* NODE_LIST header missing or malformed */
  REPLAY_E_NODE_ID = -4,    /* invalid id in NODE_LIST */
  REPLAY_E_DUP_ID = -5,     /* duplicate id in NODE_LIST */
  REPLAY_E_NO_EDGES = -6,   /* no valid contact lines */
  REPLAY_E_COUNT = -7,      /* more contacts parsed than counted (tcm:308-312) */
  REPLAY_E_NODE_RANGE = -8, /* more distinct endpoints than NODE_LIST ids */
  REPLAY_E_NOMEM = -9,      /* host or device allocation failed */
  REPLAY_E_DEVICE = -10,    /* HIP runtime error */
  REPLAY_E_OVERFLOW = -11,  /* time arithmetic overflow (cu:861-900) */
  REPLAY_E_UNSUPPORTED = -12, /* input the engine does not handle */
  REPLAY_E_STATE = -13      /* call out of sequence */
};

/* Simulation parameters: the reference's argv globals (cu:24-56, 621-699). */
typedef struct replay_params {
  int32_t M;                  /* --M, parallel simulations (default 1000) */
  float step_size;            /* --step-size seconds (default 3600) */
  int32_t iterations;         /* --iterations (default 42) */
  float initial_infected;     /* --initial-infected probability */
  float infect_prob;          /* --infect-prob per hour of contact */
  int32_t exposed_duration;   /* --exposed-duration seconds */
  int32_t infectious_duration;/* --infectious-duration seconds */
  int32_t resistant_duration; /* --resistant-duration seconds */
  int64_t start_time;         /* --start-time epoch (default 946713600) */
  int64_t static_network_duration; /* --static-network-duration (default 3600) */
  uint64_t seed;              /* Philox4x32-10 key (new: the reference seeds by clock) */
  int32_t device;             /* HIP device ordinal */
  int32_t windows_per_epoch;  /* slice packing; <= 0 picks automatically */
  int64_t max_slice_bytes;    /* device budget for packed 16-byte slices (the 8-byte
                                 copy adds half again); <= 0: 32 GiB */
  uint32_t stream;            /* Philox stream id of this network (counter word 3) */
  int32_t flags;              /* REPLAY_F_* */
  int64_t max_ring_bytes;     /* device budget for the infection ring (one bit word per
                                 iteration of the longest timer offset); beyond it the
                                 compact form keeps one flag bit per word and iteration
                                 instead.  <= 0: 16 GiB */
} replay_params;

/* replay_params.flags */
#define REPLAY_F_GRAPH 1      /* full runs from iteration 0 replay one captured hipGraph */
#define REPLAY_F_AUTOTUNE 2   /* time the kernel variants at open and keep the fastest */
#define REPLAY_F_WIDE_ENTRIES 4 /* step kernels read the 16-byte packed entries even where
                                   their 8-byte form fits (comparisons and measurement) */
#define REPLAY_F_NO_HISTORY 8 /* keep only the infection ring's flip offsets even where every
                                 iteration would fit (comparisons and measurement) */

/* One network of a batch (replay_open_batch): its contacts, as in
   replay_open_contacts, its own epidemiological parameters and its own
   number of replicas. */
typedef struct replay_network {
  int32_t n_node_list;
  const int64_t *node_ids;    /* NULL: ids 0..n_node_list-1 */
  int64_t n_contacts;
  const int32_t *src;
  const int32_t *tgt;
  const int64_t *start;
  const int64_t *end;
  float initial_infected;
  float infect_prob;
  int32_t exposed_duration;
  int32_t infectious_duration;
  int32_t resistant_duration;
  uint32_t stream;            /* Philox stream id; equal ids give equal draws */
  int32_t M;                  /* replicas of this network; 0: replay_params.M */
} replay_network;

typedef struct replay_info {
  int32_t n_nodes;            /* N (NODE_LIST count; all networks of a batch) */
  int32_t M;                  /* replicas (a batch: the most of any network) */
  int64_t n_contacts;         /* valid contacts (intervals) */
  int64_t nlists;             /* hit cap of every window query (cu:988-989) */
  int32_t iterations;
  int32_t n_windows;          /* distinct network windows */
  int32_t n_epochs;           /* packed slice epochs */
  int32_t windows_per_epoch;
  int64_t n_entries;          /* packed directed contact entries */
  int64_t slice_bytes;        /* device bytes of packed slices */
  int64_t state_bytes;        /* device bytes of per-replica state */
  int32_t steps_done;         /* iterations run since the last reset */
  int32_t truncated_windows;  /* windows whose query hits the cap */
  int64_t step_bytes;         /* algorithmic HBM bytes of one step-kernel launch */
  int32_t ring_depth;         /* iterations of infection bit words kept on device
                                 (1 in the compact form) */
  int32_t n_networks;         /* networks in the context (1 unless batched) */
  int32_t variant;            /* step-kernel variant in use (REPLAY_V_*) */
  int32_t graph_ready;        /* a captured full-run hipGraph exists */
  int32_t entry_bytes;        /* bytes per packed entry the step kernels stream:
                                 8 (the 8-byte copy) or 16 */
  int32_t replicas_total;     /* replica columns of all networks (the sum of their M) */
  int32_t ring_form;          /* REPLAY_RING_*: how the infection ring is kept */
  int32_t has_entries8;       /* 1: the context holds the 8-byte copy of the entries
                                 (what the marginal form streams); in what was
                                 padding, so the struct's size and layout are unchanged */
  int64_t ring_bytes;         /* device bytes of the infection ring (or its flags) */
  float tune_ms[4];           /* device ms of each candidate in the last replay_autotune,
                                 in the order WORD, NODE4, NODE16, SLICED; -1 where the
                                 variant does not apply or no autotune ran */
} replay_info;

/* replay_info.ring_form */
#define REPLAY_RING_OFFSETS 0 /* slots back to the longest timer offset; each infection
                                 recorded per replica as it is drawn */
#define REPLAY_RING_COMPACT 1 /* one slot plus a flag bit per word and iteration */
#define REPLAY_RING_HISTORY 2 /* every iteration; per-replica records derived when read
                                 (runs of up to 1024 iterations that fit max_ring_bytes) */

/* replay_info.variant */
#define REPLAY_V_WORD 0       /* one pass over a node's contacts per 64-replica word */
#define REPLAY_V_NODE4 4      /* one pass per 4 words, accumulated in registers */
#define REPLAY_V_NODE16 16    /* one pass per 16 words, accumulated in registers */
#define REPLAY_V_SLICED 1     /* one lane per (node, word): bit-sliced contact counts,
                                 whole-window slices only (others use NODE4/NODE16) */
#define REPLAY_V_BLOCK1 0x100 /* or'ed with NODE4 or NODE16: the exposure kernel in blocks
                                 of one wavefront instead of four (a schedule; the results
                                 are the same) */
#define REPLAY_V_XCD 0x200    /* or'ed with WORD, NODE4 or NODE16 (and BLOCK1): XCD x takes
                                 the x-th contiguous eighth of the nodes instead of every
                                 eighth block (a schedule; the results are the same) */

typedef struct replay_ctx replay_ctx;

/* ABI version (REPLAY_HIP_ABI_VERSION). */
int32_t replay_abi_version(void);

/* Fills the reference defaults of cu:25-27, 48-56; seed 0, device 0. */
void replay_params_init(replay_params *p);

/* Parse a graph file, index its contacts, plan the windows, pack the slices
   into HBM and draw the initial state.  Replaces
   parse_edgelist_build_intervals (tcm:148-321), build_interval_db_wrapper
   (tcm:377-392), the per-iteration find_intervals + build_csr_from_intervals
   (cu:973-1001) and initialize_countdown_vector (cu:73-90, 788-803). */
int32_t replay_open(const char *graph_path, const replay_params *p,
                    replay_ctx **out);

/* Same as replay_open for contacts already in memory: node_ids[n_node_list]
   is the NODE_LIST (NULL means ids 0..n_node_list-1); contact k joins ids
   src[k], tgt[k] over [start[k], end[k]) in Unix seconds.  Contacts with a
   start or end of -1 are skipped, as an unparsable time is (tcm:285-288). */
int32_t replay_open_contacts(int32_t n_node_list, const int64_t *node_ids,
                             int64_t n_contacts, const int32_t *src,
                             const int32_t *tgt, const int64_t *start,
                             const int64_t *end, const replay_params *p,
                             replay_ctx **out);

/* Several networks in one context: one step launch advances all of them
   (the batched form; the reference runs one network per process, with one
   global replica count, cu:25, 647-648).  Nodes of network n are rows
   node_off[n] .. node_off[n+1]-1 of the context and its replicas are the
   replica columns rep_off[n] .. rep_off[n+1]-1 (replay_get_replica_offsets);
   the shared parameters (step, iterations, windows, start time, seed,
   device, flags) come from p, the epidemiological ones and the replica count
   from each replay_network (M = 0 takes p->M).  A network gets the same draws
   as in a single-network context with p->stream equal to its stream id and
   p->M its replica count, so both forms give identical results. */
int32_t replay_open_batch(int32_t n_networks, const replay_network *nets,
                          const replay_params *p, replay_ctx **out);

/* node_off[0..n_networks]: first row of each network, then the total. */
int32_t replay_get_network_offsets(replay_ctx *ctx, int32_t *node_off,
                                   int32_t n);

/* rep_off[0..n_networks]: first replica column of each network (its sims
   in replay_get_counts), then the total.  No counterpart in the reference,
   which has one network per process. */
int32_t replay_get_replica_offsets(replay_ctx *ctx, int32_t *rep_off, int32_t n);

/* Times `steps` iterations of each applicable step-kernel variant (from a
   reset state, after an untimed warm-up, with eager launches even when
   REPLAY_F_GRAPH is set; resetting again afterwards), keeps the fastest and
   returns its REPLAY_V_* code in *variant; replay_info.tune_ms holds the
   times.  A captured graph is dropped (it is recaptured with the chosen
   variant by the next full run).  No counterpart in the reference: the
   tile/schedule choice cuSPARSE makes internally (cu:1039-1048). */
int32_t replay_autotune(replay_ctx *ctx, int32_t steps, int32_t *variant);

/* Forces a step-kernel variant (REPLAY_V_*); REPLAY_E_ARG if it does not
   apply to this context. */
int32_t replay_set_variant(replay_ctx *ctx, int32_t variant);

/* Sizes and counters of an open context. */
int32_t replay_get_info(replay_ctx *ctx, replay_info *out);

/* Redraws the initial state (same seed => same draw) and clears the counts. */
int32_t replay_reset(replay_ctx *ctx);

/* Enqueue `steps` more iterations of the step loop (cu:859-1150) on the
   context's stream; returns once they are enqueued.  Each iteration is two
   kernels: exposure -> probability -> draw (one wavefront per node), then
   the timer update and the running prefix of infection counts (one thread
   per 64-replica word).  SEIR counts are reduced from those prefixes on the
   device when they are read. */
int32_t replay_run(replay_ctx *ctx, int32_t steps);

/* replay_run, bracketed by HIP events recorded on the context's stream;
   waits for the steps and stores their device time in *elapsed_ms.  For
   benchmarks: the events measure exactly the enqueued step kernels. */
int32_t replay_run_timed(replay_ctx *ctx, int32_t steps, double *elapsed_ms);

/* Wait for all enqueued work of the context. */
int32_t replay_sync(replay_ctx *ctx);

/* Per-iteration SEIR counts in CSV column order (exposed, infectious,
   resistant, susceptible) -- the reduction write_simulation_state performs
   (cu:505-568).  out holds (steps_done + 1) * replicas_total * 4 int32,
   ordered [iteration][replica column][class], i.e. [iteration][network][sim]
   [class] with network n's M_n sims from column rep_off[n]; block 0 is the
   initial state.  Synchronises. */
int32_t replay_get_counts(replay_ctx *ctx, int32_t *out, int64_t n_out);

/* Countdown state, N x M int32 column-major (cu:1087-1088); for a batch,
   N counts the rows of all networks (REPLAY_E_UNSUPPORTED when its networks
   have different replica counts).  Synchronises. */
int32_t replay_get_state(replay_ctx *ctx, int32_t *out, int64_t n_out);

/* Countdown state of network `net` alone, N_net x M_net int32 column-major
   (node - node_off[net] + sim * N_net).  Synchronises. */
int32_t replay_get_network_state(replay_ctx *ctx, int32_t net, int32_t *out,
                                 int64_t n_out);

/* Timestamp written with iteration k's block (cu:876, 1093-1095); k = -1
   gives the initial block's time. */
int64_t replay_iteration_time(replay_ctx *ctx, int32_t k);

/* Inject a uniform stream for parity tests: u[(slot * M + sim) * N + node],
   slot 0 = initial draw, slot k+1 = iteration k.  n must be
   (iterations + 1) * M * N.  NULL restores Philox.  Redraws the initial
   state.  REPLAY_E_UNSUPPORTED for a batch whose networks have different
   replica counts. */
int32_t replay_set_uniforms(replay_ctx *ctx, const float *u, int64_t n);

/* CSR of iteration k's network window as build_csr_from_intervals
   (cu:238-350) would produce it, evaluated on the device from the packed
   slices.  rowptr[N+1]; colind/val hold up to cap entries; *nnz receives
   the count (also when cap is too small, then REPLAY_E_ARG).  cap < 0
   counts only: rowptr and *nnz are filled, colind/val may be NULL. */
int32_t replay_get_slice(replay_ctx *ctx, int32_t k, int32_t *rowptr,
                         int32_t *colind, float *val, int64_t cap,
                         int64_t *nnz);

/* Exposure seconds E = A_k * I of the current state for iteration k's
   window, as cusparseSpMM (cu:1062-1064) produces it before
   compute_infection_probability: out[node + sim * N], N * M floats
   (REPLAY_E_UNSUPPORTED for a batch whose networks have different replica
   counts). */
int32_t replay_get_exposure(replay_ctx *ctx, int32_t k, float *out,
                            int64_t n_out);

/* replay_get_exposure without any device allocation: the same result, bit
   for bit, staged through the library's static per-device array of 2^24
   floats (64 MiB, held by each device the library runs on until the process
   ends) instead of a buffer allocated per call.  No counterpart in the
   reference; for callers that cannot allocate device memory (code under
   hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
   REPLAY_E_ARG when N * M exceeds 2^24.  Calls on one device share the
   array, so calls that overlap on one device race. */
int32_t replay_get_exposure_noalloc(replay_ctx *ctx, int32_t k, float *out,
                                    int64_t n_out);

/* ---- Differentiable marginal-probability form (no counterpart in the
   reference, which only samples).  Individual-level SEIR probabilities per
   node and parameter set over the same duration-weighted slices as the
   stochastic step: lam = 1 - exp(-beta * phi_i * E_i / 3600) with
   E_i = sum_j A_k[i,j] iota_j, exponential exposed and infectious sojourns
   (rates sigma, gamma per hour), initial iota = initial_infected.  Forward
   curves and a hand-written adjoint (transposed-slice product plus the
   transition adjoints), with checkpoints every checkpoint_every iterations
   recomputed during the backward pass.  All data pointers of the forward and
   backward calls are DEVICE pointers (fp32); stream is the hipStream_t the
   work is ordered on (NULL: HIP's default stream), so the caller's own
   producers and consumers of those buffers on that stream need no extra
   synchronisation.  Results are deterministic. */
typedef struct replay_marginal replay_marginal;

/* n_sets parameter sets over the context's network(s) and iterations. */
int32_t replay_marginal_create(replay_ctx *ctx, int32_t n_sets,
                               int32_t checkpoint_every, replay_marginal **out);

/* beta, sigma, gamma: [n_sets]; phi: [N] per-node susceptibility; curves out:
   [(iterations + 1)][n_sets][4] expected (S, E, I, R) counts. */
int32_t replay_marginal_forward(replay_marginal *m, const float *beta,
                                const float *sigma, const float *gamma,
                                const float *phi, float *curves, void *stream);

/* Gradients of a loss with dL/dcurves = grad_curves (same shape as curves)
   with respect to the parameters of the last forward call. */
int32_t replay_marginal_backward(replay_marginal *m, const float *grad_curves,
                                 float *grad_beta, float *grad_sigma,
                                 float *grad_gamma, float *grad_phi,
                                 void *stream);

void replay_marginal_destroy(replay_marginal *m);

/* ---- Benchmark input generator ----
   Replaces the group-structured generator executable
   (generate_networks/graph_generator.c:39-157, main.c:39-65).  Writes to
   path the file that generator writes for the same arguments on a glibc
   system: nodes 1..population in round-robin groups (i % groups), one draw
   per node pair and step from the C library's rand() stream after
   srand(seed), a contact when draw / RAND_MAX < prob_inner (same group) or
   prob_inter, two start/end draws base_time + step * step_size +
   rand() % step_size in ascending order, lines "u, v, ISO, ISO;" under a
   NODE_LIST header.  base_time is UTC seconds (the reference's main() reads
   its start argument as local time).  n_contacts (may be NULL) receives the
   number of contact lines; verbose != 0 prints the reference's progress
   lines.  REPLAY_E_ARG for a negative seed, steps or population, or a
   non-positive step size or group count; REPLAY_E_IO when path cannot be
   written.  Needs no device. */
int32_t replay_generate_groups(int64_t base_time, int32_t steps, int32_t step_size,
                               int32_t population, int32_t groups, double prob_inner,
                               double prob_inter, int32_t seed, const char *path,
                               int64_t *n_contacts, int32_t verbose);

/* Replaces the command-file generator contactgraphgen/contactgraphgen.py
   (process_command_file py:316-362, write_contact_graph py:365-381, main
   py:395-418).  Reads the ADDNODE / ADDEDGE commands in commands_path and
   writes to out_path the graph that script writes: NODE_LIST of the added
   ids in ascending order, then every pair u < v of each ADDEDGE clique at
   each start START + k DELTA (k < N for OCCURRENCES N, start < T for
   UNTIL T, one start when DELTA is 0 or no clause), duplicates removed, in
   (start, u, v, end) order, lines "u, v, ISO, ISO; ".  Rejected lines
   print the script's messages on stderr (unless quiet) and are skipped.
   n_contacts (may be NULL) receives the number of contact lines written.
   REPLAY_E_IO when the command file cannot be read or the graph cannot be
   written (nothing written), REPLAY_E_FORMAT for a command file that is not
   UTF-8, REPLAY_E_OVERFLOW when a time passes 9999-12-31T23:59:59Z (the
   file then ends before that line, as the script's does), REPLAY_E_NOMEM
   beyond 2^30 contact occurrences.  Needs no device. */
int32_t replay_contactgraph(const char *commands_path, const char *out_path,
                            int64_t *n_contacts, int32_t quiet);

/* Message of the last failed call on ctx (or of the last failed open on
   this thread when ctx is NULL). */
const char *replay_last_error(replay_ctx *ctx);

/* Static description of an error code. */
const char *replay_error_string(int32_t code);

void replay_close(replay_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* REPLAY_HIP_H */
