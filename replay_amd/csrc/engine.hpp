// engine.hpp -- definitions shared by the translation units of
// libreplay_hip: launch constants, the device-side argument structs, the
// host-side context and the error helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>
#include <vector>

#include "device_math.hpp"
#include "ingest.hpp"
#include "replay_hip.h"

namespace rp {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr unsigned kXcds = 8;  // MI355X accelerator complex dies
constexpr int32_t kNever = INT32_MAX;
constexpr int kMaxSegs = 16;

struct Segs {  // runs of the tabulated class over n = 1..T for new infections
  int32_t n;
  int32_t a[kMaxSegs];    // first offset of the run
  int32_t b[kMaxSegs];    // one past the last offset
  int32_t cls[kMaxSegs];  // 1 E, 2 I, 3 R
};

// Constants of one network on the device.  Its replicas occupy W words per
// node: node i's first word is wsh + i * W (the network's words are one
// contiguous run, so wsh is its first word less node_off * W), and replica r
// is column moff + r of the per-replica count arrays.
struct NetDev {
  float log_base;    // logf(1 - infect_prob), computed on the host
  float p0;          // initial_infected
  int32_t node_off;  // first row
  uint32_t stream;   // Philox counter word 3
  int32_t n_on, n_off, n_s;  // flip offsets of a new infection (kNever: none)
  int32_t m_on, m_off, m_s;  // the same for the initial seeds
  int32_t M, W;      // replicas, and words per node
  int32_t moff;      // first replica column
  int32_t blk0;      // first block of update_kernel's grid
  int64_t wsh;       // first word - node_off * W
  int64_t woff, nwords;  // first word and word count of the network
};

struct StepArgs {
  const int32_t *rowptr;      // slice epoch rowptr (N+1), epoch-relative entries
  const int4 *entries;        // slice epoch entries {col, ts, te, rank}
  const uint2 *entries8;      // the same entries as {col | ts << 24, ts >> 8 | te << 12}, or null
  const uint64_t *I_cur;      // infectious bits at iteration k
  uint64_t *I_next;           // infectious bits at iteration k+1
  uint64_t *S;                // susceptible bits, k -> k+1 in place
  uint64_t *X;                // infection ring, depth x NW
  const uint64_t *X_init;
  int32_t *entry;             // [N][Mp]
  int32_t *inf_count;         // [T+1][n_net][M]
  int32_t *P;                 // prefix sums of inf_count, [T+2][n_net][M]
  const float *uniforms;      // injected stream for this slot, [M][N], or null
  const NetDev *nets;
  const int32_t *net_of;      // node -> network
  const int32_t *node_word;   // node -> its first word (wsh + i * W of its network)
  const uint32_t *rank_limit; // [n_net] of this window, or null (no cap bites)
  const float *p_tab;         // [n_net][tab_max+1]: P of c whole windows, or null
  const int4 *upd;            // update_kernel block -> {network (-1: none), first word,
                              // end of the network's words, 0}
  uint64_t *F;                // compact ring: [depthF][update waves] flag words, or null
  int32_t depthF;             // slots of F
  int32_t record;             // 1: each infection is written to entry[] and counted in
                              // inf_count as it is drawn; 0 (the ring keeps every
                              // iteration): both are derived from the ring when read
  int32_t tab_max;
  int32_t long_row;           // rows with more entries go to long_row_kernel
  uint64_t seed;
  int64_t NW;                 // words of all networks
  int32_t N;                  // nodes of all networks
  int32_t M;                  // replicas per network when every network has the same
                              // count (the injected-stream and exposure layouts), else 0
  int32_t W;                  // most words per node of any network
  int32_t Wg;                 // greatest common divisor of the networks' words per node
  int32_t G;                  // replica columns of all networks (sum of their M)
  int32_t n_net, depth;
  int32_t k;
  int64_t xk_off;             // (k mod depth) * NW: first word of ring slot k
  int32_t ws_rel, we_rel;
  float scale;
  uint32_t slot;
  int32_t xcd_split;          // expo_kernel: XCD x takes the x-th contiguous eighth of the nodes
  const int4 *rowinfo;        // epoch's {row start, row end, network, 0} per node, so that
                              // expo_kernel reads a node's bounds and network in one load
};

}  // namespace rp

// One network of a context, on the host.
struct NetHost {
  rp::Network net;
  rp::Plan plan;
  rp::Slices slices;
  float p0 = 0.0f, beta = 0.0f, log_base = 0.0f;
  int32_t Ed = 0, Id = 0, Rd = 0, upper = 0, medium = 0, lower = 0;
  uint32_t stream = 0;
  int32_t node_off = 0;
  // timer tables
  std::vector<int32_t> lut_new, lut_init;  // c(n), n = 0..T
  int32_t M = 0, W = 0;      // replicas and words per node
  int32_t moff = 0;          // first replica column
  int64_t woff = 0;          // first word
  int32_t n_on = rp::kNever, n_off = rp::kNever, n_s = rp::kNever;
  int32_t m_on = rp::kNever, m_off = rp::kNever, m_s = rp::kNever;
  rp::Segs segs{};
  std::vector<int8_t> class_init;  // class of c_init(j), j = 0..T
};

struct replay_ctx {
  replay_params p{};
  std::vector<NetHost> nets;
  int32_t n_net = 0;
  // M: largest replica count of a network; W: its words per node; G: replica
  // columns of all networks; NW: words of all networks; uniform: every
  // network has M replicas (then node i's words start at i * W)
  int32_t N = 0, M = 0, W = 0, Wg = 0, G = 0, T = 0, depth = 1;
  int64_t NW = 0;
  bool uniform = true;
  int32_t upd_blocks = 0;  // blocks of update_kernel's grid that hold words
  // compact ring: the ring keeps one slot (depth 1) and d_F one flag bit per
  // word and iteration over depthF slots
  bool compact = false;
  int32_t depthF = 0;
  uint64_t *d_F = nullptr;
  // history ring: X keeps every iteration (depth T + 1), the step kernels
  // record nothing per infection, and inf_count / entry are derived from X
  // when counts or states are read; tallied = iterations whose inf_count is
  // derived
  bool history = false;
  int32_t tallied = 0;
  int32_t rebuilt = 0;  // iterations whose infections entry[] reflects (history ring)
  // history ring: entry[] holds the reset's seeds (0) and -1 elsewhere; the
  // reset leaves it to the first rebuild, which state reads start
  bool entry_ready = true;
  int32_t *d_tally = nullptr;  // tally blocks: network << 16 | first word of 4
  // merged plan and slices
  std::vector<int64_t> iter_time;
  std::vector<int32_t> iter_window;
  std::vector<int64_t> win_ws, win_we;
  std::vector<float> win_scale;
  std::vector<int32_t> win_epoch;
  std::vector<uint8_t> win_int;
  std::vector<int64_t> epoch_t0, epoch_entry0;
  std::vector<int32_t> epoch_maxrow;  // longest row of each epoch that is not long
  // long rows: epoch e lists long_off[e] .. long_off[e+1] in d_long_rows,
  // the first long_n_int[e] of them with exact integer sums
  std::vector<int64_t> long_off;
  std::vector<int32_t> long_n_int;
  // segments of the long rows: epoch e's segments are hub_seg_off[e] ..
  // hub_seg_off[e+1] of d_hub_seg_row (each segment's long row, relative to
  // the epoch's list); d_hub_seg0 + long_off[e] + e holds the first segment
  // of each of the epoch's long rows (relative to the epoch) and their total
  std::vector<int64_t> hub_seg_off;
  int64_t n_entries = 0;
  int32_t windows_per_epoch = 1;
  bool any_cut = false;
  int truncated = 0;
  int64_t n_contacts = 0, nlists = 0;
  int32_t variant = REPLAY_V_WORD;
  float tune_ms[4] = {-1.0f, -1.0f, -1.0f, -1.0f};  // last autotune: WORD, NODE4, NODE16, SLICED
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  // device
  int4 *d_entries = nullptr;
  uint2 *d_entries8 = nullptr;  // 8-byte copy of d_entries, where every entry fits one
  int32_t *d_rowptr = nullptr;
  int4 *d_rowinfo = nullptr;  // [n_epochs][N] {row start, row end, network, 0}
  uint64_t *d_S = nullptr;
  uint64_t *d_I[2] = {nullptr, nullptr};
  uint64_t *d_X = nullptr;
  uint64_t *d_Xinit = nullptr;
  int32_t *d_entry = nullptr;
  int32_t *d_inf = nullptr;
  int32_t *d_P = nullptr;
  int32_t *d_init_count = nullptr;
  int32_t *d_counts = nullptr;
  int8_t *d_class_init = nullptr;
  rp::Segs *d_segs = nullptr;
  rp::NetDev *d_nets = nullptr;
  int32_t *d_net_of = nullptr;
  int4 *d_upd = nullptr;          // update_kernel blocks: {network, first word, end, 0}
  int32_t *d_node_word = nullptr; // first word of each node
  int32_t *d_col_net = nullptr;   // network of each replica column
  uint32_t *d_rank_limit = nullptr;  // [n_windows][n_net] when a cap bites
  float *d_p_tab = nullptr;          // [n_net][tab_max+1]
  int32_t *d_long_rows = nullptr;    // long rows of every epoch
  int32_t *d_hub_seg0 = nullptr;
  int32_t *d_hub_seg_row = nullptr;
  uint64_t *d_hub_total = nullptr;   // [most long rows of an epoch][W][64]
  uint32_t *d_hub_part = nullptr;    // [most segments of an epoch][W][64]
  int64_t hub_total_n = 0;           // elements of d_hub_total
  int32_t tab_max = 0;
  float *d_uniforms = nullptr;
  int cur = 0;
  int32_t steps_done = 0;
  int32_t counts_done = -1;
  std::string err;
};

namespace rp {

// Message of the last failed open on this thread (no context exists yet).
extern thread_local std::string g_open_error;

inline int fail(replay_ctx *ctx, int code, const std::string &msg) {
  if (ctx)
    ctx->err = msg;
  else
    g_open_error = msg;
  return code;
}

#define HIP_OR_FAIL(ctx, call)                                                         \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail((ctx), REPLAY_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dev_alloc(replay_ctx *ctx, T **ptr, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(ptr), count * sizeof(T));
  if (e != hipSuccess)
    return fail(ctx, REPLAY_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return REPLAY_OK;
}

template <typename T>
int upload(replay_ctx *ctx, T **ptr, const T *src, size_t count) {
  int rc = dev_alloc(ctx, ptr, count);
  if (rc) return rc;
  if (count)
    HIP_OR_FAIL(ctx, hipMemcpyAsync(*ptr, src, count * sizeof(T), hipMemcpyHostToDevice,
                                    ctx->stream));
  return REPLAY_OK;
}


// Device arguments of iteration k: the slice epoch and window, the state
// buffers of the current parity, the draw slot.
StepArgs step_args(replay_ctx *ctx, int32_t k);

typedef int v4i __attribute__((ext_vector_type(4)));

// Contact entries e .. e+63 of a row that ends at e1 (e < e1), one per lane;
// lanes past the row re-read its last entry and are masked by the reader.
// Streamed once per iteration: kept out of the caches' way.
__device__ __forceinline__ v4i load_chunk(const int4 *entries, int e, int e1, int lane) {
  return __builtin_nontemporal_load(reinterpret_cast<const v4i *>(entries + min(e + lane, e1 - 1)));
}

typedef unsigned int v2u __attribute__((ext_vector_type(2)));

// The same chunk from the 8-byte copy of the entries.
__device__ __forceinline__ v2u load_chunk(const uint2 *entries8, int e, int e1, int lane) {
  return __builtin_nontemporal_load(reinterpret_cast<const v2u *>(entries8 + min(e + lane, e1 - 1)));
}

// Fields of a packed entry: {col, ts, te, rank} from the 16-byte form; from
// the 8-byte form {col | ts << 24, ts >> 8 | te << 12} (24-bit column,
// 20-bit epoch-relative times), whose contexts have no rank cap.
struct EntryFields {
  int col, ts, te;
  uint32_t rank;
};
__device__ __forceinline__ EntryFields fields(const v4i &en) {
  return EntryFields{en.x, en.y, en.z, static_cast<uint32_t>(en.w)};
}
__device__ __forceinline__ EntryFields fields(const v2u &en) {
  return EntryFields{static_cast<int>(en.x & 0xFFFFFFu),
                     static_cast<int>((en.x >> 24) | ((en.y & 0xFFFu) << 8)),
                     static_cast<int>(en.y >> 12), 0u};
}


}  // namespace rp
