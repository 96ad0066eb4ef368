// replay_hip.hip -- libreplay_hip: HIP kernels for gfx950 and the C ABI
// declared in include/replay_hip.h.
//
// Replica state is bit-packed.  A word (node i, q) holds 64 replicas
// 64q .. 64q+63; per word the device keeps
//   S       susceptible bits at the current iteration     (countdown == 0)
//   I[2]    infectious bits, double-buffered               (lower < c <= medium)
//   X ring  infection bits of each past iteration, depth D
//   X_init  bits seeded infected by the initial draw
// plus, per replica, entry[i][r] (iteration of the last infection, written
// only when one happens) and inf_count[k][r] (infections of iteration k).
//
// Why this is exact: after an infection the countdown follows one fixed
// sequence, c(n) = f^n(c0) with f(c) = (int)((float)c - step) clamped at 0
// (tool/gpu_cpu_temporal_sim.cu:153-175), the same for every node and
// replica of that kind (new infection: c0 = (int)(upper + step); initial
// seed: c0 = (int)(medium + step), cu:85).  The host tabulates c(n) once,
// with the same float operations, and derives the offsets at which a
// replica turns infectious (n_on), stops being infectious (n_off) and
// becomes susceptible again (n_S).  So
//   I(k+1) = I(k) ^ X(k+1-n_on) ^ X(k+1-n_off),  S(k+1) = S(k) & ~X(k) | X(k+1-n_S)
// and the exposed/infectious/resistant counts of iteration k are sums of
// inf_count over the offset ranges where the tabulated class
// (cu:530-550) is E, I or R.
//
// One step kernel per iteration: a block owns 256 consecutive words.  Every
// thread loads its word's S; words of nodes that have contacts in the current
// slice and still hold susceptible replicas are queued in LDS, and the
// block's wavefronts take them one at a time with one lane per replica:
// each walks the node's packed contact entries in ascending column order,
// adds each contact's overlap with the window where the neighbour's
// infectious bit is set, converts exposure to 1 - exp(log(1-p) * E/3600),
// draws a Philox uniform and ballots the infections.  Then every thread
// applies the flips of its word.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "device_math.hpp"
#include "ingest.hpp"
#include "replay_hip.h"

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int32_t kNever = INT32_MAX;
constexpr int kMaxSegs = 16;

struct Segs {  // runs of the tabulated class over n = 1..T for new infections
  int32_t n;
  int32_t a[kMaxSegs];    // first offset of the run
  int32_t b[kMaxSegs];    // one past the last offset
  int32_t cls[kMaxSegs];  // 1 E, 2 I, 3 R
};

struct StepArgs {
  const int32_t *rowptr;   // slice epoch rowptr (N+1), epoch-relative entries
  const int4 *entries;     // slice epoch entries {col, ts, te, rank}
  const uint64_t *I_cur;   // infectious bits at iteration k
  uint64_t *I_next;        // infectious bits at iteration k+1
  uint64_t *S;             // susceptible bits, k -> k+1 in place
  uint64_t *X_k;           // ring slot of iteration k
  const uint64_t *X_on;    // X(k+1-n_on) or null
  const uint64_t *X_off;   // X(k+1-n_off) or null
  const uint64_t *X_s;     // X(k+1-n_S) or null
  const uint64_t *X_init;
  int32_t *entry;          // [N][Mp]
  int32_t *inf_count;      // [T+1][M]
  int32_t *P;              // prefix sums of inf_count, [T+2][M]
  const float *uniforms;   // injected stream for this slot, [M][N], or null
  uint64_t seed;
  int64_t NW;              // N * W words
  int32_t N, M, W, Mp;
  int32_t k;
  int32_t ws_rel, we_rel;
  uint32_t rank_limit;
  float scale;
  float log_base;          // logf(1 - infect_prob), computed on the host
  int32_t c_new;           // countdown of a new infection, (int)(upper + step)
  uint32_t slot;
  // flips of local x (n == 1) and of the initial seeds at iteration k+1
  int32_t on_is_local, off_is_local, s_is_local;
  int32_t init_flip_I, init_set_S;
};

// Exposure seconds of (row i, replica lane) for one window.
// Integer mode (divisor 1, row weight sums < 2^24): overlaps are integers and
// every partial sum is exact, so the integer sum equals the reference's fp32
// sum in any order.  Float mode: contributions fl(ov * scale) of one column
// are merged in rank order (cu:299-309), and columns are added in ascending
// order from zero (Eigen SparseDenseProduct.h:64-70).
template <bool kIntMode>
__device__ __forceinline__ float row_exposure(const int32_t *rowptr, const int4 *entries,
                                              const uint64_t *I, int32_t W, int32_t ws_rel,
                                              int32_t we_rel, uint32_t rank_limit, float scale,
                                              int i, int q, int lane) {
  const int e0 = rowptr[i], e1 = rowptr[i + 1];
  if (kIntMode) {
    uint32_t acc = 0;
#pragma unroll 4
    for (int e = e0; e < e1; ++e) {
      const int4 en = entries[e];
      const int ov = min(en.z, we_rel) - max(en.y, ws_rel);
      if (ov > 0 && static_cast<uint32_t>(en.w) <= rank_limit) {
        const uint64_t m = I[static_cast<size_t>(en.x) * W + q];
        if (m != 0ull && ((m >> lane) & 1ull)) acc += static_cast<uint32_t>(ov);
      }
    }
    return static_cast<float>(acc);
  } else {
    float acc = 0.0f;
    int e = e0;
    while (e < e1) {
      const int col = entries[e].x;
      float w = 0.0f;
      bool any = false;
      for (; e < e1 && entries[e].x == col; ++e) {
        const int4 en = entries[e];
        const int ov = min(en.z, we_rel) - max(en.y, ws_rel);
        if (ov > 0 && static_cast<uint32_t>(en.w) <= rank_limit) {
          const float wc = static_cast<float>(ov) * scale;
          w = any ? w + wc : wc;
          any = true;
        }
      }
      if (any) {
        const uint64_t m = I[static_cast<size_t>(col) * W + q];
        if ((m >> lane) & 1ull) acc += w;
      }
    }
    return acc;
  }
}

// Integer-mode exposure of the words q0 .. q0+nq-1 (nq <= WT) of node i at
// once: each contact entry is read once and the neighbour's infectious words
// for the whole range are fetched together, so a node costs a few memory
// round trips instead of a few per word.
template <int WT>
__device__ __forceinline__ void node_exposure(int e0, int e1, const int4 *entries,
                                              const uint64_t *I, int32_t W, int32_t ws_rel,
                                              int32_t we_rel, uint32_t rank_limit, int q0,
                                              int nq, int lane, uint32_t (&acc)[WT]) {
#pragma unroll
  for (int q = 0; q < WT; ++q) acc[q] = 0;
  for (int e = e0; e < e1; ++e) {
    const int4 en = entries[e];
    const int ov = min(en.z, we_rel) - max(en.y, ws_rel);
    if (ov > 0 && static_cast<uint32_t>(en.w) <= rank_limit) {
      const uint64_t *mrow = I + static_cast<size_t>(en.x) * W + q0;
#pragma unroll
      for (int q = 0; q < WT; ++q) {
        if (q < nq) {
          const uint64_t m = mrow[q];
          acc[q] += static_cast<uint32_t>((m >> lane) & 1ull) * static_cast<uint32_t>(ov);
        }
      }
    }
  }
}

// Draw of one word's susceptible lanes given their exposure; returns the
// ballot of new infections and records them.
template <bool kInjected>
__device__ __forceinline__ uint64_t draw_word(const StepArgs &a, int ii, int qq, int lane,
                                              uint64_t ss, float E) {
  const int r = qq * kWave + lane;
  const bool sus = (ss >> lane) & 1ull;
  bool infect = false;
  // E == 0 gives P = 0 (or NaN), never above a uniform in [0, 1)
  if (sus && (kInjected || E > 0.0f)) {
    const float P = 1.0f - rp::rp_expf(a.log_base * (E / 3600.0f));
    const float u = kInjected ? a.uniforms[static_cast<size_t>(r) * a.N + ii]
                              : rp::philox_uniform(a.seed, static_cast<uint32_t>(ii),
                                                   static_cast<uint32_t>(r), a.slot);
    infect = P > u;
  }
  if (infect) {
    a.entry[static_cast<size_t>(ii) * a.Mp + r] = 2 * a.k + 1;
    atomicAdd(a.inf_count + static_cast<size_t>(a.k) * a.M + r, 1);
  }
  return __ballot(infect);
}

// WT > 0: integer mode with node-level items (W <= WT); WT == 0: one item per
// word (float mode or wide nodes).
template <bool kIntMode, bool kInjected, int WT>
__global__ __launch_bounds__(kBlock) void step_kernel(StepArgs a) {
  __shared__ uint64_t s_sus[kBlock];
  __shared__ uint64_t s_x[kBlock];
  __shared__ int32_t s_list[kBlock];
  __shared__ uint8_t s_need[kBlock];
  __shared__ int2 s_rp[kBlock];
  __shared__ int32_t s_n;
  const int t = threadIdx.x;
  const int64_t w = static_cast<int64_t>(blockIdx.x) * kBlock + t;
  const bool wv = w < a.NW;
  const int i = wv ? static_cast<int>(w / a.W) : 0;
  const uint64_t s = wv ? a.S[w] : 0ull;
  // the word-local flips do not depend on this iteration's draws: issue
  // their loads now so they overlap the exposure phase
  uint64_t inf_pre = 0ull, sus_add = 0ull;
  if (wv) {
    inf_pre = a.I_cur[w];
    if (a.X_on) inf_pre ^= a.X_on[w];
    if (a.X_off) inf_pre ^= a.X_off[w];
    if (a.X_s) sus_add = a.X_s[w];
    if (a.init_flip_I || a.init_set_S) {
      const uint64_t xi = a.X_init[w];
      if (a.init_flip_I & 1) inf_pre ^= xi;
      if (a.init_flip_I & 2) inf_pre ^= xi;
      if (a.init_set_S) sus_add |= xi;
    }
  }
  const int32_t rp0 = wv ? a.rowptr[i] : 0, rp1 = wv ? a.rowptr[i + 1] : 0;
  s_sus[t] = s;
  s_x[t] = 0ull;
  s_rp[t] = make_int2(rp0, rp1);
  if (t == 0) s_n = 0;
  const bool need = s != 0ull && rp1 > rp0;
  const int lane = t & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(t / kWave);
  if constexpr (WT > 0) {
    s_need[t] = need;
    __syncthreads();
    // one item per node: its first word in this block collects the others
    const int q0 = wv ? static_cast<int>(w - static_cast<int64_t>(i) * a.W) : 0;
    if (wv && (t == 0 || q0 == 0)) {
      const int nq = min(a.W - q0, kBlock - t);
      bool any = false;
      for (int j = 0; j < nq; ++j) any |= s_need[t + j] != 0;
      if (any) s_list[atomicAdd(&s_n, 1)] = t;
    }
    __syncthreads();
    const int n_list = s_n;
    for (int it = wave; it < n_list; it += kWavesPerBlock) {
      const int tt = __builtin_amdgcn_readfirstlane(s_list[it]);
      const int64_t ww = static_cast<int64_t>(blockIdx.x) * kBlock + tt;
      const int ii = static_cast<int>(ww / a.W);
      const int qq0 = static_cast<int>(ww - static_cast<int64_t>(ii) * a.W);
      const int nq = min(a.W - qq0, kBlock - tt);
      uint32_t acc[WT];
      const int2 rr = s_rp[tt];
      node_exposure<WT>(__builtin_amdgcn_readfirstlane(rr.x), __builtin_amdgcn_readfirstlane(rr.y),
                        a.entries, a.I_cur, a.W, a.ws_rel, a.we_rel, a.rank_limit, qq0, nq, lane,
                        acc);
#pragma unroll
      for (int q = 0; q < WT; ++q) {
        if (q < nq) {
          const uint64_t ss = s_sus[tt + q];
          if (ss != 0ull) {
            const uint64_t x =
                draw_word<kInjected>(a, ii, qq0 + q, lane, ss, static_cast<float>(acc[q]));
            if (lane == 0) s_x[tt + q] = x;
          }
        }
      }
    }
    __syncthreads();
  } else {
  __syncthreads();
  if (need) s_list[atomicAdd(&s_n, 1)] = t;
  __syncthreads();
  const int n_list = s_n;
  for (int it = wave; it < n_list; it += kWavesPerBlock) {
    const int tt = __builtin_amdgcn_readfirstlane(s_list[it]);
    const int64_t ww = static_cast<int64_t>(blockIdx.x) * kBlock + tt;
    const int ii = static_cast<int>(ww / a.W);
    const int qq = static_cast<int>(ww - static_cast<int64_t>(ii) * a.W);
    const uint64_t ss = s_sus[tt];
    const float E = row_exposure<kIntMode>(a.rowptr, a.entries, a.I_cur, a.W, a.ws_rel, a.we_rel,
                                           a.rank_limit, a.scale, ii, qq, lane);
    const uint64_t x = draw_word<kInjected>(a, ii, qq, lane, ss, E);
    if (lane == 0) s_x[tt] = x;
  }
  __syncthreads();
  }
  if (wv) {
    const uint64_t x = s_x[t];
    a.X_k[w] = x;
    uint64_t inf = inf_pre;
    uint64_t sus = (s & ~x) | sus_add;
    if (a.on_is_local) inf ^= x;
    if (a.off_is_local) inf ^= x;
    if (a.s_is_local) sus |= x;
    a.I_next[w] = inf;
    a.S[w] = sus;
  }
  // running prefix of the infection counts: P[k] = P[k-1] + inf_count[k-1]
  const int64_t rr = w;
  if (a.k >= 1 && rr < a.M)
    a.P[static_cast<size_t>(a.k) * a.M + rr] =
        a.P[static_cast<size_t>(a.k - 1) * a.M + rr] +
        a.inf_count[static_cast<size_t>(a.k - 1) * a.M + rr];
}

struct InitArgs {
  uint64_t *S;
  uint64_t *I;
  uint64_t *X_init;
  int32_t *entry;
  int32_t *init_count;  // [M]
  const float *uniforms;  // slot 0 or null
  uint64_t seed;
  int32_t N, M, W, Mp;
  float p0;
  int32_t init_susceptible;  // the seed countdown is 0 (susceptible at once)
  int32_t init_infectious;   // the seed countdown is infectious at iteration 0
};

// cu:73-90 with a counter-based draw (slot 0): seeds, initial bit words,
// entries and per-replica seed counts.
__global__ __launch_bounds__(kBlock) void init_kernel(InitArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int q = blockIdx.y;
  const int r = q * kWave + lane;
  const bool rv = r < a.M;
  const int i = blockIdx.x * kWavesPerBlock + wave;
  if (i >= a.N) return;
  bool seed = false;
  if (rv) {
    const float u = a.uniforms ? a.uniforms[static_cast<size_t>(r) * a.N + i]
                               : rp::philox_uniform(a.seed, static_cast<uint32_t>(i),
                                                    static_cast<uint32_t>(r), 0u);
    seed = u < a.p0;
  }
  const uint64_t xi = __ballot(seed);
  const uint64_t valid = __ballot(rv);
  a.entry[static_cast<size_t>(i) * a.Mp + r] = seed ? 0 : -1;
  if (seed) atomicAdd(a.init_count + r, 1);
  if (lane == 0) {
    const size_t w = static_cast<size_t>(i) * a.W + q;
    a.X_init[w] = xi;
    a.S[w] = a.init_susceptible ? valid : (valid & ~xi);
    a.I[w] = a.init_infectious ? xi : 0ull;
  }
}

// P[k] for the last completed iteration.
__global__ void prefix_tail_kernel(int32_t *P, const int32_t *inf_count, int32_t M, int32_t k) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < M && k >= 1)
    P[static_cast<size_t>(k) * M + r] =
        P[static_cast<size_t>(k - 1) * M + r] + inf_count[static_cast<size_t>(k - 1) * M + r];
}

// Counts of blocks j0..j1: exposed, infectious, resistant of the state
// after iteration j-1 (block 0: the initial state).
__global__ void counts_kernel(int32_t *counts, const int32_t *P, const int32_t *init_count,
                              const int8_t *class_init, Segs segs, int32_t M, int32_t j0,
                              int32_t j1) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = j0 + blockIdx.y;
  if (r >= M || j > j1) return;
  int32_t c[4] = {0, 0, 0, 0};
  for (int g = 0; g < segs.n; ++g) {
    const int hi = j - segs.a[g] + 1;  // sum over e in [j-b+1, j-a]
    const int lo = j - segs.b[g] + 1;
    if (hi <= 0) continue;
    const int32_t ph = P[static_cast<size_t>(hi) * M + r];
    const int32_t pl = lo > 0 ? P[static_cast<size_t>(lo) * M + r] : 0;
    c[segs.cls[g]] += ph - pl;
  }
  const int ci = class_init[j];
  if (ci) c[ci] += init_count[r];
  int32_t *dst = counts + (static_cast<size_t>(j) * M + r) * 4;
  dst[0] = c[1];
  dst[1] = c[2];
  dst[2] = c[3];
  dst[3] = 0;
}

// Exposure of the current state for one window (debug / parity entry point).
template <bool kIntMode>
__global__ __launch_bounds__(kBlock) void exposure_kernel(StepArgs a, float *out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int q = blockIdx.y;
  const int r = q * kWave + lane;
  const int i = blockIdx.x * kWavesPerBlock + wave;
  if (i >= a.N) return;
  const float E = row_exposure<kIntMode>(a.rowptr, a.entries, a.I_cur, a.W, a.ws_rel, a.we_rel,
                                         a.rank_limit, a.scale, i, q, lane);
  if (r < a.M) out[static_cast<size_t>(r) * a.N + i] = E;
}

// Per-row CSR of one window: pass 0 counts merged columns, pass 1 fills.
__global__ void slice_kernel(StepArgs a, int pass, const int32_t *out_rowptr, int32_t *row_nnz,
                             int32_t *colind, float *val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const int e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
  int n = 0;
  int e = e0;
  while (e < e1) {
    const int col = a.entries[e].x;
    float w = 0.0f;
    bool any = false;
    for (; e < e1 && a.entries[e].x == col; ++e) {
      const int4 en = a.entries[e];
      const int ov = min(en.z, a.we_rel) - max(en.y, a.ws_rel);
      if (ov > 0 && static_cast<uint32_t>(en.w) <= a.rank_limit) {
        const float wc = static_cast<float>(ov) * a.scale;
        w = any ? w + wc : wc;
        any = true;
      }
    }
    if (any) {
      if (pass == 1) {
        colind[out_rowptr[i] + n] = col;
        val[out_rowptr[i] + n] = w;
      }
      ++n;
    }
  }
  if (pass == 0) row_nnz[i] = n;
}

thread_local std::string g_open_error;

// cu:530-550 on the host.
int host_class(int32_t v, int32_t upper, int32_t Ed, int32_t Id, int32_t Rd) {
  const int32_t elapsed = (v <= 0 || v > upper) ? 0 : (upper - v);
  if (v == 0) return 0;
  if (Ed > 0 && elapsed < Ed) return 1;
  if (Id > 0 && elapsed < Ed + Id) return 2;
  if (Rd > 0 && elapsed < Ed + Id + Rd) return 3;
  return 0;
}

}  // namespace

// Timer tables of one context (see the header comment).
struct TimerPlan {
  std::vector<int32_t> lut_new, lut_init;  // c(n), n = 0..T
  int32_t n_on = kNever, n_off = kNever, n_s = kNever;
  int32_t m_on = kNever, m_off = kNever, m_s = kNever;
  int32_t depth = 1;
  Segs segs{};
  std::vector<int8_t> class_init;  // class of c_init(j), j = 0..T
};

struct replay_ctx {
  replay_params p{};
  rp::Network net;
  rp::Plan plan;
  rp::Slices slices;
  TimerPlan tp;
  hipStream_t stream = nullptr;
  int4 *d_entries = nullptr;
  int32_t *d_rowptr = nullptr;
  uint64_t *d_S = nullptr;
  uint64_t *d_I[2] = {nullptr, nullptr};
  uint64_t *d_X = nullptr;       // ring, depth x N x W
  uint64_t *d_Xinit = nullptr;
  int32_t *d_entry = nullptr;
  int32_t *d_inf = nullptr;      // [T+1][M]
  int32_t *d_P = nullptr;        // [T+2][M]
  int32_t *d_init_count = nullptr;
  int32_t *d_counts = nullptr;   // [T+1][M][4]
  int8_t *d_class_init = nullptr;
  float *d_uniforms = nullptr;
  int32_t N = 0, M = 0, W = 0, Mp = 0;
  int cur = 0;
  int32_t steps_done = 0;
  int32_t counts_done = -1;      // last counts block computed
  float log_base = 0.0f;
  int32_t upper = 0, medium = 0, lower = 0;
  int truncated = 0;
  std::string err;
};

namespace {

int fail(replay_ctx *ctx, int code, const std::string &msg) {
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

void release(replay_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  void *bufs[] = {ctx->d_entries, ctx->d_rowptr, ctx->d_S, ctx->d_I[0], ctx->d_I[1],
                  ctx->d_X, ctx->d_Xinit, ctx->d_entry, ctx->d_inf, ctx->d_P,
                  ctx->d_init_count, ctx->d_counts, ctx->d_class_init, ctx->d_uniforms};
  for (void *b : bufs) hipFree(b);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  ctx->stream = nullptr;
}

// Tabulates the countdown sequences and the offsets derived from them.
int plan_timers(replay_ctx *ctx) {
  TimerPlan &tp = ctx->tp;
  const int32_t T = ctx->p.iterations;
  const float step = ctx->p.step_size;
  auto f = [step](int32_t c) {  // cu:169-174
    if (c > 0) c = static_cast<int32_t>(static_cast<float>(c) - step);
    if (c < 0) c = 0;
    return c;
  };
  const int32_t c_new = static_cast<int32_t>(static_cast<float>(ctx->upper) + step);
  const int32_t c_init = static_cast<int32_t>(static_cast<float>(ctx->medium) + step);
  tp.lut_new.assign(static_cast<size_t>(T) + 1, 0);
  tp.lut_init.assign(static_cast<size_t>(T) + 1, 0);
  tp.lut_new[0] = c_new;
  tp.lut_init[0] = c_init;
  for (int32_t n = 1; n <= T; ++n) {
    tp.lut_new[n] = f(tp.lut_new[n - 1]);
    tp.lut_init[n] = f(tp.lut_init[n - 1]);
  }
  auto infectious = [ctx](int32_t c) { return c > ctx->lower && c <= ctx->medium; };
  // contiguous infectious range [on, off) and first zero, over n >= n0
  auto ranges = [&](const std::vector<int32_t> &lut, int32_t n0, int32_t *on, int32_t *off,
                    int32_t *zero) {
    *on = *off = *zero = kNever;
    for (int32_t n = n0; n <= T; ++n) {
      const bool inf = infectious(lut[n]);
      if (inf && *on == kNever) *on = n;
      if (!inf && *on != kNever && *off == kNever) *off = n;
      if (inf && *off != kNever) return false;  // not contiguous
      if (lut[n] == 0 && *zero == kNever) *zero = n;
    }
    return true;
  };
  if (!ranges(tp.lut_new, 1, &tp.n_on, &tp.n_off, &tp.n_s) ||
      !ranges(tp.lut_init, 0, &tp.m_on, &tp.m_off, &tp.m_s))
    return fail(ctx, REPLAY_E_UNSUPPORTED, "infectious period is not contiguous");
  int32_t dmax = 1;
  for (int32_t n : {tp.n_on, tp.n_off, tp.n_s})
    if (n != kNever) dmax = std::max(dmax, n);
  tp.depth = dmax;  // slots k+1-n for 2 <= n <= dmax and slot k: dmax slots
  // class runs of new infections over n = 1..T, and the seeds' class table
  tp.segs.n = 0;
  int32_t prev = -1;
  for (int32_t n = 1; n <= T; ++n) {
    const int cls = host_class(tp.lut_new[n], ctx->upper, ctx->p.exposed_duration,
                               ctx->p.infectious_duration, ctx->p.resistant_duration);
    if (cls != prev) {
      if (prev > 0) tp.segs.b[tp.segs.n - 1] = n;
      if (cls > 0) {
        if (tp.segs.n >= kMaxSegs) return fail(ctx, REPLAY_E_UNSUPPORTED, "class schedule too long");
        tp.segs.a[tp.segs.n] = n;
        tp.segs.cls[tp.segs.n] = cls;
        tp.segs.b[tp.segs.n] = T + 1;
        tp.segs.n++;
      }
      prev = cls;
    }
  }
  tp.class_init.assign(static_cast<size_t>(T) + 1, 0);
  for (int32_t j = 0; j <= T; ++j)
    tp.class_init[j] = static_cast<int8_t>(host_class(tp.lut_init[j], ctx->upper,
                                                      ctx->p.exposed_duration,
                                                      ctx->p.infectious_duration,
                                                      ctx->p.resistant_duration));
  return REPLAY_OK;
}

StepArgs step_args(replay_ctx *ctx, int32_t k) {
  const rp::Window &w = ctx->plan.windows[ctx->plan.iter_window[k]];
  const int e = w.epoch;
  const int64_t t0 = ctx->slices.epoch_t0[e];
  const TimerPlan &tp = ctx->tp;
  const size_t NW = static_cast<size_t>(ctx->N) * ctx->W;
  auto ring = [&](int32_t n) -> const uint64_t * {  // X(k+1-n) for n >= 2
    if (n == kNever || n < 2 || k + 1 - n < 0) return nullptr;
    return ctx->d_X + static_cast<size_t>((k + 1 - n) % tp.depth) * NW;
  };
  StepArgs a;
  std::memset(&a, 0, sizeof a);
  a.rowptr = ctx->d_rowptr + static_cast<size_t>(e) * (ctx->N + 1);
  a.entries = ctx->d_entries + ctx->slices.epoch_entry0[e];
  a.I_cur = ctx->d_I[ctx->cur];
  a.I_next = ctx->d_I[ctx->cur ^ 1];
  a.S = ctx->d_S;
  a.X_k = ctx->d_X + static_cast<size_t>(k % tp.depth) * NW;
  a.X_on = ring(tp.n_on);
  a.X_off = ring(tp.n_off);
  a.X_s = ring(tp.n_s);
  a.X_init = ctx->d_Xinit;
  a.entry = ctx->d_entry;
  a.inf_count = ctx->d_inf;
  a.P = ctx->d_P;
  a.uniforms = ctx->d_uniforms
                   ? ctx->d_uniforms + static_cast<size_t>(k + 1) * ctx->M * ctx->N
                   : nullptr;
  a.seed = ctx->p.seed;
  a.NW = static_cast<int64_t>(NW);
  a.N = ctx->N;
  a.M = ctx->M;
  a.W = ctx->W;
  a.Mp = ctx->Mp;
  a.k = k;
  a.ws_rel = static_cast<int32_t>(w.ws - t0);
  a.we_rel = static_cast<int32_t>(w.we - t0);
  a.rank_limit = w.rank_limit;
  a.scale = w.scale;
  a.log_base = ctx->log_base;
  a.slot = static_cast<uint32_t>(k + 1);
  a.on_is_local = tp.n_on == 1;
  a.off_is_local = tp.n_off == 1;
  a.s_is_local = tp.n_s == 1;
  a.init_flip_I = (tp.m_on != kNever && tp.m_on >= 1 && k + 1 == tp.m_on ? 1 : 0) |
                  (tp.m_off != kNever && tp.m_off >= 1 && k + 1 == tp.m_off ? 2 : 0);
  a.init_set_S = (tp.m_s != kNever && tp.m_s >= 1 && k + 1 == tp.m_s) ? 1 : 0;
  return a;
}

template <bool kInt, bool kInj>
void launch_step_w(const StepArgs &a, dim3 grid, hipStream_t st) {
  if (kInt && a.W <= 4)
    hipLaunchKernelGGL((step_kernel<kInt, kInj, 4>), grid, dim3(kBlock), 0, st, a);
  else if (kInt && a.W <= 16)
    hipLaunchKernelGGL((step_kernel<kInt, kInj, 16>), grid, dim3(kBlock), 0, st, a);
  else
    hipLaunchKernelGGL((step_kernel<kInt, kInj, 0>), grid, dim3(kBlock), 0, st, a);
}

void launch_step(const StepArgs &a, bool int_mode, bool injected, dim3 grid, hipStream_t st) {
  if (int_mode && !injected)
    launch_step_w<true, false>(a, grid, st);
  else if (int_mode)
    launch_step_w<true, true>(a, grid, st);
  else if (!injected)
    launch_step_w<false, false>(a, grid, st);
  else
    launch_step_w<false, true>(a, grid, st);
}

int do_reset(replay_ctx *ctx) {
  const int32_t T = ctx->p.iterations;
  const size_t M = static_cast<size_t>(ctx->M);
  HIP_OR_FAIL(ctx, hipMemsetAsync(ctx->d_inf, 0, (T + 1) * M * sizeof(int32_t), ctx->stream));
  HIP_OR_FAIL(ctx, hipMemsetAsync(ctx->d_P, 0, (T + 2) * M * sizeof(int32_t), ctx->stream));
  HIP_OR_FAIL(ctx, hipMemsetAsync(ctx->d_init_count, 0, M * sizeof(int32_t), ctx->stream));
  InitArgs a;
  std::memset(&a, 0, sizeof a);
  a.S = ctx->d_S;
  a.I = ctx->d_I[0];
  a.X_init = ctx->d_Xinit;
  a.entry = ctx->d_entry;
  a.init_count = ctx->d_init_count;
  a.uniforms = ctx->d_uniforms;
  a.seed = ctx->p.seed;
  a.N = ctx->N;
  a.M = ctx->M;
  a.W = ctx->W;
  a.Mp = ctx->Mp;
  a.p0 = ctx->p.initial_infected;
  a.init_susceptible = ctx->tp.m_s == 0;
  a.init_infectious = ctx->tp.m_on == 0;
  dim3 grid((ctx->N + kWavesPerBlock - 1) / kWavesPerBlock, ctx->W);
  hipLaunchKernelGGL(init_kernel, grid, dim3(kBlock), 0, ctx->stream, a);
  HIP_OR_FAIL(ctx, hipGetLastError());
  ctx->cur = 0;
  ctx->steps_done = 0;
  ctx->counts_done = -1;
  return REPLAY_OK;
}

// Counts blocks up to steps_done (enqueued on the stream).
int enqueue_counts(replay_ctx *ctx) {
  const int32_t j0 = ctx->counts_done + 1, j1 = ctx->steps_done;
  if (j0 > j1) return REPLAY_OK;
  const int tb = 256, nb = (ctx->M + tb - 1) / tb;
  hipLaunchKernelGGL(prefix_tail_kernel, dim3(nb), dim3(tb), 0, ctx->stream, ctx->d_P, ctx->d_inf,
                     ctx->M, j1);
  hipLaunchKernelGGL(counts_kernel, dim3(nb, j1 - j0 + 1), dim3(tb), 0, ctx->stream,
                     ctx->d_counts, ctx->d_P, ctx->d_init_count, ctx->d_class_init, ctx->tp.segs,
                     ctx->M, j0, j1);
  HIP_OR_FAIL(ctx, hipGetLastError());
  ctx->counts_done = j1;
  return REPLAY_OK;
}

int finish_open(replay_ctx *ctx, const replay_params *p) {
  std::string err;
  int rc = rp::finalize(ctx->net, err);
  if (rc != REPLAY_OK) return fail(ctx, rc, err);
  rc = rp::plan_windows(ctx->net, *p, ctx->plan, err);
  if (rc != REPLAY_OK) return fail(ctx, rc, err);
  int64_t budget = p->max_slice_bytes > 0 ? p->max_slice_bytes : (int64_t(32) << 30);
  rc = rp::build_slices(ctx->net, ctx->plan, p->windows_per_epoch, budget, ctx->slices, err);
  if (rc != REPLAY_OK) return fail(ctx, rc, err);
  for (const auto &w : ctx->plan.windows) ctx->truncated += (w.rank_limit != UINT32_MAX);
  ctx->N = ctx->net.n_nodes;
  ctx->M = p->M;
  ctx->W = (p->M + kWave - 1) / kWave;
  ctx->Mp = ctx->W * kWave;
  ctx->upper = p->exposed_duration + p->infectious_duration + p->resistant_duration;
  ctx->medium = p->infectious_duration + p->resistant_duration;
  ctx->lower = p->resistant_duration;
  ctx->log_base = logf(1.0f - p->infect_prob);
  rc = plan_timers(ctx);
  if (rc != REPLAY_OK) return rc;
  HIP_OR_FAIL(ctx, hipSetDevice(p->device));
  HIP_OR_FAIL(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  const size_t nE = ctx->slices.entries.size();
  const size_t nR = ctx->slices.rowptr.size();
  const size_t NW = static_cast<size_t>(ctx->N) * ctx->W;
  const size_t T = static_cast<size_t>(p->iterations);
  const size_t M = static_cast<size_t>(ctx->M);
  if ((rc = dev_alloc(ctx, &ctx->d_entries, nE)) || (rc = dev_alloc(ctx, &ctx->d_rowptr, nR)) ||
      (rc = dev_alloc(ctx, &ctx->d_S, NW)) || (rc = dev_alloc(ctx, &ctx->d_I[0], NW)) ||
      (rc = dev_alloc(ctx, &ctx->d_I[1], NW)) ||
      (rc = dev_alloc(ctx, &ctx->d_X, NW * static_cast<size_t>(ctx->tp.depth))) ||
      (rc = dev_alloc(ctx, &ctx->d_Xinit, NW)) ||
      (rc = dev_alloc(ctx, &ctx->d_entry, static_cast<size_t>(ctx->N) * ctx->Mp)) ||
      (rc = dev_alloc(ctx, &ctx->d_inf, (T + 1) * M)) ||
      (rc = dev_alloc(ctx, &ctx->d_P, (T + 2) * M)) ||
      (rc = dev_alloc(ctx, &ctx->d_init_count, M)) ||
      (rc = dev_alloc(ctx, &ctx->d_counts, (T + 1) * M * 4)) ||
      (rc = dev_alloc(ctx, &ctx->d_class_init, T + 1)))
    return rc;
  if (nE)
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->d_entries, ctx->slices.entries.data(), nE * sizeof(int4),
                                    hipMemcpyHostToDevice, ctx->stream));
  if (nR)
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->d_rowptr, ctx->slices.rowptr.data(),
                                    nR * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->d_class_init, ctx->tp.class_init.data(), T + 1,
                                  hipMemcpyHostToDevice, ctx->stream));
  rc = do_reset(ctx);
  if (rc != REPLAY_OK) return rc;
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  return REPLAY_OK;
}

int check_params(const replay_params *p) {
  if (!p) return fail(nullptr, REPLAY_E_ARG, "null params");
  if (p->M <= 0) return fail(nullptr, REPLAY_E_ARG, "M must be positive");
  if (p->iterations < 0) return fail(nullptr, REPLAY_E_ARG, "iterations must be >= 0");
  if (p->static_network_duration <= 0)
    return fail(nullptr, REPLAY_E_ARG, "static network duration must be positive");
  if (!(p->step_size > 0.0f)) return fail(nullptr, REPLAY_E_ARG, "step size must be positive");
  if (p->exposed_duration < 0 || p->infectious_duration < 0 || p->resistant_duration < 0)
    return fail(nullptr, REPLAY_E_ARG, "durations must be >= 0");
  const int64_t up = static_cast<int64_t>(p->exposed_duration) + p->infectious_duration +
                     p->resistant_duration;
  if (up > INT32_MAX) return fail(nullptr, REPLAY_E_ARG, "durations overflow int32");
  return REPLAY_OK;
}

int open_common(replay_ctx *ctx, const replay_params *p, replay_ctx **out) {
  int rc = finish_open(ctx, p);
  if (rc != REPLAY_OK) {
    g_open_error = ctx->err;
    release(ctx);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return REPLAY_OK;
}

}  // namespace

extern "C" {

int32_t replay_abi_version(void) { return REPLAY_HIP_ABI_VERSION; }

void replay_params_init(replay_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->M = 1000;
  p->step_size = 3600.0f;
  p->iterations = 42;
  p->start_time = 946713600;
  p->static_network_duration = 3600;
}

int32_t replay_open(const char *graph_path, const replay_params *p, replay_ctx **out) {
  if (!out || !graph_path) return fail(nullptr, REPLAY_E_ARG, "null argument");
  *out = nullptr;
  int rc = check_params(p);
  if (rc != REPLAY_OK) return rc;
  replay_ctx *ctx = new (std::nothrow) replay_ctx();
  if (!ctx) return fail(nullptr, REPLAY_E_NOMEM, "out of host memory");
  ctx->p = *p;
  std::string err;
  rc = rp::parse_file(graph_path, ctx->net, err);
  if (rc != REPLAY_OK) {
    delete ctx;
    return fail(nullptr, rc, err);
  }
  return open_common(ctx, p, out);
}

int32_t replay_open_contacts(int32_t n_node_list, const int64_t *node_ids, int64_t n_contacts,
                             const int32_t *src, const int32_t *tgt, const int64_t *start,
                             const int64_t *end, const replay_params *p, replay_ctx **out) {
  if (!out) return fail(nullptr, REPLAY_E_ARG, "null argument");
  *out = nullptr;
  int rc = check_params(p);
  if (rc != REPLAY_OK) return rc;
  replay_ctx *ctx = new (std::nothrow) replay_ctx();
  if (!ctx) return fail(nullptr, REPLAY_E_NOMEM, "out of host memory");
  ctx->p = *p;
  std::string err;
  rc = rp::from_contacts(n_node_list, node_ids, n_contacts, src, tgt, start, end, ctx->net, err);
  if (rc != REPLAY_OK) {
    delete ctx;
    return fail(nullptr, rc, err);
  }
  return open_common(ctx, p, out);
}

int32_t replay_get_info(replay_ctx *ctx, replay_info *out) {
  if (!ctx || !out) return REPLAY_E_ARG;
  std::memset(out, 0, sizeof *out);
  out->n_nodes = ctx->N;
  out->M = ctx->M;
  out->n_contacts = static_cast<int64_t>(ctx->net.contacts.size());
  out->nlists = ctx->net.nlists;
  out->iterations = ctx->p.iterations;
  out->n_windows = static_cast<int32_t>(ctx->plan.windows.size());
  out->n_epochs = static_cast<int32_t>(ctx->slices.epoch_t0.size());
  out->windows_per_epoch = ctx->slices.windows_per_epoch;
  out->n_entries = static_cast<int64_t>(ctx->slices.entries.size());
  out->slice_bytes = static_cast<int64_t>(ctx->slices.entries.size() * sizeof(rp::Entry) +
                                          ctx->slices.rowptr.size() * sizeof(int32_t));
  const int64_t NW = static_cast<int64_t>(ctx->N) * ctx->W;
  out->state_bytes = NW * 8 * (4 + ctx->tp.depth) + static_cast<int64_t>(ctx->N) * ctx->Mp * 4;
  out->steps_done = ctx->steps_done;
  out->truncated_windows = ctx->truncated;
  // per word: S read + write, own I read, I_next write, X_k write, and one
  // ring read per flip offset kept in the ring; plus the slice epoch's
  // entries and row pointers, and the per-replica prefix update
  int ring_reads = 0;
  for (int32_t n : {ctx->tp.n_on, ctx->tp.n_off, ctx->tp.n_s})
    ring_reads += (n != kNever && n >= 2 && n <= ctx->p.iterations) ? 1 : 0;
  const double ent = ctx->slices.epoch_t0.empty()
                         ? 0.0
                         : static_cast<double>(ctx->slices.entries.size()) /
                               static_cast<double>(ctx->slices.epoch_t0.size());
  out->step_bytes = NW * 8 * (5 + ring_reads) + static_cast<int64_t>(16.0 * ent) +
                    4 * (static_cast<int64_t>(ctx->N) + 1) + 12 * static_cast<int64_t>(ctx->M);
  out->ring_depth = ctx->tp.depth;
  return REPLAY_OK;
}

int32_t replay_reset(replay_ctx *ctx) {
  if (!ctx) return REPLAY_E_ARG;
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  return do_reset(ctx);
}

int32_t replay_run(replay_ctx *ctx, int32_t steps) {
  if (!ctx) return REPLAY_E_ARG;
  if (steps < 0 || ctx->steps_done + steps > ctx->p.iterations)
    return fail(ctx, REPLAY_E_ARG, "steps exceed the planned iterations");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  // threads cover every word and, for the running prefix, every replica
  const int64_t NW = static_cast<int64_t>(ctx->N) * ctx->W;
  const int64_t cover = std::max<int64_t>(NW, ctx->M);
  const dim3 grid(static_cast<unsigned>((cover + kBlock - 1) / kBlock));
  for (int32_t s = 0; s < steps; ++s) {
    const int32_t k = ctx->steps_done;
    StepArgs a = step_args(ctx, k);
    const bool int_mode = ctx->plan.windows[ctx->plan.iter_window[k]].int_mode != 0;
    launch_step(a, int_mode, a.uniforms != nullptr, grid, ctx->stream);
    HIP_OR_FAIL(ctx, hipGetLastError());
    ctx->cur ^= 1;
    ctx->steps_done = k + 1;
  }
  return enqueue_counts(ctx);
}

int32_t replay_run_timed(replay_ctx *ctx, int32_t steps, double *elapsed_ms) {
  if (!ctx || !elapsed_ms) return REPLAY_E_ARG;
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  hipEvent_t ev[2];
  HIP_OR_FAIL(ctx, hipEventCreate(&ev[0]));
  HIP_OR_FAIL(ctx, hipEventCreate(&ev[1]));
  HIP_OR_FAIL(ctx, hipEventRecord(ev[0], ctx->stream));
  int rc = replay_run(ctx, steps);
  if (rc == REPLAY_OK) {
    HIP_OR_FAIL(ctx, hipEventRecord(ev[1], ctx->stream));
    HIP_OR_FAIL(ctx, hipEventSynchronize(ev[1]));
    float ms = 0.0f;
    HIP_OR_FAIL(ctx, hipEventElapsedTime(&ms, ev[0], ev[1]));
    *elapsed_ms = ms;
  }
  hipEventDestroy(ev[0]);
  hipEventDestroy(ev[1]);
  return rc;
}

int32_t replay_sync(replay_ctx *ctx) {
  if (!ctx) return REPLAY_E_ARG;
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  return REPLAY_OK;
}

int32_t replay_get_counts(replay_ctx *ctx, int32_t *out, int64_t n_out) {
  if (!ctx || !out) return REPLAY_E_ARG;
  const int64_t n = static_cast<int64_t>(ctx->steps_done + 1) * ctx->M * 4;
  if (n_out < n) return fail(ctx, REPLAY_E_ARG, "counts buffer too small");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  int rc = enqueue_counts(ctx);
  if (rc) return rc;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(out, ctx->d_counts, static_cast<size_t>(n) * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  for (int64_t b = 0; b < n; b += 4) out[b + 3] = ctx->N - out[b] - out[b + 1] - out[b + 2];
  return REPLAY_OK;
}

int32_t replay_get_state(replay_ctx *ctx, int32_t *out, int64_t n_out) {
  if (!ctx || !out) return REPLAY_E_ARG;
  const int64_t n = static_cast<int64_t>(ctx->N) * ctx->M;
  if (n_out < n) return fail(ctx, REPLAY_E_ARG, "state buffer too small");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  std::vector<int32_t> ent(static_cast<size_t>(ctx->N) * ctx->Mp);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(ent.data(), ctx->d_entry, ent.size() * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  const int32_t k = ctx->steps_done;
  const auto &ln = ctx->tp.lut_new, &li = ctx->tp.lut_init;
  for (int32_t i = 0; i < ctx->N; ++i)
    for (int32_t r = 0; r < ctx->M; ++r) {
      const int32_t s = ent[static_cast<size_t>(i) * ctx->Mp + r];
      int32_t c = 0;
      if (s == 0)
        c = li[k];
      else if (s > 0)
        c = ln[k - (s - 1) / 2];
      out[static_cast<size_t>(r) * ctx->N + i] = c;
    }
  return REPLAY_OK;
}

int64_t replay_iteration_time(replay_ctx *ctx, int32_t k) {
  if (!ctx) return 0;
  if (k < 0 || k >= static_cast<int32_t>(ctx->plan.iter_time.size())) return ctx->p.start_time;
  return ctx->plan.iter_time[k];
}

int32_t replay_set_uniforms(replay_ctx *ctx, const float *u, int64_t n) {
  if (!ctx) return REPLAY_E_ARG;
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->d_uniforms) {
    hipFree(ctx->d_uniforms);
    ctx->d_uniforms = nullptr;
  }
  if (u) {
    const int64_t need = static_cast<int64_t>(ctx->p.iterations + 1) * ctx->M * ctx->N;
    if (n != need) return fail(ctx, REPLAY_E_ARG, "uniform stream has the wrong length");
    int rc = dev_alloc(ctx, &ctx->d_uniforms, static_cast<size_t>(n));
    if (rc) return rc;
    HIP_OR_FAIL(ctx, hipMemcpy(ctx->d_uniforms, u, static_cast<size_t>(n) * sizeof(float),
                               hipMemcpyHostToDevice));
  }
  return do_reset(ctx);
}

int32_t replay_get_slice(replay_ctx *ctx, int32_t k, int32_t *rowptr, int32_t *colind, float *val,
                         int64_t cap, int64_t *nnz) {
  if (!ctx || !rowptr || !nnz || k < 0 || k >= ctx->p.iterations) return REPLAY_E_ARG;
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  StepArgs a = step_args(ctx, k);
  int32_t *d_rp = nullptr, *d_cnt = nullptr, *d_col = nullptr;
  float *d_val = nullptr;
  int rc;
  if ((rc = dev_alloc(ctx, &d_cnt, static_cast<size_t>(ctx->N))) ||
      (rc = dev_alloc(ctx, &d_rp, static_cast<size_t>(ctx->N) + 1)))
    return rc;
  const int tb = 256, nb = (ctx->N + tb - 1) / tb;
  hipLaunchKernelGGL(slice_kernel, dim3(nb), dim3(tb), 0, ctx->stream, a, 0, d_rp, d_cnt,
                     nullptr, nullptr);
  std::vector<int32_t> cnt(static_cast<size_t>(ctx->N));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(cnt.data(), d_cnt, cnt.size() * 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  rowptr[0] = 0;
  for (int32_t i = 0; i < ctx->N; ++i) rowptr[i + 1] = rowptr[i] + cnt[i];
  *nnz = rowptr[ctx->N];
  if (*nnz > cap || (*nnz > 0 && (!colind || !val))) {
    hipFree(d_cnt);
    hipFree(d_rp);
    return fail(ctx, REPLAY_E_ARG, "slice buffer too small");
  }
  if ((rc = dev_alloc(ctx, &d_col, static_cast<size_t>(*nnz))) ||
      (rc = dev_alloc(ctx, &d_val, static_cast<size_t>(*nnz))))
    return rc;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(d_rp, rowptr, (ctx->N + 1) * sizeof(int32_t),
                                  hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(slice_kernel, dim3(nb), dim3(tb), 0, ctx->stream, a, 1, d_rp, d_cnt, d_col,
                     d_val);
  if (*nnz) {
    HIP_OR_FAIL(ctx, hipMemcpyAsync(colind, d_col, *nnz * sizeof(int32_t), hipMemcpyDeviceToHost,
                                    ctx->stream));
    HIP_OR_FAIL(ctx, hipMemcpyAsync(val, d_val, *nnz * sizeof(float), hipMemcpyDeviceToHost,
                                    ctx->stream));
  }
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  hipFree(d_cnt);
  hipFree(d_rp);
  hipFree(d_col);
  hipFree(d_val);
  return REPLAY_OK;
}

int32_t replay_get_exposure(replay_ctx *ctx, int32_t k, float *out, int64_t n_out) {
  if (!ctx || !out || k < 0 || k >= ctx->p.iterations) return REPLAY_E_ARG;
  const int64_t n = static_cast<int64_t>(ctx->N) * ctx->M;
  if (n_out < n) return fail(ctx, REPLAY_E_ARG, "exposure buffer too small");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  StepArgs a = step_args(ctx, k);
  float *d_out = nullptr;
  int rc = dev_alloc(ctx, &d_out, static_cast<size_t>(n));
  if (rc) return rc;
  dim3 grid((ctx->N + kWavesPerBlock - 1) / kWavesPerBlock, ctx->W);
  if (ctx->plan.windows[ctx->plan.iter_window[k]].int_mode)
    hipLaunchKernelGGL(exposure_kernel<true>, grid, dim3(kBlock), 0, ctx->stream, a, d_out);
  else
    hipLaunchKernelGGL(exposure_kernel<false>, grid, dim3(kBlock), 0, ctx->stream, a, d_out);
  HIP_OR_FAIL(ctx, hipGetLastError());
  HIP_OR_FAIL(ctx, hipMemcpyAsync(out, d_out, static_cast<size_t>(n) * sizeof(float),
                                  hipMemcpyDeviceToHost, ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  hipFree(d_out);
  return REPLAY_OK;
}

const char *replay_last_error(replay_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_open_error.c_str();
}

const char *replay_error_string(int32_t code) {
  switch (code) {
    case REPLAY_OK: return "ok";
    case REPLAY_E_ARG: return "invalid argument";
    case REPLAY_E_IO: return "i/o error";
    case REPLAY_E_FORMAT: return "malformed NODE_LIST header";
    case REPLAY_E_NODE_ID: return "invalid node id";
    case REPLAY_E_DUP_ID: return "duplicate node id";
    case REPLAY_E_NO_EDGES: return "no valid contacts";
    case REPLAY_E_COUNT: return "more contacts parsed than counted";
    case REPLAY_E_NODE_RANGE: return "more distinct endpoints than NODE_LIST ids";
    case REPLAY_E_NOMEM: return "out of memory";
    case REPLAY_E_DEVICE: return "HIP runtime error";
    case REPLAY_E_OVERFLOW: return "time arithmetic overflow";
    case REPLAY_E_UNSUPPORTED: return "unsupported input";
    case REPLAY_E_STATE: return "call out of sequence";
    default: return "unknown error";
  }
}

void replay_close(replay_ctx *ctx) {
  if (!ctx) return;
  release(ctx);
  delete ctx;
}

}  // extern "C"
