// replay_hip.hip -- libreplay_hip: HIP kernels for gfx950 and the C ABI
// declared in include/replay_hip.h.
//
// Device data (one context = one network, M replicas):
//   state     int32 [N][Mp]   countdown seconds, replica-fastest, Mp = 64*W
//   mask[2]   u64   [N][W]    infectious bit per (node, replica), W = M/64
//                             words; double-buffered across iterations
//   slices    epoch-packed contact entries {col, ts, te, rank} + rowptr
//   counts    int32 [T+1][M][4]  exposed, infectious, resistant (device);
//                             susceptible is filled in on the host
//
// One fused kernel per iteration replaces the reference's five launches
// and cuSPARSE call (tool/gpu_cpu_temporal_sim.cu:1053-1096): a wavefront
// owns 64 replicas of one node; it reads the node's countdowns, and when any
// of them is susceptible it walks the node's packed entries in ascending
// column order, evaluating each contact's overlap with the window and adding
// it where the neighbour's infectious bit is set.  Then it draws, updates the
// countdown, writes the next mask word with a ballot and accumulates counts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "device_math.hpp"
#include "ingest.hpp"
#include "replay_hip.h"

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

struct StepArgs {
  const int32_t *rowptr;   // epoch rowptr (N+1), epoch-relative entry offsets
  const int4 *entries;     // epoch entries
  const uint64_t *mask_cur;
  uint64_t *mask_next;
  int32_t *state;
  int32_t *counts;         // this iteration's block [M][4]
  const float *uniforms;   // injected stream for this slot, [M][N], or null
  uint64_t seed;
  int32_t N, M, W, Mp;
  int32_t nodes_per_block;
  int32_t ws_rel, we_rel;
  uint32_t rank_limit;
  float scale;
  float step;
  float log_base;          // logf(1 - infect_prob), computed on the host
  int32_t upper, medium, lower;
  int32_t Ed, Id, Rd;
  uint32_t slot;
};

// cu:530-550 classification: 0 S, 1 E, 2 I, 3 R.
__device__ __forceinline__ int state_class(int32_t v, int32_t upper, int32_t Ed, int32_t Id,
                                           int32_t Rd) {
  const int32_t elapsed = (v <= 0 || v > upper) ? 0 : (upper - v);
  if (v == 0) return 0;
  if (Ed > 0 && elapsed < Ed) return 1;
  if (Id > 0 && elapsed < Ed + Id) return 2;
  if (Rd > 0 && elapsed < Ed + Id + Rd) return 3;
  return 0;
}

// Exposure seconds of (row i, replica lane) for one window.
// Integer mode (scale 1, row weight sums < 2^24): overlaps are integers and
// every partial sum is exact, so the integer sum equals the reference's fp32
// sum in any order.  Float mode: contributions fl(ov * scale) of one column
// are merged in rank order (cu:299-309), and columns are added in ascending
// order from zero (Eigen SparseDenseProduct.h:64-70).
template <bool kIntMode>
__device__ __forceinline__ float row_exposure(const StepArgs &a, int i, int q, int lane) {
  const int e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
  if (kIntMode) {
    uint32_t acc = 0;
    for (int e = e0; e < e1; ++e) {
      const int4 en = a.entries[e];
      const int ov = min(en.z, a.we_rel) - max(en.y, a.ws_rel);
      if (ov > 0 && static_cast<uint32_t>(en.w) <= a.rank_limit) {
        const uint64_t m = a.mask_cur[static_cast<size_t>(en.x) * a.W + q];
        if (m != 0ull && ((m >> lane) & 1ull)) acc += static_cast<uint32_t>(ov);
      }
    }
    return static_cast<float>(acc);
  } else {
    float acc = 0.0f;
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
        const uint64_t m = a.mask_cur[static_cast<size_t>(col) * a.W + q];
        if ((m >> lane) & 1ull) acc += w;
      }
    }
    return acc;
  }
}

template <bool kIntMode, bool kInjected>
__global__ __launch_bounds__(kBlock) void step_kernel(StepArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int q = blockIdx.y;
  const int r = q * kWave + lane;
  const bool rv = r < a.M;
  int cE = 0, cI = 0, cR = 0;
  const int i_begin = blockIdx.x * a.nodes_per_block;
  const int i_end = min(a.N, i_begin + a.nodes_per_block);
  for (int i = i_begin + wave; i < i_end; i += kWavesPerBlock) {
    int32_t *sp = a.state + static_cast<size_t>(i) * a.Mp + r;
    int32_t c = *sp;
    const bool sus = rv && c == 0;
    if (__ballot(sus) != 0ull) {
      const bool has_row = a.rowptr[i + 1] > a.rowptr[i];
      if (kInjected) {
        // evaluate every susceptible lane, as the reference does
        const float E = has_row ? row_exposure<kIntMode>(a, i, q, lane) : 0.0f;
        if (sus) {
          const float P = 1.0f - rp::rp_expf(a.log_base * (E / 3600.0f));
          const float u = a.uniforms[static_cast<size_t>(r) * a.N + i];
          if (P > u) c = static_cast<int32_t>(static_cast<float>(a.upper) + a.step);
        }
      } else if (has_row) {
        const float E = row_exposure<kIntMode>(a, i, q, lane);
        // E == 0 gives P = 0 (or NaN), never above a uniform in [0, 1)
        if (sus && E > 0.0f) {
          const float P = 1.0f - rp::rp_expf(a.log_base * (E / 3600.0f));
          const float u = rp::philox_uniform(a.seed, static_cast<uint32_t>(i),
                                             static_cast<uint32_t>(r), a.slot);
          if (P > u) c = static_cast<int32_t>(static_cast<float>(a.upper) + a.step);
        }
      }
    }
    // cu:161-176
    if (c > 0) c = static_cast<int32_t>(static_cast<float>(c) - a.step);
    if (c < 0) c = 0;
    if (rv) *sp = c;
    const uint64_t bits = __ballot(rv && c > a.lower && c <= a.medium);
    if (lane == 0) a.mask_next[static_cast<size_t>(i) * a.W + q] = bits;
    const int cls = state_class(c, a.upper, a.Ed, a.Id, a.Rd);
    cE += cls == 1;
    cI += cls == 2;
    cR += cls == 3;
  }
  __shared__ int red[3][kWavesPerBlock][kWave];
  red[0][wave][lane] = cE;
  red[1][wave][lane] = cI;
  red[2][wave][lane] = cR;
  __syncthreads();
  if (wave == 0 && rv) {
    int s[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int w = 0; w < kWavesPerBlock; ++w) s[k] += red[k][w][lane];
    int32_t *dst = a.counts + static_cast<size_t>(r) * 4;
    if (s[0]) atomicAdd(dst + 0, s[0]);
    if (s[1]) atomicAdd(dst + 1, s[1]);
    if (s[2]) atomicAdd(dst + 2, s[2]);
  }
}

struct InitArgs {
  int32_t *state;
  uint64_t *mask;
  int32_t *counts;  // block 0
  const float *uniforms;  // slot 0 or null
  uint64_t seed;
  int32_t N, M, W, Mp;
  float p0;
  float step;
  int32_t upper, medium, lower, Ed, Id, Rd;
};

// cu:73-90 with a counter-based draw (slot 0), plus the initial mask and
// the initial counts block.
__global__ __launch_bounds__(kBlock) void init_kernel(InitArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int q = blockIdx.y;
  const int r = q * kWave + lane;
  const bool rv = r < a.M;
  const int i = blockIdx.x * kWavesPerBlock + wave;
  if (i >= a.N) return;
  int32_t c = 0;
  if (rv) {
    const float u = a.uniforms ? a.uniforms[static_cast<size_t>(r) * a.N + i]
                               : rp::philox_uniform(a.seed, static_cast<uint32_t>(i),
                                                    static_cast<uint32_t>(r), 0u);
    if (u < a.p0) c = static_cast<int32_t>(static_cast<float>(a.medium) + a.step);
  }
  a.state[static_cast<size_t>(i) * a.Mp + r] = c;
  const uint64_t bits = __ballot(rv && c > a.lower && c <= a.medium);
  if (lane == 0) a.mask[static_cast<size_t>(i) * a.W + q] = bits;
  if (rv) {
    const int cls = state_class(c, a.upper, a.Ed, a.Id, a.Rd);
    if (cls) atomicAdd(a.counts + static_cast<size_t>(r) * 4 + (cls - 1), 1);
  }
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
  const float E = row_exposure<kIntMode>(a, i, q, lane);
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

}  // namespace

struct replay_ctx {
  replay_params p{};
  rp::Network net;
  rp::Plan plan;
  rp::Slices slices;
  hipStream_t stream = nullptr;
  int4 *d_entries = nullptr;
  int32_t *d_rowptr = nullptr;
  int32_t *d_state = nullptr;
  uint64_t *d_mask[2] = {nullptr, nullptr};
  int32_t *d_counts = nullptr;
  float *d_uniforms = nullptr;
  int32_t N = 0, M = 0, W = 0, Mp = 0;
  int32_t nodes_per_block = 64;
  int cur = 0;
  int32_t steps_done = 0;
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
  hipFree(ctx->d_entries);
  hipFree(ctx->d_rowptr);
  hipFree(ctx->d_state);
  hipFree(ctx->d_mask[0]);
  hipFree(ctx->d_mask[1]);
  hipFree(ctx->d_counts);
  hipFree(ctx->d_uniforms);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  ctx->stream = nullptr;
}

StepArgs step_args(replay_ctx *ctx, int32_t k) {
  const rp::Window &w = ctx->plan.windows[ctx->plan.iter_window[k]];
  const int e = w.epoch;
  const int64_t t0 = ctx->slices.epoch_t0[e];
  StepArgs a;
  std::memset(&a, 0, sizeof a);
  a.rowptr = ctx->d_rowptr + static_cast<size_t>(e) * (ctx->N + 1);
  a.entries = ctx->d_entries + ctx->slices.epoch_entry0[e];
  a.mask_cur = ctx->d_mask[ctx->cur];
  a.mask_next = ctx->d_mask[ctx->cur ^ 1];
  a.state = ctx->d_state;
  a.counts = ctx->d_counts + static_cast<size_t>(k + 1) * ctx->M * 4;
  a.uniforms = ctx->d_uniforms
                   ? ctx->d_uniforms + static_cast<size_t>(k + 1) * ctx->M * ctx->N
                   : nullptr;
  a.seed = ctx->p.seed;
  a.N = ctx->N;
  a.M = ctx->M;
  a.W = ctx->W;
  a.Mp = ctx->Mp;
  a.nodes_per_block = ctx->nodes_per_block;
  a.ws_rel = static_cast<int32_t>(w.ws - t0);
  a.we_rel = static_cast<int32_t>(w.we - t0);
  a.rank_limit = w.rank_limit;
  a.scale = w.scale;
  a.step = ctx->p.step_size;
  a.log_base = ctx->log_base;
  a.upper = ctx->upper;
  a.medium = ctx->medium;
  a.lower = ctx->lower;
  a.Ed = ctx->p.exposed_duration;
  a.Id = ctx->p.infectious_duration;
  a.Rd = ctx->p.resistant_duration;
  a.slot = static_cast<uint32_t>(k + 1);
  return a;
}

int do_reset(replay_ctx *ctx) {
  const size_t counts_n = static_cast<size_t>(ctx->p.iterations + 1) * ctx->M * 4;
  HIP_OR_FAIL(ctx, hipMemsetAsync(ctx->d_counts, 0, counts_n * sizeof(int32_t), ctx->stream));
  HIP_OR_FAIL(ctx, hipMemsetAsync(ctx->d_state, 0,
                                  static_cast<size_t>(ctx->N) * ctx->Mp * sizeof(int32_t),
                                  ctx->stream));
  InitArgs a;
  std::memset(&a, 0, sizeof a);
  a.state = ctx->d_state;
  a.mask = ctx->d_mask[0];
  a.counts = ctx->d_counts;
  a.uniforms = ctx->d_uniforms;
  a.seed = ctx->p.seed;
  a.N = ctx->N;
  a.M = ctx->M;
  a.W = ctx->W;
  a.Mp = ctx->Mp;
  a.p0 = ctx->p.initial_infected;
  a.step = ctx->p.step_size;
  a.upper = ctx->upper;
  a.medium = ctx->medium;
  a.lower = ctx->lower;
  a.Ed = ctx->p.exposed_duration;
  a.Id = ctx->p.infectious_duration;
  a.Rd = ctx->p.resistant_duration;
  dim3 grid((ctx->N + kWavesPerBlock - 1) / kWavesPerBlock, ctx->W);
  hipLaunchKernelGGL(init_kernel, grid, dim3(kBlock), 0, ctx->stream, a);
  HIP_OR_FAIL(ctx, hipGetLastError());
  ctx->cur = 0;
  ctx->steps_done = 0;
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
  HIP_OR_FAIL(ctx, hipSetDevice(p->device));
  HIP_OR_FAIL(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  const size_t nE = ctx->slices.entries.size();
  const size_t nR = ctx->slices.rowptr.size();
  if ((rc = dev_alloc(ctx, &ctx->d_entries, nE)) ||
      (rc = dev_alloc(ctx, &ctx->d_rowptr, nR)) ||
      (rc = dev_alloc(ctx, &ctx->d_state, static_cast<size_t>(ctx->N) * ctx->Mp)) ||
      (rc = dev_alloc(ctx, &ctx->d_mask[0], static_cast<size_t>(ctx->N) * ctx->W)) ||
      (rc = dev_alloc(ctx, &ctx->d_mask[1], static_cast<size_t>(ctx->N) * ctx->W)) ||
      (rc = dev_alloc(ctx, &ctx->d_counts, static_cast<size_t>(p->iterations + 1) * ctx->M * 4)))
    return rc;
  if (nE)
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->d_entries, ctx->slices.entries.data(), nE * sizeof(int4),
                                    hipMemcpyHostToDevice, ctx->stream));
  if (nR)
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->d_rowptr, ctx->slices.rowptr.data(),
                                    nR * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
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
  out->state_bytes = static_cast<int64_t>(ctx->N) * ctx->Mp * 4 +
                     2 * static_cast<int64_t>(ctx->N) * ctx->W * 8;
  out->steps_done = ctx->steps_done;
  out->truncated_windows = ctx->truncated;
  return REPLAY_OK;
}

int32_t replay_reset(replay_ctx *ctx) {
  if (!ctx) return REPLAY_E_ARG;
  return do_reset(ctx);
}

int32_t replay_run(replay_ctx *ctx, int32_t steps) {
  if (!ctx) return REPLAY_E_ARG;
  if (steps < 0 || ctx->steps_done + steps > ctx->p.iterations)
    return fail(ctx, REPLAY_E_ARG, "steps exceed the planned iterations");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  dim3 grid((ctx->N + ctx->nodes_per_block - 1) / ctx->nodes_per_block, ctx->W);
  for (int32_t s = 0; s < steps; ++s) {
    const int32_t k = ctx->steps_done;
    StepArgs a = step_args(ctx, k);
    const bool int_mode = ctx->plan.windows[ctx->plan.iter_window[k]].int_mode != 0;
    const bool inj = a.uniforms != nullptr;
    if (int_mode && !inj)
      hipLaunchKernelGGL((step_kernel<true, false>), grid, dim3(kBlock), 0, ctx->stream, a);
    else if (!int_mode && !inj)
      hipLaunchKernelGGL((step_kernel<false, false>), grid, dim3(kBlock), 0, ctx->stream, a);
    else if (int_mode && inj)
      hipLaunchKernelGGL((step_kernel<true, true>), grid, dim3(kBlock), 0, ctx->stream, a);
    else
      hipLaunchKernelGGL((step_kernel<false, true>), grid, dim3(kBlock), 0, ctx->stream, a);
    HIP_OR_FAIL(ctx, hipGetLastError());
    ctx->cur ^= 1;
    ctx->steps_done = k + 1;
  }
  return REPLAY_OK;
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
  std::vector<int32_t> tmp(static_cast<size_t>(ctx->N) * ctx->Mp);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(tmp.data(), ctx->d_state, tmp.size() * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  for (int32_t i = 0; i < ctx->N; ++i)
    for (int32_t r = 0; r < ctx->M; ++r)
      out[static_cast<size_t>(r) * ctx->N + i] = tmp[static_cast<size_t>(i) * ctx->Mp + r];
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
