// marginal.hip -- the deterministic individual-level marginal-probability
// form of the replay and its hand-written adjoint (C ABI: replay_marginal_*).
//
// Model, per parameter set b and node i (probabilities of S, E, I, R):
//   E_i(k)  = sum_j A_k[i,j] * iota_j(k)      exposure seconds in window k,
//             A_k the duration-weighted slice of the stochastic step
//             (weight (float)overlap * scale, cut contacts excluded)
//   lam     = 1 - exp(-beta_b * phi_i * E_i / 3600)    beta per hour of contact
//   s'      = s (1 - lam)
//   e'      = e (1 - a_b) + s lam                    a_b = 1 - exp(-sigma_b dt)
//   iota'   = iota (1 - c_b) + e a_b                 c_b = 1 - exp(-gamma_b dt)
//   r'      = r + iota c_b                            dt = step in hours
// from s = 1 - p0, iota = p0, e = r = 0 (p0 = initial_infected of the node's
// network).  Curves C[k][b] = (sum_i s, sum_i e, sum_i iota, sum_i r).
// The stochastic replay's fixed sojourns become exponential ones with the
// same means, which makes the curves differentiable in sigma and gamma.
//
// Adjoint of one step, with bars the gradients of the loss at k+1 (plus the
// curve gradient of iteration k+1) and D = sbar - ebar:
//   G_i      = dL/dE_i = -kappa X s D,  kappa = beta phi / 3600, X = 1 - lam
//   dL/ds    = sbar X + ebar lam
//   dL/de    = ebar (1 - a) + ibar a
//   dL/diota = ibar (1 - c) + rbar c + sum_i A_k[i,j] G_i
//   dL/dr    = rbar
//   dL/dkappa = -s D E X  ->  dL/dbeta_b += dL/dkappa phi/3600,
//                               dL/dphi_i += sum_b dL/dkappa beta_b/3600
//   dL/da_b  += e (ibar - ebar),  dL/dc_b += iota (rbar - ibar)
// and da/dsigma = dt (1 - a), dc/dgamma = dt (1 - c).  The slices hold both
// directions of every contact with the same weight, so A_k is symmetric and
// the transposed product sum_i A_k[i,j] G_i is row j of the same packed CSR.
//
// Layout: state and checkpoint planes are [3][N][Bp] fp32 (s, e, iota: r
// never enters the dynamics, and s + e + iota + r = 1 for every node and
// set, so r is 1 - s - e - iota where its curve is formed), adjoint planes
// [3][N][Bp] (s, e, iota; the adjoint of r is the same for every node and is
// kept per set), with the parameter sets contiguous (Bp = B rounded up to
// 64): a wavefront owns a
// node, lane = parameter set, and every neighbour row is one coalesced
// 256-byte load.  Reductions over nodes go through per-wavefront partials
// in a fixed order, so results do not depend on scheduling.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

using namespace rp;

namespace {

struct MfStep {
  const int32_t *rowptr;
  const int4 *entries;
  const uint2 *entries8;  // the same entries in 8 bytes, or null
  const uint32_t *rank_limit;
  const int32_t *net_of;
  int32_t ws_rel, we_rel;
  float scale;
};

MfStep mf_step(replay_ctx *ctx, int32_t k) {
  const StepArgs s = step_args(ctx, k);
  return MfStep{s.rowptr, s.entries, s.entries8, s.rank_limit, s.net_of, s.ws_rel, s.we_rel,
                s.scale};
}

constexpr int kRowBatch = 8;  // neighbour rows requested together
constexpr int kMfPlanes = 3;  // state planes per node and set: s, e, iota

// sum_j A_k[row i, j] * v[j][b] in entry order (ascending column).  The
// wavefront loads 64 of the row's entries at a time (one per lane, from the
// 8-byte copy where EnT is v2u), then takes the counting ones in order,
// kRowBatch at a time: their neighbour rows (one coalesced load of the Bp
// sets each) are requested together and then added in entry order, so the
// sum is the sequential one, term for term.
//
// The row's bounds and query cap come in a RowBounds read by the caller
// (row_bounds), so that a caller can request them before its other inputs:
// vector loads complete in order, and waiting for bounds requested after
// those inputs would wait for the inputs as well.
struct RowBounds {
  int e0, e1;
  uint32_t rlim;
};

__device__ __forceinline__ RowBounds row_bounds(const MfStep &g, int i) {
  return RowBounds{g.rowptr[i], g.rowptr[i + 1],
                   g.rank_limit ? g.rank_limit[g.net_of[i]] : UINT32_MAX};
}

template <typename EnT>
__device__ __forceinline__ float row_product(const MfStep &g, const RowBounds &rb, const float *v,
                                             int Bp, int b, int lane) {
  const int e0 = rb.e0, e1 = rb.e1;
  const uint32_t rlim = rb.rlim;
  const auto *src = [&] {
    if constexpr (sizeof(EnT) == 8)
      return g.entries8;
    else
      return g.entries;
  }();
  float acc = 0.0f;
  for (int c0 = e0; c0 < e1; c0 += kWave) {
    const EntryFields en = fields(load_chunk(src, c0, e1, lane));
    const int ov = min(en.te, g.we_rel) - max(en.ts, g.ws_rel);
    const bool on = c0 + lane < e1 && ov > 0 && en.rank <= rlim;
    const float w = on ? static_cast<float>(ov) * g.scale : 0.0f;
    const int col = on ? en.col : 0;
    for (uint64_t act = __ballot(on); act != 0ull;) {
      float wj[kRowBatch], vj[kRowBatch];
      int n = 0;
#pragma unroll
      for (int u = 0; u < kRowBatch; ++u) {
        wj[u] = 0.0f;
        vj[u] = 0.0f;
        if (act != 0ull) {
          const int j = __builtin_ctzll(act);
          act &= act - 1;
          const int cj = __builtin_amdgcn_readlane(col, j);
          wj[u] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w), j));
          vj[u] = v[static_cast<size_t>(cj) * Bp + b];
          n = u + 1;
        }
      }
#pragma unroll
      for (int u = 0; u < kRowBatch; ++u)
        if (u < n) acc += wj[u] * vj[u];
    }
  }
  return acc;
}

struct MfParams {
  const float *kb;    // [Bp] beta / 3600
  const float *ac;    // [2][Bp] a, c
  const float *phi;   // [N]
  int32_t N, Bp;
};

__global__ __launch_bounds__(kBlock) void mf_prep_kernel(const float *beta, const float *sigma,
                                                         const float *gamma, const float *phi_in,
                                                         float *kb, float *ac, float *phi, int B,
                                                         int Bp, int N, float dt_h) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < Bp) {
    const bool on = t < B;
    kb[t] = on ? beta[t] / 3600.0f : 0.0f;
    ac[t] = on ? -expm1f(-sigma[t] * dt_h) : 0.0f;
    ac[Bp + t] = on ? -expm1f(-gamma[t] * dt_h) : 0.0f;
  }
  if (t < N) phi[t] = phi_in[t];
}

// Blocks of the forward and initial-state kernels: 16 wavefronts, so that
// many nodes are in flight while the per-block partials of the curves stay
// as few as with 4-wavefront blocks over a quarter of the wavefronts.
// (forward ms per calibration step on the 20k-node workload with 4, 8 and
// 16 wavefronts per block: 11.07, 9.47, 8.61)
#ifndef REPLAY_MF_BLOCK_WAVES
#define REPLAY_MF_BLOCK_WAVES 16
#endif
constexpr int kMfWaves = REPLAY_MF_BLOCK_WAVES;
constexpr int kMfBlock = kMfWaves * kWave;
// Wavefronts of the forward and transposed-product kernels (at most one per
// node): 8,192, one round of the wavefront slots of 256 CUs at 8 per SIMD,
// each taking every 8,192nd node with the next node's row bounds requested
// ahead.  Measured on the 20k-node workload (forward ms per calibration
// step): 4,096 -> 11.06, 8,192 -> 10.64, 12,288 -> 11.35, 16,384 -> 12.21.
#ifndef REPLAY_MF_MAX_WAVES
#define REPLAY_MF_MAX_WAVES 8192
#endif
constexpr int kMfMaxWaves = REPLAY_MF_MAX_WAVES;
// Forward steps whose curve partials are kept before one reduction forms
// their curves (a reduction per step cost a launch of about 4 us per step).
constexpr int kPartSteps = 64;

// Adds the four compartment sums of the block's wavefronts in wavefront
// order and writes them as the block's partial part[block][b][4].
__device__ __forceinline__ void block_partial(const float (&sum)[4], int b, int Bp, float *part) {
  __shared__ float red[kMfWaves][kWave][4];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
#pragma unroll
  for (int c = 0; c < 4; ++c) red[wv][lane][c] = sum[c];
  __syncthreads();
  if (wv == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float acc = red[0][lane][c];
#pragma unroll
      for (int w = 1; w < kMfWaves; ++w) acc += red[w][lane][c];
      part[(static_cast<size_t>(blockIdx.x) * Bp + b) * 4 + c] = acc;
    }
  }
  // the rows are written again only for a further chunk of parameter sets;
  // after the last one the other wavefronts need not wait for the sum
  if (b - lane + kWave < Bp) __syncthreads();
}

// Initial state and its curve partials.
__global__ __launch_bounds__(kMfBlock) void mf_init_kernel(float *st, const NetDev *nets,
                                                           const int32_t *net_of, float *part,
                                                           int N, int Bp) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = blockIdx.x * kMfWaves + static_cast<int>(threadIdx.x / kWave);
  const int n_waves = gridDim.x * kMfWaves;
  const size_t NB = static_cast<size_t>(N) * Bp;
  for (int cb = 0; cb < Bp; cb += kWave) {
    const int b = cb + lane;
    float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = wave; i < N; i += n_waves) {
      const float p0 = nets[net_of[i]].p0;
      const size_t o = static_cast<size_t>(i) * Bp + b;
      st[o] = 1.0f - p0;
      st[NB + o] = 0.0f;
      st[2 * NB + o] = p0;
      sum[0] += 1.0f - p0;
      sum[2] += p0;
    }
    block_partial(sum, b, Bp, part);
  }
}

// One forward step k -> k+1; optionally records the exposure of k.
template <typename EnT>
__global__ __launch_bounds__(kMfBlock) void mf_forward_kernel(MfStep g, MfParams p,
                                                              const float *cur, float *next,
                                                              float *E_out, float *part) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = blockIdx.x * kMfWaves + static_cast<int>(threadIdx.x / kWave);
  const int n_waves = gridDim.x * kMfWaves;
  const size_t NB = static_cast<size_t>(p.N) * p.Bp;
  for (int cb = 0; cb < p.Bp; cb += kWave) {
    const int b = cb + lane;
    const float kb = p.kb[b], a = p.ac[b], c = p.ac[p.Bp + b];
    float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // the row's bounds, then the node's state and susceptibility, are
    // requested first, so that the state's round trip overlaps the row
    // product's (entries, neighbour rows) instead of following it; the next
    // node's bounds are requested with the state, so that they are in when
    // its turn comes
    RowBounds rb = wave < p.N ? row_bounds(g, wave) : RowBounds{0, 0, 0u};
    for (int i = wave; i < p.N; i += n_waves) {
      const size_t o = static_cast<size_t>(i) * p.Bp + b;
      const float s = cur[o], e = cur[NB + o], io = cur[2 * NB + o];
      const float phi = p.phi[i];
      const RowBounds rb_next = i + n_waves < p.N ? row_bounds(g, i + n_waves) : rb;
      const float E = row_product<EnT>(g, rb, cur + 2 * NB, p.Bp, b, lane);
      rb = rb_next;
      const float lam = -expm1f(-(kb * phi) * E);
      const float s1 = s * (1.0f - lam);
      const float e1 = e * (1.0f - a) + s * lam;
      const float i1 = io * (1.0f - c) + e * a;
      const float r1 = 1.0f - s1 - e1 - i1;  // (r + iota c in exact arithmetic)
      next[o] = s1;
      next[NB + o] = e1;
      next[2 * NB + o] = i1;
      if (E_out) E_out[o] = E;
      sum[0] += s1;
      sum[1] += e1;
      sum[2] += i1;
      sum[3] += r1;
    }
    // (the recomputation in the backward pass needs no curves)
    if (part) block_partial(sum, b, p.Bp, part);
  }
}

// curves[s][b][c] = sum over blocks of part[s][blk][b][c] for the n_steps
// steps of a group: one block per output, strided per-thread sums and a
// fixed-order tree, so the result does not depend on scheduling.  The steps'
// partials lie part_stride floats apart and their curves B * 4 apart.
__global__ __launch_bounds__(kBlock) void mf_reduce_kernel(const float *part, int n_parts, int Bp,
                                                           int B, size_t part_stride,
                                                           float *curves) {
  __shared__ float red[kBlock];
  const int t = blockIdx.x % (B * 4);  // b * 4 + c
  const int step = blockIdx.x / (B * 4);
  const int b = t >> 2, c = t & 3;
  const float *ps = part + static_cast<size_t>(step) * part_stride;
  float acc = 0.0f;
  for (int w = threadIdx.x; w < n_parts; w += kBlock)
    acc += ps[(static_cast<size_t>(w) * Bp + b) * 4 + c];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) curves[static_cast<size_t>(step) * B * 4 + t] = red[0];
}

__global__ void mf_zero_kernel(float *p, size_t n) {
  for (size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[t] = 0.0f;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int w = 1; w < kWave; w <<= 1) v += __shfl_xor(v, w);
  return v;
}

// Adjoint phase 1 of step k: transition adjoints, dL/dE and the parameter
// partials.  adj_in holds the adjoint at k+1 without the curve term, gc the
// curve gradient of iteration k+1 ([B][4], sets >= B contribute nothing).
//
// The adjoint of r is the same for every node (r feeds nothing back): the
// sum of the R-curve gradients of the later iterations.  It is kept as one
// value per parameter set, rbar_in at k+1 and rbar_out at k (written by
// wavefront 0), which is the per-node sequence of fp32 additions without
// its N-fold plane.
__global__ __launch_bounds__(kBlock) void mf_adjoint_kernel(MfParams p, const float *st,
                                                            const float *E, const float *adj_in,
                                                            const float *rbar_in,
                                                            const float *gc, int B, float *adj_out,
                                                            float *rbar_out, float *G, float *pg,
                                                            float *phig) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = blockIdx.x * kWavesPerBlock + static_cast<int>(threadIdx.x / kWave);
  const int n_waves = gridDim.x * kWavesPerBlock;
  const size_t NB = static_cast<size_t>(p.N) * p.Bp;
  for (int cb = 0; cb < p.Bp; cb += kWave) {
    const int b = cb + lane;
    const float kb = p.kb[b], a = p.ac[b], c = p.ac[p.Bp + b];
    float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (b < B) {
#pragma unroll
      for (int q = 0; q < 4; ++q) g[q] = gc[b * 4 + q];
    }
    float d_beta = 0.0f, d_a = 0.0f, d_c = 0.0f;
    const float rb = rbar_in[b] + g[3];
    if (wave == 0) rbar_out[b] = rb;
    for (int i = wave; i < p.N; i += n_waves) {
      const size_t o = static_cast<size_t>(i) * p.Bp + b;
      const float s = st[o], e = st[NB + o], io = st[2 * NB + o];
      const float sb = adj_in[o] + g[0], eb = adj_in[NB + o] + g[1];
      const float ib = adj_in[2 * NB + o] + g[2];
      const float phi = p.phi[i];
      const float kappa = kb * phi;
      const float Ei = E[o];
      const float lam = -expm1f(-kappa * Ei);
      const float X = 1.0f - lam;
      const float D = sb - eb;
      G[o] = -kappa * X * s * D;
      adj_out[o] = sb * X + eb * lam;
      adj_out[NB + o] = eb * (1.0f - a) + ib * a;
      adj_out[2 * NB + o] = ib * (1.0f - c) + rb * c;
      const float dk = -s * D * Ei * X;
      d_beta += dk * (phi / 3600.0f);
      d_a += e * (ib - eb);
      d_c += io * (rb - ib);
      const float dphi = wave_sum(dk * kb);
      if (lane == 0) phig[i] += dphi;
    }
    float *q = pg + (static_cast<size_t>(wave) * p.Bp + b) * 3;
    q[0] += d_beta;
    q[1] += d_a;
    q[2] += d_c;
  }
}

// Adjoint phase 2 of step k: adj_out[iota][j] += sum_i A_k[j,i] G_i.
template <typename EnT>
__global__ __launch_bounds__(kMfBlock) void mf_spmm_adj_kernel(MfStep g, MfParams p,
                                                               const float *G, float *adj_out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = blockIdx.x * kMfWaves + static_cast<int>(threadIdx.x / kWave);
  const int n_waves = gridDim.x * kMfWaves;
  const size_t NB = static_cast<size_t>(p.N) * p.Bp;
  for (int cb = 0; cb < p.Bp; cb += kWave) {
    const int b = cb + lane;
    // as in the forward kernel: the next row's bounds are requested with
    // this row's output word, before this row's product
    RowBounds rb = wave < p.N ? row_bounds(g, wave) : RowBounds{0, 0, 0u};
    for (int j = wave; j < p.N; j += n_waves) {
      float *out = adj_out + 2 * NB + static_cast<size_t>(j) * p.Bp + b;
      const float prev = *out;
      const RowBounds rb_next = j + n_waves < p.N ? row_bounds(g, j + n_waves) : rb;
      *out = prev + row_product<EnT>(g, rb, G, p.Bp, b, lane);
      rb = rb_next;
    }
  }
}

// Final parameter gradients from the per-wavefront partials: one block per
// (set, parameter), strided sums and a fixed-order tree.
__global__ __launch_bounds__(kBlock) void mf_grad_kernel(const float *pg, int n_waves, int Bp,
                                                         const float *ac, float dt_h,
                                                         float *g_beta, float *g_sigma,
                                                         float *g_gamma) {
  __shared__ float red[kBlock];
  const int b = blockIdx.x / 3, q = blockIdx.x % 3;
  float acc = 0.0f;
  for (int w = threadIdx.x; w < n_waves; w += kBlock)
    acc += pg[(static_cast<size_t>(w) * Bp + b) * 3 + q];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (q == 0)
    g_beta[b] = red[0];
  else if (q == 1)
    g_sigma[b] = red[0] * dt_h * (1.0f - ac[b]);
  else
    g_gamma[b] = red[0] * dt_h * (1.0f - ac[Bp + b]);
}

}  // namespace

struct replay_marginal {
  replay_ctx *ctx = nullptr;
  int32_t B = 0, Bp = 0, T = 0, every = 0, N = 0, n_waves = 0, n_ckpt = 0;
  int32_t f_blocks = 0;  // blocks of the forward kernel (and curve partials per step)
  float dt_h = 0.0f;
  bool have_forward = false;
  float *d_st[2] = {nullptr, nullptr};  // [4][N][Bp]
  float *d_ckpt = nullptr;              // [n_ckpt][4][N][Bp]
  float *d_seg = nullptr;               // [every + 1][4][N][Bp]
  float *d_segE = nullptr;              // [every][N][Bp]
  float *d_adj[2] = {nullptr, nullptr}; // [3][N][Bp]: adjoints of s, e, iota
  float *d_rbar[2] = {nullptr, nullptr}; // [Bp]: the adjoint of r (the same for every node)
  float *d_G = nullptr;                 // [N][Bp]
  float *d_part = nullptr;              // [kPartSteps][n_blocks][Bp][4]
  float *d_pg = nullptr;                // [n_waves][Bp][3]
  float *d_phig = nullptr;              // [N]
  float *d_kb = nullptr, *d_ac = nullptr, *d_phi = nullptr;
};

namespace {

void mf_release(replay_marginal *m) {
  float *bufs[] = {m->d_st[0], m->d_st[1], m->d_ckpt, m->d_seg,  m->d_segE, m->d_adj[0],
                   m->d_adj[1], m->d_G,    m->d_part, m->d_pg,  m->d_phig, m->d_kb,
                   m->d_ac,     m->d_phi,  m->d_rbar[0], m->d_rbar[1]};
  for (float *b : bufs) (void)hipFree(b);
}

MfParams mf_params(const replay_marginal *m) {
  return MfParams{m->d_kb, m->d_ac, m->d_phi, m->N, m->Bp};
}

// The caller's stream, used as given: NULL is HIP's default (null) stream,
// which is also what a torch process passes for its default stream.
hipStream_t mf_stream(void *stream) { return static_cast<hipStream_t>(stream); }

// The forward step and the transposed product, streaming the 8-byte
// entries where the context holds them.
void launch_forward(const MfStep &g, const MfParams &p, const float *cur, float *next, float *E_out,
                    float *part, unsigned blocks, hipStream_t st) {
  if (g.entries8)
    hipLaunchKernelGGL(mf_forward_kernel<v2u>, dim3(blocks), dim3(kMfBlock), 0, st, g, p, cur,
                       next, E_out, part);
  else
    hipLaunchKernelGGL(mf_forward_kernel<v4i>, dim3(blocks), dim3(kMfBlock), 0, st, g, p, cur,
                       next, E_out, part);
}

void launch_spmm_adj(const MfStep &g, const MfParams &p, const float *G, float *adj_out,
                     unsigned blocks, hipStream_t st) {
  if (g.entries8)
    hipLaunchKernelGGL(mf_spmm_adj_kernel<v2u>, dim3(blocks), dim3(kMfBlock), 0, st, g, p, G,
                       adj_out);
  else
    hipLaunchKernelGGL(mf_spmm_adj_kernel<v4i>, dim3(blocks), dim3(kMfBlock), 0, st, g, p, G,
                       adj_out);
}

void zero(float *p, size_t n, hipStream_t st) {
  const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(mf_zero_kernel, dim3(std::max(1u, blocks)), dim3(256), 0, st, p, n);
}

}  // namespace

extern "C" {

int32_t replay_marginal_create(replay_ctx *ctx, int32_t n_sets, int32_t checkpoint_every,
                               replay_marginal **out) {
  if (!out) return REPLAY_E_ARG;
  *out = nullptr;
  if (!ctx || n_sets <= 0 || checkpoint_every <= 0) return REPLAY_E_ARG;
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  replay_marginal *m = new replay_marginal();
  m->ctx = ctx;
  m->B = n_sets;
  m->Bp = (n_sets + kWave - 1) / kWave * kWave;
  m->T = ctx->T;
  m->N = ctx->N;
  m->every = std::min(checkpoint_every, std::max(ctx->T, 1));
  m->n_ckpt = (m->T + m->every - 1) / m->every;
  m->dt_h = ctx->p.step_size / 3600.0f;
  m->n_waves = std::max(1, std::min(ctx->N, 4096)) / kWavesPerBlock * kWavesPerBlock;
  if (m->n_waves == 0) m->n_waves = kWavesPerBlock;
  m->f_blocks = std::max(1, std::min(ctx->N, kMfMaxWaves) / kMfWaves);
  const size_t plane = static_cast<size_t>(m->N) * m->Bp;
  int rc = REPLAY_OK;
  if ((rc = dev_alloc(ctx, &m->d_st[0], kMfPlanes * plane)) ||
      (rc = dev_alloc(ctx, &m->d_st[1], kMfPlanes * plane)) ||
      (rc = dev_alloc(ctx, &m->d_ckpt, kMfPlanes * plane * std::max(m->n_ckpt, 1))) ||
      (rc = dev_alloc(ctx, &m->d_seg, kMfPlanes * plane * (m->every + 1))) ||
      (rc = dev_alloc(ctx, &m->d_segE, plane * m->every)) ||
      (rc = dev_alloc(ctx, &m->d_adj[0], 3 * plane)) || (rc = dev_alloc(ctx, &m->d_adj[1], 3 * plane)) ||
      (rc = dev_alloc(ctx, &m->d_rbar[0], static_cast<size_t>(m->Bp))) ||
      (rc = dev_alloc(ctx, &m->d_rbar[1], static_cast<size_t>(m->Bp))) ||
      (rc = dev_alloc(ctx, &m->d_G, plane)) ||
      (rc = dev_alloc(ctx, &m->d_part,
                      static_cast<size_t>(kPartSteps) * m->f_blocks * m->Bp * 4)) ||
      (rc = dev_alloc(ctx, &m->d_pg, static_cast<size_t>(m->n_waves) * m->Bp * 3)) ||
      (rc = dev_alloc(ctx, &m->d_phig, static_cast<size_t>(m->N))) ||
      (rc = dev_alloc(ctx, &m->d_kb, static_cast<size_t>(m->Bp))) ||
      (rc = dev_alloc(ctx, &m->d_ac, static_cast<size_t>(2 * m->Bp))) ||
      (rc = dev_alloc(ctx, &m->d_phi, static_cast<size_t>(m->N)))) {
    mf_release(m);
    delete m;
    return rc;
  }
  *out = m;
  return REPLAY_OK;
}

int32_t replay_marginal_forward(replay_marginal *m, const float *beta, const float *sigma,
                                const float *gamma, const float *phi, float *curves,
                                void *stream) {
  if (!m || !beta || !sigma || !gamma || !phi || !curves) return REPLAY_E_ARG;
  replay_ctx *ctx = m->ctx;
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  const hipStream_t st = mf_stream(stream);
  const size_t plane = static_cast<size_t>(m->N) * m->Bp;
  const unsigned blocks = static_cast<unsigned>(m->f_blocks);
  const unsigned red_blocks = static_cast<unsigned>(m->B * 4);
  const int n_parts = m->f_blocks;
  const int prep_n = std::max(m->Bp, m->N);
  hipLaunchKernelGGL(mf_prep_kernel, dim3((prep_n + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                     beta, sigma, gamma, phi, m->d_kb, m->d_ac, m->d_phi, m->B, m->Bp, m->N,
                     m->dt_h);
  hipLaunchKernelGGL(mf_init_kernel, dim3(blocks), dim3(kMfBlock), 0, st, m->d_st[0], ctx->d_nets,
                     ctx->d_net_of, m->d_part, m->N, m->Bp);
  const size_t part_stride = static_cast<size_t>(n_parts) * m->Bp * 4;
  hipLaunchKernelGGL(mf_reduce_kernel, dim3(red_blocks), dim3(kBlock), 0, st, m->d_part, n_parts,
                     m->Bp, m->B, part_stride, curves);
  const MfParams p = mf_params(m);
  int cur = 0;
  // each step leaves its block partials in slot k mod kPartSteps; one
  // reduction per group of kPartSteps steps forms their curves
  for (int32_t k = 0; k < m->T; ++k) {
    if (k % m->every == 0)
      HIP_OR_FAIL(ctx, hipMemcpyAsync(m->d_ckpt + static_cast<size_t>(k / m->every) * kMfPlanes * plane,
                                      m->d_st[cur], kMfPlanes * plane * sizeof(float),
                                      hipMemcpyDeviceToDevice, st));
    launch_forward(mf_step(ctx, k), p, m->d_st[cur], m->d_st[cur ^ 1], nullptr,
                   m->d_part + static_cast<size_t>(k % kPartSteps) * part_stride, blocks, st);
    if (k % kPartSteps == kPartSteps - 1 || k == m->T - 1) {
      const int32_t k0 = k - k % kPartSteps;
      hipLaunchKernelGGL(mf_reduce_kernel, dim3(red_blocks * static_cast<unsigned>(k - k0 + 1)),
                         dim3(kBlock), 0, st, m->d_part, n_parts, m->Bp, m->B, part_stride,
                         curves + static_cast<size_t>(k0 + 1) * m->B * 4);
    }
    cur ^= 1;
  }
  HIP_OR_FAIL(ctx, hipGetLastError());
  m->have_forward = true;
  return REPLAY_OK;
}

int32_t replay_marginal_backward(replay_marginal *m, const float *grad_curves, float *grad_beta,
                                 float *grad_sigma, float *grad_gamma, float *grad_phi,
                                 void *stream) {
  if (!m || !grad_curves || !grad_beta || !grad_sigma || !grad_gamma || !grad_phi)
    return REPLAY_E_ARG;
  replay_ctx *ctx = m->ctx;
  if (!m->have_forward) return fail(ctx, REPLAY_E_STATE, "replay_marginal_backward before forward");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  const hipStream_t st = mf_stream(stream);
  const size_t plane = static_cast<size_t>(m->N) * m->Bp;
  const unsigned blocks = static_cast<unsigned>(m->n_waves / kWavesPerBlock);
  const MfParams p = mf_params(m);
  zero(m->d_adj[0], 3 * plane, st);
  zero(m->d_rbar[0], static_cast<size_t>(m->Bp), st);
  zero(m->d_pg, static_cast<size_t>(m->n_waves) * m->Bp * 3, st);
  zero(m->d_phig, static_cast<size_t>(m->N), st);
  int a = 0;  // adjoint at k+1 lives in d_adj[a]
  for (int32_t sgi = m->n_ckpt - 1; sgi >= 0; --sgi) {
    const int32_t k0 = sgi * m->every, k1 = std::min(k0 + m->every, m->T);
    // recompute the segment's states and exposures from its checkpoint
    HIP_OR_FAIL(ctx, hipMemcpyAsync(m->d_seg, m->d_ckpt + static_cast<size_t>(sgi) * kMfPlanes * plane,
                                    kMfPlanes * plane * sizeof(float), hipMemcpyDeviceToDevice, st));
    for (int32_t k = k0; k < k1; ++k) {
      const size_t j = static_cast<size_t>(k - k0);
      launch_forward(mf_step(ctx, k), p, m->d_seg + j * kMfPlanes * plane,
                     m->d_seg + (j + 1) * kMfPlanes * plane, m->d_segE + j * plane, nullptr,
                     static_cast<unsigned>(m->f_blocks), st);
    }
    for (int32_t k = k1 - 1; k >= k0; --k) {
      const size_t j = static_cast<size_t>(k - k0);
      hipLaunchKernelGGL(mf_adjoint_kernel, dim3(blocks), dim3(kBlock), 0, st, p,
                         m->d_seg + j * kMfPlanes * plane, m->d_segE + j * plane, m->d_adj[a],
                         m->d_rbar[a], grad_curves + static_cast<size_t>(k + 1) * m->B * 4, m->B,
                         m->d_adj[a ^ 1], m->d_rbar[a ^ 1], m->d_G, m->d_pg, m->d_phig);
      launch_spmm_adj(mf_step(ctx, k), p, m->d_G, m->d_adj[a ^ 1],
                      static_cast<unsigned>(m->f_blocks), st);
      a ^= 1;
    }
  }
  hipLaunchKernelGGL(mf_grad_kernel, dim3(3 * m->B), dim3(kBlock), 0, st, m->d_pg, m->n_waves,
                     m->Bp, m->d_ac, m->dt_h, grad_beta, grad_sigma, grad_gamma);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(grad_phi, m->d_phig, static_cast<size_t>(m->N) * sizeof(float),
                                  hipMemcpyDeviceToDevice, st));
  HIP_OR_FAIL(ctx, hipGetLastError());
  return REPLAY_OK;
}

void replay_marginal_destroy(replay_marginal *m) {
  if (!m) return;
  if (m->ctx && m->ctx->stream) (void)hipStreamSynchronize(m->ctx->stream);
  (void)hipDeviceSynchronize();
  mf_release(m);
  delete m;
}

}  // extern "C"
