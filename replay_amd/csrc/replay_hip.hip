// replay_hip.hip -- libreplay_hip: HIP kernels for gfx950 and the C ABI
// declared in include/replay_hip.h.
//
// A context holds one or more networks (a batch) as a disjoint union of node
// rows; one step launch advances every network.  Replica state is
// bit-packed: a word (node i, q) holds replicas 64q .. 64q+63, and per word
// the device keeps
//   S       susceptible bits at the current iteration     (countdown == 0)
//   I[2]    infectious bits, double-buffered               (lower < c <= medium)
//   X ring  infection bits of each past iteration, depth D
//   X_init  bits seeded infected by the initial draw
// plus, per replica, entry[i][r] (iteration of the last infection, written
// only when one happens) and inf_count[k][net][r] (infections of iteration k).
//
// Why this is exact: after an infection the countdown follows one fixed
// sequence, c(n) = f^n(c0) with f(c) = (int)((float)c - step) clamped at 0
// (tool/gpu_cpu_temporal_sim.cu:153-175), the same for every node and
// replica of a network (new infection: c0 = (int)(upper + step); initial
// seed: c0 = (int)(medium + step), cu:85).  The host tabulates c(n) once per
// network with the same float operations and derives the offsets at which a
// replica turns infectious (n_on), stops being infectious (n_off) and becomes
// susceptible again (n_S).  So
//   I(k+1) = I(k) ^ X(k+1-n_on) ^ X(k+1-n_off),  S(k+1) = S(k) & ~X(k) | X(k+1-n_S)
// and the exposed/infectious/resistant counts of iteration k are sums of
// inf_count over the offset ranges where the tabulated class
// (cu:530-550) is E, I or R.
//
// Each iteration is two kernels.  expo_kernel gives a wavefront to one node:
// the lanes load 64 of the node's packed contact entries at a time
// (coalesced), gather the neighbours' infectious words and reduce them to
// per-replica exposure -- whole-window contacts by a butterfly column count
// across the lanes, partial overlaps one contact at a time -- then draw
// Philox uniforms, compare each with 1 - exp(log(1-p) * E/3600) (from a
// hardware-exponential estimate, or the exact expression when the uniform
// lies within 1e-5 of it) and ballot the infections into the ring slot of
// the iteration.  update_kernel then gives a
// thread to each word and applies the flips above.  Variants NODE4/NODE16
// keep 4 or 16 of a node's words in registers per pass over its contacts
// (WORD: one); SLICED gives each lane one (node, word) and counts contacts in
// bit planes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <climits>
#include <numeric>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "engine.hpp"

namespace rp {
thread_local std::string g_open_error;
}  // namespace rp

// Staging of replay_get_exposure_noalloc: the exposure matrix of one call,
// written by exposure_kernel and copied to the caller's host buffer.  One
// array per device, held until the process ends; its device address comes
// from hipGetSymbolAddress (in host code the name is only a stand-in).
constexpr int64_t kNoAllocFloats = int64_t(1) << 24;  // 64 MiB
static __device__ __align__(256) float scratch[kNoAllocFloats];

using namespace rp;

namespace {

#ifndef REPLAY_FEW_CONTACTS
#define REPLAY_FEW_CONTACTS 2
#endif
// Chunk-words with at most this many whole-window contacts whose words are
// nonzero are added one contact at a time instead of by the column count.
constexpr int kFewContacts = REPLAY_FEW_CONTACTS;

// Half-width of the band around a uniform inside which an infection draw
// evaluates the reference's probability expression instead of its estimate.
constexpr float kDecideMargin = 1.0e-5f;


// Device buffers freed when the scope ends, whichever way it is left.
struct DeviceScratch {
  std::vector<void *> bufs;
  ~DeviceScratch() {
    for (void *b : bufs) (void)hipFree(b);
  }
};

__device__ __forceinline__ const uint64_t *ring_word(const StepArgs &a, int32_t n, int64_t w) {
  // X(k+1-n) for n >= 2 (n == 1 is the local infection word)
  const int32_t e = a.k + 1 - n;
  if (n == kNever || n < 2 || e < 0) return nullptr;
  return a.X + static_cast<size_t>(e % a.depth) * a.NW + w;
}

// Exposure seconds of (row i, replica lane) for one window.
// Integer mode (divisor 1, row weight sums < 2^24): overlaps are integers and
// every partial sum is exact, so the integer sum equals the reference's fp32
// sum in any order.  Float mode: contributions fl(ov * scale) of one column
// are merged in rank order (cu:299-309), and columns are added in ascending
// order from zero (Eigen SparseDenseProduct.h:64-70).
template <bool kIntMode>
__device__ __forceinline__ float row_exposure(int e0, int e1, const int4 *entries,
                                              const uint64_t *I, int32_t W, int32_t ws_rel,
                                              int32_t we_rel, uint32_t rank_limit, float scale,
                                              int q, int lane) {
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

// Word of lane (lane ^ w) within a 16-lane row, by DPP lane swizzles in
// the vector pipe: quad permutations for w = 1, 2; for w = 4, the quad
// reversal (lane ^ 3) followed by the half-row mirror (lane ^ 7); for w = 8
// the row rotation by 8.
template <int W>
__device__ __forceinline__ uint32_t row_xor(uint32_t v) {
  const int x = static_cast<int>(v);
  if (W == 1) return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true));
  if (W == 2) return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true));
  if (W == 4)
    return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(
        __builtin_amdgcn_mov_dpp(x, 0x1B, 0xF, 0xF, true), 0x141, 0xF, 0xF, true));
  return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, true));
}

template <int K>
__device__ __forceinline__ void count_stage(uint64_t &x, int lane) {
  constexpr uint64_t kLow[5] = {0x5555555555555555ull, 0x3333333333333333ull,
                                0x0F0F0F0F0F0F0F0Full, 0x00FF00FF00FF00FFull,
                                0x0000FFFF0000FFFFull};
  constexpr int w = 1 << K;
  uint64_t y;
  if (K < 4) {
    const uint32_t lo = row_xor<w>(static_cast<uint32_t>(x));
    const uint32_t hi = row_xor<w>(static_cast<uint32_t>(x >> 32));
    y = (static_cast<uint64_t>(hi) << 32) | lo;
  } else {
    y = __shfl_xor(x, w);
  }
  const int sh = (lane & w) ? w : 0;
  x = ((x >> sh) & kLow[K]) + ((y >> sh) & kLow[K]);
}

// Column counts of a 64 x 64 bit matrix held one row per lane: on exit
// lane r holds the number of rows with bit r set.  A butterfly over the lane
// bits: at stage k, lanes differing in bit k exchange their words, each keeps
// the fields (of width 2^k) whose column has its own bit k and adds the
// partner's, so field widths double while the fields per lane halve.
//  - stages 0, 1: 1- and 2-bit fields are full, so each side is masked
//    before the add;
//  - stage 2: nibbles hold at most 4, so one 64-bit add sums every nibble
//    pair without a carry, and the kept nibbles are masked after it;
//  - stage 3: bytes hold at most 8; after the add, v_perm_b32 gathers the
//    four kept bytes into 16-bit fields, columns 16j + (lane & 15) with
//    j = 0, 2 in `a` and j = 1, 3 in `b`;
//  - stage 4 (rows 2m <-> 2m+1): v_permlane16_swap trades the odd rows of
//    `a` for the even rows of `b`, so an even-row lane holds `a` of itself
//    and its partner (j = 0, 2) and an odd-row lane `b` of both (j = 1, 3);
//  - stage 5 (lanes l <-> l^32): v_permlane32_swap gives every lane the
//    words of l and l^32; their sum holds the totals of columns
//    16j + (lane & 15) for j = (lane >> 4) & 1 (low half) and 2 + that
//    (high half), and the lane takes the half its bit 5 names, which is
//    column `lane`.
// No LDS traffic: every exchange is a DPP or permlane move in the VALU.
__device__ __forceinline__ uint32_t column_count(uint64_t x, int lane) {
  count_stage<0>(x, lane);
  count_stage<1>(x, lane);
  {
    const uint32_t ylo = row_xor<4>(static_cast<uint32_t>(x));
    const uint32_t yhi = row_xor<4>(static_cast<uint32_t>(x >> 32));
    const uint64_t t = x + ((static_cast<uint64_t>(yhi) << 32) | ylo);
    x = (t >> ((lane & 4) ? 4 : 0)) & 0x0F0F0F0F0F0F0F0Full;
  }
  const uint32_t ylo = row_xor<8>(static_cast<uint32_t>(x));
  const uint32_t yhi = row_xor<8>(static_cast<uint32_t>(x >> 32));
  const uint64_t t = x + ((static_cast<uint64_t>(yhi) << 32) | ylo);
  const uint32_t b3 = (static_cast<uint32_t>(lane) >> 3) & 1u;
  // selector bytes 0-3 pick bytes of the low word, 4-7 of the high word,
  // 0x0C a zero byte
  const uint32_t a = __builtin_amdgcn_perm(static_cast<uint32_t>(t >> 32), static_cast<uint32_t>(t),
                                           0x0C000C00u | ((4u + b3) << 16) | b3);
  const uint32_t b = __builtin_amdgcn_perm(static_cast<uint32_t>(t >> 32), static_cast<uint32_t>(t),
                                           0x0C000C00u | ((6u + b3) << 16) | (2u + b3));
  const auto r16 = __builtin_amdgcn_permlane16_swap(a, b, false, false);
  const uint32_t s = static_cast<uint32_t>(r16[0]) + static_cast<uint32_t>(r16[1]);
  const auto r32 = __builtin_amdgcn_permlane32_swap(s, s, false, false);
  const uint32_t u = static_cast<uint32_t>(r32[0]) + static_cast<uint32_t>(r32[1]);
  return __builtin_amdgcn_ubfe(u, (lane & 32) ? 16u : 0u, 16u);
}

// Word of lane (lane ^ 7) within an 8-lane half-row (DPP row_half_mirror).
__device__ __forceinline__ uint32_t half_row_mirror(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x141, 0xF, 0xF, true));
}

// count_merge4's lane l holds column l ^ (3 * bit 2 of l): lanes with bit 2
// set (DPP banks 1 and 3 of each row) take the value of lane l ^ 3 (quad
// reversal); the others keep their own.  Linear, so sums of merged counts
// are fixed once.
__device__ __forceinline__ uint32_t count_fix(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(static_cast<int>(v), static_cast<int>(v), 0x1B, 0xF, 0xA, false));
}

// The same column count split for groups of four words, so that the last
// two stages, where fields have room to spare, serve four words at once.
//
// count_rows16 runs stages 0-3 of one word's block.  Stages 0 and 1 are
// written so that the exchanged value feeds a 32-bit add: each lane sends
// the fields its partner keeps (x >> (w - sh), masked) and keeps its own
// (x >> sh, masked), so the DPP move folds into the add's operand.  Stage 0's
// mask m0 is 0x55555555 for rows that count and 0 for rows that do not (one
// mask per chunk instead of a select per word).  The 64-bit sums are two
// 32-bit adds: the halves split at a field boundary and no field overflows.
//
// Stage 2 pairs lane l with l ^ 7 (the half-row mirror, one DPP operand)
// rather than l ^ 4, which DPP reaches only in two moves.  For the two to
// hold the same columns, stages 0 and 1 keep the columns whose bit 0 is
// bit 0 ^ bit 2 of the lane, and whose bit 1 is bit 1 ^ bit 2 (their
// partners l ^ 1, l ^ 2 keep the others, as before).  The result holds
// four 8-bit counts of the lane's 16-lane row (sel3 = 0x06040200 +
// 0x01010101 * bit 3 of the lane picks the kept bytes of stage 3, at most
// 16 each); after count_merge4 lane l holds column l ^ (3 * bit 2 of l),
// which count_fix moves to lane l.
__device__ __forceinline__ uint32_t count_rows16(uint64_t x, uint32_t m0, uint32_t sel3,
                                                 int lane) {
  {  // stage 0: 1-bit fields
    const uint32_t sh = static_cast<uint32_t>(lane ^ (lane >> 2)) & 1u;
    const uint64_t keep = x >> sh, send = x >> (1u - sh);
    const uint32_t lo = row_xor<1>(static_cast<uint32_t>(send) & m0) +
                        (static_cast<uint32_t>(keep) & m0);
    const uint32_t hi = row_xor<1>(static_cast<uint32_t>(send >> 32) & m0) +
                        (static_cast<uint32_t>(keep >> 32) & m0);
    x = (static_cast<uint64_t>(hi) << 32) | lo;
  }
  {  // stage 1: 2-bit fields holding at most 2
    constexpr uint32_t m = 0x33333333u;
    const uint32_t sh = static_cast<uint32_t>(lane ^ (lane >> 1)) & 2u;
    const uint64_t keep = x >> sh, send = x >> (2u - sh);
    const uint32_t lo = row_xor<2>(static_cast<uint32_t>(send) & m) +
                        (static_cast<uint32_t>(keep) & m);
    const uint32_t hi = row_xor<2>(static_cast<uint32_t>(send >> 32) & m) +
                        (static_cast<uint32_t>(keep >> 32) & m);
    x = (static_cast<uint64_t>(hi) << 32) | lo;
  }
  {  // stage 2: nibbles hold at most 4, so add first and keep after
    const uint32_t lo = static_cast<uint32_t>(x), hi = static_cast<uint32_t>(x >> 32);
    const uint64_t t =
        (static_cast<uint64_t>(half_row_mirror(hi) + hi) << 32) | (half_row_mirror(lo) + lo);
    x = (t >> (static_cast<uint32_t>(lane) & 4u)) & 0x0F0F0F0F0F0F0F0Full;
  }
  // stage 3: bytes hold at most 8
  const uint32_t lo = static_cast<uint32_t>(x), hi = static_cast<uint32_t>(x >> 32);
  return __builtin_amdgcn_perm(row_xor<8>(hi) + hi, row_xor<8>(lo) + lo, sel3);
}

// Stages 4 and 5 for four words A..D given their count_rows16 results: byte q
// of the result is the count of column `lane` of word q over all 64 rows.
//  - stage 4 (lane ^ 16), per pair of words: P = [A0, A2, B0, B2] and
//    Q = [A1, A3, B1, B3] (byte j of A is Aj); v_permlane16_swap trades the
//    odd rows of P for the even rows of Q, so the sum holds
//    [A_b, A_b+2, B_b, B_b+2] for b = bit 4 of the lane, at most 32 each;
//  - stage 5 (lane ^ 32): U = [A_b, B_b, C_b, D_b], V = [A_b+2, .., D_b+2];
//    v_permlane32_swap trades the upper half of U for the lower half of V,
//    so the sum holds index b + 2 * (bit 5 of the lane), i.e. column lane,
//    of every word, at most 64 each.
__device__ __forceinline__ uint32_t count_merge4(uint32_t ra, uint32_t rb, uint32_t rc,
                                                 uint32_t rd) {
  const auto ab = __builtin_amdgcn_permlane16_swap(__builtin_amdgcn_perm(rb, ra, 0x06040200u),
                                                   __builtin_amdgcn_perm(rb, ra, 0x07050301u),
                                                   false, false);
  const auto cd = __builtin_amdgcn_permlane16_swap(__builtin_amdgcn_perm(rd, rc, 0x06040200u),
                                                   __builtin_amdgcn_perm(rd, rc, 0x07050301u),
                                                   false, false);
  const uint32_t tab = static_cast<uint32_t>(ab[0]) + static_cast<uint32_t>(ab[1]);
  const uint32_t tcd = static_cast<uint32_t>(cd[0]) + static_cast<uint32_t>(cd[1]);
  const auto r = __builtin_amdgcn_permlane32_swap(__builtin_amdgcn_perm(tcd, tab, 0x06040200u),
                                                  __builtin_amdgcn_perm(tcd, tab, 0x07050301u),
                                                  false, false);
  return static_cast<uint32_t>(r[0]) + static_cast<uint32_t>(r[1]);
}

// Lane j's 64-bit word, broadcast to the wavefront (wave-uniform).
__device__ __forceinline__ uint64_t lane_word(uint64_t m, int j) {
  const uint32_t lo = static_cast<uint32_t>(
      __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(m)), j));
  const uint32_t hi = static_cast<uint32_t>(
      __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(m >> 32)), j));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Exposure sums of the words q0 .. q0+nq-1 (nq <= WT) of a node over its
// entries e0 .. e1 for one window.
// The wavefront takes 64 contact entries at a time (one per lane, coalesced)
// together with the neighbours' infectious words of the range, then takes
// the entries in order, broadcasting each one's overlap and words from its
// lane.  en0, en1 are the row's first two chunks of entries, loaded by the
// caller together with the node's other inputs.  No load sits under an exec
// mask: lanes without a counting contact read row col0, a row of the node's
// own network (its words are at wsh + col0 * W like every neighbour's), and
// are masked after the loads.  Integer mode adds overlaps into exact integer sums;
// float mode merges the contributions fl(ov * scale) of one column in rank
// order and adds columns in ascending order (cu:299-309, Eigen
// SparseDenseProduct.h:64-70), visiting only entries that overlap the window
// with a nonzero word: a column's entries share one neighbour, so a column is
// skipped whole or not at all, and adding nothing leaves the fp32 sum as it is.
//
// The sums start from iacc (integer mode) or facc (float mode) as given.
// kStop (integer mode): the walk stops before the first step (one or two
// chunks) after which a susceptible replica's sum reaches 2^24; iacc then
// holds the sums before that step, and the step's first entry is returned
// (e1 when no step reaches it).
//
// Groups of a multiple of four words (integer mode, no kStop) take the
// split column count (count_rows16 per word, count_merge4 per four words).
// A 4-word group keeps the whole-window counts of the walk in 16-bit fields,
// two words per register, multiplied by the window length once at the end;
// a walk covers at most max(kLongRow, kHubSeg) entries, so no field
// overflows.  Wider groups add each merged count times the window length
// into iacc at once.
template <bool kIntMode, int WT, bool kFull, typename EnT, bool kStop = false, bool kLive = false>
__device__ __forceinline__ int group_sums(const StepArgs &a, int64_t wsh, int W, int col0, int e0,
                                          int e1, uint32_t rank_limit, int q0, int nq, int lane,
                                          const uint64_t (&ss)[WT], const EnT &en0,
                                          const EnT &en1, uint32_t (&iacc)[WT],
                                          float (&facc)[WT]) {
  static_assert(kIntMode || !kStop, "the stop test reads integer sums");
  constexpr bool kPacked = kIntMode && !kStop && WT % 4 == 0;
  // 4-word groups: whole-window counts of words (0, 1) and (2, 3), 16 bits
  // each (wider groups add each merged count into iacc, which takes fewer
  // registers than 16-bit fields for all their words)
  uint32_t cnt[2] = {0u, 0u};
  const uint32_t sel3 = 0x06040200u + 0x01010101u * ((static_cast<uint32_t>(lane) >> 3) & 1u);
  // float mode: the column being merged, its weight and its words
  int cur_col = -1;
  float cur_w = 0.0f;
  bool cur_any = false;
  uint32_t cur_lo[WT], cur_hi[WT];
#pragma unroll
  for (int q = 0; q < WT; ++q) cur_lo[q] = cur_hi[q] = 0;
  // overlap of this lane's entry of chunk c0 (0 when it does not count) and
  // the column whose words it reads (col0 when it does not count)
  auto decode = [&](int c0, const EnT &raw, int &col, int &ov) {
    const EntryFields en = fields(raw);
    int o = min(en.te, a.we_rel) - max(en.ts, a.ws_rel);
    if (!(c0 + lane < e1 && o > 0 && en.rank <= rank_limit)) o = 0;
    ov = o;
    col = o > 0 ? en.col : col0;
  };
  // kFull: every network's W is a multiple of WT, so every group is whole
  // and, WT being even, the word offset wsh + col * W + q0 is even (every
  // network's words start at an even word) and 16-byte aligned.  Such a
  // group is read two words per load: a divergent gather costs the vector
  // L1 one tag lookup per lane and load, so 16-byte loads halve the lookups
  // of 8-byte ones for the same bytes.  The neighbour col belongs to the
  // node's own network, so its words are at wsh + col * W.
  auto fetch = [&](int col, uint64_t (&raw)[WT]) {
    const uint64_t *mrow = a.I_cur + (wsh + static_cast<int64_t>(col) * W + q0);
    if (kFull && WT % 2 == 0) {
      const ulonglong2 *m2 = reinterpret_cast<const ulonglong2 *>(mrow);
#pragma unroll
      for (int h = 0; h < WT / 2; ++h) {
        const ulonglong2 v = m2[h];
        raw[2 * h] = v.x;
        raw[2 * h + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int q = 0; q < WT; ++q) raw[q] = mrow[q < nq ? q : nq - 1];
    }
  };
  // float mode: the words of entries that count, zero elsewhere
  auto mask = [&](int ov, const uint64_t (&raw)[WT], uint32_t (&lo)[WT], uint32_t (&hi)[WT]) {
#pragma unroll
    for (int q = 0; q < WT; ++q) {
      const uint64_t m = (ov > 0 && q < nq && ss[q] != 0ull) ? raw[q] : 0ull;
      lo[q] = static_cast<uint32_t>(m);
      hi[q] = static_cast<uint32_t>(m >> 32);
    }
  };
  // one chunk of 64 entries: this lane's entry has column col, overlap ov
  // (0 where the entry does not count) and gathered words raw
  auto reduce = [&](int col, int ov, const uint64_t (&raw)[WT]) {
    if (kIntMode) {
      // Whole-window contacts: count, per replica lane, the contacts whose
      // neighbour word has that replica's bit set (column counts of the
      // 64 x 64 block of gathered words); partial overlaps are added one
      // contact at a time from the broadcast word.  The words are not
      // zeroed where the entry does not count: every lane read below is in
      // `counting` (ov > 0), and a whole-window lane has ov == L > 0.
      const int L = a.we_rel - a.ws_rel;
      const bool full = L > 0 && ov == L;
      const uint64_t counting = __ballot(ov > 0);
      const uint64_t partial = __ballot(ov > 0 && ov != L);
      const uint32_t m0 = full ? 0x55555555u : 0u;
      // Word q: `count` (the column count) when enough whole-window contacts
      // contribute, then the other contacts one at a time.  The word of
      // contact j is broadcast and read as a lane mask (inverse ballot):
      // replica lane r tests bit r in one select.
      auto word = [&](int q, uint64_t m, auto &&count) {
        const uint64_t nz = __ballot(m != 0ull) & counting;
        if (nz == 0ull) return;
        // few contributing contacts: add them one at a time
        const bool few = __popcll(nz & ~partial) <= kFewContacts;
        if (!few && (nz & ~partial) != 0ull) count();
        for (uint64_t pm = few ? nz : (partial & nz); pm != 0ull; pm &= pm - 1) {
          const int j = __builtin_ctzll(pm);
          const uint32_t ovj = static_cast<uint32_t>(__builtin_amdgcn_readlane(ov, j));
          iacc[q] += __builtin_amdgcn_inverse_ballot_w64(lane_word(m, j)) ? ovj : 0u;
        }
      };
      if constexpr (kPacked) {
        // four words at a time: stage-3 bytes of each counted word (zero
        // for the others), then the shared last stages
#pragma unroll
        for (int g = 0; g < WT / 4; ++g) {
          uint32_t r3[4];
          bool counted = false;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int q = 4 * g + k;
            r3[k] = 0u;
            // words without susceptible replicas need no exposure (kLive:
            // the caller found every word of the group live)
            if (!kLive && (q >= nq || ss[q] == 0ull)) continue;
            word(q, raw[q], [&] {
              r3[k] = count_rows16(raw[q], m0, sel3, lane);
              counted = true;
            });
          }
          if (!counted) continue;
          const uint32_t u = count_merge4(r3[0], r3[1], r3[2], r3[3]);
          if constexpr (WT == 4) {
            // bytes 0, 1 and 2, 3 of u into 16-bit fields (in count_merge4's
            // lane order, fixed once at the end)
            cnt[0] += __builtin_amdgcn_perm(0u, u, 0x0C010C00u);
            cnt[1] += __builtin_amdgcn_perm(0u, u, 0x0C030C02u);
          } else {
            const uint32_t uf = count_fix(u);
#pragma unroll
            for (int k = 0; k < 4; ++k)
              iacc[4 * g + k] += __umul24(static_cast<uint32_t>(L),
                                          __builtin_amdgcn_ubfe(uf, 8u * k, 8u));
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < WT; ++q) {
          // words without susceptible replicas need no exposure
          if (!kLive && (q >= nq || ss[q] == 0ull)) continue;
          const uint64_t m = raw[q];
          word(q, m, [&] {
            iacc[q] += static_cast<uint32_t>(L) * column_count(full ? m : 0ull, lane);
          });
        }
      }
    } else {
      uint32_t mlo[WT], mhi[WT];
      mask(ov, raw, mlo, mhi);
      uint64_t any = 0ull;
#pragma unroll
      for (int q = 0; q < WT; ++q) any |= (static_cast<uint64_t>(mhi[q]) << 32) | mlo[q];
      for (uint64_t act = __ballot(any != 0ull); act != 0ull; act &= act - 1) {
        const int j = __builtin_ctzll(act);
        const int colj = __builtin_amdgcn_readlane(col, j);
        if (colj != cur_col) {  // flush the finished column
          if (cur_any) {
#pragma unroll
            for (int q = 0; q < WT; ++q)
              if (__builtin_amdgcn_inverse_ballot_w64((static_cast<uint64_t>(cur_hi[q]) << 32) |
                                                      cur_lo[q]))
                facc[q] += cur_w;
          }
          cur_col = colj;
          cur_w = 0.0f;
          cur_any = false;
        }
        const float wc = static_cast<float>(__builtin_amdgcn_readlane(ov, j)) * a.scale;
        cur_w = cur_any ? cur_w + wc : wc;
        cur_any = true;
#pragma unroll
        for (int q = 0; q < WT; ++q) {
          cur_lo[q] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mlo[q]), j));
          cur_hi[q] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mhi[q]), j));
        }
      }
    }
  };
  // Chunks are taken in pairs for WT <= 4 (one at a time for wider groups,
  // whose words take twice the registers): every neighbour-word load of the
  // pair is issued before the first is waited for, then the pair is reduced
  // and the next pair's entries are loaded.
  constexpr bool kPair = WT <= 4;
  // the chunks come from the copy of the entries whose form EnT is
  const auto *src = [&] {
    if constexpr (sizeof(EnT) == 8)
      return a.entries8;
    else
      return a.entries;
  }();
  EnT enA = en0, enB = en1;
  for (int c0 = e0; c0 < e1;) {
    int colA, ovA, colB = 0, ovB = 0;
    uint64_t rawA[WT], rawB[WT];
    decode(c0, enA, colA, ovA);
    fetch(colA, rawA);
    const bool two = kPair && c0 + kWave < e1;
    if (two) {
      decode(c0 + kWave, enB, colB, ovB);
      fetch(colB, rawB);
    }
    uint32_t before[WT];
    if constexpr (kStop) {
#pragma unroll
      for (int q = 0; q < WT; ++q) before[q] = iacc[q];
    }
    reduce(colA, ovA, rawA);
    if (two) reduce(colB, ovB, rawB);
    if constexpr (kStop) {
      bool over = false;
#pragma unroll
      for (int q = 0; q < WT; ++q)
        over |= q < nq && __builtin_amdgcn_inverse_ballot_w64(ss[q]) && iacc[q] >= (1u << 24);
      if (__ballot(over) != 0ull) {
#pragma unroll
        for (int q = 0; q < WT; ++q) iacc[q] = before[q];
        return c0;
      }
    }
    c0 += kPair ? 2 * kWave : kWave;
    if (c0 < e1) {
      enA = load_chunk(src, c0, e1, lane);
      if (kPair && c0 + kWave < e1) enB = load_chunk(src, c0 + kWave, e1, lane);
    }
  }
  if (!kIntMode && cur_any) {
#pragma unroll
    for (int q = 0; q < WT; ++q)
      if (__builtin_amdgcn_inverse_ballot_w64((static_cast<uint64_t>(cur_hi[q]) << 32) | cur_lo[q]))
        facc[q] += cur_w;
  }
  if constexpr (kPacked && WT == 4) {
    const uint32_t L = static_cast<uint32_t>(a.we_rel - a.ws_rel);
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const uint32_t c = count_fix(cnt[g]);
      iacc[2 * g] += L * (c & 0xFFFFu);
      iacc[2 * g + 1] += L * (c >> 16);
    }
  }
  return e1;
}

// Exposure of the group's words from zero sums (see group_sums).
template <bool kIntMode, int WT, bool kFull, typename EnT, bool kLive = false>
__device__ __forceinline__ void group_exposure(const StepArgs &a, int64_t wsh, int W, int col0,
                                               int e0, int e1, uint32_t rank_limit, int q0,
                                               int nq, int lane, const uint64_t (&ss)[WT],
                                               const EnT &en0, const EnT &en1, float (&E)[WT]) {
  uint32_t iacc[WT];
  float facc[WT];
#pragma unroll
  for (int q = 0; q < WT; ++q) {
    iacc[q] = 0;
    facc[q] = 0.0f;
  }
  group_sums<kIntMode, WT, kFull, EnT, false, kLive>(a, wsh, W, col0, e0, e1, rank_limit, q0, nq,
                                                    lane, ss, en0, en1, iacc, facc);
#pragma unroll
  for (int q = 0; q < WT; ++q) E[q] = kIntMode ? static_cast<float>(iacc[q]) : facc[q];
}

// Whether rows e0 .. e1 of a window of integer weights (divisor 1) may be
// summed in integers first: every overlap below 2^24 (so the reference's
// (float)overlap is exact) and no 32-bit sum overflow.  Weights being
// non-negative, while a replica's sum stays below 2^24 every partial sum of
// the reference's fp32 order is an integer that fp32 holds exactly, so the
// integer sum is that order's result; the caller falls back to the
// reference order when a susceptible replica's sum reaches 2^24.
constexpr int kTryIntMinEntries = 32;

__device__ __forceinline__ bool try_integer_sums(const StepArgs &a, int e0, int e1) {
  const int64_t L = static_cast<int64_t>(a.we_rel) - a.ws_rel;
  return L < (int64_t(1) << 24) && static_cast<int64_t>(e1 - e0) * L < (int64_t(1) << 32);
}

// Timer flips of word w for iteration k+1 (see the header comment): x is
// the word's infections of iteration k, inf and sus its infectious and
// susceptible words with the ring's flips at offsets n >= 2 applied.
__device__ __forceinline__ void apply_flips(const StepArgs &a, const NetDev &nd, int64_t w,
                                            uint64_t x, uint64_t inf, uint64_t sus) {
  const int32_t k1 = a.k + 1;
  if (nd.n_on == 1) inf ^= x;
  if (nd.n_off == 1) inf ^= x;
  if (nd.n_s == 1) sus |= x;
  const bool fl_on = nd.m_on != kNever && nd.m_on >= 1 && k1 == nd.m_on;
  const bool fl_off = nd.m_off != kNever && nd.m_off >= 1 && k1 == nd.m_off;
  const bool fl_s = nd.m_s != kNever && nd.m_s >= 1 && k1 == nd.m_s;
  if (fl_on || fl_off || fl_s) {
    const uint64_t xi = a.X_init[w];
    if (fl_on) inf ^= xi;
    if (fl_off) inf ^= xi;
    if (fl_s) sus |= xi;
  }
  a.I_next[w] = inf;
  a.S[w] = sus;
}

// Draw of one word (word Q of node ii) given its lanes' exposure; records the
// infections and returns their ballot.  The probability is computed, not
// looked up: a table read is one more dependent memory round trip per node,
// which costs more than the arithmetic.  The Philox block serving words
// 4g .. 4g+3 is computed once per lane and kept in pw / pw_grp.
template <bool kInjected>
__device__ __forceinline__ uint64_t draw_word(const StepArgs &a, const NetDev &nd, int64_t wi,
                                              int ii, int Q, int lane, uint64_t ss, float E,
                                              uint32_t (&pw)[4], int &pw_grp) {
  const int r = Q * kWave + lane;
  // ss is wave-uniform: read as a lane mask, replica lane r tests bit r
  const bool sus = __builtin_amdgcn_inverse_ballot_w64(ss);
  // E == 0 gives P = 0 (or NaN), never above a uniform in [0, 1)
  const bool need = sus && (kInjected || E > 0.0f);
  // The block serving words 4g .. 4g+3 is computed once per wavefront, by
  // every lane together, when the first word of the group that some lane
  // draws comes up (a uniform branch; pw_grp is the wavefront's).  Counter
  // words 0, 2, 3 are the wavefront's (node, slot, stream): made opaque,
  // the products of the first round are formed here by the scalar unit
  // rather than hoisted to the kernel's start and held in VGPR lanes.
  if (!kInjected && pw_grp != (Q >> 2) && __ballot(need) != 0ull) {
    uint32_t c0 = static_cast<uint32_t>(ii - nd.node_off), c2 = a.slot, c3 = nd.stream;
    asm volatile("" : "+s"(c0), "+s"(c2), "+s"(c3));
    rp::philox4(a.seed, c0, (static_cast<uint32_t>(Q >> 2) << 6) | static_cast<uint32_t>(lane),
                c2, c3, pw);
    pw_grp = Q >> 2;
  }
  bool infect = false;
  if (need) {
    float u;
    if (kInjected) {
      u = a.uniforms[static_cast<size_t>(r) * a.N + ii];
    } else {
      const int k = Q & 3;
      u = rp::u24(k == 0 ? pw[0] : k == 1 ? pw[1] : k == 2 ? pw[2] : pw[3]);
    }
    // The reference's P = 1 - expf(log_base * (E / 3600)) is needed only to
    // compare it with u.  An estimate from the hardware base-2 exponential
    // (v_exp_f32, within 1 ulp) of fl(log_base * fl(E * fl(1/3600))) differs
    // from it by less than 5e-7: the relative error of the exponent, a few
    // ulp, moves e^x by at most |x| e^x <= 1/e times that, and the two
    // roundings of 1 - e^x add 1.2e-7.  Outside a band of kDecideMargin
    // (20 times that bound) around u the estimate's comparison is the
    // reference's; inside it, or when the estimate is NaN, the reference
    // expression is evaluated.
    const float x_est = nd.log_base * (E * (1.0f / 3600.0f));
    const float P_est = 1.0f - __builtin_amdgcn_exp2f(x_est * 1.44269504088896341f);
    if (fabsf(P_est - u) > kDecideMargin)
      infect = P_est > u;
    else
      infect = 1.0f - rp::rp_expf(nd.log_base * (E / 3600.0f)) > u;
  }
  if (infect && a.record) {
    // entry is laid out [word][64 replicas]; wi is the node's first word
    a.entry[(static_cast<size_t>(wi) + Q) * kWave + lane] = 2 * a.k + 1;
    atomicAdd(a.inf_count + static_cast<size_t>(a.k) * a.G + nd.moff + r, 1);
  }
  return __ballot(infect);
}

// Phase 1 of an iteration: one wavefront per node computes the exposure of
// its susceptible replicas, draws, and writes the node's infection words
// (zero words included) into ring slot k.
//
// Occupancy floors, measured: 8 wavefronts per SIMD for groups of up to 4
// words (SGPR values parked in VGPR lanes, no scratch); 6 for whole 16-word
// groups (a few VGPRs to scratch, outside the chunk loop: -2.3% per step on
// the dense workload); the compiler's choice otherwise.
//
// kTryInt (with kIntMode false: a window of integer weights, divisor 1,
// whose short-row weight bound reaches 2^24): each group is summed in
// integers first and in the reference order only when a susceptible
// replica's sum reaches 2^24 (see try_integer_sums).
template <bool kIntMode, bool kInjected, int WT, bool kFull, bool kCompact, bool kTryInt>
__global__ __launch_bounds__(kBlock) __attribute__((
    amdgpu_waves_per_eu(WT <= 4 ? 8 : (WT == 16 && kFull && !kInjected ? 6 : 1))))
void expo_kernel(StepArgs a) {
  // kCompact: the chunks are read from the 8-byte copy of the entries
  using EnT = std::conditional_t<kCompact, v2u, v4i>;
  const int lane = threadIdx.x & (kWave - 1);
  // Workgroups are dealt to the 8 XCDs round-robin.  Block b takes the
  // b-th run of nodes: the XCDs walk the node range together, each taking
  // every eighth block, so their work stays even however it varies along
  // the node range and over a run (measured: XCDs given fixed contiguous
  // eighths of 64 networks ended a launch up to 20% apart), and the
  // networks in flight are the same few on every XCD.  With xcd_split
  // (REPLAY_V_XCD) block b is remapped so that XCD x takes the x-th
  // contiguous eighth of the nodes instead: where contacts stay within
  // short runs of node ids (villages), a neighbour word is then fetched
  // into one XCD's L2 rather than into all of them.  gridDim.x is a
  // multiple of 8.  The choice is a mask, not a branch, so that the flag,
  // the grid size and the block size are requested together (a branch put
  // two more rounds of scalar loads before the node's first load: 1.2% per
  // iteration on 10,000 short rows).
  const unsigned split = 0u - static_cast<unsigned>(a.xcd_split != 0);
  const unsigned remap = (blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds;
  const unsigned lb = (remap & split) | (blockIdx.x & ~split);
  // blocks of four wavefronts, or of one (REPLAY_V_BLOCK1)
  const int i = static_cast<int>(lb) * static_cast<int>(blockDim.x / kWave) +
                __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (i >= a.N) return;
  // The node's inputs are requested together before any is waited for: the
  // row bounds and network (scalar), then the network's constants and the
  // node's first susceptible words and, once the bounds are in, the first
  // two chunks of contact entries.
  // the row bounds and network come in one 16-byte scalar load: read from
  // two arrays, the network's load went out a round after the bounds'
  // (config3 84.9 -> 83.4 ms per run, village 68.6 -> 67.7 ms)
  const int4 ri = a.rowinfo[i];
  const int e0 = ri.x, e1 = ri.y;
  const int net = ri.z;
  // the node's first word: i * W when every network has the same replica
  // count (a.M != 0: no table read before the state load), else its entry
  // in the word table
  const int64_t wi = a.M != 0 ? static_cast<int64_t>(i) * a.W : a.node_word[i];
  // lanes past the node's words read the next node's (masked below), never
  // past the last word
  const uint64_t s_first = a.S[min(wi + min(lane, WT - 1), a.NW - 1)];
  const NetDev nd = a.nets[net];
  const int W = a.M != 0 ? a.W : nd.W;
  uint64_t *X_k = a.X + a.xk_off + wi;
  const uint64_t *S = a.S + wi;
  // 4-word groups take the row exits after the node's loads are requested
  // (the compiler would otherwise defer those loads below the exits):
  // config3 83.4 -> 82.9 ms per run; with 16-word groups, where rows
  // without a contact are common (the village's), the exits stay first
  constexpr bool kLateExits = WT == 4;
  if constexpr (!kLateExits) {
    if (e1 - e0 > a.long_row) return;  // long_row_kernel takes it
    if (!kInjected && e1 <= e0) {
      // no contact in the epoch: no exposure, so no infection (a draw
      // needs E > 0); the node's words of ring slot k are zero
      for (int w = lane; w < W; w += kWave) X_k[w] = 0ull;
      return;
    }
  }
  EnT en0{}, en1{};
  if (e1 > e0) {
    if constexpr (kCompact) {
      en0 = load_chunk(a.entries8, e0, e1, lane);
      if (WT <= 4 && e1 - e0 > kWave) en1 = load_chunk(a.entries8, e0 + kWave, e1, lane);
    } else {
      en0 = load_chunk(a.entries, e0, e1, lane);
      if (WT <= 4 && e1 - e0 > kWave) en1 = load_chunk(a.entries, e0 + kWave, e1, lane);
    }
  }
  if constexpr (kLateExits) {
    if (e1 - e0 > a.long_row) return;  // long_row_kernel takes it
    if (!kInjected && e1 <= e0) {
      for (int w = lane; w < W; w += kWave) X_k[w] = 0ull;
      return;
    }
  }
  const uint32_t rlim = a.rank_limit ? a.rank_limit[net] : UINT32_MAX;
  for (int q0 = 0; q0 < W; q0 += WT) {
    const int nq = kFull ? WT : min(WT, W - q0);
    // lane q holds susceptible word q0 + q (one load for the group)
    const uint64_t sv = q0 == 0 ? s_first : S[q0 + min(lane, nq - 1)];
    uint64_t ss[WT];
    uint64_t any_sus = 0ull;
#pragma unroll
    for (int q = 0; q < WT; ++q) {
      const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(sv), q));
      const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(sv >> 32), q));
      ss[q] = q < nq ? (static_cast<uint64_t>(hi) << 32) | lo : 0ull;
      any_sus |= ss[q];
    }
    uint64_t x[WT];
#pragma unroll
    for (int q = 0; q < WT; ++q) x[q] = 0ull;
    if (any_sus != 0ull && (e1 > e0 || kInjected)) {
      float E[WT];
      bool done = false;
      if constexpr (kTryInt) {
        // rows of a few contacts keep the float path, which visits only the
        // active entries and costs less there than a column count per word
        if (e1 - e0 > kTryIntMinEntries && try_integer_sums(a, e0, e1)) {
          group_exposure<true, WT, kFull, EnT>(a, nd.wsh, W, nd.node_off, e0, e1, rlim, q0, nq,
                                               lane, ss, en0, en1, E);
          bool big = false;
#pragma unroll
          for (int q = 0; q < WT; ++q)
            big |= __builtin_amdgcn_inverse_ballot_w64(ss[q]) && E[q] >= 16777216.0f;
          done = __ballot(big) == 0ull;
        }
      }
      if (!done) {
        // integer sums over whole 16-word groups in which every word has a
        // susceptible replica (the common case) skip the per-chunk test of
        // each word.  Measured: village 72.8 -> 68.9 ms and dense 819 ->
        // 790 ms per run (the 16-word instance also stops spilling VGPRs);
        // for 4-word groups the second copy of the walk cost 0.4%.
        constexpr bool kSpecial = kIntMode && kFull && WT == 16;
        bool live = kSpecial;
#pragma unroll
        for (int q = 0; q < WT; ++q) live &= ss[q] != 0ull;
        if (live)
          group_exposure<kIntMode, WT, kFull, EnT, kSpecial>(
              a, nd.wsh, W, nd.node_off, e0, e1, rlim, q0, nq, lane, ss, en0, en1, E);
        else
          group_exposure<kIntMode, WT, kFull, EnT>(a, nd.wsh, W, nd.node_off, e0, e1, rlim, q0, nq,
                                                   lane, ss, en0, en1, E);
      }
      uint32_t pw[4] = {0u, 0u, 0u, 0u};
      int pw_grp = -1;
#pragma unroll
      for (int q = 0; q < WT; ++q)
        if (q < nq && ss[q] != 0ull)
          x[q] = draw_word<kInjected>(a, nd, wi, i, q0 + q, lane, ss[q], E[q], pw, pw_grp);
    }
    // lane q stores word q of the group
    uint64_t mine = 0ull;
#pragma unroll
    for (int q = 0; q < WT; ++q)
      if (lane == q) mine = x[q];
    if (lane < nq) X_k[q0 + lane] = mine;
  }
}

// Phase 1 for long rows (more than long_row entries in the epoch, e.g. the
// hubs of a power-law network): one wavefront per (listed row, 64-replica
// word), so a hub's contact list is walked by W wavefronts in parallel
// instead of one.  Integer mode where the row's own weight bound is below
// 2^24; otherwise the float path, which follows the reference order.
// kTryInt (rows of a window whose weights are integers, divisor 1, but whose
// bound reaches 2^24): the row is first summed in integers.  Weights are
// non-negative, so while a replica's sum stays below 2^24 every partial sum
// of the reference's fp32 order is an integer that fp32 holds exactly, and
// the integer sum is that order's result.  Only when a susceptible
// replica's sum reaches 2^24, or the row could overflow 32 bits, is it
// summed again in the reference order.
template <bool kIntMode, bool kInjected, bool kTryInt>
__global__ __launch_bounds__(kBlock) void long_row_kernel(StepArgs a, const int32_t *rows,
                                                          int n_rows) {
  // a.W (the most words per node of any network) wavefronts per listed row;
  // those past the row's own network's W have nothing to do
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = static_cast<int>(blockIdx.x) * kWavesPerBlock +
                  __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (wid >= n_rows * a.W) return;
  const int h = wid / a.W;
  const int q = wid - h * a.W;
  const int i = rows[h];
  const int net = a.net_of[i];
  const NetDev nd = a.nets[net];
  if (q >= nd.W) return;
  const int64_t wi = nd.wsh + static_cast<int64_t>(i) * nd.W;
  uint64_t *X_k = a.X + a.xk_off + wi;
  const uint64_t ss[1] = {a.S[wi + q]};
  uint64_t x = 0ull;
  if (ss[0] != 0ull) {
    const int e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    const uint32_t rlim = a.rank_limit ? a.rank_limit[net] : UINT32_MAX;
    float E[1];
    const v4i en0 = load_chunk(a.entries, e0, e1, lane);  // rows here hold > 2 chunks
    const v4i en1 = load_chunk(a.entries, e0 + kWave, e1, lane);
    bool done = false;
    if (kTryInt && try_integer_sums(a, e0, e1)) {
      group_exposure<true, 1, false, v4i>(a, nd.wsh, nd.W, nd.node_off, e0, e1, rlim, q, 1, lane,
                                          ss, en0, en1, E);
      // E is the integer sum converted to float: >= 2^24 exactly when the
      // sum is
      const bool sus = __builtin_amdgcn_inverse_ballot_w64(ss[0]);
      done = __ballot(sus && E[0] >= 16777216.0f) == 0ull;
    }
    if (!done)
      group_exposure<kIntMode, 1, false, v4i>(a, nd.wsh, nd.W, nd.node_off, e0, e1, rlim, q, 1,
                                              lane, ss, en0, en1, E);
    uint32_t pw[4] = {0u, 0u, 0u, 0u};
    int pw_grp = -1;
    x = draw_word<kInjected>(a, nd, wi, i, q, lane, ss[0], E[0], pw, pw_grp);
  }
  if (lane == 0) X_k[q] = x;
}

// Long rows in windows of integer weights (divisor 1), summed in segments.
// A long row of an epoch (a hub) is cut into segments of kHubSeg entries;
// hub_segment_kernel gives one wavefront to each (segment, word) and adds
// the segment's integer sums into the (row, word) totals, so a hub's
// contacts are walked by all its segments at once.  hub_finish_kernel then
// gives one wavefront to each (row, word).  The reference adds each
// replica's columns in one fp32 chain; with non-negative integer weights
// every partial sum of that chain is exact while the sum stays below 2^24,
// so for a replica whose total stays below 2^24 the integer total is the
// chain's result.  A replica whose total reaches 2^24 has a chain result of
// at least 2^24 (fp32 rounding is monotone), and when
// fl(log_base * fl(2^24 / 3600)) < -104 every such exposure gives
// rp_expf(...) = 0, so P = 1 > u: the draw infects whatever the exact sum.
// Otherwise the finisher finds the step of entries where a susceptible
// replica's sum first reaches 2^24, backs up to the first entry of the
// column the step starts in, and runs the reference's fp32 chain from there,
// starting from the exact integer sums before it.
#ifndef REPLAY_HUB_SEGMENT
#define REPLAY_HUB_SEGMENT 2048
#endif
// entries per segment (1024 and 4096 measured on the dense workload: 1,032
// and 1,062 ms per run against 1,036)
constexpr int kHubSeg = REPLAY_HUB_SEGMENT;
// group_sums keeps a 4-word group's counts of one walk in 16-bit fields
static_assert(rp::kLongRow < 65536 && kHubSeg < 65536, "a walk's counts must fit 16 bits");

struct HubArgs {
  const int32_t *rows;     // the epoch's long rows
  const int32_t *seg0;     // first segment of each long row (n_rows + 1 entries)
  const int32_t *seg_row;  // long row of each segment
  uint64_t *total;         // [n_rows][W][64] integer sums; the finisher reads and clears them
  uint32_t *part;          // [n_seg][W][64] sums of each segment, for the chain
  int32_t n_rows, n_seg;
};

// Whether every exposure of 2^24 s or more makes the draw infect (the
// exponent computed as draw_word does is then below -104, where rp_expf is
// 0); injected uniforms are not assumed to be below 1.
__device__ __forceinline__ bool saturation_infects(float log_base, bool injected) {
  return !injected && log_base * (16777216.0f / 3600.0f) < -104.0f;
}

// One wavefront per (segment, group of WT words); kFull: every network's W
// is a multiple of WT.
template <typename EnT, int WT, bool kFull>
__global__ __launch_bounds__(kBlock) void hub_segment_kernel(StepArgs a, HubArgs h,
                                                             int injected) {
  const int lane = threadIdx.x & (kWave - 1);
  // Block b takes the b-th run of (segment, word group) pairs, so the XCDs
  // (dealt blocks round-robin) walk the segments together and share their
  // work evenly.  Measured against keeping a segment's consecutive blocks
  // on one XCD, for the entries they share: 2.5% less per iteration on
  // 500 hubs of a 50,000-node network.
  const int wid = static_cast<int>(blockIdx.x) * kWavesPerBlock +
                  __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int groups = (a.W + WT - 1) / WT;
  if (wid >= h.n_seg * groups) return;
  const int sg = wid / groups;
  const int q0 = (wid - sg * groups) * WT;
  const int hr = h.seg_row[sg];
  const int i = h.rows[hr];
  const int net = a.net_of[i];
  const NetDev nd = a.nets[net];
  if (q0 >= nd.W) return;
  const int nq = kFull ? WT : min(WT, nd.W - q0);
  const int64_t wi = nd.wsh + static_cast<int64_t>(i) * nd.W;
  // lane q holds susceptible word q0 + q
  const uint64_t sv = a.S[wi + q0 + min(lane, nq - 1)];
  uint64_t ss[WT];
  uint64_t any_sus = 0ull;
#pragma unroll
  for (int q = 0; q < WT; ++q) {
    ss[q] = q < nq ? lane_word(sv, q) : 0ull;
    any_sus |= ss[q];
  }
  if (any_sus == 0ull) return;  // the finisher skips these words too
  const int e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
  const int c0 = e0 + (sg - h.seg0[hr]) * kHubSeg;
  const int c1 = min(c0 + kHubSeg, e1);
  const uint32_t rlim = a.rank_limit ? a.rank_limit[net] : UINT32_MAX;
  const auto *src = [&] {
    if constexpr (sizeof(EnT) == 8)
      return a.entries8;
    else
      return a.entries;
  }();
  const EnT en0 = load_chunk(src, c0, c1, lane);
  const EnT en1 = c1 - c0 > kWave ? load_chunk(src, c0 + kWave, c1, lane) : EnT{};
  uint32_t iacc[WT];
  float facc[WT];
#pragma unroll
  for (int q = 0; q < WT; ++q) {
    iacc[q] = 0u;
    facc[q] = 0.0f;
  }
  group_sums<true, WT, kFull, EnT>(a, nd.wsh, nd.W, nd.node_off, c0, c1, rlim, q0, nq, lane, ss,
                                   en0, en1, iacc, facc);
  const bool keep = !saturation_infects(nd.log_base, injected != 0);
#pragma unroll
  for (int q = 0; q < WT; ++q) {
    if (q >= nq || ss[q] == 0ull) continue;
    if (keep) h.part[(static_cast<size_t>(sg) * a.W + q0 + q) * kWave + lane] = iacc[q];
    if (iacc[q] != 0u)
      atomicAdd(reinterpret_cast<unsigned long long *>(
                    h.total + (static_cast<size_t>(hr) * a.W + q0 + q) * kWave + lane),
                static_cast<unsigned long long>(iacc[q]));
  }
}

// First entry of the run of row entries with the column of entry c
// (columns ascend within a row), and in part the overlaps of the run's
// counting entries before c; col is that column.  c == e0 or c == e1: c
// itself, nothing before it.
__device__ int column_run_start(const StepArgs &a, int e0, int e1, int c, uint32_t rlim,
                                int lane, uint32_t &part, int &col) {
  part = 0u;
  col = 0;
  if (c <= e0 || c >= e1) return c;
  col = a.entries[c].x;
  int p = c;
  while (p > e0) {
    const int idx = p - kWave + lane;
    int cl = -1, ov = 0;
    uint32_t rk = 0u;
    if (idx >= e0) {
      const EntryFields f = fields(reinterpret_cast<const v4i *>(a.entries)[idx]);
      cl = f.col;
      ov = min(f.te, a.we_rel) - max(f.ts, a.ws_rel);
      rk = f.rank;
    }
    // lanes before the row read nothing and count as another column
    const uint64_t other = __ballot(cl != col);
    const int hi = other != 0ull ? 63 - __builtin_clzll(other) : -1;
    uint32_t v = (lane > hi && ov > 0 && rk <= rlim) ? static_cast<uint32_t>(ov) : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    part += v;
    p = p - kWave + hi + 1;
    if (other != 0ull) break;
  }
  return p;
}

// Exposure of word q of long row hr in the reference's fp32 order, for the
// rows where a susceptible replica's integer sum reaches 2^24 and the draw
// is not settled by it.
__device__ float hub_chain(const StepArgs &a, const HubArgs &h, int hr, int i, const NetDev &nd,
                           uint32_t rlim, int q, int lane, uint64_t ssw) {
  const int e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
  const int s0 = h.seg0[hr], s1 = h.seg0[hr + 1];
  const bool sus = __builtin_amdgcn_inverse_ballot_w64(ssw);
  // the first segment after which a susceptible replica's sum reaches 2^24,
  // and the sums before it (below 2^24 for the susceptible replicas)
  uint64_t acc = 0ull;
  int sc = -1;
  for (int s = s0; s < s1; ++s) {
    const uint32_t v = h.part[(static_cast<size_t>(s) * a.W + q) * kWave + lane];
    if (__ballot(sus && acc + v >= (uint64_t(1) << 24)) != 0ull) {
      sc = s;
      break;
    }
    acc += v;
  }
  if (sc < 0) return static_cast<float>(acc);  // every sum exact
  const uint64_t ss[1] = {ssw};
  const int cs = e0 + (sc - s0) * kHubSeg, ce = min(cs + kHubSeg, e1);
  uint32_t iacc[1] = {static_cast<uint32_t>(acc)};
  float facc[1] = {0.0f};
  v4i en0 = load_chunk(a.entries, cs, ce, lane);
  v4i en1 = ce - cs > kWave ? load_chunk(a.entries, cs + kWave, ce, lane) : v4i{};
  const int c = group_sums<true, 1, false, v4i, true>(a, nd.wsh, nd.W, nd.node_off, cs, ce, rlim,
                                                      q, 1, lane, ss, en0, en1, iacc, facc);
  // every susceptible replica's sum before entry c is exact and below 2^24;
  // the chain restarts at the first entry of c's column, from the sums
  // before that column
  uint32_t part;
  int col;
  const int p = column_run_start(a, e0, e1, c, rlim, lane, part, col);
  uint32_t start = iacc[0];
  if (part != 0u) {
    const uint64_t m = a.I_cur[nd.wsh + static_cast<int64_t>(col) * nd.W + q];
    if ((m >> lane) & 1ull) start -= part;
  }
  facc[0] = static_cast<float>(start);
  iacc[0] = 0u;
  if (p < e1) {
    en0 = load_chunk(a.entries, p, e1, lane);
    en1 = e1 - p > kWave ? load_chunk(a.entries, p + kWave, e1, lane) : v4i{};
    group_sums<false, 1, false, v4i>(a, nd.wsh, nd.W, nd.node_off, p, e1, rlim, q, 1, lane, ss,
                                     en0, en1, iacc, facc);
  }
  return facc[0];
}

template <bool kInjected>
__global__ __launch_bounds__(kBlock) void hub_finish_kernel(StepArgs a, HubArgs h) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = static_cast<int>(blockIdx.x) * kWavesPerBlock +
                  __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (wid >= h.n_rows * a.W) return;
  const int hr = wid / a.W;
  const int q = wid - hr * a.W;
  const int i = h.rows[hr];
  const int net = a.net_of[i];
  const NetDev nd = a.nets[net];
  if (q >= nd.W) return;
  const int64_t wi = nd.wsh + static_cast<int64_t>(i) * nd.W;
  uint64_t *X_k = a.X + a.xk_off + wi;
  const uint64_t ss = a.S[wi + q];
  uint64_t x = 0ull;
  if (ss != 0ull) {
    uint64_t *tp = h.total + (static_cast<size_t>(hr) * a.W + q) * kWave + lane;
    const uint64_t tot = *tp;
    *tp = 0ull;  // ready for the next iteration
    const bool sus = __builtin_amdgcn_inverse_ballot_w64(ss);
    const bool sat = sus && tot >= (uint64_t(1) << 24);
    float E;
    if (__ballot(sat) == 0ull) {
      E = static_cast<float>(tot);  // exact below 2^24
    } else if (saturation_infects(nd.log_base, kInjected)) {
      E = sat ? 16777216.0f : static_cast<float>(tot);
    } else {
      const uint32_t rlim = a.rank_limit ? a.rank_limit[net] : UINT32_MAX;
      E = hub_chain(a, h, hr, i, nd, rlim, q, lane, ss);
    }
    uint32_t pw[4] = {0u, 0u, 0u, 0u};
    int pw_grp = -1;
    x = draw_word<kInjected>(a, nd, wi, i, q, lane, ss, E, pw, pw_grp);
  }
  if (lane == 0) X_k[q] = x;
}

// Carry-save adder of three 64-replica words (Harley and Seal): per replica
// bit, a + b + c = 2 * hi + lo.
__device__ __forceinline__ void csa(uint64_t &hi, uint64_t &lo, uint64_t a, uint64_t b,
                                    uint64_t c) {
  const uint64_t u = a ^ b;
  hi = (u & c) | (~u & a);  // majority of a, b, c
  lo = u ^ c;
}

constexpr int kPartialWords = 256;  // partial-overlap words recorded per wavefront
constexpr int kTabLds = 1024;       // probability-table entries staged in LDS per block

// Phase 1, bit-sliced form (integer windows only; contexts whose networks
// all have the same replica count, so node i's words start at i * W).  A
// wavefront covers
// G = 64 / W nodes with one lane per (node, 64-replica word).  Each lane walks
// its node's contact entries eight at a time, gathers the neighbour's word for
// every whole-window contact and adds the eight words into bit-sliced
// counters (plane b holds bit b of every replica's count) with a carry-save
// adder tree: about one 32-bit operation per replica word and contact, and no
// cross-lane traffic.  Contacts that overlap only part of the window keep
// their overlap and gathered words in LDS (every lane of a node walks the same
// entries, so all of them agree on the slot).  Then the wavefront takes its
// nodes one at a time with one lane per replica: it reads each replica's count
// out of the owning lane's planes, adds the recorded partial overlaps, and
// draws as expo_kernel does.  The exposure is the same exact integer,
// L * count + sum of partial overlaps.
template <int NB, bool kInjected>
__global__ __launch_bounds__(kBlock) void sliced_kernel(StepArgs a, int G) {
  extern __shared__ uint64_t lds[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int W = a.W;
  const int K = kPartialWords / (G * W);  // partial slots per node, >= 4
  uint64_t *part_words = lds + static_cast<size_t>(wv) * kPartialWords;  // [node slot][K][W]
  int32_t *ints = reinterpret_cast<int32_t *>(lds + kWavesPerBlock * kPartialWords);
  float *tab_lds = reinterpret_cast<float *>(ints);  // [kTabLds]
  int32_t *n_part = ints + kTabLds + wv * (kWave + kPartialWords);  // [node slot]
  int32_t *part_ov = n_part + kWave;                                // [node slot][K]
  const unsigned per_xcd = gridDim.x / kXcds;
  const unsigned lb = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  const int block_base = static_cast<int>(lb) * kWavesPerBlock * G;
  const int base = block_base + wv * G;
  const int L = a.we_rel - a.ws_rel;

  // the block's first network's probability table, staged in LDS
  const int tab_net = block_base < a.N ? a.net_of[block_base] : 0;
  const int n_tab = a.p_tab ? min(a.tab_max + 1, kTabLds) : 0;
  for (int t = threadIdx.x; t < n_tab; t += kBlock)
    tab_lds[t] = a.p_tab[static_cast<size_t>(tab_net) * (a.tab_max + 1) + t];

  // ---- counting: lane = (node slot g, word q)
  const int g_l = lane / W;
  const int q_l = lane - g_l * W;
  const int i_l = base + g_l;
  int e0_l = 0, e1_l = 0, net_l = 0, off_l = 0;
  uint32_t rlim_l = UINT32_MAX, stream_l = 0;
  float logb_l = 0.0f;
  uint64_t ss_l = 0ull;
  if (g_l < G && i_l < a.N) {
    e0_l = a.rowptr[i_l];
    e1_l = a.rowptr[i_l + 1];
    net_l = a.net_of[i_l];
    if (a.rank_limit) rlim_l = a.rank_limit[net_l];
    ss_l = a.S[static_cast<size_t>(i_l) * W + q_l];
    if (e1_l - e0_l > a.long_row) e1_l = e0_l = -1;  // long_row_kernel takes it
    const NetDev &nd = a.nets[net_l];
    logb_l = nd.log_base;
    off_l = nd.node_off;
    stream_l = nd.stream;
  }
  n_part[lane] = 0;
  __builtin_amdgcn_wave_barrier();
  uint64_t p[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) p[b] = 0ull;
  {
    int np = 0;
    const uint64_t *col_words = a.I_cur + q_l;
    // words whose replicas are all non-susceptible need no counts
    const int e_end = (ss_l != 0ull || kInjected) ? e1_l : e0_l;
    int4 en[8];
    if (e0_l < e_end) {
#pragma unroll
      for (int t = 0; t < 8; ++t) en[t] = a.entries[min(e0_l + t, e_end - 1)];
    }
    for (int e = e0_l; e < e_end; e += 8) {
      uint64_t m[8];
      int ov[8];
      bool ok[8], part_any = false;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        ov[t] = min(en[t].z, a.we_rel) - max(en[t].y, a.ws_rel);
        ok[t] = (e + t < e_end) & (ov[t] > 0) & (static_cast<uint32_t>(en[t].w) <= rlim_l);
        part_any |= ok[t] & (ov[t] != L);
        // N < 2^24 and N * W < 2^32 (checked on the host): 24-bit multiply
        m[t] = col_words[__umul24(static_cast<uint32_t>(ok[t] ? en[t].x : 0), static_cast<uint32_t>(W))];
      }
      // the next eight entries load while these words are in flight
      if (e + 8 < e_end) {
#pragma unroll
        for (int t = 0; t < 8; ++t) en[t] = a.entries[min(e + 8 + t, e_end - 1)];
      }
      uint64_t w[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) w[t] = (ok[t] & (ov[t] == L)) ? m[t] : 0ull;
      if (__ballot(part_any) != 0ull) {  // partial overlaps: rare
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          if (ok[t] & (ov[t] != L)) {
            if (np < K) {
              part_words[(g_l * K + np) * W + q_l] = m[t];
              part_ov[g_l * K + np] = ov[t];  // the same value from every lane of the node
            }
            ++np;
          }
        }
      }
      uint64_t twosA, twosB, foursA, foursB, eights;
      csa(twosA, p[0], p[0], w[0], w[1]);
      csa(twosB, p[0], p[0], w[2], w[3]);
      csa(foursA, p[1], p[1], twosA, twosB);
      csa(twosA, p[0], p[0], w[4], w[5]);
      csa(twosB, p[0], p[0], w[6], w[7]);
      csa(foursB, p[1], p[1], twosA, twosB);
      csa(eights, p[2], p[2], foursA, foursB);
#pragma unroll
      for (int b = 3; b < NB; ++b) {  // counts stay below 2^NB: no carry out
        const uint64_t c = p[b] & eights;
        p[b] ^= eights;
        eights = c;
      }
    }
    // lanes of a word without susceptible replicas skip the walk; the others
    // all found the same np
    if (np > 0) n_part[g_l] = np;
  }
  int hb_l = -1;
#pragma unroll
  for (int b = 0; b < NB; ++b)
    if (p[b] != 0ull) hb_l = b;
  __syncthreads();

  // ---- draws: lane = replica within a word, one node at a time
  const bool hi_half = lane >= 32;
  const int bitpos = lane & 31;
  for (int g = 0; g < G; ++g) {
    const int i = base + g;
    if (i >= a.N) break;
    const int slot0 = g * W;
    if (__builtin_amdgcn_readlane(e0_l, slot0) < 0) continue;  // a long row
    const int net = __builtin_amdgcn_readlane(net_l, slot0);
    const float log_base = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(logb_l), slot0));
    const int node_off = __builtin_amdgcn_readlane(off_l, slot0);
    const uint32_t stream = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(stream_l), slot0));
    uint64_t *X_k = a.X + a.xk_off + static_cast<size_t>(i) * W;
    const int np = __builtin_amdgcn_readfirstlane(n_part[g]);
    const bool tab_here = net == tab_net;
    const float *tab = a.p_tab ? a.p_tab + static_cast<size_t>(net) * (a.tab_max + 1) : nullptr;
    uint64_t mine = 0ull;
    for (int q0 = 0; q0 < W; q0 += 4) {
      const int nq = min(4, W - q0);
      float E[4];
      uint32_t cnt[4], part[4];
      uint64_t ss[4];
      bool need_any = false;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        ss[t] = 0ull;
        E[t] = 0.0f;
        cnt[t] = part[t] = 0u;
        if (t >= nq) continue;
        const int q = q0 + t;
        const int slot = slot0 + q;
        ss[t] = (static_cast<uint64_t>(static_cast<uint32_t>(
                     __builtin_amdgcn_readlane(static_cast<int>(ss_l >> 32), slot)))
                 << 32) |
                static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(ss_l), slot));
        if (ss[t] == 0ull) continue;
        const int hb = __builtin_amdgcn_readlane(hb_l, slot);
        uint32_t c = 0;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (b <= hb) {
            const uint32_t lo = static_cast<uint32_t>(
                __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(p[b])), slot));
            const uint32_t hi = static_cast<uint32_t>(
                __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(p[b] >> 32)), slot));
            c |= (((hi_half ? hi : lo) >> bitpos) & 1u) << b;
          }
        }
        uint32_t add = 0;
        if (np > 0 && np <= K) {
          for (int j = 0; j < np; ++j) {
            const uint64_t mw = part_words[(g * K + j) * W + q];
            add += static_cast<uint32_t>((mw >> lane) & 1ull) *
                   static_cast<uint32_t>(part_ov[g * K + j]);
          }
        } else if (np > K) {  // slots overflowed: scan the node's entries again
          const int e0 = __builtin_amdgcn_readlane(e0_l, slot0);
          const int e1 = __builtin_amdgcn_readlane(e1_l, slot0);
          const uint32_t rlim = static_cast<uint32_t>(
              __builtin_amdgcn_readlane(static_cast<int>(rlim_l), slot0));
          const uint64_t *cw = a.I_cur + q;
          for (int ee = e0; ee < e1; ++ee) {
            const int4 en = a.entries[ee];
            const int ov = min(en.z, a.we_rel) - max(en.y, a.ws_rel);
            if (!(ov > 0 && ov != L && static_cast<uint32_t>(en.w) <= rlim)) continue;
            const uint64_t mw = cw[static_cast<size_t>(en.x) * W];
            add += static_cast<uint32_t>((mw >> lane) & 1ull) * static_cast<uint32_t>(ov);
          }
        }
        cnt[t] = c;
        part[t] = add;
        E[t] = static_cast<float>(static_cast<uint32_t>(L) * c + add);
        need_any |= ((ss[t] >> lane) & 1ull) && (kInjected || E[t] > 0.0f);
      }
      if (__ballot(need_any) == 0ull) continue;  // no draws in these words
      uint32_t pw[4] = {0u, 0u, 0u, 0u};
      if (!kInjected)
        rp::philox4(a.seed, static_cast<uint32_t>(i - node_off),
                    (static_cast<uint32_t>(q0 >> 2) << 6) | static_cast<uint32_t>(lane), a.slot,
                    stream, pw);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t >= nq) continue;
        const int q = q0 + t;
        const int r = q * kWave + lane;
        bool infect = false;
        if (((ss[t] >> lane) & 1ull) && (kInjected || E[t] > 0.0f)) {
          float P;
          if (part[t] == 0u && tab_here && cnt[t] < static_cast<uint32_t>(n_tab))
            P = tab_lds[cnt[t]];
          else if (tab && part[t] == 0u && cnt[t] <= static_cast<uint32_t>(a.tab_max))
            P = tab[cnt[t]];
          else
            P = 1.0f - rp::rp_expf(log_base * (E[t] / 3600.0f));
          const float u = kInjected ? a.uniforms[static_cast<size_t>(r) * a.N + i] : rp::u24(pw[t]);
          infect = P > u;
        }
        if (infect && a.record) {
          // every network has a.M replicas here: node i's words start at i * W
          a.entry[static_cast<size_t>(i) * W * kWave + r] = 2 * a.k + 1;
          atomicAdd(a.inf_count + static_cast<size_t>(a.k) * a.G + static_cast<size_t>(net) * a.M + r,
                    1);
        }
        const uint64_t x = __ballot(infect);
        if (lane == q) mine = x;
      }
    }
    if (lane < W) X_k[lane] = mine;
  }
}

constexpr size_t sliced_lds_bytes(int nb) {
  (void)nb;
  return static_cast<size_t>(kWavesPerBlock) * kPartialWords * sizeof(uint64_t) +
         (kTabLds + static_cast<size_t>(kWavesPerBlock) * (kWave + kPartialWords)) *
             sizeof(int32_t);
}

// Replicas of word w (entry[w*64 .. w*64+63]) whose last infection was at
// iteration j (entry value 2j+1): the word X(j) held in the full ring.
__device__ __forceinline__ uint64_t infected_at(const int32_t *entry, int64_t w, int32_t j) {
  const int4 *e = reinterpret_cast<const int4 *>(entry + w * kWave);
  const int32_t v = 2 * j + 1;
  uint64_t m = 0ull;
#pragma unroll
  for (int t = 0; t < kWave / 4; ++t) {
    const int4 q = e[t];
    m |= (static_cast<uint64_t>(q.x == v) | (static_cast<uint64_t>(q.y == v) << 1) |
          (static_cast<uint64_t>(q.z == v) << 2) | (static_cast<uint64_t>(q.w == v) << 3))
         << (4 * t);
  }
  return m;
}

// Phase 2 of an iteration: one thread per word applies the timer flips for
// iteration k+1 and advances the running prefix of the infection counts,
// P[k] = P[k-1] + inf_count[k-1].  Each block's words belong to one network
// (a.upd names it; a network's words start at its own block), so the
// flip offsets are the block's.
//
// The flip at offset n >= 2 needs X(k+1-n), the replicas infected n-1
// iterations ago.  Full ring: read from slot (k+1-n) mod depth.  Compact
// form: each wavefront of the grid owns one flag word per iteration (bit l:
// its lane l's word had an infection), written here from X(k); a flagged
// word's X(j) is rebuilt from entry[], which still holds 2j+1 for those
// replicas -- a replica is infected again only once susceptible, and that
// comes at offset n_S, the last of the three.
__global__ __launch_bounds__(kBlock) void update_kernel(StepArgs a) {
  const int4 ub = a.upd[blockIdx.x];
  const int net = ub.x;
  if (net >= 0) {
    // the word loads need only the block's entry; the flip offsets come
    // from the network's constants, loaded alongside
    const int64_t w = static_cast<int64_t>(ub.y) + threadIdx.x;
    const bool valid = w < ub.z;
    const NetDev nd = a.nets[net];
    const uint64_t x = valid ? a.X[a.xk_off + w] : 0ull;
    uint64_t inf = valid ? a.I_cur[w] : 0ull;
    uint64_t sus = valid ? a.S[w] & ~x : 0ull;
    const int32_t k1 = a.k + 1;
    if (!a.F) {
      if (valid) {
        if (const uint64_t *p = ring_word(a, nd.n_on, w)) inf ^= *p;
        if (const uint64_t *p = ring_word(a, nd.n_off, w)) inf ^= *p;
        if (const uint64_t *p = ring_word(a, nd.n_s, w)) sus |= *p;
      }
    } else {
      const int lane = threadIdx.x & (kWave - 1);
      const size_t wave = static_cast<size_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
      const size_t n_waves = static_cast<size_t>(gridDim.x) * kWavesPerBlock;
      const int32_t offs[3] = {nd.n_on, nd.n_off, nd.n_s};
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        const int32_t n = offs[t];
        if (n == kNever || n < 2 || k1 - n < 0) continue;  // uniform over the block
        const int32_t jj = k1 - n;
        const uint64_t f = a.F[static_cast<size_t>(jj % a.depthF) * n_waves + wave];
        const uint64_t xr = (valid && ((f >> lane) & 1ull)) ? infected_at(a.entry, w, jj) : 0ull;
        if (t < 2)
          inf ^= xr;
        else
          sus |= xr;
      }
      const uint64_t fk = __ballot(x != 0ull);
      if (lane == 0) a.F[static_cast<size_t>(a.k % a.depthF) * n_waves + wave] = fk;
    }
    if (valid) apply_flips(a, nd, w, x, inf, sus);
  }
  const int64_t g = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t G = a.G;
  if (a.record && a.k >= 1 && g < G)
    a.P[static_cast<size_t>(a.k) * G + g] =
        a.P[static_cast<size_t>(a.k - 1) * G + g] + a.inf_count[static_cast<size_t>(a.k - 1) * G + g];
}

// History ring: infections of iteration j0 + blockIdx.y per replica.  A
// block takes one network's words q0 .. q0+3 (tally[blockIdx.x] = net << 16
// | q0) and its 8 wavefronts walk the network's nodes in tiles of 64: a
// wavefront loads a tile's 64 x 4 words with consecutive lanes on
// consecutive words, turns it in its own LDS rows so that lane l holds node
// l's word q, and the column count of that row gives each lane the number of
// the tile's nodes whose replica 64 q + lane was infected.  The wavefronts'
// sums meet in LDS and replica column moff + 64 q + r is written once: no
// atomics, every column of the iteration written.
constexpr int kTallyWaves = 8;
constexpr int32_t kHistoryMaxIterations = 1024;  // longest run kept in the history form
__global__ __launch_bounds__(kTallyWaves * kWave) void tally_kernel(StepArgs a,
                                                                    const int32_t *tally,
                                                                    int32_t j0) {
  constexpr int kRow = kWave + 1;  // padded LDS row: the transposing writes spread over banks
  __shared__ uint64_t tile[kTallyWaves][4][kRow];
  __shared__ uint32_t part[kTallyWaves][4][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  const int32_t t = tally[blockIdx.x];
  const int net = t >> 16, q0 = t & 0xFFFF;
  const NetDev nd = a.nets[net];
  const int nq = min(4, nd.W - q0);
  const int32_t j = j0 + static_cast<int32_t>(blockIdx.y);
  const uint64_t *X = a.X + static_cast<size_t>(j % a.depth) * a.NW + nd.woff;
  const int32_t n_nodes = static_cast<int32_t>(nd.nwords / nd.W);
  uint32_t acc[4] = {0u, 0u, 0u, 0u};
  const uint32_t sel3 = 0x06040200u + 0x01010101u * ((static_cast<uint32_t>(lane) >> 3) & 1u);
  // Two tiles per pass: all eight loads are requested before the first is
  // used, then each tile is turned in LDS and counted (the split column
  // count: stages 0-3 per word, 4-5 shared by the four words).
  constexpr int kStride = kTallyWaves * kWave;
  for (int32_t n0 = wv * kWave; n0 < n_nodes; n0 += 2 * kStride) {
    uint64_t v[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int u = 0; u < 4; ++u) {  // tile word u * 64 + lane: node / 4, word % 4
        const int idx = u * kWave + lane;
        const int32_t node = n0 + h * kStride + (idx >> 2);
        const int qq = idx & 3;
        v[h][u] = (node < n_nodes && qq < nq) ? X[static_cast<int64_t>(node) * nd.W + q0 + qq]
                                              : 0ull;
      }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int idx = u * kWave + lane;
        tile[wv][idx & 3][idx >> 2] = v[h][u];
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      uint32_t r3[4];
      bool any = false;
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const uint64_t x = tile[wv][qq][lane];
        r3[qq] = 0u;
        if (__ballot(x != 0ull) != 0ull) {
          r3[qq] = count_rows16(x, 0x55555555u, sel3, lane);
          any = true;
        }
      }
      if (any) {
        const uint32_t c = count_merge4(r3[0], r3[1], r3[2], r3[3]);
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) acc[qq] += __builtin_amdgcn_ubfe(c, 8u * qq, 8u);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
#pragma unroll
  for (int qq = 0; qq < 4; ++qq) part[wv][qq][lane] = count_fix(acc[qq]);
  __syncthreads();
  if (wv < nq) {  // wavefront qq sums word q0 + qq over the 8 wavefronts
    uint32_t sum = 0;
#pragma unroll
    for (int v = 0; v < kTallyWaves; ++v) sum += part[v][wv][lane];
    const int r = (q0 + wv) * kWave + lane;
    if (r < nd.M) a.inf_count[static_cast<size_t>(j) * a.G + nd.moff + r] = static_cast<int32_t>(sum);
  }
}

// Running prefix P[j] = P[j-1] + inf_count[j-1] for j = j0+1 .. j1, one
// thread per replica column.
__global__ void prefix_scan_kernel(int32_t *P, const int32_t *inf_count, int32_t G, int32_t j0,
                                   int32_t j1) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  int32_t p = P[static_cast<size_t>(j0) * G + g];
  for (int32_t j = j0 + 1; j <= j1; ++j) {
    p += inf_count[static_cast<size_t>(j - 1) * G + g];
    P[static_cast<size_t>(j) * G + g] = p;
  }
}

// History ring: brings entry[] (each replica's last infection, 2j+1; 0
// seeded, -1 never, as init_kernel leaves it) from iteration j0 up to k by
// walking X forward over j0 .. k-1: a later infection overwrites an earlier
// one.  One thread per word.
__global__ void rebuild_entry_kernel(StepArgs a, int32_t j0, int32_t k, int32_t init) {
  const int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w >= a.NW) return;
  int32_t *e = a.entry + w * kWave;
  if (init) {  // the reset's entries: 0 for a seed, -1 otherwise
    const uint64_t xi = a.X_init[w];
    for (int r = 0; r < kWave; ++r) e[r] = ((xi >> r) & 1ull) ? 0 : -1;
  }
  for (int32_t j = j0; j < k; ++j)
    for (uint64_t x = a.X[static_cast<size_t>(j % a.depth) * a.NW + w]; x != 0ull; x &= x - 1)
      e[__builtin_ctzll(x)] = 2 * j + 1;
}

struct InitArgs {
  uint64_t *S;
  uint64_t *I;
  uint64_t *X_init;
  int32_t *entry;
  int32_t *init_count;    // [G] replica columns
  const float *uniforms;  // slot 0 or null ([sim][node], every network with M replicas)
  const NetDev *nets;
  const int32_t *net_of;
  uint64_t seed;
  int32_t N;
  int32_t write_entry;    // 0: entry[] is left to the first rebuild (history ring)
};

// cu:73-90 with a counter-based draw (slot 0): seeds, initial bit words,
// entries and per-replica seed counts.  One wavefront per (node, group of
// four words), blockIdx.y being the group: the four words' replicas 64
// apart share one Philox block (replica r takes word (r >> 6) & 3 of
// counter {node, (r >> 8) * 64 + (r & 63), 0, stream}), which each lane
// computes once.  Groups past the node's own network's words have nothing
// to do.
__global__ __launch_bounds__(kBlock) void init_kernel(InitArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int q0 = 4 * static_cast<int>(blockIdx.y);
  const int i = blockIdx.x * kWavesPerBlock + wave;
  if (i >= a.N) return;
  const int net = a.net_of[i];
  const NetDev nd = a.nets[net];
  if (q0 >= nd.W) return;
  uint32_t pw[4] = {0u, 0u, 0u, 0u};
  if (!a.uniforms)
    rp::philox4(a.seed, static_cast<uint32_t>(i - nd.node_off),
                (static_cast<uint32_t>(q0 >> 2) << 6) | static_cast<uint32_t>(lane), 0u, nd.stream,
                pw);
  const int nq = min(4, nd.W - q0);
  for (int k = 0; k < nq; ++k) {
    const int q = q0 + k;
    const int r = q * kWave + lane;
    const bool rv = r < nd.M;
    const size_t w = static_cast<size_t>(nd.wsh + static_cast<int64_t>(i) * nd.W + q);
    bool seed = false;
    if (rv) {
      const float u = a.uniforms ? a.uniforms[static_cast<size_t>(r) * a.N + i]
                                 : rp::u24(k == 0 ? pw[0] : k == 1 ? pw[1] : k == 2 ? pw[2] : pw[3]);
      seed = u < nd.p0;
    }
    const uint64_t xi = __ballot(seed);
    const uint64_t valid = __ballot(rv);
    if (a.write_entry) a.entry[w * kWave + lane] = seed ? 0 : -1;
    if (seed) atomicAdd(a.init_count + nd.moff + r, 1);
    if (lane == 0) {
      a.X_init[w] = xi;
      a.S[w] = nd.m_s == 0 ? valid : (valid & ~xi);
      a.I[w] = nd.m_on == 0 ? xi : 0ull;
    }
  }
}

// P of an exposure of c whole windows, per network: the value the draw
// would compute from E = c * window, 1 - exp(log(1-p) * (E / 3600)).
// The packed entries in 8 bytes, {col | ts << 24, ts >> 8 | te << 12}: the
// form the exposure kernels stream when every entry fits it.  An entry that
// does not (a column of 2^24 or more, a time outside [0, 2^20)) sets
// *misfit; the copy is then discarded.
__global__ void pack8_kernel(const int4 *in, uint2 *out, int64_t n, int32_t *misfit) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += stride) {
    const int4 en = in[e];
    const bool fits = static_cast<uint32_t>(en.x) < (1u << 24) &&
                      static_cast<uint32_t>(en.y) < (1u << 20) &&
                      static_cast<uint32_t>(en.z) < (1u << 20);
    if (!fits) atomicOr(misfit, 1);
    out[e] = make_uint2(static_cast<uint32_t>(en.x) | (static_cast<uint32_t>(en.y) << 24),
                        (static_cast<uint32_t>(en.y) >> 8) | (static_cast<uint32_t>(en.z) << 12));
  }
}

__global__ void p_table_kernel(float *tab, const NetDev *nets, int32_t n_net, int32_t tab_max,
                               int32_t window) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_net * (tab_max + 1)) return;
  const int net = g / (tab_max + 1), c = g - net * (tab_max + 1);
  const float E = static_cast<float>(c * window);
  tab[g] = 1.0f - rp::rp_expf(nets[net].log_base * (E / 3600.0f));
}

// P[k] for the last completed iteration.
__global__ void prefix_tail_kernel(int32_t *P, const int32_t *inf_count, int32_t G, int32_t k) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G && k >= 1)
    P[static_cast<size_t>(k) * G + g] =
        P[static_cast<size_t>(k - 1) * G + g] + inf_count[static_cast<size_t>(k - 1) * G + g];
}

// Counts of blocks j0..j1: exposed, infectious, resistant of the state after
// iteration j-1 (block 0: the initial state), per replica column g (network
// col_net[g], replica g - its first column).
__global__ void counts_kernel(int32_t *counts, const int32_t *P, const int32_t *init_count,
                              const int8_t *class_init, const Segs *segs, const int32_t *col_net,
                              int32_t G, int32_t T, int32_t j0, int32_t j1) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = j0 + blockIdx.y;
  if (g >= G || j > j1) return;
  const int net = col_net[g];
  const Segs &sg = segs[net];
  int32_t c[4] = {0, 0, 0, 0};
  for (int h = 0; h < sg.n; ++h) {
    const int hi = j - sg.a[h] + 1;  // sum over e in [j-b+1, j-a]
    const int lo = j - sg.b[h] + 1;
    if (hi <= 0) continue;
    const int32_t ph = P[static_cast<size_t>(hi) * G + g];
    const int32_t pl = lo > 0 ? P[static_cast<size_t>(lo) * G + g] : 0;
    c[sg.cls[h]] += ph - pl;
  }
  const int ci = class_init[static_cast<size_t>(net) * (T + 1) + j];
  if (ci) c[ci] += init_count[g];
  int32_t *dst = counts + (static_cast<size_t>(j) * G + g) * 4;
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
  const uint32_t rlim = a.rank_limit ? a.rank_limit[a.net_of[i]] : UINT32_MAX;
  const int e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
  // the float order is the reference's and always exact; integer sums only
  // where the window's short-row bound covers the row
  const float E = (kIntMode && e1 - e0 <= a.long_row)
                      ? row_exposure<true>(e0, e1, a.entries, a.I_cur, a.W, a.ws_rel, a.we_rel,
                                           rlim, a.scale, q, lane)
                      : row_exposure<false>(e0, e1, a.entries, a.I_cur, a.W, a.ws_rel,
                                            a.we_rel, rlim, a.scale, q, lane);
  if (r < a.M) out[static_cast<size_t>(r) * a.N + i] = E;
}

// Per-row CSR of one window: pass 0 counts merged columns, pass 1 fills.
__global__ void slice_kernel(StepArgs a, int pass, const int32_t *out_rowptr, int32_t *row_nnz,
                             int32_t *colind, float *val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const uint32_t rlim = a.rank_limit ? a.rank_limit[a.net_of[i]] : UINT32_MAX;
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
      if (ov > 0 && static_cast<uint32_t>(en.w) <= rlim) {
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

namespace {

void release(replay_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
  ctx->graph = nullptr;
  void *bufs[] = {ctx->d_entries, ctx->d_entries8, ctx->d_rowptr, ctx->d_rowinfo, ctx->d_S,      ctx->d_I[0],        ctx->d_I[1],
                  ctx->d_X,       ctx->d_Xinit,  ctx->d_entry,  ctx->d_inf,         ctx->d_P,
                  ctx->d_init_count, ctx->d_counts, ctx->d_class_init, ctx->d_segs, ctx->d_nets,
                  ctx->d_net_of,  ctx->d_rank_limit, ctx->d_uniforms, ctx->d_p_tab,
                  ctx->d_long_rows, ctx->d_upd, ctx->d_col_net, ctx->d_F, ctx->d_tally,
                  ctx->d_node_word, ctx->d_hub_seg0, ctx->d_hub_seg_row, ctx->d_hub_total,
                  ctx->d_hub_part};
  for (void *b : bufs) (void)hipFree(b);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  ctx->stream = nullptr;
}

// Tabulates one network's countdown sequences and the offsets derived from
// them (see the header comment).
int plan_timers(int32_t T, float step, NetHost &nh, std::string &err) {
  auto f = [step](int32_t c) {  // cu:169-174
    if (c > 0) c = static_cast<int32_t>(static_cast<float>(c) - step);
    if (c < 0) c = 0;
    return c;
  };
  nh.lut_new.assign(static_cast<size_t>(T) + 1, 0);
  nh.lut_init.assign(static_cast<size_t>(T) + 1, 0);
  nh.lut_new[0] = static_cast<int32_t>(static_cast<float>(nh.upper) + step);   // cu:155
  nh.lut_init[0] = static_cast<int32_t>(static_cast<float>(nh.medium) + step); // cu:85
  for (int32_t n = 1; n <= T; ++n) {
    nh.lut_new[n] = f(nh.lut_new[n - 1]);
    nh.lut_init[n] = f(nh.lut_init[n - 1]);
  }
  auto infectious = [&nh](int32_t c) { return c > nh.lower && c <= nh.medium; };
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
  if (!ranges(nh.lut_new, 1, &nh.n_on, &nh.n_off, &nh.n_s) ||
      !ranges(nh.lut_init, 0, &nh.m_on, &nh.m_off, &nh.m_s)) {
    err = "infectious period is not contiguous";
    return REPLAY_E_UNSUPPORTED;
  }
  nh.segs.n = 0;
  int32_t prev = -1;
  for (int32_t n = 1; n <= T; ++n) {
    const int cls = host_class(nh.lut_new[n], nh.upper, nh.Ed, nh.Id, nh.Rd);
    if (cls != prev) {
      if (prev > 0) nh.segs.b[nh.segs.n - 1] = n;
      if (cls > 0) {
        if (nh.segs.n >= kMaxSegs) {
          err = "class schedule too long";
          return REPLAY_E_UNSUPPORTED;
        }
        nh.segs.a[nh.segs.n] = n;
        nh.segs.cls[nh.segs.n] = cls;
        nh.segs.b[nh.segs.n] = T + 1;
        nh.segs.n++;
      }
      prev = cls;
    }
  }
  nh.class_init.assign(static_cast<size_t>(T) + 1, 0);
  for (int32_t j = 0; j <= T; ++j)
    nh.class_init[j] = static_cast<int8_t>(host_class(nh.lut_init[j], nh.upper, nh.Ed, nh.Id, nh.Rd));
  return REPLAY_OK;
}

}  // namespace

StepArgs rp::step_args(replay_ctx *ctx, int32_t k) {
  const int32_t win = ctx->iter_window[k];
  const int e = ctx->win_epoch[win];
  const int64_t t0 = ctx->epoch_t0[e];
  StepArgs a;
  std::memset(&a, 0, sizeof a);
  a.rowptr = ctx->d_rowptr + static_cast<size_t>(e) * (ctx->N + 1);
  a.rowinfo = ctx->d_rowinfo + static_cast<size_t>(e) * ctx->N;
  a.entries = ctx->d_entries + ctx->epoch_entry0[e];
  a.entries8 = ctx->d_entries8 ? ctx->d_entries8 + ctx->epoch_entry0[e] : nullptr;
  a.I_cur = ctx->d_I[ctx->cur];
  a.I_next = ctx->d_I[ctx->cur ^ 1];
  a.S = ctx->d_S;
  a.X = ctx->d_X;
  a.X_init = ctx->d_Xinit;
  a.entry = ctx->d_entry;
  a.inf_count = ctx->d_inf;
  a.P = ctx->d_P;
  a.uniforms = ctx->d_uniforms
                   ? ctx->d_uniforms + static_cast<size_t>(k + 1) * ctx->M * ctx->N
                   : nullptr;
  a.nets = ctx->d_nets;
  a.net_of = ctx->d_net_of;
  a.node_word = ctx->d_node_word;
  a.rank_limit = ctx->any_cut ? ctx->d_rank_limit + static_cast<size_t>(win) * ctx->n_net : nullptr;
  a.p_tab = ctx->d_p_tab;
  a.upd = ctx->d_upd;
  a.node_word = ctx->d_node_word;
  a.F = ctx->d_F;
  a.depthF = ctx->depthF;
  a.record = ctx->history ? 0 : 1;
  a.tab_max = ctx->tab_max;
  a.long_row = rp::kLongRow;
  a.seed = ctx->p.seed;
  a.NW = ctx->NW;
  a.N = ctx->N;
  a.M = ctx->uniform ? ctx->M : 0;
  a.W = ctx->W;
  a.Wg = ctx->Wg;
  a.G = ctx->G;
  a.n_net = ctx->n_net;
  a.depth = ctx->depth;
  a.k = k;
  a.xk_off = static_cast<int64_t>(k % ctx->depth) * ctx->NW;
  a.xcd_split = (ctx->variant & REPLAY_V_XCD) != 0;
  a.ws_rel = static_cast<int32_t>(ctx->win_ws[win] - t0);
  a.we_rel = static_cast<int32_t>(ctx->win_we[win] - t0);
  a.scale = ctx->win_scale[win];
  a.slot = static_cast<uint32_t>(k + 1);
  return a;
}

namespace {

template <bool kInt, bool kInj, bool kTry>
void launch_expo_t(const StepArgs &a, int variant, hipStream_t st) {
  // REPLAY_V_BLOCK1: one wavefront per block, which refills wavefront slots
  // as each wavefront ends instead of four at a time (the autotuner picks)
  const int waves = (variant & REPLAY_V_BLOCK1) ? 1 : kWavesPerBlock;
  variant &= ~(REPLAY_V_BLOCK1 | REPLAY_V_XCD);  // (the mapping is a.xcd_split)
  const unsigned nb = static_cast<unsigned>((a.N + waves - 1) / waves);
  const dim3 grid((nb + kXcds - 1) / kXcds * kXcds);
  const dim3 block(static_cast<unsigned>(waves * kWave));
  // whole groups (W a multiple of the group width) take the kernel whose
  // neighbour-word loads are 16 bytes wide
  // and, where the context holds the 8-byte copy of the entries, the one
  // that streams it
  const bool c8 = a.entries8 != nullptr;
  // (every network's W a multiple of the group width: a.Wg, their gcd)
  if (variant == REPLAY_V_NODE16) {
    if (a.Wg % 16 == 0 && c8)
      hipLaunchKernelGGL((expo_kernel<kInt, kInj, 16, true, true, kTry>), grid, block, 0, st, a);
    else if (a.Wg % 16 == 0)
      hipLaunchKernelGGL((expo_kernel<kInt, kInj, 16, true, false, kTry>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((expo_kernel<kInt, kInj, 16, false, false, kTry>), grid, block, 0, st, a);
  } else if (variant == REPLAY_V_NODE4) {
    if (a.Wg % 4 == 0 && c8)
      hipLaunchKernelGGL((expo_kernel<kInt, kInj, 4, true, true, kTry>), grid, block, 0, st, a);
    else if (a.Wg % 4 == 0)
      hipLaunchKernelGGL((expo_kernel<kInt, kInj, 4, true, false, kTry>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((expo_kernel<kInt, kInj, 4, false, false, kTry>), grid, block, 0, st, a);
  } else {
    hipLaunchKernelGGL((expo_kernel<kInt, kInj, 1, false, false, kTry>), grid, block, 0, st, a);
  }
}

// try_int: a float window whose weights are integers (divisor 1), summed in
// integers first
template <bool kInt, bool kInj>
void launch_expo(const StepArgs &a, int variant, hipStream_t st, bool try_int) {
  if constexpr (!kInt) {
    if (try_int) {
      launch_expo_t<false, kInj, true>(a, variant, st);
      return;
    }
  }
  launch_expo_t<kInt, kInj, false>(a, variant, st);
}

// Planes the bit-sliced counters need for rows of up to maxrow contacts
// (0: rows too long for the bit-sliced form).
int sliced_planes(int32_t maxrow) {
  if (maxrow < (1 << 8)) return 8;
  if (maxrow < (1 << 12)) return 12;
  if (maxrow < (1 << 16)) return 16;
  return 0;
}

template <int NB>
void launch_sliced_nb(const StepArgs &a, bool injected, hipStream_t st) {
  const int G = kWave / a.W;
  const int nodes_per_block = kWavesPerBlock * G;
  const unsigned nb = (a.N + nodes_per_block - 1) / nodes_per_block;
  const dim3 grid((nb + kXcds - 1) / kXcds * kXcds);
  if (injected)
    hipLaunchKernelGGL((sliced_kernel<NB, true>), grid, dim3(kBlock), sliced_lds_bytes(NB), st, a, G);
  else
    hipLaunchKernelGGL((sliced_kernel<NB, false>), grid, dim3(kBlock), sliced_lds_bytes(NB), st, a,
                       G);
}

// Long rows of the current epoch: rows[0 .. n_int) allow integer sums,
// rows[n_int .. n_all) need the float order.
struct LongRows {
  const int32_t *rows;
  int32_t n_int, n_all;
  HubArgs hub;  // the same rows cut into segments
};

template <int WT, bool kFull>
void launch_hub_segments(const StepArgs &a, const HubArgs &h, int inj, hipStream_t st) {
  const int64_t waves = static_cast<int64_t>(h.n_seg) * ((a.W + WT - 1) / WT);
  const int64_t blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
  const dim3 grid(static_cast<unsigned>((blocks + kXcds - 1) / kXcds * kXcds));
  if (a.entries8)
    hipLaunchKernelGGL((hub_segment_kernel<v2u, WT, kFull>), grid, dim3(kBlock), 0, st, a, h, inj);
  else
    hipLaunchKernelGGL((hub_segment_kernel<v4i, WT, kFull>), grid, dim3(kBlock), 0, st, a, h, inj);
}

// The segmented long-row kernels: (segment, word group) sums, then (row,
// word) draws.  The words of a wavefront share each chunk of entries:
// 16 per wavefront where every network's W is a multiple of 16 (measured on
// 4096 replicas: 25 us per iteration against 32 with 4 and 64 with 1),
// else 4.
template <bool kInj>
void launch_hub(const StepArgs &a, const HubArgs &h, hipStream_t st) {
  const int inj = kInj ? 1 : 0;
  if (a.Wg % 16 == 0)
    launch_hub_segments<16, true>(a, h, inj, st);
  else if (a.Wg % 4 == 0)
    launch_hub_segments<4, true>(a, h, inj, st);
  else
    launch_hub_segments<4, false>(a, h, inj, st);
  const int64_t fin_waves = static_cast<int64_t>(h.n_rows) * a.W;
  const dim3 fgrid(static_cast<unsigned>((fin_waves + kWavesPerBlock - 1) / kWavesPerBlock));
  hipLaunchKernelGGL(hub_finish_kernel<kInj>, fgrid, dim3(kBlock), 0, st, a, h);
}

template <bool kInt, bool kInj, bool kTryInt>
void launch_long_part(const StepArgs &a, const int32_t *rows, int32_t n, hipStream_t st) {
  if (n <= 0) return;
  const int64_t waves = static_cast<int64_t>(n) * a.W;
  const dim3 grid(static_cast<unsigned>((waves + kWavesPerBlock - 1) / kWavesPerBlock));
  hipLaunchKernelGGL((long_row_kernel<kInt, kInj, kTryInt>), grid, dim3(kBlock), 0, st, a, rows,
                     n);
}

// Windows of integer weights (divisor 1) whose segment sums fit 32 bits take
// the segmented kernels.  Otherwise one wavefront walks each (row, word):
// in integers where the row's weight bound is below 2^24; integers first
// for the other rows of integer weights; the reference order elsewhere.
template <bool kInj>
void launch_long(const StepArgs &a, bool int_mode, const LongRows &lr, hipStream_t st) {
  if (lr.n_all <= 0) return;
  const int64_t L = static_cast<int64_t>(a.we_rel) - a.ws_rel;
  if (a.scale == 1.0f && L < (int64_t(1) << 24) && L * kHubSeg < (int64_t(1) << 32)) {
    launch_hub<kInj>(a, lr.hub, st);
    return;
  }
  if (int_mode) {
    launch_long_part<true, kInj, false>(a, lr.rows, lr.n_int, st);
    launch_long_part<false, kInj, true>(a, lr.rows + lr.n_int, lr.n_all - lr.n_int, st);
  } else if (a.scale == 1.0f) {
    launch_long_part<false, kInj, true>(a, lr.rows, lr.n_all, st);
  } else {
    launch_long_part<false, kInj, false>(a, lr.rows, lr.n_all, st);
  }
}

// One iteration: exposure + draws per node (long rows by the hub kernels or
// long_row_kernel), then the word flips.  nb_planes > 0 selects the
// bit-sliced exposure kernel (integer windows only).
void launch_step(const StepArgs &a, int variant, bool int_mode, bool injected, dim3 grid,
                 hipStream_t st, int nb_planes, const LongRows &lr) {
  if (variant == REPLAY_V_SLICED && !(int_mode && nb_planes > 0))
    variant = a.W <= 4 ? REPLAY_V_NODE4 : REPLAY_V_NODE16;
  if (variant == REPLAY_V_SLICED) {
    if (nb_planes == 8)
      launch_sliced_nb<8>(a, injected, st);
    else if (nb_planes == 12)
      launch_sliced_nb<12>(a, injected, st);
    else
      launch_sliced_nb<16>(a, injected, st);
  } else if (int_mode && !injected) {
    launch_expo<true, false>(a, variant, st, false);
  } else if (int_mode) {
    launch_expo<true, true>(a, variant, st, false);
  } else if (!injected) {
    launch_expo<false, false>(a, variant, st, a.scale == 1.0f);
  } else {
    launch_expo<false, true>(a, variant, st, a.scale == 1.0f);
  }
  if (injected)
    launch_long<true>(a, int_mode, lr, st);
  else
    launch_long<false>(a, int_mode, lr, st);
  hipLaunchKernelGGL(update_kernel, grid, dim3(kBlock), 0, st, a);
}

dim3 step_grid(const replay_ctx *ctx) {
  // blocks cover every network's words (upd_blocks) and, for the running
  // prefix, every replica column
  const int64_t cover = std::max<int64_t>(ctx->upd_blocks, (ctx->G + kBlock - 1) / kBlock);
  return dim3(static_cast<unsigned>(cover));
}

bool variant_applies(const replay_ctx *ctx, int v) {
  // the XCD mapping is a schedule of the exposure kernel (not of the
  // bit-sliced one)
  if (v & REPLAY_V_XCD) {
    v &= ~REPLAY_V_XCD;
    if (v == REPLAY_V_SLICED) return false;
  }
  if (v == (REPLAY_V_NODE4 | REPLAY_V_BLOCK1) || v == (REPLAY_V_NODE16 | REPLAY_V_BLOCK1)) return true;
  if (v == REPLAY_V_SLICED)
    return ctx->uniform && ctx->W <= kWave && ctx->N < (1 << 24) &&
           static_cast<int64_t>(ctx->N) * ctx->W < (int64_t(1) << 32);
  return v == REPLAY_V_WORD || v == REPLAY_V_NODE4 || v == REPLAY_V_NODE16;
}

int enqueue_steps(replay_ctx *ctx, int32_t steps) {
  const dim3 grid = step_grid(ctx);
  for (int32_t s = 0; s < steps; ++s) {
    const int32_t k = ctx->steps_done;
    StepArgs a = step_args(ctx, k);
    const int32_t win = ctx->iter_window[k];
    const int32_t ep = ctx->win_epoch[win];
    const bool int_mode = ctx->win_int[win] != 0;
    const int32_t n_long = static_cast<int32_t>(ctx->long_off[ep + 1] - ctx->long_off[ep]);
    const LongRows lr{ctx->d_long_rows + ctx->long_off[ep], ctx->long_n_int[ep], n_long,
                      HubArgs{ctx->d_long_rows + ctx->long_off[ep],
                              ctx->d_hub_seg0 + ctx->long_off[ep] + ep,
                              ctx->d_hub_seg_row + ctx->hub_seg_off[ep], ctx->d_hub_total,
                              ctx->d_hub_part, n_long,
                              static_cast<int32_t>(ctx->hub_seg_off[ep + 1] -
                                                   ctx->hub_seg_off[ep])}};
    launch_step(a, ctx->variant, int_mode, a.uniforms != nullptr, grid, ctx->stream,
                sliced_planes(ctx->epoch_maxrow[ep]), lr);
    HIP_OR_FAIL(ctx, hipGetLastError());
    ctx->cur ^= 1;
    ctx->steps_done = k + 1;
  }
  return REPLAY_OK;
}

int do_reset(replay_ctx *ctx) {
  const size_t G = static_cast<size_t>(ctx->G);
  const size_t T = static_cast<size_t>(ctx->T);
  HIP_OR_FAIL(ctx, hipMemsetAsync(ctx->d_inf, 0, (T + 1) * G * sizeof(int32_t), ctx->stream));
  HIP_OR_FAIL(ctx, hipMemsetAsync(ctx->d_P, 0, (T + 2) * G * sizeof(int32_t), ctx->stream));
  HIP_OR_FAIL(ctx, hipMemsetAsync(ctx->d_init_count, 0, G * sizeof(int32_t), ctx->stream));
  if (ctx->hub_total_n > 0)
    HIP_OR_FAIL(ctx, hipMemsetAsync(ctx->d_hub_total, 0, ctx->hub_total_n * sizeof(uint64_t),
                                    ctx->stream));
  InitArgs a;
  std::memset(&a, 0, sizeof a);
  a.S = ctx->d_S;
  a.I = ctx->d_I[0];
  a.X_init = ctx->d_Xinit;
  a.entry = ctx->d_entry;
  // the history ring rebuilds entry[] when states are read, so the reset
  // need not write it (826 MB on the 64-network workload)
  a.write_entry = ctx->history ? 0 : 1;
  ctx->entry_ready = !ctx->history;
  a.init_count = ctx->d_init_count;
  a.uniforms = ctx->d_uniforms;
  a.nets = ctx->d_nets;
  a.net_of = ctx->d_net_of;
  a.seed = ctx->p.seed;
  a.N = ctx->N;
  dim3 grid((ctx->N + kWavesPerBlock - 1) / kWavesPerBlock, (ctx->W + 3) / 4);
  hipLaunchKernelGGL(init_kernel, grid, dim3(kBlock), 0, ctx->stream, a);
  HIP_OR_FAIL(ctx, hipGetLastError());
  ctx->cur = 0;
  ctx->steps_done = 0;
  ctx->counts_done = -1;
  ctx->tallied = 0;
  ctx->rebuilt = 0;
  return REPLAY_OK;
}

// Counts blocks up to steps_done (enqueued on the stream).
int enqueue_counts(replay_ctx *ctx) {
  const int32_t j0 = ctx->counts_done + 1, j1 = ctx->steps_done;
  if (j0 > j1) return REPLAY_OK;
  const int G = ctx->G;
  const int tb = 256, nb = (G + tb - 1) / tb;
  if (ctx->history) {
    // infections of the iterations not yet tallied, from the ring, and the
    // running prefix over them
    const int32_t t0 = ctx->tallied;
    if (j1 > t0) {
      StepArgs a = step_args(ctx, 0);
      const dim3 grid(static_cast<unsigned>(std::accumulate(
                          ctx->nets.begin(), ctx->nets.end(), 0,
                          [](int acc, const NetHost &nh) { return acc + (nh.W + 3) / 4; })),
                      static_cast<unsigned>(j1 - t0));
      hipLaunchKernelGGL(tally_kernel, grid, dim3(kTallyWaves * kWave), 0, ctx->stream, a,
                         ctx->d_tally, t0);
      hipLaunchKernelGGL(prefix_scan_kernel, dim3(nb), dim3(tb), 0, ctx->stream, ctx->d_P,
                         ctx->d_inf, G, t0, j1);
      ctx->tallied = j1;
    }
  } else {
    hipLaunchKernelGGL(prefix_tail_kernel, dim3(nb), dim3(tb), 0, ctx->stream, ctx->d_P,
                       ctx->d_inf, G, j1);
  }
  hipLaunchKernelGGL(counts_kernel, dim3(nb, j1 - j0 + 1), dim3(tb), 0, ctx->stream,
                     ctx->d_counts, ctx->d_P, ctx->d_init_count, ctx->d_class_init, ctx->d_segs,
                     ctx->d_col_net, G, ctx->T, j0, j1);
  HIP_OR_FAIL(ctx, hipGetLastError());
  ctx->counts_done = j1;
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

int check_durations(int32_t Ed, int32_t Id, int32_t Rd) {
  if (Ed < 0 || Id < 0 || Rd < 0) return fail(nullptr, REPLAY_E_ARG, "durations must be >= 0");
  if (static_cast<int64_t>(Ed) + Id + Rd > INT32_MAX)
    return fail(nullptr, REPLAY_E_ARG, "durations overflow int32");
  return REPLAY_OK;
}

// Runs fn(0..n-1) on up to 16 host threads.
template <typename F>
void parallel_for(size_t n, F fn) {
  const size_t nt = std::min<size_t>(
      n, std::max<size_t>(1, std::min<unsigned>(16, std::thread::hardware_concurrency())));
  if (nt <= 1) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; ++t)
    th.emplace_back([&] {
      for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
    });
  for (auto &x : th) x.join();
}

// Picks the epoch size for all networks together.
int32_t pick_epoch(replay_ctx *ctx) {
  if (ctx->p.windows_per_epoch > 0) return ctx->p.windows_per_epoch;
  // 32 GiB of 16-byte entries by default (48 GiB with the 8-byte copy) of
  // the 288 GB: shorter epochs stream fewer entries per step (measured:
  // 32 -> 8 windows per epoch, -2.7% per step on the 64-network workload;
  // 128 -> 32 on the dense one, -8.7%), at the cost of packing more.
  const int64_t budget = ctx->p.max_slice_bytes > 0 ? ctx->p.max_slice_bytes : (int64_t(32) << 30);
  const int32_t W = static_cast<int32_t>(ctx->win_ws.size());
  std::vector<int64_t> cand;
  for (int64_t S = 1; S < W; S *= 2) cand.push_back(S);
  cand.push_back(std::max<int32_t>(W, 1));
  std::vector<std::vector<int64_t>> per(ctx->nets.size());
  parallel_for(ctx->nets.size(),
               [&](size_t n) { per[n] = rp::slice_bytes(ctx->nets[n].net, ctx->nets[n].plan, cand); });
  std::vector<int64_t> total(cand.size(), 0);
  for (const auto &b : per)
    for (size_t j = 0; j < cand.size(); ++j) total[j] += b[j];
  for (size_t j = 0; j < cand.size(); ++j)
    if (total[j] <= budget) return static_cast<int32_t>(cand[j]);
  return static_cast<int32_t>(cand.back());
}

int finish_open(replay_ctx *ctx) {
  std::string err;
  const replay_params &p = ctx->p;
  ctx->T = p.iterations;
  ctx->n_net = static_cast<int32_t>(ctx->nets.size());
  // per network, independent, on host threads: rank order, cap, windows,
  // timers
  const size_t n_nets = ctx->nets.size();
  std::vector<int> rcs(n_nets, REPLAY_OK);
  std::vector<std::string> errs(n_nets);
  parallel_for(n_nets, [&](size_t n) {
    NetHost &nh = ctx->nets[n];
    int rc = rp::finalize(nh.net, errs[n]);
    if (rc == REPLAY_OK) rc = rp::plan_windows(nh.net, p, nh.plan, errs[n]);
    if (rc == REPLAY_OK) {
      nh.upper = nh.Ed + nh.Id + nh.Rd;
      nh.medium = nh.Id + nh.Rd;
      nh.lower = nh.Rd;
      nh.log_base = logf(1.0f - nh.beta);
      rc = plan_timers(ctx->T, p.step_size, nh, errs[n]);
    }
    rcs[n] = rc;
  });
  for (size_t n = 0; n < n_nets; ++n)
    if (rcs[n] != REPLAY_OK)
      return fail(ctx, rcs[n], (n_nets > 1 ? "network " + std::to_string(n) + ": " : "") + errs[n]);
  // node rows, words and replica columns: network after network
  int32_t off = 0;
  int64_t woff = 0, moff = 0;
  for (auto &nh : ctx->nets) {
    nh.node_off = off;
    off += nh.net.n_nodes;
    nh.W = (nh.M + kWave - 1) / kWave;
    nh.woff = woff;
    woff += static_cast<int64_t>(nh.net.n_nodes) * nh.W;
    nh.moff = static_cast<int32_t>(std::min<int64_t>(moff, INT32_MAX));
    moff += nh.M;
    ctx->M = std::max(ctx->M, nh.M);
    ctx->W = std::max(ctx->W, nh.W);
    ctx->Wg = std::gcd(ctx->Wg, nh.W);
    ctx->uniform = ctx->uniform && nh.M == ctx->nets[0].M;
    for (int32_t n : {nh.n_on, nh.n_off, nh.n_s})
      if (n != kNever) ctx->depth = std::max(ctx->depth, n);
    ctx->n_contacts += static_cast<int64_t>(nh.net.contacts.size());
    ctx->nlists = std::max(ctx->nlists, nh.net.nlists);
  }
  ctx->N = off;
  ctx->NW = woff;
  if (woff > INT32_MAX) return fail(ctx, REPLAY_E_UNSUPPORTED, "more than 2^31 replica words");
  if (moff > INT32_MAX / 4) return fail(ctx, REPLAY_E_UNSUPPORTED, "more than 2^29 replicas");
  ctx->G = static_cast<int32_t>(moff);
  // update_kernel's blocks: each network's words start at a block of their own
  std::vector<int32_t> blk0(static_cast<size_t>(ctx->n_net));
  std::vector<int4> upd;
  for (int32_t n = 0; n < ctx->n_net; ++n) {
    const NetHost &nh = ctx->nets[n];
    blk0[n] = static_cast<int32_t>(upd.size());
    const int64_t end = nh.woff + static_cast<int64_t>(nh.net.n_nodes) * nh.W;
    for (int64_t w = nh.woff; w < end; w += kBlock)
      upd.push_back(make_int4(n, static_cast<int32_t>(w), static_cast<int32_t>(end), 0));
  }
  ctx->upd_blocks = static_cast<int32_t>(upd.size());
  upd.resize(std::max<size_t>(upd.size(), static_cast<size_t>((ctx->G + kBlock - 1) / kBlock)),
             make_int4(-1, 0, 0, 0));
  std::vector<int32_t> node_word(static_cast<size_t>(ctx->N));
  for (const NetHost &nh : ctx->nets)
    for (int32_t r = 0; r < nh.net.n_nodes; ++r)
      node_word[static_cast<size_t>(nh.node_off) + r] =
          static_cast<int32_t>(nh.woff + static_cast<int64_t>(r) * nh.W);

  std::vector<int32_t> col_net(static_cast<size_t>(ctx->G));
  for (int32_t n = 0; n < ctx->n_net; ++n)
    std::fill(col_net.begin() + ctx->nets[n].moff, col_net.begin() + ctx->nets[n].moff + ctx->nets[n].M,
              n);
  // windows are shared: every network planned the same iterations
  const rp::Plan &p0 = ctx->nets[0].plan;
  ctx->iter_time = p0.iter_time;
  ctx->iter_window = p0.iter_window;
  const int32_t nW = static_cast<int32_t>(p0.windows.size());
  for (const auto &w : p0.windows) {
    ctx->win_ws.push_back(w.ws);
    ctx->win_we.push_back(w.we);
    ctx->win_scale.push_back(w.scale);
  }
  // slices with one epoch size for all networks, merged row-wise
  const int32_t S = pick_epoch(ctx);
  parallel_for(n_nets, [&](size_t n) {
    NetHost &nh = ctx->nets[n];
    rcs[n] = rp::build_slices(nh.net, nh.plan, S, INT64_MAX, nh.slices, errs[n]);
  });
  for (size_t n = 0; n < n_nets; ++n)
    if (rcs[n] != REPLAY_OK) return fail(ctx, rcs[n], errs[n]);
  ctx->windows_per_epoch = nW ? ctx->nets[0].slices.windows_per_epoch : 1;
  ctx->epoch_t0 = ctx->nets[0].slices.epoch_t0;
  const int32_t n_ep = static_cast<int32_t>(ctx->epoch_t0.size());
  ctx->win_epoch.assign(static_cast<size_t>(nW), 0);
  ctx->win_int.assign(static_cast<size_t>(nW), 1);
  std::vector<uint32_t> limits(static_cast<size_t>(nW) * ctx->n_net, UINT32_MAX);
  for (int32_t w = 0; w < nW; ++w) {
    ctx->win_epoch[w] = ctx->nets[0].plan.windows[w].epoch;
    bool cut = false;
    for (int32_t n = 0; n < ctx->n_net; ++n) {
      const rp::Window &nw = ctx->nets[n].plan.windows[w];
      ctx->win_int[w] &= static_cast<uint8_t>(nw.int_mode != 0);
      limits[static_cast<size_t>(w) * ctx->n_net + n] = nw.rank_limit;
      cut |= nw.rank_limit != UINT32_MAX;
    }
    ctx->truncated += cut;
  }
  ctx->any_cut = ctx->truncated > 0;
  std::vector<int32_t> rowptr(static_cast<size_t>(n_ep) * (ctx->N + 1));
  ctx->epoch_entry0.assign(static_cast<size_t>(n_ep) + 1, 0);
  ctx->epoch_maxrow.assign(static_cast<size_t>(n_ep), 0);
  int64_t total = 0;
  for (int32_t e = 0; e < n_ep; ++e) {
    ctx->epoch_entry0[e] = total;
    int64_t base = 0;  // entries of earlier networks in this epoch
    for (auto &nh : ctx->nets) {
      const int32_t Nn = nh.net.n_nodes;
      const int32_t *rp = &nh.slices.rowptr[static_cast<size_t>(e) * (Nn + 1)];
      for (int32_t r = 0; r < Nn; ++r)
        rowptr[static_cast<size_t>(e) * (ctx->N + 1) + nh.node_off + r] =
            static_cast<int32_t>(base + rp[r]);
      base += rp[Nn];
      if (base > INT32_MAX) return fail(ctx, REPLAY_E_UNSUPPORTED, "epoch holds more than 2^31 entries");
    }
    rowptr[static_cast<size_t>(e) * (ctx->N + 1) + ctx->N] = static_cast<int32_t>(base);
    total += base;
    int32_t mr = 0;
    const int32_t *rpe = &rowptr[static_cast<size_t>(e) * (ctx->N + 1)];
    for (int32_t r = 0; r < ctx->N; ++r) {
      const int32_t len = rpe[r + 1] - rpe[r];
      if (len <= rp::kLongRow) mr = std::max(mr, len);
    }
    ctx->epoch_maxrow[e] = mr;
  }
  // long rows of every epoch, integer-capable ones first
  std::vector<int32_t> long_rows;
  ctx->long_off.assign(static_cast<size_t>(n_ep) + 1, 0);
  ctx->long_n_int.assign(static_cast<size_t>(n_ep), 0);
  for (int32_t e = 0; e < n_ep; ++e) {
    for (int pass = 0; pass < 2; ++pass) {
      for (auto &nh : ctx->nets) {
        const auto &sl = nh.slices;
        if (sl.long_off.empty()) continue;
        for (int64_t q = sl.long_off[e]; q < sl.long_off[e + 1]; ++q) {
          if ((sl.long_int[static_cast<size_t>(q)] != 0) != (pass == 0)) continue;
          long_rows.push_back(nh.node_off + sl.long_row[static_cast<size_t>(q)]);
        }
      }
      if (pass == 0)
        ctx->long_n_int[e] = static_cast<int32_t>(long_rows.size() - ctx->long_off[e]);
    }
    ctx->long_off[static_cast<size_t>(e) + 1] = static_cast<int64_t>(long_rows.size());
  }
  // the long rows' segments of kHubSeg entries, epoch by epoch
  std::vector<int32_t> hub_seg0, hub_seg_row;
  ctx->hub_seg_off.assign(static_cast<size_t>(n_ep) + 1, 0);
  int64_t hub_rows_max = 0, hub_seg_max = 0;
  for (int32_t e = 0; e < n_ep; ++e) {
    const int32_t *rpe = &rowptr[static_cast<size_t>(e) * (ctx->N + 1)];
    int32_t ns = 0;
    for (int64_t g = ctx->long_off[e]; g < ctx->long_off[e + 1]; ++g) {
      hub_seg0.push_back(ns);
      const int32_t r = long_rows[static_cast<size_t>(g)];
      const int32_t nseg = (rpe[r + 1] - rpe[r] + kHubSeg - 1) / kHubSeg;
      hub_seg_row.insert(hub_seg_row.end(), static_cast<size_t>(nseg),
                         static_cast<int32_t>(g - ctx->long_off[e]));
      ns += nseg;
    }
    hub_seg0.push_back(ns);
    ctx->hub_seg_off[static_cast<size_t>(e) + 1] = static_cast<int64_t>(hub_seg_row.size());
    hub_rows_max = std::max<int64_t>(hub_rows_max, ctx->long_off[e + 1] - ctx->long_off[e]);
    hub_seg_max = std::max<int64_t>(hub_seg_max, ns);
  }
  ctx->epoch_entry0[n_ep] = total;
  ctx->n_entries = total;
  // Epoch e holds the networks' epoch-e entries one network after another,
  // columns shifted by the network's first node.  One network: its array is
  // already that layout (shift 0) and is taken over; several: each network
  // copies its ranges on its own thread.
  std::vector<rp::Entry> entries;
  if (ctx->nets.size() == 1) {
    entries.swap(ctx->nets[0].slices.entries);
  } else {
    entries.resize(static_cast<size_t>(total));
    std::vector<int64_t> dst0(static_cast<size_t>(n_ep) * ctx->nets.size());
    for (int32_t e = 0; e < n_ep; ++e) {
      int64_t dst = ctx->epoch_entry0[e];
      for (size_t n = 0; n < ctx->nets.size(); ++n) {
        dst0[static_cast<size_t>(e) * ctx->nets.size() + n] = dst;
        const auto &sl = ctx->nets[n].slices;
        dst += sl.epoch_entry0[e + 1] - sl.epoch_entry0[e];
      }
    }
    parallel_for(ctx->nets.size(), [&](size_t n) {
      const NetHost &nh = ctx->nets[n];
      for (int32_t e = 0; e < n_ep; ++e) {
        int64_t dst = dst0[static_cast<size_t>(e) * ctx->nets.size() + n];
        for (int64_t q = nh.slices.epoch_entry0[e]; q < nh.slices.epoch_entry0[e + 1]; ++q) {
          rp::Entry en = nh.slices.entries[static_cast<size_t>(q)];
          en.col += nh.node_off;
          entries[static_cast<size_t>(dst++)] = en;
        }
      }
    });
  }
  for (auto &nh : ctx->nets) {  // per-network copies are no longer needed
    std::vector<rp::Entry>().swap(nh.slices.entries);
    std::vector<int32_t>().swap(nh.slices.rowptr);
  }
  // device tables
  std::vector<NetDev> nd(static_cast<size_t>(ctx->n_net));
  std::vector<Segs> segs(static_cast<size_t>(ctx->n_net));
  std::vector<int8_t> class_init(static_cast<size_t>(ctx->n_net) * (ctx->T + 1));
  std::vector<int32_t> net_of(static_cast<size_t>(ctx->N));
  for (int32_t n = 0; n < ctx->n_net; ++n) {
    const NetHost &nh = ctx->nets[n];
    nd[n] = NetDev{nh.log_base, nh.p0,      nh.node_off, nh.stream, nh.n_on,
                   nh.n_off,    nh.n_s,     nh.m_on,     nh.m_off,  nh.m_s,
                   nh.M,        nh.W,       nh.moff,     blk0[n],
                   nh.woff - static_cast<int64_t>(nh.node_off) * nh.W,
                   nh.woff,     static_cast<int64_t>(nh.net.n_nodes) * nh.W};
    segs[n] = nh.segs;
    std::copy(nh.class_init.begin(), nh.class_init.end(),
              class_init.begin() + static_cast<size_t>(n) * (ctx->T + 1));
    std::fill(net_of.begin() + nh.node_off, net_of.begin() + nh.node_off + nh.net.n_nodes, n);
  }
  HIP_OR_FAIL(ctx, hipSetDevice(p.device));
  HIP_OR_FAIL(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  const size_t NW = static_cast<size_t>(ctx->NW);
  const size_t T = static_cast<size_t>(ctx->T);
  const size_t G = static_cast<size_t>(ctx->G);
  // The infection ring holds depth slots of every word.  Beyond the budget
  // the compact form keeps one slot and, over depth slots, one flag bit per
  // word: a flag word per wavefront of update_kernel's grid.
  // Where every iteration fits (T + 1 slots), the ring keeps them all and
  // the step kernels record nothing per infection (history form).
  const int64_t ring_budget = p.max_ring_bytes > 0 ? p.max_ring_bytes : (int64_t(16) << 30);
  const size_t upd_waves = static_cast<size_t>(step_grid(ctx).x) * kWavesPerBlock;
  std::vector<int32_t> tally;
  // The tally's running prefix walks the iterations in sequence, which on
  // long runs (thousands of iterations) costs more than the per-infection
  // records it saves: those keep the offsets ring.
  if (ctx->T >= 1 && ctx->T <= kHistoryMaxIterations && !(p.flags & REPLAY_F_NO_HISTORY) &&
      static_cast<int64_t>(NW) * 8 * (ctx->T + 1) <= ring_budget) {
    ctx->history = true;
    ctx->depth = ctx->T + 1;
    for (int32_t n = 0; n < ctx->n_net; ++n)
      for (int32_t q = 0; q < ctx->nets[n].W; q += 4) tally.push_back(n << 16 | q);
    if (ctx->n_net >= (1 << 15) || ctx->W >= (1 << 16))
      return fail(ctx, REPLAY_E_UNSUPPORTED, "too many networks or words for the history tally");
  } else if (static_cast<int64_t>(NW) * 8 * ctx->depth > ring_budget && ctx->depth > 1) {
    ctx->compact = true;
    ctx->depthF = ctx->depth;
    ctx->depth = 1;
  }
  // per epoch and node: row bounds and network together (expo_kernel)
  std::vector<int4> rowinfo(rowptr.size() / (static_cast<size_t>(ctx->N) + 1) * ctx->N);
  for (size_t e = 0; e < rowinfo.size() / ctx->N; ++e)
    for (int32_t r = 0; r < ctx->N; ++r) {
      const int32_t *rp = &rowptr[e * (static_cast<size_t>(ctx->N) + 1) + r];
      rowinfo[e * ctx->N + r] = make_int4(rp[0], rp[1], net_of[r], 0);
    }
  int rc;
  if ((rc = upload(ctx, &ctx->d_entries, reinterpret_cast<const int4 *>(entries.data()),
                   entries.size())) ||
      (rc = upload(ctx, &ctx->d_rowptr, rowptr.data(), rowptr.size())) ||
      (rc = upload(ctx, &ctx->d_rowinfo, rowinfo.data(), rowinfo.size())) ||
      (rc = upload(ctx, &ctx->d_nets, nd.data(), nd.size())) ||
      (rc = upload(ctx, &ctx->d_segs, segs.data(), segs.size())) ||
      (rc = upload(ctx, &ctx->d_class_init, class_init.data(), class_init.size())) ||
      (rc = upload(ctx, &ctx->d_net_of, net_of.data(), net_of.size())) ||
      (rc = upload(ctx, &ctx->d_upd, upd.data(), upd.size())) ||
      (rc = upload(ctx, &ctx->d_node_word, node_word.data(), node_word.size())) ||
      (rc = upload(ctx, &ctx->d_col_net, col_net.data(), col_net.size())) ||
      (rc = (ctx->any_cut ? upload(ctx, &ctx->d_rank_limit, limits.data(), limits.size())
                          : REPLAY_OK)) ||
      (rc = upload(ctx, &ctx->d_long_rows, long_rows.data(), long_rows.size())) ||
      (rc = upload(ctx, &ctx->d_hub_seg0, hub_seg0.data(), hub_seg0.size())) ||
      (rc = upload(ctx, &ctx->d_hub_seg_row, hub_seg_row.data(), hub_seg_row.size())) ||
      (rc = dev_alloc(ctx, &ctx->d_hub_total, hub_rows_max * ctx->W * kWave)) ||
      (rc = dev_alloc(ctx, &ctx->d_hub_part, hub_seg_max * ctx->W * kWave)) ||
      (rc = dev_alloc(ctx, &ctx->d_S, NW)) || (rc = dev_alloc(ctx, &ctx->d_I[0], NW)) ||
      (rc = dev_alloc(ctx, &ctx->d_I[1], NW)) ||
      (rc = dev_alloc(ctx, &ctx->d_X, NW * static_cast<size_t>(ctx->depth))) ||
      (rc = (ctx->compact ? dev_alloc(ctx, &ctx->d_F, upd_waves * static_cast<size_t>(ctx->depthF))
                          : REPLAY_OK)) ||
      (rc = (ctx->history ? upload(ctx, &ctx->d_tally, tally.data(), tally.size()) : REPLAY_OK)) ||
      (rc = dev_alloc(ctx, &ctx->d_Xinit, NW)) ||
      (rc = dev_alloc(ctx, &ctx->d_entry, NW * kWave)) ||
      (rc = dev_alloc(ctx, &ctx->d_inf, (T + 1) * G)) ||
      (rc = dev_alloc(ctx, &ctx->d_P, (T + 2) * G)) ||
      (rc = dev_alloc(ctx, &ctx->d_init_count, G)) ||
      (rc = dev_alloc(ctx, &ctx->d_counts, (T + 1) * G * 4)))
    return rc;
  // The 8-byte copy of the entries, for contexts without a rank cap.  It is
  // optional: when the device has no room for it or an entry does not fit
  // it, the step kernels read the 16-byte entries.
  if (!ctx->any_cut && !(p.flags & REPLAY_F_WIDE_ENTRIES) && !entries.empty()) {
    const int64_t n = static_cast<int64_t>(entries.size());
    int32_t *d_misfit = nullptr;
    DeviceScratch scratch;
    if (hipMalloc(reinterpret_cast<void **>(&ctx->d_entries8), static_cast<size_t>(n) * 8) !=
            hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&d_misfit), sizeof(int32_t)) != hipSuccess) {
      (void)hipGetLastError();  // out of room: not an error of the open
      (void)hipFree(ctx->d_entries8);
      ctx->d_entries8 = nullptr;
    } else {
      scratch.bufs.push_back(d_misfit);
      int32_t misfit = 0;
      HIP_OR_FAIL(ctx, hipMemcpyAsync(d_misfit, &misfit, sizeof misfit, hipMemcpyHostToDevice,
                                      ctx->stream));
      const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
      hipLaunchKernelGGL(pack8_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                         ctx->stream, ctx->d_entries, ctx->d_entries8, n, d_misfit);
      HIP_OR_FAIL(ctx, hipGetLastError());
      HIP_OR_FAIL(ctx, hipMemcpyAsync(&misfit, d_misfit, sizeof misfit, hipMemcpyDeviceToHost,
                                      ctx->stream));
      HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
      if (misfit) {
        (void)hipFree(ctx->d_entries8);
        ctx->d_entries8 = nullptr;
      }
    }
  }
  // probabilities of whole-window exposures, computed by the device's own exp
  if (p.static_network_duration < (int64_t(1) << 24)) {
    const int32_t D = static_cast<int32_t>(p.static_network_duration);
    ctx->tab_max = std::min<int32_t>(1023, ((1 << 24) - 1) / D);
    const int32_t n = ctx->n_net * (ctx->tab_max + 1);
    if ((rc = dev_alloc(ctx, &ctx->d_p_tab, static_cast<size_t>(n)))) return rc;
    hipLaunchKernelGGL(p_table_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       ctx->d_p_tab, ctx->d_nets, ctx->n_net, ctx->tab_max, D);
    HIP_OR_FAIL(ctx, hipGetLastError());
  }
  ctx->hub_total_n = hub_rows_max * ctx->W * kWave;
  // measured defaults: 4-word groups up to 1024 replicas, 16-word groups above
  ctx->variant = ctx->W <= 16 ? REPLAY_V_NODE4 : REPLAY_V_NODE16;
  rc = do_reset(ctx);
  if (rc != REPLAY_OK) return rc;
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  if (p.flags & REPLAY_F_AUTOTUNE) {
    int32_t v;
    rc = replay_autotune(ctx, std::min<int32_t>(ctx->T, 32), &v);
    if (rc != REPLAY_OK) return rc;
  }
  return REPLAY_OK;
}

int open_common(replay_ctx *ctx, replay_ctx **out) {
  int rc = finish_open(ctx);
  if (rc != REPLAY_OK) {
    g_open_error = ctx->err;
    release(ctx);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return REPLAY_OK;
}

NetHost single_net(const replay_params *p) {
  NetHost nh;
  nh.M = p->M;
  nh.p0 = p->initial_infected;
  nh.beta = p->infect_prob;
  nh.Ed = p->exposed_duration;
  nh.Id = p->infectious_duration;
  nh.Rd = p->resistant_duration;
  nh.stream = p->stream;
  return nh;
}

// Device address of `scratch` on the context's device (current when
// called): looked up once per device with hipGetSymbolAddress and kept;
// later calls on that device reuse it.  Nothing is allocated.
int noalloc_scratch(replay_ctx *ctx, float **ptr) {
  static std::mutex mu;
  static std::vector<float *> addr;  // per device ordinal, null until looked up
  std::lock_guard<std::mutex> lock(mu);
  const int dev = ctx->p.device;
  if (dev < 0) return fail(ctx, REPLAY_E_ARG, "invalid device ordinal");
  if (static_cast<size_t>(dev) >= addr.size()) addr.resize(static_cast<size_t>(dev) + 1, nullptr);
  if (!addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ctx, REPLAY_E_DEVICE,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    addr[dev] = static_cast<float *>(p);
  }
  *ptr = addr[dev];
  return REPLAY_OK;
}

// History ring: derive entry[] (each replica's last infection) from the
// ring before a state read.
int rebuild_entries(replay_ctx *ctx) {
  if (!ctx->history || (ctx->entry_ready && ctx->rebuilt >= ctx->steps_done)) return REPLAY_OK;
  StepArgs a = step_args(ctx, 0);
  const int64_t nb = (ctx->NW + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(rebuild_entry_kernel, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0,
                     ctx->stream, a, ctx->rebuilt, ctx->steps_done, ctx->entry_ready ? 0 : 1);
  HIP_OR_FAIL(ctx, hipGetLastError());
  ctx->entry_ready = true;
  ctx->rebuilt = ctx->steps_done;
  return REPLAY_OK;
}

// Countdown of a replica after k iterations from its entry value (-1 never
// infected, 0 initial seed, 2j+1 infected at iteration j), by the network's
// tabulated sequences.
int32_t countdown_of(const NetHost &nh, int32_t s, int32_t k) {
  if (s == 0) return nh.lut_init[k];
  if (s > 0) return nh.lut_new[k - (s - 1) / 2];
  return 0;
}

// Captures the step and counts kernels of a full run (iterations 0..T-1
// from a reset state) into one executable graph.
int build_graph(replay_ctx *ctx) {
  const int32_t saved_cur = ctx->cur, saved_done = ctx->steps_done, saved_counts = ctx->counts_done,
                saved_tallied = ctx->tallied;
  ctx->cur = 0;
  ctx->steps_done = 0;
  ctx->counts_done = -1;
  ctx->tallied = 0;
  hipGraph_t g = nullptr;
  HIP_OR_FAIL(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_steps(ctx, ctx->T);
  if (rc == REPLAY_OK) rc = enqueue_counts(ctx);
  hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  ctx->cur = saved_cur;
  ctx->steps_done = saved_done;
  ctx->counts_done = saved_counts;
  ctx->tallied = saved_tallied;
  if (rc != REPLAY_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(ctx, REPLAY_E_DEVICE, std::string("capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    ctx->graph = nullptr;
    return fail(ctx, REPLAY_E_DEVICE, std::string("instantiate: ") + hipGetErrorString(e));
  }
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
  if (rc == REPLAY_OK)
    rc = check_durations(p->exposed_duration, p->infectious_duration, p->resistant_duration);
  if (rc != REPLAY_OK) return rc;
  replay_ctx *ctx = new (std::nothrow) replay_ctx();
  if (!ctx) return fail(nullptr, REPLAY_E_NOMEM, "out of host memory");
  ctx->p = *p;
  ctx->nets.push_back(single_net(p));
  std::string err;
  rc = rp::parse_file(graph_path, ctx->nets[0].net, err);
  if (rc != REPLAY_OK) {
    delete ctx;
    return fail(nullptr, rc, err);
  }
  return open_common(ctx, out);
}

int32_t replay_open_contacts(int32_t n_node_list, const int64_t *node_ids, int64_t n_contacts,
                             const int32_t *src, const int32_t *tgt, const int64_t *start,
                             const int64_t *end, const replay_params *p, replay_ctx **out) {
  if (!out) return fail(nullptr, REPLAY_E_ARG, "null argument");
  replay_network n;
  std::memset(&n, 0, sizeof n);
  n.n_node_list = n_node_list;
  n.node_ids = node_ids;
  n.n_contacts = n_contacts;
  n.src = src;
  n.tgt = tgt;
  n.start = start;
  n.end = end;
  if (p) {
    n.initial_infected = p->initial_infected;
    n.infect_prob = p->infect_prob;
    n.exposed_duration = p->exposed_duration;
    n.infectious_duration = p->infectious_duration;
    n.resistant_duration = p->resistant_duration;
    n.stream = p->stream;
  }
  return replay_open_batch(1, &n, p, out);
}

int32_t replay_open_batch(int32_t n_networks, const replay_network *nets, const replay_params *p,
                          replay_ctx **out) {
  if (!out || !nets || n_networks <= 0) return fail(nullptr, REPLAY_E_ARG, "null argument");
  *out = nullptr;
  int rc = check_params(p);
  if (rc != REPLAY_OK) return rc;
  for (int32_t n = 0; n < n_networks; ++n) {
    rc = check_durations(nets[n].exposed_duration, nets[n].infectious_duration,
                         nets[n].resistant_duration);
    if (rc != REPLAY_OK) return rc;
    if (nets[n].M < 0)
      return fail(nullptr, REPLAY_E_ARG, "network " + std::to_string(n) + ": M must be >= 0");
  }
  replay_ctx *ctx = new (std::nothrow) replay_ctx();
  if (!ctx) return fail(nullptr, REPLAY_E_NOMEM, "out of host memory");
  ctx->p = *p;
  ctx->nets.resize(static_cast<size_t>(n_networks));
  for (int32_t n = 0; n < n_networks; ++n) {
    const replay_network &d = nets[n];
    NetHost &nh = ctx->nets[n];
    nh.M = d.M > 0 ? d.M : p->M;
    nh.p0 = d.initial_infected;
    nh.beta = d.infect_prob;
    nh.Ed = d.exposed_duration;
    nh.Id = d.infectious_duration;
    nh.Rd = d.resistant_duration;
    nh.stream = d.stream;
    std::string err;
    rc = rp::from_contacts(d.n_node_list, d.node_ids, d.n_contacts, d.src, d.tgt, d.start, d.end,
                           nh.net, err);
    if (rc != REPLAY_OK) {
      delete ctx;
      return fail(nullptr, rc, "network " + std::to_string(n) + ": " + err);
    }
  }
  return open_common(ctx, out);
}

int32_t replay_get_network_offsets(replay_ctx *ctx, int32_t *node_off, int32_t n) {
  if (!ctx || !node_off || n < ctx->n_net + 1) return REPLAY_E_ARG;
  for (int32_t k = 0; k < ctx->n_net; ++k) node_off[k] = ctx->nets[k].node_off;
  node_off[ctx->n_net] = ctx->N;
  return REPLAY_OK;
}

int32_t replay_get_info(replay_ctx *ctx, replay_info *out) {
  if (!ctx || !out) return REPLAY_E_ARG;
  std::memset(out, 0, sizeof *out);
  out->n_nodes = ctx->N;
  out->M = ctx->M;
  out->replicas_total = ctx->G;
  out->n_contacts = ctx->n_contacts;
  out->nlists = ctx->nlists;
  out->iterations = ctx->T;
  out->n_windows = static_cast<int32_t>(ctx->win_ws.size());
  out->n_epochs = static_cast<int32_t>(ctx->epoch_t0.size());
  out->windows_per_epoch = ctx->windows_per_epoch;
  out->n_entries = ctx->n_entries;
  out->slice_bytes = ctx->n_entries * static_cast<int64_t>(sizeof(rp::Entry)) +
                     static_cast<int64_t>(ctx->epoch_t0.size()) * (ctx->N + 1) * 4;
  const int64_t NW = ctx->NW;
  const int64_t flag_bytes =
      ctx->compact ? static_cast<int64_t>(step_grid(ctx).x) * kWavesPerBlock * 8 * ctx->depthF : 0;
  out->state_bytes = NW * 8 * (4 + ctx->depth) + NW * kWave * 4 + flag_bytes;
  out->steps_done = ctx->steps_done;
  out->truncated_windows = ctx->truncated;
  // per word: S read + write, own I read, I_next write, X_k write, and one
  // ring read per flip offset kept in the ring (the first network's offsets);
  // plus the slice epoch's entries and row pointers, and the prefix update
  int ring_reads = 0;
  const NetHost &n0 = ctx->nets[0];
  for (int32_t n : {n0.n_on, n0.n_off, n0.n_s})
    ring_reads += (n != kNever && n >= 2 && n <= ctx->T) ? 1 : 0;
  const double ent = ctx->epoch_t0.empty() ? 0.0
                                           : static_cast<double>(ctx->n_entries) /
                                                 static_cast<double>(ctx->epoch_t0.size());
  // entries at the width the NODE4 / NODE16 kernels of whole groups stream
  const int base = ctx->variant & ~(REPLAY_V_BLOCK1 | REPLAY_V_XCD);
  const bool c8 = ctx->d_entries8 && ((base == REPLAY_V_NODE4 && ctx->Wg % 4 == 0) ||
                                      (base == REPLAY_V_NODE16 && ctx->Wg % 16 == 0));
  // compact form: a flag bit per word read per offset and one written, in
  // place of the ring reads (the entry reads of flagged words depend on the
  // epidemic and are not counted)
  const int64_t ring_term = ctx->compact ? (NW / 8) * (ring_reads + 1) : NW * 8 * ring_reads;
  out->step_bytes = NW * 8 * 5 + ring_term + static_cast<int64_t>((c8 ? 8.0 : 16.0) * ent) +
                    4 * (static_cast<int64_t>(ctx->N) + 1) +
                    12 * static_cast<int64_t>(ctx->G);
  out->ring_depth = ctx->depth;
  out->ring_form = ctx->history   ? REPLAY_RING_HISTORY
                  : ctx->compact ? REPLAY_RING_COMPACT
                                 : REPLAY_RING_OFFSETS;
  out->ring_bytes = ctx->compact ? static_cast<int64_t>(step_grid(ctx).x) * kWavesPerBlock * 8 *
                                       ctx->depthF
                                 : NW * 8 * ctx->depth;
  out->n_networks = ctx->n_net;
  out->variant = ctx->variant;
  out->graph_ready = ctx->graph != nullptr;
  out->entry_bytes = c8 ? 8 : 16;
  out->has_entries8 = ctx->d_entries8 != nullptr ? 1 : 0;
  for (int v = 0; v < 4; ++v) out->tune_ms[v] = ctx->tune_ms[v];
  return REPLAY_OK;
}

int32_t replay_reset(replay_ctx *ctx) {
  if (!ctx) return REPLAY_E_ARG;
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  return do_reset(ctx);
}

int32_t replay_run(replay_ctx *ctx, int32_t steps) {
  if (!ctx) return REPLAY_E_ARG;
  if (steps < 0 || ctx->steps_done + steps > ctx->T)
    return fail(ctx, REPLAY_E_ARG, "steps exceed the planned iterations");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  if ((ctx->p.flags & REPLAY_F_GRAPH) && ctx->steps_done == 0 && steps == ctx->T && steps > 0 &&
      !ctx->d_uniforms) {
    if (!ctx->graph) {
      int rc = build_graph(ctx);
      if (rc != REPLAY_OK) return rc;
    }
    HIP_OR_FAIL(ctx, hipGraphLaunch(ctx->graph, ctx->stream));
    ctx->steps_done = ctx->T;
    ctx->cur = ctx->T & 1;
    ctx->counts_done = ctx->T;
    ctx->tallied = ctx->T;
    return REPLAY_OK;
  }
  int rc = enqueue_steps(ctx, steps);
  if (rc != REPLAY_OK) return rc;
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

int32_t replay_autotune(replay_ctx *ctx, int32_t steps, int32_t *variant) {
  if (!ctx || !variant) return REPLAY_E_ARG;
  steps = std::min(steps, ctx->T);
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  // A graph captured with one variant would be replayed for every candidate
  // of a full run: drop it, and time eager launches (the flag is restored
  // on every exit; the next full run recaptures with the chosen variant).
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->graph) {
    (void)hipGraphExecDestroy(ctx->graph);
    ctx->graph = nullptr;
  }
  struct FlagRestore {
    replay_ctx *c;
    int32_t flags;
    ~FlagRestore() { c->p.flags = flags; }
  } restore{ctx, ctx->p.flags};
  ctx->p.flags &= ~REPLAY_F_GRAPH;
  // each variant, NODE4 and NODE16 also in blocks of one wavefront, and the
  // exposure kernel's variants with either XCD mapping; tune_ms[slot] keeps
  // the fastest schedule of a variant
  const int candidates[] = {REPLAY_V_WORD,   REPLAY_V_NODE4,
                            REPLAY_V_NODE4 | REPLAY_V_BLOCK1, REPLAY_V_NODE16,
                            REPLAY_V_NODE16 | REPLAY_V_BLOCK1, REPLAY_V_SLICED};
  const int slot_of[] = {0, 1, 1, 2, 2, 3};
  constexpr int kCand = static_cast<int>(sizeof candidates / sizeof candidates[0]);
  static_assert(sizeof slot_of / sizeof slot_of[0] == kCand, "one slot per candidate");
  for (int c = 0; c < 4; ++c) ctx->tune_ms[c] = -1.0f;
  int best = ctx->variant;
  double best_ms = 1e30;
  // every candidate with plain blocks, then with contiguous XCD eighths
  // (variant_applies rejects SLICED | REPLAY_V_XCD)
  for (int c = 0; c < 2 * kCand; ++c) {
    const int v = candidates[c % kCand] | (c >= kCand ? REPLAY_V_XCD : 0);
    if (!variant_applies(ctx, v) || steps <= 0) continue;
    ctx->variant = v;
    // untimed warm-up: the first launch of a kernel also loads its code
    int rc = do_reset(ctx);
    if (rc == REPLAY_OK) rc = replay_run(ctx, std::min<int32_t>(steps, 2));
    if (rc == REPLAY_OK) rc = do_reset(ctx);
    if (rc != REPLAY_OK) return rc;
    double ms = 0.0;
    rc = replay_run_timed(ctx, steps, &ms);
    if (rc != REPLAY_OK) return rc;
    float &t = ctx->tune_ms[slot_of[c % kCand]];
    if (t < 0.0f || static_cast<float>(ms) < t) t = static_cast<float>(ms);
    if (ms < best_ms) {
      best_ms = ms;
      best = v;
    }
  }
  ctx->variant = best;
  *variant = best;
  return do_reset(ctx);
}

int32_t replay_set_variant(replay_ctx *ctx, int32_t variant) {
  if (!ctx || !variant_applies(ctx, variant)) return REPLAY_E_ARG;
  if (ctx->variant != variant && ctx->graph) {
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    hipGraphExecDestroy(ctx->graph);
    ctx->graph = nullptr;
  }
  ctx->variant = variant;
  return REPLAY_OK;
}

int32_t replay_sync(replay_ctx *ctx) {
  if (!ctx) return REPLAY_E_ARG;
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  return REPLAY_OK;
}

int32_t replay_get_counts(replay_ctx *ctx, int32_t *out, int64_t n_out) {
  if (!ctx || !out) return REPLAY_E_ARG;
  const int64_t G = ctx->G;
  const int64_t n = static_cast<int64_t>(ctx->steps_done + 1) * G * 4;
  if (n_out < n) return fail(ctx, REPLAY_E_ARG, "counts buffer too small");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  int rc = enqueue_counts(ctx);
  if (rc) return rc;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(out, ctx->d_counts, static_cast<size_t>(n) * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<int32_t> nodes(static_cast<size_t>(G));  // nodes of each column's network
  for (const NetHost &nh : ctx->nets)
    std::fill(nodes.begin() + nh.moff, nodes.begin() + nh.moff + nh.M, nh.net.n_nodes);
  for (int64_t b = 0; b < n; b += 4)
    out[b + 3] = nodes[static_cast<size_t>((b / 4) % G)] - out[b] - out[b + 1] - out[b + 2];
  return REPLAY_OK;
}

int32_t replay_get_state(replay_ctx *ctx, int32_t *out, int64_t n_out) {
  if (!ctx || !out) return REPLAY_E_ARG;
  if (!ctx->uniform)
    return fail(ctx, REPLAY_E_UNSUPPORTED,
                "networks have different replica counts: use replay_get_network_state");
  const int64_t n = static_cast<int64_t>(ctx->N) * ctx->M;
  if (n_out < n) return fail(ctx, REPLAY_E_ARG, "state buffer too small");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  if (int rc = rebuild_entries(ctx)) return rc;
  // entry is [word][64 replicas]; with one replica count node i's words
  // start at i * W, so this is [node][W * 64]
  std::vector<int32_t> ent(static_cast<size_t>(ctx->NW) * kWave);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(ent.data(), ctx->d_entry, ent.size() * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  const int32_t k = ctx->steps_done;
  const size_t Mp = static_cast<size_t>(ctx->W) * kWave;
  for (const NetHost &nh : ctx->nets) {
    for (int32_t i = nh.node_off; i < nh.node_off + nh.net.n_nodes; ++i)
      for (int32_t r = 0; r < ctx->M; ++r)
        out[static_cast<size_t>(r) * ctx->N + i] =
            countdown_of(nh, ent[static_cast<size_t>(i) * Mp + r], k);
  }
  return REPLAY_OK;
}

int32_t replay_get_network_state(replay_ctx *ctx, int32_t net, int32_t *out, int64_t n_out) {
  if (!ctx || !out || net < 0 || net >= ctx->n_net) return REPLAY_E_ARG;
  const NetHost &nh = ctx->nets[net];
  const int32_t Nn = nh.net.n_nodes;
  const int64_t n = static_cast<int64_t>(Nn) * nh.M;
  if (n_out < n) return fail(ctx, REPLAY_E_ARG, "state buffer too small");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  if (int rc = rebuild_entries(ctx)) return rc;
  // the network's words are one run of Nn * W words from woff
  std::vector<int32_t> ent(static_cast<size_t>(Nn) * nh.W * kWave);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(ent.data(), ctx->d_entry + static_cast<size_t>(nh.woff) * kWave,
                                  ent.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                                  ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  const int32_t k = ctx->steps_done;
  const size_t Mp = static_cast<size_t>(nh.W) * kWave;
  for (int32_t i = 0; i < Nn; ++i)
    for (int32_t r = 0; r < nh.M; ++r)
      out[static_cast<size_t>(r) * Nn + i] = countdown_of(nh, ent[static_cast<size_t>(i) * Mp + r], k);
  return REPLAY_OK;
}

int32_t replay_get_replica_offsets(replay_ctx *ctx, int32_t *rep_off, int32_t n) {
  if (!ctx || !rep_off || n < ctx->n_net + 1) return REPLAY_E_ARG;
  for (int32_t k = 0; k < ctx->n_net; ++k) rep_off[k] = ctx->nets[k].moff;
  rep_off[ctx->n_net] = ctx->G;
  return REPLAY_OK;
}

int64_t replay_iteration_time(replay_ctx *ctx, int32_t k) {
  if (!ctx) return 0;
  if (k < 0 || k >= static_cast<int32_t>(ctx->iter_time.size())) return ctx->p.start_time;
  return ctx->iter_time[k];
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
    if (!ctx->uniform)
      return fail(ctx, REPLAY_E_UNSUPPORTED,
                  "an injected stream needs every network to have the same replica count");
    const int64_t need = static_cast<int64_t>(ctx->T + 1) * ctx->M * ctx->N;
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
  if (!ctx || !rowptr || !nnz || k < 0 || k >= ctx->T) return REPLAY_E_ARG;
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  StepArgs a = step_args(ctx, k);
  int32_t *d_rp = nullptr, *d_cnt = nullptr, *d_col = nullptr;
  float *d_val = nullptr;
  DeviceScratch scratch;
  int rc;
  if ((rc = dev_alloc(ctx, &d_cnt, static_cast<size_t>(ctx->N)))) return rc;
  scratch.bufs.push_back(d_cnt);
  if ((rc = dev_alloc(ctx, &d_rp, static_cast<size_t>(ctx->N) + 1))) return rc;
  scratch.bufs.push_back(d_rp);
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
  if (cap < 0) return REPLAY_OK;  // count only
  if (*nnz > cap || (*nnz > 0 && (!colind || !val)))
    return fail(ctx, REPLAY_E_ARG, "slice buffer too small");
  if ((rc = dev_alloc(ctx, &d_col, static_cast<size_t>(*nnz)))) return rc;
  scratch.bufs.push_back(d_col);
  if ((rc = dev_alloc(ctx, &d_val, static_cast<size_t>(*nnz)))) return rc;
  scratch.bufs.push_back(d_val);
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
  return REPLAY_OK;
}

int32_t replay_get_exposure(replay_ctx *ctx, int32_t k, float *out, int64_t n_out) {
  if (!ctx || !out || k < 0 || k >= ctx->T) return REPLAY_E_ARG;
  if (!ctx->uniform)
    return fail(ctx, REPLAY_E_UNSUPPORTED, "networks have different replica counts");
  const int64_t n = static_cast<int64_t>(ctx->N) * ctx->M;
  if (n_out < n) return fail(ctx, REPLAY_E_ARG, "exposure buffer too small");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  StepArgs a = step_args(ctx, k);
  float *d_out = nullptr;
  int rc = dev_alloc(ctx, &d_out, static_cast<size_t>(n));
  if (rc) return rc;
  DeviceScratch scratch{{d_out}};
  dim3 grid((ctx->N + kWavesPerBlock - 1) / kWavesPerBlock, ctx->W);
  if (ctx->win_int[ctx->iter_window[k]])
    hipLaunchKernelGGL(exposure_kernel<true>, grid, dim3(kBlock), 0, ctx->stream, a, d_out);
  else
    hipLaunchKernelGGL(exposure_kernel<false>, grid, dim3(kBlock), 0, ctx->stream, a, d_out);
  HIP_OR_FAIL(ctx, hipGetLastError());
  HIP_OR_FAIL(ctx, hipMemcpyAsync(out, d_out, static_cast<size_t>(n) * sizeof(float),
                                  hipMemcpyDeviceToHost, ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  return REPLAY_OK;
}

int32_t replay_get_exposure_noalloc(replay_ctx *ctx, int32_t k, float *out, int64_t n_out) {
  if (!ctx || !out || k < 0 || k >= ctx->T) return REPLAY_E_ARG;
  if (!ctx->uniform)
    return fail(ctx, REPLAY_E_UNSUPPORTED, "networks have different replica counts");
  const int64_t n = static_cast<int64_t>(ctx->N) * ctx->M;
  if (n_out < n) return fail(ctx, REPLAY_E_ARG, "exposure buffer too small");
  if (n > kNoAllocFloats)
    return fail(ctx, REPLAY_E_ARG,
                "N * M = " + std::to_string(n) + " exceeds the no-allocation staging limit of " +
                    std::to_string(kNoAllocFloats) + " floats (64 MiB); use replay_get_exposure");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->p.device));
  float *d_out = nullptr;
  const int rc = noalloc_scratch(ctx, &d_out);
  if (rc != REPLAY_OK) return rc;
  StepArgs a = step_args(ctx, k);
  dim3 grid((ctx->N + kWavesPerBlock - 1) / kWavesPerBlock, ctx->W);
  // every (node, sim < M) element is written: nothing of an earlier call's
  // contents is read back
  if (ctx->win_int[ctx->iter_window[k]])
    hipLaunchKernelGGL(exposure_kernel<true>, grid, dim3(kBlock), 0, ctx->stream, a, d_out);
  else
    hipLaunchKernelGGL(exposure_kernel<false>, grid, dim3(kBlock), 0, ctx->stream, a, d_out);
  HIP_OR_FAIL(ctx, hipGetLastError());
  HIP_OR_FAIL(ctx, hipMemcpyAsync(out, d_out, static_cast<size_t>(n) * sizeof(float),
                                  hipMemcpyDeviceToHost, ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
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
