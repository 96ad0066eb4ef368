// device_math.hpp -- Philox4x32-10 and the project exponential, for gfx950.
//
// Philox4x32-10 follows Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC'11): ten rounds of two 32x32->64
// multiplies (0xD2511F53, 0xCD9E8D57) with Weyl key increments
// (0x9E3779B9, 0xBB67AE85).
//
// rp_expf evaluates e^x in double and rounds once to float: Cody-Waite
// reduction x = k ln2 + r with |r| <= ln2/2, the degree-13 Taylor polynomial
// of e^r by fused multiply-adds, scaling by 2^k.  Every step is an IEEE
// operation with one rounding (rint, fma, ldexp, convert), so host and device
// builds of the same sequence agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rp {

__device__ __forceinline__ uint32_t philox_word0(uint64_t seed, uint32_t c0, uint32_t c1,
                                                 uint32_t c2, uint32_t c3) {
  uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c0;
}

// All four output words of Philox4x32-10.  Each round's two 32x32->64
// products are written as 64-bit multiplies of zero-extended words, which
// gfx950 executes as one v_mad_u64_u32 each (high and low halves together).
__device__ __forceinline__ void philox4(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2,
                                        uint32_t c3, uint32_t (&out)[4]) {
  uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
  // The key is the same for every lane and lives in SGPRs.  Made opaque
  // here, its schedule is computed by the scalar unit where a block is
  // drawn, instead of being hoisted to the kernel's start and held in VGPR
  // lanes until use (a spill and a reload, both vector instructions, per
  // round key).
  asm volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t lo0 = static_cast<uint32_t>(p0), hi0 = static_cast<uint32_t>(p0 >> 32);
    const uint32_t lo1 = static_cast<uint32_t>(p1), hi1 = static_cast<uint32_t>(p1 >> 32);
    // Three-way XORs as one v_bitop3_b32 (truth table 0x96) once both
    // counter words feeding them vary across the wavefront, which they do
    // from round 4 (c0 from round 3) when only c1 varies, as in the draws.
    // Earlier, a word that is still wavefront-uniform is combined with the
    // key by the scalar unit, and one vector XOR finishes the job; a bitop3
    // there would read two scalar operands, more than one instruction may.
    c0 = r >= 3 ? __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96) : hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = r >= 4 ? __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96) : hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

__device__ __forceinline__ float u24(uint32_t w) {
  return static_cast<float>(w >> 8) * 0x1p-24f;
}

// Uniform in [0, 1) of replica `sim`: one Philox block serves four replicas
// 64 apart, replica r taking word (r >> 6) & 3 of counter
// {node, (r >> 8) * 64 + (r & 63), slot, stream}.
__device__ __forceinline__ float philox_uniform(uint64_t seed, uint32_t node, uint32_t sim,
                                                uint32_t slot, uint32_t stream) {
  uint32_t w[4];
  philox4(seed, node, ((sim >> 8) << 6) | (sim & 63u), slot, stream, w);
  const uint32_t k = (sim >> 6) & 3u;
  return u24(k == 0 ? w[0] : k == 1 ? w[1] : k == 2 ? w[2] : w[3]);
}

// A VGPR written with kBits where it is used.  gfx950's VOP3 encodings take
// no literal, so a double coefficient other than 0.5, 1, 2 or 4 must sit in a
// register pair; written by an instruction with side effects it is neither
// hoisted out of loops nor kept live across the code between its uses (the
// exponential's coefficients would otherwise hold about 20 VGPRs for the
// whole of a kernel that calls it).
template <uint32_t kBits>
__device__ __forceinline__ uint32_t vgpr_const() {
  uint32_t v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "i"(kBits));
  return v;
}

template <uint64_t kBits>
__device__ __forceinline__ double dconst() {
  const uint64_t lo = vgpr_const<static_cast<uint32_t>(kBits)>();
  const uint64_t hi = vgpr_const<static_cast<uint32_t>(kBits >> 32)>();
  return __builtin_bit_cast(double, (hi << 32) | lo);
}

#define RP_DCONST(x) dconst<__builtin_bit_cast(uint64_t, static_cast<double>(x))>()

__device__ __forceinline__ float rp_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return __builtin_huge_valf();
  if (x < -104.0f) return 0.0f;
  const double xd = static_cast<double>(x);
  const double kd = rint(xd * RP_DCONST(1.4426950408889634074));
  double r = fma(-kd, RP_DCONST(6.93147180369123816490e-01), xd);
  r = fma(-kd, RP_DCONST(1.90821492927058770002e-10), r);
  double p = RP_DCONST(1.0 / 6227020800.0);
  p = fma(p, r, RP_DCONST(1.0 / 479001600.0));
  p = fma(p, r, RP_DCONST(1.0 / 39916800.0));
  p = fma(p, r, RP_DCONST(1.0 / 3628800.0));
  p = fma(p, r, RP_DCONST(1.0 / 362880.0));
  p = fma(p, r, RP_DCONST(1.0 / 40320.0));
  p = fma(p, r, RP_DCONST(1.0 / 5040.0));
  p = fma(p, r, RP_DCONST(1.0 / 720.0));
  p = fma(p, r, RP_DCONST(1.0 / 120.0));
  p = fma(p, r, RP_DCONST(1.0 / 24.0));
  p = fma(p, r, RP_DCONST(1.0 / 6.0));
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return static_cast<float>(ldexp(p, static_cast<int>(kd)));
}

#undef RP_DCONST

}  // namespace rp
