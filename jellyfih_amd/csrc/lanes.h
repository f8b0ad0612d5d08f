// Cross-lane helpers of the LDS scatters (fused engine, differentiable forward, renderer):
// wave reductions, groups of same-cell lanes that add their contributions before the LDS
// atomics, and the 64-bit fixed-point encoding of those atomics. Device code for gfx950
// (wave64, DPP rows of 16).
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>

namespace jm {

// Minimum over the 64 lanes (all active), uniform in the wave: DPP row_shr by 1, 2, 4, 8
// leaves each row's minimum in its last lane (a lane with no source keeps its own value),
// then the four row results combine through readlane.
__device__ __forceinline__ int wave_min(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x111, 0xF, 0xF, false));  // row_shr:1
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x112, 0xF, 0xF, false));  // row_shr:2
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x114, 0xF, 0xF, false));  // row_shr:4
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x118, 0xF, 0xF, false));  // row_shr:8
  return min(min(__builtin_amdgcn_readlane(v, 15), __builtin_amdgcn_readlane(v, 31)),
             min(__builtin_amdgcn_readlane(v, 47), __builtin_amdgcn_readlane(v, 63)));
}

// Lanes of one DPP row whose particles share a base cell form runs (cell-sorted particles
// sit in adjacent lanes; a lane that scatters nothing through LDS carries a unique key).
// Each run adds its contributions in two steps: lane l takes lane l + 1 when it shares
// l's cell (f1), then lane l + 2 (holding l + 2 and l + 3) when l + 1 and l + 2 both
// share it (f2: between sorts a cell can recur after a gap, A B A A, and the second A run
// emits by itself). The first lane of every group of 4 in a run then emits. Called by
// every lane of the wave.
struct LaneGroups {
  float f1, f2;
  bool emit;
};

__device__ __forceinline__ LaneGroups lane_groups(int key, bool scatters) {
  const int lane = threadIdx.x & 63;
  const unsigned long long lanes_le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int k1 = __builtin_amdgcn_update_dpp(INT_MIN, key, 0x101, 0xF, 0xF, false);  // l+1
  const int k2 = __builtin_amdgcn_update_dpp(INT_MIN, key, 0x102, 0xF, 0xF, false);  // l+2
  const int kp = __builtin_amdgcn_update_dpp(INT_MIN, key, 0x111, 0xF, 0xF, false);  // l-1
  const unsigned long long heads = __ballot(kp != key);
  const int runstart = 63 - __clzll(heads & lanes_le);
  LaneGroups g;
  g.f1 = k1 == key ? 1.0f : 0.0f;
  g.f2 = (k1 == key && k2 == key) ? 1.0f : 0.0f;
  g.emit = scatters && (((lane - runstart) & 3) == 0);
  return g;
}

// The two pre-sum steps on three values: c += (c of lane + 1) * f1, then c += (c of
// lane + 2) * f2, within DPP rows (bound_ctrl: a lane past the end of its row reads 0).
// One v_fmac_f32_dpp per value and step (the compiler keeps a separate v_mov_b32_dpp
// before each FMA). Runs with every lane active, on finite values (f = 0 must cancel).
// The leading s_nop 1 gives the two wait states a DPP read of a VGPR needs after a VALU
// write of it (the three values are written just before); inside the block each DPP
// source was written at least two instructions earlier.
__device__ __forceinline__ void presum3(float& c0, float& c1, float& c2, float f1, float f2) {
  asm volatile(
      "s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %0, %3 row_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f32_dpp %1, %1, %3 row_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f32_dpp %2, %2, %3 row_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f32_dpp %0, %0, %4 row_shl:2 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f32_dpp %1, %1, %4 row_shl:2 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f32_dpp %2, %2, %4 row_shl:2 row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "+v"(c0), "+v"(c1), "+v"(c2)
      : "v"(f1), "v"(f2));
}

// presum3's two steps on two values. With two values the second step's DPP source was
// written one instruction earlier, so an s_nop 0 completes its two wait states.
__device__ __forceinline__ void presum2(float& c0, float& c1, float f1, float f2) {
  asm volatile(
      "s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %0, %2 row_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f32_dpp %1, %1, %2 row_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 0\n\t"
      "v_fmac_f32_dpp %0, %0, %3 row_shl:2 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f32_dpp %1, %1, %3 row_shl:2 row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "+v"(c0), "+v"(c1)
      : "v"(f1), "v"(f2));
}

// The same steps in fp64: each lane's value moves as two 32-bit DPP moves (a lane past the
// end of its DPP row reads 0), then one FMA.
template <int CTRL>
__device__ __forceinline__ double dpp_row_shl(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void presum2(double& c0, double& c1, float f1, float f2) {
  c0 += double(f1) * dpp_row_shl<0x101>(c0);
  c1 += double(f1) * dpp_row_shl<0x101>(c1);
  c0 += double(f2) * dpp_row_shl<0x102>(c0);
  c1 += double(f2) * dpp_row_shl<0x102>(c1);
}

// LDS scatter sums are 64-bit fixed point with 50 fractional bits: integer LDS atomics
// run at full rate on gfx950 where fp32 LDS atomics do not, and integer sums are exact and
// independent of the order of the contributions. fixed_bits encodes |c| < 2 exactly (below);
// with encoded values below 2 and at most 4096 of them per node, every node sum fits int64
// (partial sums may wrap; the final two's-complement sum is still exact).
constexpr double kFixInv = 1.0 / 1125899906842624.0;   // 2^-50
constexpr float kFixPre = 1125899906842624.0f;          // 2^50

// y = c * 2^50 (|y| < 2^51; callers pre-scale their per-particle factors by 2^50, which is
// exact) rounded to the nearest integer, as two's-complement int64 bits, in three VALU
// operations: y + 1.5 * 2^52 in double is exact up to that rounding and lands in
// [2^52, 2^53), where the 52 mantissa bits hold y + 2^51; the low word is then y's low word
// and the high word, minus the exponent and the 2^51 bit (0x43380000), is y's high word.
__device__ __forceinline__ unsigned long long fixed_bits(float y) {
  const unsigned long long b = __double_as_longlong((double)y + 6755399441055744.0);
  const unsigned hi = (unsigned)(b >> 32) - 0x43380000u;
  return ((unsigned long long)hi << 32) | (unsigned)b;
}

__device__ __forceinline__ float from_fixed_bits(unsigned long long v) {
  return (float)((double)(long long)v * kFixInv);
}

// The same encoding without the high-word correction: the bits of y + 1.5 * 2^52 are
// kFixBias + y as an int64, so a sum of n such values is n * kFixBias + sum(y). Two VALU
// operations per value. When one of the summed planes holds non-negative values whose sum
// stays below 2^48 (and n < 8192), its top bits hold n * 0x4338 mod 2^16, from which
// biased_count recovers n; from_biased_bits then removes n * kFixBias from every plane that
// received the same n values.
constexpr unsigned long long kFixBias = 0x4338000000000000ull;

__device__ __forceinline__ unsigned long long fixed_bits_biased(float y) {
  return __double_as_longlong((double)y + 6755399441055744.0);
}

// n from a plane of non-negative biased values: (v >> 51) = n * 2151 mod 2^13 (0x4338 =
// 8 * 2151), and 4951 is the inverse of 2151 modulo 2^13.
__device__ __forceinline__ unsigned biased_count(unsigned long long v) {
  return ((unsigned)(v >> 51) * 4951u) & 8191u;
}

__device__ __forceinline__ float from_biased_bits(unsigned long long v, unsigned n) {
  return from_fixed_bits(v - ((unsigned long long)(n * 0x43380000u) << 32));
}

// The same encoding with 40 fractional bits, for sums whose terms reach ~2^11 (frame
// pixels): every pixel sum fits int64 with room to spare, at 2^-40 resolution.
constexpr float kFix40Pre = 1099511627776.0f;          // 2^40
constexpr double kFix40Inv = 1.0 / 1099511627776.0;    // 2^-40

__device__ __forceinline__ float from_fixed40_bits(unsigned long long v) {
  return (float)((double)(long long)v * kFix40Inv);
}

}  // namespace jm
