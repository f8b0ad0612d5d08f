// HIP kernels for gfx950 (MI355X): plain per-stage path, fused G2P->P2G path, cell sort,
// statistics. Physics lives in physics.h; this file is about data movement.
//
// Device layout
//   particles: structure of arrays in per-instance segments of `cap` slots; instance m's
//     live particles occupy [m*cap, m*cap + n_live[m]) in cell-sorted order.
//     x, v, fiber: float2; C, F: float4 (row-major 2x2); Jp, glow: float; mat, slot: int.
//   grids: per instance three planes (px, py, mass) of nx*ny floats, node (i, j) at j*nx + i.
//     The fused path keeps three such grid buffers and rotates them: each substep gathers
//     from one, scatters into the next and zeroes the third.
#pragma once

#include <limits.h>

#include "lanes.h"
#include "physics.h"

namespace jm {

struct State {
  float2* x;
  float2* v;
  float4* C;
  float4* F;
  float* Jp;
  int* mat;
  float2* fiber;
  float* glow;
  int* slot;
  float* r_ref;  // spawn radius per particle (axisymmetric variant only, else null)
};

__device__ __forceinline__ Mat2 as_mat(float4 a) { return {a.x, a.y, a.z, a.w}; }
__device__ __forceinline__ float4 as_f4(const Mat2& m) {
  return make_float4(m.a00, m.a01, m.a10, m.a11);
}

// ---------------------------------------------------------------------------------------
// Small utility kernels (graph-safe replacements for memset).

__global__ void k_fill_f32(float* p, int64_t n, float v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_i32(int* p, int64_t n, int v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Reset the dynamic state of particles [begin, end): v=0, F=I, C=0, Jp=1, glow=0
// (mpm_sim.py:722-725).
__global__ void k_init_particles(State s, int64_t begin, int64_t end) {
  for (int64_t i = begin + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < end;
       i += (int64_t)gridDim.x * blockDim.x) {
    s.v[i] = make_float2(0.f, 0.f);
    s.C[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    s.F[i] = make_float4(1.f, 0.f, 0.f, 1.f);
    s.Jp[i] = 1.f;
    if (s.glow) s.glow[i] = 0.f;
    if (s.r_ref) s.r_ref[i] = s.x[i].x;  // mpm_sim.py:727
  }
}

// ---------------------------------------------------------------------------------------
// Plain path: one launch per stage, global atomics, any particle order.

struct PlainArgs {
  State s;
  InstParams ip;
  float* acc;          // one grid buffer: [m][3][G]
  float* hoop;         // axisymmetric variant: [m][2][G] mass-weighted s_rr, s_tt; else null
  float* t_buf;        // t_buf[0] = sim_time
  const int* n_live;
  int cap;
  int m0;              // first instance of this launch (blockIdx.y adds to it)
  int G;
  int glow;
  PhysConsts k;
};

__global__ void __launch_bounds__(256) k_plain_p2g(PlainArgs a) {
  const int m = a.m0 + blockIdx.y;
  const int local = blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= a.n_live[m]) return;
  const PhysConsts& k = a.k;
  const size_t i = (size_t)m * a.cap + local;
  const int mt = a.s.mat[i];
  if (mt < 0) return;
  float2 x = a.s.x[i];
  int b0, b1;
  float fx0, fx1;
  base_frac(x.x, k.inv_dx, b0, fx0);
  base_frac(x.y, k.inv_dx, b1, fx1);
  if (!in_range(b0, b1, k)) return;
  const float act = activation(a.t_buf[0], a.ip.freq[m], a.ip.contraction[m], k);
  float2 fib = make_float2(0.f, 1.f);
  if (mt == 3) fib = a.s.fiber[i];
  const Mat2 C = as_mat(a.s.C[i]);
  P2GOut o = p2g_update(mt, as_mat(a.s.F[i]), C, fib.x, fib.y, a.ip.amp[m] * act, a.ip.mu_j[m],
                        a.ip.la_j[m], a.ip.pdens[m], k);
  a.s.F[i] = as_f4(o.F);
  if (o.jp_changes) a.s.Jp[i] = (a.s.Jp[i] * o.q0) * o.q1;
  const float2 v = a.s.v[i];
  const float mv0 = o.mass * v.x, mv1 = o.mass * v.y;
  float wx[3], wy[3];
  bspline(fx0, wx);
  bspline(fx1, wy);
  float* px = a.acc + (size_t)m * 3 * a.G;
  float* py = px + a.G;
  float* pm = py + a.G;
#pragma unroll
  for (int ii = 0; ii < 3; ++ii) {
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
      const float d0 = ((float)ii - fx0) * k.dx, d1 = ((float)jj - fx1) * k.dx;
      const float w = wx[ii] * wy[jj];
      const int node = (b1 + jj) * k.nx + (b0 + ii);
      unsafeAtomicAdd(px + node, w * (mv0 + (o.affine.a00 * d0 + o.affine.a01 * d1)));
      unsafeAtomicAdd(py + node, w * (mv1 + (o.affine.a10 * d0 + o.affine.a11 * d1)));
      unsafeAtomicAdd(pm + node, w * o.mass);
    }
  }
}

// LDS accumulators of an axisymmetric P2G workgroup's node box: 5 planes x up to 819
// nodes, 64-bit fixed point (32 KB).
constexpr int kAxiTile = 4096;

// P2G of the axisymmetric variant (mpm_sim.py:167-271 with axisym, 236-247,258-267), one
// launch for the batch (blockIdx.y = instance): each particle is a ring of radius r = x, so
// mass and particle momentum are weighted by r, and two more grids gather the mass-weighted
// Kirchhoff s_rr and the hoop stress of the circumferential stretch r / r_ref. The 256
// particles of a workgroup are consecutive in the engine's order (fill order, or cell order
// after a sort) and reach a small box of nodes: when the box's five planes fit kAxiTile,
// the workgroup adds into an LDS copy, as 64-bit fixed point with 50 fractional bits as in
// k_fused (integer LDS atomics run at full rate on gfx950, fp32 ones do not; the r-weighted
// node masses near the axis are ~1e-8, so they need the fine resolution), and then adds
// each non-zero box value to the grid once with an fp32 atomic; a particle whose
// contributions could reach 1 in magnitude (or are not finite: a diverged state) and every
// particle of a box that does not fit add to the grid directly. The box bounds are wave minima
// combined by one LDS atomic per wave.
__global__ void __launch_bounds__(256) k_axisym_p2g(PlainArgs a) {
  __shared__ unsigned long long tile[kAxiTile];
  __shared__ int box[4];  // min node column, -max, min node row, -max
  const int m = a.m0 + blockIdx.y;
  const int local = blockIdx.x * blockDim.x + threadIdx.x;
  const PhysConsts& k = a.k;
  bool act = false;
  int b0 = 0, b1 = 0;
  float fx0 = 0.f, fx1 = 0.f;
  size_t i = 0;
  int mt = -1;
  float2 x = make_float2(0.f, 0.f);
  if (local < a.n_live[m]) {
    i = (size_t)m * a.cap + local;
    mt = a.s.mat[i];
    if (mt >= 0) {
      x = a.s.x[i];
      base_frac(x.x, k.inv_dx, b0, fx0);
      base_frac(x.y, k.inv_dx, b1, fx1);
      act = in_range(b0, b1, k);
    }
  }
  P2GOut o{};
  float khoop = 0.f, mr = 0.f;
  float2 v = make_float2(0.f, 0.f);
  if (act) {
    const float actv = activation(a.t_buf[0], a.ip.freq[m], a.ip.contraction[m], k);
    float2 fib = make_float2(0.f, 1.f);
    if (mt == 3) fib = a.s.fiber[i];
    o = p2g_update(mt, as_mat(a.s.F[i]), as_mat(a.s.C[i]), fib.x, fib.y, a.ip.amp[m] * actv,
                   a.ip.mu_j[m], a.ip.la_j[m], a.ip.pdens[m], k);
    a.s.F[i] = as_f4(o.F);
    if (o.jp_changes) a.s.Jp[i] = (a.s.Jp[i] * o.q0) * o.q1;
    v = a.s.v[i];
    const float rp = x.x;
    const float rr = fmaxf(a.s.r_ref[i], k.r_ref_floor);
    const float lam = fmaxf(rp, k.r_floor) / rr;
    khoop = ((2.0f * o.mu) * (lam - 1.0f)) + ((o.la * o.J) * (o.J - 1.0f));
    mr = o.mass * rp;
  }
  if (threadIdx.x < 4) box[threadIdx.x] = INT_MAX;
  __syncthreads();
  {
    const int w0 = wave_min(act ? b0 : INT_MAX), w1 = wave_min(act ? -(b0 + 2) : INT_MAX);
    const int w2 = wave_min(act ? b1 : INT_MAX), w3 = wave_min(act ? -(b1 + 2) : INT_MAX);
    if ((threadIdx.x & 63) == 0 && w0 != INT_MAX) {
      atomicMin(&box[0], w0);
      atomicMin(&box[1], w1);
      atomicMin(&box[2], w2);
      atomicMin(&box[3], w3);
    }
  }
  __syncthreads();
  const int bx0 = box[0], by0 = box[2];
  if (bx0 == INT_MAX) return;  // no in-range particle in this workgroup (uniform)
  const int bw = -box[1] - bx0 + 1, area = bw * (-box[3] - by0 + 1);
  const bool in_lds = 5 * area <= kAxiTile;  // uniform
  if (in_lds) {
    for (int q = threadIdx.x; q < 5 * area; q += blockDim.x) tile[q] = 0ull;
    __syncthreads();
  }
  float* px = a.acc + (size_t)m * 3 * a.G;
  float* psr = a.hoop + (size_t)m * 2 * a.G;
  if (act) {
    // |w| <= 1 and |d| <= 1.5 dx bound every contribution by this sum, which a NaN or
    // infinity anywhere turns into NaN or inf, failing the comparison (no fmaxf: it drops
    // NaN operands)
    const float aff = fabsf(o.affine.a00) + fabsf(o.affine.a01) + fabsf(o.affine.a10) +
                      fabsf(o.affine.a11);
    const float bnd = fabsf(mr) * (fabsf(v.x) + fabsf(v.y)) + aff * (1.5f * k.dx) +
                      o.mass * (fabsf(x.x) + fabsf(o.s00) + fabsf(khoop));
    const bool to_lds = in_lds && bnd < 1.0f;
    float wx[3], wy[3];
    bspline(fx0, wx);
    bspline(fx1, wy);
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        const float d0 = ((float)ii - fx0) * k.dx, d1 = ((float)jj - fx1) * k.dx;
        const float w = wx[ii] * wy[jj];
        const float wm = w * o.mass;
        const float c[5] = {w * (mr * v.x + (o.affine.a00 * d0 + o.affine.a01 * d1)),
                            w * (mr * v.y + (o.affine.a10 * d0 + o.affine.a11 * d1)),
                            wm * x.x, wm * o.s00, wm * khoop};
        if (to_lds) {
          const int q = (b1 + jj - by0) * bw + (b0 + ii - bx0);
#pragma unroll
          for (int pl = 0; pl < 5; ++pl)
            atomicAdd(&tile[pl * area + q], fixed_bits(c[pl] * kFixPre));
        } else {
          const int node = (b1 + jj) * k.nx + (b0 + ii);
#pragma unroll
          for (int pl = 0; pl < 3; ++pl) unsafeAtomicAdd(px + pl * a.G + node, c[pl]);
          unsafeAtomicAdd(psr + node, c[3]);
          unsafeAtomicAdd(psr + a.G + node, c[4]);
        }
      }
    }
  }
  if (!in_lds) return;
  __syncthreads();
  for (int q = threadIdx.x; q < area; q += blockDim.x) {
    const int node = (by0 + q / bw) * k.nx + bx0 + q % bw;
#pragma unroll
    for (int pl = 0; pl < 5; ++pl) {
      const float val = from_fixed_bits(tile[pl * area + q]);
      if (val != 0.f) {
        float* dst = pl < 3 ? px + pl * a.G : psr + (pl - 3) * a.G;
        unsafeAtomicAdd(dst + node, val);
      }
    }
  }
}

// Grid op in place: momentum planes become velocity (mpm_sim.py:274-309).
__global__ void __launch_bounds__(256) k_plain_gridop(PlainArgs a) {
  const int m = a.m0 + blockIdx.y;
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= a.G) return;
  float* px = a.acc + (size_t)m * 3 * a.G;
  float* py = px + a.G;
  const float* pm = py + a.G;
  const int j = node / a.k.nx, i = node - j * a.k.nx;
  float2 g;
  if (a.hoop) {
    const float* psr = a.hoop + (size_t)m * 2 * a.G;
    g = grid_node_axisym(px[node], py[node], pm[node], psr[node], psr[a.G + node], i, j, a.k);
  } else {
    g = grid_node(px[node], py[node], pm[node], i, j, a.k);
  }
  px[node] = g.x;
  py[node] = g.y;
}

// G2P, gravity, advection, clamp (mpm_sim.py:311-333) and the optional glow pass (:335-338).
__global__ void __launch_bounds__(256) k_plain_g2p(PlainArgs a) {
  const int m = a.m0 + blockIdx.y;
  const int local = blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= a.n_live[m]) return;
  const PhysConsts& k = a.k;
  const size_t i = (size_t)m * a.cap + local;
  if (a.s.mat[i] < 0) return;
  float2 x = a.s.x[i];
  float2 v = a.s.v[i];
  int b0, b1;
  float fx0, fx1;
  base_frac(x.x, k.inv_dx, b0, fx0);
  base_frac(x.y, k.inv_dx, b1, fx1);
  if (in_range(b0, b1, k)) {
    const float* px = a.acc + (size_t)m * 3 * a.G;
    const float* py = px + a.G;
    float wx[3], wy[3];
    bspline(fx0, wx);
    bspline(fx1, wy);
    float nv0 = 0.f, nv1 = 0.f;
    Mat2 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        const int node = (b1 + jj) * k.nx + (b0 + ii);
        const float g0 = px[node], g1 = py[node];
        const float d0 = (float)ii - fx0, d1 = (float)jj - fx1;
        const float w = wx[ii] * wy[jj];
        nv0 += w * g0;
        nv1 += w * g1;
        const float ws = k.apic_scale * w;
        c.a00 += ws * (g0 * d0);
        c.a01 += ws * (g0 * d1);
        c.a10 += ws * (g1 * d0);
        c.a11 += ws * (g1 * d1);
      }
    }
    v = make_float2(nv0, nv1);
    a.s.C[i] = as_f4(c);
  }
  advect(x, v, k);
  a.s.v[i] = v;
  a.s.x[i] = x;
  if (a.glow) a.s.glow[i] = fmaxf(a.s.glow[i] * 0.9f, fabsf(1.0f - a.s.Jp[i]));
}

// sim_time += dt (mpm_sim.py:340)
__global__ void k_time_advance(float* t_buf, float dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) t_buf[0] = t_buf[0] + dt;
}

// n substeps' worth of sim_time += dt, one fp32 addition each (mpm_sim.py:340): the clock of
// a fused-path batch with no live particles, where no substep kernel runs.
__global__ void k_time_advance_n(float* t_buf, float dt, int64_t n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float t = t_buf[0];
  for (int64_t s = 0; s < n; ++s) t = t + dt;
  t_buf[0] = t;
}

// ---------------------------------------------------------------------------------------
// Fused path. One launch per substep: gather the previous substep's grid through LDS (the
// grid op is applied on the fly while staging), G2P + advection, then P2G of the next
// substep scattered into an LDS window and flushed with one global atomic per touched
// node. DO_G2P=false is the block's first kernel (P2G only), DO_P2G=false its last.

struct FusedArgs {
  State s;
  InstParams ip;
  const float* acc_r;  // gather buffer (null when !DO_G2P)
  float* acc_w;        // scatter buffer (null when !DO_P2G)
  float* acc_z;        // buffer zeroed for the substep after next (may be null)
  // axisymmetric variant: each grid buffer's two stress planes ([m][2][G], mass-weighted
  // s_rr and s_tt), gathered, scattered and zeroed with acc_r, acc_w, acc_z; else null
  const float* hoop_r;
  float* hoop_w;
  float* hoop_z;
  float* t_buf;
  int t_rd, t_wr, t_adv;  // t_buf[t_wr] = t_buf[t_rd] (+dt if t_adv), when t_wr >= 0
  const int4* chunks;     // (instance, first particle, count, chunk index in instance)
  const int* nchunks;     // chunks per instance
  int xcd_off[8];
  int xcd_cnt[8];
  int affinity;           // 1: blockIdx -> (XCD = blockIdx % 8, local = blockIdx / 8)
  int total_chunks;
  int G;
  int lazy_vc;            // v, C of in-range particles are not stored between fused substeps
  int glow;
  PhysConsts k;
};

template <int BLOCK>
__device__ __forceinline__ void block_bbox(int& mn0, int& mx0, int& mn1, int& mx1,
                                           int (*red)[4]) {
  int v[4] = {mn0, -mx0, mn1, -mx1};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[c] = min(v[c], __shfl_xor(v[c], off, 64));
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) red[wave][c] = v[c];
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    int r = red[0][c];
#pragma unroll
    for (int w = 1; w < BLOCK / 64; ++w) r = min(r, red[w][c]);
    v[c] = r;
  }
  __syncthreads();
  mn0 = v[0];
  mx0 = -v[1];
  mn1 = v[2];
  mx1 = -v[3];
}

constexpr int kBoxEmpty = 1 << 29;

// P2G accumulates in LDS as 64-bit fixed point (fixed_bits, lanes.h). A particle's 27
// contributions are w_ij (m v + A d) and w_ij m with w_ij <= 0.5625 and |d| <= 1.5 dx, so
// bounding |m v_c| + 1.5 dx (|A_c0| + |A_c1|) and m by kFixMax bounds each contribution by
// 0.5625 kFixMax. The same-cell pre-sums add up to 4 lanes before one encode, so an encoded
// value stays below 2.25 kFixMax = 1.6875 < 2, inside fixed_bits' exact range (|c| < 2);
// a window gathers at most 4096 particles, so every node sum fits int64. A particle outside
// that bound, or with a NaN or infinite term (the comparisons below fail on NaN), scatters
// straight to the global fp32 grid instead, so non-finite values reach exactly its nine
// nodes, as in fp32.
constexpr float kFixMax = 0.75f;
// The window's planes are summed as biased fixed point (fixed_bits_biased, lanes.h): the
// mass plane recovers each node's emission count from its top bits, which needs its sums
// below 2^48 units of 2^-50, i.e. below 0.25. A window adds at most 4096 particles' terms
// of at most 0.5625 m each, so the fixed-point path takes particles of mass below 2^-14
// (water and jelly 1.5e-5, the payload at density 2.5 3.8e-5; up to density 4); heavier
// ones scatter to the global grid like out-of-range values.
constexpr float kFixMassMax = 6.103515625e-05f;  // 2^-14

__device__ __forceinline__ void add_fixed_scaled(unsigned long long* slot, float y) {
  atomicAdd(slot, fixed_bits_biased(y));
}

// Row of node q of a window ww nodes wide, for q < 2^21: (q + 0.5) / ww lies at least
// 0.5 / ww from an integer, and the product with the approximate reciprocal (relative
// error below 2^-22) is off by less than that.
__device__ __forceinline__ int window_row(int q, float rcp_ww) {
  return (int)(((float)q + 0.5f) * rcp_ww);
}

// Rare path: a particle whose contributions leave the fixed-point range scatters its nine
// nodes with fp32 global atomics. Node (i, j) receives g_j P0_i + e_j Q0_i,
// g_j P1_i + e_j Q1_i, g_j M_i.
struct Axis3 {
  float v0, v1, v2;
  __device__ __forceinline__ float operator[](int i) const { return i == 0 ? v0 : (i == 1 ? v1 : v2); }
};

__device__ __forceinline__ void scatter_global(float* gw, int G, int nx, int b0, int b1, Axis3 wy,
                                            Axis3 ey, Axis3 P0, Axis3 Q0, Axis3 P1, Axis3 Q1,
                                            Axis3 Mi) {
#pragma unroll
  for (int jj = 0; jj < 3; ++jj)
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
      const int node = (b1 + jj) * nx + (b0 + ii);
      unsafeAtomicAdd(gw + node, wy[jj] * P0[ii] + ey[jj] * Q0[ii]);
      unsafeAtomicAdd(gw + G + node, wy[jj] * P1[ii] + ey[jj] * Q1[ii]);
      unsafeAtomicAdd(gw + 2 * G + node, wy[jj] * Mi[ii]);
    }
}


// G2P of one in-range particle from a 3x3 block of grid velocities (row stride `ld`):
// v = sum w g, C = 4 inv_dx sum w g (x) (node - x) (mpm_sim.py:317-327), evaluated in the
// separable form sum_i wx_i [sum_j wy_j g_ij (x) (d0_i, d1_j)].
__device__ __forceinline__ void g2p_gather(const float2* g, int ld, float fx0, float fx1,
                                           const PhysConsts& k, float2& v, float4& C) {
  float wx[3], wy[3];
  bspline(fx0, wx);
  bspline(fx1, wy);
  float ey[3];
#pragma unroll
  for (int jj = 0; jj < 3; ++jj) ey[jj] = wy[jj] * ((float)jj - fx1);
  float v0 = 0.f, v1 = 0.f, c00 = 0.f, c01 = 0.f, c10 = 0.f, c11 = 0.f;
#pragma unroll
  for (int ii = 0; ii < 3; ++ii) {
    float s0 = 0.f, s1 = 0.f, t0 = 0.f, t1 = 0.f;
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
      const float2 n = g[jj * ld + ii];
      s0 += wy[jj] * n.x;
      s1 += wy[jj] * n.y;
      t0 += ey[jj] * n.x;
      t1 += ey[jj] * n.y;
    }
    const float d0 = (float)ii - fx0;
    v0 += wx[ii] * s0;
    v1 += wx[ii] * s1;
    c00 += (wx[ii] * d0) * s0;
    c10 += (wx[ii] * d0) * s1;
    c01 += wx[ii] * t0;
    c11 += wx[ii] * t1;
  }
  v = make_float2(v0, v1);
  C = make_float4(k.apic_scale * c00, k.apic_scale * c01, k.apic_scale * c10,
                  k.apic_scale * c11);
}

// G2P + advection of one particle straight from the global grid (grid op applied per
// node, the axisymmetric one when hp, the instance's two stress planes, is given); the
// slow path for chunks whose window does not fit in LDS. Reads and writes the particle's
// x, v, C in memory.
__device__ __forceinline__ void g2p_global(const float* gp, const float* hp, int G,
                                        const PhysConsts* kp, float2* xs, float2* vs,
                                        float4* Cs, int gi) {
  const PhysConsts& k = *kp;
  float2 x = xs[gi], v = vs[gi];
  float4 C = Cs[gi];
  int b0, b1;
  float fx0, fx1;
  base_frac(x.x, k.inv_dx, b0, fx0);
  base_frac(x.y, k.inv_dx, b1, fx1);
  if (in_range(b0, b1, k)) {
    float wx[3], wy[3];
    bspline(fx0, wx);
    bspline(fx1, wy);
    float v0 = 0.f, v1 = 0.f, c00 = 0.f, c01 = 0.f, c10 = 0.f, c11 = 0.f;
    for (int ii = 0; ii < 3; ++ii)
      for (int jj = 0; jj < 3; ++jj) {
        const int i = b0 + ii, j = b1 + jj, node = j * k.nx + i;
        const float2 n =
            hp ? grid_node_axisym(gp[node], gp[G + node], gp[2 * G + node], hp[node],
                                  hp[G + node], i, j, k)
               : grid_node(gp[node], gp[G + node], gp[2 * G + node], i, j, k);
        const float w = wx[ii] * wy[jj], d0 = (float)ii - fx0, d1 = (float)jj - fx1;
        v0 += w * n.x;
        v1 += w * n.y;
        c00 += w * n.x * d0;
        c01 += w * n.x * d1;
        c10 += w * n.y * d0;
        c11 += w * n.y * d1;
      }
    v = make_float2(v0, v1);
    C = make_float4(k.apic_scale * c00, k.apic_scale * c01, k.apic_scale * c10,
                    k.apic_scale * c11);
  }
  advect(x, v, k);
  xs[gi] = x;
  vs[gi] = v;
  Cs[gi] = C;
}

// Element i of an array, addressed as base + 32-bit byte offset: with a uniform base the
// load or store takes the scalar-base form (one 32-bit offset per access instead of 64-bit
// address arithmetic). Arrays hold fewer than 2^32 bytes (jm_create checks the slot count).
template <typename T>
__device__ __forceinline__ T& elem(T* base, uint32_t i) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T& elem(const T* base, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                     i * (uint32_t)sizeof(T));
}

// P2G base cell of a particle, packed as b0 | b1 << 16 (both in [0, 2^15) when in range),
// with its fractional offsets; -1 when the particle is dead or outside the transfer range.
// in_range is the test of mpm_sim.py:184,316.
__device__ __forceinline__ int packed_cell(float2 x, const PhysConsts& k, float& fx0,
                                           float& fx1) {
  int b0, b1;
  base_frac(x.x, k.inv_dx, b0, fx0);
  base_frac(x.y, k.inv_dx, b1, fx1);
  return in_range(b0, b1, k) ? (b0 | (b1 << 16)) : -1;
}

// Register budget: five waves per SIMD where the LDS window allows it (chunks of up to 1024
// particles in workgroups of at least 256 threads); at 32 x 80K, 256 x 4 runs 39.5 us per
// launch at five waves, 42.2 us at four and 53.2 us at six (which spills).
//
// AXI: the axisymmetric variant (mpm_sim.py:163-165,236-304 with axisym): P2G weights the
// particle-velocity momentum and the mass by the radius r = x (a particle is a ring) and
// scatters two more planes, the mass-weighted Kirchhoff s_rr and the hoop stress of the
// circumferential stretch r / r_ref, into the stress grids; the grid op adds the hoop
// correction and the axis condition (grid_node_axisym).
template <int BLOCK, int PPT, bool DO_G2P, bool DO_P2G, bool AXI = false>
__global__ void __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(!AXI && BLOCK >= 256 && BLOCK * PPT <= 1024 ? 5 : 1)))
k_fused(FusedArgs a) {
  // Grid-window capacity (nodes): a chunk of P sorted particles (4 per cell in water)
  // covers about P/4 cells (P/4/S columns of an S-row strip, or P/512 + 2 rows of 128 in
  // row-major order), plus the 2-node stencil halo and drift since the last sort.
  // 1280 nodes (30 KB of LDS) for 1024-particle chunks: five workgroups per CU, as many as
  // the kernel's registers allow, with twice the typical window (~400-650 nodes) as
  // headroom before a chunk falls back to the slow global path. The axisymmetric variant's
  // five planes keep the same 30 KB with 768 nodes.
  constexpr int NPL = AXI ? 5 : 3;
  constexpr int WMAX = AXI ? ((BLOCK * PPT <= 1024) ? 768 : 1024)
                           : (BLOCK * PPT <= 1024) ? 1280 : (BLOCK * PPT <= 2048) ? 1536 : 2560;
  // One LDS buffer: the gather window as float2, then the scatter window as NPL planes of
  // 64-bit fixed-point accumulators.
  __shared__ unsigned long long lds64[NPL * WMAX];
  float* lds = reinterpret_cast<float*>(lds64);
  __shared__ int red[BLOCK / 64][4];
  const int tid = threadIdx.x;
  const PhysConsts& k = a.k;

  if (a.t_wr >= 0 && blockIdx.x == 0 && tid == 0) {
    const float t = a.t_buf[a.t_rd];
    a.t_buf[a.t_wr] = a.t_adv ? t + k.dt : t;  // mpm_sim.py:340
  }

  int ci;
  if (a.affinity) {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    if (local >= a.xcd_cnt[xcd]) return;
    ci = a.xcd_off[xcd] + local;
  } else {
    ci = blockIdx.x;
    if (ci >= a.total_chunks) return;
  }
  const int4 ch = a.chunks[ci];
  const int m = ch.x, start = ch.y, count = ch.z, cidx = ch.w;
  const int nx = k.nx, G = a.G;

  // Zero this chunk's share of the grid that the substep after next scatters into.
  if (a.acc_z) {
    const int nch = a.nchunks[m];
    const int64_t n4 = (int64_t)3 * G / 4;
    const int64_t lo = n4 * cidx / nch, hi = n4 * (cidx + 1) / nch;
    float4* z = reinterpret_cast<float4*>(a.acc_z + (size_t)m * 3 * G);
    for (int64_t q = lo + tid; q < hi; q += BLOCK) z[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if constexpr (AXI) {
    if (a.hoop_z) {
      const int nch = a.nchunks[m];
      const int64_t n4 = (int64_t)2 * G / 4;
      const int64_t lo = n4 * cidx / nch, hi = n4 * (cidx + 1) / nch;
      float4* z = reinterpret_cast<float4*>(a.hoop_z + (size_t)m * 2 * G);
      for (int64_t q = lo + tid; q < hi; q += BLOCK) z[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }

  // This chunk's slices of the particle arrays (uniform bases; lanes add 32-bit offsets).
  float2* const xs = a.s.x + start;
  float2* const vs = a.s.v + start;
  float4* const Cs = a.s.C + start;
  float4* const Fs = a.s.F + start;
  float* const Jps = a.s.Jp + start;
  const int* const mats = a.s.mat + start;
  const float2* const fibs = a.s.fiber + start;

  float2 x[PPT], v[PPT];
  float4 C[PPT];
  int mt[PPT];
  // Base cell (packed_cell) and fractional offsets of each particle, computed once per phase:
  // at the gathered positions for G2P, then at the advected positions for P2G.
  int cell[PPT];
  float f0[PPT], f1[PPT];
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    const uint32_t li = p * BLOCK + tid;
    mt[p] = -1;
    if ((int)li < count) {
      x[p] = elem(xs, li);
      mt[p] = elem(mats, li);
    }
  }

  if constexpr (DO_G2P) {
    // Window of grid nodes touched by this chunk's in-range particles.
    int mn0 = kBoxEmpty, mx0 = -kBoxEmpty, mn1 = kBoxEmpty, mx1 = -kBoxEmpty;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      cell[p] = -1;
      f0[p] = f1[p] = 0.5f;
      if (mt[p] < 0) continue;
      cell[p] = packed_cell(x[p], k, f0[p], f1[p]);
      if (cell[p] < 0) continue;
      const int b0 = cell[p] & 0xffff, b1 = cell[p] >> 16;
      mn0 = min(mn0, b0);
      mx0 = max(mx0, b0);
      mn1 = min(mn1, b1);
      mx1 = max(mx1, b1);
    }
    block_bbox<BLOCK>(mn0, mx0, mn1, mx1, red);
    const int ww = mx0 - mn0 + 3, wh = mx1 - mn1 + 3;
    const bool any = mn0 <= mx0;
    const bool use_lds = any && ww * wh <= WMAX;
    const float* gp = a.acc_r + (size_t)m * 3 * G;
    const float* hp = AXI ? a.hoop_r + (size_t)m * 2 * G : nullptr;
    float2* lv = reinterpret_cast<float2*>(lds);
    if (use_lds) {
      const int nodes = ww * wh;
      const float rcp_ww = __builtin_amdgcn_rcpf((float)ww);
      for (int q = tid; q < nodes; q += BLOCK) {
        const int jj = window_row(q, rcp_ww), ii = q - jj * ww;
        const int i = mn0 + ii, j = mn1 + jj;
        const int node = j * nx + i;
        if constexpr (AXI)
          lv[q] = grid_node_axisym(gp[node], gp[G + node], gp[2 * G + node], hp[node],
                                   hp[G + node], i, j, k);
        else
          lv[q] = grid_node(gp[node], gp[G + node], gp[2 * G + node], i, j, k);
      }
    }
    __syncthreads();
    if (use_lds) {
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        if (mt[p] < 0) continue;
        const uint32_t li = p * BLOCK + tid;
        if (cell[p] >= 0) {
          const int b0 = cell[p] & 0xffff, b1 = cell[p] >> 16;
          g2p_gather(lv + __mul24(b1 - mn1, ww) + (b0 - mn0), ww, f0[p], f1[p], k, v[p], C[p]);
        } else {
          v[p] = elem(vs, li);
          C[p] = elem(Cs, li);
        }
        advect(x[p], v[p], k);
      }
    } else {
      // Window too large for LDS (particles far from their sorted cells): gather from the
      // global grid, one particle at a time, through memory (not unrolled: rare path).
#pragma unroll 1
      for (int p = 0; p < PPT; ++p) {
        const int li = p * BLOCK + tid;
        if (li >= count) continue;
        g2p_global(gp, hp, G, &a.k, a.s.x, a.s.v, a.s.C, start + li);
      }
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        const uint32_t li = p * BLOCK + tid;
        if ((int)li < count) {
          x[p] = elem(xs, li);
          v[p] = elem(vs, li);
          C[p] = elem(Cs, li);
        }
      }
    }
    if (a.glow) {
#pragma unroll 1
      for (int p = 0; p < PPT; ++p) {
        const int li = p * BLOCK + tid;
        if (li >= count) continue;
        const int gi = start + li;
        a.s.glow[gi] = fmaxf(a.s.glow[gi] * 0.9f, fabsf(1.0f - a.s.Jp[gi]));
      }
    }
  } else {
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      if (mt[p] < 0) continue;
      const uint32_t li = p * BLOCK + tid;
      v[p] = elem(vs, li);
      C[p] = elem(Cs, li);
    }
  }

  if constexpr (DO_P2G) {
    const int mi = m;
    // The activation is uniform over the workgroup: the first wave evaluates it and the
    // others read it after block_bbox's barrier.
    __shared__ float s_act;
    if (tid < 64) {
      const float act = activation(a.t_buf[a.t_rd], a.ip.freq[mi], a.ip.contraction[mi], k);
      if (tid == 0) s_act = act;
    }
    const float mu_j = a.ip.mu_j[mi], la_j = a.ip.la_j[mi], pdens = a.ip.pdens[mi];

    int mn0 = kBoxEmpty, mx0 = -kBoxEmpty, mn1 = kBoxEmpty, mx1 = -kBoxEmpty;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      cell[p] = -1;
      f0[p] = f1[p] = 0.5f;
      if (mt[p] < 0) continue;
      cell[p] = packed_cell(x[p], k, f0[p], f1[p]);
      if (cell[p] < 0) continue;
      const int b0 = cell[p] & 0xffff, b1 = cell[p] >> 16;
      mn0 = min(mn0, b0);
      mx0 = max(mx0, b0);
      mn1 = min(mn1, b1);
      mx1 = max(mx1, b1);
    }
    block_bbox<BLOCK>(mn0, mx0, mn1, mx1, red);
    const float amp_act = a.ip.amp[mi] * s_act;
    const int ww = mx0 - mn0 + 3, wh = mx1 - mn1 + 3;
    const bool any = mn0 <= mx0;
    const bool use_lds = any && ww * wh <= WMAX;
    const int nodes = use_lds ? ww * wh : 0;
    const float rcp_ww = __builtin_amdgcn_rcpf((float)ww);
    for (int q = tid; q < nodes; q += BLOCK) {
#pragma unroll
      for (int pl = 0; pl < NPL; ++pl) lds64[pl * WMAX + q] = 0ull;
    }
    __syncthreads();
    float* gw = a.acc_w + (size_t)m * 3 * G;
    float* hw = AXI ? a.hoop_w + (size_t)m * 2 * G : nullptr;
    const float* const rrefs = AXI ? a.s.r_ref + start : nullptr;
    const int lane = tid & 63;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const uint32_t li = p * BLOCK + tid;
      const bool inr = cell[p] >= 0;
      const int b0 = cell[p] & 0xffff, b1 = cell[p] >> 16;
      const float fx0 = f0[p], fx1 = f1[p];
      if (mt[p] >= 0 && !inr && DO_G2P && a.lazy_vc) {
        elem(vs, li) = v[p];
        elem(Cs, li) = C[p];
      }
      // Particle update for in-range particles; every other lane scatters zeros, computed
      // from finite stand-ins (mass 0, affine 0, v 0, offsets 0.5) rather than by clearing
      // the 21 per-particle factors.
      float mass = 0.f;
      Mat2 aff = {0.f, 0.f, 0.f, 0.f};
      // axisymmetric: mass * r (the ring's weight of momentum and mass), mass * s_rr and
      // mass * s_tt (mpm_sim.py:236-247,258-267)
      float mr = 0.f, ms_rr = 0.f, ms_tt = 0.f;
      if (inr) {
        float2 fib = make_float2(0.f, 1.f);
        if (mt[p] == 3) fib = elem(fibs, li);
        const Mat2 Cm = as_mat(C[p]);
        P2GOut o = p2g_update(mt[p], as_mat(elem(Fs, li)), Cm, fib.x, fib.y, amp_act, mu_j,
                              la_j, pdens, k);
        elem(Fs, li) = as_f4(o.F);
        if (o.jp_changes) elem(Jps, li) = (elem(Jps, li) * o.q0) * o.q1;
        mass = o.mass;
        aff = o.affine;
        if constexpr (AXI) {
          const float rp = x[p].x;
          const float rr = fmaxf(elem(rrefs, li), k.r_ref_floor);
          const float lam = fmaxf(rp, k.r_floor) / rr;
          const float khoop = ((2.0f * o.mu) * (lam - 1.0f)) + ((o.la * o.J) * (o.J - 1.0f));
          mr = o.mass * rp;
          ms_rr = o.mass * o.s00;
          ms_tt = o.mass * khoop;
        }
      }
      // the particle-velocity momentum's weight: m, or m r for a ring
      const float mw = AXI ? mr : mass;
      const float mv0 = mw * (inr ? v[p].x : 0.f), mv1 = mw * (inr ? v[p].y : 0.f);
      float wx[3], wy[3], ey[3], P0[3], Q0[3], P1[3], Q1[3], Mi[3], Si[3], Ti[3];
      bspline(fx0, wx);
      bspline(fx1, wy);
      const float bnd = 1.5f * k.dx;
      // No fmaxf here: it drops NaN operands. Each comparison fails on NaN, and the
      // fractional offsets are tested too (a NaN position truncates to cell 0 with NaN
      // weights). The axisymmetric mass plane holds m r < m (r < 1); its stress planes'
      // terms are bounded like the momentum's.
      const bool fixed_ok = inr && use_lds &&
                            fabsf(mv0) + bnd * (fabsf(aff.a00) + fabsf(aff.a01)) < kFixMax &&
                            fabsf(mv1) + bnd * (fabsf(aff.a10) + fabsf(aff.a11)) < kFixMax &&
                            mass < kFixMassMax && fx0 + fx1 < 4.0f &&
                            (!AXI || (fabsf(ms_rr) < kFixMax && fabsf(ms_tt) < kFixMax));
      const float sc = fixed_ok ? kFixPre : (inr ? 1.0f : 0.0f);
#pragma unroll
      for (int ii = 0; ii < 3; ++ii) {
        const float ai = wx[ii] * sc;
        const float bi = ai * (((float)ii - fx0) * k.dx);
        P0[ii] = ai * mv0 + bi * aff.a00;
        Q0[ii] = ai * aff.a01;
        P1[ii] = ai * mv1 + bi * aff.a10;
        Q1[ii] = ai * aff.a11;
        Mi[ii] = ai * mw;
        if constexpr (AXI) {
          Si[ii] = ai * ms_rr;
          Ti[ii] = ai * ms_tt;
        }
      }
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) ey[jj] = wy[jj] * (((float)jj - fx1) * k.dx);
      if (inr && !fixed_ok) {
        scatter_global(gw, G, nx, b0, b1, {wy[0], wy[1], wy[2]}, {ey[0], ey[1], ey[2]},
                       {P0[0], P0[1], P0[2]}, {Q0[0], Q0[1], Q0[2]}, {P1[0], P1[1], P1[2]},
                       {Q1[0], Q1[1], Q1[2]}, {Mi[0], Mi[1], Mi[2]});
        if constexpr (AXI) {
#pragma unroll
          for (int jj = 0; jj < 3; ++jj)
#pragma unroll
            for (int ii = 0; ii < 3; ++ii) {
              const int node = (b1 + jj) * nx + (b0 + ii);
              unsafeAtomicAdd(hw + node, wy[jj] * Si[ii]);
              unsafeAtomicAdd(hw + G + node, wy[jj] * Ti[ii]);
            }
        }
        // done: its neighbour lane adds this lane's (now zero, finite) terms times 0
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) {
          P0[ii] = Q0[ii] = P1[ii] = Q1[ii] = Mi[ii] = 0.f;
          if constexpr (AXI) Si[ii] = Ti[ii] = 0.f;
          wy[ii] = ey[ii] = 0.f;
        }
      }
      // Cell-sorted particles of one base cell sit in adjacent lanes (water: 4 per cell).
      // Same-cell runs add their 27 contributions across lanes in groups of 4
      // (lane_groups, presum3), so only the first lane of each group issues the 27 64-bit
      // LDS atomics: a quarter of the atomic traffic of one set per particle, no two lanes
      // of a water cell hitting one address, in a fixed order (deterministic). A
      // neighbour's term enters as t * f with f = 1 or 0 (one FMA, exact either way: every
      // lane's terms are finite, since out-of-range lanes carry zeros). Lanes outside the
      // fixed-point LDS path carry unique keys and add nothing. Measured at 32 x 80K:
      // 44.3 us per substep against 45.7 us with pairs only.
      const int key = fixed_ok ? (cell[p] + 1) : -(1 + lane);
      const LaneGroups lg = lane_groups(key, fixed_ok);
      const int q0 = __mul24(b1 - mn1, ww) + (b0 - mn0);
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) {
          float c[3] = {wy[jj] * P0[ii] + ey[jj] * Q0[ii], wy[jj] * P1[ii] + ey[jj] * Q1[ii],
                        wy[jj] * Mi[ii]};
          presum3(c[0], c[1], c[2], lg.f1, lg.f2);
          float s[2] = {0.f, 0.f};
          if constexpr (AXI) {
            s[0] = wy[jj] * Si[ii];
            s[1] = wy[jj] * Ti[ii];
            presum2(s[0], s[1], lg.f1, lg.f2);
          }
          if (lg.emit) {
            const int q = q0 + jj * ww + ii;
            add_fixed_scaled(&lds64[q], c[0]);
            add_fixed_scaled(&lds64[WMAX + q], c[1]);
            add_fixed_scaled(&lds64[2 * WMAX + q], c[2]);
            if constexpr (AXI) {
              add_fixed_scaled(&lds64[3 * WMAX + q], s[0]);
              add_fixed_scaled(&lds64[4 * WMAX + q], s[1]);
            }
          }
        }
      }
    }
    __syncthreads();
    for (int q = tid; q < nodes; q += BLOCK) {
      const int jj = window_row(q, rcp_ww), ii = q - jj * ww;
      const int node = (mn1 + jj) * nx + (mn0 + ii);
      const unsigned long long vm = lds64[2 * WMAX + q];
      const unsigned n = biased_count(vm);  // emissions into this node (all planes)
      const float c0 = from_biased_bits(lds64[q], n), c1 = from_biased_bits(lds64[WMAX + q], n),
                  cm = from_biased_bits(vm, n);
      if (c0 != 0.f) unsafeAtomicAdd(gw + node, c0);
      if (c1 != 0.f) unsafeAtomicAdd(gw + G + node, c1);
      if (cm != 0.f) unsafeAtomicAdd(gw + 2 * G + node, cm);
      if constexpr (AXI) {
        const float s0 = from_biased_bits(lds64[3 * WMAX + q], n),
                    s1 = from_biased_bits(lds64[4 * WMAX + q], n);
        if (s0 != 0.f) unsafeAtomicAdd(hw + node, s0);
        if (s1 != 0.f) unsafeAtomicAdd(hw + G + node, s1);
      }
    }
  }

  if constexpr (DO_G2P) {
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      if (mt[p] < 0) continue;
      const uint32_t li = p * BLOCK + tid;
      elem(xs, li) = x[p];
      if (!DO_P2G || !a.lazy_vc) {
        elem(vs, li) = v[p];
        elem(Cs, li) = C[p];
      }
    }
  }
}

// Grid velocity of one buffer as the G2P that read it saw it: (M, 2, G) planes.
__global__ void k_grid_velocity(const float* acc, float* out, int M, int G, PhysConsts k) {
  const int m = blockIdx.y;
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M || node >= G) return;
  const float* gp = acc + (size_t)m * 3 * G;
  const int j = node / k.nx, i = node - j * k.nx;
  float2 g = grid_node(gp[node], gp[G + node], gp[2 * G + node], i, j, k);
  out[(size_t)m * 2 * G + node] = g.x;
  out[(size_t)m * 2 * G + G + node] = g.y;
}

// Axisymmetric fused path: the grid buffer the last G2P gathered from becomes velocities in
// place (planes 0-1 through grid_node_axisym with the buffer's stress planes; the mass
// plane stays), as the plain path's grid op leaves its grid, so that grid reads and the
// renderer use it without the axisymmetric grid op. No later substep reads that buffer
// (the next block's first kernel zeroes it).
__global__ void k_axisym_velocity(float* acc, const float* hoop, int M, int G, PhysConsts k) {
  const int m = blockIdx.y;
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M || node >= G) return;
  float* gp = acc + (size_t)m * 3 * G;
  const float* hp = hoop + (size_t)m * 2 * G;
  const int j = node / k.nx, i = node - j * k.nx;
  const float2 g =
      grid_node_axisym(gp[node], gp[G + node], gp[2 * G + node], hp[node], hp[G + node], i, j, k);
  gp[node] = g.x;
  gp[G + node] = g.y;
}

// ---------------------------------------------------------------------------------------
// Stable sort of each instance's particles by P2G base cell: an LSD radix sort with 8-bit
// digits over per-instance tiles of 1024 particles. Each pass ranks a tile's elements per
// digit in their current order (wave ballots + a per-wave prefix), scans the (digit, tile)
// histogram of the instance, and scatters; no atomics, so the permutation is deterministic
// and equals a stable counting sort by cell.

constexpr int kSortTile = 1024;

struct SortArgs {
  State s;          // live state (sorted in place by the gather)
  State tmp;        // copy of the state before the sort
  int* key_a;       // keys / values, ping-pong between a and b
  int* val_a;
  int* key_b;
  int* val_b;
  int* rank;        // per element: stable rank within (tile, digit)
  int* hist;        // [m][256][tmax]: digit counts per tile, then exclusive offsets
  const int* n_live;
  int cap;
  int M;
  int tmax;         // tiles per instance segment (cap / 1024, rounded up)
  int strip;        // cell order: strips of `strip` rows (0: row-major), see cell_key
  PhysConsts k;
};

// Sort key of a particle's P2G base cell (clamped into the grid). strip = 0: row-major
// cell index. strip = S > 0: the grid is cut into strips of S cell rows; strips follow
// each other upwards, the columns of even strips run left to right and those of odd
// strips right to left, and the S cells of a column run upwards. A chunk of consecutive
// sorted particles then covers a compact block of about S rows (its LDS window is about
// half the nodes of a row-major chunk's, which spans full rows), and the chunk that turns
// at the end of a strip continues right above. Every order keeps the particles of one
// cell adjacent.
__device__ __forceinline__ int cell_key(float2 x, const PhysConsts& k, int strip) {
  int b0, b1;
  float f0, f1;
  base_frac(x.x, k.inv_dx, b0, f0);
  base_frac(x.y, k.inv_dx, b1, f1);
  b0 = min(max(b0, 0), k.nx - 1);
  b1 = min(max(b1, 0), k.ny - 1);
  if (strip <= 0) return b1 * k.nx + b0;
  const int s = b1 / strip, r = b1 - s * strip;
  const int xs = (s & 1) ? (k.nx - 1 - b0) : b0;
  return (s * k.nx + xs) * strip + r;
}

// Keys of the current order, identity values, and a copy of the state to gather from.
__global__ void __launch_bounds__(256) k_sort_keys(SortArgs a) {
  const int m = blockIdx.y;
  const int local = blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= a.n_live[m]) return;
  const size_t i = (size_t)m * a.cap + local;
  const float2 x = a.s.x[i];
  a.key_a[i] = cell_key(x, a.k, a.strip);
  a.val_a[i] = local;
  a.tmp.x[i] = x;
  a.tmp.v[i] = a.s.v[i];
  a.tmp.C[i] = a.s.C[i];
  a.tmp.F[i] = a.s.F[i];
  a.tmp.Jp[i] = a.s.Jp[i];
  a.tmp.mat[i] = a.s.mat[i];
  a.tmp.fiber[i] = a.s.fiber[i];
  a.tmp.slot[i] = a.s.slot[i];
  if (a.s.glow) a.tmp.glow[i] = a.s.glow[i];
  if (a.s.r_ref) a.tmp.r_ref[i] = a.s.r_ref[i];
}

__global__ void __launch_bounds__(kSortTile) k_sort_rank(SortArgs a, const int* keys, int shift) {
  __shared__ int wcnt[kSortTile / 64][256];
  const int m = blockIdx.y, t = blockIdx.x, tid = threadIdx.x;
  const int n = a.n_live[m];
  if (t * kSortTile >= n) return;
  for (int q = tid; q < (kSortTile / 64) * 256; q += kSortTile) (&wcnt[0][0])[q] = 0;
  __syncthreads();
  const int local = t * kSortTile + tid;
  const bool valid = local < n;
  const size_t i = (size_t)m * a.cap + local;
  const int d = valid ? (keys[i] >> shift) & 255 : 0;
  unsigned long long same = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1;
    const unsigned long long bb = __ballot(valid && bit);
    same &= bit ? bb : ~bb;
  }
  const int lane = tid & 63, wave = tid >> 6;
  const int rw = __popcll(same & ((1ull << lane) - 1ull));
  if (valid && rw == 0) wcnt[wave][d] = __popcll(same);
  __syncthreads();
  if (tid < 256) {
    int run = 0;
    for (int w = 0; w < kSortTile / 64; ++w) {
      const int c = wcnt[w][tid];
      wcnt[w][tid] = run;
      run += c;
    }
    a.hist[((size_t)m * 256 + tid) * a.tmax + t] = run;
  }
  __syncthreads();
  if (valid) a.rank[i] = wcnt[wave][d] + rw;
}

// Exclusive scan of one instance's (digit, tile) counts in digit-major order.
__global__ void __launch_bounds__(1024) k_sort_scan(SortArgs a) {
  __shared__ int part[1024];
  const int m = blockIdx.x, tid = threadIdx.x;
  const int T = (a.n_live[m] + kSortTile - 1) / kSortTile;
  const int E = 256 * T;
  const int per = (E + 1023) / 1024;
  int* h = a.hist + (size_t)m * 256 * a.tmax;
  int sum = 0;
  for (int q = 0; q < per; ++q) {
    const int f = tid * per + q;
    if (f < E) sum += h[(f / T) * a.tmax + (f % T)];
  }
  part[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int add = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += add;
    __syncthreads();
  }
  int run = part[tid] - sum;
  for (int q = 0; q < per; ++q) {
    const int f = tid * per + q;
    if (f < E) {
      int* e = &h[(f / T) * a.tmax + (f % T)];
      const int c = *e;
      *e = run;
      run += c;
    }
  }
}

__global__ void __launch_bounds__(256) k_sort_scatter(SortArgs a, const int* kin, const int* vin,
                                                       int* kout, int* vout, int shift) {
  const int m = blockIdx.y;
  const int local = blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= a.n_live[m]) return;
  const size_t i = (size_t)m * a.cap + local;
  const int key = kin[i];
  const int d = (key >> shift) & 255;
  const int dst = a.hist[((size_t)m * 256 + d) * a.tmax + local / kSortTile] + a.rank[i];
  kout[(size_t)m * a.cap + dst] = key;
  vout[(size_t)m * a.cap + dst] = vin[i];
}

// Move every field to its sorted position: s[i] = tmp[perm[i]].
__global__ void __launch_bounds__(256) k_sort_gather(SortArgs a, const int* perm) {
  const int m = blockIdx.y;
  const int local = blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= a.n_live[m]) return;
  const size_t i = (size_t)m * a.cap + local;
  const size_t src = (size_t)m * a.cap + perm[i];
  a.s.x[i] = a.tmp.x[src];
  a.s.v[i] = a.tmp.v[src];
  a.s.C[i] = a.tmp.C[src];
  a.s.F[i] = a.tmp.F[src];
  a.s.Jp[i] = a.tmp.Jp[src];
  a.s.mat[i] = a.tmp.mat[src];
  a.s.fiber[i] = a.tmp.fiber[src];
  a.s.slot[i] = a.tmp.slot[src];
  if (a.s.glow) a.s.glow[i] = a.tmp.glow[src];
  if (a.s.r_ref) a.s.r_ref[i] = a.tmp.r_ref[src];
}

// ---------------------------------------------------------------------------------------
// Live particles of each instance whose v, C or F holds a NaN or an infinity (a diverged
// state; its positions stay finite because advection clamps them into the domain,
// mpm_sim.py:331-333). out[m] must start at 0. blockIdx.y = instance.
__global__ void __launch_bounds__(256) k_count_nonfinite(State s, const int* n_live, int cap,
                                                         int* out) {
  const int m = blockIdx.y;
  const int local = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (local < n_live[m]) {
    const size_t i = (size_t)m * cap + local;
    const float2 v = s.v[i];
    const float4 C = s.C[i], F = s.F[i];
    // x * 0 is 0 for every finite x and NaN for NaN and +-inf
    const float z = ((v.x * 0.f + v.y * 0.f) + (C.x * 0.f + C.y * 0.f) + (C.z * 0.f + C.w * 0.f)) +
                    ((F.x * 0.f + F.y * 0.f) + (F.z * 0.f + F.w * 0.f));
    bad = !(z == 0.f);
  }
  const unsigned long long b = __ballot(bad);
  if ((threadIdx.x & 63) == 0 && b != 0ull) atomicAdd(&out[m], __popcll(b));
}

// ---------------------------------------------------------------------------------------
// fitness_buffer sums (mpm_sim.py:746-762) for payload and body at once, accumulated in
// double in a fixed order: out[m][0..5] = payload (sum_y, count, sum_x), body (same).

__global__ void __launch_bounds__(256) k_stats(State s, const int* n_live, int cap, float* out) {
  __shared__ double red[4][6];
  const int m = blockIdx.x, tid = threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int n = n_live[m];
  for (int local = tid; local < n; local += 256) {
    const size_t i = (size_t)m * cap + local;
    const int mt = s.mat[i];
    const float2 x = s.x[i];
    if (mt == 2) {
      acc[0] += x.y;
      acc[1] += 1.0;
      acc[2] += x.x;
    } else if (mt == 1 || mt == 3) {
      acc[3] += x.y;
      acc[4] += 1.0;
      acc[5] += x.x;
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
  }
  if ((tid & 63) == 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) red[tid >> 6][c] = acc[c];
  }
  __syncthreads();
  if (tid < 6) {
    const double r = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
    out[(size_t)m * 6 + tid] = (float)r;
  }
}

}  // namespace jm
