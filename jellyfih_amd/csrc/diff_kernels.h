// HIP kernels of the differentiable path: a batched substep and its adjoint over explicit
// device buffers (torch tensors), particles in any order with an instance index each.
// Math in adjoint.h; grids accumulate with fp32 global atomics.
//
// Per instance the grid block holds 5 planes of nx*ny floats: momentum x, momentum y, mass,
// velocity x, velocity y (after the grid op). The adjoint grid block mirrors it: momentum
// adjoints, mass adjoint, velocity adjoints.
#pragma once

#include "adjoint.h"

namespace jmd {

struct DiffInputs {
  const int* inst;          // (P,) instance of each particle
  const int* mat;           // (P,) material (< 0: dead)
  const float* fiber;       // (P, 2)
  const float* amp_scale;   // (P,) per-particle actuation multiplier
  const float* params;      // (M, 5): amp, contraction, freq, payload density, E_jelly
};

struct AtomicAddF {
  __device__ __forceinline__ void operator()(float* p, float v) const { unsafeAtomicAdd(p, v); }
};

__device__ __forceinline__ void inst_params(const DiffInputs& in, int m, float t,
                                            const Consts<float>& k, Act<float>& act, float& mu,
                                            float& la, float& amp, float& pdens) {
  const float* pr = in.params + 5 * m;
  act = activation(t, pr[2], pr[1], k);
  const float E = pr[4], nu = k.nu_jelly;
  mu = E / (2.0f * (1.0f + nu));
  la = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
  amp = pr[0];
  pdens = pr[3];
}

__global__ void k_d_fill(float* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// P2G of substep (time t): scatter from `cur`, write the updated F (and copy x, v, C) to `nxt`.
__global__ void __launch_bounds__(256) k_d_p2g(DiffInputs in, const float* cur, float* nxt,
                                               float* grid, int P, int G, float t,
                                               Consts<float> k) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  PState<float> s = load_state(cur + 12 * (size_t)p);
  const int mt = in.mat[p];
  if (mt >= 0) {
    const int m = in.inst[p];
    Act<float> act;
    float mu, la, amp, pdens;
    inst_params(in, m, t, k, act, mu, la, amp, pdens);
    float* g = grid + (size_t)m * 5 * G;
    p2g_scatter(s, mt, in.fiber[2 * p], in.fiber[2 * p + 1], act.a * amp * in.amp_scale[p], mu,
                la, pdens, k, g, g + G, g + 2 * G, AtomicAddF());
  }
  if (nxt) store_state(nxt + 12 * (size_t)p, s);
}

__global__ void __launch_bounds__(256) k_d_gridop(float* grid, int M, int G, Consts<float> k) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (node >= G || m >= M) return;
  float* g = grid + (size_t)m * 5 * G;
  const int j = node / k.nx, i = node - j * k.nx;
  gridop_forward(g[node], g[G + node], g[2 * G + node], i, j, k, g[3 * G + node],
                 g[4 * G + node]);
}

// G2P of the substep: reads x, v, C from `cur`, F from `nxt` (already updated), writes `nxt`.
__global__ void __launch_bounds__(256) k_d_g2p(DiffInputs in, const float* cur, float* nxt,
                                               const float* grid, int P, int G,
                                               Consts<float> k) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  PState<float> s = load_state(cur + 12 * (size_t)p);
  const float* fn = nxt + 12 * (size_t)p + 8;
  s.F = {fn[0], fn[1], fn[2], fn[3]};
  if (in.mat[p] >= 0) {
    const float* g = grid + (size_t)in.inst[p] * 5 * G;
    g2p_particle(s, k, g + 3 * G, g + 4 * G);
  }
  store_state(nxt + 12 * (size_t)p, s);
}

// Adjoint of G2P: from bout (adjoint of the state after the substep) to the partial adjoint
// of the state before (x, and v, C of particles outside the P2G range) in bmid, scattering
// the grid-velocity adjoint.
__global__ void __launch_bounds__(256) k_d_g2p_bwd(DiffInputs in, const float* cur,
                                                   const float* bout, float* bmid,
                                                   const float* grid, float* agrid, int P, int G,
                                                   Consts<float> k) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  PState<float> bo = load_state(bout + 12 * (size_t)p);
  PState<float> bi = {0.f, 0.f, 0.f, 0.f, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
  if (in.mat[p] < 0) {
    bi = bo;  // dead slots are never updated
  } else {
    PState<float> s = load_state(cur + 12 * (size_t)p);
    const size_t off = (size_t)in.inst[p] * 5 * G;
    const float* g = grid + off;
    float* a = agrid + off;
    g2p_backward(s, bo, k, g + 3 * G, g + 4 * G, a + 3 * G, a + 4 * G, bi, AtomicAddF());
  }
  store_state(bmid + 12 * (size_t)p, bi);
}

__global__ void __launch_bounds__(256) k_d_gridop_bwd(const float* grid, float* agrid, int M,
                                                      int G, Consts<float> k) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (node >= G || m >= M) return;
  const float* g = grid + (size_t)m * 5 * G;
  float* a = agrid + (size_t)m * 5 * G;
  const int j = node / k.nx, i = node - j * k.nx;
  gridop_backward(g[node], g[G + node], g[2 * G + node], i, j, k, a[3 * G + node],
                  a[4 * G + node], a[node], a[G + node], a[2 * G + node]);
}

// Adjoint of P2G (gather) and of the constitutive update; completes the adjoint of the
// state before the substep in `bin` and accumulates parameter adjoints per particle:
// gamp[p] (actuation multiplier), gpp[p] = (E_jelly, amp, contraction, freq).
__global__ void __launch_bounds__(256) k_d_p2g_bwd(DiffInputs in, const float* cur,
                                                   const float* bout, const float* bmid,
                                                   float* bin, const float* agrid, int P, int G,
                                                   float t, Consts<float> k, float* gamp,
                                                   float* gpp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  PState<float> bi = load_state(bmid + 12 * (size_t)p);
  const int mt = in.mat[p];
  if (mt >= 0) {
    const int m = in.inst[p];
    PState<float> s = load_state(cur + 12 * (size_t)p);
    const float* fb = bout + 12 * (size_t)p + 8;
    const Mat<float> Fnb = {fb[0], fb[1], fb[2], fb[3]};
    Act<float> act;
    float mu, la, amp, pdens;
    inst_params(in, m, t, k, act, mu, la, amp, pdens);
    const float* a = agrid + (size_t)m * 5 * G;
    float mub = 0.f, lab = 0.f, betab = 0.f;
    const float as = in.amp_scale[p];
    p2g_backward_particle(s, mt, in.fiber[2 * p], in.fiber[2 * p + 1], act.a * amp * as, mu, la,
                          pdens, k, a, a + G, a + 2 * G, Fnb, bi, mub, lab, betab);
    if (mt == 1 || mt == 3) {
      const float nu = k.nu_jelly;
      gpp[4 * (size_t)p] += mub / (2.f * (1.f + nu)) + lab * nu / ((1.f + nu) * (1.f - 2.f * nu));
    }
    if (mt == 3) {
      gamp[p] += betab * act.a * amp;
      gpp[4 * (size_t)p + 1] += betab * act.a * as;
      gpp[4 * (size_t)p + 2] += betab * amp * as * act.da_dc;
      gpp[4 * (size_t)p + 3] += betab * amp * as * act.da_df;
    }
  }
  store_state(bin + 12 * (size_t)p, bi);
}

}  // namespace jmd
