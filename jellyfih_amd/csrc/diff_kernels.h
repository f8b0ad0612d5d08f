// HIP kernels of the differentiable path: a batched substep and its adjoint over explicit
// device buffers (torch tensors), particles in any order with an instance index each.
// Math in adjoint.h.
//
// Per instance a grid holds 3 planes of nx*ny floats: momentum x, momentum y, mass. The
// grid op is node-local, so G2P evaluates it at each stencil node, and the G2P adjoint
// scatters straight into the momentum/mass adjoints: a substep is two kernels (P2G, G2P),
// its adjoint two more (G2P adjoint scatter, P2G adjoint gather).
//
// The two scatters (P2G, G2P adjoint) accumulate in an LDS tile covering the workgroup's
// stencil nodes when the workgroup's particles are of one instance and close together
// (particles packed instance by instance in spatial order, as tanks are filled), then add
// the tile to the grid with one global atomic per touched node and plane; otherwise they
// use global atomics directly. The tile is fp64 for both scalar types: gfx950's fp64 LDS
// atomic add is much faster than its fp32 one, and the sums only gain precision. Every grid a kernel accumulates into starts zeroed: the
// kernel before it clears it (grid-stride), or a fill kernel at the start of a block.
#pragma once

#include <climits>

#include "adjoint.h"

namespace jmd {

constexpr int kDiffBlock = 256;   // threads per workgroup, one particle each
constexpr int kTileNodes = 2048;  // LDS tile: 3 planes x 2048 nodes of fp64 = 48 KB

template <typename T>
struct DiffInputs {
  const int* inst;      // (P,) instance of each particle
  const int* mat;       // (P,) material (< 0: dead)
  const T* fiber;       // (P, 2)
  const T* amp_scale;   // (P,) per-particle actuation multiplier
  const T* params;      // (M, 5): amp, contraction, freq, payload density, E_jelly
};

template <typename T>
__device__ __forceinline__ void inst_params(const DiffInputs<T>& in, int m, T t,
                                            const Consts<T>& k, Act<T>& act, T& mu, T& la,
                                            T& amp, T& pdens) {
  const T* pr = in.params + 5 * m;
  act = activation(t, pr[2], pr[1], k);
  const T E = pr[4], nu = k.nu_jelly;
  mu = E / (T(2) * (T(1) + nu));
  la = E * nu / ((T(1) + nu) * (T(1) - T(2) * nu));
  amp = pr[0];
  pdens = pr[3];
}

template <typename T>
__global__ void k_d_fill(T* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = T(0);
}

template <typename T>
__device__ __forceinline__ void zero_span(T* p, int64_t n) {
  if (!p) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = T(0);
}

// ---- workgroup tile ------------------------------------------------------------------

struct Tile {
  int ox, oy, tw, th, inst;
  bool use;
};

__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

// Bounding box of the active particles' base cells over the workgroup, and whether they
// share one instance and fit the LDS tile. Called by every thread of the workgroup.
__device__ Tile block_tile(bool active, int inst, int b0, int b1) {
  __shared__ int red[kDiffBlock / 64][6];
  int v[6] = {active ? b0 : INT_MAX,   active ? -b0 : INT_MAX,   active ? b1 : INT_MAX,
              active ? -b1 : INT_MAX,  active ? inst : INT_MAX,  active ? -inst : INT_MAX};
#pragma unroll
  for (int q = 0; q < 6; ++q) v[q] = wave_min(v[q]);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 6; ++q) red[wave][q] = v[q];
  }
  __syncthreads();
  int r[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    r[q] = red[0][q];
    for (int w = 1; w < kDiffBlock / 64; ++w) r[q] = min(r[q], red[w][q]);
  }
  Tile t;
  const bool any = r[0] != INT_MAX;
  t.ox = r[0];
  t.oy = r[2];
  t.inst = r[4];
  t.tw = any ? -r[1] - r[0] + 3 : 0;
  t.th = any ? -r[3] - r[2] + 3 : 0;
  t.use = any && r[4] == -r[5] && t.tw * t.th <= kTileNodes;
  return t;
}

// Adds the LDS tile (3 planes of tw*th) into the instance's grid planes.
template <typename T>
__device__ __forceinline__ void tile_flush(const double* lds, const Tile& t, T* g, int nx,
                                           int G) {
  const int n = t.tw * t.th;
  for (int li = threadIdx.x; li < n; li += kDiffBlock) {
    const int jj = li / t.tw;
    const int node = (t.oy + jj) * nx + t.ox + (li - jj * t.tw);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T v = T(lds[c * n + li]);
      if (v != T(0)) unsafeAtomicAdd(g + (size_t)c * G + node, v);
    }
  }
}

template <typename T>
struct GlobalEmit {
  T* g;
  int nx, G;
  __device__ __forceinline__ void operator()(int i, int j, T c0, T c1, T c2) const {
    const int node = j * nx + i;
    unsafeAtomicAdd(g + node, c0);
    unsafeAtomicAdd(g + G + node, c1);
    unsafeAtomicAdd(g + 2 * (size_t)G + node, c2);
  }
};

// ---- forward ---------------------------------------------------------------------------

// P2G of the substep at time t: scatter from `cur` into `grid` (zeroed), write the updated
// F (and x, v, C unchanged) to `nxt` when non-null.
template <typename T>
__global__ void __launch_bounds__(kDiffBlock) k_d_p2g(DiffInputs<T> in, const T* cur,
                                                      T* nxt, T* grid, int P, int G,
                                                      T t, Consts<T> k) {
  __shared__ double lds[3 * kTileNodes];
  const int p = blockIdx.x * kDiffBlock + threadIdx.x;
  const bool valid = p < P;
  PState<T> s = {T(0), T(0), T(0), T(0), {T(1), T(0), T(0), T(1)}, {T(1), T(0), T(0), T(1)}};
  int mt = -1, m = 0;
  if (valid) {
    s = load_state(cur + 12 * (size_t)p);
    mt = in.mat[p];
    m = in.inst[p];
  }
  int b0 = 0, b1 = 0;
  T fx0, fx1;
  const bool active = valid && mt >= 0 && base_cell(s.x0, s.x1, k, b0, b1, fx0, fx1);
  const Tile tl = block_tile(active, m, b0, b1);
  Act<T> act = {T(0), T(0), T(0)};
  T mu = T(0), la = T(0), amp = T(0), pdens = T(0), d0 = T(0), d1 = T(0), as = T(0);
  if (active) {
    inst_params(in, m, t, k, act, mu, la, amp, pdens);
    d0 = in.fiber[2 * p];
    d1 = in.fiber[2 * p + 1];
    as = in.amp_scale[p];
  }
  const int n = tl.tw * tl.th;
  const GlobalEmit<T> direct{grid + (size_t)m * 3 * G, k.nx, G};
  if (tl.use) {
    for (int i = threadIdx.x; i < 3 * n; i += kDiffBlock) lds[i] = 0.0;
    __syncthreads();
  }
  if (active)
    p2g_particle(s, mt, d0, d1, act.a * amp * as, mu, la, pdens, k,
                 [&](int i, int j, T c0, T c1, T c2) {
                   if (tl.use) {
                     const int li = (j - tl.oy) * tl.tw + (i - tl.ox);
                     atomicAdd(&lds[li], double(c0));
                     atomicAdd(&lds[n + li], double(c1));
                     atomicAdd(&lds[2 * n + li], double(c2));
                   } else {
                     direct(i, j, c0, c1, c2);
                   }
                 });
  if (tl.use) {
    __syncthreads();
    tile_flush(lds, tl, grid + (size_t)tl.inst * 3 * G, k.nx, G);
  }
  if (valid && nxt) store_state(nxt + 12 * (size_t)p, s);
}

// G2P of the substep (with the grid op per stencil node): x, v, C from `cur`, F from `nxt`
// (already updated), result to `nxt`. Also zeroes `zero` (zero_n floats), the grid the next
// substep accumulates into.
template <typename T>
__global__ void __launch_bounds__(kDiffBlock) k_d_g2p(DiffInputs<T> in, const T* cur,
                                                      T* nxt, const T* grid,
                                                      T* zero, int64_t zero_n, int P, int G,
                                                      Consts<T> k) {
  zero_span(zero, zero_n);
  const int p = blockIdx.x * kDiffBlock + threadIdx.x;
  if (p >= P) return;
  PState<T> s = load_state(cur + 12 * (size_t)p);
  const T* fn = nxt + 12 * (size_t)p + 8;
  s.F = {fn[0], fn[1], fn[2], fn[3]};
  if (in.mat[p] >= 0) {
    const T* g = grid + (size_t)in.inst[p] * 3 * G;
    g2p_particle(s, k, g, g + G, g + 2 * (size_t)G);
  }
  store_state(nxt + 12 * (size_t)p, s);
}

// ---- adjoint ---------------------------------------------------------------------------

// Adjoint of G2P (and of the grid op): from bout (adjoint of the state after the substep)
// to the partial adjoint of the state before (x, and v, C of particles outside the P2G
// range) in bmid; scatters the momentum/mass adjoints into `agrid` (zeroed).
template <typename T>
__global__ void __launch_bounds__(kDiffBlock) k_d_g2p_bwd(DiffInputs<T> in, const T* cur,
                                                          const T* bout, T* bmid,
                                                          const T* grid, T* agrid,
                                                          int P, int G, Consts<T> k) {
  __shared__ double lds[3 * kTileNodes];
  const int p = blockIdx.x * kDiffBlock + threadIdx.x;
  const bool valid = p < P;
  PState<T> s = {T(0), T(0), T(0), T(0), {T(1), T(0), T(0), T(1)}, {T(1), T(0), T(0), T(1)}};
  PState<T> bo = {T(0), T(0), T(0), T(0), {T(0), T(0), T(0), T(0)}, {T(0), T(0), T(0), T(0)}};
  int mt = -1, m = 0;
  if (valid) {
    s = load_state(cur + 12 * (size_t)p);
    bo = load_state(bout + 12 * (size_t)p);
    mt = in.mat[p];
    m = in.inst[p];
  }
  int b0 = 0, b1 = 0;
  T fx0, fx1;
  const bool active = valid && mt >= 0 && base_cell(s.x0, s.x1, k, b0, b1, fx0, fx1);
  const Tile tl = block_tile(active, m, b0, b1);
  PState<T> bi = {T(0), T(0), T(0), T(0), {T(0), T(0), T(0), T(0)}, {T(0), T(0), T(0), T(0)}};
  if (valid && mt < 0) bi = bo;  // dead slots are never updated
  const T* g = grid + (size_t)m * 3 * G;
  const int n = tl.tw * tl.th;
  const GlobalEmit<T> direct{agrid + (size_t)m * 3 * G, k.nx, G};
  if (tl.use) {
    for (int i = threadIdx.x; i < 3 * n; i += kDiffBlock) lds[i] = 0.0;
    __syncthreads();
  }
  if (valid && mt >= 0)
    g2p_backward(s, bo, k, g, g + G, g + 2 * (size_t)G, bi,
                 [&](int i, int j, T c0, T c1, T c2) {
                   if (tl.use) {
                     const int li = (j - tl.oy) * tl.tw + (i - tl.ox);
                     atomicAdd(&lds[li], double(c0));
                     atomicAdd(&lds[n + li], double(c1));
                     atomicAdd(&lds[2 * n + li], double(c2));
                   } else {
                     direct(i, j, c0, c1, c2);
                   }
                 });
  if (tl.use) {
    __syncthreads();
    tile_flush(lds, tl, agrid + (size_t)tl.inst * 3 * G, k.nx, G);
  }
  if (valid) store_state(bmid + 12 * (size_t)p, bi);
}

// Adjoint of P2G (gather) and of the constitutive update; completes the adjoint of the
// state before the substep in `bin` and accumulates parameter adjoints per particle:
// gamp[p] (actuation multiplier), gpp[p] = (E_jelly, amp, contraction, freq). Also zeroes
// `zero` (zero_n floats), the adjoint grid of the next (earlier) substep.
template <typename T>
__global__ void __launch_bounds__(kDiffBlock) k_d_p2g_bwd(DiffInputs<T> in, const T* cur,
                                                          const T* bout, const T* bmid,
                                                          T* bin, const T* agrid,
                                                          T* zero, int64_t zero_n, int P,
                                                          int G, T t, Consts<T> k,
                                                          T* gamp, T* gpp) {
  zero_span(zero, zero_n);
  const int p = blockIdx.x * kDiffBlock + threadIdx.x;
  if (p >= P) return;
  PState<T> bi = load_state(bmid + 12 * (size_t)p);
  const int mt = in.mat[p];
  if (mt >= 0) {
    const int m = in.inst[p];
    PState<T> s = load_state(cur + 12 * (size_t)p);
    const T* fb = bout + 12 * (size_t)p + 8;
    const Mat<T> Fnb = {fb[0], fb[1], fb[2], fb[3]};
    Act<T> act;
    T mu, la, amp, pdens;
    inst_params(in, m, t, k, act, mu, la, amp, pdens);
    const T* a = agrid + (size_t)m * 3 * G;
    T mub = T(0), lab = T(0), betab = T(0);
    const T as = in.amp_scale[p];
    p2g_backward_particle(s, mt, in.fiber[2 * p], in.fiber[2 * p + 1], act.a * amp * as, mu, la,
                          pdens, k, a, a + G, a + 2 * (size_t)G, Fnb, bi, mub, lab, betab);
    if (mt == 1 || mt == 3) {
      const T nu = k.nu_jelly;
      gpp[4 * (size_t)p] += mub / (T(2) * (T(1) + nu)) + lab * nu / ((T(1) + nu) * (T(1) - T(2) * nu));
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
