// HIP kernels of the differentiable path: a batched substep and its adjoint over explicit
// device buffers (torch tensors), particles in any order with an instance index each.
// Math in adjoint.h.
//
// Per instance a grid holds 3 planes of nx*ny floats: momentum x, momentum y, mass. The
// grid op is node-local, so G2P evaluates it at each stencil node, and the G2P adjoint
// scatters straight into the momentum/mass adjoints. The G2P half of one substep and the
// P2G half of the next are particle-local, so one kernel runs both (and likewise for the
// adjoint halves): a block of n substeps is n + 1 launches forward and n + 1 backward.
//
// The two scatters (P2G, G2P adjoint) accumulate in an LDS tile covering the workgroup's
// stencil nodes when the workgroup's particles are of one instance and close together
// (particles packed instance by instance in spatial order, as tanks are filled), then add
// the tile to the grid with one global atomic per touched node and plane; otherwise they
// use global atomics directly. Adjacent lanes of one DPP row whose particles share a base
// cell add their contributions pairwise first, so only one lane of a pair issues atomics.
// The fp32 P2G tile accumulates in 64-bit fixed point (exact); the fp64 P2G tile and the
// adjoint tile (whose values have no fixed range) in fp64 (gfx950's fp64 LDS atomic add is
// much faster than its fp32 one). Every
// grid a kernel accumulates into starts zeroed: grids rotate through three buffers and
// each kernel clears the one the kernel after next accumulates into (grid-stride), or a
// fill kernel clears them at the start of a block.
#pragma once

#include <climits>

#include "adjoint.h"

namespace jmd {

constexpr int kDiffBlock = 256;   // threads per workgroup, one particle each
constexpr int kTileNodes = 1024;  // forward LDS window/tile: 3 planes x 1024 nodes of fp64 (24 KB)
constexpr int kWinB = 1024;       // adjoint LDS window: 6 staged + 3 accumulated per node

template <typename T>
struct DiffInputs {
  const int* inst;      // (P,) instance of each particle
  const int* mat;       // (P,) material (< 0: dead)
  const T* fiber;       // (P, 2)
  const T* amp_scale;   // (P,) per-particle actuation multiplier
  const T* params;      // (M, 5): amp, contraction, freq, payload density, E_jelly
};

// The per-instance parameters a particle of material `mt` uses: jelly/muscle Lame
// parameters from E_jelly (nu = 0.3), the actuation waveform and amplitude for muscle, the
// payload density for payload (zeros otherwise: water, the majority, needs none of them).
template <typename T>
__device__ __forceinline__ void inst_params(const DiffInputs<T>& in, int m, int mt, T t,
                                            const Consts<T>& k, Act<T>& act, T& mu, T& la,
                                            T& amp, T& pdens) {
  const T* pr = in.params + 5 * m;
  act = {T(0), T(0), T(0)};
  mu = la = amp = pdens = T(0);
  if (mt == 1 || mt == 3) {
    const T E = pr[4], nu = k.nu_jelly;
    mu = E / (T(2) * (T(1) + nu));
    la = E * nu / ((T(1) + nu) * (T(1) - T(2) * nu));
  }
  if (mt == 3) {
    act = activation(t, pr[2], pr[1], k);
    amp = pr[0];
  }
  if (mt == 2) pdens = pr[3];
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

// Bounding box of the active particles' base cells over the workgroup, and whether they
// share one instance and their stencil window (box + 2-node halo) fits `cap` nodes.
// Called by every thread of the workgroup; ends with a barrier, so it can be called again.
__device__ Tile block_tile(bool active, int inst, int b0, int b1, int cap = kTileNodes) {
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
  __syncthreads();
  Tile t;
  const bool any = r[0] != INT_MAX;
  t.ox = r[0];
  t.oy = r[2];
  t.inst = r[4];
  t.tw = any ? -r[1] - r[0] + 3 : 0;
  t.th = any ? -r[3] - r[2] + 3 : 0;
  t.use = any && r[4] == -r[5] && t.tw * t.th <= cap;
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

// ---- cross-lane pre-sums and fixed point ------------------------------------------------

// Pairs of adjacent lanes (within a DPP row of 16) whose particles share a base cell (keys
// equal; lanes that scatter nothing carry unique keys): the first lane of each pair in a
// run of equal keys adds its neighbour's contributions and alone issues the LDS atomics.
// Called by every lane of the wave.
struct LanePairs {
  bool same_next, head;
};

__device__ __forceinline__ LanePairs lane_pairs(int key) {
  const int lane = threadIdx.x & 63;
  const int kn = __builtin_amdgcn_update_dpp(INT_MIN, key, 0x101, 0xF, 0xF, false);  // lane+1
  const int kp = __builtin_amdgcn_update_dpp(INT_MIN, key, 0x111, 0xF, 0xF, false);  // lane-1
  const unsigned long long heads = __ballot(kp != key);
  const unsigned long long le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int runstart = 63 - __clzll(heads & le);
  return {kn == key, ((lane - runstart) & 1) == 0};
}

// The value held by lane + 1 of the same DPP row (0 past the row's end).
__device__ __forceinline__ float next_lane(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x101, 0xF, 0xF, false));
}
__device__ __forceinline__ double next_lane(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x101, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x101, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// 64-bit fixed point for the fp32 kernels' P2G tile: c * 2^50 rounded to the nearest
// integer, as two's-complement int64 bits, for |c| < 1 (an error of at most 2^-51, finer
// than fp32 rounding of any contribution above 2^-27). c * 2^50 + 1.5 * 2^52 in double
// (one FMA, exact but for that rounding) lies in [2^52, 2^53), where the mantissa holds
// c * 2^50 + 2^51: the low word is the integer's low word, the high word minus the
// exponent and the 2^51 bit (0x43380000) its high word. A node of a workgroup's tile
// receives at most 256 contributions (one per particle), so its sum stays below 2^58
// units: no overflow. Larger contributions (and NaN) go to the global grid. The fp64
// kernels accumulate in fp64 instead (fixed point would be coarser than fp64 rounding).
__device__ __forceinline__ unsigned long long to_fixed(float c) {
  const unsigned long long b =
      __double_as_longlong(__fma_rn((double)c, 1125899906842624.0, 6755399441055744.0));
  const unsigned hi = (unsigned)(b >> 32) - 0x43380000u;
  return ((unsigned long long)hi << 32) | (unsigned)b;
}
__device__ __forceinline__ float from_fixed(unsigned long long v) {
  return float((double)(long long)v * (1.0 / 1125899906842624.0));  // 2^-50
}

// Adds a 3-plane LDS tile of fixed-point sums into the instance's grid planes.
__device__ __forceinline__ void tile_flush_fixed(const unsigned long long* acc, const Tile& t,
                                                 float* g, int nx, int G) {
  const int n = t.tw * t.th;
  for (int li = threadIdx.x; li < n; li += kDiffBlock) {
    const int jj = li / t.tw;
    const int node = (t.oy + jj) * nx + t.ox + (li - jj * t.tw);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float v = from_fixed(acc[c * n + li]);
      if (v != 0.0f) unsafeAtomicAdd(g + (size_t)c * G + node, v);
    }
  }
}

// ---- forward ---------------------------------------------------------------------------
//
// One kernel per substep boundary: K_f(s) finishes substep s-1 (G2P, with the grid op per
// stencil node) and starts substep s (P2G), per particle, so the state passes through
// memory once per substep. Between kernels a particle is carried as its state (x, v, C,
// and F of that state) plus F_new, the F that substep s's P2G computed (the F of state
// s+1). K_f(0) is P2G only, K_f(n) G2P only.

template <typename T>
__device__ __forceinline__ void load_xvc(const T* p, PState<T>& s) {
  s.x0 = p[0]; s.x1 = p[1]; s.v0 = p[2]; s.v1 = p[3];
  s.C = {p[4], p[5], p[6], p[7]};
}

template <typename T, bool DO_G2P, bool DO_P2G>
__global__ void __launch_bounds__(kDiffBlock)
k_d_fwd(DiffInputs<T> in, const T* src, const T* fnew_in, const T* grid_prev, T* dst,
        T* fnew_out, T* grid_cur, T* zero, int64_t zero_n, int P, int G, T t, Consts<T> k) {
  __shared__ double lds[3 * kTileNodes];
  zero_span(zero, zero_n);
  const int p = blockIdx.x * kDiffBlock + threadIdx.x;
  const bool valid = p < P;
  PState<T> s = {T(0), T(0), T(0), T(0), {T(1), T(0), T(0), T(1)}, {T(1), T(0), T(0), T(1)}};
  int mt = -1, m = 0;
  if (valid) {
    mt = in.mat[p];
    m = in.inst[p];
    if (DO_G2P) {
      // G2P of substep s-1: x, v, C of state s-1, F = F_new of substep s-1
      load_xvc(src + 12 * (size_t)p, s);
      const T* fn = fnew_in + 4 * (size_t)p;
      s.F = {fn[0], fn[1], fn[2], fn[3]};
    } else {
      s = load_state(src + 12 * (size_t)p);
    }
  }
  if constexpr (DO_G2P) {
    // Node velocities of the workgroup's window, grid op applied once per node, in LDS.
    int b0 = 0, b1 = 0;
    T fx0, fx1;
    const bool ing = valid && mt >= 0 && base_cell(s.x0, s.x1, k, b0, b1, fx0, fx1);
    const Tile wg = block_tile(ing, m, b0, b1);
    T* vw = reinterpret_cast<T*>(lds);
    const int wn = wg.tw * wg.th;
    if (wg.use) {
      const T* gp = grid_prev + (size_t)wg.inst * 3 * G;
      for (int li = threadIdx.x; li < wn; li += kDiffBlock) {
        const int jj = li / wg.tw;
        const int i = wg.ox + (li - jj * wg.tw), j = wg.oy + jj;
        const int node = j * k.nx + i;
        gridop_forward(gp[node], gp[G + node], gp[2 * (size_t)G + node], i, j, k, vw[li],
                       vw[wn + li]);
      }
      __syncthreads();
    }
    if (valid && mt >= 0) {
      if (wg.use) {
        g2p_particle_v(s, k, [&](int i, int j, T& a0, T& a1) {
          const int li = (j - wg.oy) * wg.tw + (i - wg.ox);
          a0 = vw[li];
          a1 = vw[wn + li];
        });
      } else {
        const T* g = grid_prev + (size_t)m * 3 * G;
        g2p_particle(s, k, g, g + G, g + 2 * (size_t)G);
      }
    }
    if (valid) store_state(dst + 12 * (size_t)p, s);
    if (DO_P2G && wg.use) __syncthreads();  // the P2G tile reuses the window's LDS
  }
  if constexpr (DO_P2G) {
    int b0 = 0, b1 = 0;
    T fx0, fx1;
    const bool active = valid && mt >= 0 && base_cell(s.x0, s.x1, k, b0, b1, fx0, fx1);
    const Tile tl = block_tile(active, m, b0, b1);
    Act<T> act = {T(0), T(0), T(0)};
    T mu = T(0), la = T(0), amp = T(0), pdens = T(0), d0 = T(0), d1 = T(0), as = T(0);
    if (active) {
      inst_params(in, m, mt, t, k, act, mu, la, amp, pdens);
      if (mt == 3) {  // fibre and multiplier matter for muscle only
        d0 = in.fiber[2 * p];
        d1 = in.fiber[2 * p + 1];
        as = in.amp_scale[p];
      }
    }
    const int n = tl.tw * tl.th;
    T* gi = grid_cur + (size_t)m * 3 * G;
    const GlobalEmit<T> direct{gi, k.nx, G};
    // fp32: 64-bit fixed point in LDS (exact, order-independent sums), a contribution
    // outside the fixed-point range (or NaN) going to the global grid directly; fp64: fp64
    constexpr bool kFixed = sizeof(T) == 4;
    unsigned long long* accf = reinterpret_cast<unsigned long long*>(lds);
    const LanePairs pr = lane_pairs(active && tl.use ? b1 * k.nx + b0 + 1 : -(1 + (int)threadIdx.x));
    if (tl.use) {
      for (int i = threadIdx.x; i < 3 * n; i += kDiffBlock) {
        if (kFixed)
          accf[i] = 0ull;
        else
          lds[i] = 0.0;
      }
      __syncthreads();
    }
    if (active)
      p2g_particle(s, mt, d0, d1, act.a * amp * as, mu, la, pdens, k,
                   [&](int i, int j, T c0, T c1, T c2) {
                     if (tl.use) {
                       const T n0 = next_lane(c0), n1 = next_lane(c1), n2 = next_lane(c2);
                       if (pr.same_next) {
                         c0 += n0;
                         c1 += n1;
                         c2 += n2;
                       }
                       if (pr.head) {
                         const int li = (j - tl.oy) * tl.tw + (i - tl.ox);
                         const T cs[3] = {c0, c1, c2};
#pragma unroll
                         for (int c = 0; c < 3; ++c) {
                           if constexpr (kFixed) {
                             if (fabsf(cs[c]) < 1.0f)
                               atomicAdd(&accf[c * n + li], to_fixed(cs[c]));
                             else
                               unsafeAtomicAdd(gi + (size_t)c * G + j * k.nx + i, cs[c]);
                           } else {
                             atomicAdd(&lds[c * n + li], double(cs[c]));
                           }
                         }
                       }
                     } else {
                       direct(i, j, c0, c1, c2);
                     }
                   });
    if (tl.use) {
      __syncthreads();
      if constexpr (kFixed)
        tile_flush_fixed(accf, tl, grid_cur + (size_t)tl.inst * 3 * G, k.nx, G);
      else
        tile_flush(lds, tl, grid_cur + (size_t)tl.inst * 3 * G, k.nx, G);
    }
    if (valid) {
      T* fo = fnew_out + 4 * (size_t)p;
      fo[0] = s.F.a;  // p2g_particle leaves F unchanged outside the transfer range
      fo[1] = s.F.b;
      fo[2] = s.F.c;
      fo[3] = s.F.d;
    }
  }
}

// ---- adjoint ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Adds the particles' parameter adjoints (E_jelly, amp, contraction, freq) into
// gparams (M, 4): one set of atomics per wave whose contributing particles share an
// instance (the usual case: instances are packed one after another), else one per particle.
__device__ __forceinline__ void add_param_adjoints(const double g[4], bool has, int inst,
                                                   double* gparams) {
  const unsigned long long any = __ballot(has);
  if (any == 0ull) return;
  const int lo = wave_min(has ? inst : INT_MAX), hi = -wave_min(has ? -inst : INT_MAX);
  if (lo == hi) {
    double s[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) s[c] = wave_sum(has ? g[c] : 0.0);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) unsafeAtomicAdd(gparams + 4 * (size_t)lo + c, s[c]);
    }
  } else if (has) {
#pragma unroll
    for (int c = 0; c < 4; ++c) unsafeAtomicAdd(gparams + 4 * (size_t)inst + c, g[c]);
  }
}

//
// One kernel per substep boundary, in reverse: K_b(s) finishes the adjoint of substep s
// (P2G adjoint: gather from the momentum/mass adjoints of substep s, giving a_s, the adjoint
// of state s, in registers) and starts the adjoint of substep s-1 (G2P adjoint of state
// s-1 driven by a_s, scattering into the adjoint grid of substep s-1). Between kernels a
// particle is carried as m (the x, v, C part of the G2P adjoint, 8 scalars) and the F part
// of a_s (4 scalars). K_b(n) is the G2P adjoint only (driven by grad_out), K_b(0) the P2G
// adjoint only (writing grad_in).

template <typename T, bool DO_P2GB, bool DO_G2PB>
__global__ void __launch_bounds__(kDiffBlock)
k_d_bwd(DiffInputs<T> in, const T* st_cur, const T* st_prev, const T* aF_in, int aF_stride,
        const T* m_in, const T* agrid_cur, const T* grid_prev, T* agrid_prev,
        const T* a_full_in, T* m_out, T* aF_out, T* a_full_out, T* zero, int64_t zero_n,
        int P, int G, T t, Consts<T> k, T* gamp, double* gparams) {
  __shared__ double acc[3 * kWinB];  // G2P-adjoint scatter tile
  __shared__ T cw[6 * kWinB];        // staged NodeCoef of the window
  zero_span(zero, zero_n);
  const int p = blockIdx.x * kDiffBlock + threadIdx.x;
  const bool valid = p < P;
  int mt = -1, m = 0;
  if (valid) {
    mt = in.mat[p];
    m = in.inst[p];
  }
  // a: the adjoint of the state between the two halves (state s)
  PState<T> a = {T(0), T(0), T(0), T(0), {T(0), T(0), T(0), T(0)}, {T(0), T(0), T(0), T(0)}};
  double gp4[4] = {0.0, 0.0, 0.0, 0.0};  // this particle's share of (E_jelly, amp, c, f)
  if (DO_P2GB) {
    if (valid) {
      const T* mi = m_in + 8 * (size_t)p;
      a.x0 = mi[0]; a.x1 = mi[1]; a.v0 = mi[2]; a.v1 = mi[3];
      a.C = {mi[4], mi[5], mi[6], mi[7]};
      const T* fb = aF_in + (size_t)aF_stride * p;
      const Mat<T> Fnb = {fb[0], fb[1], fb[2], fb[3]};
      if (mt >= 0) {
        const PState<T> sc = load_state(st_cur + 12 * (size_t)p);
        Act<T> act;
        T mu, la, amp, pdens;
        inst_params(in, m, mt, t, k, act, mu, la, amp, pdens);
        const T* ag = agrid_cur + (size_t)m * 3 * G;
        T mub = T(0), lab = T(0), betab = T(0);
        T as = T(0), d0 = T(0), d1 = T(0);
        if (mt == 3) {  // fibre and multiplier matter for muscle only
          as = in.amp_scale[p];
          d0 = in.fiber[2 * p];
          d1 = in.fiber[2 * p + 1];
        }
        p2g_backward_particle(sc, mt, d0, d1, act.a * amp * as, mu, la, pdens, k, ag, ag + G,
                              ag + 2 * (size_t)G, Fnb, a, mub, lab, betab);
        if (mt == 1 || mt == 3) {
          const T nu = k.nu_jelly;
          gp4[0] = double(mub / (T(2) * (T(1) + nu)) +
                          lab * nu / ((T(1) + nu) * (T(1) - T(2) * nu)));
        }
        if (mt == 3) {
          gamp[p] += betab * act.a * amp;
          gp4[1] = double(betab * act.a * as);
          gp4[2] = double(betab * amp * as * act.da_dc);
          gp4[3] = double(betab * amp * as * act.da_df);
        }
      } else {
        a.F = Fnb;  // dead slots are never updated
      }
      if (!DO_G2PB) store_state(a_full_out + 12 * (size_t)p, a);
    }
    add_param_adjoints(gp4, valid && (mt == 1 || mt == 3), m, gparams);
  } else if (valid) {
    a = load_state(a_full_in + 12 * (size_t)p);
  }
  if constexpr (DO_G2PB) {
    PState<T> sp = {T(0), T(0), T(0), T(0), {T(1), T(0), T(0), T(1)}, {T(1), T(0), T(0), T(1)}};
    if (valid) sp = load_state(st_prev + 12 * (size_t)p);
    int b0 = 0, b1 = 0;
    T fx0, fx1;
    const bool active = valid && mt >= 0 && base_cell(sp.x0, sp.x1, k, b0, b1, fx0, fx1);
    const Tile tl = block_tile(active, m, b0, b1, kWinB);
    PState<T> mo = {T(0), T(0), T(0), T(0), {T(0), T(0), T(0), T(0)}, {T(0), T(0), T(0), T(0)}};
    if (valid && mt < 0) mo = a;  // dead slots pass the adjoint through
    const T* g = grid_prev + (size_t)m * 3 * G;
    const int n = tl.tw * tl.th;
    const GlobalEmit<T> direct{agrid_prev + (size_t)m * 3 * G, k.nx, G};
    if (tl.use) {
      // the window's NodeCoef once per node, and the zeroed scatter tile, in LDS
      const T* gw = grid_prev + (size_t)tl.inst * 3 * G;
      for (int li = threadIdx.x; li < n; li += kDiffBlock) {
        const int jj = li / tl.tw;
        const int i = tl.ox + (li - jj * tl.tw), j = tl.oy + jj;
        const int node = j * k.nx + i;
        const NodeCoef<T> c = node_coef(gw[node], gw[G + node], gw[2 * (size_t)G + node], i, j, k);
        cw[li] = c.v0;
        cw[n + li] = c.v1;
        cw[2 * n + li] = c.A0;
        cw[3 * n + li] = c.A1;
        cw[4 * n + li] = c.B0;
        cw[5 * n + li] = c.B1;
        acc[li] = 0.0;
        acc[n + li] = 0.0;
        acc[2 * n + li] = 0.0;
      }
      __syncthreads();
    }
    const LanePairs pr = lane_pairs(active && tl.use ? b1 * k.nx + b0 + 1 : -(1 + (int)threadIdx.x));
    if (valid && mt >= 0)
      g2p_backward_c(
          sp, a, k,
          [&](int i, int j) {
            if (tl.use) {
              const int li = (j - tl.oy) * tl.tw + (i - tl.ox);
              return NodeCoef<T>{cw[li],         cw[n + li],     cw[2 * n + li],
                                 cw[3 * n + li], cw[4 * n + li], cw[5 * n + li]};
            }
            const int node = j * k.nx + i;
            return node_coef(g[node], g[G + node], g[2 * (size_t)G + node], i, j, k);
          },
          mo,
          [&](int i, int j, T c0, T c1, T c2) {
            if (tl.use) {
              const T n0 = next_lane(c0), n1 = next_lane(c1), n2 = next_lane(c2);
              if (pr.same_next) {
                c0 += n0;
                c1 += n1;
                c2 += n2;
              }
              if (pr.head) {
                const int li = (j - tl.oy) * tl.tw + (i - tl.ox);
                atomicAdd(&acc[li], double(c0));
                atomicAdd(&acc[n + li], double(c1));
                atomicAdd(&acc[2 * n + li], double(c2));
              }
            } else {
              direct(i, j, c0, c1, c2);
            }
          });
    if (tl.use) {
      __syncthreads();
      tile_flush(acc, tl, agrid_prev + (size_t)tl.inst * 3 * G, k.nx, G);
    }
    if (valid) {
      T* mo_p = m_out + 8 * (size_t)p;
      mo_p[0] = mo.x0; mo_p[1] = mo.x1; mo_p[2] = mo.v0; mo_p[3] = mo.v1;
      mo_p[4] = mo.C.a; mo_p[5] = mo.C.b; mo_p[6] = mo.C.c; mo_p[7] = mo.C.d;
      T* af = aF_out + 4 * (size_t)p;
      af[0] = a.F.a; af[1] = a.F.b; af[2] = a.F.c; af[3] = a.F.d;
    }
  }
}

}  // namespace jmd
