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
// use global atomics directly. Lanes of one DPP row whose particles share a base cell add
// their contributions first (groups of 4 in the fp32 P2G, pairs elsewhere), so only one
// lane per group issues atomics. The fp32 P2G tile accumulates in 64-bit fixed point
// (exact); the fp64 P2G tile and the adjoint tile (whose values have no fixed range) in
// fp64 (gfx950's fp64 LDS atomic add is much faster than its fp32 one).
//
// Kernels take two scalar types: T, of the buffers (states, grids, inputs, gradients), and
// A, of the arithmetic inside a substep. T = A = float is the reference's precision;
// T = A = double is fp64 throughout; T = float, A = double keeps fp32 buffers (fp32 states
// and forward grids, as in the reference) but evaluates each substep, and carries the
// adjoint between kernels, in fp64: small strains (F - R of the jelly, J - 1 of the water,
// whose stiffness is 1e5) are differences of numbers near 1 that fp32 arithmetic resolves
// to only ~1e-7, while fp32 storage of the results costs the gradients little. Every
// grid a kernel accumulates into starts zeroed: grids rotate through three buffers and
// each kernel clears the one the kernel after next accumulates into (grid-stride), or a
// fill kernel clears them at the start of a block.
#pragma once

#include <climits>

#include "adjoint.h"
#include "lanes.h"

namespace jmd {

constexpr int kDiffBlock = 256;   // threads per workgroup, one particle each
constexpr int kTileNodes = 1024;  // forward LDS window/tile: 3 planes x 1024 nodes of fp64 (24 KB)
constexpr int kWinB = 512;        // adjoint LDS window: 6 staged + 2 accumulated per node

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
template <typename A, typename T>
__device__ __forceinline__ void inst_params(const DiffInputs<T>& in, int m, int mt, A t,
                                            const Consts<A>& k, Act<A>& act, A& mu, A& la,
                                            A& amp, A& pdens) {
  const T* pr = in.params + 5 * m;
  act = {A(0), A(0), A(0)};
  mu = la = amp = pdens = A(0);
  if (mt == 1 || mt == 3) {
    const A E = A(pr[4]), nu = k.nu_jelly;
    mu = E / (A(2) * (A(1) + nu));
    la = E * nu / ((A(1) + nu) * (A(1) - A(2) * nu));
  }
  if (mt == 3) {
    act = activation(t, A(pr[2]), A(pr[1]), k);
    amp = A(pr[0]);
  }
  if (mt == 2) pdens = A(pr[3]);
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

using jm::wave_min;

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


// fp32 P2G scatter of one particle per lane into the workgroup's LDS tile (3 planes of
// kTileNodes 64-bit fixed-point sums, jm::fixed_bits) and its flush, called by every thread
// of the workgroup, from the particle's mass, affine term A and velocity (p2g_forward's
// results, rounded to fp32): the contributions of p2g_particle, w (m v + A (x_node - x)) and
// w m, evaluated in the separable form wy_j P_i + ey_j Q_i (P_i = wx_i (m v + A_0 dx_i),
// Q_i = wx_i A_1, ey_j = wy_j dy_j), then summed over same-cell groups of 4 lanes
// (jm::lane_groups / presum3) so that one lane per group issues the atomics. A particle
// whose contribution bounds reach 0.75 (or are NaN), or any particle when the tile
// is not used, scatters to the global grid with fp32 atomics and then carries zeros; lanes
// without a particle carry zeros computed from finite stand-ins, so every lane's values are
// finite in the pre-sums.
__device__ __forceinline__ void p2g_scatter_fixed(bool active, float mass, const Mat<float>& A,
                                                  float v0, float v1, float dx, int nx, int b0,
                                                  int b1, float fx0, float fx1, const Tile& tl,
                                                  double* lds, float* gi, float* gtile, int G) {
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(lds);
  const int n = tl.tw * tl.th;
  if (tl.use) {
    for (int i = threadIdx.x; i < n; i += kDiffBlock) {
      acc[i] = 0ull;
      acc[kTileNodes + i] = 0ull;
      acc[2 * kTileNodes + i] = 0ull;
    }
    __syncthreads();
  }
  const float f0 = active ? fx0 : 0.5f, f1 = active ? fx1 : 0.5f;
  const float mv0 = mass * (active ? v0 : 0.f), mv1 = mass * (active ? v1 : 0.f);
  float wx[3], wy[3], dw[3], ey[3], P0[3], Q0[3], P1[3], Q1[3], Mi[3];
  weights(f0, wx, dw);
  weights(f1, wy, dw);
  // Each of the 27 contributions is at most 0.5625 times these bounds and up to 4 lanes add
  // before one encode, so values stay inside fixed_bits' exact range (|c| < 2) with a
  // bound of 0.75. No fmaxf: it drops NaN operands, and each comparison fails on NaN.
  const float bnd = 1.5f * dx;
  const bool lds_ok = active && tl.use &&
                      fabsf(mv0) + bnd * (fabsf(A.a) + fabsf(A.b)) < 0.75f &&
                      fabsf(mv1) + bnd * (fabsf(A.c) + fabsf(A.d)) < 0.75f && mass < 0.75f &&
                      f0 + f1 < 4.0f;
  const float sc = lds_ok ? jm::kFixPre : (active ? 1.0f : 0.0f);
#pragma unroll
  for (int ii = 0; ii < 3; ++ii) {
    const float ai = wx[ii] * sc;
    const float bi = ai * ((float(ii) - f0) * dx);
    P0[ii] = ai * mv0 + bi * A.a;
    Q0[ii] = ai * A.b;
    P1[ii] = ai * mv1 + bi * A.c;
    Q1[ii] = ai * A.d;
    Mi[ii] = ai * mass;
  }
#pragma unroll
  for (int jj = 0; jj < 3; ++jj) ey[jj] = wy[jj] * ((float(jj) - f1) * dx);
  if (active && !lds_ok) {
    for (int jj = 0; jj < 3; ++jj)
      for (int ii = 0; ii < 3; ++ii) {
        const int node = (b1 + jj) * nx + (b0 + ii);
        unsafeAtomicAdd(gi + node, wy[jj] * P0[ii] + ey[jj] * Q0[ii]);
        unsafeAtomicAdd(gi + G + node, wy[jj] * P1[ii] + ey[jj] * Q1[ii]);
        unsafeAtomicAdd(gi + 2 * (size_t)G + node, wy[jj] * Mi[ii]);
      }
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
      P0[ii] = Q0[ii] = P1[ii] = Q1[ii] = Mi[ii] = 0.f;
      wy[ii] = ey[ii] = 0.f;
    }
  }
  const jm::LaneGroups lg =
      jm::lane_groups(lds_ok ? b1 * nx + b0 + 1 : -(1 + (int)(threadIdx.x & 63)), lds_ok);
  const int q0 = (b1 - tl.oy) * tl.tw + (b0 - tl.ox);
#pragma unroll
  for (int jj = 0; jj < 3; ++jj) {
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
      float c0 = wy[jj] * P0[ii] + ey[jj] * Q0[ii], c1 = wy[jj] * P1[ii] + ey[jj] * Q1[ii],
            c2 = wy[jj] * Mi[ii];
      jm::presum3(c0, c1, c2, lg.f1, lg.f2);
      if (lg.emit) {
        const int q = q0 + jj * tl.tw + ii;
        atomicAdd(&acc[q], jm::fixed_bits(c0));
        atomicAdd(&acc[kTileNodes + q], jm::fixed_bits(c1));
        atomicAdd(&acc[2 * kTileNodes + q], jm::fixed_bits(c2));
      }
    }
  }
  if (tl.use) {
    __syncthreads();
    for (int li = threadIdx.x; li < n; li += kDiffBlock) {
      const int jj = li / tl.tw;
      const int node = (tl.oy + jj) * nx + tl.ox + (li - jj * tl.tw);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float v = jm::from_fixed_bits(acc[c * kTileNodes + li]);
        if (v != 0.0f) unsafeAtomicAdd(gtile + (size_t)c * G + node, v);
      }
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

// State buffers of fp32 with fp64 arithmetic (sizeof(T) < sizeof(A)) hold the deformation
// gradient as F - I: near the identity fp32 keeps F - I to its own relative precision where
// F itself would round it to ~6e-8 absolute, and a strain of 1e-5 to two digits.
template <typename T, typename A>
constexpr bool kShiftF = sizeof(T) < sizeof(A);

template <typename A, typename T>
__device__ __forceinline__ Mat<A> load_F(const T* f) {
  const A one = kShiftF<T, A> ? A(1) : A(0);
  return {A(f[0]) + one, A(f[1]), A(f[2]), A(f[3]) + one};
}

template <typename T, typename A>
__device__ __forceinline__ void store_F(T* f, const Mat<A>& F) {
  const A one = kShiftF<T, A> ? A(1) : A(0);
  f[0] = T(F.a - one);
  f[1] = T(F.b);
  f[2] = T(F.c);
  f[3] = T(F.d - one);
}

// A particle state from a state buffer (F as stored: see kShiftF), and back.
template <typename A, typename T>
__device__ __forceinline__ PState<A> load_particle(const T* p) {
  PState<A> s = load_state_as<A>(p);
  s.F = load_F<A>(p + 8);
  return s;
}

template <typename T, typename A>
__device__ __forceinline__ void store_particle(T* p, const PState<A>& s) {
  store_state(p, s);
  store_F(p + 8, s.F);
}

template <typename A, typename T>
__device__ __forceinline__ void load_xvc(const T* p, PState<A>& s) {
  s.x0 = A(p[0]); s.x1 = A(p[1]); s.v0 = A(p[2]); s.v1 = A(p[3]);
  s.C = {A(p[4]), A(p[5]), A(p[6]), A(p[7])};
}

template <typename T, typename A, bool DO_G2P, bool DO_P2G>
__global__ void __launch_bounds__(kDiffBlock)
k_d_fwd(DiffInputs<T> in, const T* src, const T* fnew_in, const T* grid_prev, T* dst,
        T* fnew_out, T* grid_cur, T* zero, int64_t zero_n, int P, int G, A t, Consts<A> k) {
  static_assert(sizeof(A) >= sizeof(T), "arithmetic at least as wide as the buffers");
  __shared__ double lds[3 * kTileNodes];
  zero_span(zero, zero_n);
  const int p = blockIdx.x * kDiffBlock + threadIdx.x;
  const bool valid = p < P;
  PState<A> s = {A(0), A(0), A(0), A(0), {A(1), A(0), A(0), A(1)}, {A(1), A(0), A(0), A(1)}};
  int mt = -1, m = 0;
  if (valid) {
    mt = in.mat[p];
    m = in.inst[p];
    if (DO_G2P) {
      // G2P of substep s-1: x, v, C of state s-1, F = F_new of substep s-1
      load_xvc(src + 12 * (size_t)p, s);
      s.F = load_F<A>(fnew_in + 4 * (size_t)p);
    } else {
      s = load_particle<A>(src + 12 * (size_t)p);
    }
  }
  if constexpr (DO_G2P) {
    // Node velocities of the workgroup's window, grid op applied once per node, in LDS.
    int b0 = 0, b1 = 0;
    A fx0, fx1;
    const bool ing = valid && mt >= 0 && base_cell(s.x0, s.x1, k, b0, b1, fx0, fx1);
    const Tile wg = block_tile(ing, m, b0, b1);
    A* vw = reinterpret_cast<A*>(lds);  // 2 x kTileNodes values of A fit the tile's bytes
    const int wn = wg.tw * wg.th;
    if (wg.use) {
      const T* gp = grid_prev + (size_t)wg.inst * 3 * G;
      for (int li = threadIdx.x; li < wn; li += kDiffBlock) {
        const int jj = li / wg.tw;
        const int i = wg.ox + (li - jj * wg.tw), j = wg.oy + jj;
        const int node = j * k.nx + i;
        gridop_forward(A(gp[node]), A(gp[G + node]), A(gp[2 * (size_t)G + node]), i, j, k,
                       vw[li], vw[wn + li]);
      }
      __syncthreads();
    }
    if (valid && mt >= 0) {
      if (wg.use) {
        g2p_particle_v(s, k, [&](int i, int j, A& a0, A& a1) {
          const int li = (j - wg.oy) * wg.tw + (i - wg.ox);
          a0 = vw[li];
          a1 = vw[wn + li];
        });
      } else {
        const T* g = grid_prev + (size_t)m * 3 * G;
        g2p_particle(s, k, g, g + G, g + 2 * (size_t)G);
      }
    }
    if (valid) store_particle(dst + 12 * (size_t)p, s);
    if (DO_P2G && wg.use) __syncthreads();  // the P2G tile reuses the window's LDS
  }
  if constexpr (DO_P2G) {
    int b0 = 0, b1 = 0;
    A fx0 = A(0.5), fx1 = A(0.5);
    const bool active = valid && mt >= 0 && base_cell(s.x0, s.x1, k, b0, b1, fx0, fx1);
    const Tile tl = block_tile(active, m, b0, b1);
    Act<A> act = {A(0), A(0), A(0)};
    A mu = A(0), la = A(0), amp = A(0), pdens = A(0), d0 = A(0), d1 = A(0), as = A(0);
    if (active) {
      inst_params(in, m, mt, t, k, act, mu, la, amp, pdens);
      if (mt == 3) {  // fibre and multiplier matter for muscle only
        d0 = A(in.fiber[2 * p]);
        d1 = A(in.fiber[2 * p + 1]);
        as = A(in.amp_scale[p]);
      }
    }
    const int n = tl.tw * tl.th;
    T* gi = grid_cur + (size_t)m * 3 * G;
    if constexpr (sizeof(T) == 4) {
      // the particle update in A, the scatter of its fp32-rounded results
      float mass = 0.f;
      Mat<float> Af = {0.f, 0.f, 0.f, 0.f};
      if (active) {
        Mat<A> Fn, Aa;
        A ma;
        p2g_forward(mt, s.F, s.C, d0, d1, act.a * amp * as, mu, la, pdens, k, Fn, Aa, ma);
        s.F = Fn;
        mass = float(ma);
        Af = {float(Aa.a), float(Aa.b), float(Aa.c), float(Aa.d)};
      }
      p2g_scatter_fixed(active, mass, Af, float(s.v0), float(s.v1), float(k.dx), k.nx, b0,
                        b1, float(fx0), float(fx1), tl, lds, gi,
                        grid_cur + (size_t)tl.inst * 3 * G, G);
    } else {
      // fp64: an fp64 LDS tile (fixed point would be coarser than fp64 rounding), lane pairs
      const GlobalEmit<T> direct{gi, k.nx, G};
      const LanePairs pr =
          lane_pairs(active && tl.use ? b1 * k.nx + b0 + 1 : -(1 + (int)threadIdx.x));
      if (tl.use) {
        for (int i = threadIdx.x; i < 3 * n; i += kDiffBlock) lds[i] = 0.0;
        __syncthreads();
      }
      if (active)
        p2g_particle(s, mt, d0, d1, act.a * amp * as, mu, la, pdens, k,
                     [&](int i, int j, A c0, A c1, A c2) {
                       if (tl.use) {
                         const A n0 = next_lane(c0), n1 = next_lane(c1), n2 = next_lane(c2);
                         if (pr.same_next) {
                           c0 += n0;
                           c1 += n1;
                           c2 += n2;
                         }
                         if (pr.head) {
                           const int li = (j - tl.oy) * tl.tw + (i - tl.ox);
                           atomicAdd(&lds[li], double(c0));
                           atomicAdd(&lds[n + li], double(c1));
                           atomicAdd(&lds[2 * n + li], double(c2));
                         }
                       } else {
                         direct(i, j, T(c0), T(c1), T(c2));
                       }
                     });
      if (tl.use) {
        __syncthreads();
        tile_flush(lds, tl, grid_cur + (size_t)tl.inst * 3 * G, k.nx, G);
      }
    }
    if (valid) store_F(fnew_out + 4 * (size_t)p, s.F);  // unchanged outside the range
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
//
// A launch covers particles [p0, P). WATER: every particle of the range is water
// (material 0; the caller packs the water after the body's particles), so the material is
// a compile-time constant and the constitutive adjoint of the other materials (the SVD
// derivative of jelly, muscle and payload) drops out of the kernel: 126 instead of 155
// VGPRs with fp32 arithmetic (four waves per SIMD instead of three), 246 instead of 256
// with fp64 arithmetic (two instead of one). Water whose F is degenerate still takes the
// full SVD path, as in the general kernel.

template <typename T, typename A, bool DO_P2GB, bool DO_G2PB, bool WATER = false>
__global__ void __launch_bounds__(kDiffBlock)
k_d_bwd(DiffInputs<T> in, const T* st_cur, const T* st_prev, const A* aF_in, int aF_stride,
        const A* m_in, const A* agrid_cur, const T* grid_prev, A* agrid_prev,
        const T* a_full_in, A* m_out, A* aF_out, T* a_full_out, A* zero, int64_t zero_n,
        int p0, int P, int G, A t, Consts<A> k, T* gamp, double* gparams) {
  static_assert(sizeof(A) >= sizeof(T), "arithmetic at least as wide as the buffers");
  __shared__ double acc[2 * kWinB];  // G2P-adjoint tile: node-velocity adjoints (x, y)
  __shared__ A cw[6 * kWinB];        // staged NodeCoef of the window
  zero_span(zero, zero_n);
  const int p = p0 + blockIdx.x * kDiffBlock + threadIdx.x;
  const bool valid = p < P;
  int mt = -1, m = 0;
  if (valid) {
    mt = WATER ? 0 : in.mat[p];
    m = in.inst[p];
  }
  // x and v of state s-1, all that the G2P adjoint reads of it, loaded before the P2G
  // adjoint so that their latency hides behind it (tall batch: 237 -> 229 us per substep
  // with fp32 arithmetic, 434 -> 407 us with fp64 over fp32 buffers), except in the fp32
  // water-only kernel, where the 4 registers held across it cost a wave per SIMD (126 ->
  // 130 VGPRs; 242 -> 263 us) and the load stays where it is used
  constexpr bool kEarlyPrev = !(WATER && sizeof(A) == sizeof(float));
  PState<A> sp = {A(0), A(0), A(0), A(0), {A(1), A(0), A(0), A(1)}, {A(1), A(0), A(0), A(1)}};
  const auto load_prev = [&]() {
    const T* q = st_prev + 12 * (size_t)p;
    sp.x0 = A(q[0]); sp.x1 = A(q[1]); sp.v0 = A(q[2]); sp.v1 = A(q[3]);
  };
  if (DO_G2PB && kEarlyPrev && valid) load_prev();
  // a: the adjoint of the state between the two halves (state s)
  PState<A> a = {A(0), A(0), A(0), A(0), {A(0), A(0), A(0), A(0)}, {A(0), A(0), A(0), A(0)}};
  double gp4[4] = {0.0, 0.0, 0.0, 0.0};  // this particle's share of (E_jelly, amp, c, f)
  if (DO_P2GB) {
    if (valid) {
      const A* mi = m_in + 8 * (size_t)p;
      a.x0 = mi[0]; a.x1 = mi[1]; a.v0 = mi[2]; a.v1 = mi[3];
      a.C = {mi[4], mi[5], mi[6], mi[7]};
      const A* fb = aF_in + (size_t)aF_stride * p;
      const Mat<A> Fnb = {fb[0], fb[1], fb[2], fb[3]};
      if (mt >= 0) {
        const PState<A> sc = load_particle<A>(st_cur + 12 * (size_t)p);
        Act<A> act;
        A mu, la, amp, pdens;
        inst_params(in, m, mt, t, k, act, mu, la, amp, pdens);
        const A* ag = agrid_cur + (size_t)m * 3 * G;
        A mub = A(0), lab = A(0), betab = A(0);
        A as = A(0), d0 = A(0), d1 = A(0);
        if (mt == 3) {  // fibre and multiplier matter for muscle only
          as = A(in.amp_scale[p]);
          d0 = A(in.fiber[2 * p]);
          d1 = A(in.fiber[2 * p + 1]);
        }
        p2g_backward_particle(sc, mt, d0, d1, act.a * amp * as, mu, la, pdens, k, ag, ag + G,
                              ag + 2 * (size_t)G, Fnb, a, mub, lab, betab);
        if (mt == 1 || mt == 3) {
          const A nu = k.nu_jelly;
          gp4[0] = double(mub / (A(2) * (A(1) + nu)) +
                          lab * nu / ((A(1) + nu) * (A(1) - A(2) * nu)));
        }
        if (mt == 3) {
          gamp[p] += T(betab * act.a * amp);
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
    a = load_state_as<A>(a_full_in + 12 * (size_t)p);
  }
  if constexpr (DO_G2PB) {
    if (!kEarlyPrev && valid) load_prev();
    int b0 = 0, b1 = 0;
    A fx0, fx1;
    const bool active = valid && mt >= 0 && base_cell(sp.x0, sp.x1, k, b0, b1, fx0, fx1);
    const Tile tl = block_tile(active, m, b0, b1, kWinB);
    PState<A> mo = {A(0), A(0), A(0), A(0), {A(0), A(0), A(0), A(0)}, {A(0), A(0), A(0), A(0)}};
    if (valid && mt < 0) mo = a;  // dead slots pass the adjoint through
    const T* g = grid_prev + (size_t)m * 3 * G;
    const int n = tl.tw * tl.th;
    if (tl.use) {
      // the window's NodeCoef once per node, and the zeroed velocity-adjoint tile, in LDS
      const T* gw = grid_prev + (size_t)tl.inst * 3 * G;
      for (int li = threadIdx.x; li < n; li += kDiffBlock) {
        const int jj = li / tl.tw;
        const int i = tl.ox + (li - jj * tl.tw), j = tl.oy + jj;
        const int node = j * k.nx + i;
        const NodeCoef<A> c = node_coef(A(gw[node]), A(gw[G + node]),
                                        A(gw[2 * (size_t)G + node]), i, j, k);
        cw[li] = c.v0;
        cw[n + li] = c.v1;
        cw[2 * n + li] = c.A0;
        cw[3 * n + li] = c.A1;
        cw[4 * n + li] = c.B0;
        cw[5 * n + li] = c.B1;
        acc[li] = 0.0;
        acc[n + li] = 0.0;
      }
      __syncthreads();
    }
    SepFactors<A> sf;
    sf.inr = false;
#pragma unroll
    for (int q = 0; q < 3; ++q)
      sf.wy[q] = sf.ey[q] = sf.P0[q] = sf.Q0[q] = sf.P1[q] = sf.Q1[q] = A(0);
    if (valid && mt >= 0) {
      if (tl.use)
        g2p_backward_sep(
            sp, a, k,
            [&](int i, int j, A& a0, A& a1) {
              const int li = (j - tl.oy) * tl.tw + (i - tl.ox);
              a0 = cw[li];
              a1 = cw[n + li];
            },
            mo, sf);
      else
        g2p_backward_sep(
            sp, a, k,
            [&](int i, int j, A& a0, A& a1) {
              node_velocity(g, g + G, g + 2 * (size_t)G, i, j, k, a0, a1);
            },
            mo, sf);
    }
    if (tl.use) {
      // Every lane runs the same-cell pre-sums (lanes out of range carry zero factors and
      // unique keys); the first lane of each group of 4 adds the two node-velocity
      // adjoints of each stencil node into the tile. A non-finite adjoint can reach the
      // other lanes of its DPP row through f = 0, all of them lanes of this instance,
      // whose gradient it makes non-finite anyway (the tile holds one instance).
      const int lane = threadIdx.x & 63;
      const int key = sf.inr ? sf.b1 * k.nx + sf.b0 + 1 : -(1 + lane);
      const jm::LaneGroups lg = jm::lane_groups(key, sf.inr);
      const int q0 = (sf.b1 - tl.oy) * tl.tw + (sf.b0 - tl.ox);
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) {
          A c0 = sf.wy[jj] * sf.P0[ii] + sf.ey[jj] * sf.Q0[ii];
          A c1 = sf.wy[jj] * sf.P1[ii] + sf.ey[jj] * sf.Q1[ii];
          jm::presum2(c0, c1, lg.f1, lg.f2);
          if (lg.emit) {
            const int li = q0 + jj * tl.tw + ii;
            atomicAdd(&acc[li], double(c0));
            atomicAdd(&acc[n + li], double(c1));
          }
        }
      }
      __syncthreads();
      // through each node's grid op, once per node, into the instance's adjoint grid
      A* ga = agrid_prev + (size_t)tl.inst * 3 * G;
      for (int li = threadIdx.x; li < n; li += kDiffBlock) {
        const double g0 = acc[li], g1 = acc[n + li];
        if (g0 == 0.0 && g1 == 0.0) continue;
        const int jj = li / tl.tw;
        const int node = (tl.oy + jj) * k.nx + tl.ox + (li - jj * tl.tw);
        const A pb0 = A(double(cw[2 * n + li]) * g0), pb1 = A(double(cw[3 * n + li]) * g1);
        const A mb = A(double(cw[4 * n + li]) * g0 + double(cw[5 * n + li]) * g1);
        if (pb0 != A(0)) unsafeAtomicAdd(ga + node, pb0);
        if (pb1 != A(0)) unsafeAtomicAdd(ga + G + node, pb1);
        if (mb != A(0)) unsafeAtomicAdd(ga + 2 * (size_t)G + node, mb);
      }
    } else if (sf.inr) {
      const GlobalEmit<A> direct{agrid_prev + (size_t)m * 3 * G, k.nx, G};
      emit_sep(sf, [&](int i, int j, A g0, A g1) {
        const int node = j * k.nx + i;
        const NodeCoef<A> c = node_coef(A(g[node]), A(g[G + node]), A(g[2 * (size_t)G + node]),
                                        i, j, k);
        A pb0, pb1, mb;
        through_gridop(c, g0, g1, pb0, pb1, mb);
        direct(i, j, pb0, pb1, mb);
      });
    }
    if (valid) {
      A* mo_p = m_out + 8 * (size_t)p;
      mo_p[0] = mo.x0; mo_p[1] = mo.x1; mo_p[2] = mo.v0; mo_p[3] = mo.v1;
      mo_p[4] = mo.C.a; mo_p[5] = mo.C.b; mo_p[6] = mo.C.c; mo_p[7] = mo.C.d;
      A* af = aF_out + 4 * (size_t)p;
      af[0] = a.F.a; af[1] = a.F.b; af[2] = a.F.c; af[3] = a.F.d;
    }
  }
}

}  // namespace jmd
