// TEST HARNESS: the engine's adjoint math (jellyfih_amd/csrc/adjoint.h) compiled for the
// host, single-threaded, for small particle sets; in fp64 by default, in fp32 with
// -DHARNESS_FP32 (arrays stay double, holding the fp32 values exactly; the clock is fp32).
// tests/test_adjoint_cpu.py checks its forward against the oracle and its gradients
// against central differences; the GPU tests then check the HIP kernels (same header)
// against it.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../jellyfih_amd/csrc/adjoint.h"

extern "C" {
struct OracleConsts {  // same layout as oracle/mpm_oracle.cpp
  int32_t nx, ny, track_glow, reserved;
  double dt, dx, inv_dx, p_vol, p_mass, gravity, mu_base, la_base, mu_payload, la_payload,
      water_la, act_freq, domain_height;
};
}

namespace {

#ifdef HARNESS_FP32
using T = float;
#else
using T = double;
#endif

jmd::PState<T> load(const double* s) {
  return {T(s[0]), T(s[1]), T(s[2]), T(s[3]), {T(s[4]), T(s[5]), T(s[6]), T(s[7])},
          {T(s[8]), T(s[9]), T(s[10]), T(s[11])}};
}
void store(double* s, const jmd::PState<T>& p) {
  s[0] = p.x0; s[1] = p.x1; s[2] = p.v0; s[3] = p.v1;
  s[4] = p.C.a; s[5] = p.C.b; s[6] = p.C.c; s[7] = p.C.d;
  s[8] = p.F.a; s[9] = p.F.b; s[10] = p.F.c; s[11] = p.F.d;
}

jmd::Consts<T> make(const OracleConsts& c) {
  jmd::Consts<T> k;
  k.nx = c.nx;
  k.ny = c.ny;
  k.damp_x = c.nx / 20;
  k.damp_y = c.ny / 20;
  k.dt = c.dt;
  k.dx = c.dx;
  k.inv_dx = c.inv_dx;
  k.p_mass = c.p_mass;
  k.gravity_dt = c.dt * c.gravity;
  k.stress_scale = -c.dt * c.p_vol * 4.0 * c.inv_dx * c.inv_dx;
  k.apic_scale = 4.0 * c.inv_dx;
  k.mu_base = c.mu_base;
  k.la_base = c.la_base;
  k.mu_pay = c.mu_payload;
  k.la_pay = c.la_payload;
  k.water_la = c.water_la;
  k.act_freq = c.act_freq;
  k.x_lo = 0.001;
  k.x_hi = 0.999;
  k.y_hi = c.domain_height - 0.001;
  k.sig_floor = 1e-6;
  k.pay_lo = 1.0 - 5e-3;
  k.pay_hi = 1.0 + 5e-3;
  k.pi = 3.14159265;
  k.nu_jelly = 0.3;
  return k;
}

// Plain accumulation into the three planes (momentum x, momentum y, mass) of one grid.
struct Planes {
  T* g;
  size_t G;
  int nx;
  void operator()(int i, int j, T c0, T c1, T c2) const {
    const size_t n = (size_t)j * nx + i;
    g[n] += c0;
    g[G + n] += c1;
    g[2 * G + n] += c2;
  }
};

void lame(T E, T nu, T& mu, T& la) {
  mu = E / (2 * (1 + nu));
  la = E * nu / ((1 + nu) * (1 - 2 * nu));
}

std::vector<T> tv(const double* a, size_t n) { return std::vector<T>(a, a + n); }

// Momentum and mass of every instance's grid from the states `st` at time t.
void p2g_all(const jmd::Consts<T>& k, int M, int P, const int32_t* inst, const int32_t* mat,
             const T* fiber, const T* amp_scale, const T* params, T t,
             std::vector<jmd::PState<T>>& st, std::vector<T>& grid, size_t G) {
  std::fill(grid.begin(), grid.end(), 0.0);
  for (int p = 0; p < P; ++p) {
    if (mat[p] < 0) continue;
    const int m = inst[p];
    const T* pr = params + 5 * m;
    jmd::Act<T> act = jmd::activation(t, pr[2], pr[1], k);
    T mu, la;
    lame(pr[4], T(0.3), mu, la);
    jmd::p2g_particle(st[p], mat[p], fiber[2 * p], fiber[2 * p + 1],
                      act.a * pr[0] * amp_scale[p], mu, la, pr[3], k,
                      Planes{grid.data() + (size_t)m * 3 * G, G, k.nx});
  }
}

}  // namespace

extern "C" {

// params (M,5): amp, contraction, freq, payload density, E_jelly. traj (nsteps+1, P, 12).
int adj_forward(const OracleConsts* c, int M, int P, const int32_t* inst, const int32_t* mat,
                const double* fiber, const double* amp_scale, const double* params,
                const double* state_in, double t0, int nsteps, double* traj) {
  jmd::Consts<T> k = make(*c);
  const size_t G = (size_t)k.nx * k.ny;
  std::vector<T> grid((size_t)M * 3 * G);
  const std::vector<T> fib = tv(fiber, 2 * (size_t)P), amps = tv(amp_scale, P),
                       par = tv(params, 5 * (size_t)M);
  for (size_t i = 0; i < 12 * (size_t)P; ++i) traj[i] = T(state_in[i]);
  T t = T(t0);
  for (int s = 0; s < nsteps; ++s) {
    const double* cur = traj + (size_t)s * P * 12;
    double* nxt = traj + (size_t)(s + 1) * P * 12;
    std::vector<jmd::PState<T>> st(P);
    for (int p = 0; p < P; ++p) st[p] = load(cur + 12 * p);
    p2g_all(k, M, P, inst, mat, fib.data(), amps.data(), par.data(), t, st, grid, G);
    for (int p = 0; p < P; ++p) {
      if (mat[p] >= 0) {
        T* g = grid.data() + (size_t)inst[p] * 3 * G;
        jmd::g2p_particle(st[p], k, g, g + G, g + 2 * G);
      }
      store(nxt + 12 * p, st[p]);
    }
    t = t + k.dt;
  }
  return 0;
}

// Reverse pass over traj (nsteps+1 states). grad_out (P,12) is the adjoint of the last
// state; grad_in (P,12) receives the adjoint of the first. grad_amp_scale (P) and
// grad_pp (P,4: E, amp, contraction, freq per particle, to be summed per instance)
// accumulate.
int adj_backward(const OracleConsts* c, int M, int P, const int32_t* inst, const int32_t* mat,
                 const double* fiber, const double* amp_scale, const double* params,
                 const double* traj, double t0, int nsteps, const double* grad_out,
                 double* grad_in, double* grad_amp_scale, double* grad_pp) {
  jmd::Consts<T> k = make(*c);
  const size_t G = (size_t)k.nx * k.ny;
  const std::vector<T> fib = tv(fiber, 2 * (size_t)P), amps = tv(amp_scale, P),
                       par = tv(params, 5 * (size_t)M);
  std::vector<T> times(nsteps + 1);
  times[0] = T(t0);
  for (int s = 0; s < nsteps; ++s) times[s + 1] = times[s] + k.dt;
  std::vector<T> grid((size_t)M * 3 * G), adj((size_t)M * 3 * G), vadj((size_t)M * 2 * G);
  std::vector<jmd::PState<T>> bout(P), bin(P);
  for (int p = 0; p < P; ++p) bout[p] = load(grad_out + 12 * p);
  for (int s = nsteps - 1; s >= 0; --s) {
    const T t = times[s];
    const double* cur = traj + (size_t)s * P * 12;
    std::fill(adj.begin(), adj.end(), 0.0);
    std::vector<jmd::PState<T>> st(P), tmp(P);
    for (int p = 0; p < P; ++p) st[p] = tmp[p] = load(cur + 12 * p);
    p2g_all(k, M, P, inst, mat, fib.data(), amps.data(), par.data(), t, tmp, grid, G);
    // G2P adjoint: node-velocity adjoints summed per node (two planes per instance), then
    // through each node's grid op into the momentum/mass adjoints, as the GPU kernel does
    std::fill(vadj.begin(), vadj.end(), 0.0);
    for (int p = 0; p < P; ++p) {
      bin[p] = jmd::PState<T>{0, 0, 0, 0, {0, 0, 0, 0}, {0, 0, 0, 0}};
      if (mat[p] < 0) {
        bin[p] = bout[p];
        continue;
      }
      const T* g = grid.data() + (size_t)inst[p] * 3 * G;
      T* va = vadj.data() + (size_t)inst[p] * 2 * G;
      jmd::SepFactors<T> sf;
      jmd::g2p_backward_sep(
          st[p], bout[p], k,
          [&](int i, int j, T& a0, T& a1) {
            jmd::node_velocity(g, g + G, g + 2 * G, i, j, k, a0, a1);
          },
          bin[p], sf);
      jmd::emit_sep(sf, [&](int i, int j, T g0, T g1) {
        const size_t n = (size_t)j * k.nx + i;
        va[n] += g0;
        va[G + n] += g1;
      });
    }
    for (int m = 0; m < M; ++m)
      for (int j = 0; j < k.ny; ++j)
        for (int i = 0; i < k.nx; ++i) {
          const size_t n = (size_t)j * k.nx + i;
          const T* g = grid.data() + (size_t)m * 3 * G;
          const T* va = vadj.data() + (size_t)m * 2 * G;
          T* a = adj.data() + (size_t)m * 3 * G;
          const jmd::NodeCoef<T> c = jmd::node_coef(g[n], g[G + n], g[2 * G + n], i, j, k);
          jmd::through_gridop(c, va[n], va[G + n], a[n], a[G + n], a[2 * G + n]);
        }
    // P2G adjoint (gather)
    for (int p = 0; p < P; ++p) {
      if (mat[p] < 0) continue;
      const int m = inst[p];
      const T* pr = par.data() + 5 * m;
      jmd::Act<T> act = jmd::activation(t, pr[2], pr[1], k);
      T mu, la;
      lame(pr[4], T(0.3), mu, la);
      const T* a = adj.data() + (size_t)m * 3 * G;
      T mub = 0, lab = 0, betab = 0;
      jmd::p2g_backward_particle(st[p], mat[p], fib[2 * p], fib[2 * p + 1],
                                 act.a * pr[0] * amps[p], mu, la, pr[3], k, a, a + G,
                                 a + 2 * G, bout[p].F, bin[p], mub, lab, betab);
      if (mat[p] == 1 || mat[p] == 3) {
        const T nu = T(0.3);
        grad_pp[4 * p] += mub / (2 * (1 + nu)) + lab * nu / ((1 + nu) * (1 - 2 * nu));
      }
      if (mat[p] == 3) {
        grad_amp_scale[p] += betab * act.a * pr[0];
        grad_pp[4 * p + 1] += betab * act.a * amps[p];
        grad_pp[4 * p + 2] += betab * pr[0] * amps[p] * act.da_dc;
        grad_pp[4 * p + 3] += betab * pr[0] * amps[p] * act.da_df;
      }
    }
    bout.swap(bin);
  }
  for (int p = 0; p < P; ++p) store(grad_in + 12 * p, bout[p]);
  return 0;
}

}  // extern "C"
