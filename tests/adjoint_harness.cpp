// TEST HARNESS: the engine's adjoint math (jellyfih_amd/csrc/adjoint.h) compiled for the
// host in fp64, single-threaded, for small particle sets. tests/test_adjoint_cpu.py checks
// its forward against the oracle and its gradients against central differences; the GPU
// tests then check the HIP kernels (same header, fp32) against it.
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

using T = double;

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

struct Add {
  void operator()(T* p, T v) const { *p += v; }
};

void lame(T E, T nu, T& mu, T& la) {
  mu = E / (2 * (1 + nu));
  la = E * nu / ((1 + nu) * (1 - 2 * nu));
}

}  // namespace

extern "C" {

// params (M,5): amp, contraction, freq, payload density, E_jelly. traj (nsteps+1, P, 12).
int adj_forward(const OracleConsts* c, int M, int P, const int32_t* inst, const int32_t* mat,
                const double* fiber, const double* amp_scale, const double* params,
                const double* state_in, double t0, int nsteps, double* traj) {
  jmd::Consts<T> k = make(*c);
  const size_t G = (size_t)k.nx * k.ny;
  std::vector<T> grid((size_t)M * 5 * G);
  std::memcpy(traj, state_in, sizeof(double) * 12 * P);
  T t = t0;
  for (int s = 0; s < nsteps; ++s) {
    const T* cur = traj + (size_t)s * P * 12;
    T* nxt = traj + (size_t)(s + 1) * P * 12;
    std::fill(grid.begin(), grid.end(), 0.0);
    std::vector<jmd::PState<T>> st(P);
    for (int p = 0; p < P; ++p) st[p] = jmd::load_state(cur + 12 * p);
    for (int p = 0; p < P; ++p) {
      if (mat[p] < 0) continue;
      const int m = inst[p];
      const T* pr = params + 5 * m;
      jmd::Act<T> act = jmd::activation(t, pr[2], pr[1], k);
      T mu, la;
      lame(pr[4], 0.3, mu, la);
      T* g = grid.data() + (size_t)m * 5 * G;
      jmd::p2g_scatter(st[p], mat[p], fiber[2 * p], fiber[2 * p + 1],
                       act.a * pr[0] * amp_scale[p], mu, la, pr[3], k, g, g + G, g + 2 * G,
                       Add());
    }
    for (int m = 0; m < M; ++m) {
      T* g = grid.data() + (size_t)m * 5 * G;
      for (int j = 0; j < k.ny; ++j)
        for (int i = 0; i < k.nx; ++i) {
          size_t n = (size_t)j * k.nx + i;
          jmd::gridop_forward(g[n], g[G + n], g[2 * G + n], i, j, k, g[3 * G + n], g[4 * G + n]);
        }
    }
    for (int p = 0; p < P; ++p) {
      if (mat[p] >= 0) {
        T* g = grid.data() + (size_t)inst[p] * 5 * G;
        jmd::g2p_particle(st[p], k, g + 3 * G, g + 4 * G);
      }
      jmd::store_state(nxt + 12 * p, st[p]);
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
  std::vector<T> times(nsteps + 1);
  times[0] = t0;
  for (int s = 0; s < nsteps; ++s) times[s + 1] = times[s] + k.dt;
  std::vector<T> grid((size_t)M * 5 * G), adj((size_t)M * 5 * G);
  std::vector<jmd::PState<T>> bout(P), bin(P);
  for (int p = 0; p < P; ++p) bout[p] = jmd::load_state(grad_out + 12 * p);
  for (int s = nsteps - 1; s >= 0; --s) {
    const T t = times[s];
    const T* cur = traj + (size_t)s * P * 12;
    std::fill(grid.begin(), grid.end(), 0.0);
    std::fill(adj.begin(), adj.end(), 0.0);
    std::vector<jmd::PState<T>> st(P), tmp(P);
    for (int p = 0; p < P; ++p) st[p] = tmp[p] = jmd::load_state(cur + 12 * p);
    // recompute the grid of substep s
    for (int p = 0; p < P; ++p) {
      if (mat[p] < 0) continue;
      const int m = inst[p];
      const T* pr = params + 5 * m;
      jmd::Act<T> act = jmd::activation(t, pr[2], pr[1], k);
      T mu, la;
      lame(pr[4], 0.3, mu, la);
      T* g = grid.data() + (size_t)m * 5 * G;
      jmd::p2g_scatter(tmp[p], mat[p], fiber[2 * p], fiber[2 * p + 1],
                       act.a * pr[0] * amp_scale[p], mu, la, pr[3], k, g, g + G, g + 2 * G,
                       Add());
    }
    for (int m = 0; m < M; ++m) {
      T* g = grid.data() + (size_t)m * 5 * G;
      for (int j = 0; j < k.ny; ++j)
        for (int i = 0; i < k.nx; ++i) {
          size_t n = (size_t)j * k.nx + i;
          jmd::gridop_forward(g[n], g[G + n], g[2 * G + n], i, j, k, g[3 * G + n], g[4 * G + n]);
        }
    }
    // G2P adjoint (scatter into the grid-velocity adjoint)
    for (int p = 0; p < P; ++p) {
      bin[p] = jmd::PState<T>{0, 0, 0, 0, {0, 0, 0, 0}, {0, 0, 0, 0}};
      if (mat[p] < 0) {
        bin[p] = bout[p];
        continue;
      }
      T* g = grid.data() + (size_t)inst[p] * 5 * G;
      T* a = adj.data() + (size_t)inst[p] * 5 * G;
      jmd::g2p_backward(st[p], bout[p], k, g + 3 * G, g + 4 * G, a + 3 * G, a + 4 * G, bin[p],
                        Add());
    }
    // grid-op adjoint
    for (int m = 0; m < M; ++m) {
      T* g = grid.data() + (size_t)m * 5 * G;
      T* a = adj.data() + (size_t)m * 5 * G;
      for (int j = 0; j < k.ny; ++j)
        for (int i = 0; i < k.nx; ++i) {
          size_t n = (size_t)j * k.nx + i;
          jmd::gridop_backward(g[n], g[G + n], g[2 * G + n], i, j, k, a[3 * G + n], a[4 * G + n],
                               a[n], a[G + n], a[2 * G + n]);
        }
    }
    // P2G adjoint (gather)
    for (int p = 0; p < P; ++p) {
      if (mat[p] < 0) continue;
      const int m = inst[p];
      const T* pr = params + 5 * m;
      jmd::Act<T> act = jmd::activation(t, pr[2], pr[1], k);
      T mu, la;
      lame(pr[4], 0.3, mu, la);
      const T* a = adj.data() + (size_t)m * 5 * G;
      T mub = 0, lab = 0, betab = 0;
      jmd::p2g_backward_particle(st[p], mat[p], (T)fiber[2 * p], (T)fiber[2 * p + 1],
                                 act.a * pr[0] * amp_scale[p], mu, la, pr[3], k, a, a + G,
                                 a + 2 * G, bout[p].F, bin[p], mub, lab, betab);
      if (mat[p] == 1 || mat[p] == 3) {
        const T nu = 0.3;
        grad_pp[4 * p] += mub / (2 * (1 + nu)) + lab * nu / ((1 + nu) * (1 - 2 * nu));
      }
      if (mat[p] == 3) {
        grad_amp_scale[p] += betab * act.a * pr[0];
        grad_pp[4 * p + 1] += betab * act.a * amp_scale[p];
        grad_pp[4 * p + 2] += betab * pr[0] * amp_scale[p] * act.da_dc;
        grad_pp[4 * p + 3] += betab * pr[0] * amp_scale[p] * act.da_df;
      }
    }
    bout.swap(bin);
  }
  for (int p = 0; p < P; ++p) jmd::store_state(grad_in + 12 * p, bout[p]);
  return 0;
}

}  // extern "C"
