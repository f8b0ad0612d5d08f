// TEST INFRASTRUCTURE — the oracle. Not part of the product path.
//
// CPU restatement of the reference's batched 2D MLS-MPM simulation core, templated on
// float (fp32, the reference's arithmetic type) and double (fp64, used to attribute fp32
// error). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the reference results / CPU baseline, never as the thing
// measured.
//
// Every function follows the reference line by line (citations are file:line in
// c-m-x-r/jellyfih):
//   constants            mpm_sim.py:18-58
//   substep()            mpm_sim.py:153-340, in exactly this order:
//     grid reset         mpm_sim.py:159-165
//     P2G                mpm_sim.py:167-271  (actuation :171-181, range skip :183-184,
//                        F update :189, materials :191-203, SVD+plasticity :205-223,
//                        stress :225-233, scaling :250, APIC :254, scatter :256-271)
//     grid op            mpm_sim.py:273-309
//     G2P                mpm_sim.py:311-333
//     glow (optional)    mpm_sim.py:335-338
//     time advance       mpm_sim.py:340
//   axisymmetric variant mpm_sim.py:127-133 (fields), :163-165 (reset), :236-247 (sigma_rr,
//                        hoop stress), :258-267 (ring-weighted scatter), :292-304 (hoop
//                        correction, axis BC); r_ref = spawn x (:727)
//   payload/body stats   mpm_sim.py:740-762
//   2x2 SVD              taichi 1.7.4 python/taichi/_funcs.py `_polar_decompose2d` and
//                        `_svd2d` (third-party, not vendored in the reference; the
//                        reference calls it at mpm_sim.py:206). Restated from its published
//                        algorithm: polar rotation from normalize(F00+F11, F10-F01), then a
//                        one-sided Jacobi rotation of the symmetric factor.
//
// Arithmetic notes (hazard list of the reference):
//   * `.cast(int)` truncates toward zero (mpm_sim.py:183,315); NaN casts to 0.
//   * Taichi float `%` is floored: a - b*floor(a/b) (mpm_sim.py:174).
//   * Constants that the reference folds in Python f64 (e.g. -dt*p_vol*4*inv_dx*inv_dx at
//     mpm_sim.py:250, dt*gravity at :329, 4*inv_dx at :326) are folded here in double and
//     then rounded to T once.
//   * ti.max / ti.min follow IEEE maxNum/minNum (the non-NaN operand wins): fmax / fmin.
//   * Build with -ffp-contract=off: no fused multiply-add, every operation rounds once.
//
// Parallelism: instances are independent, so the multi-threaded entry points run one
// instance per OpenMP thread; each instance is processed sequentially in slot order, so
// results are bit-identical for any thread count.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// Physical constants of the reference (mpm_sim.py:11-58), always given in double;
// they are rounded to the arithmetic type T where the reference rounds them.
struct OracleConsts {
  int32_t nx;             // n_grid (x cells), mpm_sim.py:14
  int32_t ny;             // n_grid_y (y cells), mpm_sim.py:15
  int32_t track_glow;     // run the render-only glow pass (mpm_sim.py:335-338)
  int32_t axisym;  // JELLY_AXISYM (mpm_sim.py:17)
  double dt;              // 2e-5 / quality, mpm_sim.py:19
  double dx;              // 1 / n_grid, mpm_sim.py:18
  double inv_dx;          // float(n_grid), mpm_sim.py:18
  double p_vol;           // (dx*0.5)^2, mpm_sim.py:20
  double p_mass;          // p_vol * p_rho, mpm_sim.py:22
  double gravity;         // 10.0, mpm_sim.py:49
  double mu_base;         // mpm_sim.py:27
  double la_base;         // mpm_sim.py:28
  double mu_payload;      // mpm_sim.py:45
  double la_payload;      // mpm_sim.py:46
  double water_la;        // 100000.0, mpm_sim.py:48
  double act_freq;        // actuation_freq (1 Hz), mpm_sim.py:52
  double domain_height;   // mpm_sim.py:16
};

}  // extern "C"

namespace {

template <typename T>
struct M2 {
  T a00, a01, a10, a11;
};

template <typename T>
inline M2<T> matmul(const M2<T>& a, const M2<T>& b) {
  // Taichi's 2x2 `@`: sum over k in order k = 0, 1.
  return {a.a00 * b.a00 + a.a01 * b.a10, a.a00 * b.a01 + a.a01 * b.a11,
          a.a10 * b.a00 + a.a11 * b.a10, a.a10 * b.a01 + a.a11 * b.a11};
}

template <typename T>
inline M2<T> transpose(const M2<T>& a) {
  return {a.a00, a.a10, a.a01, a.a11};
}

// Truncation toward zero with the GPU's saturating conversion (NaN -> 0).
template <typename T>
inline int32_t cast_int(T v) {
  if (!(v == v)) return 0;
  if (v >= T(2147483647.0)) return INT32_MAX;
  if (v <= T(-2147483648.0)) return INT32_MIN;
  return static_cast<int32_t>(v);
}

template <typename T>
struct Consts {
  int nx, ny, damp_x, damp_y;
  bool glow;
  T dt, dx, inv_dx, p_mass, gravity_dt, stress_scale, apic_scale;
  T mu_base, la_base, mu_pay, la_pay, water_la, act_freq;
  T x_lo, x_hi, y_hi;
  T sig_floor, pay_lo, pay_hi;
  T pi;
  bool axisym;
  T r_ref_floor, r_floor;

  explicit Consts(const OracleConsts& c) {
    nx = c.nx;
    ny = c.ny;
    damp_x = c.nx / 20;  // mpm_sim.py:283 (Python floor division)
    damp_y = c.ny / 20;  // mpm_sim.py:284
    glow = c.track_glow != 0;
    dt = T(c.dt);
    dx = T(c.dx);
    inv_dx = T(c.inv_dx);
    p_mass = T(c.p_mass);
    gravity_dt = T(c.dt * c.gravity);                                // mpm_sim.py:329
    stress_scale = T(-c.dt * c.p_vol * 4.0 * c.inv_dx * c.inv_dx);  // mpm_sim.py:250
    apic_scale = T(4.0 * c.inv_dx);                                  // mpm_sim.py:326
    mu_base = T(c.mu_base);
    la_base = T(c.la_base);
    mu_pay = T(c.mu_payload);
    la_pay = T(c.la_payload);
    water_la = T(c.water_la);
    act_freq = T(c.act_freq);
    x_lo = T(0.001);
    x_hi = T(0.999);
    y_hi = T(c.domain_height - 0.001);  // mpm_sim.py:333
    sig_floor = T(1e-6);                // mpm_sim.py:208
    pay_lo = T(1.0 - 5e-3);             // mpm_sim.py:215
    pay_hi = T(1.0 + 5e-3);
    pi = T(3.14159265);  // mpm_sim.py:180
    axisym = c.axisym != 0;
    r_ref_floor = T(c.dx * 0.1);  // mpm_sim.py:246 (Python f64 product, then fp)
    r_floor = T(c.dx * 0.05);     // mpm_sim.py:247
  }
};

// Raised-cosine activation of one instance at time t (mpm_sim.py:173-181).
template <typename T>
inline T activation(const Consts<T>& k, T t, T freq, T contraction) {
  T period = T(1) / (k.act_freq * freq);
  T q = std::floor(t / period);
  T phase = (t - period * q) / period;
  T a = T(0);
  if (phase < contraction) a = T(0.5) * (T(1) - std::cos(phase / contraction * k.pi));
  return a;
}

// taichi 1.7.4 `_polar_decompose2d` + `_svd2d` (see file header).
template <typename T>
inline void svd2d(const M2<T>& A, M2<T>& U, T& s1, T& s2, M2<T>& V, M2<T>* R_out) {
  T x = A.a00 + A.a11, y = A.a10 - A.a01;
  T scale = T(1) / std::sqrt(x * x + y * y);
  T c = x * scale, s = y * scale;
  M2<T> R = {c, -s, s, c};
  M2<T> S = matmul(transpose(R), A);
  T cj, sj;
  if (std::fabs(S.a01) < T(1e-5)) {
    cj = T(1);
    sj = T(0);
    s1 = S.a00;
    s2 = S.a11;
  } else {
    T tao = T(0.5) * (S.a00 - S.a11);
    T w = std::sqrt(tao * tao + S.a01 * S.a01);
    T tt;
    if (tao > T(0))
      tt = S.a01 / (tao + w);
    else
      tt = S.a01 / (tao - w);
    cj = T(1) / std::sqrt(tt * tt + T(1));
    sj = -tt * cj;
    s1 = cj * cj * S.a00 - T(2) * cj * sj * S.a01 + sj * sj * S.a11;
    s2 = sj * sj * S.a00 + T(2) * cj * sj * S.a01 + cj * cj * S.a11;
  }
  if (s1 < s2) {
    std::swap(s1, s2);
    V = {-sj, cj, -cj, -sj};
  } else {
    V = {cj, sj, -sj, cj};
  }
  U = matmul(R, V);
  if (R_out) *R_out = R;
}

template <typename T>
inline void bspline_weights(T fx, T w[3]) {
  // mpm_sim.py:187 (`** 2` is a product)
  T a = T(1.5) - fx, b = fx - T(1), c = fx - T(0.5);
  w[0] = T(0.5) * (a * a);
  w[1] = T(0.75) - b * b;
  w[2] = T(0.5) * (c * c);
}

template <typename T>
struct Fields {
  int M, N;
  T* x;           // (M,N,2)
  T* v;           // (M,N,2)
  T* C;           // (M,N,2,2)
  T* F;           // (M,N,2,2)
  const int32_t* mat;  // (M,N)
  T* Jp;          // (M,N)
  const T* fiber; // (M,N,2)
  T* glow;        // (M,N) or null
  T* grid_v;      // (M,nx,ny,2)
  T* grid_m;      // (M,nx,ny)
  const T* amp;          // (M,) instance_actuation
  const T* contraction;  // (M,) instance_act_contraction
  const T* freq;         // (M,) instance_freq
  const T* pdens;        // (M,) instance_payload_density
  const T* mu_j;         // (M,) mu_jelly (per instance; E_jelly promoted to runtime)
  const T* la_j;         // (M,) lambda_jelly
  const T* r_ref;        // (M,N) spawn radius (axisymmetric variant only, else null)
  T* g_srr;              // (M,nx,ny) mass-weighted Kirchhoff sigma_rr (axisymmetric only)
  T* g_shoop;            // (M,nx,ny) mass-weighted Kirchhoff sigma_theta_theta
};

enum Stage : int {
  ST_RESET = 1,
  ST_P2G = 2,
  ST_GRIDOP = 4,
  ST_G2P = 8,
  ST_GLOW = 16,
};

template <typename T>
void p2g_instance(const Consts<T>& k, const Fields<T>& f, int m, T t) {
  const int N = f.N, nx = k.nx, ny = k.ny;
  T* gv = f.grid_v + (size_t)m * nx * ny * 2;
  T* gm = f.grid_m + (size_t)m * nx * ny;
  T* gsr = k.axisym ? f.g_srr + (size_t)m * nx * ny : nullptr;
  T* gsh = k.axisym ? f.g_shoop + (size_t)m * nx * ny : nullptr;
  const T act = activation(k, t, f.freq[m], f.contraction[m]);
  for (int p = 0; p < N; ++p) {
    const size_t ip = (size_t)m * N + p;
    const int32_t mat = f.mat[ip];
    if (mat < 0) continue;  // mpm_sim.py:169
    T x0 = f.x[2 * ip], x1 = f.x[2 * ip + 1];
    int32_t b0 = cast_int(x0 * k.inv_dx - T(0.5));
    int32_t b1 = cast_int(x1 * k.inv_dx - T(0.5));
    if (b0 < 0 || b1 < 0 || b0 >= nx - 2 || b1 >= ny - 2) continue;  // mpm_sim.py:184
    T fx0 = x0 * k.inv_dx - T(b0), fx1 = x1 * k.inv_dx - T(b1);
    T wx[3], wy[3];
    bspline_weights(fx0, wx);
    bspline_weights(fx1, wy);

    M2<T> C = {f.C[4 * ip], f.C[4 * ip + 1], f.C[4 * ip + 2], f.C[4 * ip + 3]};
    M2<T> F = {f.F[4 * ip], f.F[4 * ip + 1], f.F[4 * ip + 2], f.F[4 * ip + 3]};
    // mpm_sim.py:189  F <- (I + dt*C) @ F
    M2<T> A = {T(1) + k.dt * C.a00, T(0) + k.dt * C.a01, T(0) + k.dt * C.a10,
               T(1) + k.dt * C.a11};
    F = matmul(A, F);

    // mpm_sim.py:192-203
    T mu = k.mu_base, la = k.la_base, mass = k.p_mass;
    if (mat == 1 || mat == 3) {
      mu = f.mu_j[m];
      la = f.la_j[m];
    } else if (mat == 2) {
      mu = k.mu_pay;
      la = k.la_pay;
      mass = mass * f.pdens[m];
    } else if (mat == 0) {
      mu = T(0);
      la = k.water_la;
    }

    // mpm_sim.py:206-218
    M2<T> U, V;
    T s[2];
    svd2d(F, U, s[0], s[1], V, (M2<T>*)nullptr);
    for (int d = 0; d < 2; ++d) s[d] = std::fmax(s[d], k.sig_floor);
    T J = T(1);
    T jp = f.Jp[ip];
    for (int d = 0; d < 2; ++d) {
      T ns = s[d];
      if (mat == 2) ns = std::fmin(std::fmax(s[d], k.pay_lo), k.pay_hi);
      jp = jp * (s[d] / ns);
      s[d] = ns;
      J = J * ns;
    }
    f.Jp[ip] = jp;

    // mpm_sim.py:220-223
    if (mat == 0) {
      T r = std::sqrt(J);
      F = {T(1) * r, T(0) * r, T(0) * r, T(1) * r};
    } else if (mat == 2) {
      M2<T> Sg = {s[0], T(0), T(0), s[1]};
      F = matmul(matmul(U, Sg), transpose(V));
    }

    // mpm_sim.py:226-227  stress = 2*mu*(F - U V^T) F^T + I*la*J*(J-1)
    M2<T> R = matmul(U, transpose(V));
    T mu2 = T(2) * mu;
    M2<T> D = {mu2 * (F.a00 - R.a00), mu2 * (F.a01 - R.a01), mu2 * (F.a10 - R.a10),
               mu2 * (F.a11 - R.a11)};
    M2<T> st = matmul(D, transpose(F));
    T dg = ((T(1) * la) * J) * (J - T(1));
    T og = ((T(0) * la) * J) * (J - T(1));
    st = {st.a00 + dg, st.a01 + og, st.a10 + og, st.a11 + dg};

    // mpm_sim.py:230-233 active muscle stress along the fibre
    if (mat == 3) {
      T cp = f.amp[m] * act;
      T d0 = f.fiber[2 * ip], d1 = f.fiber[2 * ip + 1];
      st.a00 = st.a00 + ((d0 * d0) * cp) * J;
      st.a01 = st.a01 + ((d0 * d1) * cp) * J;
      st.a10 = st.a10 + ((d1 * d0) * cp) * J;
      st.a11 = st.a11 + ((d1 * d1) * cp) * J;
    }

    // mpm_sim.py:236-247 axisymmetric variant: Kirchhoff sigma_rr and the hoop stress of
    // the circumferential stretch r / r_ref, taken before the stress is scaled
    T krr = T(0), khoop = T(0), rp = T(0);
    if (k.axisym) {
      krr = st.a00;
      rp = x0;
      const T rr = std::fmax(f.r_ref[ip], k.r_ref_floor);
      const T lam = std::fmax(rp, k.r_floor) / rr;
      khoop = ((T(2) * mu) * (lam - T(1))) + ((la * J) * (J - T(1)));
    }

    // mpm_sim.py:250,254
    st = {k.stress_scale * st.a00, k.stress_scale * st.a01, k.stress_scale * st.a10,
          k.stress_scale * st.a11};
    M2<T> af = {st.a00 + mass * C.a00, st.a01 + mass * C.a01, st.a10 + mass * C.a10,
                st.a11 + mass * C.a11};
    f.F[4 * ip] = F.a00;
    f.F[4 * ip + 1] = F.a01;
    f.F[4 * ip + 2] = F.a10;
    f.F[4 * ip + 3] = F.a11;

    // mpm_sim.py:256-271 scatter
    T v0 = f.v[2 * ip], v1 = f.v[2 * ip + 1];
    T mv0 = mass * v0, mv1 = mass * v1;
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) {
        T d0 = (T(i) - fx0) * k.dx, d1 = (T(j) - fx1) * k.dx;
        T wgt = wx[i] * wy[j];
        size_t node = (size_t)(b0 + i) * ny + (b1 + j);
        if (k.axisym) {  // mpm_sim.py:260-267: ring mass m r, stresses weighted by m
          T mr = mass * rp;
          gv[2 * node] += wgt * ((mr * v0) + (af.a00 * d0 + af.a01 * d1));
          gv[2 * node + 1] += wgt * ((mr * v1) + (af.a10 * d0 + af.a11 * d1));
          gm[node] += (wgt * mass) * rp;
          gsr[node] += (wgt * mass) * krr;
          gsh[node] += (wgt * mass) * khoop;
          continue;
        }
        T c0 = wgt * (mv0 + (af.a00 * d0 + af.a01 * d1));
        T c1 = wgt * (mv1 + (af.a10 * d0 + af.a11 * d1));
        gv[2 * node] += c0;
        gv[2 * node + 1] += c1;
        gm[node] += wgt * mass;
      }
    }
  }
}

template <typename T>
void gridop_instance(const Consts<T>& k, const Fields<T>& f, int m) {
  const int nx = k.nx, ny = k.ny;
  T* gv = f.grid_v + (size_t)m * nx * ny * 2;
  T* gm = f.grid_m + (size_t)m * nx * ny;
  const T* gsr = k.axisym ? f.g_srr + (size_t)m * nx * ny : nullptr;
  const T* gsh = k.axisym ? f.g_shoop + (size_t)m * nx * ny : nullptr;
  for (int i = 0; i < nx; ++i) {
    for (int j = 0; j < ny; ++j) {
      size_t node = (size_t)i * ny + j;
      T mass = gm[node];
      if (!(mass > T(0))) continue;  // mpm_sim.py:275
      T g0 = gv[2 * node] / mass, g1 = gv[2 * node + 1] / mass;
      // mpm_sim.py:283-290
      T damp = T(1);
      if (i < k.damp_x) damp = damp * (T(0.95) + T(0.05) * T(i) / T(k.damp_x));
      if (i > nx - k.damp_x) damp = damp * (T(0.95) + T(0.05) * T(nx - i) / T(k.damp_x));
      if (j < k.damp_y) damp = damp * (T(0.95) + T(0.05) * T(j) / T(k.damp_y));
      if (j > ny - k.damp_y) damp = damp * (T(0.95) + T(0.05) * T(ny - j) / T(k.damp_y));
      g0 = g0 * damp;
      g1 = g1 * damp;
      if (k.axisym) {
        // mpm_sim.py:295-299 hoop correction (sigma_rr - sigma_tt) / r, off the axis cell
        const T r_node = (T(i) + T(0.5)) * k.dx;
        if (r_node > k.dx) {
          const T hoop = (gsr[node] - gsh[node]) / mass;
          g0 = g0 + (k.dt * hoop) / r_node;
        }
        if (i == 0) g0 = T(0);  // mpm_sim.py:302-304 axis: no radial velocity
      } else if (i < 3 && g0 < T(0)) {
        g0 = T(0);  // mpm_sim.py:306
      }
      // mpm_sim.py:307-309
      if (i > nx - 3 && g0 > T(0)) g0 = T(0);
      if (j < 3 && g1 < T(0)) g1 = T(0);
      if (j > ny - 3 && g1 > T(0)) g1 = T(0);
      gv[2 * node] = g0;
      gv[2 * node + 1] = g1;
    }
  }
}

template <typename T>
void g2p_instance(const Consts<T>& k, const Fields<T>& f, int m) {
  const int N = f.N, nx = k.nx, ny = k.ny;
  const T* gv = f.grid_v + (size_t)m * nx * ny * 2;
  for (int p = 0; p < N; ++p) {
    const size_t ip = (size_t)m * N + p;
    if (f.mat[ip] < 0) continue;  // mpm_sim.py:313
    T x0 = f.x[2 * ip], x1 = f.x[2 * ip + 1];
    int32_t b0 = cast_int(x0 * k.inv_dx - T(0.5));
    int32_t b1 = cast_int(x1 * k.inv_dx - T(0.5));
    T v0 = f.v[2 * ip], v1 = f.v[2 * ip + 1];
    if (b0 >= 0 && b1 >= 0 && b0 < nx - 2 && b1 < ny - 2) {  // mpm_sim.py:316
      T fx0 = x0 * k.inv_dx - T(b0), fx1 = x1 * k.inv_dx - T(b1);
      T wx[3], wy[3];
      bspline_weights(fx0, wx);
      bspline_weights(fx1, wy);
      T nv0 = T(0), nv1 = T(0);
      M2<T> nc = {T(0), T(0), T(0), T(0)};
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
          T d0 = T(i) - fx0, d1 = T(j) - fx1;
          size_t node = (size_t)(b0 + i) * ny + (b1 + j);
          T g0 = gv[2 * node], g1 = gv[2 * node + 1];
          T wgt = wx[i] * wy[j];
          nv0 = nv0 + wgt * g0;
          nv1 = nv1 + wgt * g1;
          T ws = k.apic_scale * wgt;
          nc.a00 = nc.a00 + ws * (g0 * d0);
          nc.a01 = nc.a01 + ws * (g0 * d1);
          nc.a10 = nc.a10 + ws * (g1 * d0);
          nc.a11 = nc.a11 + ws * (g1 * d1);
        }
      }
      v0 = nv0;
      v1 = nv1;
      f.C[4 * ip] = nc.a00;
      f.C[4 * ip + 1] = nc.a01;
      f.C[4 * ip + 2] = nc.a10;
      f.C[4 * ip + 3] = nc.a11;
    }
    // mpm_sim.py:329-333
    v1 = v1 - k.gravity_dt;
    x0 = x0 + k.dt * v0;
    x1 = x1 + k.dt * v1;
    x0 = std::fmax(std::fmin(x0, k.x_hi), k.x_lo);
    x1 = std::fmax(std::fmin(x1, k.y_hi), k.x_lo);
    f.v[2 * ip] = v0;
    f.v[2 * ip + 1] = v1;
    f.x[2 * ip] = x0;
    f.x[2 * ip + 1] = x1;
  }
}

template <typename T>
void glow_instance(const Consts<T>& k, const Fields<T>& f, int m) {
  (void)k;
  for (int p = 0; p < f.N; ++p) {
    const size_t ip = (size_t)m * f.N + p;
    f.glow[ip] = std::fmax(f.glow[ip] * T(0.9), std::fabs(T(1) - f.Jp[ip]));  // :338
  }
}

template <typename T>
void run_instance(const Consts<T>& k, const Fields<T>& f, int m, T t0, int nsteps, int stages) {
  const size_t G = (size_t)k.nx * k.ny;
  T t = t0;
  for (int s = 0; s < nsteps; ++s) {
    if (stages & ST_RESET) {
      std::fill(f.grid_v + (size_t)m * G * 2, f.grid_v + (size_t)(m + 1) * G * 2, T(0));
      std::fill(f.grid_m + (size_t)m * G, f.grid_m + (size_t)(m + 1) * G, T(0));
      if (k.axisym) {  // mpm_sim.py:163-165
        std::fill(f.g_srr + (size_t)m * G, f.g_srr + (size_t)(m + 1) * G, T(0));
        std::fill(f.g_shoop + (size_t)m * G, f.g_shoop + (size_t)(m + 1) * G, T(0));
      }
    }
    if (stages & ST_P2G) p2g_instance(k, f, m, t);
    if (stages & ST_GRIDOP) gridop_instance(k, f, m);
    if (stages & ST_G2P) g2p_instance(k, f, m);
    if ((stages & ST_GLOW) && f.glow) glow_instance(k, f, m);
    t = t + k.dt;  // mpm_sim.py:340
  }
}

template <typename T>
int run(const OracleConsts* c, int stages, int M, int N, T* x, T* v, T* C, T* F,
        const int32_t* mat, T* Jp, const T* fiber, T* glow, T* grid_v, T* grid_m, T* t,
        const T* amp, const T* contraction, const T* freq, const T* pdens, const T* mu_j,
        const T* la_j, const T* r_ref, T* g_srr, T* g_shoop, int nsteps, int nthreads) {
  if (!c || M <= 0 || N < 0 || nsteps < 0 || c->nx < 3 || c->ny < 3) return 1;
  if (c->axisym && (!r_ref || !g_srr || !g_shoop)) return 2;
  Consts<T> k(*c);
  Fields<T> f{M,   N,       x,    v,     C,     F,     mat, Jp,  fiber, glow,
              grid_v, grid_m, amp, contraction, freq, pdens, mu_j, la_j, r_ref, g_srr, g_shoop};
  const T t0 = *t;
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
  for (int m = 0; m < M; ++m) run_instance(k, f, m, t0, nsteps, stages);
  T tt = t0;
  for (int s = 0; s < nsteps; ++s) tt = tt + k.dt;
  *t = tt;
  (void)nthreads;
  return 0;
}

template <typename T>
void stats(int M, int N, const T* x, const int32_t* mat, int body_mode, T* out) {
  for (int m = 0; m < M; ++m) {
    T sy = T(0), cnt = T(0), sx = T(0);
    for (int p = 0; p < N; ++p) {
      size_t ip = (size_t)m * N + p;
      int32_t mt = mat[ip];
      bool take = body_mode ? (mt == 1 || mt == 3) : (mt == 2);
      if (!take) continue;
      sy = sy + x[2 * ip + 1];
      cnt = cnt + T(1);
      sx = sx + x[2 * ip];
    }
    out[3 * m] = sy;
    out[3 * m + 1] = cnt;
    out[3 * m + 2] = sx;
  }
}

}  // namespace

extern "C" {

// Run `nsteps` substeps (or the stages selected by `stages`) for all M instances.
// Layout is the reference's: x,v,fiber (M,N,2); C,F (M,N,2,2) row-major; mat,Jp,glow (M,N);
// grid_v (M,nx,ny,2); grid_m (M,nx,ny). `t` is the fp time accumulator (in/out).
int oracle_run_f32(const OracleConsts* c, int stages, int M, int N, float* x, float* v,
                   float* C, float* F, const int32_t* mat, float* Jp, const float* fiber,
                   float* glow, float* grid_v, float* grid_m, float* t, const float* amp,
                   const float* contraction, const float* freq, const float* pdens,
                   const float* mu_j, const float* la_j, int nsteps, int nthreads) {
  return run<float>(c, stages, M, N, x, v, C, F, mat, Jp, fiber, glow, grid_v, grid_m, t, amp,
                    contraction, freq, pdens, mu_j, la_j, nullptr, nullptr, nullptr, nsteps,
                    nthreads);
}

int oracle_run_f64(const OracleConsts* c, int stages, int M, int N, double* x, double* v,
                   double* C, double* F, const int32_t* mat, double* Jp, const double* fiber,
                   double* glow, double* grid_v, double* grid_m, double* t, const double* amp,
                   const double* contraction, const double* freq, const double* pdens,
                   const double* mu_j, const double* la_j, int nsteps, int nthreads) {
  return run<double>(c, stages, M, N, x, v, C, F, mat, Jp, fiber, glow, grid_v, grid_m, t, amp,
                     contraction, freq, pdens, mu_j, la_j, nullptr, nullptr, nullptr, nsteps,
                     nthreads);
}

// The axisymmetric variant (JELLY_AXISYM=1, mpm_sim.py:127-133,163-165,236-304): as above
// with the spawn radii r_ref (M,N) and the two stress grids (M,nx,ny) of the hoop term;
// consts->axisym must be set.
int oracle_run_axisym_f32(const OracleConsts* c, int stages, int M, int N, float* x, float* v,
                          float* C, float* F, const int32_t* mat, float* Jp,
                          const float* fiber, float* glow, float* grid_v, float* grid_m,
                          float* t, const float* amp, const float* contraction,
                          const float* freq, const float* pdens, const float* mu_j,
                          const float* la_j, const float* r_ref, float* g_srr, float* g_shoop,
                          int nsteps, int nthreads) {
  return run<float>(c, stages, M, N, x, v, C, F, mat, Jp, fiber, glow, grid_v, grid_m, t, amp,
                    contraction, freq, pdens, mu_j, la_j, r_ref, g_srr, g_shoop, nsteps,
                    nthreads);
}

int oracle_run_axisym_f64(const OracleConsts* c, int stages, int M, int N, double* x,
                          double* v, double* C, double* F, const int32_t* mat, double* Jp,
                          const double* fiber, double* glow, double* grid_v, double* grid_m,
                          double* t, const double* amp, const double* contraction,
                          const double* freq, const double* pdens, const double* mu_j,
                          const double* la_j, const double* r_ref, double* g_srr,
                          double* g_shoop, int nsteps, int nthreads) {
  return run<double>(c, stages, M, N, x, v, C, F, mat, Jp, fiber, glow, grid_v, grid_m, t, amp,
                     contraction, freq, pdens, mu_j, la_j, r_ref, g_srr, g_shoop, nsteps,
                     nthreads);
}

// Activation value per instance (mpm_sim.py:173-181).
void oracle_activation_f32(const OracleConsts* c, int n, const float* t, const float* freq,
                           const float* contraction, float* out) {
  Consts<float> k(*c);
  for (int i = 0; i < n; ++i) out[i] = activation(k, t[i], freq[i], contraction[i]);
}
void oracle_activation_f64(const OracleConsts* c, int n, const double* t, const double* freq,
                           const double* contraction, double* out) {
  Consts<double> k(*c);
  for (int i = 0; i < n; ++i) out[i] = activation(k, t[i], freq[i], contraction[i]);
}

// Batched 2x2 SVD (taichi `_svd2d`): F (n,4) -> U (n,4), sig (n,2), V (n,4), R (n,4).
void oracle_svd2d_f32(int n, const float* F, float* U, float* sig, float* V, float* R) {
  for (int i = 0; i < n; ++i) {
    M2<float> A = {F[4 * i], F[4 * i + 1], F[4 * i + 2], F[4 * i + 3]}, u, vv, r;
    svd2d(A, u, sig[2 * i], sig[2 * i + 1], vv, &r);
    std::memcpy(U + 4 * i, &u, sizeof u);
    std::memcpy(V + 4 * i, &vv, sizeof vv);
    std::memcpy(R + 4 * i, &r, sizeof r);
  }
}
void oracle_svd2d_f64(int n, const double* F, double* U, double* sig, double* V, double* R) {
  for (int i = 0; i < n; ++i) {
    M2<double> A = {F[4 * i], F[4 * i + 1], F[4 * i + 2], F[4 * i + 3]}, u, vv, r;
    svd2d(A, u, sig[2 * i], sig[2 * i + 1], vv, &r);
    std::memcpy(U + 4 * i, &u, sizeof u);
    std::memcpy(V + 4 * i, &vv, sizeof vv);
    std::memcpy(R + 4 * i, &r, sizeof r);
  }
}

// P2G cell index of each position (mpm_sim.py:183): trunc(x*inv_dx - 0.5), per axis.
void oracle_base_cell_f32(const OracleConsts* c, int n, const float* x, int32_t* base) {
  Consts<float> k(*c);
  for (int i = 0; i < n; ++i) {
    base[2 * i] = cast_int(x[2 * i] * k.inv_dx - 0.5f);
    base[2 * i + 1] = cast_int(x[2 * i + 1] * k.inv_dx - 0.5f);
  }
}

// fitness_buffer (M,3) = [sum_y, count, sum_x] over payload (body_mode=0, mpm_sim.py:746-753)
// or over jelly+muscle (body_mode=1, mpm_sim.py:755-762), summed in slot order.
void oracle_stats_f32(int M, int N, const float* x, const int32_t* mat, int body_mode,
                      float* out) {
  stats<float>(M, N, x, mat, body_mode, out);
}
void oracle_stats_f64(int M, int N, const double* x, const int32_t* mat, int body_mode,
                      double* out) {
  stats<double>(M, N, x, mat, body_mode, out);
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
