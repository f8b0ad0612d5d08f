// Differentiable substep: forward and reverse-mode (adjoint) math per particle and per node.
//
// Templated on the scalar type and usable on the host and the device: the HIP kernels in
// diff_kernels.h instantiate it in fp32 and fp64 (fp64 also over fp32 buffers: state and
// grid loads widen, stores round); the same functions compiled for the host in fp64 let
// the CPU tests compare the adjoint with central differences of the independent oracle.
//
// Forward semantics follow the reference substep (c-m-x-r/jellyfih mpm_sim.py:153-340) with
// two parameters promoted to runtime values (jelly Young's modulus per instance, and a
// per-particle multiplier on the actuation amplitude). The adjoint differentiates:
//   * P2G (mpm_sim.py:183-271): B-spline weights in x, F <- (I + dt C) F, the constitutive
//     update (J = product of the clamped singular values of taichi `_svd2d`, including its
//     |S01| < 1e-5 diagonal branch; water: F <- sqrt(J) I; jelly/muscle: fixed corotated
//     stress with the polar rotation R; payload: the spectral clamp of the singular values)
//     and the APIC affine scatter;
//   * the grid op (mpm_sim.py:274-309) with the wall clamps and damping as fixed masks,
//     evaluated per stencil node from momentum and mass inside G2P (it is node-local);
//   * G2P and advection (mpm_sim.py:311-333) with the position clamp as a fixed mask.
// Non-smooth points (clamps, the range skip, the phase branch) take the derivative of the
// branch the forward took.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define JMD_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define JMD_HD inline
#endif

namespace jmd {

using std_int = int;

template <typename T>
struct Mat {
  T a, b, c, d;  // [[a, b], [c, d]]
};

template <typename T>
JMD_HD Mat<T> mmul(const Mat<T>& x, const Mat<T>& y) {
  return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c,
          x.c * y.b + x.d * y.d};
}
template <typename T>
JMD_HD Mat<T> mmul_bt(const Mat<T>& x, const Mat<T>& y) {  // x y^T
  return {x.a * y.a + x.b * y.b, x.a * y.c + x.b * y.d, x.c * y.a + x.d * y.b,
          x.c * y.c + x.d * y.d};
}
template <typename T>
JMD_HD Mat<T> mmul_at(const Mat<T>& x, const Mat<T>& y) {  // x^T y
  return {x.a * y.a + x.c * y.c, x.a * y.b + x.c * y.d, x.b * y.a + x.d * y.c,
          x.b * y.b + x.d * y.d};
}
template <typename T>
JMD_HD Mat<T> madd(const Mat<T>& x, const Mat<T>& y) {
  return {x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d};
}
template <typename T>
JMD_HD Mat<T> msub(const Mat<T>& x, const Mat<T>& y) {
  return {x.a - y.a, x.b - y.b, x.c - y.c, x.d - y.d};
}
template <typename T>
JMD_HD Mat<T> mscale(T s, const Mat<T>& x) {
  return {s * x.a, s * x.b, s * x.c, s * x.d};
}
template <typename T>
JMD_HD Mat<T> mtrans(const Mat<T>& x) {
  return {x.a, x.c, x.b, x.d};
}
template <typename T>
JMD_HD T mdot(const Mat<T>& x, const Mat<T>& y) {  // Frobenius x : y
  return x.a * y.a + x.b * y.b + x.c * y.c + x.d * y.d;
}
template <typename T>
JMD_HD T mdet(const Mat<T>& x) {
  return x.a * x.d - x.b * x.c;
}
template <typename T>
JMD_HD Mat<T> mcof(const Mat<T>& x) {  // d det / dX
  return {x.d, -x.c, -x.b, x.a};
}

// Constants of mpm_sim.py:18-58 in the arithmetic type (see physics.h for the folding).
template <typename T>
struct Consts {
  int nx, ny, damp_x, damp_y;
  T dt, dx, inv_dx, p_mass, gravity_dt, stress_scale, apic_scale;
  T mu_base, la_base, mu_pay, la_pay, water_la, act_freq;
  T x_lo, x_hi, y_hi, sig_floor, pay_lo, pay_hi, pi;
  T nu_jelly;  // Poisson ratio of jelly/muscle (0.3): mu = E/(2(1+nu)), la = E nu/((1+nu)(1-2nu))
};

// Raised-cosine activation (mpm_sim.py:173-181) and its derivatives in the contraction
// fraction and the frequency multiplier (phase = frac(t * f), d phase / d f = t between
// wraps).
template <typename T>
struct Act {
  T a, da_dc, da_df;
};

template <typename T>
JMD_HD Act<T> activation(T t, T freq, T cf, const Consts<T>& k) {
  Act<T> r{T(0), T(0), T(0)};
  T period = T(1) / (k.act_freq * freq);
  T q = floor(t / period);
  T phase = (t - period * q) / period;
  if (phase < cf) {
    T arg = phase / cf * k.pi;
    r.a = T(0.5) * (T(1) - cos(arg));
    T da_dphase = T(0.5) * sin(arg) * k.pi / cf;
    r.da_dc = -T(0.5) * sin(arg) * k.pi * phase / (cf * cf);
    r.da_df = da_dphase * t * k.act_freq;
  }
  return r;
}

template <typename T>
JMD_HD void weights(T fx, T w[3], T dw[3]) {
  T a = T(1.5) - fx, b = fx - T(1), c = fx - T(0.5);
  w[0] = T(0.5) * (a * a);
  w[1] = T(0.75) - b * b;
  w[2] = T(0.5) * (c * c);
  dw[0] = -a;
  dw[1] = T(-2) * b;
  dw[2] = c;
}

template <typename T>
JMD_HD int trunc_int(T v) {
  if (!(v == v)) return 0;
  if (v >= T(2147483647.0)) return 2147483647;
  if (v <= T(-2147483648.0)) return -2147483647 - 1;
  return static_cast<int>(v);
}

template <typename T>
JMD_HD bool base_cell(T x0, T x1, const Consts<T>& k, int& b0, int& b1, T& fx0, T& fx1) {
  b0 = trunc_int(x0 * k.inv_dx - T(0.5));
  b1 = trunc_int(x1 * k.inv_dx - T(0.5));
  fx0 = x0 * k.inv_dx - T(b0);
  fx1 = x1 * k.inv_dx - T(b1);
  return b0 >= 0 && b1 >= 0 && b0 < k.nx - 2 && b1 < k.ny - 2;
}

// taichi `_svd2d` (python/taichi/_funcs.py): polar rotation (c, s) from
// normalize(F00 + F11, F10 - F01), S = R^T F, then a Jacobi rotation of S unless
// |S01| < 1e-5. Returns R, S, singular values, V and whether the diagonal branch was taken.
template <typename T>
struct Svd {
  Mat<T> R, S, U, V;
  T s0, s1;
  bool diag;
};

template <typename T>
JMD_HD Svd<T> svd2d(const Mat<T>& A) {
  Svd<T> o;
  T x = A.a + A.d, y = A.c - A.b;
  T scale = T(1) / sqrt(x * x + y * y);
  T c = x * scale, s = y * scale;
  o.R = {c, -s, s, c};
  o.S = mmul_at(o.R, A);
  T cj, sj;
  T s1, s2;
  o.diag = fabs(o.S.b) < T(1e-5);
  if (o.diag) {
    cj = T(1);
    sj = T(0);
    s1 = o.S.a;
    s2 = o.S.d;
  } else {
    T tao = T(0.5) * (o.S.a - o.S.d);
    T w = sqrt(tao * tao + o.S.b * o.S.b);
    T tt = tao > T(0) ? o.S.b / (tao + w) : o.S.b / (tao - w);
    cj = T(1) / sqrt(tt * tt + T(1));
    sj = -tt * cj;
    s1 = cj * cj * o.S.a - T(2) * cj * sj * o.S.b + sj * sj * o.S.d;
    s2 = sj * sj * o.S.a + T(2) * cj * sj * o.S.b + cj * cj * o.S.d;
  }
  if (s1 < s2) {
    T tmp = s1;
    s1 = s2;
    s2 = tmp;
    o.V = {-sj, cj, -cj, -sj};
  } else {
    o.V = {cj, sj, -sj, cj};
  }
  o.U = mmul(o.R, o.V);
  o.s0 = s1;
  o.s1 = s2;
  return o;
}

// Material parameters (mpm_sim.py:192-203); mass of payload scaled by the density.
template <typename T>
JMD_HD void material(int mat, T mu_j, T la_j, T pdens, const Consts<T>& k, T& mu, T& la, T& mass) {
  mu = k.mu_base;
  la = k.la_base;
  mass = k.p_mass;
  if (mat == 1 || mat == 3) {
    mu = mu_j;
    la = la_j;
  } else if (mat == 2) {
    mu = k.mu_pay;
    la = k.la_pay;
    mass = mass * pdens;
  } else if (mat == 0) {
    mu = T(0);
    la = k.water_la;
  }
}

template <typename T>
JMD_HD T clampv(T v, T lo, T hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Clamped singular values (mpm_sim.py:207-218): the 1e-6 floor, then for payload the
// [0.995, 1.005] plasticity clamp; d0, d1 are the derivatives of each clamp (0 or 1).
template <typename T>
JMD_HD void clamp_sigma(int mat, const Svd<T>& sv, const Consts<T>& k, T& n0, T& n1, T& g0,
                        T& g1) {
  n0 = sv.s0 > k.sig_floor ? sv.s0 : k.sig_floor;
  n1 = sv.s1 > k.sig_floor ? sv.s1 : k.sig_floor;
  g0 = sv.s0 > k.sig_floor ? T(1) : T(0);
  g1 = sv.s1 > k.sig_floor ? T(1) : T(0);
  if (mat == 2) {
    if (n0 <= k.pay_lo || n0 >= k.pay_hi) g0 = T(0);
    if (n1 <= k.pay_lo || n1 >= k.pay_hi) g1 = T(0);
    n0 = clampv(n0, k.pay_lo, k.pay_hi);
    n1 = clampv(n1, k.pay_lo, k.pay_hi);
  }
}

// Water's J without the Jacobi step. With the polar rotation R of taichi's `_svd2d`
// (c, s) = normalize(F00 + F11, F10 - F01) and S = R^T F, tr S = |(F00 + F11, F10 - F01)|
// >= 0; when det F > 1e-6 |F|_F (> 0) both singular values exceed the 1e-6 floor, so the
// clamped product is the unclamped one: S00 * S11 in the diagonal branch (|S01| < 1e-5),
// det S = det F otherwise, equal to the SVD path up to rounding. Returns false when the
// fast form does not apply (the general path then runs).
template <typename T>
struct WaterJ {
  T J, c, s, S00, S11;
  bool diag;
};

template <typename T>
JMD_HD bool water_j(const Mat<T>& F, WaterJ<T>& w) {
  const T det = F.a * F.d - F.b * F.c;
  const T fro2 = F.a * F.a + F.b * F.b + F.c * F.c + F.d * F.d;
  if (!(det > T(0) && det * det > T(1e-12) * fro2)) return false;
  const T x = F.a + F.d, y = F.c - F.b;
  const T scale = T(1) / sqrt(x * x + y * y);
  w.c = x * scale;
  w.s = y * scale;
  const T S01 = w.c * F.b + w.s * F.d;
  w.diag = fabs(S01) < T(1e-5);
  w.S00 = w.c * F.a + w.s * F.c;
  w.S11 = -w.s * F.b + w.c * F.d;
  w.J = w.diag ? w.S00 * w.S11 : det;
  return true;
}

// Forward constitutive update of one in-range particle: new F and the affine term
// A = stress_scale * tau + mass * C (mpm_sim.py:189-254), with J the product of the clamped
// singular values of taichi's SVD as in the reference. `beta` = activation *
// instance_actuation * per-particle multiplier (muscle only).
template <typename T>
JMD_HD void p2g_forward(int mat, const Mat<T>& F0, const Mat<T>& C, T d0, T d1, T beta, T mu_j,
                        T la_j, T pdens, const Consts<T>& k, Mat<T>& Fn, Mat<T>& A, T& mass) {
  Mat<T> M = {T(1) + k.dt * C.a, k.dt * C.b, k.dt * C.c, T(1) + k.dt * C.d};
  Mat<T> F = mmul(M, F0);
  T mu, la;
  material(mat, mu_j, la_j, pdens, k, mu, la, mass);
  WaterJ<T> wj;
  if (mat == 0 && water_j(F, wj)) {
    // mu = 0: tau = la J (J - 1) I, F <- sqrt(J) I
    const T r = sqrt(wj.J), dg = (la * wj.J) * (wj.J - T(1));
    Fn = {r, T(0), T(0), r};
    A = madd(mscale(k.stress_scale, Mat<T>{dg, T(0), T(0), dg}), mscale(mass, C));
    return;
  }
  Svd<T> sv = svd2d(F);
  T n0, n1, g0, g1;
  clamp_sigma(mat, sv, k, n0, n1, g0, g1);
  T J = n0 * n1;
  if (mat == 0) {
    T r = sqrt(J);
    Fn = {r, T(0), T(0), r};
  } else if (mat == 2) {
    Mat<T> Us = {sv.U.a * n0, sv.U.b * n1, sv.U.c * n0, sv.U.d * n1};
    Fn = mmul_bt(Us, sv.V);
  } else {
    Fn = F;
  }
  Mat<T> tau = mmul_bt(mscale(T(2) * mu, msub(Fn, sv.R)), Fn);
  T dg = (la * J) * (J - T(1));
  tau.a += dg;
  tau.d += dg;
  if (mat == 3) {
    tau.a += d0 * d0 * beta * J;
    tau.b += d0 * d1 * beta * J;
    tau.c += d1 * d0 * beta * J;
    tau.d += d1 * d1 * beta * J;
  }
  A = madd(mscale(k.stress_scale, tau), mscale(mass, C));
}

// Adjoint of p2g_forward. Inputs: the same arguments plus the adjoints Ab (of A) and Fnb
// (of the new F). Outputs (accumulated): F0b, Cb, and the adjoints of mu_j, la_j, beta.
// The forward reads S = R^T F through S00, S01, S11 only (taichi's Jacobi step), so the
// adjoint of S is expressed in those three entries (S_b stored with S_b10 = 0).
template <typename T>
JMD_HD void p2g_backward(int mat, const Mat<T>& F0, const Mat<T>& C, T d0, T d1, T beta, T mu_j,
                         T la_j, T pdens, const Consts<T>& k, const Mat<T>& Ab,
                         const Mat<T>& Fnb, Mat<T>& F0b, Mat<T>& Cb, T& mub, T& lab,
                         T& betab) {
  Mat<T> M = {T(1) + k.dt * C.a, k.dt * C.b, k.dt * C.c, T(1) + k.dt * C.d};
  Mat<T> F = mmul(M, F0);
  T mu, la, mass;
  material(mat, mu_j, la_j, pdens, k, mu, la, mass);
  WaterJ<T> wj;
  if (mat == 0 && water_j(F, wj)) {
    // A = s la J (J - 1) I + mass C, Fn = sqrt(J) I
    const Mat<T> taub = mscale(k.stress_scale, Ab);
    Cb = madd(Cb, mscale(mass, Ab));
    const T J = wj.J;
    const T Jb = la * (T(2) * J - T(1)) * (taub.a + taub.d) + (Fnb.a + Fnb.d) / (T(2) * sqrt(J));
    Mat<T> Fb;
    if (wj.diag) {
      // J = S00 S11 with S = R^T F: F_b += R S_b, R_b += F S_b^T, then the polar angle
      const Mat<T> R = {wj.c, -wj.s, wj.s, wj.c};
      const Mat<T> Sb = {Jb * wj.S11, T(0), T(0), Jb * wj.S00};
      Fb = mmul(R, Sb);
      const Mat<T> Rb = mmul_bt(F, Sb);
      const Mat<T> RK = {R.b, -R.a, R.d, -R.c};
      const T thb = mdot(Rb, RK);
      const T x = F.a + F.d, y = F.c - F.b;
      Fb = madd(Fb, mscale(thb / (x * x + y * y), Mat<T>{-y, -x, x, -y}));
    } else {
      Fb = mscale(Jb, mcof(F));  // d det F / dF
    }
    F0b = madd(F0b, mmul_at(M, Fb));
    Cb = madd(Cb, mscale(k.dt, mmul_bt(Fb, F0)));
    return;
  }
  Svd<T> sv = svd2d(F);
  const Mat<T>& R = sv.R;
  T n0, n1, g0, g1;
  clamp_sigma(mat, sv, k, n0, n1, g0, g1);
  const T J = n0 * n1;
  Mat<T> Fn;
  if (mat == 0) {
    T r = sqrt(J);
    Fn = {r, T(0), T(0), r};
  } else if (mat == 2) {
    Mat<T> Us = {sv.U.a * n0, sv.U.b * n1, sv.U.c * n0, sv.U.d * n1};
    Fn = mmul_bt(Us, sv.V);
  } else {
    Fn = F;
  }
  // A = s tau + mass C
  Mat<T> taub = mscale(k.stress_scale, Ab);
  Cb = madd(Cb, mscale(mass, Ab));
  const T trt = taub.a + taub.d;
  // tau = 2 mu (Fn - R) Fn^T + la J (J-1) I + beta J d d^T
  T Jb = la * (T(2) * J - T(1)) * trt;
  if (mat == 3) {
    T ddt = d0 * d0 * taub.a + d0 * d1 * (taub.b + taub.c) + d1 * d1 * taub.d;  // d^T taub d
    Jb += beta * ddt;
    betab += J * ddt;
  }
  mub += T(2) * mdot(mmul_bt(msub(Fn, R), Fn), taub);
  lab += J * (J - T(1)) * trt;
  Mat<T> Fnb_t = madd(Fnb, mscale(T(2) * mu, msub(madd(mmul(taub, Fn), mmul_at(taub, Fn)),
                                                   mmul_at(taub, R))));
  Mat<T> Rb = mscale(-T(2) * mu, mmul(taub, Fn));
  Mat<T> Fb = {T(0), T(0), T(0), T(0)};  // adjoint of F (after the (I + dt C) update)
  Mat<T> Sb = {T(0), T(0), T(0), T(0)};  // in (S00, S01, -, S11)
  if (mat == 0) {
    Jb += (Fnb_t.a + Fnb_t.d) / (T(2) * sqrt(J));  // Fn = sqrt(J) I
  } else if (mat == 2) {
    // Fn = R h, h = V diag(n) V^T: R_b += Fn_b h^T, then the spectral adjoint of h
    Mat<T> h = mmul_at(R, Fn);
    Rb = madd(Rb, mmul_bt(Fnb_t, h));
    Mat<T> hb = mmul_at(R, Fnb_t);
    if (sv.diag) {
      // h = diag(n(S00), n(S11)) whatever the ordering
      const bool sw = !(sv.V.a == T(1));
      const T ga = sw ? g1 : g0, gd = sw ? g0 : g1;
      Sb.a += hb.a * ga;
      Sb.d += hb.d * gd;
    } else {
      T ds = sv.s0 - sv.s1;
      T g01 = fabs(ds) > T(1e-12) * (fabs(sv.s0) + fabs(sv.s1)) + T(1e-30)
                  ? (n0 - n1) / ds
                  : T(0.5) * (g0 + g1);
      Mat<T> Q = mmul(mmul_at(sv.V, hb), sv.V);
      Q = {g0 * Q.a, g01 * Q.b, g01 * Q.c, g1 * Q.d};
      Mat<T> Mm = mmul_bt(mmul(sv.V, Q), sv.V);
      Sb.a += Mm.a;
      Sb.b += Mm.b + Mm.c;
      Sb.d += Mm.d;
    }
  } else {
    Fb = madd(Fb, Fnb_t);  // Fn = F
  }
  // J = n0(s0) n1(s1): the singular values are the eigenvalues of S (Jacobi branch) or its
  // diagonal (diagonal branch).
  const T c0 = Jb * n1 * g0, c1 = Jb * n0 * g1;
  if (sv.diag) {
    const bool sw = !(sv.V.a == T(1));
    Sb.a += sw ? c1 : c0;
    Sb.d += sw ? c0 : c1;
  } else {
    // d s_i = v_i^T dS v_i with v_i the columns of V
    const T v00 = sv.V.a, v10 = sv.V.c, v01 = sv.V.b, v11 = sv.V.d;
    Sb.a += c0 * v00 * v00 + c1 * v01 * v01;
    Sb.b += T(2) * (c0 * v00 * v10 + c1 * v01 * v11);
    Sb.d += c0 * v10 * v10 + c1 * v11 * v11;
  }
  // S = R^T F (entries 00, 01, 11): F_b += R S_b, R_b += F S_b^T
  Sb.c = T(0);
  Fb = madd(Fb, mmul(R, Sb));
  Rb = madd(Rb, mmul_bt(F, Sb));
  // R = rot(theta), theta = atan2(F10 - F01, F00 + F11): dR/dtheta = R K, K = [[0,-1],[1,0]]
  Mat<T> RK = {R.b, -R.a, R.d, -R.c};
  T thb = mdot(Rb, RK);
  T x = F.a + F.d, y = F.c - F.b;
  T r2 = x * x + y * y;
  Fb = madd(Fb, mscale(thb / r2, Mat<T>{-y, -x, x, -y}));
  // F = (I + dt C) F0: F0_b += (I + dt C)^T F_b, C_b += dt F_b F0^T
  F0b = madd(F0b, mmul_at(M, Fb));
  Cb = madd(Cb, mscale(k.dt, mmul_bt(Fb, F0)));
}

// Grid op of one node (mpm_sim.py:274-309), with the damping factor of the node. Returns the
// velocity and the wall-clamp mask (1 where the component was kept).
//
// Damping along one axis of n nodes with a band of d (mpm_sim.py:287-294): the factor is
// 0.95 + 0.05 q / d with q = i below the band's end and q = n - i above its start, 1 in
// between. The reference multiplies damp = 1 by the x factor, then the y factor; the two
// conditions on one axis never hold together (d = n/20), so damp = fx(i) fy(j) exactly,
// each factor with the reference's operations and one division.
template <typename T>
JMD_HD T band_factor(int i, int n, int d) {
  const int q = i < d ? i : n - i;
  return (i < d || i > n - d) ? T(0.95) + T(0.05) * T(q) / T(d) : T(1);
}

template <typename T>
JMD_HD T damping(int i, int j, const Consts<T>& k) {
  return band_factor<T>(i, k.nx, k.damp_x) * band_factor<T>(j, k.ny, k.damp_y);
}

// The normalisation multiplies by 1/m, one division per node instead of one per component
// (within an ulp of p / m; the adjoint below differentiates the same products).
template <typename T>
JMD_HD void gridop_forward(T px, T py, T m, int i, int j, const Consts<T>& k, T& g0, T& g1) {
  if (!(m > T(0))) {
    g0 = px;
    g1 = py;
    return;
  }
  const T damp = damping(i, j, k), rm = T(1) / m;
  g0 = (px * rm) * damp;
  g1 = (py * rm) * damp;
  if (i < 3 && g0 < T(0)) g0 = T(0);
  if (i > k.nx - 3 && g0 > T(0)) g0 = T(0);
  if (j < 3 && g1 < T(0)) g1 = T(0);
  if (j > k.ny - 3 && g1 > T(0)) g1 = T(0);
}

// ---------------------------------------------------------------------------------------
// Per-particle transfer loops. State layout per particle (12 values): x0 x1 v0 v1 C00 C01
// C10 C11 F00 F01 F10 F11. Grids of one instance: three planes (momentum x, momentum y,
// mass) of nx*ny values, node (i, j) at j*nx + i. Scatters hand per-node contributions to
// an `emit(i, j, c0, c1, c2)` callable, which accumulates them (LDS tile, global atomics,
// or plain adds on the host).

template <typename T>
struct PState {
  T x0, x1, v0, v1;
  Mat<T> C, F;
};

// A state of arithmetic type A from a buffer of type S (fp32 buffers with fp64 arithmetic:
// exact widening on load, one rounding on store).
template <typename A, typename S>
JMD_HD PState<A> load_state_as(const S* s) {
  return {A(s[0]), A(s[1]), A(s[2]), A(s[3]), {A(s[4]), A(s[5]), A(s[6]), A(s[7])},
          {A(s[8]), A(s[9]), A(s[10]), A(s[11])}};
}
template <typename T>
JMD_HD PState<T> load_state(const T* s) {
  return load_state_as<T>(s);
}
template <typename S, typename A>
JMD_HD void store_state(S* s, const PState<A>& p) {
  s[0] = S(p.x0); s[1] = S(p.x1); s[2] = S(p.v0); s[3] = S(p.v1);
  s[4] = S(p.C.a); s[5] = S(p.C.b); s[6] = S(p.C.c); s[7] = S(p.C.d);
  s[8] = S(p.F.a); s[9] = S(p.F.b); s[10] = S(p.F.c); s[11] = S(p.F.d);
}

// P2G of one particle (mpm_sim.py:183-271): updates F in `p` and hands the momentum and
// mass contributions of each stencil node to `emit(i, j, momentum_x, momentum_y, mass)`.
// Returns false (F unchanged, nothing emitted) outside the transfer range.
template <typename T, typename Emit>
JMD_HD bool p2g_particle(PState<T>& p, int mat, T d0, T d1, T beta, T mu_j, T la_j, T pdens,
                         const Consts<T>& k, Emit&& emit) {
  int b0, b1;
  T fx0, fx1;
  if (!base_cell(p.x0, p.x1, k, b0, b1, fx0, fx1)) return false;
  T wx[3], wy[3], dwx[3], dwy[3];
  weights(fx0, wx, dwx);
  weights(fx1, wy, dwy);
  Mat<T> Fn, A;
  T mass;
  p2g_forward(mat, p.F, p.C, d0, d1, beta, mu_j, la_j, pdens, k, Fn, A, mass);
  p.F = Fn;
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) {
      T dp0 = (T(ii) - fx0) * k.dx, dp1 = (T(jj) - fx1) * k.dx;
      T w = wx[ii] * wy[jj];
      emit(b0 + ii, b1 + jj, w * (mass * p.v0 + (A.a * dp0 + A.b * dp1)),
           w * (mass * p.v1 + (A.c * dp0 + A.d * dp1)), w * mass);
    }
  return true;
}

// Velocity of node (i, j) after the grid op, from its momentum and mass planes.
template <typename T, typename S>
JMD_HD void node_velocity(const S* gp0, const S* gp1, const S* gm, int i, int j,
                          const Consts<T>& k, T& a0, T& a1) {
  const int node = j * k.nx + i;
  gridop_forward(T(gp0[node]), T(gp1[node]), T(gm[node]), i, j, k, a0, a1);
}

// What the G2P adjoint needs of one node: its velocity after the grid op (v0, v1) and the
// grid op's adjoint as a linear map from the velocity adjoint (g0, g1) to the momentum and
// mass adjoints: q0 = A0 g0, q1 = A1 g1, qm = B0 g0 + B1 g1 (with the wall-clamp masks
// folded into A and B: the clamped components carry no adjoint).
template <typename T>
struct NodeCoef {
  T v0, v1, A0, A1, B0, B1;
};

template <typename T>
JMD_HD NodeCoef<T> node_coef(T px, T py, T m, int i, int j, const Consts<T>& k) {
  NodeCoef<T> c;
  if (!(m > T(0))) {
    c.v0 = px;
    c.v1 = py;
    c.A0 = T(1);
    c.A1 = T(1);
    c.B0 = T(0);
    c.B1 = T(0);
    return c;
  }
  // gridop_forward's products, then its wall clamps as masks
  const T damp = damping(i, j, k), rm = T(1) / m;
  const T u0 = (px * rm) * damp, u1 = (py * rm) * damp;  // before the wall clamps
  const bool k0 = !((i < 3 && u0 < T(0)) || (i > k.nx - 3 && u0 > T(0)));
  const bool k1 = !((j < 3 && u1 < T(0)) || (j > k.ny - 3 && u1 > T(0)));
  c.v0 = k0 ? u0 : T(0);
  c.v1 = k1 ? u1 : T(0);
  c.A0 = k0 ? damp * rm : T(0);
  c.A1 = k1 ? damp * rm : T(0);
  c.B0 = k0 ? -u0 * rm : T(0);
  c.B1 = k1 ? -u1 * rm : T(0);
  return c;
}

// G2P + gravity + advection + clamp of one particle (mpm_sim.py:311-333). `vel(i, j, a0, a1)`
// gives the velocity of node (i, j) after the grid op (mpm_sim.py:274-309).
template <typename T, typename Vel>
JMD_HD void g2p_particle_v(PState<T>& p, const Consts<T>& k, Vel&& vel) {
  int b0, b1;
  T fx0, fx1;
  if (base_cell(p.x0, p.x1, k, b0, b1, fx0, fx1)) {
    T wx[3], wy[3], dwx[3], dwy[3];
    weights(fx0, wx, dwx);
    weights(fx1, wy, dwy);
    T v0 = T(0), v1 = T(0);
    Mat<T> C = {T(0), T(0), T(0), T(0)};
    for (int ii = 0; ii < 3; ++ii)
      for (int jj = 0; jj < 3; ++jj) {
        T a0, a1;
        vel(b0 + ii, b1 + jj, a0, a1);
        T w = wx[ii] * wy[jj];
        T e0 = T(ii) - fx0, e1 = T(jj) - fx1;
        v0 += w * a0;
        v1 += w * a1;
        T ws = k.apic_scale * w;
        C.a += ws * (a0 * e0);
        C.b += ws * (a0 * e1);
        C.c += ws * (a1 * e0);
        C.d += ws * (a1 * e1);
      }
    p.v0 = v0;
    p.v1 = v1;
    p.C = C;
  }
  p.v1 = p.v1 - k.gravity_dt;
  p.x0 = p.x0 + k.dt * p.v0;
  p.x1 = p.x1 + k.dt * p.v1;
  p.x0 = p.x0 < k.x_lo ? k.x_lo : (p.x0 > k.x_hi ? k.x_hi : p.x0);
  p.x1 = p.x1 < k.x_lo ? k.x_lo : (p.x1 > k.y_hi ? k.y_hi : p.x1);
}

// The same, with the grid op evaluated per stencil node from momentum gp0, gp1, mass gm.
template <typename T, typename S>
JMD_HD void g2p_particle(PState<T>& p, const Consts<T>& k, const S* gp0, const S* gp1,
                         const S* gm) {
  g2p_particle_v(p, k, [&](int i, int j, T& a0, T& a1) {
    node_velocity(gp0, gp1, gm, i, j, k, a0, a1);
  });
}

// The grid-side result of the G2P adjoint of one particle, in separable form: the adjoint
// of node (b0 + i, b1 + j)'s velocity (after the grid op) is
//   g0 = wy[j] P0[i] + ey[j] Q0[i],  g1 = wy[j] P1[i] + ey[j] Q1[i],
// and through that node's grid op (NodeCoef) it is d momentum = (A0 g0, A1 g1) and
// d mass = B0 g0 + B1 g1. These are linear in g with per-node coefficients, so a caller
// that sums g over the particles of a node applies the coefficients once per node. Out of
// range (inr false), every factor is 0.
template <typename T>
struct SepFactors {
  int b0, b1;
  bool inr;
  T wy[3], ey[3], P0[3], Q0[3], P1[3], Q1[3];
};

// Adjoint of g2p_particle. `s` is the particle's state before the substep (its x sets the
// weights; v, C pass through when out of range), `b_out` the adjoint of the state after.
// Accumulates the x adjoint and the pass-through v, C adjoints into `b_in` and returns the
// node-velocity adjoints as SepFactors. `vel(i, j, a0, a1)` gives node (i, j)'s velocity
// after the grid op. One pass over the stencil: the forward velocity (summed in
// g2p_particle_v's order, so that the wall-clamp decisions match the forward) and, per
// stencil column i, the moments over rows j of the node velocities a against wy, ey =
// wy e1, dwy and dwy e1 (e1 = j - fx1), from which the x adjoint follows in closed form:
//   d L/d fx0 = sum_i dwx_i (p0_i U0_i + q0 E0_i + p1_i U1_i + q1 E1_i)
//               - apic wx_i (Cb00 U0_i + Cb10 U1_i)
//   d L/d fx1 = sum_i wx_i (p0_i D0_i + q0 H0_i + p1_i D1_i + q1 H1_i)
//               - apic wx_i (Cb01 U0_i + Cb11 U1_i)
// with p0_i = vb0 + apic Cb00 e0_i, p1_i = vb1 + apic Cb10 e0_i, q0 = apic Cb01,
// q1 = apic Cb11 (Cb the adjoint of C, vb of v after G2P, e0_i = i - fx0), U/E/D/H the
// moments of a against wy/ey/dwy/dwy e1. This is the node-by-node sum of
// w (vb . a + apic a^T Cb e) differentiated in fx, regrouped by stencil column.
template <typename T, typename Vel>
JMD_HD void g2p_backward_sep(const PState<T>& s, const PState<T>& b_out, const Consts<T>& k,
                             Vel&& vel, PState<T>& b_in, SepFactors<T>& sf) {
  T v0n = s.v0, v1n = s.v1;
  T fx0, fx1;
  sf.inr = base_cell(s.x0, s.x1, k, sf.b0, sf.b1, fx0, fx1);
  T wx[3], dwx[3], dwy[3], dey[3];
  T U0[3], U1[3], E0[3], E1[3], D0[3], D1[3], H0[3], H1[3];
  if (sf.inr) {
    weights(fx0, wx, dwx);
    weights(fx1, sf.wy, dwy);
    for (int jj = 0; jj < 3; ++jj) {
      const T e1 = T(jj) - fx1;
      sf.ey[jj] = sf.wy[jj] * e1;
      dey[jj] = dwy[jj] * e1;
    }
    v0n = T(0);
    v1n = T(0);
    for (int ii = 0; ii < 3; ++ii) {
      U0[ii] = U1[ii] = E0[ii] = E1[ii] = D0[ii] = D1[ii] = H0[ii] = H1[ii] = T(0);
      for (int jj = 0; jj < 3; ++jj) {
        T a0, a1;
        vel(sf.b0 + ii, sf.b1 + jj, a0, a1);
        const T w = wx[ii] * sf.wy[jj];
        v0n += w * a0;
        v1n += w * a1;
        U0[ii] += sf.wy[jj] * a0;
        U1[ii] += sf.wy[jj] * a1;
        E0[ii] += sf.ey[jj] * a0;
        E1[ii] += sf.ey[jj] * a1;
        D0[ii] += dwy[jj] * a0;
        D1[ii] += dwy[jj] * a1;
        H0[ii] += dey[jj] * a0;
        H1[ii] += dey[jj] * a1;
      }
    }
  }
  v1n = v1n - k.gravity_dt;
  T xu0 = s.x0 + k.dt * v0n, xu1 = s.x1 + k.dt * v1n;
  T k0 = (xu0 < k.x_lo || xu0 > k.x_hi) ? T(0) : T(1);
  T k1 = (xu1 < k.x_lo || xu1 > k.y_hi) ? T(0) : T(1);
  T xb0 = k0 * b_out.x0, xb1 = k1 * b_out.x1;
  b_in.x0 += xb0;
  b_in.x1 += xb1;
  T vb0 = b_out.v0 + k.dt * xb0, vb1 = b_out.v1 + k.dt * xb1;  // adjoint of v after G2P
  if (!sf.inr) {
    b_in.v0 += vb0;
    b_in.v1 += vb1;
    b_in.C = madd(b_in.C, b_out.C);
    for (int q = 0; q < 3; ++q)
      sf.wy[q] = sf.ey[q] = sf.P0[q] = sf.Q0[q] = sf.P1[q] = sf.Q1[q] = T(0);
    return;
  }
  const Mat<T>& Cb = b_out.C;
  const T as = k.apic_scale;
  const T q0 = as * Cb.b, q1 = as * Cb.d;
  T f0 = T(0), f1 = T(0);  // d L / d fx
  for (int ii = 0; ii < 3; ++ii) {
    const T e0 = T(ii) - fx0;
    const T p0 = vb0 + as * Cb.a * e0, p1 = vb1 + as * Cb.c * e0;
    f0 += dwx[ii] * (p0 * U0[ii] + q0 * E0[ii] + p1 * U1[ii] + q1 * E1[ii]);
    f1 += wx[ii] * (p0 * D0[ii] + q0 * H0[ii] + p1 * D1[ii] + q1 * H1[ii]);
    f0 -= as * wx[ii] * (Cb.a * U0[ii] + Cb.c * U1[ii]);
    f1 -= as * wx[ii] * (Cb.b * U0[ii] + Cb.d * U1[ii]);
    sf.P0[ii] = wx[ii] * p0;
    sf.Q0[ii] = wx[ii] * q0;
    sf.P1[ii] = wx[ii] * p1;
    sf.Q1[ii] = wx[ii] * q1;
  }
  b_in.x0 += k.inv_dx * f0;
  b_in.x1 += k.inv_dx * f1;
}

// The node-velocity adjoints of SepFactors, one `emit(i, j, g0, g1)` per stencil node
// (nothing when out of range).
template <typename T, typename Emit>
JMD_HD void emit_sep(const SepFactors<T>& sf, Emit&& emit) {
  if (!sf.inr) return;
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj)
      emit(sf.b0 + ii, sf.b1 + jj, sf.wy[jj] * sf.P0[ii] + sf.ey[jj] * sf.Q0[ii],
           sf.wy[jj] * sf.P1[ii] + sf.ey[jj] * sf.Q1[ii]);
}

// Node (i, j)'s momentum and mass adjoints from its velocity adjoint (g0, g1), through the
// node's grid op.
template <typename T>
JMD_HD void through_gridop(const NodeCoef<T>& c, T g0, T g1, T& pb0, T& pb1, T& mb) {
  pb0 = c.A0 * g0;
  pb1 = c.A1 * g1;
  mb = c.B0 * g0 + c.B1 * g1;
}

// Adjoint of p2g_particle (gather from the momentum/mass adjoints pb0, pb1, mb). `b_out`
// holds the adjoint of the new F (b_out.F); results accumulate into `b_in` and the
// parameter adjoints (mu_j, la_j, beta).
template <typename T>
JMD_HD void p2g_backward_particle(const PState<T>& s, int mat, T d0, T d1, T beta, T mu_j, T la_j,
                                  T pdens, const Consts<T>& k, const T* pb0, const T* pb1,
                                  const T* mb, const Mat<T>& Fnb, PState<T>& b_in, T& mub,
                                  T& lab, T& betab) {
  int b0, b1;
  T fx0, fx1;
  if (!base_cell(s.x0, s.x1, k, b0, b1, fx0, fx1)) {
    b_in.F = madd(b_in.F, Fnb);  // F unchanged when P2G skips the particle
    return;
  }
  T wx[3], wy[3], dwx[3], dwy[3];
  weights(fx0, wx, dwx);
  weights(fx1, wy, dwy);
  Mat<T> Fn, A;
  T mass;
  p2g_forward(mat, s.F, s.C, d0, d1, beta, mu_j, la_j, pdens, k, Fn, A, mass);
  // The gather of the node adjoints (q0, q1, qm) = (pb0, pb1, mb) at node (b0 + i, b1 + j),
  // with w = wx_i wy_j and dpos = dx (e0_i, e1_j), e0_i = i - fx0, e1_j = j - fx1:
  //   vb = mass sum w q,  Ab = sum w q dpos^T,
  //   d L/d fx0 = sum dwx_i wy_j lw - dx w (A^T q)_0,  d L/d fx1 = sum wx_i dwy_j lw - dx w (A^T q)_1,
  //   lw = q0 (mass v0 + (A dpos)_0) + q1 (mass v1 + (A dpos)_1) + qm mass,
  // regrouped by stencil column i: per column, the moments over rows j of q0, q1, qm against
  // wy, ey = wy e1, dwy and dwy e1 (ten sums), folded into the results before the next column.
  T ey[3], dey[3];
  for (int jj = 0; jj < 3; ++jj) {
    const T e1 = T(jj) - fx1;
    ey[jj] = wy[jj] * e1;
    dey[jj] = dwy[jj] * e1;
  }
  const T s0 = k.dx * A.b, s1 = k.dx * A.d;  // the e1 parts of (A dpos)
  Mat<T> Ab = {T(0), T(0), T(0), T(0)};
  T vb0 = T(0), vb1 = T(0), f0 = T(0), f1 = T(0);
  for (int ii = 0; ii < 3; ++ii) {
    T U0 = T(0), U1 = T(0), Um = T(0), E0 = T(0), E1 = T(0);
    T D0 = T(0), D1 = T(0), Dm = T(0), H0 = T(0), H1 = T(0);
    for (int jj = 0; jj < 3; ++jj) {
      const int node = (b1 + jj) * k.nx + (b0 + ii);
      const T q0 = pb0[node], q1 = pb1[node], qm = mb[node];
      U0 += wy[jj] * q0;
      U1 += wy[jj] * q1;
      Um += wy[jj] * qm;
      E0 += ey[jj] * q0;
      E1 += ey[jj] * q1;
      D0 += dwy[jj] * q0;
      D1 += dwy[jj] * q1;
      Dm += dwy[jj] * qm;
      H0 += dey[jj] * q0;
      H1 += dey[jj] * q1;
    }
    const T e0 = T(ii) - fx0, dp0 = e0 * k.dx;
    const T r0 = mass * s.v0 + A.a * dp0, r1 = mass * s.v1 + A.c * dp0;  // (A dpos) at e1 = 0
    vb0 += wx[ii] * U0;
    vb1 += wx[ii] * U1;
    Ab.a += wx[ii] * dp0 * U0;
    Ab.b += wx[ii] * E0;
    Ab.c += wx[ii] * dp0 * U1;
    Ab.d += wx[ii] * E1;
    f0 += dwx[ii] * (r0 * U0 + s0 * E0 + r1 * U1 + s1 * E1 + mass * Um);
    f1 += wx[ii] * (r0 * D0 + s0 * H0 + r1 * D1 + s1 * H1 + mass * Dm);
    f0 -= k.dx * wx[ii] * (A.a * U0 + A.c * U1);
    f1 -= k.dx * wx[ii] * (A.b * U0 + A.d * U1);
  }
  Ab.b *= k.dx;
  Ab.d *= k.dx;
  b_in.v0 += mass * vb0;
  b_in.v1 += mass * vb1;
  b_in.x0 += k.inv_dx * f0;
  b_in.x1 += k.inv_dx * f1;
  Mat<T> F0b = {T(0), T(0), T(0), T(0)}, Cb = {T(0), T(0), T(0), T(0)};
  p2g_backward(mat, s.F, s.C, d0, d1, beta, mu_j, la_j, pdens, k, Ab, Fnb, F0b, Cb, mub, lab,
               betab);
  b_in.F = madd(b_in.F, F0b);
  b_in.C = madd(b_in.C, Cb);
}

}  // namespace jmd
