// Device-side physics of one MLS-MPM substep, fp32, shared by every kernel path.
//
// Each function restates a piece of the reference's `substep()` (c-m-x-r/jellyfih
// mpm_sim.py:153-340); the comment on each names the lines. The 2x2 SVD restates taichi
// 1.7.4's `_polar_decompose2d` / `_svd2d` (python/taichi/_funcs.py), which the reference
// calls at mpm_sim.py:206.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jm {

// Constants of mpm_sim.py:18-58 as the kernels use them. Products that the reference
// folds in Python f64 (stress scale :250, dt*gravity :329, 4*inv_dx :326) are folded on
// the host in double and rounded to fp32 once.
struct PhysConsts {
  int nx, ny, damp_x, damp_y;
  // Boundary damping factor per node column (nx floats) then per node row (ny floats),
  // filled on the host from damping_factor(): the grid op's damp = damp[i] * damp[nx + j].
  const float* damp;
  float dt, dx, inv_dx, p_mass, gravity_dt, stress_scale, apic_scale;
  float mu_base, la_base, mu_pay, la_pay, water_la, act_freq;
  float x_lo, x_hi, y_hi, sig_floor, pay_lo, pay_hi, pi;
  // axisymmetric variant (mpm_sim.py:246-247): floors of the spawn and current radius
  float r_ref_floor, r_floor;
};

struct Mat2 {
  float a00, a01, a10, a11;
};

__device__ __forceinline__ Mat2 mul(const Mat2& a, const Mat2& b) {
  return {a.a00 * b.a00 + a.a01 * b.a10, a.a00 * b.a01 + a.a01 * b.a11,
          a.a10 * b.a00 + a.a11 * b.a10, a.a10 * b.a01 + a.a11 * b.a11};
}

__device__ __forceinline__ Mat2 mul_bt(const Mat2& a, const Mat2& b) {  // a @ b^T
  return {a.a00 * b.a00 + a.a01 * b.a01, a.a00 * b.a10 + a.a01 * b.a11,
          a.a10 * b.a00 + a.a11 * b.a01, a.a10 * b.a10 + a.a11 * b.a11};
}

// Raised-cosine activation (mpm_sim.py:173-181). Taichi's float `%` is the floored
// remainder t - P*floor(t/P). No FMA contraction, so the phase matches the fp32 reference.
__device__ __forceinline__ float activation(float t, float freq, float contraction,
                                            const PhysConsts& k) {
#pragma clang fp contract(off)
  float period = 1.0f / (k.act_freq * freq);
  float q = floorf(t / period);
  float phase = (t - period * q) / period;
  float a = 0.0f;
  if (phase < contraction) a = 0.5f * (1.0f - cosf(phase / contraction * k.pi));
  return a;
}

// P2G base cell and fractional offset along one axis (mpm_sim.py:183,186): the cast
// truncates toward zero; NaN converts to 0.
__device__ __forceinline__ void base_frac(float x, float inv_dx, int& b, float& fx) {
#pragma clang fp contract(off)
  float s = x * inv_dx;
  b = __float2int_rz(s - 0.5f);
  fx = s - (float)b;
}

// Quadratic B-spline weights (mpm_sim.py:187).
__device__ __forceinline__ void bspline(float fx, float w[3]) {
  float a = 1.5f - fx, b = fx - 1.0f, c = fx - 0.5f;
  w[0] = 0.5f * (a * a);
  w[1] = 0.75f - b * b;
  w[2] = 0.5f * (c * c);
}

__device__ __forceinline__ bool in_range(int b0, int b1, const PhysConsts& k) {
  return b0 >= 0 && b1 >= 0 && b0 < k.nx - 2 && b1 < k.ny - 2;  // mpm_sim.py:184,316
}

// taichi `_svd2d`: A = U diag(s1,s2) V^T with det U = det V = +1 and s1 >= s2.
__device__ __forceinline__ void svd2d(const Mat2& A, Mat2& U, float& s1, float& s2, Mat2& V) {
  float x = A.a00 + A.a11, y = A.a10 - A.a01;
  float scale = 1.0f / sqrtf(x * x + y * y);
  float c = x * scale, s = y * scale;
  Mat2 R = {c, -s, s, c};
  Mat2 S = {c * A.a00 + s * A.a10, c * A.a01 + s * A.a11, -s * A.a00 + c * A.a10,
            -s * A.a01 + c * A.a11};  // R^T @ A
  float cj, sj;
  if (fabsf(S.a01) < 1e-5f) {
    cj = 1.0f;
    sj = 0.0f;
    s1 = S.a00;
    s2 = S.a11;
  } else {
    float tao = 0.5f * (S.a00 - S.a11);
    float w = sqrtf(tao * tao + S.a01 * S.a01);
    float tt = tao > 0.0f ? S.a01 / (tao + w) : S.a01 / (tao - w);
    cj = 1.0f / sqrtf(tt * tt + 1.0f);
    sj = -tt * cj;
    s1 = cj * cj * S.a00 - 2.0f * cj * sj * S.a01 + sj * sj * S.a11;
    s2 = sj * sj * S.a00 + 2.0f * cj * sj * S.a01 + cj * cj * S.a11;
  }
  if (s1 < s2) {
    float tmp = s1;
    s1 = s2;
    s2 = tmp;
    V = {-sj, cj, -cj, -sj};
  } else {
    V = {cj, sj, -sj, cj};
  }
  U = mul(R, V);
}

// Per-particle material parameters (mpm_sim.py:192-203).
struct InstParams {
  const float* amp;          // instance_actuation
  const float* contraction;  // instance_act_contraction
  const float* freq;         // instance_freq
  const float* pdens;        // instance_payload_density
  const float* mu_j;         // jelly/muscle Lame mu (from E_jelly, nu=0.3)
  const float* la_j;         // jelly/muscle Lame lambda
};

// Result of the P2G particle update: new F, the scatter's affine term and mass, and
// the plastic volume ratio factors of mpm_sim.py:216.
struct P2GOut {
  Mat2 F;
  Mat2 affine;
  float mass;
  float q0, q1;      // Jp <- (Jp * q0) * q1
  bool jp_changes;   // false when q0 == q1 == 1 (every non-payload particle with finite F)
  // for the axisymmetric variant's hoop term: the unscaled Kirchhoff stress's (0,0)
  // entry, the material's Lame parameters and J after plasticity
  float s00, mu, la, J;
};

// F update, SVD, plasticity, fixed-corotated + active stress, APIC affine term for one
// in-range particle (mpm_sim.py:189-254). `amp_act` = instance_actuation * activation.
__device__ __forceinline__ P2GOut p2g_update(int mat, const Mat2& F0, const Mat2& C, float fib0,
                                             float fib1, float amp_act, float mu_j, float la_j,
                                             float pdens, const PhysConsts& k) {
  P2GOut o;
  // mpm_sim.py:189  F <- (I + dt*C) @ F
  Mat2 A = {1.0f + k.dt * C.a00, k.dt * C.a01, k.dt * C.a10, 1.0f + k.dt * C.a11};
  Mat2 F = mul(A, F0);
  // mpm_sim.py:192-203
  float mu = k.mu_base, la = k.la_base, mass = k.p_mass;
  if (mat == 1 || mat == 3) {
    mu = mu_j;
    la = la_j;
  } else if (mat == 2) {
    mu = k.mu_pay;
    la = k.la_pay;
    mass = mass * pdens;
  } else if (mat == 0) {
    mu = 0.0f;
    la = k.water_la;
  }
  // Water (mu = 0) needs only J = s0*s1 from the SVD (the deviatoric term is 0*(F - R)F^T,
  // and F is reset to sqrt(J) I). When det F > 1e-6 * |F|_F both singular values exceed
  // the 1e-6 floor (s_min = det/s_max >= det/|F|_F), so no clamp applies and s0*s1 = det F:
  // the SVD is skipped. Degenerate F (det near 0 or negative) takes the full path below.
  if (mat == 0) {
    const float det = F.a00 * F.a11 - F.a01 * F.a10;
    const float fro2 = F.a00 * F.a00 + F.a01 * F.a01 + F.a10 * F.a10 + F.a11 * F.a11;
    if (det > 0.0f && det * det > 1e-12f * fro2) {
      const float J = det;
      const float r = __builtin_amdgcn_sqrtf(J);  // v_sqrt_f32 (1 ulp; J is near 1 here)
      const float dg = (la * J) * (J - 1.0f);
      o.q0 = o.q1 = 1.0f;
      o.jp_changes = false;
      o.F = {r, 0.0f, 0.0f, r};
      o.affine = {k.stress_scale * dg + mass * C.a00, mass * C.a01, mass * C.a10,
                  k.stress_scale * dg + mass * C.a11};
      o.mass = mass;
      o.s00 = dg;
      o.mu = mu;
      o.la = la;
      o.J = J;
      return o;
    }
  }
  // mpm_sim.py:206-218
  Mat2 U, V;
  float s0, s1;
  svd2d(F, U, s0, s1, V);
  s0 = fmaxf(s0, k.sig_floor);
  s1 = fmaxf(s1, k.sig_floor);
  float n0 = s0, n1 = s1;
  if (mat == 2) {
    n0 = fminf(fmaxf(s0, k.pay_lo), k.pay_hi);
    n1 = fminf(fmaxf(s1, k.pay_lo), k.pay_hi);
  }
  if (mat == 2) {
    o.q0 = s0 / n0;
    o.q1 = s1 / n1;
    o.jp_changes = true;
  } else {  // n == s: the ratio is exactly 1, or NaN when s = +inf (inf/inf)
    o.q0 = s0 < INFINITY ? 1.0f : __int_as_float(0x7fc00000);
    o.q1 = s1 < INFINITY ? 1.0f : __int_as_float(0x7fc00000);
    o.jp_changes = !(s0 < INFINITY && s1 < INFINITY);
  }
  float J = n0 * n1;
  // mpm_sim.py:220-223
  if (mat == 0) {
    float r = sqrtf(J);
    F = {r, 0.0f * r, 0.0f * r, r};
  } else if (mat == 2) {
    Mat2 US = {U.a00 * n0, U.a01 * n1, U.a10 * n0, U.a11 * n1};
    F = mul_bt(US, V);
  }
  // mpm_sim.py:226-227  stress = 2 mu (F - U V^T) F^T + I la J (J-1)
  Mat2 R = mul_bt(U, V);
  float mu2 = 2.0f * mu;
  Mat2 D = {mu2 * (F.a00 - R.a00), mu2 * (F.a01 - R.a01), mu2 * (F.a10 - R.a10),
            mu2 * (F.a11 - R.a11)};
  Mat2 st = mul_bt(D, F);
  float dg = (la * J) * (J - 1.0f);
  st.a00 += dg;
  st.a11 += dg;
  // mpm_sim.py:230-233  active muscle stress along the fibre
  if (mat == 3) {
    st.a00 += ((fib0 * fib0) * amp_act) * J;
    st.a01 += ((fib0 * fib1) * amp_act) * J;
    st.a10 += ((fib1 * fib0) * amp_act) * J;
    st.a11 += ((fib1 * fib1) * amp_act) * J;
  }
  // mpm_sim.py:250,254
  o.affine = {k.stress_scale * st.a00 + mass * C.a00, k.stress_scale * st.a01 + mass * C.a01,
              k.stress_scale * st.a10 + mass * C.a10, k.stress_scale * st.a11 + mass * C.a11};
  o.F = F;
  o.mass = mass;
  o.s00 = st.a00;
  o.mu = mu;
  o.la = la;
  o.J = J;
  return o;
}

// Damping factor of mpm_sim.py:287-294 along one axis of n nodes with a band of d: the
// reference multiplies damp = 1 by the x factor, then the y factor; the two x (or y)
// conditions never hold together (d = n/20), so damp = fx(i) * fy(j) with each factor
// computed exactly as the reference's fp32 expression. Host and device share this code.
__host__ __device__ inline float damping_factor(int i, int n, int d) {
#pragma clang fp contract(off)
  float f = 1.0f;
  if (i < d) f = f * (0.95f + 0.05f * (float)i / (float)d);
  if (i > n - d) f = f * (0.95f + 0.05f * (float)(n - i) / (float)d);
  return f;
}

// Grid update of one node (mpm_sim.py:274-309): normalise by mass, boundary damping band
// of n/20 cells on all four sides (factors from the table in k.damp), then the wall
// clamps. Nodes with no mass keep their (zero) momentum, as in the reference. The
// normalisation multiplies by the hardware reciprocal (v_rcp_f32, 1 ulp) instead of two
// IEEE divisions of ~10 instructions each: the node sums it divides already differ from a
// sequential fp32 sum in their last bits, and the reference's own Taichi build divides
// with fast math.
__device__ __forceinline__ float2 grid_node(float px, float py, float m, int i, int j,
                                            const PhysConsts& k) {
  if (!(m > 0.0f)) return make_float2(px, py);
  const float rm = __builtin_amdgcn_rcpf(m);
  float g0 = px * rm, g1 = py * rm;
  const float damp = k.damp[i] * k.damp[k.nx + j];
  g0 = g0 * damp;
  g1 = g1 * damp;
  if (i < 3 && g0 < 0.0f) g0 = 0.0f;
  if (i > k.nx - 3 && g0 > 0.0f) g0 = 0.0f;
  if (j < 3 && g1 < 0.0f) g1 = 0.0f;
  if (j > k.ny - 3 && g1 > 0.0f) g1 = 0.0f;
  return make_float2(g0, g1);
}

// Grid update of one node in the axisymmetric variant (mpm_sim.py:274-309 with axisym):
// the same normalisation and damping, then the hoop correction dt (s_rr - s_tt) / (m r) off
// the axis cell (s_rr, s_tt: the node's mass-weighted Kirchhoff stresses, r = (i + 1/2) dx)
// and the axis condition v_r = 0 at i = 0 in place of the left wall clamp.
__device__ __forceinline__ float2 grid_node_axisym(float px, float py, float m, float srr,
                                                   float shoop, int i, int j,
                                                   const PhysConsts& k) {
  if (!(m > 0.0f)) return make_float2(px, py);
  const float rm = __builtin_amdgcn_rcpf(m);
  float g0 = px * rm, g1 = py * rm;
  const float damp = k.damp[i] * k.damp[k.nx + j];
  g0 = g0 * damp;
  g1 = g1 * damp;
  const float r_node = ((float)i + 0.5f) * k.dx;
  if (r_node > k.dx) {
    const float hoop = (srr - shoop) / m;
    g0 = g0 + (k.dt * hoop) / r_node;
  }
  if (i == 0) g0 = 0.0f;
  if (i > k.nx - 3 && g0 > 0.0f) g0 = 0.0f;
  if (j < 3 && g1 < 0.0f) g1 = 0.0f;
  if (j > k.ny - 3 && g1 > 0.0f) g1 = 0.0f;
  return make_float2(g0, g1);
}

// Gravity, advection and the position clamp applied to every live particle after G2P
// (mpm_sim.py:329-333).
__device__ __forceinline__ void advect(float2& x, float2& v, const PhysConsts& k) {
  v.y = v.y - k.gravity_dt;
  x.x = x.x + k.dt * v.x;
  x.y = x.y + k.dt * v.y;
  x.x = fmaxf(fminf(x.x, k.x_hi), k.x_lo);
  x.y = fmaxf(fminf(x.y, k.y_hi), k.x_lo);
}

}  // namespace jm
