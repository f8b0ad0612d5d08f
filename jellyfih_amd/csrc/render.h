// Frame renderer of the reference (mpm_sim.py:343-712): particles splatted into an HDR frame
// buffer (additive, fp32 atomics), flat-colour passes, Reinhard tone mapping, and the grid
// vorticity overlays. Frame buffer: res x res x 3 fp32, row-major [py][px][c] like the
// reference's frame_buffer field. Instance m is drawn in tile (row, col) = (m / grid_side,
// m % grid_side) of res_sub pixels. Each kernel restates one reference kernel; the float
// arithmetic is fp32, as in the reference, with every product and sum rounded on its own (no
// fused multiply-add), so that disc coverage, pixel-to-cell maps and thresholds are the ones
// plain fp32 evaluation of the reference's expressions gives.
#pragma once

#include "kernels.h"

#pragma clang fp contract(off)

namespace jm {

// hsv2rgb (mpm_sim.py:345-359); i % 6 as Taichi's integer modulo of a non-negative i.
__device__ __forceinline__ float3 hsv2rgb(float h, float s, float v) {
  const int i = (int)(h * 6.0f);
  const float f = (h * 6.0f) - (float)i;
  const float p = v * (1.0f - s), q = v * (1.0f - f * s), t = v * (1.0f - (1.0f - f) * s);
  const int r6 = ((i % 6) + 6) % 6;
  switch (r6) {
    case 0: return make_float3(v, t, p);
    case 1: return make_float3(q, v, p);
    case 2: return make_float3(p, v, t);
    case 3: return make_float3(p, q, v);
    case 4: return make_float3(t, p, v);
    default: return make_float3(v, p, q);
  }
}

struct RenderArgs {
  State s;
  const int* n_live;
  int cap;
  int M;
  const float* freq;         // instance_freq
  const float* contraction;  // instance_act_contraction
  const float* hue;          // instance_hue
  const float* muscle_hue;   // instance_muscle_hue
  const float* t_buf;        // sim_time
  float* fb;
  int res, res_sub, grid_side;
  float radius, domain_height;
  int water_mode;            // water_angle_color: 0 ghost, 1 plain, 2 rainbow
  int mat_sel;               // render_flat_pass: target material
  float3 flat;               // render_flat_pass: colour
  PhysConsts k;
};

__global__ void k_fb_fill(float* fb, int64_t n, float v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    fb[i] = v;
}

// Tile-space centre of a particle (mpm_sim.py:463-464,500-501).
__device__ __forceinline__ float2 splat_centre(float2 pos, int m, const RenderArgs& a) {
  const int row = m / a.grid_side, col = m % a.grid_side;
  return make_float2((pos.x / a.domain_height + (float)col) * (float)a.res_sub,
                     ((a.domain_height - pos.y) / a.domain_height + (float)row) * (float)a.res_sub);
}

// Colour x intensity of one particle's glow (mpm_sim.py:390-459); false where the
// reference draws nothing (dead slot, x < 0, intensity <= 0.001).
__device__ __forceinline__ bool abyss_colour(const RenderArgs& a, int m, size_t i, float2& pos,
                                             float3& ci) {
  pos = a.s.x[i];
  const int mat = a.s.mat[i];
  if (mat < 0 || pos.x < 0.0f) return false;
  const PhysConsts& k = a.k;
  const float act = activation(a.t_buf[0], a.freq[m], a.contraction[m], k);
  const float2 vv = a.s.v[i];
  const float vel = sqrtf(vv.x * vv.x + vv.y * vv.y);
  const float stress = 1.0f - a.s.Jp[i];
  float3 color = make_float3(0.0f, 0.0f, 0.0f);
  float intensity = 0.0f;
  const float gnf = a.grid_side > 1 ? (float)(a.grid_side - 1) : 1.0f;
  const float norm_y = (float)(m / a.grid_side) / gnf;
  if (mat == 0) {
    if (vel > 0.02f) {
      if (a.water_mode == 0) {
        color = make_float3(0.15f, 0.15f, 0.18f);
        intensity = 0.008f;
      } else {
        const float brightness = fminf(vel / 2.0f, 1.0f);
        if (a.water_mode == 1) {
          color = hsv2rgb(0.60f, 0.55f, brightness);
        } else {
          const float ang = atan2f(vv.y, vv.x);
          const float h = ang / (2.0f * 3.14159265f) + 0.5f;
          color = hsv2rgb(h - floorf(h / 1.0f) * 1.0f, 0.8f, brightness);  // float % 1.0
        }
        intensity = 0.025f + 0.075f * brightness;
      }
    }
  } else if (mat == 1) {
    const float3 base = hsv2rgb(a.hue[m], 0.4f + (norm_y * 0.4f), 0.9f);
    const float glow = fabsf(stress) * 2.0f + 0.1f;
    color = make_float3(base.x + glow, base.y + glow, base.z + glow);
    intensity = 0.05f + (glow * 0.2f);
  } else if (mat == 3) {
    const float3 base = hsv2rgb(a.muscle_hue[m], 0.2f, 1.0f);
    color = make_float3(base.x + act, base.y + act, base.z + act);
    intensity = 0.4f + (act * 0.4f);
  } else if (mat == 2) {
    color = make_float3(1.0f, 0.2f, 0.0f);
    intensity = 0.8f;
  }
  if (!(intensity > 0.001f)) return false;
  ci = make_float3(color.x * intensity, color.y * intensity, color.z * intensity);
  return true;
}

// LDS accumulators of a workgroup's splat tile: up to 1365 pixels x 3 channels, 64-bit
// fixed point (32 KB).
constexpr int kSplatTile = 4096;

// render_frame_abyss (mpm_sim.py:385-477): one thread per particle slot (blockIdx.y =
// instance); colour and intensity by material, then an additive splat with a (1 - d^2)^4
// falloff over a disc of radius 1.5 * radius. The particles of a workgroup are consecutive
// in the engine's cell-sorted order, so their discs cover a small pixel box: when the box
// fits kSplatTile, the workgroup adds its splats into an LDS copy of the box, as 64-bit
// fixed point with 40 fractional bits (integer LDS atomics run at full rate on gfx950; with
// fp32 ones the kernel took 3.5x as long), and then adds the box's non-zero pixels to the frame
// with one fp32 global atomic per channel, instead of one per disc pixel (sums in another
// order than the reference's atomics, rounded to 2^-40 in LDS). The box bounds are wave
// minima combined by one LDS atomic per wave.
__global__ void __launch_bounds__(256) k_render_abyss(RenderArgs a) {
  __shared__ unsigned long long tile[kSplatTile];
  __shared__ int box[4];  // min px, -max px, min py, -max py of the workgroup's discs
  const int m = blockIdx.y;
  const int local = blockIdx.x * blockDim.x + threadIdx.x;
  float2 pos = make_float2(0.0f, 0.0f);
  float3 ci = make_float3(0.0f, 0.0f, 0.0f);
  bool draw = m < a.M && local < a.n_live[m] &&
              abyss_colour(a, m, (size_t)m * a.cap + local, pos, ci);
  float2 c = make_float2(0.0f, 0.0f);
  const float draw_r = a.radius * 1.5f;
  int lx = 0, hx = -1, ly = 0, hy = -1;
  if (draw) {
    c = splat_centre(pos, m, a);
    // the reference's loops over [int(c - r), int(c + r)] skip pixels outside the frame;
    // iterate over the intersection only
    lx = max((int)(c.x - draw_r), 0);
    hx = min((int)(c.x + draw_r), a.res - 1);
    ly = max((int)(c.y - draw_r), 0);
    hy = min((int)(c.y + draw_r), a.res - 1);
    draw = lx <= hx && ly <= hy;
  }
  if (threadIdx.x < 4) box[threadIdx.x] = INT_MAX;
  __syncthreads();
  {
    const int w0 = wave_min(draw ? lx : INT_MAX), w1 = wave_min(draw ? -hx : INT_MAX);
    const int w2 = wave_min(draw ? ly : INT_MAX), w3 = wave_min(draw ? -hy : INT_MAX);
    if ((threadIdx.x & 63) == 0 && w0 != INT_MAX) {
      atomicMin(&box[0], w0);
      atomicMin(&box[1], w1);
      atomicMin(&box[2], w2);
      atomicMin(&box[3], w3);
    }
  }
  __syncthreads();
  const int bx0 = box[0], by0 = box[2];
  if (bx0 == INT_MAX) return;  // no disc in this workgroup (uniform)
  const int bw = -box[1] - bx0 + 1, bh = -box[3] - by0 + 1;
  const bool in_lds = (int64_t)bw * bh * 3 <= kSplatTile;  // uniform
  if (in_lds) {
    for (int q = threadIdx.x; q < bw * bh * 3; q += blockDim.x) tile[q] = 0ull;
    __syncthreads();
  }
  if (draw) {
    const float r2 = draw_r * draw_r;
    // the fixed-point tile takes contributions below 2^11 (falloff <= 1); a larger or
    // non-finite colour (a diverged particle) adds to the frame directly (NaN fails the
    // comparison)
    const bool to_lds = in_lds && fabsf(ci.x) + fabsf(ci.y) + fabsf(ci.z) < 2048.0f;
    for (int px = lx; px <= hx; ++px)
      for (int py = ly; py <= hy; ++py) {
        const float dxp = (float)px - c.x, dyp = (float)py - c.y;
        const float dn = (dxp * dxp + dyp * dyp) / r2;
        if (dn < 1.0f) {
          const float u = 1.0f - dn, u2 = u * u, falloff = u2 * u2;
          if (to_lds) {
            unsigned long long* t = tile + ((py - by0) * bw + (px - bx0)) * 3;
            atomicAdd(t + 0, fixed_bits(ci.x * falloff * kFix40Pre));
            atomicAdd(t + 1, fixed_bits(ci.y * falloff * kFix40Pre));
            atomicAdd(t + 2, fixed_bits(ci.z * falloff * kFix40Pre));
          } else {
            float* f = a.fb + ((size_t)py * a.res + px) * 3;
            unsafeAtomicAdd(f + 0, ci.x * falloff);
            unsafeAtomicAdd(f + 1, ci.y * falloff);
            unsafeAtomicAdd(f + 2, ci.z * falloff);
          }
        }
      }
  }
  if (!in_lds) return;
  __syncthreads();
  for (int q = threadIdx.x; q < bw * bh; q += blockDim.x) {
    const int py = by0 + q / bw, px = bx0 + q % bw;
    float* f = a.fb + ((size_t)py * a.res + px) * 3;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const float val = from_fixed40_bits(tile[3 * q + ch]);
      if (val != 0.0f) unsafeAtomicAdd(f + ch, val);
    }
  }
}

// render_flat_pass (mpm_sim.py:487-517): particles of one material set the pixels of their
// disc of radius `radius` to one colour (every writer stores the same value).
__global__ void __launch_bounds__(256) k_render_flat(RenderArgs a) {
  const int m = blockIdx.y;
  const int local = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.M || local >= a.n_live[m]) return;
  const size_t i = (size_t)m * a.cap + local;
  const float2 pos = a.s.x[i];
  if (a.s.mat[i] != a.mat_sel || pos.x < 0.0f) return;
  const float2 c = splat_centre(pos, m, a);
  const float draw_r = a.radius;
  const int lx = (int)(c.x - draw_r), hx = (int)(c.x + draw_r);
  const int ly = (int)(c.y - draw_r), hy = (int)(c.y + draw_r);
  for (int px = max(lx, 0); px <= min(hx, a.res - 1); ++px)
    for (int py = max(ly, 0); py <= min(hy, a.res - 1); ++py) {
      const float dxp = (float)px - c.x, dyp = (float)py - c.y;
      if (dxp * dxp + dyp * dyp <= draw_r * draw_r) {
        float* f = a.fb + ((size_t)py * a.res + px) * 3;
        f[0] = a.flat.x;
        f[1] = a.flat.y;
        f[2] = a.flat.z;
      }
    }
}

// tone_map_and_encode (mpm_sim.py:367-383): Reinhard with exposure 2.5, a 0.15 knee, then
// gamma 1/2.2, per value in place; four values per thread and load (the frame buffer is a
// hipMalloc allocation, so 16-byte aligned; launched with one float4 per thread), the
// n % 4 tail one at a time.
__device__ __forceinline__ float tone_map(float x) {
  const float e = x * 2.5f;
  float mapped = e / (e + 1.0f);
  mapped = mapped / (mapped + 0.15f);
  return powf(mapped, (float)(1.0 / 2.2));  // the reference's exponent: 1/2.2 rounded once
}

__global__ void k_tone_map(float* fb, int64_t n) {
  const int64_t n4 = n / 4;
  float4* fb4 = reinterpret_cast<float4*>(fb);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float4 q = fb4[i];
    fb4[i] = make_float4(tone_map(q.x), tone_map(q.y), tone_map(q.z), tone_map(q.w));
  }
  if (blockIdx.x == 0 && threadIdx.x < n - 4 * n4) fb[4 * n4 + threadIdx.x] =
      tone_map(fb[4 * n4 + threadIdx.x]);
}

// The callers' 8-bit encode of a frame, (np.clip(img, 0, 1) * 255).astype(np.uint8)
// (evolve.py:397, helpers/view_single.py:257): clip, one fp32 product by 255, truncation
// toward zero. NaN encodes as 0 (fmaxf returns the number; numpy's cast of NaN is
// unspecified and gives 0 on x86-64).
__device__ __forceinline__ unsigned char encode_u8(float v) {
  const float c = fminf(fmaxf(v, 0.0f), 1.0f);
  return (unsigned char)(unsigned)(c * 255.0f);
}

// Four values per thread: one float4 load, one 4-byte store (both buffers are hipMalloc
// allocations); the n % 4 tail one at a time.
__global__ void k_encode_u8(const float* fb, unsigned char* out, int64_t n) {
  const int64_t n4 = n / 4;
  const float4* fb4 = reinterpret_cast<const float4*>(fb);
  uchar4* out4 = reinterpret_cast<uchar4*>(out);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float4 q = fb4[i];
    out4[i] = make_uchar4(encode_u8(q.x), encode_u8(q.y), encode_u8(q.z), encode_u8(q.w));
  }
  if (blockIdx.x == 0 && threadIdx.x < n - 4 * n4)
    out[4 * n4 + threadIdx.x] = encode_u8(fb[4 * n4 + threadIdx.x]);
}

// Curl of the post-grid-op velocity at interior node (i, j) of instance m with mass, from
// the grid planes gp the last G2P gathered: momentum x, momentum y, mass (the grid op is
// applied per node), or with vel_planes velocities already (mpm_sim.py:528-533,575-579);
// returns false where the reference skips the node.
__device__ __forceinline__ bool node_curl(const float* gp, int G, int i, int j, int vel_planes,
                                          const PhysConsts& k, float& curl) {
  if (i < 1 || i >= k.nx - 1 || j < 1 || j >= k.ny - 1) return false;
  const int n = j * k.nx + i;
  if (!(gp[2 * G + n] > 0.0f)) return false;
  auto vel = [&](int ii, int jj) {
    const int q = jj * k.nx + ii;
    if (vel_planes) return make_float2(gp[q], gp[G + q]);
    return grid_node(gp[q], gp[G + q], gp[2 * G + q], ii, jj, k);
  };
  const float dvx_dy = (vel(i, j + 1).x - vel(i, j - 1).x) / (2.0f * k.dx);
  const float dvy_dx = (vel(i + 1, j).y - vel(i - 1, j).y) / (2.0f * k.dx);
  curl = dvx_dy - dvy_dx;
  return true;
}

// compute_vorticity_grid (mpm_sim.py:560-578): vort = tanh(curl * scale), 0 elsewhere.
__global__ void k_vorticity_grid(const float* acc, float* vort, int M, int G, float scale,
                                 int vel_planes, PhysConsts k) {
  const int m = blockIdx.y;
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M || node >= G) return;
  const int j = node / k.nx, i = node - j * k.nx;
  float curl = 0.0f, v = 0.0f;
  if (node_curl(acc + (size_t)m * 3 * G, G, i, j, vel_planes, k, curl)) v = tanhf(curl * scale);
  vort[(size_t)m * G + node] = v;
}

// render_vorticity_overlay (mpm_sim.py:520-558): additive warm/cool squares per node.
__global__ void k_vorticity_overlay(const float* acc, int M, int G, float scale, int vel_planes,
                                    RenderArgs a) {
  const int m = blockIdx.y;
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M || node >= G) return;
  const PhysConsts& k = a.k;
  const int j = node / k.nx, i = node - j * k.nx;
  float curl;
  if (!node_curl(acc + (size_t)m * 3 * G, G, i, j, vel_planes, k, curl)) return;
  const float vi = fminf(fabsf(curl) * scale, 1.0f);
  if (vi < 0.01f) return;
  const float3 col = curl > 0.0f ? make_float3(1.0f * vi, 0.4f * vi, 0.0f * vi)
                                 : make_float3(0.0f * vi, 0.4f * vi, 1.0f * vi);
  const int row = m / a.grid_side, c = m % a.grid_side;
  const float ng = (float)k.nx;
  const int cx = (int)(((float)i / ng / a.domain_height + (float)c) * (float)a.res_sub);
  const int cy =
      (int)(((a.domain_height - (float)j / ng) / a.domain_height + (float)row) * (float)a.res_sub);
  const int cpx = max(1, a.res_sub / k.nx);
  for (int px = max(cx - cpx / 2, 0); px <= min(cx + cpx / 2, a.res - 1); ++px)
    for (int py = max(cy - cpx / 2, 0); py <= min(cy + cpx / 2, a.res - 1); ++py) {
      float* f = a.fb + ((size_t)py * a.res + px) * 3;
      unsafeAtomicAdd(f + 0, col.x * 0.5f);
      unsafeAtomicAdd(f + 1, col.y * 0.5f);
      unsafeAtomicAdd(f + 2, col.z * 0.5f);
    }
}

// render_vorticity_rdbu / _hueshift / _white (mpm_sim.py:581-712): per pixel, the vorticity
// of the grid cell under it alpha-composited toward red (CCW) or blue (CW); `white` composites
// over white (the caller fills the buffer with 1.0 first).
enum VortMode { kVortRdbu = 1, kVortHueshift = 2, kVortWhite = 3 };

__global__ void k_vorticity_pixels(const float* vort, int M, int G, int mode, float amax,
                                   RenderArgs a) {
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y;
  if (px >= a.res || py >= a.res) return;
  const PhysConsts& k = a.k;
  const int col = px * a.grid_side / a.res, row = py * a.grid_side / a.res;
  const int m = row * a.grid_side + col;
  if (m >= M) return;
  const float sub_x = (float)(px - col * a.res_sub) / (float)a.res_sub;
  const float sub_y = 1.0f - (float)(py - row * a.res_sub) / (float)a.res_sub;
  const int gi = (int)(sub_x * (float)k.nx * a.domain_height);
  const int gj = (int)(sub_y * (float)k.ny);
  if (gi < 0 || gi >= k.nx || gj < 0 || gj >= k.ny) return;
  const float v = vort[(size_t)m * G + gj * k.nx + gi];
  const float mag = fabsf(v);
  if (mag < 0.02f) return;
  const float alpha = amax * mag;
  const float t = 0.5f + 0.5f * (v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f));
  float3 c;
  if (mode == kVortHueshift)
    c = make_float3(t * 0.90f + (1.0f - t) * 0.15f, t * 0.18f + (1.0f - t) * 0.35f,
                    t * 0.15f + (1.0f - t) * 0.90f);
  else
    c = make_float3(t * 0.84f + (1.0f - t) * 0.13f, t * 0.10f + (1.0f - t) * 0.31f,
                    t * 0.11f + (1.0f - t) * 0.80f);
  float* f = a.fb + ((size_t)py * a.res + px) * 3;
  if (mode == kVortWhite) {
    f[0] = (1.0f - alpha) + alpha * c.x;
    f[1] = (1.0f - alpha) + alpha * c.y;
    f[2] = (1.0f - alpha) + alpha * c.z;
  } else {
    f[0] = (1.0f - alpha) * f[0] + alpha * c.x;
    f[1] = (1.0f - alpha) * f[1] + alpha * c.y;
    f[2] = (1.0f - alpha) * f[2] + alpha * c.z;
  }
}

}  // namespace jm

#pragma clang fp contract(fast)
