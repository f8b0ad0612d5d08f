// jellymx engine: host side of the C ABI declared in include/jellymx.h.
//
// Owns the device state of a batch of instances, the per-instance segment layout and the
// chunk tables the fused kernel walks, launches the substep paths (capturing each block of
// substeps as a HIP graph), and converts between the engine's cell-sorted compact layout
// and the reference's (instance, slot) layout for field access.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/jellymx.h"
#include "kernels.h"
#include "diff_kernels.h"
#include "render.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(JM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define JM_TRY(expr)            \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != JM_OK) return rc_; \
  } while (0)

constexpr int kNumXcd = 8;
constexpr int kMaxGroups = 4;  // concurrent instance groups of the fused path

// One instance group's chunk table: its chunks are chunks_d[off, off + cnt), laid out for
// its own launch (blockIdx -> chunk, XCD-affine when the group has >= 8 instances).
struct GroupChunks {
  int off = 0, cnt = 0, grid_blocks = 0;
  bool affinity = false;
  int xcd_off[kNumXcd] = {0}, xcd_cnt[kNumXcd] = {0};
};

struct DevState {
  jm::State s;
  void* base = nullptr;
};

}  // namespace

struct jm_handle {
  jm_config cfg{};
  jm_tuning tun{};
  int M = 0, N = 0, nx = 0, ny = 0, G = 0, cap = 0;
  jm::PhysConsts k{};
  hipStream_t stream = nullptr;

  // device buffers
  DevState st, tmp;
  int* sortbuf = nullptr;  // key_a, val_a, key_b, val_b, rank: 5 * M * cap ints
  int* hist = nullptr;     // radix histogram: M * 256 * tmax ints
  int tmax = 0;            // radix tiles per instance segment
  int key_bits = 0;        // bits of the largest cell key
  float* acc = nullptr;       // 3 grid buffers of M*3*G floats
  float* t_buf = nullptr;     // 2 floats per instance group (clock ping-pong), group 0 first
  float* params = nullptr;    // 6 * M floats: amp, contraction, freq, pdens, mu_j, la_j
  int* n_live_d = nullptr;
  int4* chunks_d = nullptr;
  int* nchunks_d = nullptr;
  float* stats_d = nullptr;   // M * 6
  int* nonfinite_d = nullptr; // M: live particles with a non-finite v, C or F (nan_invalid)
  float* staging = nullptr;   // M * 2 * G floats (grid velocity readout)
  float* damp = nullptr;      // nx + ny floats: boundary damping factors (k.damp)
  float* frame = nullptr;     // renderer: frame_res^2 x 3 floats (allocated on first use)
  int frame_res = 0;
  unsigned char* frame_u8 = nullptr;  // its 8-bit encode (allocated on first jm_read_frame_u8)
  float* vort = nullptr;      // renderer: M * G floats, node (i, j) at j * nx + i
  float* rhue = nullptr;      // renderer: 2 * M floats, instance_hue then instance_muscle_hue
  std::vector<float> rhue_host;  // what rhue holds
  float* hoop = nullptr;      // axisymmetric variant: 3 buffers of [m][2][G] stress grids of
                              // the hoop term, one beside each grid buffer (plain: buffer 0)

  // host mirrors
  std::vector<float> h_param[5];  // amp, contraction, freq, pdens, E_jelly
  std::vector<int> n_live;
  std::vector<float> h_pos, h_fiber;  // loaded values (dead-slot contents), (M,N,2)
  std::vector<int> h_mat;             // (M,N)
  std::vector<char> inst_dirty;
  bool layout_dirty = true;
  bool params_dirty = true;

  // chunk tables
  int chunk_size = 0;
  int total_chunks = 0;
  int n_groups = 1;                 // instance groups of the current chunk tables
  GroupChunks grp[kMaxGroups];
  // Group g > 0 runs on gstream[g]; group 0 on `stream`, which forks to the others
  // (fork_ev) and joins them (gdone) around every fused block.
  hipStream_t gstream[kMaxGroups] = {nullptr};
  hipEvent_t gdone[kMaxGroups] = {nullptr};
  hipEvent_t fork_ev = nullptr;

  // rotation / bookkeeping
  int acc_base = 0;
  int last_read = -1;        // grid buffer the last G2P gathered from (-1: none yet)
  bool grid_is_velocity = false;  // plain path leaves velocity in the planes
  bool tuning_user = false;  // jm_set_tuning was called: keep its launch shape
  int64_t since_sort = 0;

  std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;

  ~jm_handle() { release(); }

  void drop_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }

  void release() {
    drop_graphs();
    if (stream) (void)hipStreamSynchronize(stream);
    void* ptrs[] = {st.base, tmp.base, sortbuf, hist, acc, t_buf, params, n_live_d,
                    chunks_d, nchunks_d, stats_d, staging, damp, frame, frame_u8, vort, rhue,
                    hoop, nonfinite_d};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    st.base = tmp.base = nullptr;
    sortbuf = hist = nullptr;
    acc = t_buf = params = stats_d = staging = damp = frame = vort = rhue = hoop = nullptr;
    frame_u8 = nullptr;
    frame_res = 0;
    rhue_host.clear();
    n_live_d = nullptr;
    nonfinite_d = nullptr;
    chunks_d = nullptr;
    nchunks_d = nullptr;
    for (int g = 1; g < kMaxGroups; ++g) {
      if (gdone[g]) (void)hipEventDestroy(gdone[g]);
      if (gstream[g]) (void)hipStreamDestroy(gstream[g]);
      gdone[g] = nullptr;
      gstream[g] = nullptr;
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    fork_ev = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};

namespace {

jm::PhysConsts make_consts(const jm_config& c, int nx, int ny) {
  // mpm_sim.py:12-58, evaluated in double like the reference's Python, then rounded.
  const double n_grid = 128.0 * c.quality;
  const double dx = 1.0 / n_grid, inv_dx = n_grid;
  const double dt = 2e-5 / c.quality;
  const double p_vol = (dx * 0.5) * (dx * 0.5);
  const double p_mass = p_vol * 1;
  const double E_base = 0.1e4, nu_base = 0.2, E_pay = 2.0e5, nu_pay = 0.2;
  jm::PhysConsts k{};
  k.nx = nx;
  k.ny = ny;
  k.damp_x = nx / 20;
  k.damp_y = ny / 20;
  k.dt = (float)dt;
  k.dx = (float)dx;
  k.inv_dx = (float)inv_dx;
  k.p_mass = (float)p_mass;
  k.gravity_dt = (float)(dt * 10.0);
  k.stress_scale = (float)(-dt * p_vol * 4 * inv_dx * inv_dx);
  k.apic_scale = (float)(4 * inv_dx);
  k.mu_base = (float)(E_base / (2 * (1 + nu_base)));
  k.la_base = (float)(E_base * nu_base / ((1 + nu_base) * (1 - 2 * nu_base)));
  k.mu_pay = (float)(E_pay / (2 * (1 + nu_pay)));
  k.la_pay = (float)(E_pay * nu_pay / ((1 + nu_pay) * (1 - 2 * nu_pay)));
  k.water_la = 100000.0f;
  k.act_freq = 1.0f;
  k.x_lo = 0.001f;
  k.x_hi = 0.999f;
  k.y_hi = (float)(c.domain_height - 0.001);
  k.sig_floor = 1e-6f;
  k.pay_lo = (float)(1 - 5e-3);
  k.pay_hi = (float)(1 + 5e-3);
  k.pi = 3.14159265f;
  k.r_ref_floor = (float)(dx * 0.1);  // mpm_sim.py:246
  k.r_floor = (float)(dx * 0.05);     // mpm_sim.py:247
  return k;
}

template <typename T>
T* carve(char*& p, size_t n) {
  size_t bytes = (n * sizeof(T) + 255) & ~size_t(255);
  T* r = reinterpret_cast<T*>(p);
  p += bytes;
  return r;
}

size_t state_bytes(size_t n, bool glow, bool axisym) {
  auto rnd = [](size_t b) { return (b + 255) & ~size_t(255); };
  size_t b = rnd(n * 8) * 3 + rnd(n * 16) * 2 + rnd(n * 4) * 3;  // x v fiber; C F; Jp mat slot
  if (glow) b += rnd(n * 4);
  if (axisym) b += rnd(n * 4);
  return b;
}

int alloc_state(DevState& d, size_t n, bool glow, bool axisym) {
  void* base = nullptr;
  size_t bytes = state_bytes(n, glow, axisym);
  if (hipMalloc(&base, bytes) != hipSuccess) return fail(JM_ERR_NOMEM, "hipMalloc(state)");
  d.base = base;
  char* p = static_cast<char*>(base);
  d.s.x = carve<float2>(p, n);
  d.s.v = carve<float2>(p, n);
  d.s.fiber = carve<float2>(p, n);
  d.s.C = carve<float4>(p, n);
  d.s.F = carve<float4>(p, n);
  d.s.Jp = carve<float>(p, n);
  d.s.mat = carve<int>(p, n);
  d.s.slot = carve<int>(p, n);
  d.s.glow = glow ? carve<float>(p, n) : nullptr;
  d.s.r_ref = axisym ? carve<float>(p, n) : nullptr;
  return JM_OK;
}

inline unsigned blocks_for(int64_t n, int threads) {
  return (unsigned)std::max<int64_t>(1, (n + threads - 1) / threads);
}

// Largest live count over the instances (after commit): the per-particle launches cover
// [0, max_live) of every instance's slots.
int max_live(const jm_handle* h) {
  int n = 0;
  for (int m = 0; m < h->M; ++m) n = std::max(n, h->n_live[m]);
  return n;
}

jm::InstParams inst_params(const jm_handle* h) {
  const int M = h->M;
  return {h->params, h->params + M, h->params + 2 * M, h->params + 3 * M, h->params + 4 * M,
          h->params + 5 * M};
}

int upload_params(jm_handle* h) {
  if (!h->params_dirty) return JM_OK;
  const int M = h->M;
  std::vector<float> p(6 * M);
  for (int m = 0; m < M; ++m) {
    p[m] = h->h_param[JM_PARAM_ACTUATION][m];
    p[M + m] = h->h_param[JM_PARAM_CONTRACTION][m];
    p[2 * M + m] = h->h_param[JM_PARAM_FREQ][m];
    p[3 * M + m] = h->h_param[JM_PARAM_PAYLOAD_DENSITY][m];
    // jelly/muscle Lame parameters from E (nu = 0.3), mpm_sim.py:31-34, in double
    const double E = h->h_param[JM_PARAM_E_JELLY][m], nu = 0.3;
    p[4 * M + m] = (float)(E / (2 * (1 + nu)));
    p[5 * M + m] = (float)(E * nu / ((1 + nu) * (1 - 2 * nu)));
  }
  HIP_TRY(hipMemcpyAsync(h->params, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->params_dirty = false;
  return JM_OK;
}

// Assign chunks of `chunk_size` sorted particles to workgroups. Instances are split into
// `instance_groups` groups balanced by chunk count (longest first); each group is one
// launch per substep, and the groups' launches run concurrently. Within a group, with XCD
// affinity every chunk of one instance runs on the same XCD (workgroup i is dispatched to
// XCD i % 8), so its grid's atomics and gathers stay in that XCD's L2; the group's
// instances are spread over XCDs longest-first to balance the chunk counts.
int build_chunks(jm_handle* h) {
  const int M = h->M;
  if (!h->tuning_user) {
    // Default launch shape by batch size: with fewer than 512 chunks of 1024 live particles
    // (two per CU), 256 x 2 chunks give the small batch twice the workgroups (one tank:
    // 16.1 -> 10.8 us per substep; four: 17.2 -> 12.8 us); larger batches keep 256 x 4
    // (the autotuner's choice at 32 x 80K).
    int64_t chunks = 0;
    for (int m = 0; m < M; ++m) chunks += (h->n_live[m] + 1023) / 1024;
    h->tun.threads = 256;
    h->tun.ppt = chunks < 512 ? 2 : 4;
  }
  const int cs = h->tun.threads * h->tun.ppt;
  h->chunk_size = cs;
  std::vector<int> nch(M);
  for (int m = 0; m < M; ++m) nch[m] = (h->n_live[m] + cs - 1) / cs;
  std::vector<int> order;  // instances with live particles, longest first
  for (int m = 0; m < M; ++m)
    if (nch[m] > 0) order.push_back(m);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nch[a] > nch[b]; });
  const int ng = std::max(1, std::min<int>(h->tun.instance_groups, (int)order.size()));
  std::vector<std::vector<int>> members(ng);
  {
    std::vector<int64_t> load(ng, 0);
    for (int m : order) {
      const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[g] += nch[m];
      members[g].push_back(m);  // longest first within the group too
    }
  }
  auto chunks_of = [&](int m) {
    std::vector<int4> mine;
    for (int c = 0; c < nch[m]; ++c) {
      const int start = m * h->cap + c * cs;
      const int count = std::min(cs, h->n_live[m] - c * cs);
      mine.push_back(make_int4(m, start, count, c));
    }
    return mine;
  };
  std::vector<int4> all;
  h->n_groups = ng;
  for (int g = 0; g < kMaxGroups; ++g) h->grp[g] = GroupChunks{};
  for (int g = 0; g < ng; ++g) {
    GroupChunks& gc = h->grp[g];
    gc.off = (int)all.size();
    gc.affinity = h->tun.xcd_affinity && (int)members[g].size() >= kNumXcd;
    if (gc.affinity) {
      std::vector<std::vector<int4>> per_xcd(kNumXcd);
      std::vector<int64_t> load(kNumXcd, 0);
      for (int m : members[g]) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += nch[m];
        std::vector<int4> mine = chunks_of(m);
        per_xcd[x].insert(per_xcd[x].end(), mine.begin(), mine.end());
      }
      int mx = 0;
      for (int x = 0; x < kNumXcd; ++x) {
        gc.xcd_off[x] = (int)all.size() - gc.off;
        gc.xcd_cnt[x] = (int)per_xcd[x].size();
        mx = std::max(mx, gc.xcd_cnt[x]);
        all.insert(all.end(), per_xcd[x].begin(), per_xcd[x].end());
      }
      gc.grid_blocks = kNumXcd * mx;
    } else {
      std::vector<int> ms = members[g];
      std::sort(ms.begin(), ms.end());  // instance order for the non-affine mapping
      for (int m : ms) {
        std::vector<int4> mine = chunks_of(m);
        all.insert(all.end(), mine.begin(), mine.end());
      }
      gc.grid_blocks = (int)all.size() - gc.off;
    }
    gc.cnt = (int)all.size() - gc.off;
  }
  h->total_chunks = (int)all.size();
  // the old table may still be read by queued kernels of every group (joined into the
  // main stream)
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->chunks_d) (void)hipFree(h->chunks_d);
  h->chunks_d = nullptr;
  if (!all.empty()) {
    HIP_TRY(hipMalloc(&h->chunks_d, all.size() * sizeof(int4)));
    HIP_TRY(hipMemcpyAsync(h->chunks_d, all.data(), all.size() * sizeof(int4),
                           hipMemcpyHostToDevice, h->stream));
  }
  HIP_TRY(hipMemcpyAsync(h->nchunks_d, nch.data(), M * sizeof(int), hipMemcpyHostToDevice,
                         h->stream));
  // Every group runs every block, so the groups' clocks stay equal from here on; a new
  // grouping starts them all from group 0's clock.
  for (int g = 1; g < ng; ++g)
    HIP_TRY(hipMemcpyAsync(h->t_buf + 2 * g, h->t_buf, 2 * sizeof(float),
                           hipMemcpyDeviceToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return JM_OK;
}

// Upload instances loaded since the last commit, rebuild chunk tables, clear the grids.
int commit(jm_handle* h) {
  JM_TRY(upload_params(h));
  if (!h->layout_dirty) return JM_OK;
  const int M = h->M, N = h->N;
  std::vector<float2> px;
  std::vector<float2> pf;
  std::vector<int> pm, ps;
  for (int m = 0; m < M; ++m) {
    if (!h->inst_dirty[m]) continue;
    px.clear();
    pf.clear();
    pm.clear();
    ps.clear();
    for (int p = 0; p < N; ++p) {
      size_t q = (size_t)m * N + p;
      if (h->h_mat[q] < 0) continue;
      px.push_back(make_float2(h->h_pos[2 * q], h->h_pos[2 * q + 1]));
      pf.push_back(make_float2(h->h_fiber[2 * q], h->h_fiber[2 * q + 1]));
      pm.push_back(h->h_mat[q]);
      ps.push_back(p);
    }
    const int n = (int)px.size();
    h->n_live[m] = n;
    const size_t off = (size_t)m * h->cap;
    if (n > 0) {
      HIP_TRY(hipMemcpyAsync(h->st.s.x + off, px.data(), n * sizeof(float2),
                             hipMemcpyHostToDevice, h->stream));
      HIP_TRY(hipMemcpyAsync(h->st.s.fiber + off, pf.data(), n * sizeof(float2),
                             hipMemcpyHostToDevice, h->stream));
      HIP_TRY(hipMemcpyAsync(h->st.s.mat + off, pm.data(), n * sizeof(int),
                             hipMemcpyHostToDevice, h->stream));
      HIP_TRY(hipMemcpyAsync(h->st.s.slot + off, ps.data(), n * sizeof(int),
                             hipMemcpyHostToDevice, h->stream));
      hipLaunchKernelGGL(jm::k_init_particles, dim3(blocks_for(n, 256)), dim3(256), 0,
                         h->stream, h->st.s, (int64_t)off, (int64_t)(off + n));
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(h->stream));  // host vectors are reused next iteration
    h->inst_dirty[m] = 0;
  }
  HIP_TRY(hipMemcpyAsync(h->n_live_d, h->n_live.data(), M * sizeof(int), hipMemcpyHostToDevice,
                         h->stream));
  JM_TRY(build_chunks(h));
  HIP_TRY(hipMemsetAsync(h->acc, 0, (size_t)3 * M * 3 * h->G * sizeof(float), h->stream));
  if (h->hoop)
    HIP_TRY(hipMemsetAsync(h->hoop, 0, (size_t)3 * M * 2 * h->G * sizeof(float), h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->acc_base = 0;
  h->last_read = -1;
  h->grid_is_velocity = false;
  h->since_sort = h->tun.sort_interval;  // sort before the next fused block
  h->drop_graphs();
  h->layout_dirty = false;
  return JM_OK;
}

// Radix bits (a multiple of 8) that cover every sort key of cell_key: row-major keys are
// below nx * ny, strip keys below (strips) * nx * strip.
int sort_key_bits(int nx, int ny, int strip) {
  const int64_t rows = strip > 0 ? (int64_t)(ny + strip - 1) / strip * strip : ny;
  int bits = 8;
  while ((1LL << bits) < rows * nx) bits += 8;
  return bits;
}

jm::SortArgs sort_args(jm_handle* h) {
  jm::SortArgs a;
  const size_t n = (size_t)h->M * h->cap;
  a.s = h->st.s;
  a.tmp = h->tmp.s;
  a.key_a = h->sortbuf;
  a.val_a = h->sortbuf + n;
  a.key_b = h->sortbuf + 2 * n;
  a.val_b = h->sortbuf + 3 * n;
  a.rank = h->sortbuf + 4 * n;
  a.hist = h->hist;
  a.n_live = h->n_live_d;
  a.cap = h->cap;
  a.M = h->M;
  a.tmax = h->tmax;
  a.strip = h->tun.sort_strip;
  a.k = h->k;
  return a;
}

// Stable radix sort of every instance's particles by cell key (8 bits per pass), then the
// gather of all fields into sorted order.
int enqueue_sort(jm_handle* h) {
  jm::SortArgs a = sort_args(h);
  const dim3 pg(blocks_for(h->cap, 256), h->M);
  hipLaunchKernelGGL(jm::k_sort_keys, pg, dim3(256), 0, h->stream, a);
  int* kin = a.key_a;
  int* vin = a.val_a;
  int* kout = a.key_b;
  int* vout = a.val_b;
  for (int shift = 0; shift < h->key_bits; shift += 8) {
    hipLaunchKernelGGL(jm::k_sort_rank, dim3(h->tmax, h->M), dim3(jm::kSortTile), 0, h->stream,
                       a, (const int*)kin, shift);
    hipLaunchKernelGGL(jm::k_sort_scan, dim3(h->M), dim3(1024), 0, h->stream, a);
    hipLaunchKernelGGL(jm::k_sort_scatter, pg, dim3(256), 0, h->stream, a, (const int*)kin,
                       (const int*)vin, kout, vout, shift);
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  hipLaunchKernelGGL(jm::k_sort_gather, pg, dim3(256), 0, h->stream, a, (const int*)vin);
  HIP_TRY(hipGetLastError());
  return JM_OK;
}

// Kernel arguments of instance group g: its chunk table and its own clock pair.
jm::FusedArgs fused_args(jm_handle* h, int g) {
  const GroupChunks& gc = h->grp[g];
  jm::FusedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.s = h->st.s;
  a.ip = inst_params(h);
  a.t_buf = h->t_buf + 2 * g;
  a.chunks = h->chunks_d + gc.off;
  a.nchunks = h->nchunks_d;
  for (int x = 0; x < kNumXcd; ++x) {
    a.xcd_off[x] = gc.xcd_off[x];
    a.xcd_cnt[x] = gc.xcd_cnt[x];
  }
  a.affinity = gc.affinity ? 1 : 0;
  a.total_chunks = gc.cnt;
  a.G = h->G;
  a.lazy_vc = h->tun.lazy_vc;
  a.glow = h->cfg.track_glow;
  a.k = h->k;
  return a;
}

hipStream_t group_stream(jm_handle* h, int g) { return g == 0 ? h->stream : h->gstream[g]; }

template <int B, int P, bool AXI = false>
void launch_fused_t(jm_handle* h, int g, bool g2p, bool p2g, const jm::FusedArgs& a) {
  dim3 grid(std::max(1, h->grp[g].grid_blocks)), block(B);
  hipStream_t st = group_stream(h, g);
  if (g2p && p2g)
    hipLaunchKernelGGL((jm::k_fused<B, P, true, true, AXI>), grid, block, 0, st, a);
  else if (g2p)
    hipLaunchKernelGGL((jm::k_fused<B, P, true, false, AXI>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((jm::k_fused<B, P, false, true, AXI>), grid, block, 0, st, a);
}

// Launch shapes of the fused kernel; the axisymmetric variant is built for 256 x 2 and
// 256 x 4 only.
bool tuning_supported(int threads, int ppt, bool axisym) {
  if (axisym) return threads == 256 && (ppt == 2 || ppt == 4);
  static const int combos[][2] = {{128, 8}, {128, 16}, {256, 2}, {256, 4}, {256, 8},
                                  {256, 16}, {512, 2}, {512, 4}, {512, 8}};
  for (auto& c : combos)
    if (c[0] == threads && c[1] == ppt) return true;
  return false;
}

void launch_fused(jm_handle* h, int g, bool g2p, bool p2g, const jm::FusedArgs& a) {
  const int T = h->tun.threads, P = h->tun.ppt;
  if (h->cfg.axisym) {
    if (P == 2) return launch_fused_t<256, 2, true>(h, g, g2p, p2g, a);
    return launch_fused_t<256, 4, true>(h, g, g2p, p2g, a);
  }
#define JM_CASE(TT, PP) \
  if (T == TT && P == PP) return launch_fused_t<TT, PP>(h, g, g2p, p2g, a);
  JM_CASE(128, 8)
  JM_CASE(128, 16)
  JM_CASE(256, 2)
  JM_CASE(256, 4)
  JM_CASE(256, 8)
  JM_CASE(256, 16)
  JM_CASE(512, 2)
  JM_CASE(512, 4)
  JM_CASE(512, 8)
#undef JM_CASE
}

float* grid_buf(jm_handle* h, int b) { return h->acc + (size_t)b * h->M * 3 * h->G; }

// The stress planes beside grid buffer b (axisymmetric variant), else null.
float* hoop_buf(jm_handle* h, int b) {
  return h->hoop ? h->hoop + (size_t)b * h->M * 2 * h->G : nullptr;
}

// Kernel kk (0..n) of a block of n fused substeps with grid-buffer base b, for instance
// group g: it gathers from buffer (b+kk-1)%3, scatters into (b+kk)%3 and zeroes (b+kk+1)%3
// — the next kernel's scatter target, last read two kernels ago (each workgroup touches only
// its own instance's planes, so the groups never share memory). Kernel 0 only scatters
// (P2G of substep 0), kernel n only gathers (G2P of substep n-1). The group's clock
// ping-pongs between its t_buf[0] and t_buf[1] and is back in t_buf[0] after the block;
// every group advances its clock by the same fp32 additions, so the clocks stay equal.
void enqueue_fused_kernel(jm_handle* h, int g, jm::FusedArgs& a, int kk, int n, int b) {
  const bool g2p = kk > 0, p2g = kk < n;
  a.acc_r = g2p ? grid_buf(h, (b + kk - 1) % 3) : nullptr;
  a.acc_w = p2g ? grid_buf(h, (b + kk) % 3) : nullptr;
  a.acc_z = grid_buf(h, (b + kk + 1) % 3);
  a.hoop_r = g2p ? hoop_buf(h, (b + kk - 1) % 3) : nullptr;
  a.hoop_w = p2g ? hoop_buf(h, (b + kk) % 3) : nullptr;
  a.hoop_z = hoop_buf(h, (b + kk + 1) % 3);
  if (p2g) {  // substep kk reads t_kk, writes t_{kk+1}
    a.t_rd = kk % 2;
    a.t_wr = (kk + 1) % 2;
    a.t_adv = 1;
  } else if (n % 2 == 1) {
    a.t_rd = 1;
    a.t_wr = 0;
    a.t_adv = 0;
  } else {
    a.t_rd = 0;
    a.t_wr = -1;
    a.t_adv = 0;
  }
  launch_fused(h, g, g2p, p2g, a);
}

// Kernels kk0..kk1 of a block for group g, on the group's stream.
int enqueue_group_kernels(jm_handle* h, int g, int kk0, int kk1, int n, int b) {
  if (h->grp[g].cnt == 0) return JM_OK;
  jm::FusedArgs a = fused_args(h, g);
  for (int kk = kk0; kk <= kk1; ++kk) enqueue_fused_kernel(h, g, a, kk, n, b);
  HIP_TRY(hipGetLastError());
  return JM_OK;
}

// Groups g > 0 start after the work already on the main stream.
int fork_groups(jm_handle* h) {
  if (h->n_groups < 2) return JM_OK;
  HIP_TRY(hipEventRecord(h->fork_ev, h->stream));
  for (int g = 1; g < h->n_groups; ++g) HIP_TRY(hipStreamWaitEvent(h->gstream[g], h->fork_ev, 0));
  return JM_OK;
}

// The main stream continues after every group's work.
int join_groups(jm_handle* h) {
  for (int g = 1; g < h->n_groups; ++g) {
    HIP_TRY(hipEventRecord(h->gdone[g], h->gstream[g]));
    HIP_TRY(hipStreamWaitEvent(h->stream, h->gdone[g], 0));
  }
  return JM_OK;
}

void finish_block(jm_handle* h, int n, int b, bool sort) {
  if (sort) h->since_sort = 0;
  h->since_sort += n;
  h->last_read = (b + n - 1) % 3;
  h->acc_base = (b + n + 1) % 3;
  h->grid_is_velocity = false;
}

// Axisymmetric fused path, after its blocks: the grid the last G2P gathered from becomes
// velocities in place (k_axisym_velocity), so grid reads and the renderer see what the
// plain path leaves.
int finish_axisym(jm_handle* h) {
  if (!h->hoop || h->cfg.path != JM_PATH_FUSED || h->last_read < 0 || h->grid_is_velocity)
    return JM_OK;
  hipLaunchKernelGGL(jm::k_axisym_velocity, dim3(blocks_for(h->G, 256), h->M), dim3(256), 0,
                     h->stream, grid_buf(h, h->last_read), hoop_buf(h, h->last_read), h->M,
                     h->G, h->k);
  HIP_TRY(hipGetLastError());
  h->grid_is_velocity = true;
  return JM_OK;
}

// Captured graph of group g's n + 1 kernels of a block with grid-buffer base b.
int group_graph(jm_handle* h, int g, int n, int b, hipGraphExec_t* out) {
  auto key = std::make_tuple(n, b, g);
  auto it = h->graphs.find(key);
  if (it == h->graphs.end()) {
    hipStream_t st = group_stream(h, g);
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_group_kernels(h, g, 0, n, n, b);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(st, &gr);
    if (rc != JM_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess)
      return fail(JM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess)
      return fail(JM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    it = h->graphs.emplace(key, ex).first;
  }
  *out = it->second;
  return JM_OK;
}

// Blocks of up to block_substeps substeps: the sort (every sort_interval substeps) on the
// main stream, then each instance group's n + 1 kernels on its own stream, concurrently,
// joined back into the main stream. With graphs, each group's kernels of a block are one
// captured graph per (n, grid-buffer base, group), captured before the block is enqueued.
int run_fused(jm_handle* h, int64_t steps) {
  if (h->total_chunks == 0) {
    // no live particle: only the clock moves, as the reference's substep() still advances it
    hipLaunchKernelGGL(jm::k_time_advance_n, dim3(1), dim3(64), 0, h->stream, h->t_buf,
                       h->k.dt, steps);
    HIP_TRY(hipGetLastError());
    return JM_OK;
  }
  int64_t left = steps;
  while (left > 0) {
    const int n = (int)std::min<int64_t>(left, h->tun.block_substeps);
    const bool sort = h->since_sort >= h->tun.sort_interval;
    const int b = h->acc_base;
    hipGraphExec_t ex[kMaxGroups] = {nullptr};
    if (h->tun.use_graphs && h->total_chunks > 0)
      for (int g = 0; g < h->n_groups; ++g)
        if (h->grp[g].cnt > 0) JM_TRY(group_graph(h, g, n, b, &ex[g]));
    if (sort) JM_TRY(enqueue_sort(h));
    if (h->total_chunks > 0) {
      JM_TRY(fork_groups(h));
      for (int g = 0; g < h->n_groups; ++g) {
        if (h->grp[g].cnt == 0) continue;
        if (ex[g]) HIP_TRY(hipGraphLaunch(ex[g], group_stream(h, g)));
      }
      if (!h->tun.use_graphs)  // interleaved, as in jm_timed_block
        for (int kk = 0; kk <= n; ++kk)
          for (int g = 0; g < h->n_groups; ++g) JM_TRY(enqueue_group_kernels(h, g, kk, kk, n, b));
      JM_TRY(join_groups(h));
    }
    finish_block(h, n, b, sort);
    left -= n;
  }
  return finish_axisym(h);
}

jm::PlainArgs plain_args(jm_handle* h, int m0) {
  jm::PlainArgs a;
  a.s = h->st.s;
  a.ip = inst_params(h);
  a.acc = grid_buf(h, 0);
  a.hoop = h->hoop;
  a.t_buf = h->t_buf;
  a.n_live = h->n_live_d;
  a.cap = h->cap;
  a.m0 = m0;
  a.G = h->G;
  a.glow = h->cfg.track_glow;
  a.k = h->k;
  return a;
}

// Plain path: per substep, for each instance separately: reset its grid, P2G, grid op, G2P;
// then the time advance. The axisymmetric variant, which runs only here, launches each
// stage once for the whole batch instead (blockIdx.y = instance).
int run_plain(jm_handle* h, int64_t steps) {
  const int M = h->M, G = h->G;
  if (h->hoop) {
    jm::PlainArgs a = plain_args(h, 0);
    const dim3 pg(blocks_for(max_live(h), 256), M), gg(blocks_for(G, 256), M);
    for (int64_t s = 0; s < steps; ++s) {
      hipLaunchKernelGGL(jm::k_fill_f32, dim3(blocks_for((int64_t)3 * M * G, 256 * 4)),
                         dim3(256), 0, h->stream, grid_buf(h, 0), (int64_t)3 * M * G, 0.f);
      hipLaunchKernelGGL(jm::k_fill_f32, dim3(blocks_for((int64_t)2 * M * G, 256 * 4)),
                         dim3(256), 0, h->stream, h->hoop, (int64_t)2 * M * G, 0.f);
      hipLaunchKernelGGL(jm::k_axisym_p2g, pg, dim3(256), 0, h->stream, a);
      hipLaunchKernelGGL(jm::k_plain_gridop, gg, dim3(256), 0, h->stream, a);
      hipLaunchKernelGGL(jm::k_plain_g2p, pg, dim3(256), 0, h->stream, a);
      hipLaunchKernelGGL(jm::k_time_advance, dim3(1), dim3(64), 0, h->stream, h->t_buf,
                         h->k.dt);
    }
    HIP_TRY(hipGetLastError());
    h->last_read = 0;
    h->grid_is_velocity = true;
    return JM_OK;
  }
  for (int64_t s = 0; s < steps; ++s) {
    for (int m = 0; m < M; ++m) {
      if (h->n_live[m] == 0) continue;
      jm::PlainArgs a = plain_args(h, m);
      hipLaunchKernelGGL(jm::k_fill_f32, dim3(blocks_for(3 * G, 256 * 4)), dim3(256), 0,
                         h->stream, grid_buf(h, 0) + (size_t)m * 3 * G, (int64_t)3 * G, 0.f);
      const dim3 pg(blocks_for(h->n_live[m], 256), 1);
      hipLaunchKernelGGL(jm::k_plain_p2g, pg, dim3(256), 0, h->stream, a);
      hipLaunchKernelGGL(jm::k_plain_gridop, dim3(blocks_for(G, 256), 1), dim3(256), 0,
                         h->stream, a);
      hipLaunchKernelGGL(jm::k_plain_g2p, pg, dim3(256), 0, h->stream, a);
    }
    hipLaunchKernelGGL(jm::k_time_advance, dim3(1), dim3(64), 0, h->stream, h->t_buf, h->k.dt);
  }
  HIP_TRY(hipGetLastError());
  h->last_read = 0;
  h->grid_is_velocity = true;
  return JM_OK;
}

bool valid_field(int f) { return f >= JM_FIELD_X && f <= JM_FIELD_FITNESS; }

int64_t field_bytes(const jm_handle* h, int f) {
  const int64_t MN = (int64_t)h->M * h->N, MG = (int64_t)h->M * h->G;
  switch (f) {
    case JM_FIELD_X:
    case JM_FIELD_V:
    case JM_FIELD_FIBER:
      return MN * 8;
    case JM_FIELD_C:
    case JM_FIELD_F:
      return MN * 16;
    case JM_FIELD_MATERIAL:
    case JM_FIELD_JP:
    case JM_FIELD_GLOW:
      return MN * 4;
    case JM_FIELD_GRID_V:
      return MG * 8;
    case JM_FIELD_GRID_M:
      return MG * 4;
    case JM_FIELD_FITNESS:
      return (int64_t)h->M * 12;
  }
  return -1;
}

// Per-particle field: device pointer and floats (or ints) per particle.
bool particle_field(jm_handle* h, int f, void** dptr, int* comps) {
  jm::State& s = h->st.s;
  switch (f) {
    case JM_FIELD_X: *dptr = s.x; *comps = 2; return true;
    case JM_FIELD_V: *dptr = s.v; *comps = 2; return true;
    case JM_FIELD_C: *dptr = s.C; *comps = 4; return true;
    case JM_FIELD_F: *dptr = s.F; *comps = 4; return true;
    case JM_FIELD_MATERIAL: *dptr = s.mat; *comps = 1; return true;
    case JM_FIELD_JP: *dptr = s.Jp; *comps = 1; return true;
    case JM_FIELD_FIBER: *dptr = s.fiber; *comps = 2; return true;
    case JM_FIELD_GLOW: *dptr = s.glow; *comps = 1; return s.glow != nullptr;
  }
  return false;
}

int read_slots(jm_handle* h, std::vector<int>& slots) {
  slots.assign((size_t)h->M * h->cap, -1);
  for (int m = 0; m < h->M; ++m) {
    const int n = h->n_live[m];
    if (n == 0) continue;
    HIP_TRY(hipMemcpyAsync(slots.data() + (size_t)m * h->cap, h->st.s.slot + (size_t)m * h->cap,
                           n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return JM_OK;
}

int fill_stats(jm_handle* h) {
  hipLaunchKernelGGL(jm::k_stats, dim3(h->M), dim3(256), 0, h->stream, h->st.s, h->n_live_d,
                     h->cap, h->stats_d);
  HIP_TRY(hipGetLastError());
  return JM_OK;
}


int check_handle(jm_handle* h) {
  if (!h) return fail(JM_ERR_ARG, "null handle");
  return JM_OK;
}

}  // namespace

// =========================================================================================
// C ABI

extern "C" {

const char* jm_last_error(void) { return g_err.c_str(); }

int jm_default_config(jm_config* cfg) {
  if (!cfg) return fail(JM_ERR_ARG, "null config");
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->abi_version = JM_ABI_VERSION;
  cfg->n_instances = 16;    // mpm_sim.py:11
  cfg->n_particles = 80000; // mpm_sim.py:13
  cfg->quality = 1;         // mpm_sim.py:12
  cfg->n_grid_y = 0;
  cfg->track_glow = 0;
  cfg->path = JM_PATH_FUSED;
  cfg->domain_height = 1.0; // mpm_sim.py:16
  return JM_OK;
}

int jm_create(const jm_config* cfg, jm_handle** out) {
  if (!cfg || !out) return fail(JM_ERR_ARG, "null argument");
  *out = nullptr;
  if (cfg->abi_version != JM_ABI_VERSION) return fail(JM_ERR_ARG, "abi_version mismatch");
  if (cfg->n_instances <= 0 || cfg->n_particles <= 0 || cfg->quality <= 0)
    return fail(JM_ERR_ARG, "n_instances, n_particles and quality must be positive");
  if (cfg->path != JM_PATH_FUSED && cfg->path != JM_PATH_PLAIN)
    return fail(JM_ERR_ARG, "unknown path");
  if (!(cfg->domain_height > 0.002)) return fail(JM_ERR_ARG, "domain_height too small");
  if (cfg->axisym != 0 && cfg->axisym != 1) return fail(JM_ERR_ARG, "axisym must be 0 or 1");
  if (cfg->nan_invalid != 0 && cfg->nan_invalid != 1)
    return fail(JM_ERR_ARG, "nan_invalid must be 0 or 1");
  if (cfg->reserved != 0) return fail(JM_ERR_ARG, "reserved must be 0");
  const int nx = 128 * cfg->quality;
  const int ny = cfg->n_grid_y > 0 ? cfg->n_grid_y : nx;
  if (ny < 8 || ny % 4 != 0) return fail(JM_ERR_ARG, "n_grid_y must be a multiple of 4, >= 8");
  // the kernels address particle arrays with 32-bit byte offsets (16-byte elements at most)
  // and pack a base cell's two indices into 16 bits each
  if ((int64_t)cfg->n_instances * cfg->n_particles > ((int64_t)1 << 28))
    return fail(JM_ERR_ARG, "n_instances * n_particles above 2^28 slots");
  if (cfg->quality > 255 || (cfg->n_grid_y > 0 && cfg->n_grid_y >= 32768))
    return fail(JM_ERR_ARG, "grid above 32767 nodes per axis");

  jm_handle* h = new jm_handle();
  h->cfg = *cfg;
  h->M = cfg->n_instances;
  h->N = cfg->n_particles;
  h->nx = nx;
  h->ny = ny;
  h->G = nx * ny;
  h->cap = cfg->n_particles;
  h->k = make_consts(*cfg, nx, ny);
  h->tun = {100, 100, 256, 4, 1, 1, 1, 8, 2};
  const int M = h->M;
  h->h_param[JM_PARAM_ACTUATION].assign(M, 500.0f);       // mpm_sim.py:146
  h->h_param[JM_PARAM_CONTRACTION].assign(M, 0.20f);      // :148
  h->h_param[JM_PARAM_FREQ].assign(M, 1.0f);              // :149
  h->h_param[JM_PARAM_PAYLOAD_DENSITY].assign(M, 2.5f);   // :147
  h->h_param[JM_PARAM_E_JELLY].assign(M, 0.7e3f);         // :31
  h->n_live.assign(M, 0);
  h->h_pos.assign((size_t)M * h->N * 2, -1.0f);
  h->h_fiber.assign((size_t)M * h->N * 2, 0.0f);
  for (size_t q = 0; q < (size_t)M * h->N; ++q) h->h_fiber[2 * q + 1] = 1.0f;
  h->h_mat.assign((size_t)M * h->N, -1);
  h->inst_dirty.assign(M, 0);

  auto cleanup = [&](int rc) {
    delete h;
    return rc;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming) != hipSuccess)
    return cleanup(fail(JM_ERR_HIP, "hipStreamCreateWithFlags"));
  for (int g = 1; g < kMaxGroups; ++g)  // default priority, like the main stream
    if (hipStreamCreateWithFlags(&h->gstream[g], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->gdone[g], hipEventDisableTiming) != hipSuccess)
      return cleanup(fail(JM_ERR_HIP, "instance group streams"));
  const size_t n = (size_t)M * h->cap;
  const bool glow = cfg->track_glow != 0;
  const bool axisym = cfg->axisym != 0;
  int rc = alloc_state(h->st, n, glow, axisym);
  if (rc == JM_OK) rc = alloc_state(h->tmp, n, glow, axisym);
  if (rc != JM_OK) return cleanup(rc);
  auto dm = [&](void** p, size_t bytes) {
    return hipMalloc(p, std::max<size_t>(bytes, 256)) == hipSuccess;
  };
  h->tmax = (h->cap + jm::kSortTile - 1) / jm::kSortTile;
  h->key_bits = sort_key_bits(nx, ny, h->tun.sort_strip);
  bool ok = dm((void**)&h->sortbuf, n * 5 * 4) &&
            dm((void**)&h->hist, (size_t)M * 256 * h->tmax * 4) &&
            dm((void**)&h->acc, (size_t)3 * M * 3 * h->G * 4) && dm((void**)&h->t_buf, 8 * kMaxGroups) &&
            dm((void**)&h->params, (size_t)6 * M * 4) && dm((void**)&h->n_live_d, M * 4) &&
            dm((void**)&h->nchunks_d, M * 4) && dm((void**)&h->stats_d, (size_t)M * 6 * 4) &&
            dm((void**)&h->nonfinite_d, M * 4) &&
            dm((void**)&h->staging, (size_t)M * 2 * h->G * 4) &&
            dm((void**)&h->damp, (size_t)(nx + ny) * 4) &&
            (!axisym || dm((void**)&h->hoop, (size_t)3 * M * 2 * h->G * 4));
  if (!ok) return cleanup(fail(JM_ERR_NOMEM, "hipMalloc(engine buffers)"));
  std::vector<float> damp(nx + ny);
  for (int i = 0; i < nx; ++i) damp[i] = jm::damping_factor(i, nx, h->k.damp_x);
  for (int j = 0; j < ny; ++j) damp[nx + j] = jm::damping_factor(j, ny, h->k.damp_y);
  if (hipMemcpy(h->damp, damp.data(), damp.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(JM_ERR_HIP, "damping table upload"));
  h->k.damp = h->damp;
  if (hipMemsetAsync(h->acc, 0, (size_t)3 * M * 3 * h->G * 4, h->stream) != hipSuccess ||
      (h->hoop &&
       hipMemsetAsync(h->hoop, 0, (size_t)3 * M * 2 * h->G * 4, h->stream) != hipSuccess) ||
      hipMemsetAsync(h->t_buf, 0, 8 * kMaxGroups, h->stream) != hipSuccess ||
      hipMemsetAsync(h->stats_d, 0, (size_t)M * 6 * 4, h->stream) != hipSuccess ||
      hipMemsetAsync(h->n_live_d, 0, M * 4, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return cleanup(fail(JM_ERR_HIP, "initial clears"));
  h->layout_dirty = true;
  *out = h;
  g_err.clear();
  return JM_OK;
}

int jm_destroy(jm_handle* h) {
  if (!h) return JM_OK;
  delete h;
  return JM_OK;
}

int jm_set_tuning(jm_handle* h, const jm_tuning* t) {
  JM_TRY(check_handle(h));
  if (!t) return fail(JM_ERR_ARG, "null tuning");
  if (t->block_substeps < 1 || t->block_substeps > 100000)
    return fail(JM_ERR_ARG, "block_substeps out of range");
  if (t->sort_interval < 1) return fail(JM_ERR_ARG, "sort_interval must be >= 1");
  if (!tuning_supported(t->threads, t->ppt, h->cfg.axisym != 0))
    return fail(JM_ERR_ARG, h->cfg.axisym ? "the axisymmetric variant runs 256 x 2 or 256 x 4"
                                          : "unsupported (threads, ppt) combination");
  if (t->sort_strip < 0 || t->sort_strip > h->ny)
    return fail(JM_ERR_ARG, "sort_strip must be in [0, grid rows]");
  if (t->instance_groups < 1 || t->instance_groups > kMaxGroups)
    return fail(JM_ERR_ARG, "instance_groups must be in [1, 4]");
  JM_TRY(jm_sync(h));
  const bool relayout = t->threads * t->ppt != h->tun.threads * h->tun.ppt ||
                        t->xcd_affinity != h->tun.xcd_affinity ||
                        t->instance_groups != h->tun.instance_groups;
  if (t->sort_strip != h->tun.sort_strip) h->since_sort = t->sort_interval;  // re-sort next
  h->tun = *t;
  h->tuning_user = true;
  h->key_bits = sort_key_bits(h->nx, h->ny, t->sort_strip);
  h->drop_graphs();
  if (relayout && !h->layout_dirty) JM_TRY(build_chunks(h));
  return JM_OK;
}

int jm_get_tuning(jm_handle* h, jm_tuning* t) {
  JM_TRY(check_handle(h));
  if (!t) return fail(JM_ERR_ARG, "null tuning");
  *t = h->tun;
  return JM_OK;
}

int jm_load_particles(jm_handle* h, int32_t instance, const float* pos, const int32_t* mat,
                      const float* fiber, int64_t n) {
  JM_TRY(check_handle(h));
  if (instance < 0 || instance >= h->M) return fail(JM_ERR_ARG, "instance out of range");
  if (!pos || !mat) return fail(JM_ERR_ARG, "pos and mat are required");
  if (n < 0 || n > h->N) return fail(JM_ERR_ARG, "n must be in [0, n_particles]");
  const size_t base = (size_t)instance * h->N;
  for (int64_t p = 0; p < n; ++p) {
    const int32_t mt = mat[p];
    const float x0 = pos[2 * p], x1 = pos[2 * p + 1];
    if (mt >= 0 && !(std::isfinite(x0) && std::isfinite(x1)))
      return fail(JM_ERR_ARG, "non-finite position of a live particle");
  }
  for (int64_t p = 0; p < h->N; ++p) {
    const size_t q = base + p;
    if (p < n) {
      h->h_pos[2 * q] = pos[2 * p];
      h->h_pos[2 * q + 1] = pos[2 * p + 1];
      h->h_mat[q] = mat[p];
      h->h_fiber[2 * q] = fiber ? fiber[2 * p] : 0.0f;  // default (0,1), mpm_sim.py:733-735
      h->h_fiber[2 * q + 1] = fiber ? fiber[2 * p + 1] : 1.0f;
    } else {
      h->h_pos[2 * q] = -1.0f;
      h->h_pos[2 * q + 1] = -1.0f;
      h->h_mat[q] = -1;
      h->h_fiber[2 * q] = 0.0f;
      h->h_fiber[2 * q + 1] = 1.0f;
    }
  }
  h->inst_dirty[instance] = 1;
  h->layout_dirty = true;
  return JM_OK;
}

int jm_set_instance_param(jm_handle* h, int32_t param, const float* values) {
  JM_TRY(check_handle(h));
  if (param < 0 || param > JM_PARAM_E_JELLY) return fail(JM_ERR_ARG, "unknown param");
  if (!values) return fail(JM_ERR_ARG, "null values");
  for (int m = 0; m < h->M; ++m) h->h_param[param][m] = values[m];
  h->params_dirty = true;
  return upload_params(h);
}

int jm_get_instance_param(jm_handle* h, int32_t param, float* values) {
  JM_TRY(check_handle(h));
  if (param < 0 || param > JM_PARAM_E_JELLY) return fail(JM_ERR_ARG, "unknown param");
  if (!values) return fail(JM_ERR_ARG, "null values");
  for (int m = 0; m < h->M; ++m) values[m] = h->h_param[param][m];
  return JM_OK;
}

int jm_set_time(jm_handle* h, float t) {
  JM_TRY(check_handle(h));
  float clocks[2 * kMaxGroups] = {0.f};
  for (int g = 0; g < kMaxGroups; ++g) clocks[2 * g] = t;  // every instance group's clock
  HIP_TRY(hipMemcpyAsync(h->t_buf, clocks, sizeof(clocks), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return JM_OK;
}

int jm_get_time(jm_handle* h, float* t) {
  JM_TRY(check_handle(h));
  if (!t) return fail(JM_ERR_ARG, "null t");
  HIP_TRY(hipMemcpyAsync(t, h->t_buf, sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return JM_OK;
}

int jm_substep(jm_handle* h, int64_t n) {
  JM_TRY(check_handle(h));
  if (n < 0) return fail(JM_ERR_ARG, "negative step count");
  if (n == 0) return JM_OK;
  JM_TRY(commit(h));
  if (h->cfg.path == JM_PATH_PLAIN) return run_plain(h, n);
  return run_fused(h, n);
}

int jm_sort_now(jm_handle* h) {
  JM_TRY(check_handle(h));
  JM_TRY(commit(h));
  JM_TRY(enqueue_sort(h));
  h->since_sort = 0;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return JM_OK;
}

int jm_compute_stats(jm_handle* h, int32_t mode, float* out) {
  JM_TRY(check_handle(h));
  if (!out) return fail(JM_ERR_ARG, "null out");
  if (mode < JM_STATS_PAYLOAD || mode > JM_STATS_ANY) return fail(JM_ERR_ARG, "unknown mode");
  JM_TRY(commit(h));
  JM_TRY(fill_stats(h));
  std::vector<float> raw((size_t)h->M * 6);
  HIP_TRY(hipMemcpyAsync(raw.data(), h->stats_d, raw.size() * 4, hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (int m = 0; m < h->M; ++m) {
    const float* pay = &raw[(size_t)m * 6];
    const float* body = pay + 3;
    const float* src = pay;
    if (mode == JM_STATS_BODY || (mode == JM_STATS_ANY && pay[1] == 0.0f)) src = body;
    for (int c = 0; c < 3; ++c) out[3 * m + c] = src[c];
  }
  return JM_OK;
}

int jm_run_batch_headless(jm_handle* h, int64_t steps, double* out) {
  JM_TRY(check_handle(h));
  if (!out || steps < 0) return fail(JM_ERR_ARG, "bad arguments");
  const int M = h->M;
  JM_TRY(jm_set_time(h, 0.0f));  // mpm_sim.py:808
  std::vector<float> r0((size_t)M * 3), r1((size_t)M * 3);
  JM_TRY(jm_compute_stats(h, JM_STATS_ANY, r0.data()));  // :811
  JM_TRY(jm_substep(h, steps));                            // :814-815
  JM_TRY(jm_compute_stats(h, JM_STATS_ANY, r1.data()));  // :818
  std::vector<int> nonfinite(M, 0);
  if (h->cfg.nan_invalid) {
    hipLaunchKernelGGL(jm::k_fill_i32, dim3(1), dim3(256), 0, h->stream, h->nonfinite_d,
                       (int64_t)M, 0);
    hipLaunchKernelGGL(jm::k_count_nonfinite, dim3(blocks_for(max_live(h), 256), M), dim3(256),
                       0, h->stream, h->st.s, h->n_live_d, h->cap, h->nonfinite_d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(nonfinite.data(), h->nonfinite_d, M * sizeof(int),
                           hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  // mpm_sim.py:793-800: raw[i, 0] / count divides two np.float32 values, so the quotient is
  // rounded to fp32 before it is stored in the float64 table; then :821-834
  for (int m = 0; m < M; ++m) {
    double* o = out + 5 * m;
    for (int c = 0; c < 5; ++c) o[c] = 0.0;
    const float c0 = r0[3 * m + 1], c1 = r1[3 * m + 1];
    if (c0 > 0.0f && c1 > 0.0f) {
      o[0] = (double)(r0[3 * m] / c0);
      o[1] = (double)(r0[3 * m + 2] / c0);
      o[2] = (double)(r1[3 * m] / c1);
      o[3] = (double)(r1[3 * m + 2] / c1);
      // The reference's gate `final_y < 0.01` lets a NaN CoM through as valid (H12);
      // nan_invalid = 1 marks an instance invalid when its CoM is not finite or any of its
      // live particles ended with a non-finite v, C or F.
      const bool finite = std::isfinite(o[0]) && std::isfinite(o[1]) && std::isfinite(o[2]) &&
                          std::isfinite(o[3]) && nonfinite[m] == 0;
      o[4] = (o[2] < 0.01 || (h->cfg.nan_invalid && !finite)) ? 0.0 : 1.0;
    }
  }
  return JM_OK;
}

int jm_read_field(jm_handle* h, int32_t field, void* out, int64_t nbytes) {
  JM_TRY(check_handle(h));
  if (!valid_field(field)) return fail(JM_ERR_ARG, "unknown field");
  if (!out) return fail(JM_ERR_ARG, "null out");
  if (nbytes != field_bytes(h, field)) return fail(JM_ERR_ARG, "nbytes does not match field");
  JM_TRY(commit(h));
  const int M = h->M, N = h->N, G = h->G, nx = h->nx, ny = h->ny;
  if (field == JM_FIELD_GRID_V || field == JM_FIELD_GRID_M) {
    std::vector<float> planes((size_t)M * 3 * G, 0.0f);
    if (h->last_read >= 0) {
      if (field == JM_FIELD_GRID_V && !h->grid_is_velocity) {
        hipLaunchKernelGGL(jm::k_grid_velocity, dim3(blocks_for(G, 256), M), dim3(256), 0,
                           h->stream, grid_buf(h, h->last_read), h->staging, M, G, h->k);
        HIP_TRY(hipGetLastError());
        std::vector<float> vel((size_t)M * 2 * G);
        HIP_TRY(hipMemcpyAsync(vel.data(), h->staging, vel.size() * 4, hipMemcpyDeviceToHost,
                               h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (int m = 0; m < M; ++m) {
          std::memcpy(&planes[(size_t)m * 3 * G], &vel[(size_t)m * 2 * G], (size_t)2 * G * 4);
        }
      } else {
        HIP_TRY(hipMemcpyAsync(planes.data(), grid_buf(h, h->last_read), planes.size() * 4,
                               hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
      }
    }
    float* o = static_cast<float*>(out);
    for (int m = 0; m < M; ++m) {
      const float* pl = &planes[(size_t)m * 3 * G];
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
          const size_t node = (size_t)j * nx + i;
          const size_t q = ((size_t)m * nx + i) * ny + j;
          if (field == JM_FIELD_GRID_V) {
            o[2 * q] = pl[node];
            o[2 * q + 1] = pl[G + node];
          } else {
            o[q] = pl[2 * G + node];
          }
        }
    }
    return JM_OK;
  }
  if (field == JM_FIELD_FITNESS) {
    // the buffer as compute_payload_stats leaves it (mpm_sim.py:109)
    std::vector<float> raw((size_t)M * 6);
    HIP_TRY(hipMemcpyAsync(raw.data(), h->stats_d, raw.size() * 4, hipMemcpyDeviceToHost,
                           h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    float* o = static_cast<float*>(out);
    for (int m = 0; m < M; ++m)
      for (int c = 0; c < 3; ++c) o[3 * m + c] = raw[(size_t)m * 6 + c];
    return JM_OK;
  }
  void* dptr = nullptr;
  int comps = 0;
  if (!particle_field(h, field, &dptr, &comps))
    return fail(JM_ERR_STATE, "field not tracked (glow needs track_glow)");
  // defaults for dead slots: the loaded values (x, material, fiber) or the reset state
  if (field == JM_FIELD_X) {
    std::memcpy(out, h->h_pos.data(), (size_t)nbytes);
  } else if (field == JM_FIELD_FIBER) {
    std::memcpy(out, h->h_fiber.data(), (size_t)nbytes);
  } else if (field == JM_FIELD_MATERIAL) {
    std::memcpy(out, h->h_mat.data(), (size_t)nbytes);
  } else {
    float* o = static_cast<float*>(out);
    const size_t MN = (size_t)M * N;
    for (size_t q = 0; q < MN; ++q) {
      if (field == JM_FIELD_F) {
        o[4 * q] = 1.f;
        o[4 * q + 1] = 0.f;
        o[4 * q + 2] = 0.f;
        o[4 * q + 3] = 1.f;
      } else if (field == JM_FIELD_JP) {
        o[q] = 1.f;
      } else {
        for (int c = 0; c < comps; ++c) o[comps * q + c] = 0.f;
      }
    }
  }
  std::vector<int> slots;
  JM_TRY(read_slots(h, slots));
  const size_t esz = (size_t)comps * 4;
  std::vector<char> buf;
  char* o = static_cast<char*>(out);
  for (int m = 0; m < M; ++m) {
    const int n = h->n_live[m];
    if (n == 0) continue;
    buf.resize((size_t)n * esz);
    HIP_TRY(hipMemcpyAsync(buf.data(), static_cast<char*>(dptr) + (size_t)m * h->cap * esz,
                           buf.size(), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int* sl = slots.data() + (size_t)m * h->cap;
    for (int i = 0; i < n; ++i)
      std::memcpy(o + ((size_t)m * N + sl[i]) * esz, buf.data() + (size_t)i * esz, esz);
  }
  return JM_OK;
}

int jm_write_field(jm_handle* h, int32_t field, const void* in, int64_t nbytes) {
  JM_TRY(check_handle(h));
  if (!valid_field(field)) return fail(JM_ERR_ARG, "unknown field");
  if (!in) return fail(JM_ERR_ARG, "null input");
  if (field == JM_FIELD_MATERIAL || field >= JM_FIELD_GRID_V)
    return fail(JM_ERR_ARG, "field is read-only (materials are set by jm_load_particles)");
  if (nbytes != field_bytes(h, field)) return fail(JM_ERR_ARG, "nbytes does not match field");
  JM_TRY(commit(h));
  void* dptr = nullptr;
  int comps = 0;
  if (!particle_field(h, field, &dptr, &comps))
    return fail(JM_ERR_STATE, "field not tracked (glow needs track_glow)");
  const int M = h->M, N = h->N;
  const size_t esz = (size_t)comps * 4;
  const char* src = static_cast<const char*>(in);
  if (field == JM_FIELD_X || field == JM_FIELD_FIBER) {  // keep dead-slot contents too
    std::vector<float>& hv = field == JM_FIELD_X ? h->h_pos : h->h_fiber;
    for (size_t q = 0; q < (size_t)M * N; ++q)
      if (h->h_mat[q] < 0) std::memcpy(&hv[2 * q], src + q * esz, esz);
  }
  std::vector<int> slots;
  JM_TRY(read_slots(h, slots));
  std::vector<char> buf;
  for (int m = 0; m < M; ++m) {
    const int n = h->n_live[m];
    if (n == 0) continue;
    buf.resize((size_t)n * esz);
    const int* sl = slots.data() + (size_t)m * h->cap;
    for (int i = 0; i < n; ++i)
      std::memcpy(buf.data() + (size_t)i * esz, src + ((size_t)m * N + sl[i]) * esz, esz);
    HIP_TRY(hipMemcpyAsync(static_cast<char*>(dptr) + (size_t)m * h->cap * esz, buf.data(),
                           buf.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  if (field == JM_FIELD_X) h->since_sort = h->tun.sort_interval;  // positions moved: re-sort
  return JM_OK;
}

int jm_live_count(jm_handle* h, int32_t instance, int64_t* n) {
  JM_TRY(check_handle(h));
  if (instance < 0 || instance >= h->M || !n) return fail(JM_ERR_ARG, "bad arguments");
  JM_TRY(commit(h));
  *n = h->n_live[instance];
  return JM_OK;
}

int jm_read_device_order(jm_handle* h, int32_t instance, int32_t* slot, float* x, int64_t n) {
  JM_TRY(check_handle(h));
  if (instance < 0 || instance >= h->M) return fail(JM_ERR_ARG, "instance out of range");
  JM_TRY(commit(h));
  if (n != h->n_live[instance]) return fail(JM_ERR_ARG, "n must equal the live count");
  const size_t off = (size_t)instance * h->cap;
  if (n > 0 && slot)
    HIP_TRY(hipMemcpyAsync(slot, h->st.s.slot + off, n * 4, hipMemcpyDeviceToHost, h->stream));
  if (n > 0 && x)
    HIP_TRY(hipMemcpyAsync(x, h->st.s.x + off, n * 8, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return JM_OK;
}

int jm_timed_block(jm_handle* h, int32_t n, float* fused_ms, float* block_ms) {
  JM_TRY(check_handle(h));
  if (n < 2) return fail(JM_ERR_ARG, "n must be >= 2");
  if (h->cfg.path != JM_PATH_FUSED) return fail(JM_ERR_STATE, "timed blocks need the fused path");
  JM_TRY(commit(h));
  if (h->total_chunks == 0) return fail(JM_ERR_STATE, "no live particles");
  const bool sort = h->since_sort >= h->tun.sort_interval;
  if (sort) JM_TRY(enqueue_sort(h));
  const int b = h->acc_base;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
  // Events on the main stream around kernel 0, kernels 1..n-1 and kernel n of every
  // instance group (each segment forked to the groups' streams and joined back), so the
  // middle interval is the wall time of n - 1 substeps of all instances.
  // The groups' launches are enqueued interleaved (kernel kk of every group, then kk + 1),
  // so that no group waits for the host to enqueue another group's whole segment.
  auto segment = [&](int kk0, int kk1) {
    int rc = fork_groups(h);
    for (int kk = kk0; kk <= kk1 && rc == JM_OK; ++kk)
      for (int g = 0; g < h->n_groups && rc == JM_OK; ++g)
        rc = enqueue_group_kernels(h, g, kk, kk, n, b);
    if (rc == JM_OK) rc = join_groups(h);
    return rc;
  };
  int rc = JM_OK;
  if (hipEventRecord(ev[0], h->stream) != hipSuccess) rc = JM_ERR_HIP;
  if (rc == JM_OK) rc = segment(0, 0);
  if (rc == JM_OK && hipEventRecord(ev[1], h->stream) != hipSuccess) rc = JM_ERR_HIP;
  if (rc == JM_OK) rc = segment(1, n - 1);
  if (rc == JM_OK && hipEventRecord(ev[2], h->stream) != hipSuccess) rc = JM_ERR_HIP;
  if (rc == JM_OK) rc = segment(n, n);
  if (rc == JM_OK && hipEventRecord(ev[3], h->stream) != hipSuccess) rc = JM_ERR_HIP;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipEventSynchronize(ev[3]);
  float mid = 0.f, all = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&mid, ev[1], ev[2]);
  if (e == hipSuccess) e = hipEventElapsedTime(&all, ev[0], ev[3]);
  for (auto& ee : ev) (void)hipEventDestroy(ee);
  if (e != hipSuccess) return fail(JM_ERR_HIP, std::string("timed block: ") + hipGetErrorString(e));
  if (rc != JM_OK) return rc == JM_ERR_HIP ? fail(rc, "timed block: enqueue failed") : rc;
  finish_block(h, n, b, sort);
  JM_TRY(finish_axisym(h));
  if (fused_ms) *fused_ms = mid / (float)(n - 1);
  if (block_ms) *block_ms = all;
  return JM_OK;
}

int jm_sync(jm_handle* h) {
  JM_TRY(check_handle(h));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return JM_OK;
}

void* jm_stream(jm_handle* h) { return h ? (void*)h->stream : nullptr; }

}  // extern "C"

// -----------------------------------------------------------------------------------------
// Renderer (mpm_sim.py:343-712)

namespace {

jm::RenderArgs render_args(jm_handle* h, int res_sub, int grid_side, float radius) {
  jm::RenderArgs a;
  std::memset(&a, 0, sizeof(a));
  const jm::InstParams ip = inst_params(h);
  a.s = h->st.s;
  a.n_live = h->n_live_d;
  a.cap = h->cap;
  a.M = h->M;
  a.freq = ip.freq;
  a.contraction = ip.contraction;
  a.hue = h->rhue;
  a.muscle_hue = h->rhue ? h->rhue + h->M : nullptr;
  a.t_buf = h->t_buf;
  a.fb = h->frame;
  a.res = h->frame_res;
  a.res_sub = res_sub;
  a.grid_side = grid_side;
  a.radius = radius;
  a.domain_height = (float)h->cfg.domain_height;
  a.k = h->k;
  return a;
}

int render_ready(jm_handle* h, int res_sub, int grid_side) {
  JM_TRY(check_handle(h));
  if (!h->frame) return fail(JM_ERR_STATE, "no frame buffer: call jm_render_clear first");
  if (res_sub < 1 || res_sub > 16384 || grid_side < 1 || grid_side > 16384)
    return fail(JM_ERR_ARG, "res_sub and grid_side must be in [1, 16384]");
  return commit(h);
}

// The grid the last G2P gathered and whether its planes already hold velocities (plain path).
const float* render_grid(jm_handle* h, int* vel_planes) {
  *vel_planes = h->grid_is_velocity ? 1 : 0;
  return h->last_read >= 0 ? grid_buf(h, h->last_read) : nullptr;
}

}  // namespace

extern "C" {

int jm_render_clear(jm_handle* h, int32_t res, float value) {
  JM_TRY(check_handle(h));
  if (res < 1 || res > 16384) return fail(JM_ERR_ARG, "res must be in [1, 16384]");
  if (res != h->frame_res) {
    JM_TRY(jm_sync(h));
    if (h->frame) (void)hipFree(h->frame);
    if (h->frame_u8) (void)hipFree(h->frame_u8);
    h->frame = nullptr;
    h->frame_u8 = nullptr;
    h->frame_res = 0;
    if (hipMalloc(&h->frame, (size_t)res * res * 3 * sizeof(float)) != hipSuccess)
      return fail(JM_ERR_NOMEM, "hipMalloc(frame buffer)");
    h->frame_res = res;
  }
  const int64_t n = (int64_t)res * res * 3;
  hipLaunchKernelGGL(jm::k_fb_fill, dim3(blocks_for(n, 256 * 8)), dim3(256), 0, h->stream,
                     h->frame, n, value);
  HIP_TRY(hipGetLastError());
  return JM_OK;
}

int jm_render_abyss(jm_handle* h, int32_t res_sub, int32_t grid_side, float radius,
                    const float* hue, const float* muscle_hue, int32_t water_mode) {
  JM_TRY(render_ready(h, res_sub, grid_side));
  if (!(radius >= 0.0f && radius <= 16384.0f))
    return fail(JM_ERR_ARG, "radius must be in [0, 16384]");
  if (!hue || !muscle_hue) return fail(JM_ERR_ARG, "null hue arrays");
  if (!h->rhue && hipMalloc(&h->rhue, (size_t)2 * h->M * sizeof(float)) != hipSuccess)
    return fail(JM_ERR_NOMEM, "hipMalloc(render hues)");
  // the hues live on the device; upload only when they change (then wait for the copy, so
  // that the host copy can be rewritten by the next change)
  const size_t M = h->M;
  if (h->rhue_host.size() != 2 * M ||
      std::memcmp(h->rhue_host.data(), hue, M * sizeof(float)) != 0 ||
      std::memcmp(h->rhue_host.data() + M, muscle_hue, M * sizeof(float)) != 0) {
    h->rhue_host.assign(hue, hue + M);
    h->rhue_host.insert(h->rhue_host.end(), muscle_hue, muscle_hue + M);
    HIP_TRY(hipMemcpyAsync(h->rhue, h->rhue_host.data(), 2 * M * sizeof(float),
                           hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  jm::RenderArgs a = render_args(h, res_sub, grid_side, radius);
  a.water_mode = water_mode;
  hipLaunchKernelGGL(jm::k_render_abyss, dim3(blocks_for(max_live(h), 256), h->M), dim3(256),
                     0, h->stream, a);
  HIP_TRY(hipGetLastError());
  return JM_OK;
}

int jm_render_flat(jm_handle* h, int32_t res_sub, int32_t grid_side, float radius,
                   int32_t material, float r, float g, float b) {
  JM_TRY(render_ready(h, res_sub, grid_side));
  if (!(radius >= 0.0f && radius <= 16384.0f))
    return fail(JM_ERR_ARG, "radius must be in [0, 16384]");
  jm::RenderArgs a = render_args(h, res_sub, grid_side, radius);
  a.mat_sel = material;
  a.flat = make_float3(r, g, b);
  hipLaunchKernelGGL(jm::k_render_flat, dim3(blocks_for(max_live(h), 256), h->M), dim3(256),
                     0, h->stream, a);
  HIP_TRY(hipGetLastError());
  return JM_OK;
}

int jm_render_tone_map(jm_handle* h) {
  JM_TRY(check_handle(h));
  if (!h->frame) return fail(JM_ERR_STATE, "no frame buffer: call jm_render_clear first");
  const int64_t n = (int64_t)h->frame_res * h->frame_res * 3;
  hipLaunchKernelGGL(jm::k_tone_map, dim3(blocks_for(n, 256 * 4)), dim3(256), 0, h->stream,
                     h->frame, n);
  HIP_TRY(hipGetLastError());
  return JM_OK;
}

int jm_render_vorticity_grid(jm_handle* h, float scale) {
  JM_TRY(check_handle(h));
  JM_TRY(commit(h));
  const int M = h->M, G = h->G;
  if (!h->vort && hipMalloc(&h->vort, (size_t)M * G * sizeof(float)) != hipSuccess)
    return fail(JM_ERR_NOMEM, "hipMalloc(vorticity grid)");
  int vel_planes = 0;
  const float* g = render_grid(h, &vel_planes);
  if (!g) {  // no substep yet: the reference's reset grid has no curl
    hipLaunchKernelGGL(jm::k_fb_fill, dim3(blocks_for((int64_t)M * G, 256 * 8)), dim3(256), 0,
                       h->stream, h->vort, (int64_t)M * G, 0.0f);
  } else {
    hipLaunchKernelGGL(jm::k_vorticity_grid, dim3(blocks_for(G, 256), M), dim3(256), 0,
                       h->stream, g, h->vort, M, G, scale, vel_planes, h->k);
  }
  HIP_TRY(hipGetLastError());
  return JM_OK;
}

int jm_render_vorticity(jm_handle* h, int32_t mode, int32_t res_sub, int32_t grid_side,
                        float param) {
  JM_TRY(render_ready(h, res_sub, grid_side));
  if (mode < 0 || mode > 3) return fail(JM_ERR_ARG, "mode must be 0..3");
  const int M = h->M, G = h->G;
  jm::RenderArgs a = render_args(h, res_sub, grid_side, 0.0f);
  if (mode == 0) {
    int vel_planes = 0;
    const float* g = render_grid(h, &vel_planes);
    if (g)
      hipLaunchKernelGGL(jm::k_vorticity_overlay, dim3(blocks_for(G, 256), M), dim3(256), 0,
                         h->stream, g, M, G, param, vel_planes, a);
  } else {
    if (!h->vort) return fail(JM_ERR_STATE, "no vorticity grid: call jm_render_vorticity_grid");
    if (mode == jm::kVortWhite) {
      const int64_t n = (int64_t)h->frame_res * h->frame_res * 3;
      hipLaunchKernelGGL(jm::k_fb_fill, dim3(blocks_for(n, 256 * 8)), dim3(256), 0, h->stream,
                         h->frame, n, 1.0f);
    }
    hipLaunchKernelGGL(jm::k_vorticity_pixels, dim3(blocks_for(h->frame_res, 256), h->frame_res),
                       dim3(256), 0, h->stream, h->vort, M, G, mode, param, a);
  }
  HIP_TRY(hipGetLastError());
  return JM_OK;
}

int jm_read_frame(jm_handle* h, float* out, int64_t n) {
  JM_TRY(check_handle(h));
  if (!out) return fail(JM_ERR_ARG, "null out");
  if (!h->frame) return fail(JM_ERR_STATE, "no frame buffer: call jm_render_clear first");
  if (n != (int64_t)h->frame_res * h->frame_res * 3) return fail(JM_ERR_ARG, "n != res*res*3");
  HIP_TRY(hipMemcpyAsync(out, h->frame, n * 4, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return JM_OK;
}

int jm_read_frame_u8(jm_handle* h, uint8_t* out, int64_t n) {
  JM_TRY(check_handle(h));
  if (!out) return fail(JM_ERR_ARG, "null out");
  if (!h->frame) return fail(JM_ERR_STATE, "no frame buffer: call jm_render_clear first");
  if (n != (int64_t)h->frame_res * h->frame_res * 3) return fail(JM_ERR_ARG, "n != res*res*3");
  if (!h->frame_u8 && hipMalloc(&h->frame_u8, (size_t)n) != hipSuccess)
    return fail(JM_ERR_NOMEM, "hipMalloc(8-bit frame)");
  hipLaunchKernelGGL(jm::k_encode_u8, dim3(blocks_for(n, 256 * 4)), dim3(256), 0, h->stream,
                     h->frame, h->frame_u8, n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, h->frame_u8, n, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return JM_OK;
}

int jm_read_vorticity(jm_handle* h, float* out, int64_t n) {
  JM_TRY(check_handle(h));
  if (!out) return fail(JM_ERR_ARG, "null out");
  if (!h->vort) return fail(JM_ERR_STATE, "no vorticity grid: call jm_render_vorticity_grid");
  const int M = h->M, nx = h->nx, ny = h->ny, G = h->G;
  if (n != (int64_t)M * G) return fail(JM_ERR_ARG, "n != n_instances*nx*ny");
  std::vector<float> tmp((size_t)M * G);
  HIP_TRY(hipMemcpyAsync(tmp.data(), h->vort, tmp.size() * 4, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (int m = 0; m < M; ++m)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        out[((size_t)m * nx + i) * ny + j] = tmp[(size_t)m * G + (size_t)j * nx + i];
  return JM_OK;
}

}  // extern "C"

// -----------------------------------------------------------------------------------------
// Tank building (host): the water candidates a body does not exclude (make_jelly.py:359-366).

int jm_water_mask(const double* body, int64_t n_body, const double* cand, int64_t n_cand,
                  double margin, uint8_t* keep) {
#pragma clang fp contract(off)
  if ((n_body > 0 && !body) || (n_cand > 0 && (!cand || !keep)) || n_body < 0 || n_cand < 0)
    return fail(JM_ERR_ARG, "jm_water_mask: bad arguments");
  if (n_body == 0 || !(margin > 0.0)) {
    // no body, or every distance > margin unless it is not: test all pairs
    for (int64_t i = 0; i < n_cand; ++i) {
      bool k = true;
      for (int64_t j = 0; j < n_body && k; ++j) {
        const double dx = cand[2 * i] - body[2 * j], dy = cand[2 * i + 1] - body[2 * j + 1];
        const double d2 = dx * dx;
        k = std::sqrt(d2 + dy * dy) > margin;
      }
      keep[i] = k ? 1 : 0;
    }
    return JM_OK;
  }
  // Bucket the body by cells of side `margin`: a body point within `margin` of a candidate
  // lies at most one cell away in each axis; two cells are searched, so rounding of the
  // cell index cannot hide one. The distance is the Euclidean norm in double, summed in the
  // order x then y and compared with `> margin`, as a k-d tree query followed by that
  // comparison computes it.
  double x0 = body[0], y0 = body[1], x1 = body[0], y1 = body[1];
  for (int64_t j = 1; j < n_body; ++j) {
    x0 = std::min(x0, body[2 * j]);
    x1 = std::max(x1, body[2 * j]);
    y0 = std::min(y0, body[2 * j + 1]);
    y1 = std::max(y1, body[2 * j + 1]);
  }
  if (!(std::isfinite(x0) && std::isfinite(x1) && std::isfinite(y0) && std::isfinite(y1)))
    return fail(JM_ERR_ARG, "jm_water_mask: non-finite body position");
  const double h = margin;
  const int64_t cx = std::min<int64_t>((int64_t)((x1 - x0) / h) + 1, 1 << 20);
  const int64_t cy = std::min<int64_t>((int64_t)((y1 - y0) / h) + 1, 1 << 20);
  if (cx * cy > (int64_t)1 << 26) return fail(JM_ERR_ARG, "jm_water_mask: body too sparse");
  auto cell_of = [&](double v, double lo, int64_t n) -> int64_t {
    const double f = std::floor((v - lo) / h);
    return f < 0 ? -1 : (f >= (double)n ? n : (int64_t)f);
  };
  std::vector<int64_t> start(cx * cy + 1, 0), order(n_body);
  for (int64_t j = 0; j < n_body; ++j)
    ++start[cell_of(body[2 * j + 1], y0, cy) * cx + cell_of(body[2 * j], x0, cx) + 1];
  for (int64_t c = 0; c < cx * cy; ++c) start[c + 1] += start[c];
  {
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t j = 0; j < n_body; ++j)
      order[fill[cell_of(body[2 * j + 1], y0, cy) * cx + cell_of(body[2 * j], x0, cx)]++] = j;
  }
  for (int64_t i = 0; i < n_cand; ++i) {
    const double px = cand[2 * i], py = cand[2 * i + 1];
    const int64_t ix = cell_of(px, x0, cx), iy = cell_of(py, y0, cy);
    bool k = true;
    for (int64_t yy = std::max<int64_t>(iy - 2, 0); yy <= std::min<int64_t>(iy + 2, cy - 1) && k; ++yy)
      for (int64_t xx = std::max<int64_t>(ix - 2, 0); xx <= std::min<int64_t>(ix + 2, cx - 1) && k; ++xx) {
        const int64_t c = yy * cx + xx;
        for (int64_t q = start[c]; q < start[c + 1]; ++q) {
          const int64_t j = order[q];
          const double dx = px - body[2 * j], dy = py - body[2 * j + 1];
          const double d2 = dx * dx;
          if (!(std::sqrt(d2 + dy * dy) > margin)) {
            k = false;
            break;
          }
        }
      }
    keep[i] = k ? 1 : 0;
  }
  return JM_OK;
}

namespace {

// Proper crossing of segments ab and cd (make_jelly.py:31-43): cross products in double,
// parallel below 1e-12, crossing parameters strictly inside (0, 1).
bool segments_cross(const double* a, const double* b, const double* c, const double* d) {
#pragma clang fp contract(off)
  const double r0 = b[0] - a[0], r1 = b[1] - a[1];
  const double s0 = d[0] - c[0], s1 = d[1] - c[1];
  const double rs_a = r0 * s1, rs_b = r1 * s0;
  const double rxs = rs_a - rs_b;
  if (std::fabs(rxs) < 1e-12) return false;
  const double q0 = c[0] - a[0], q1 = c[1] - a[1];
  const double qs_a = q0 * s1, qs_b = q1 * s0, qr_a = q0 * r1, qr_b = q1 * r0;
  const double t = (qs_a - qs_b) / rxs;
  const double u = (qr_a - qr_b) / rxs;
  return 0 < t && t < 1 && 0 < u && u < 1;
}

}  // namespace

// Whether a closed polygon has a crossing between non-adjacent edges (make_jelly.py:46-58).
int jm_polygon_self_intersects(const double* poly, int64_t n, int32_t* out) {
  if (!out || n < 0 || (n > 0 && !poly)) return fail(JM_ERR_ARG, "bad arguments");
  *out = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double* a = poly + 2 * i;
    const double* b = poly + 2 * ((i + 1) % n);
    for (int64_t j = i + 2; j < n; ++j) {
      if (i == 0 && j == n - 1) continue;  // the last edge is adjacent to the first
      if (segments_cross(a, b, poly + 2 * j, poly + 2 * ((j + 1) % n))) {
        *out = 1;
        return JM_OK;
      }
    }
  }
  return JM_OK;
}

// Lattice points inside a polygon (make_jelly.py:150-157,209-215: matplotlib
// Path(poly).contains_points over the meshgrid of xs and ys). The test is the crossing-number
// rule in the multiply form of E. Haines ("Point in Polygon Strategies", Graphics Gems IV):
// a point (tx, ty) toggles its inside flag on every edge a -> b (the polygon's edges in
// order, then the closing edge) whose end points lie on opposite sides of y = ty
// (a.y >= ty differs from b.y >= ty) and for which
//   ((b.y - ty) * (a.x - b.x) >= (b.x - tx) * (a.y - b.y)) == (b.y >= ty),
// each product and difference rounded in double on its own. Along one lattice row ty is
// fixed and the right-hand side is monotone in tx (fl(b.x - tx) falls as tx grows, and the
// product with the constant a.y - b.y is monotone), so each edge toggles a prefix or a
// suffix of the row's sorted xs: a binary search with the same comparison finds its end,
// and a running XOR over the row applies all edges. out[j * nx + i] is point (xs[i], ys[j]),
// the meshgrid's ravel order.
int jm_points_in_polygon_grid(const double* poly, int64_t n, const double* xs, int64_t nx,
                              const double* ys, int64_t ny, uint8_t* out) {
#pragma clang fp contract(off)
  if (n < 0 || nx < 0 || ny < 0 || (n > 0 && !poly) || (nx > 0 && (!xs || !out)) ||
      (ny > 0 && !ys))
    return fail(JM_ERR_ARG, "bad arguments");
  for (int64_t i = 1; i < nx; ++i)
    if (!(xs[i - 1] <= xs[i])) return fail(JM_ERR_ARG, "xs must be ascending");
  std::vector<uint8_t> flip((size_t)nx + 1);
  for (int64_t j = 0; j < ny; ++j) {
    const double ty = ys[j];
    std::fill(flip.begin(), flip.end(), 0);
    for (int64_t e = 0; e < n; ++e) {
      const double ax = poly[2 * e], ay = poly[2 * e + 1];
      const double bx = poly[2 * ((e + 1) % n)], by = poly[2 * ((e + 1) % n) + 1];
      const bool f0 = ay >= ty, f1 = by >= ty;
      if (f0 == f1) continue;
      const double lhs = (by - ty) * (ax - bx), c = ay - by;
      auto below = [&](double tx) { return lhs >= (bx - tx) * c; };
      // first index whose comparison differs from the row's first entry's (monotone)
      const bool first = below(xs[0]);
      int64_t lo = 0, hi = nx;
      while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (below(xs[mid]) == first) lo = mid + 1; else hi = mid;
      }
      // indices [0, lo) compare as `first`, [lo, nx) as !first; toggle those equal to f1
      if (first == f1) {
        flip[0] ^= 1;
        flip[(size_t)lo] ^= 1;
      } else {
        flip[(size_t)lo] ^= 1;
        flip[(size_t)nx] ^= 1;
      }
    }
    uint8_t in = 0;
    for (int64_t i = 0; i < nx; ++i) {
      in ^= flip[(size_t)i];
      out[(size_t)j * nx + i] = in;
    }
  }
  return JM_OK;
}

// -----------------------------------------------------------------------------------------
// The phenotype generator and tank fill (make_jelly.py:61-402) in host code. Every value is
// computed in double in the reference's operation order with FP contraction off, so that
// results are bit-identical to the numpy evaluation: linspace as numpy forms it
// (i * step + start, the last point set to stop), arange as start + i * ((start + step) -
// start), numpy's scalar ** as C pow, np.gradient with unit spacing, row norms as
// sqrt(x*x + y*y) and the norm of one 2-vector as numpy's dot product forms it,
// sqrt(fma(y, y, x*x)); the nearest spine point of a muscle particle (the k-d tree query of
// make_jelly.py:267-272) by exhaustive search of the 50 points, in squared distance summed
// x then y, the first minimum kept.

namespace {

constexpr double kPayloadW = 0.08;   // make_jelly.py:8
constexpr double kPayloadH = 0.05;   // make_jelly.py:9

struct V2 {
  double x, y;
};

std::vector<double> np_linspace(double a, double b, int64_t n) {
#pragma clang fp contract(off)
  std::vector<double> y((size_t)std::max<int64_t>(n, 0));
  if (n <= 0) return y;
  if (n == 1) {
    y[0] = a;
    return y;
  }
  const double div = (double)(n - 1), delta = b - a, step = delta / div;
  for (int64_t i = 0; i < n; ++i)
    y[(size_t)i] = step == 0.0 ? ((double)i / div) * delta + a : (double)i * step + a;
  y[(size_t)n - 1] = b;
  return y;
}

std::vector<double> np_arange(double a, double b, double step) {
#pragma clang fp contract(off)
  const double len = std::ceil((b - a) / step);
  std::vector<double> y(len > 0 ? (size_t)len : 0);
  if (y.empty()) return y;
  y[0] = a;
  const double delta = (a + step) - a;
  for (size_t i = 1; i < y.size(); ++i) y[i] = a + (double)i * delta;
  return y;
}

// make_jelly.py:61-63, term by term and summed left to right
V2 bezier(V2 p0, V2 p1, V2 p2, V2 p3, double t) {
#pragma clang fp contract(off)
  const double u = 1.0 - t;
  const double c0 = std::pow(u, 3.0), c1 = (3.0 * std::pow(u, 2.0)) * t;
  const double c2 = (3.0 * u) * std::pow(t, 2.0), c3 = std::pow(t, 3.0);
  return {((c0 * p0.x + c1 * p1.x) + c2 * p2.x) + c3 * p3.x,
          ((c0 * p0.y + c1 * p1.y) + c2 * p2.y) + c3 * p3.y};
}

void pip_grid(const std::vector<V2>& poly, const std::vector<double>& xs,
              const std::vector<double>& ys, std::vector<uint8_t>& out) {
  out.resize(xs.size() * ys.size());
  jm_points_in_polygon_grid(&poly[0].x, (int64_t)poly.size(), xs.data(), (int64_t)xs.size(),
                            ys.data(), (int64_t)ys.size(), out.data());
}

std::vector<V2> closed(const std::vector<V2>& a, const std::vector<V2>& b_reversed) {
  std::vector<V2> p(a);
  p.insert(p.end(), b_reversed.rbegin(), b_reversed.rend());
  return p;
}

struct Phenotype {
  std::vector<V2> pos;
  std::vector<int32_t> mat;
  std::vector<float> fib;  // 2 per particle
  bool self_int = false;
};

// make_jelly.py:74-334
Phenotype generate_phenotype(const double* g, V2 offset, int grid_res, bool with_payload,
                             bool axisym) {
#pragma clang fp contract(off)
  Phenotype ph;
  const V2 start = {kPayloadW / 2.0, 0.0};
  const V2 cp1 = {start.x + std::fabs(g[0]), start.y + g[1]};
  const V2 cp2 = {start.x + std::fabs(g[2]), start.y + g[3]};
  const V2 endp = {start.x + std::fabs(g[4]), start.y + g[5]};
  const double t_base = std::fabs(g[6]), t_mid = std::fabs(g[7]), t_tip = std::fabs(g[8]);

  const int S = 50;
  const std::vector<double> ts = np_linspace(0.0, 1.0, S);
  std::vector<V2> spine(S), tan(S), stan(S), nrm(S), outer(S), inner(S), medge(S);
  for (int i = 0; i < S; ++i) spine[i] = bezier(start, cp1, cp2, endp, ts[i]);
  for (int i = 0; i < S; ++i) {
    if (i == 0)
      tan[i] = {(spine[1].x - spine[0].x) / 1.0, (spine[1].y - spine[0].y) / 1.0};
    else if (i == S - 1)
      tan[i] = {(spine[i].x - spine[i - 1].x) / 1.0, (spine[i].y - spine[i - 1].y) / 1.0};
    else
      tan[i] = {(spine[i + 1].x - spine[i - 1].x) / 2.0, (spine[i + 1].y - spine[i - 1].y) / 2.0};
  }
  std::vector<double> half(S);
  for (int i = 0; i < S; ++i) {
    double tl = std::sqrt(tan[i].x * tan[i].x + tan[i].y * tan[i].y);
    if (tl < 1e-10) tl = 1.0;
    stan[i] = {tan[i].x / tl, tan[i].y / tl};
    const V2 n = {-tan[i].y, tan[i].x};
    double nl = std::sqrt(n.x * n.x + n.y * n.y);
    if (nl == 0.0) nl = 1.0;
    nrm[i] = {n.x / nl, n.y / nl};
    // np.interp(ts, [0, 0.5, 1], [t_base, t_mid, t_tip])
    const double x = ts[i];
    double tp;
    if (x == 1.0) {
      tp = t_tip;
    } else {
      const int j = x < 0.5 ? 0 : 1;
      const double xj = j == 0 ? 0.0 : 0.5, f0 = j == 0 ? t_base : t_mid;
      const double f1 = j == 0 ? t_mid : t_tip;
      tp = x == xj ? f0 : ((f1 - f0) / ((j == 0 ? 0.5 : 1.0) - xj)) * (x - xj) + f0;
    }
    half[i] = tp / 2.0;
  }
  const double spacing = 1.0 / (grid_res * 2.0), min_base_muscle = spacing * 1.2;
  const double muscle_ratio = 0.2;
  for (int i = 0; i < S; ++i) {
    outer[i] = {spine[i].x + nrm[i].x * half[i], spine[i].y + nrm[i].y * half[i]};
    inner[i] = {spine[i].x - nrm[i].x * half[i], spine[i].y - nrm[i].y * half[i]};
    const double mt = std::min(min_base_muscle + half[i] * muscle_ratio, half[i] * 0.9);
    const double d = half[i] - mt;
    medge[i] = {spine[i].x - nrm[i].x * d, spine[i].y - nrm[i].y * d};
  }
  const std::vector<V2> body = closed(outer, inner), muscle = closed(medge, inner);
  int32_t si0 = 0, si1 = 0;
  jm_polygon_self_intersects(&body[0].x, (int64_t)body.size(), &si0);
  if (!si0) jm_polygon_self_intersects(&muscle[0].x, (int64_t)muscle.size(), &si1);
  ph.self_int = si0 || si1;

  const int raster = grid_res * 2;
  double bx0 = body[0].x, bx1 = body[0].x, by0 = body[0].y, by1 = body[0].y;
  for (const V2& p : body) {
    bx0 = std::min(bx0, p.x);
    bx1 = std::max(bx1, p.x);
    by0 = std::min(by0, p.y);
    by1 = std::max(by1, p.y);
  }
  const double lo_x = std::min(bx0, -kPayloadW), hi_x = std::max(bx1, kPayloadW);
  const double lo_y = std::min(by0, -0.05), hi_y = std::max(by1, kPayloadH);
  const double pad = 0.05;
  const std::vector<double> xs =
      np_linspace(lo_x - pad, hi_x + pad, (int64_t)((hi_x - lo_x + 2 * pad) * raster));
  const std::vector<double> ys =
      np_linspace(lo_y - pad, hi_y + pad, (int64_t)((hi_y - lo_y + 2 * pad) * raster));
  const size_t nx = xs.size(), nc = xs.size() * ys.size();
  std::vector<uint8_t> in_bell, in_muscle, in_collar(nc, 0), in_collar_mus(nc, 0);
  pip_grid(body, xs, ys, in_bell);
  pip_grid(muscle, xs, ys, in_muscle);
  auto payload_box = [&](size_t c) {
    return xs[c % nx] <= kPayloadW / 2.0 && ys[c / nx] <= kPayloadH;
  };

  // collar around the payload socket (make_jelly.py:164-223)
  const double collar_top = kPayloadH * 0.50, collar_thick = t_base * 0.35;
  const std::vector<double> cts = np_linspace(0.0, 1.0, 20);
  const V2 q0 = {kPayloadW / 2.0, collar_top}, q3 = inner[0];
  const V2 q1 = {kPayloadW / 2.0, collar_top - collar_top * 0.4};
  const V2 q2 = {q3.x + 0.0, q3.y + std::min(collar_top * 0.3, std::fabs(q3.y - q0.y) * 0.4)};
  std::vector<V2> collar_in(20);
  for (int i = 0; i < 20; ++i) collar_in[i] = bezier(q0, q1, q2, q3, cts[i]);
  const V2 outer0 = outer[0];
  double bridge_hw;
  if (outer0.x > kPayloadW / 2.0) {
    const V2 p0 = {kPayloadW / 2.0 + collar_thick, collar_top}, p3 = outer0;
    const V2 p1 = {p0.x, p0.y - collar_top * 0.5};
    const V2 bt = {outer[2].x - outer[0].x, outer[2].y - outer[0].y};
    const double bt_len = std::sqrt(std::fma(bt.y, bt.y, bt.x * bt.x));
    const double reach = collar_top * 0.6;
    const V2 p2 = bt_len > 1e-10 ? V2{p3.x - (bt.x / bt_len) * reach, p3.y - (bt.y / bt_len) * reach}
                                 : V2{p3.x + 0.0, p3.y + 0.01};
    std::vector<V2> collar_out(20), collar_mus(20);
    for (int i = 0; i < 20; ++i) {
      collar_out[i] = bezier(p0, p1, p2, p3, cts[i]);
      collar_mus[i] = {collar_in[i].x + muscle_ratio * (collar_out[i].x - collar_in[i].x),
                       collar_in[i].y + muscle_ratio * (collar_out[i].y - collar_in[i].y)};
    }
    pip_grid(closed(collar_out, collar_in), xs, ys, in_collar);
    pip_grid(closed(collar_mus, collar_in), xs, ys, in_collar_mus);
    for (size_t c = 0; c < nc; ++c)
      if (payload_box(c)) in_collar[c] = in_collar_mus[c] = 0;
    bridge_hw = std::max(outer0.x, kPayloadW / 2.0 + collar_thick);
  } else {
    bridge_hw = kPayloadW / 2.0 + 0.015;
  }
  // transverse bridge (make_jelly.py:225-238)
  if (!with_payload) bridge_hw = std::max(bridge_hw, kPayloadW / 2.0 + 0.02);
  const double bridge_top = with_payload ? 0.0 : kPayloadH;

  // materials, later masks win (make_jelly.py:240-248); the right side (x >= 0)
  std::vector<V2> rpos;
  std::vector<int32_t> rmat;
  for (size_t c = 0; c < nc; ++c) {
    const double x = xs[c % nx], y = ys[c / nx];
    int m = 0;
    if (x >= -bridge_hw && x <= bridge_hw && y >= -0.03 && y <= bridge_top) m = 1;
    if (in_collar[c]) m = 1;
    if (in_collar_mus[c]) m = 3;
    if (in_bell[c]) m = 1;
    if (in_bell[c] && in_muscle[c]) m = 3;
    if (x >= 0 && m > 0) {
      rpos.push_back({x, y});
      rmat.push_back(m);
    }
  }
  const size_t nr = rpos.size();
  std::vector<float> rfib(2 * nr);
  for (size_t q = 0; q < nr; ++q) {
    rfib[2 * q] = 0.0f;
    rfib[2 * q + 1] = 1.0f;
    if (rmat[q] != 3) continue;
    int best = 0;
    double bd = 0.0;
    for (int i = 0; i < S; ++i) {
      const double dx = rpos[q].x - spine[i].x, dy = rpos[q].y - spine[i].y;
      const double d2 = dx * dx + dy * dy;
      if (i == 0 || d2 < bd) {
        bd = d2;
        best = i;
      }
    }
    rfib[2 * q] = (float)stan[best].x;
    rfib[2 * q + 1] = (float)stan[best].y;
  }

  // right half mirrored (Cartesian), or the right half alone (axisymmetric,
  // make_jelly.py:274-296); then the payload block and the spawn offset (make_jelly.py:298-325)
  auto add = [&](V2 p, int32_t m, float f0, float f1) {
    ph.pos.push_back({p.x + offset.x, p.y + offset.y});
    ph.mat.push_back(m);
    ph.fib.push_back(f0);
    ph.fib.push_back(f1);
  };
  for (size_t q = 0; q < nr; ++q) add(rpos[q], rmat[q], rfib[2 * q], rfib[2 * q + 1]);
  if (!axisym)
    for (size_t q = 0; q < nr; ++q)
      add({-rpos[q].x, rpos[q].y}, rmat[q], -rfib[2 * q], rfib[2 * q + 1]);
  if (with_payload) {
    const std::vector<double> px =
        axisym ? np_linspace(0.0, kPayloadW / 2, (int64_t)(kPayloadW / 2 * raster))
               : np_linspace(-kPayloadW / 2, kPayloadW / 2, (int64_t)(kPayloadW * raster));
    const std::vector<double> py = np_linspace(0.0, kPayloadH, (int64_t)(kPayloadH * raster));
    for (double y : py)
      for (double x : px) add({x, y}, 2, 0.0f, 1.0f);
  }
  return ph;
}

}  // namespace

extern "C" {

int jm_generate_phenotype(const double* genome, int32_t n_genes, const double* spawn,
                          int32_t grid_res, int32_t with_payload, int32_t axisym, double* pos,
                          int32_t* mat, float* fiber, int64_t capacity, int64_t* n_out,
                          int32_t* self_intersecting) {
  if (!genome || n_genes < 9 || !spawn || grid_res <= 0 || grid_res > 4096 || !n_out ||
      !self_intersecting || capacity < 0 || (capacity > 0 && (!pos || !mat || !fiber)))
    return fail(JM_ERR_ARG, "jm_generate_phenotype: bad arguments");
  for (int i = 0; i < 9; ++i)
    if (!std::isfinite(genome[i])) return fail(JM_ERR_ARG, "non-finite genome");
  if (!std::isfinite(spawn[0]) || !std::isfinite(spawn[1]))
    return fail(JM_ERR_ARG, "non-finite spawn offset");
  const Phenotype ph =
      generate_phenotype(genome, {spawn[0], spawn[1]}, grid_res, with_payload != 0, axisym != 0);
  const int64_t n = (int64_t)ph.pos.size();
  *n_out = n;
  *self_intersecting = ph.self_int ? 1 : 0;
  if (n > capacity) return fail(JM_ERR_ARG, "jm_generate_phenotype: capacity below *n_out");
  for (int64_t i = 0; i < n; ++i) {
    pos[2 * i] = ph.pos[(size_t)i].x;
    pos[2 * i + 1] = ph.pos[(size_t)i].y;
    mat[i] = ph.mat[(size_t)i];
    fiber[2 * i] = ph.fib[2 * (size_t)i];
    fiber[2 * i + 1] = ph.fib[2 * (size_t)i + 1];
  }
  return JM_OK;
}

int jm_fill_tank(const double* genome, int32_t n_genes, int64_t max_particles, int32_t grid_res,
                 const double* spawn, double water_margin, int32_t with_payload,
                 double domain_height, int32_t axisym, float* pos, int32_t* mat, float* fiber,
                 int64_t* info) {
#pragma clang fp contract(off)
  if (!genome || n_genes < 9 || max_particles < 0 || grid_res <= 0 || grid_res > 4096 ||
      !info || (max_particles > 0 && (!pos || !mat || !fiber)) ||
      !(domain_height > 0.0 && domain_height <= 64.0) || !std::isfinite(water_margin))
    return fail(JM_ERR_ARG, "jm_fill_tank: bad arguments");
  for (int i = 0; i < 9; ++i)
    if (!std::isfinite(genome[i])) return fail(JM_ERR_ARG, "non-finite genome");
  // make_jelly.py:341-343: the default spawn is the axis in axisymmetric mode
  const V2 off = spawn ? V2{spawn[0], spawn[1]} : (axisym ? V2{0.0, 0.40} : V2{0.5, 0.40});
  if (!std::isfinite(off.x) || !std::isfinite(off.y))
    return fail(JM_ERR_ARG, "non-finite spawn offset");
  const Phenotype ph = generate_phenotype(genome, off, grid_res, with_payload != 0, axisym != 0);
  const int64_t n_robot = (int64_t)ph.pos.size();
  if (n_robot > max_particles)
    return fail(JM_ERR_ARG, "jm_fill_tank: the robot alone exceeds max_particles");
  // water lattice, 4 per cell with a 3-spacing margin, minus the exclusion zone around the
  // robot (make_jelly.py:349-366)
  const double spacing = 1.0 / (grid_res * 2.0), margin = spacing * 3;
  const std::vector<double> wx = np_arange(margin, 1.0 - margin, spacing);
  const std::vector<double> wy = np_arange(margin, domain_height - margin, spacing);
  std::vector<double> cand(2 * wx.size() * wy.size());
  size_t c = 0;
  for (double y : wy)
    for (double x : wx) {
      cand[c++] = x;
      cand[c++] = y;
    }
  const int64_t n_cand = (int64_t)(cand.size() / 2);
  std::vector<uint8_t> keep((size_t)n_cand, 1);
  if (n_robot > 0)
    JM_TRY(jm_water_mask(&ph.pos[0].x, n_robot, cand.data(), n_cand, water_margin, keep.data()));
  int64_t n_water = 0;
  for (uint8_t k : keep) n_water += k;
  if (n_robot + n_water > max_particles) n_water = max_particles - n_robot;  // truncate water
  for (int64_t i = 0; i < max_particles; ++i) {
    pos[2 * i] = -1.0f;
    pos[2 * i + 1] = -1.0f;
    mat[i] = -1;
    fiber[2 * i] = 0.0f;
    fiber[2 * i + 1] = 1.0f;
  }
  int64_t muscle = 0, jelly = 0;
  for (int64_t i = 0; i < n_robot; ++i) {
    pos[2 * i] = (float)ph.pos[(size_t)i].x;
    pos[2 * i + 1] = (float)ph.pos[(size_t)i].y;
    mat[i] = ph.mat[(size_t)i];
    fiber[2 * i] = ph.fib[2 * (size_t)i];
    fiber[2 * i + 1] = ph.fib[2 * (size_t)i + 1];
    muscle += ph.mat[(size_t)i] == 3;
    jelly += ph.mat[(size_t)i] == 1;
  }
  int64_t w = n_robot;
  for (int64_t q = 0; q < n_cand && w < n_robot + n_water; ++q) {
    if (!keep[(size_t)q]) continue;
    pos[2 * w] = (float)cand[2 * q];
    pos[2 * w + 1] = (float)cand[2 * q + 1];
    mat[w] = 0;
    ++w;
  }
  info[0] = n_robot;
  info[1] = n_water;
  info[2] = muscle;
  info[3] = jelly;
  info[4] = ph.self_int ? 1 : 0;
  return JM_OK;
}

}  // extern "C"

// -----------------------------------------------------------------------------------------
// Differentiable path (no reference counterpart: the reference has no autodiff).

struct jm_diff {
  int M = 0, nx = 0, ny = 0, G = 0;
  int64_t P = 0;
  int64_t water_start = 0;       // particles [water_start, P) are water (jm_diff_set_water_start)
  int bits = 32;                 // scalar type of the buffers: states, grids, parameters
  int math = 32;                 // scalar type of the arithmetic and the carried adjoint
  jmd::Consts<float> kf{};
  jmd::Consts<double> kd{};
  void* grid = nullptr;   // 3 * M * 3 * G buffer scalars (forward grid ring)
  void* agrid = nullptr;  // 3 * M * 3 * G arithmetic scalars (adjoint grid ring)
  void* sbuf = nullptr;   // 2 * P * 12 buffer scalars (forward states, ping-pong)
  void* pbuf = nullptr;   // F_new (2 x 4P buffer scalars), then adjoint m (2 x 8P) and adjoint
                          // F (2 x 4P) in arithmetic scalars
  ~jm_diff() {
    for (void* p : {grid, agrid, sbuf, pbuf})
      if (p) (void)hipFree(p);
  }
};

namespace {

// mpm_sim.py:12-58 evaluated in double like the reference's Python, rounded to T (for
// float, the same values as make_consts).
template <typename T>
jmd::Consts<T> diff_consts(const jm_config& c, int nx, int ny) {
  const double n_grid = 128.0 * c.quality;
  const double dx = 1.0 / n_grid, inv_dx = n_grid;
  const double dt = 2e-5 / c.quality;
  const double p_vol = (dx * 0.5) * (dx * 0.5);
  const double E_base = 0.1e4, nu_base = 0.2, E_pay = 2.0e5, nu_pay = 0.2;
  jmd::Consts<T> k;
  k.nx = nx;
  k.ny = ny;
  k.damp_x = nx / 20;
  k.damp_y = ny / 20;
  k.dt = T(dt);
  k.dx = T(dx);
  k.inv_dx = T(inv_dx);
  k.p_mass = T(p_vol);
  k.gravity_dt = T(dt * 10.0);
  k.stress_scale = T(-dt * p_vol * 4 * inv_dx * inv_dx);
  k.apic_scale = T(4 * inv_dx);
  k.mu_base = T(E_base / (2 * (1 + nu_base)));
  k.la_base = T(E_base * nu_base / ((1 + nu_base) * (1 - 2 * nu_base)));
  k.mu_pay = T(E_pay / (2 * (1 + nu_pay)));
  k.la_pay = T(E_pay * nu_pay / ((1 + nu_pay) * (1 - 2 * nu_pay)));
  k.water_la = T(100000.0);
  k.act_freq = T(1.0);
  k.x_lo = T(0.001);
  k.x_hi = T(0.999);
  k.y_hi = T(c.domain_height - 0.001);
  k.sig_floor = T(1e-6);
  k.pay_lo = T(1 - 5e-3);
  k.pay_hi = T(1 + 5e-3);
  k.pi = T(3.14159265);
  k.nu_jelly = T(0.3);  // mpm_sim.py:32
  return k;
}

// scalar_bits of jm_diff_create: the buffer and arithmetic scalar widths.
bool diff_precision(int32_t scalar_bits, int* bits, int* math) {
  if (scalar_bits == 32 || scalar_bits == 64) {
    *bits = *math = scalar_bits;
    return true;
  }
  if (scalar_bits == JM_DIFF_MIXED) {
    *bits = 32;
    *math = 64;
    return true;
  }
  return false;
}

int diff_check(jm_diff* d, const jm_diff_inputs* in) {
  if (!d) return fail(JM_ERR_ARG, "null diff handle");
  if (!in || !in->inst || !in->mat || !in->fiber || !in->amp_scale || !in->params)
    return fail(JM_ERR_ARG, "jm_diff_inputs has a null pointer");
  return JM_OK;
}

template <typename T>
jmd::DiffInputs<T> diff_inputs(const jm_diff_inputs* in) {
  return {in->inst, in->mat, static_cast<const T*>(in->fiber),
          static_cast<const T*>(in->amp_scale), static_cast<const T*>(in->params)};
}

// D: jm_diff, or a DiffScratch view (same members) over caller-free device scratch.
template <typename T, typename D>
const jmd::Consts<T>& consts_of(const D* d) {
  if constexpr (sizeof(T) == 4) return d->kf;
  else return d->kd;
}

// T: buffer scalar, A: arithmetic scalar (see diff_kernels.h). The clock advances in T, as
// the reference's fp32 sim_time does in its fp32 build (mpm_sim.py:340).
template <typename T, typename A, typename D>
int fwd_block(D* d, const jm_diff_inputs* in, const T* state_in, double t0_in,
              int nsteps, T* out, T* grid_out, hipStream_t st) {
  const jmd::DiffInputs<T> din = diff_inputs<T>(in);
  const jmd::Consts<A>& k = consts_of<A>(d);
  const T dt = consts_of<T>(d).dt;
  const int P = (int)d->P, G = d->G, n = nsteps;
  const size_t S = (size_t)P * 12;
  const int64_t MG = (int64_t)d->M * 3 * G;  // one substep's grids
  const dim3 pg(blocks_for(P, jmd::kDiffBlock)), tb(jmd::kDiffBlock);
  T* ring = static_cast<T*>(d->grid);
  T* sbuf = static_cast<T*>(d->sbuf);
  T* fnew[2] = {static_cast<T*>(d->pbuf), static_cast<T*>(d->pbuf) + 4 * (size_t)P};
  const bool store = grid_out != nullptr;
  // grid of substep s; state s (s >= 1) as the kernel that finishes substep s-1 writes it
  auto grid = [&](int s) { return store ? grid_out + (size_t)s * MG : ring + (size_t)(s % 3) * MG; };
  auto state = [&](int s) -> T* {
    if (store) return out + (size_t)(s - 1) * S;
    return s == n ? out : sbuf + (size_t)(s % 2) * S;
  };
  if (store)
    hipLaunchKernelGGL(jmd::k_d_fill<T>, dim3(blocks_for(MG * n, 256 * 8)), dim3(256), 0, st,
                       grid_out, MG * n);
  else
    hipLaunchKernelGGL(jmd::k_d_fill<T>, dim3(blocks_for(MG, 256 * 8)), dim3(256), 0, st, ring,
                       MG);
  T t = T(t0_in);
  // K_f(0): P2G of substep 0
  hipLaunchKernelGGL((jmd::k_d_fwd<T, A, false, true>), pg, tb, 0, st, din, state_in,
                     (const T*)nullptr, (const T*)nullptr, (T*)nullptr, fnew[0], grid(0),
                     (!store && n > 1) ? ring + (size_t)MG : (T*)nullptr, MG, P, G, A(t), k);
  for (int s = 1; s < n; ++s) {
    t = t + dt;  // mpm_sim.py:340 (fp32 in the reference)
    // K_f(s): G2P of substep s-1, P2G of substep s
    hipLaunchKernelGGL((jmd::k_d_fwd<T, A, true, true>), pg, tb, 0, st, din,
                       (const T*)(s == 1 ? state_in : state(s - 1)), (const T*)fnew[(s - 1) % 2],
                       (const T*)grid(s - 1), state(s), fnew[s % 2], grid(s),
                       (!store && s + 1 < n) ? ring + (size_t)((s + 1) % 3) * MG : (T*)nullptr,
                       MG, P, G, A(t), k);
  }
  // K_f(n): G2P of substep n-1
  hipLaunchKernelGGL((jmd::k_d_fwd<T, A, true, false>), pg, tb, 0, st, din,
                     (const T*)(n == 1 ? state_in : state(n - 1)), (const T*)fnew[(n - 1) % 2],
                     (const T*)grid(n - 1), state(n), (T*)nullptr, (T*)nullptr, (T*)nullptr,
                     MG, P, G, A(t), k);
  HIP_TRY(hipGetLastError());
  return JM_OK;
}

// One adjoint kernel of a block (template flags P2GB, G2PB): the general kernel over the
// particles before the water range (rounded up to whole workgroups) and the water-only
// instantiation over the rest, one after the other on the stream; the water launch, the
// larger, clears the zero span.
template <typename T, typename A, bool P2GB, bool G2PB, typename D>
void launch_bwd(const D* d, hipStream_t st, const jmd::DiffInputs<T>& din, const T* st_cur,
                const T* st_prev, const A* aF_in, const A* m_in, const A* ag_cur,
                const T* g_prev, A* ag_prev, const T* a_full_in, A* m_out, A* aF_out,
                T* a_full_out, A* zero, int64_t zero_n, A t, const jmd::Consts<A>& k,
                T* gamp, double* gparams) {
  const int P = (int)d->P, G = d->G;
  const int64_t ws = std::min<int64_t>(d->water_start, d->P);
  const int W = (int)std::min<int64_t>(P, (ws + jmd::kDiffBlock - 1) / jmd::kDiffBlock *
                                              jmd::kDiffBlock);
  const dim3 tb(jmd::kDiffBlock);
  if (W > 0)
    hipLaunchKernelGGL((jmd::k_d_bwd<T, A, P2GB, G2PB, false>), dim3(blocks_for(W, jmd::kDiffBlock)),
                       tb, 0, st, din, st_cur, st_prev, aF_in, 4, m_in, ag_cur, g_prev, ag_prev,
                       a_full_in, m_out, aF_out, a_full_out, W < P ? (A*)nullptr : zero, zero_n,
                       0, W, G, t, k, gamp, gparams);
  if (W < P)
    hipLaunchKernelGGL((jmd::k_d_bwd<T, A, P2GB, G2PB, true>),
                       dim3(blocks_for(P - W, jmd::kDiffBlock)), tb, 0, st, din, st_cur, st_prev,
                       aF_in, 4, m_in, ag_cur, g_prev, ag_prev, a_full_in, m_out, aF_out,
                       a_full_out, zero, zero_n, W, P, G, t, k, gamp, gparams);
}

template <typename T, typename A, typename D>
int bwd_block(D* d, const jm_diff_inputs* in, const T* traj, const T* grids,
              double t0_in, int nsteps, const T* grad_out, T* grad_in, T* grad_amp,
              double* grad_params, hipStream_t st) {
  const jmd::DiffInputs<T> din = diff_inputs<T>(in);
  const jmd::Consts<A>& k = consts_of<A>(d);
  const T dt = consts_of<T>(d).dt;
  const int P = (int)d->P, n = nsteps;
  const size_t S = (size_t)P * 12;
  const int64_t MG = (int64_t)d->M * 3 * d->G;
  A* ring = static_cast<A*>(d->agrid);
  A* pb = reinterpret_cast<A*>(static_cast<char*>(d->pbuf) + 8 * (size_t)P * sizeof(T));
  A* mb[2] = {pb, pb + 8 * (size_t)P};                        // m_s, by s % 2
  A* aF[2] = {pb + 16 * (size_t)P, pb + 20 * (size_t)P};      // F part of a_s, by s % 2
  auto agrid = [&](int s) { return ring + (size_t)(s % 3) * MG; };
  std::vector<T> times(n + 1);
  times[0] = T(t0_in);
  for (int s = 0; s < n; ++s) times[s + 1] = times[s] + dt;
  hipLaunchKernelGGL(jmd::k_d_fill<A>, dim3(blocks_for(MG, 256 * 8)), dim3(256), 0, st,
                     agrid(n - 1), MG);
  // K_b(n): G2P adjoint of substep n-1, driven by grad_out
  launch_bwd<T, A, false, true>(d, st, din, (const T*)nullptr, traj + (size_t)(n - 1) * S,
                                (const A*)nullptr, (const A*)nullptr, (const A*)nullptr,
                                grids + (size_t)(n - 1) * MG, agrid(n - 1), grad_out,
                                mb[(n - 1) % 2], aF[n % 2], (T*)nullptr,
                                n >= 2 ? agrid(n + 1) : (A*)nullptr, MG, A(times[n]), k,
                                grad_amp, grad_params);
  for (int s = n - 1; s >= 1; --s) {
    // K_b(s): P2G adjoint of substep s, G2P adjoint of substep s-1
    launch_bwd<T, A, true, true>(d, st, din, traj + (size_t)s * S, traj + (size_t)(s - 1) * S,
                                 (const A*)aF[(s + 1) % 2], (const A*)mb[s % 2],
                                 (const A*)agrid(s), grids + (size_t)(s - 1) * MG,
                                 agrid(s - 1), (const T*)nullptr, mb[(s - 1) % 2], aF[s % 2],
                                 (T*)nullptr, s >= 2 ? agrid(s + 1) : (A*)nullptr, MG,
                                 A(times[s]), k, grad_amp, grad_params);
  }
  // K_b(0): P2G adjoint of substep 0, writing grad_in
  launch_bwd<T, A, true, false>(d, st, din, traj, (const T*)nullptr, (const A*)aF[1],
                                (const A*)mb[0], (const A*)agrid(0), (const T*)nullptr,
                                (A*)nullptr, (const T*)nullptr, (A*)nullptr, (A*)nullptr,
                                grad_in, (A*)nullptr, MG, A(times[0]), k, grad_amp,
                                grad_params);
  HIP_TRY(hipGetLastError());
  return JM_OK;
}

// ---- fixed device scratch of jm_fwd_block_noalloc ---------------------------------------
// Every GPU the library runs on holds one copy of this array until the process ends. Host
// code never passes its name to a kernel (on the host the name is a stand-in, not a device
// address): the entry point looks up each device's address once, with hipGetSymbolAddress,
// and uses that.
static __device__ __align__(256) unsigned char scratch[JM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

// The current device's address of `scratch`, looked up on the first call on that device.
int noalloc_scratch(void** out) {
  static std::mutex mu;
  static void* ptrs[kMaxDevices] = {nullptr};
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(JM_ERR_STATE, "device index beyond 63");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(JM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ptrs[dev] = p;
  }
  *out = ptrs[dev];
  return JM_OK;
}

// The buffers fwd_block reads from a jm_diff, here carved from `scratch` (no destructor:
// nothing is owned).
struct DiffScratch {
  int M = 0, G = 0;
  int64_t P = 0;
  jmd::Consts<float> kf{};
  jmd::Consts<double> kd{};
  void* grid = nullptr;  // 3 * M * 3 * G scalars: the forward grid ring
  void* sbuf = nullptr;  // 2 * P * 12 scalars: states between kernels (ping-pong)
  void* pbuf = nullptr;  // 8 * P scalars: F_new of the last two substeps
};

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int jm_diff_create(const jm_config* cfg, int64_t n_particles, int32_t scalar_bits,
                   jm_diff** out) {
  if (!cfg || !out) return fail(JM_ERR_ARG, "null argument");
  *out = nullptr;
  if (cfg->abi_version != JM_ABI_VERSION) return fail(JM_ERR_ARG, "abi_version mismatch");
  if (cfg->n_instances <= 0 || cfg->quality <= 0 || n_particles <= 0 ||
      n_particles > (int64_t)INT32_MAX / 16)
    return fail(JM_ERR_ARG, "bad sizes");
  int bits = 0, math = 0;
  if (!diff_precision(scalar_bits, &bits, &math))
    return fail(JM_ERR_ARG, "scalar_bits: 32, 64 or JM_DIFF_MIXED");
  if (cfg->axisym) return fail(JM_ERR_ARG, "the differentiable path has no axisymmetric variant");
  const int nx = 128 * cfg->quality;
  const int ny = cfg->n_grid_y > 0 ? cfg->n_grid_y : nx;
  if (ny < 8) return fail(JM_ERR_ARG, "n_grid_y too small");
  jm_diff* d = new jm_diff();
  d->M = cfg->n_instances;
  d->nx = nx;
  d->ny = ny;
  d->G = nx * ny;
  d->P = n_particles;
  d->water_start = n_particles;
  d->bits = bits;
  d->math = math;
  d->kf = diff_consts<float>(*cfg, nx, ny);
  d->kd = diff_consts<double>(*cfg, nx, ny);
  const size_t es = bits / 8, ea = math / 8;
  const size_t grid_nodes = 3 * (size_t)d->M * 3 * d->G;
  const size_t sbytes = (size_t)d->P * 12 * es;
  if (hipMalloc(&d->grid, grid_nodes * es) != hipSuccess ||
      hipMalloc(&d->agrid, grid_nodes * ea) != hipSuccess ||
      hipMalloc(&d->sbuf, 2 * sbytes) != hipSuccess ||
      hipMalloc(&d->pbuf, 8 * (size_t)d->P * es + 24 * (size_t)d->P * ea) != hipSuccess) {
    delete d;
    return fail(JM_ERR_NOMEM, "hipMalloc(diff buffers)");
  }
  *out = d;
  return JM_OK;
}

int jm_diff_destroy(jm_diff* d) {
  delete d;
  return JM_OK;
}

int jm_diff_set_water_start(jm_diff* d, int64_t water_start) {
  if (!d) return fail(JM_ERR_ARG, "null diff handle");
  if (water_start < 0 || water_start > d->P)
    return fail(JM_ERR_ARG, "water_start must be in [0, n_particles]");
  d->water_start = water_start;
  return JM_OK;
}

int jm_fwd_block(jm_diff* d, const jm_diff_inputs* in, const void* state_in, double t0,
                 int32_t nsteps, void* out, void* grid_out, void* stream) {
  JM_TRY(diff_check(d, in));
  if (!state_in || !out || nsteps < 1) return fail(JM_ERR_ARG, "bad state or step count");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (d->bits == 64)
    return fwd_block<double, double>(d, in, static_cast<const double*>(state_in), t0, nsteps,
                                     static_cast<double*>(out),
                                     static_cast<double*>(grid_out), st);
  if (d->math == 64)
    return fwd_block<float, double>(d, in, static_cast<const float*>(state_in), t0, nsteps,
                                    static_cast<float*>(out), static_cast<float*>(grid_out),
                                    st);
  return fwd_block<float, float>(d, in, static_cast<const float*>(state_in), t0, nsteps,
                                 static_cast<float*>(out), static_cast<float*>(grid_out), st);
}

int jm_fwd_block_noalloc(const jm_config* cfg, int64_t n_particles, int32_t scalar_bits,
                         const jm_diff_inputs* in, const void* state_in, double t0,
                         int32_t nsteps, void* out, void* stream) {
  if (!cfg || !in) return fail(JM_ERR_ARG, "null argument");
  if (cfg->abi_version != JM_ABI_VERSION) return fail(JM_ERR_ARG, "abi_version mismatch");
  if (cfg->n_instances <= 0 || cfg->quality <= 0 || n_particles <= 0 ||
      n_particles > (int64_t)INT32_MAX / 16)
    return fail(JM_ERR_ARG, "bad sizes");
  int bits = 0, math = 0;
  if (!diff_precision(scalar_bits, &bits, &math))
    return fail(JM_ERR_ARG, "scalar_bits: 32, 64 or JM_DIFF_MIXED");
  if (cfg->axisym) return fail(JM_ERR_ARG, "the differentiable path has no axisymmetric variant");
  if (!in->inst || !in->mat || !in->fiber || !in->amp_scale || !in->params)
    return fail(JM_ERR_ARG, "jm_diff_inputs has a null pointer");
  if (!state_in || !out || nsteps < 1) return fail(JM_ERR_ARG, "bad state or step count");
  const int nx = 128 * cfg->quality;
  const int ny = cfg->n_grid_y > 0 ? cfg->n_grid_y : nx;
  if (ny < 8) return fail(JM_ERR_ARG, "n_grid_y too small");
  DiffScratch d;
  d.M = cfg->n_instances;
  d.G = nx * ny;
  d.P = n_particles;
  const size_t es = bits / 8;
  const size_t gbytes = align256(3 * (size_t)d.M * 3 * d.G * es);
  const size_t sbytes = align256(2 * (size_t)d.P * 12 * es);
  const size_t pbytes = align256(8 * (size_t)d.P * es);
  if (gbytes + sbytes + pbytes > (size_t)JM_NOALLOC_SCRATCH_BYTES)
    return fail(JM_ERR_ARG, "jm_fwd_block_noalloc: needs " +
                                std::to_string(gbytes + sbytes + pbytes) +
                                " bytes of scratch (9*M*nx*ny + 32*P scalars), beyond its " +
                                std::to_string((size_t)JM_NOALLOC_SCRATCH_BYTES) +
                                "; use jm_diff_create + jm_fwd_block");
  void* base = nullptr;
  JM_TRY(noalloc_scratch(&base));
  char* p = static_cast<char*>(base);
  d.grid = p;
  d.sbuf = p + gbytes;
  d.pbuf = p + gbytes + sbytes;
  hipStream_t st = static_cast<hipStream_t>(stream);
  d.kf = diff_consts<float>(*cfg, nx, ny);
  d.kd = diff_consts<double>(*cfg, nx, ny);
  if (bits == 64)
    return fwd_block<double, double>(&d, in, static_cast<const double*>(state_in), t0, nsteps,
                                     static_cast<double*>(out), (double*)nullptr, st);
  if (math == 64)
    return fwd_block<float, double>(&d, in, static_cast<const float*>(state_in), t0, nsteps,
                                    static_cast<float*>(out), (float*)nullptr, st);
  return fwd_block<float, float>(&d, in, static_cast<const float*>(state_in), t0, nsteps,
                                 static_cast<float*>(out), (float*)nullptr, st);
}

int jm_bwd_block(jm_diff* d, const jm_diff_inputs* in, const void* traj, const void* grids,
                 double t0, int32_t nsteps, const void* grad_out, void* grad_in, void* grad_amp,
                 double* grad_params, void* stream) {
  JM_TRY(diff_check(d, in));
  if (!traj || !grids || !grad_out || !grad_in || !grad_amp || !grad_params || nsteps < 1)
    return fail(JM_ERR_ARG, "bad backward arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (d->bits == 64)
    return bwd_block<double, double>(d, in, static_cast<const double*>(traj),
                                     static_cast<const double*>(grids), t0, nsteps,
                                     static_cast<const double*>(grad_out),
                                     static_cast<double*>(grad_in),
                                     static_cast<double*>(grad_amp), grad_params, st);
  if (d->math == 64)
    return bwd_block<float, double>(d, in, static_cast<const float*>(traj),
                                    static_cast<const float*>(grids), t0, nsteps,
                                    static_cast<const float*>(grad_out),
                                    static_cast<float*>(grad_in),
                                    static_cast<float*>(grad_amp), grad_params, st);
  return bwd_block<float, float>(d, in, static_cast<const float*>(traj),
                                 static_cast<const float*>(grids), t0, nsteps,
                                 static_cast<const float*>(grad_out),
                                 static_cast<float*>(grad_in), static_cast<float*>(grad_amp),
                                 grad_params, st);
}

}  // extern "C"
