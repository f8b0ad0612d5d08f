"""The fp64 host build of the engine's adjoint math (tests/adjoint_harness.cpp including
jellyfih_amd/csrc/adjoint.h) and the small perturbed-state cases the adjoint tests share.
Test infrastructure: built with g++ into tests/_build/."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "_build")
SRC = os.path.join(ROOT, "tests", "adjoint_harness.cpp")


def build_harness(force: bool = False, fp32: bool = False) -> str:
    """Compile the harness (fp64, or fp32 arithmetic with ``fp32``) into tests/_build."""
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, "libadjharness32.so" if fp32 else "libadjharness.so")
    deps = [SRC, os.path.join(ROOT, "jellyfih_amd", "csrc", "adjoint.h")]
    if force or not os.path.exists(out) or \
            os.path.getmtime(out) < max(os.path.getmtime(d) for d in deps):
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                        *(["-DHARNESS_FP32"] if fp32 else []), "-o", out + ".tmp", SRC],
                       check=True)
        os.replace(out + ".tmp", out)
    return out


def load_harness(fp32: bool = False):
    L = ctypes.CDLL(build_harness(fp32=fp32))
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int32)
    L.adj_forward.argtypes = [ctypes.POINTER(oracle.OracleConsts), ctypes.c_int, ctypes.c_int,
                              I, I, D, D, D, D, ctypes.c_double, ctypes.c_int, D]
    L.adj_backward.argtypes = [ctypes.POINTER(oracle.OracleConsts), ctypes.c_int, ctypes.c_int,
                               I, I, D, D, D, D, ctypes.c_double, ctypes.c_int, D, D, D, D]
    return L


def _p(a, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class Case:
    """M instances of N all-live particles of every material, perturbed state."""

    def __init__(self, seed, M=2, N=160, steps=12, t0=0.05):
        rng = np.random.default_rng(seed)
        self.c = oracle.reference_consts()
        self.M, self.N, self.P, self.steps, self.t0 = M, N, M * N, steps, t0
        pos = np.concatenate([rng.uniform([0.40, 0.40], [0.47, 0.46], size=(N, 2))
                              for _ in range(M)])
        self.mat = rng.choice([0, 0, 1, 2, 3, 3], size=self.P).astype(np.int32)
        ang = rng.uniform(0, 2 * np.pi, self.P)
        self.fiber = np.stack([np.cos(ang), np.sin(ang)], 1)
        self.inst = np.repeat(np.arange(M), N).astype(np.int32)
        self.state = np.zeros((self.P, 12))
        self.state[:, 0:2] = pos
        self.state[:, 2:4] = rng.normal(size=(self.P, 2)) * 0.05
        self.state[:, 4:8] = rng.normal(size=(self.P, 4)) * 0.5
        F = np.eye(2) + rng.normal(size=(self.P, 2, 2)) * 0.003
        self.state[:, 8:12] = F.reshape(self.P, 4)
        self.amp_scale = rng.uniform(0.5, 1.5, self.P)
        # amp, contraction, freq, payload density, E_jelly
        self.params = np.tile([500.0, 0.25, 1.1, 2.5, 700.0], (M, 1))
        self.params[:, 0] *= rng.uniform(0.8, 1.2, M)
        self.params[:, 4] *= rng.uniform(0.8, 1.2, M)

    @classmethod
    def from_arrays(cls, inst, mat, fiber, state, amp_scale, params, steps, t0=0.0):
        """A case over given packed particles (``inst``, ``mat``, ``fiber`` per particle,
        ``state`` (P, 12), ``params`` (M, 5))."""
        cs = cls.__new__(cls)
        cs.c = oracle.reference_consts()
        cs.inst = np.ascontiguousarray(inst, np.int32)
        cs.mat = np.ascontiguousarray(mat, np.int32)
        cs.fiber = np.ascontiguousarray(fiber, np.float64)
        cs.state = np.ascontiguousarray(state, np.float64)
        cs.amp_scale = np.ascontiguousarray(amp_scale, np.float64)
        cs.params = np.ascontiguousarray(params, np.float64)
        cs.M, cs.P = cs.params.shape[0], cs.state.shape[0]
        cs.N = cs.P // cs.M
        cs.steps, cs.t0 = steps, t0
        return cs

    def round_f32(self):
        """Round every input to fp32 so an fp32 engine and this fp64 case start alike."""
        for name in ("fiber", "state", "amp_scale", "params"):
            setattr(self, name, getattr(self, name).astype(np.float32).astype(np.float64))
        self.t0 = float(np.float32(self.t0))
        return self

    def forward(self, L, state=None, amp_scale=None, params=None):
        state = self.state if state is None else state
        amp_scale = self.amp_scale if amp_scale is None else amp_scale
        params = self.params if params is None else params
        traj = np.zeros((self.steps + 1, self.P, 12))
        L.adj_forward(ctypes.byref(self.c), self.M, self.P, _p(self.inst, ctypes.c_int32),
                      _p(self.mat, ctypes.c_int32), _p(np.ascontiguousarray(self.fiber)),
                      _p(np.ascontiguousarray(amp_scale)), _p(np.ascontiguousarray(params)),
                      _p(np.ascontiguousarray(state)), self.t0, self.steps, _p(traj))
        return traj

    def backward(self, L, traj, gout):
        gin = np.zeros((self.P, 12))
        gamp = np.zeros(self.P)
        gpp = np.zeros((self.P, 4))
        L.adj_backward(ctypes.byref(self.c), self.M, self.P, _p(self.inst, ctypes.c_int32),
                       _p(self.mat, ctypes.c_int32), _p(np.ascontiguousarray(self.fiber)),
                       _p(self.amp_scale), _p(self.params), _p(traj), self.t0, self.steps,
                       _p(np.ascontiguousarray(gout)), _p(gin), _p(gamp), _p(gpp))
        gpar = np.zeros((self.M, 5))
        for m in range(self.M):
            sel = self.inst == m
            gpar[m, 4] = gpp[sel, 0].sum()
            gpar[m, 0] = gpp[sel, 1].sum()
            gpar[m, 1] = gpp[sel, 2].sum()
            gpar[m, 2] = gpp[sel, 3].sum()
        return gin, gamp, gpar

    def oracle_forward(self, state=None, params=None, dtype=np.float64):
        """The oracle's substeps (instances of equal size N, amp_scale taken as 1)."""
        state = self.state if state is None else state
        params = self.params if params is None else params
        o = oracle.OracleSim(self.M, self.N, self.c, dtype=dtype)
        o.x[...] = state[:, 0:2].reshape(self.M, self.N, 2)
        o.v[...] = state[:, 2:4].reshape(self.M, self.N, 2)
        o.C[...] = state[:, 4:8].reshape(self.M, self.N, 2, 2)
        o.F[...] = state[:, 8:12].reshape(self.M, self.N, 2, 2)
        o.material[...] = self.mat.reshape(self.M, self.N)
        o.fiber[...] = self.fiber.reshape(self.M, self.N, 2)
        o.Jp[...] = 1
        o.amp[:] = params[:, 0]
        o.contraction[:] = params[:, 1]
        o.freq[:] = params[:, 2]
        o.pdens[:] = params[:, 3]
        o.set_e_jelly(params[:, 4])
        o.t[0] = self.t0
        o.run(self.steps)
        return np.concatenate([o.x.reshape(-1, 2), o.v.reshape(-1, 2), o.C.reshape(-1, 4),
                               o.F.reshape(-1, 4)], 1)


# --- the reference-tank gradient case (one seeded tank, payload CoM y objective) ----------

def reference_tank_case(steps: int, seed: int = 0):
    """One seeded reference tank (fill_tank of genome ``seed``, 80K slots) packed to its
    live particles, with the genome's contraction and frequency, amp 500, pdens 2.5,
    E_jelly 700, all inputs rounded to fp32. Returns (case, pos, mat, fiber) with the
    slot-layout arrays for the oracle."""
    from tests.cases import reference_tank
    pos, mat, fib, g, info = reference_tank(seed)
    live = mat >= 0
    P = int(live.sum())
    state = np.zeros((P, 12))
    state[:, 0:2] = pos[live]
    state[:, 8] = state[:, 11] = 1.0
    params = np.array([[500.0, np.clip(g[9], 0.05, 0.6), np.clip(g[10], 0.5, 2.0), 2.5, 700.0]])
    cs = Case.from_arrays(np.zeros(P, np.int32), mat[live], fib[live], state, np.ones(P),
                          params, steps=steps).round_f32()
    return cs, pos, mat, fib


GRAD_KINDS = ("v", "v", "amp", "contraction", "freq", "E", "v+params", "v+params")
PARAM_SCALES = np.array([500.0, 0.2, 1.0, 0.0, 700.0])   # amp, contraction, freq, pdens, E


def gradient_directions(mat_live, seed: int = 2024):
    """Eight seeded directions: initial velocities of the body particles (jelly, payload,
    muscle; N(0, 0.01) m/s each), single parameters, and mixtures. Water starts at rest
    against the walls, where the wall clamps (mpm_sim.py:305-309) have a kink at exactly
    zero velocity, so its initial velocity has no unique derivative and is left out.
    Returns (kinds, dvs (P,2) list, dps (5,) list)."""
    rng = np.random.default_rng(seed)
    body = (np.asarray(mat_live) > 0)[:, None]
    P = body.shape[0]
    dvs, dps = [], []
    for kind in GRAD_KINDS:
        dv = np.zeros((P, 2))
        dp = np.zeros(5)
        if kind.startswith("v"):
            dv = rng.normal(size=(P, 2)) * 0.01 * body
        if kind in ("amp", "contraction", "freq", "E"):
            dp[["amp", "contraction", "freq", "-", "E"].index(kind)] = 1.0
        if kind == "v+params":
            dp = rng.normal(size=5)
        dvs.append(dv)
        dps.append(dp * PARAM_SCALES)
    return list(GRAD_KINDS), dvs, dps


def amp_scale_direction(mat_live, seed: int = 11):
    """A seeded direction in the per-particle actuation multiplier (muscle particles)."""
    rng = np.random.default_rng(seed)
    return rng.normal(size=len(mat_live)) * (np.asarray(mat_live) == 3)


def oracle_payload_fd(cs, pos, mat, fib, dvs, dps, eps_list, nthreads=None):
    """fp64 central differences of the oracle's payload CoM y after ``cs.steps`` substeps,
    along each (dv, dp) direction and each eps; all runs are instances of one oracle call.
    Returns an array (n_directions, n_eps)."""
    live = mat >= 0
    K, E = len(dvs), len(eps_list)
    R = 2 * K * E
    o = oracle.OracleSim(R, mat.shape[0], cs.c, dtype=np.float64)
    base = cs.params[0]
    Es = np.empty(R)
    r = 0
    for k in range(K):
        for eps in eps_list:
            for sg in (1.0, -1.0):
                o.load(r, pos, mat, fib)
                o.v[r][live] += sg * eps * dvs[k]
                p = base + sg * eps * dps[k]
                o.amp[r], o.contraction[r], o.freq[r], o.pdens[r] = p[0], p[1], p[2], p[3]
                Es[r] = p[4]
                r += 1
    o.set_e_jelly(Es)
    o.t[:] = cs.t0
    o.run(cs.steps, **({"nthreads": nthreads} if nthreads else {}))
    pay = mat == 2
    ys = np.array([o.x[i][pay][:, 1].mean() for i in range(R)]).reshape(K, E, 2)
    return (ys[:, :, 0] - ys[:, :, 1]) / (2 * np.asarray(eps_list))[None, :]


def harness_amp_scale_fd(L, cs, d, eps_list):
    """fp64 central differences of the harness forward's payload CoM y along the
    per-particle actuation multiplier direction ``d``."""
    pay = cs.mat == 2
    out = []
    for eps in eps_list:
        v = [cs.forward(L, amp_scale=cs.amp_scale + sg * eps * d)[-1][pay, 1].mean()
             for sg in (1, -1)]
        out.append((v[0] - v[1]) / (2 * eps))
    return np.array(out)


def harness_payload_gradient(L, cs, block: int = 100):
    """fp64 adjoint of the payload CoM y after ``cs.steps`` substeps (harness build of the
    engine's adjoint math), with checkpoints every ``block`` substeps so that memory stays
    at one block's trajectory. Returns (grad state (P,12), grad amp_scale (P,),
    grad params (M,5))."""
    pay = cs.mat == 2
    n_blocks = (cs.steps + block - 1) // block
    starts, t0s = [cs.state.copy()], [cs.t0]
    steps_total, t0_total = cs.steps, cs.t0
    t = cs.t0
    try:
        for b in range(n_blocks):
            n = min(block, steps_total - b * block)
            cs.steps, cs.t0 = n, t0s[-1]
            tr = cs.forward(L, state=starts[-1])
            for _ in range(n):
                t = t + cs.c.dt
            if b + 1 < n_blocks:
                starts.append(tr[-1].copy())
                t0s.append(t)
        gout = np.zeros((cs.P, 12))
        gout[pay, 1] = 1.0 / pay.sum()
        gamp = np.zeros(cs.P)
        gpp = np.zeros((cs.P, 4))
        for b in reversed(range(n_blocks)):
            n = min(block, steps_total - b * block)
            cs.steps, cs.t0 = n, t0s[b]
            tr = cs.forward(L, state=starts[b])
            gin = np.zeros((cs.P, 12))
            L.adj_backward(ctypes.byref(cs.c), cs.M, cs.P, _p(cs.inst, ctypes.c_int32),
                           _p(cs.mat, ctypes.c_int32), _p(np.ascontiguousarray(cs.fiber)),
                           _p(cs.amp_scale), _p(cs.params), _p(tr), cs.t0, n,
                           _p(np.ascontiguousarray(gout)), _p(gin), _p(gamp), _p(gpp))
            gout = gin
    finally:
        cs.steps, cs.t0 = steps_total, t0_total
    gpar = np.zeros((cs.M, 5))
    for m in range(cs.M):
        sel = cs.inst == m
        gpar[m, 4] = gpp[sel, 0].sum()
        gpar[m, 0] = gpp[sel, 1].sum()
        gpar[m, 1] = gpp[sel, 2].sum()
        gpar[m, 2] = gpp[sel, 3].sum()
    return gout, gamp, gpar
