"""Differentiable batched substeps: reverse-mode gradients through the MLS-MPM substep.

The reference simulator (``mpm_sim.py:153-340``) has no autodiff. This module puts the same
substep behind ``torch.autograd``: the forward and the adjoint are HIP kernels in
``libjellymx.so`` (``jm_fwd_block`` / ``jm_bwd_block``), and :class:`_Block` is one
``torch.autograd.Function`` per block of ``checkpoint`` substeps. A block keeps only its
input state (the checkpoint); its backward re-runs the block's forward storing every
substep's state, then runs the adjoint substeps in reverse.

State layout, one row of 12 fp32 per live particle:
``x0 x1 v0 v1 C00 C01 C10 C11 F00 F01 F10 F11``. Particles of all instances share one
array; ``inst`` gives each particle's instance. Dead slots are dropped when a tank is
packed, so every row is a live particle.

Differentiable inputs: the state (initial positions, velocities, C, F), the per-particle
actuation multiplier ``amp_scale`` (muscle particles only), and per instance the actuation
amplitude (``instance_actuation``, ``mpm_sim.py:116,231``), ``contraction``
(``instance_act_contraction``, ``mpm_sim.py:118-121,177``; the refractory fraction is
``1 - contraction``), ``freq`` (``instance_freq``, ``mpm_sim.py:123-125,173``) and the jelly
Young's modulus ``E_jelly`` (``mpm_sim.py:31-34``, Poisson ratio 0.3; a module constant in
the reference, a per-instance input here). The payload density is an input without a
gradient.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import JmConfig, JmDiffInputs, call, load

STATE_COLS = 12
DT = np.float32(2e-5)                 # mpm_sim.py:19 (quality 1), as the fp32 clock adds it


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def clock(t0: float, n: int, dt=None) -> float:
    """The fp32 clock after ``n`` substeps from ``t0`` (``sim_time += dt`` in fp32,
    ``mpm_sim.py:336-340``)."""
    dt = DT if dt is None else np.float32(dt)
    t = np.float32(t0)
    for _ in range(int(n)):
        t = np.float32(t + dt)
    return float(t)


class DiffTank:
    """A batch of ``n_instances`` tanks packed as one particle array, with the device grids
    and scratch the forward and adjoint substeps need.

    ``inst`` (P,) instance index, ``mat`` (P,) material (0 water, 1 jelly, 2 payload,
    3 muscle), ``fiber`` (P, 2) muscle fibre direction."""

    def __init__(self, inst, mat, fiber, n_instances: int, quality: int = 1,
                 n_grid_y: int | None = None, domain_height: float = 1.0,
                 device: str | torch.device = "cuda"):
        load()
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.inst = torch.as_tensor(inst, dtype=torch.int32).to(dev).contiguous()
        self.mat = torch.as_tensor(mat, dtype=torch.int32).to(dev).contiguous()
        self.fiber = torch.as_tensor(fiber, dtype=torch.float32).to(dev).reshape(-1, 2)
        self.fiber = self.fiber.contiguous()
        self.M = int(n_instances)
        self.P = int(self.inst.numel())
        if self.P == 0 or self.M <= 0:
            raise ValueError("DiffTank needs at least one particle and one instance")
        if self.mat.numel() != self.P or self.fiber.shape[0] != self.P:
            raise ValueError("inst, mat and fiber lengths differ")
        inst_np = self.inst.cpu().numpy()
        mat_np = self.mat.cpu().numpy()
        if inst_np.min() < 0 or inst_np.max() >= self.M:
            raise ValueError("instance index out of range")
        if mat_np.min() < 0 or mat_np.max() > 3:
            raise ValueError("materials must be 0..3 (dead slots are dropped when packing)")
        self._inst_long = self.inst.long()
        self.payload_count = torch.zeros(self.M, dtype=torch.float64, device=dev).index_add_(
            0, self._inst_long, (self.mat == 2).double())
        cfg = JmConfig()
        call("jm_default_config", ctypes.byref(cfg))
        cfg.n_instances = self.M
        cfg.n_particles = 1
        cfg.quality = int(quality)
        cfg.n_grid_y = int(n_grid_y or 0)
        cfg.domain_height = float(domain_height)
        self.dt = np.float32(2e-5 / quality)
        h = ctypes.c_void_p()
        call("jm_diff_create", ctypes.byref(cfg), ctypes.c_int64(self.P), ctypes.byref(h))
        self._h = h
        self._L = load()

    @classmethod
    def from_tanks(cls, pos, mat, fiber, **kw):
        """Pack reference-layout tanks (``pos (M,N,2)``, ``mat (M,N)`` with -1 for dead
        slots, ``fiber (M,N,2)``) into a DiffTank and its initial state (v = 0, C = 0,
        F = I, as ``load_particles`` sets them, ``mpm_sim.py:715-736``)."""
        pos = np.asarray(pos, np.float32)
        mat = np.asarray(mat, np.int32)
        fiber = np.asarray(fiber, np.float32)
        M = mat.shape[0]
        live = mat >= 0
        inst = np.repeat(np.arange(M, dtype=np.int32)[:, None], mat.shape[1], 1)[live]
        tank = cls(inst, mat[live], fiber[live], M, **kw)
        state = np.zeros((int(live.sum()), STATE_COLS), np.float32)
        state[:, 0:2] = pos[live]
        state[:, 8] = 1.0
        state[:, 11] = 1.0
        return tank, torch.from_numpy(state).to(tank.device)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.jm_diff_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- raw calls -----------------------------------------------------------------------
    def _inputs(self, amp_scale: torch.Tensor, params: torch.Tensor) -> JmDiffInputs:
        return JmDiffInputs(self.inst.data_ptr(), self.mat.data_ptr(), self.fiber.data_ptr(),
                            amp_scale.data_ptr(), params.data_ptr())

    def _check(self, state, amp_scale, params):
        if state.shape != (self.P, STATE_COLS) or state.dtype != torch.float32:
            raise ValueError(f"state must be ({self.P}, {STATE_COLS}) float32")
        if amp_scale.shape != (self.P,) or params.shape != (self.M, 5):
            raise ValueError("amp_scale must be (P,) and params (M, 5)")
        for t in (state, amp_scale, params):
            if t.device != self.device or not t.is_contiguous():
                raise ValueError("tensors must be contiguous and on the tank's device")

    def forward_raw(self, state, amp_scale, params, t0: float, n: int, out=None,
                    store_all: bool = False) -> torch.Tensor:
        """``n`` substeps; returns the final state, or (n, P, 12) every state if
        ``store_all``. ``params`` (M, 5) = amp, contraction, freq, payload density, E."""
        self._check(state, amp_scale, params)
        if out is None:
            shape = (n, self.P, STATE_COLS) if store_all else (self.P, STATE_COLS)
            out = torch.empty(shape, dtype=torch.float32, device=self.device)
        ins = self._inputs(amp_scale, params)
        call("jm_fwd_block", self._h, ctypes.byref(ins), _ptr(state), ctypes.c_float(t0),
             int(n), _ptr(out), int(bool(store_all)), _stream())
        return out

    def backward_raw(self, state, amp_scale, params, t0: float, n: int, grad_out):
        """Adjoint of ``n`` substeps from ``state``: returns the adjoint of the input state
        (P, 12), of ``amp_scale`` (P,) and of ``params`` (M, 5; column 3 is zero)."""
        self._check(state, amp_scale, params)
        traj = torch.empty((n + 1, self.P, STATE_COLS), dtype=torch.float32,
                           device=self.device)
        traj[0].copy_(state)
        self.forward_raw(state, amp_scale, params, t0, n, out=traj[1:], store_all=True)
        grad_out = grad_out.to(torch.float32).contiguous()
        grad_in = torch.empty_like(state)
        gamp = torch.zeros(self.P, dtype=torch.float32, device=self.device)
        gpp = torch.zeros((self.P, 4), dtype=torch.float32, device=self.device)
        ins = self._inputs(amp_scale, params)
        call("jm_bwd_block", self._h, ctypes.byref(ins), _ptr(traj), ctypes.c_float(t0),
             int(n), _ptr(grad_out), _ptr(grad_in), _ptr(gamp), _ptr(gpp), _stream())
        per_inst = torch.zeros((self.M, 4), dtype=torch.float64, device=self.device)
        per_inst.index_add_(0, self._inst_long, gpp.double())
        gpar = torch.zeros((self.M, 5), dtype=torch.float32, device=self.device)
        gpar[:, 4] = per_inst[:, 0].float()    # E_jelly
        gpar[:, 0] = per_inst[:, 1].float()    # amplitude
        gpar[:, 1] = per_inst[:, 2].float()    # contraction
        gpar[:, 2] = per_inst[:, 3].float()    # freq
        return grad_in, gamp, gpar

    # -- autograd ---------------------------------------------------------------------
    def run(self, state0: torch.Tensor, amp_scale=None, amp=500.0, contraction=0.2, freq=1.0,
            E_jelly=700.0, payload_density=2.5, t0: float = 0.0, nsteps: int = 100,
            checkpoint: int = 100) -> torch.Tensor:
        """``nsteps`` substeps from ``state0`` at time ``t0``, as autograd blocks of
        ``checkpoint`` substeps. Per-instance arguments are scalars or (M,) tensors."""
        if amp_scale is None:
            amp_scale = torch.ones(self.P, dtype=torch.float32, device=self.device)

        def inst_vec(v):
            if isinstance(v, torch.Tensor):
                return v.to(self.device, torch.float32).expand(self.M)
            return torch.full((self.M,), float(v), dtype=torch.float32, device=self.device)

        amp, contraction, freq, E_jelly = (inst_vec(v) for v in (amp, contraction, freq,
                                                                 E_jelly))
        pdens = inst_vec(payload_density).detach()
        state = state0
        t = float(np.float32(t0))
        done = 0
        while done < nsteps:
            n = min(int(checkpoint), nsteps - done)
            state = _Block.apply(state, amp_scale, amp, contraction, freq, E_jelly, pdens,
                                 self, t, n)
            t = clock(t, n, self.dt)
            done += n
        return state

    def payload_com_y(self, state: torch.Tensor) -> torch.Tensor:
        """Per-instance mean y of the payload particles (the payload CoM the reference's
        fitness reads, ``mpm_sim.py:746-777``), differentiable in ``state``."""
        w = (self.mat == 2).to(state.dtype)
        s = torch.zeros(self.M, dtype=state.dtype, device=self.device)
        s = s.index_add(0, self._inst_long, state[:, 1] * w)
        return s / self.payload_count.clamp(min=1).to(state.dtype)


class _Block(torch.autograd.Function):
    """``n`` substeps as one autograd node; saves only its input state."""

    @staticmethod
    def forward(ctx, state, amp_scale, amp, contraction, freq, E_jelly, pdens, tank, t0, n):
        params = torch.stack([amp, contraction, freq, pdens, E_jelly], 1).detach().contiguous()
        state_c = state.detach().contiguous()
        amp_c = amp_scale.detach().to(torch.float32).contiguous()
        out = tank.forward_raw(state_c, amp_c, params, t0, n)
        ctx.save_for_backward(state_c, amp_c, params)
        ctx.tank, ctx.t0, ctx.n = tank, t0, n
        return out

    @staticmethod
    def backward(ctx, grad_out):
        state, amp_c, params = ctx.saved_tensors
        gin, gamp, gpar = ctx.tank.backward_raw(state, amp_c, params, ctx.t0, ctx.n, grad_out)
        return (gin, gamp, gpar[:, 0], gpar[:, 1], gpar[:, 2], gpar[:, 4], None, None, None,
                None)
