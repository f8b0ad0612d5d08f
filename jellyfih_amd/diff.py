"""Differentiable batched substeps: reverse-mode gradients through the MLS-MPM substep.

The reference simulator (``mpm_sim.py:153-340``) has no autodiff. This module puts the same
substep behind ``torch.autograd``: the forward and the adjoint are HIP kernels in
``libjellymx.so`` (``jm_fwd_block`` / ``jm_bwd_block``), and :class:`_Block` is one
``torch.autograd.Function`` per block of ``checkpoint`` substeps. A block keeps its input
state (the checkpoint); its backward re-runs the block's forward storing every substep's
state and grid, then runs the adjoint substeps in reverse. The last blocks of a run keep
their stored states and grids from the forward pass instead, as many as device memory
allows (``DiffTank.run``'s ``keep_memory`` / ``keep_last``).

State layout, one row of 12 fp32 per live particle:
``x0 x1 v0 v1 C00 C01 C10 C11 F00 F01 F10 F11``; with fp32 buffers and fp64 arithmetic
(the default for fp32 tanks) the last four hold F - I (``DiffTank.encode_state`` /
``decode_state`` convert). Particles of all instances share one
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


_SIDE_STREAMS: dict = {}


def _side_stream(device: torch.device) -> "torch.cuda.Stream":
    """One extra stream per device (default priority) for the backward pass's recomputation
    of the previous block."""
    s = _SIDE_STREAMS.get(device.index)
    if s is None:
        s = _SIDE_STREAMS[device.index] = torch.cuda.Stream(device=device)
    return s


def clock(t0: float, n: int, dt=None, fp64: bool = False) -> float:
    """The clock after ``n`` substeps from ``t0``: ``sim_time += dt`` in fp32 as the
    reference does (``mpm_sim.py:336-340``), or in fp64 for the fp64 path."""
    if fp64:
        t = float(t0)
        step = 2e-5 if dt is None else float(dt)
        for _ in range(int(n)):
            t = t + step
        return t
    dt = DT if dt is None else np.float32(dt)
    t = np.float32(t0)
    for _ in range(int(n)):
        t = np.float32(t + dt)
    return float(t)


def packing_order(pos, mat, order: str, nx: int, ny: int):
    """(instance, slot) of every live particle (``mat >= 0``) in packing order: instance
    by instance, and within an instance by slot (``order="slot"``) or by grid cell in
    snaking strips of 8 cell rows, ties by slot (``order="cell"``); or (``order="split"``)
    the body's particles (material > 0) of every instance first and then the water of every
    instance, each instance by instance in that cell order."""
    N = mat.shape[1]
    inst, slot = np.divmod(np.flatnonzero(mat >= 0), N)
    if order == "slot":
        return inst, slot
    if order not in ("cell", "split"):
        raise ValueError("order must be 'cell', 'split' or 'slot'")
    x = np.asarray(pos, np.float32)[inst, slot]
    b = np.trunc(x * np.float32(nx) - np.float32(0.5)).astype(np.int64)
    bx, by = np.clip(b[:, 0], 0, nx - 1), np.clip(b[:, 1], 0, ny - 1)
    s, r = by // 8, by % 8
    key = (s * nx + np.where(s % 2 == 1, nx - 1 - bx, bx)) * 8 + r
    if order == "split":
        water = (mat[inst, slot] == 0).astype(np.int64)
        perm = np.lexsort((slot, key, inst, water))
    else:
        perm = np.lexsort((slot, key, inst))
    return inst[perm], slot[perm]


class DiffTank:
    """A batch of ``n_instances`` tanks packed as one particle array, with the device grids
    and scratch the forward and adjoint substeps need.

    ``inst`` (P,) instance index, ``mat`` (P,) material (0 water, 1 jelly, 2 payload,
    3 muscle), ``fiber`` (P, 2) muscle fibre direction. ``dtype`` is the type of the states,
    grids, inputs and gradients: float32 (the reference's storage) or float64. ``math`` is
    the arithmetic inside each substep: "fp64" (default; with float32 buffers each substep
    widens its fp32 inputs, computes and carries the adjoint in fp64, and rounds its results
    to fp32) or "fp32" (float32 only: the reference's arithmetic). fp32 arithmetic resolves
    the small strains near F = I (the jelly's F - R, the water's J - 1 at stiffness 1e5) to
    a few digits only, which moves gradients by percents within a few hundred substeps; fp32
    storage of the results does not (the fp64 arithmetic over fp32 buffers agrees with fp64
    finite differences to ~1e-5 over 200 substeps).

    ``water_start``: when the particles from that index on are all water (as
    ``from_tanks(order="split")`` packs them), the adjoint runs them with a water-only
    kernel, which holds more waves per SIMD; None (default) runs every particle with the
    general kernel. Results are the same."""

    def __init__(self, inst, mat, fiber, n_instances: int, quality: int = 1,
                 n_grid_y: int | None = None, domain_height: float = 1.0,
                 device: str | torch.device = "cuda", dtype: torch.dtype = torch.float32,
                 math: str = "fp64", water_start: int | None = None):
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("dtype must be torch.float32 or torch.float64")
        if math not in ("fp32", "fp64") or (dtype == torch.float64 and math != "fp64"):
            raise ValueError("math must be 'fp64', or 'fp32' with float32 buffers")
        self.dtype = dtype
        self.math = math
        load()
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.inst = torch.as_tensor(inst, dtype=torch.int32).to(dev).contiguous()
        self.mat = torch.as_tensor(mat, dtype=torch.int32).to(dev).contiguous()
        self.fiber = torch.as_tensor(fiber, dtype=dtype).to(dev).reshape(-1, 2)
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
        # per-instance sums of per-particle values: a stable sort by instance (identity when
        # packed instance by instance) and the segment bounds, for cumulative-sum reductions
        order = np.argsort(inst_np, kind="stable")
        self._order = None if np.all(order == np.arange(self.P)) else \
            torch.from_numpy(order).to(dev)
        bounds = np.searchsorted(inst_np[order], np.arange(self.M + 1))
        self._seg = torch.from_numpy(bounds.astype(np.int64)).to(dev)
        self.payload_count = self.instance_sums((self.mat == 2).double())
        cfg = JmConfig()
        call("jm_default_config", ctypes.byref(cfg))
        cfg.n_instances = self.M
        cfg.n_particles = 1
        cfg.quality = int(quality)
        cfg.n_grid_y = int(n_grid_y or 0)
        cfg.domain_height = float(domain_height)
        self.dt = 2e-5 / quality
        self.nx = 128 * int(quality)
        self.ny = int(n_grid_y or self.nx)
        self.G = self.nx * self.ny
        h = ctypes.c_void_p()
        bits = 64 if dtype == torch.float64 else (_lib.JM_DIFF_MIXED if math == "fp64" else 32)
        call("jm_diff_create", ctypes.byref(cfg), ctypes.c_int64(self.P), bits, ctypes.byref(h))
        self.scalar_bits = bits
        # fp32 buffers with fp64 arithmetic carry F - I in the state's F columns
        self.f_shift = bits == _lib.JM_DIFF_MIXED
        self._h = h
        self._L = load()
        self._owner = None      # a view (see _view) keeps the tank that owns the handle
        self.last_kept = 0      # blocks whose trajectories the last run() kept
        self.water_start = self.P
        if water_start is not None:
            ws = int(water_start)
            if not 0 <= ws <= self.P or np.any(mat_np[ws:] != 0):
                self.close()
                raise ValueError("water_start: particles from it on must all be water")
            call("jm_diff_set_water_start", self._h, ctypes.c_int64(ws))
            self.water_start = ws

    @classmethod
    def from_tanks(cls, pos, mat, fiber, order: str = "auto", **kw):
        """Pack reference-layout tanks (``pos (M,N,2)``, ``mat (M,N)`` with -1 for dead
        slots, ``fiber (M,N,2)``) into a DiffTank and its initial state (v = 0, C = 0,
        F = I, as ``load_particles`` sets them, ``mpm_sim.py:715-736``).

        ``order="cell"`` packs each instance's live particles sorted by grid cell, in
        snaking strips of 8 cell rows (the fused engine's sort order): a workgroup's
        particles then touch a compact window of nodes and a cell's particles sit in
        adjacent lanes, which the kernels' LDS tiles and cross-lane pre-sums exploit.
        ``order="split"`` packs every instance's body particles that way first and then
        every instance's water, and sets ``water_start`` to the first water particle, so
        the adjoint runs the water with its water-only kernel. ``order="slot"`` keeps slot
        order. ``order="auto"`` (default) is "split" for batches of at least 2^20 live
        particles with fp64 arithmetic, where the general adjoint kernel holds one wave per
        SIMD and the water-only one two (the 32-tank tall batch's forward + backward with
        fp32 buffers: 3.98 -> 3.11 s), and "cell" otherwise: with fp32 arithmetic the split
        does not pay (2.02 against 2.07 s there), and a batch that leaves the chip half
        empty (one reference tank is 242 workgroups) waits for its slowest workgroup, which
        the split makes one full of the body's SVD work (one tank, fp32 buffers with fp64
        arithmetic: forward 20.4 -> 33.6 ms). ``tank.slots`` (P, 2) gives the (instance,
        slot) of every packed particle either way."""
        pos = np.asarray(pos, np.float32)
        mat = np.asarray(mat, np.int32)
        if order == "auto":
            big = int(np.count_nonzero(mat >= 0)) >= (1 << 20)
            order = "split" if big and kw.get("math", "fp64") == "fp64" else "cell"
        fiber = np.asarray(fiber, np.float32)
        dtype = kw.get("dtype", torch.float32)
        M = mat.shape[0]
        nx = 128 * int(kw.get("quality", 1))
        inst_all, slot_all = packing_order(pos, mat, order, nx, int(kw.get("n_grid_y") or nx))
        mat_all = mat[inst_all, slot_all]
        if order == "split":
            kw["water_start"] = int(np.count_nonzero(mat_all != 0))
        tank = cls(inst_all.astype(np.int32), mat_all, fiber[inst_all, slot_all], M, **kw)
        tank.slots = np.stack([inst_all, slot_all], 1)
        state = np.zeros((inst_all.size, STATE_COLS), np.float32)
        state[:, 0:2] = pos[inst_all, slot_all]
        state[:, 8] = state[:, 11] = 0.0 if tank.f_shift else 1.0   # F = I
        return tank, torch.from_numpy(state).to(tank.device, dtype)

    # -- state format ---------------------------------------------------------------------
    def _identity_cols(self, like: torch.Tensor) -> torch.Tensor:
        e = torch.zeros(STATE_COLS, dtype=like.dtype, device=like.device)
        e[8] = e[11] = 1.0
        return e

    def encode_state(self, state: torch.Tensor) -> torch.Tensor:
        """A state with F in columns 8-11 in this tank's buffer format (F - I when
        ``f_shift``); differentiable, a no-op otherwise."""
        return state - self._identity_cols(state) if self.f_shift else state

    def decode_state(self, state: torch.Tensor) -> torch.Tensor:
        """The inverse of :meth:`encode_state`: columns 8-11 hold F."""
        return state + self._identity_cols(state) if self.f_shift else state

    def close(self) -> None:
        if getattr(self, "_owner", None) is not None:   # a view: the owner frees the handle
            self._h = None
            return
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.jm_diff_destroy(self._h)
            self._h.value = None   # views share this object: their calls now fail cleanly
        self._h = None

    # -- re-sorting between blocks ----------------------------------------------------------
    def cell_permutation(self, state: torch.Tensor) -> torch.Tensor:
        """The permutation (P,) that orders the particles of each instance by grid cell at
        their current positions, in the ``order="cell"`` packing order (snaking strips of 8
        cell rows), ties by current position in the array; instances keep their blocks (and,
        packed with ``order="split"``, the body and the water keep theirs)."""
        split = self.water_start < self.P
        if self._order is not None and not split:
            raise ValueError("re-sorting needs particles packed instance by instance")
        x = state[:, :2].detach().to(torch.float32)
        b = torch.trunc(x * float(self.nx) - 0.5).nan_to_num(0.0).long()
        bx = b[:, 0].clamp(0, self.nx - 1)
        by = b[:, 1].clamp(0, self.ny - 1)
        s, r = by // 8, by % 8
        key = (s * self.nx + torch.where(s % 2 == 1, self.nx - 1 - bx, bx)) * 8 + r
        span = ((self.ny + 7) // 8) * 8 * self.nx
        group = self.inst.long()
        if split:
            group = group + self.M * (self.mat == 0).long()
        return torch.argsort(group * span + key, stable=True)

    def _view(self, perm: torch.Tensor) -> "DiffTank":
        """This tank with its particles in the order ``perm`` (a within-instance permutation
        of this tank's order), sharing the device handle."""
        v = object.__new__(DiffTank)
        v.__dict__.update(self.__dict__)
        v._owner = self if self._owner is None else self._owner
        v.inst = self.inst[perm].contiguous()
        v.mat = self.mat[perm].contiguous()
        v.fiber = self.fiber[perm].contiguous()
        return v

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
        if state.shape != (self.P, STATE_COLS):
            raise ValueError(f"state must be ({self.P}, {STATE_COLS})")
        if amp_scale.shape != (self.P,) or params.shape != (self.M, 5):
            raise ValueError("amp_scale must be (P,) and params (M, 5)")
        for t in (state, amp_scale, params):
            if t.dtype != self.dtype:
                raise ValueError(f"tensors must be {self.dtype}")
        for t in (state, amp_scale, params):
            if t.device != self.device or not t.is_contiguous():
                raise ValueError("tensors must be contiguous and on the tank's device")

    def forward_raw(self, state, amp_scale, params, t0: float, n: int, out=None,
                    grids=None) -> torch.Tensor:
        """``n`` substeps; returns the final state (P, 12). With ``grids`` (a tensor
        (n, M, 3, nx*ny) of the tank's dtype), ``out`` must be (n, P, 12) and receives every
        substep's state, ``grids`` every substep's momentum and mass planes. ``params``
        (M, 5) = amp, contraction, freq, payload density, E_jelly."""
        self._check(state, amp_scale, params)
        if grids is not None:
            if out is None or tuple(out.shape) != (n, self.P, STATE_COLS) or \
                    tuple(grids.shape) != (n, self.M, 3, self.G) or not grids.is_contiguous():
                raise ValueError("stored trajectory needs out (n, P, 12) and grids "
                                 "(n, M, 3, nx*ny)")
        elif out is None:
            out = torch.empty((self.P, STATE_COLS), dtype=self.dtype, device=self.device)
        ins = self._inputs(amp_scale, params)
        call("jm_fwd_block", self._h, ctypes.byref(ins), _ptr(state), ctypes.c_double(t0),
             int(n), _ptr(out), None if grids is None else _ptr(grids), _stream())
        return out

    def instance_sums(self, values: torch.Tensor) -> torch.Tensor:
        """Per-instance fp64 sums of per-particle values, (P,) -> (M,) or (P, C) -> (M, C):
        differences of a cumulative sum at the instance bounds (the scan runs along the
        contiguous particle axis)."""
        v = values.double()
        if self._order is not None:
            v = v[self._order]
        vt = v.reshape(self.P, -1).t().contiguous()                  # (C, P)
        cs = torch.cat([torch.zeros_like(vt[:, :1]), torch.cumsum(vt, 1)], 1)
        out = (cs[:, self._seg[1:]] - cs[:, self._seg[:-1]]).t()     # (M, C)
        return out.reshape((self.M,) + tuple(values.shape[1:]))

    def block_bytes(self, n: int) -> int:
        """Device bytes of one block's stored states and grids (:meth:`trajectory_buffers`)."""
        item = 8 if self.dtype == torch.float64 else 4
        return ((n + 1) * self.P * STATE_COLS + n * self.M * 3 * self.G) * item

    def blocks_to_keep(self, nsteps: int, checkpoint: int, keep_memory: float) -> int:
        """How many of the last blocks of :meth:`run` keep their trajectories: as many as
        fit in ``keep_memory`` times the device memory free now (the caching allocator's
        unused reserve counted as free) less 2 GiB, at most all of them; 0 off the GPU."""
        nb = -(-int(nsteps) // int(checkpoint))
        if keep_memory <= 0 or nb == 0 or self.device.type != "cuda":
            return 0
        free, _ = torch.cuda.mem_get_info(self.device)
        free += torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
        room = float(keep_memory) * free - float(2 << 30)
        return max(0, min(nb, int(room // self.block_bytes(int(checkpoint)))))

    def trajectory_buffers(self, n: int):
        """Uninitialised (traj (n+1, P, 12), grids (n, M, 3, nx*ny)) for :meth:`trajectory`."""
        return (torch.empty((n + 1, self.P, STATE_COLS), dtype=self.dtype, device=self.device),
                torch.empty((n, self.M, 3, self.G), dtype=self.dtype, device=self.device))

    def trajectory(self, state, amp_scale, params, t0: float, n: int, side=None, bufs=None):
        """``n`` substeps from ``state`` with every state and grid stored: returns
        (traj (n+1, P, 12) starting with ``state``, grids (n, M, 3, nx*ny)), written into
        the leading rows of ``bufs`` (from :meth:`trajectory_buffers`) when given.

        With ``side`` (a stream), the substeps run on ``side`` after the current stream's
        work so far; the returned event marks their end (wait on it before reading the
        trajectory). The store-mode forward uses none of the handle's scratch that the
        adjoint uses, so it may run beside :meth:`backward_raw` of another block."""
        if bufs is None:
            bufs = self.trajectory_buffers(n)
        if bufs[0].shape[0] < n + 1 or bufs[1].shape[0] < n:
            raise ValueError("trajectory buffers hold fewer substeps than the block")
        traj, grids = bufs[0][:n + 1], bufs[1][:n]
        if side is None:
            traj[0].copy_(state)
            self.forward_raw(state, amp_scale, params, t0, n, out=traj[1:], grids=grids)
            return traj, grids
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            traj[0].copy_(state)
            self.forward_raw(state, amp_scale, params, t0, n, out=traj[1:], grids=grids)
            done = torch.cuda.Event()
            done.record(side)
        # the allocator must not hand these to other work before the side stream is done
        for t in (traj, grids, state, amp_scale, params, self.inst, self.mat, self.fiber):
            t.record_stream(side)
        return traj, grids, done

    def backward_raw(self, state, amp_scale, params, t0: float, n: int, grad_out,
                     stored=None):
        """Adjoint of ``n`` substeps from ``state``: returns the adjoint of the input state
        (P, 12), of ``amp_scale`` (P,) and of ``params`` (M, 5; column 3 is zero). The
        trajectory is recomputed unless ``stored`` = trajectory(...) of the same block."""
        self._check(state, amp_scale, params)
        traj, grids = stored if stored is not None else \
            self.trajectory(state, amp_scale, params, t0, n)
        grad_out = grad_out.to(self.dtype).contiguous()
        grad_in = torch.empty_like(state)
        gamp = torch.zeros(self.P, dtype=self.dtype, device=self.device)
        gparams = torch.zeros((self.M, 4), dtype=torch.float64, device=self.device)
        ins = self._inputs(amp_scale, params)
        call("jm_bwd_block", self._h, ctypes.byref(ins), _ptr(traj), _ptr(grids),
             ctypes.c_double(t0), int(n), _ptr(grad_out), _ptr(grad_in), _ptr(gamp),
             _ptr(gparams), _stream())
        gpar = torch.zeros((self.M, 5), dtype=self.dtype, device=self.device)
        gpar[:, 4] = gparams[:, 0]    # E_jelly
        gpar[:, 0] = gparams[:, 1]    # amplitude
        gpar[:, 1] = gparams[:, 2]    # contraction
        gpar[:, 2] = gparams[:, 3]    # freq
        return grad_in, gamp, gpar

    # -- autograd ---------------------------------------------------------------------
    def run(self, state0: torch.Tensor, amp_scale=None, amp=500.0, contraction=0.2, freq=1.0,
            E_jelly=700.0, payload_density=2.5, t0: float = 0.0, nsteps: int = 100,
            checkpoint: int = 100, recompute: bool = True, resort_every: int = 0,
            prefetch: bool = True, keep_memory: float = 0.8,
            keep_last: int | None = None) -> torch.Tensor:
        """``nsteps`` substeps from ``state0`` at time ``t0``, as autograd blocks of
        ``checkpoint`` substeps. Per-instance arguments are scalars or (M,) tensors.

        ``recompute`` (default): a block keeps only its input state and re-runs its
        substeps in the backward pass. ``recompute=False`` keeps every block's states and
        grids from the forward pass instead: no second forward, at ((checkpoint + 1) * P * 12
        + checkpoint * M * 3 * nx * ny) scalars of device memory per block.

        ``prefetch`` (with ``recompute``): the backward pass of block b starts the
        recomputation of block b-1 on a second stream before it runs its own adjoint, so
        the two run side by side (the recomputation does not depend on the adjoint); at
        most two blocks' trajectories are held at once. Same kernels, same results.

        ``keep_memory`` (with ``recompute``): the last blocks keep their states and grids
        from the forward pass, as with ``recompute=False``, as many as fit in this fraction
        of the device memory free when the run starts (:meth:`blocks_to_keep`; 0 keeps
        none), so the backward pass recomputes only the blocks before them. Storing costs
        the forward pass no extra work (it writes every substep's state and grid either
        way); the kept blocks' memory is returned as the backward pass consumes them, and
        the recomputation of the earlier blocks reuses it. One reference tank keeps its
        whole episode (~0.3 GB per 100-substep block); the 32-tank tall batch keeps 6 of
        its 20 blocks (41 GB each) on a 288 GB device. ``keep_last`` = K (an int) keeps
        exactly the last K blocks instead (at most all of them; 0 recomputes every block).

        ``resort_every`` = K > 0: before every K-th block after the first, the particles
        are re-ordered by their current cells (``cell_permutation``), as the fused engine
        re-sorts every K substeps, so the kernels' LDS tiles stay compact over long
        episodes. The reordering is a gather of the state and of ``amp_scale`` (autograd
        scatters their adjoints back); the returned state is in this tank's order. Results
        change only by the order of floating-point sums."""
        if amp_scale is None:
            amp_scale = torch.ones(self.P, dtype=self.dtype, device=self.device)

        def inst_vec(v):
            if isinstance(v, torch.Tensor):
                return v.to(self.device, self.dtype).expand(self.M)
            return torch.full((self.M,), float(v), dtype=self.dtype, device=self.device)

        amp, contraction, freq, E_jelly = (inst_vec(v) for v in (amp, contraction, freq,
                                                                 E_jelly))
        pdens = inst_vec(payload_density).detach()
        state = state0
        fp64 = self.dtype == torch.float64
        t = float(t0) if fp64 else float(np.float32(t0))
        done = blocks = 0
        cur, perm, amp_cur = self, None, amp_scale
        episode = _Episode() if recompute and prefetch and self.device.type == "cuda" \
            else None
        nblocks = -(-int(nsteps) // int(checkpoint))
        if not recompute:
            kept = nblocks          # every block keeps its trajectory
        elif keep_last is not None:
            kept = max(0, min(nblocks, int(keep_last)))
        else:
            kept = self.blocks_to_keep(nsteps, checkpoint, keep_memory)
        keep_from = nblocks - kept
        self.last_kept = kept
        while done < nsteps:
            if resort_every > 0 and blocks > 0 and blocks % int(resort_every) == 0:
                p = cur.cell_permutation(state)
                state, amp_cur = state[p], amp_cur[p]
                perm = p if perm is None else perm[p]
                cur = self._view(perm)
            n = min(int(checkpoint), nsteps - done)
            state = _Block.apply(state, amp_cur, amp, contraction, freq, E_jelly, pdens,
                                 cur, t, n, recompute and blocks < keep_from, episode)
            t = clock(t, n, self.dt, fp64)
            done += n
            blocks += 1
        if perm is not None:
            inv = torch.empty_like(perm)
            inv[perm] = torch.arange(perm.numel(), device=perm.device)
            state = state[inv]
        return state

    def payload_com_y(self, state: torch.Tensor) -> torch.Tensor:
        """Per-instance mean y of the payload particles (the payload CoM the reference's
        fitness reads, ``mpm_sim.py:746-777``), differentiable in ``state``."""
        w = (self.mat == 2).to(state.dtype)
        s = self.instance_sums(state[:, 1] * w)
        return (s / self.payload_count.clamp(min=1)).to(state.dtype)


class _Episode:
    """The blocks of one ``DiffTank.run`` with recomputation, so that the backward pass of
    block b can start block b-1's recomputation (on a side stream) before its own adjoint.

    ``blocks[b]`` = (tank, state, amp_scale, params, t0, n) as block b's forward saw them,
    or None for a block that kept its trajectory or whose recomputation was taken;
    ``pending[b]`` = (traj, grids, event, buffers) of a recomputation in flight. Two sets
    of trajectory buffers, sized for the longest block, serve every block in turn: one is
    read by the current adjoint while the other is written by the next recomputation."""

    def __init__(self):
        self.blocks: list = []
        self.pending: dict = {}
        self.free: list = []
        self.n_max = 0

    def add(self, rec) -> int:
        """Registers the next block: its inputs, or None for a block that keeps its
        trajectory (nothing to recompute)."""
        self.blocks.append(rec)
        if rec is not None:
            self.n_max = max(self.n_max, rec[5])
        return len(self.blocks) - 1

    def _buffers(self, tank):
        return self.free.pop() if self.free else tank.trajectory_buffers(self.n_max)

    def start(self, b: int) -> None:
        """Block b's recomputation on the side stream, after the current stream's work."""
        if b < 0 or b in self.pending or self.blocks[b] is None:
            return
        tank, state, amp_c, params, t0, n = self.blocks[b]
        bufs = self._buffers(tank)
        self.pending[b] = tank.trajectory(state, amp_c, params, t0, n,
                                          side=_side_stream(tank.device), bufs=bufs) + (bufs,)

    def take(self, b: int):
        """(traj, grids, buffers) of block b: the recomputation start(b) launched, ordered
        before the current stream's next work, or else run now on the current stream; None
        once block b has been taken (a second backward pass through a retained graph
        recomputes from the block's own saved inputs)."""
        got = self.pending.pop(b, None)
        if got is not None:
            traj, grids, done, bufs = got
            torch.cuda.current_stream(traj.device).wait_event(done)
        elif self.blocks[b] is not None:
            tank, state, amp_c, params, t0, n = self.blocks[b]
            bufs = self._buffers(tank)
            traj, grids = tank.trajectory(state, amp_c, params, t0, n, bufs=bufs)
        else:
            return None
        self.blocks[b] = None     # its inputs are no longer needed here
        return traj, grids, bufs

    def release(self, bufs) -> None:
        """Buffers whose readers are all on the current stream, queued: a later
        recomputation on the side stream waits for them (start orders it after the current
        stream's work)."""
        self.free.append(bufs)


class _Block(torch.autograd.Function):
    """``n`` substeps as one autograd node; keeps its input state (and, without
    recomputation, the block's stored states and grids)."""

    @staticmethod
    def forward(ctx, state, amp_scale, amp, contraction, freq, E_jelly, pdens, tank, t0, n,
                recompute, episode=None):
        params = torch.stack([amp, contraction, freq, pdens, E_jelly], 1).detach().contiguous()
        state_c = state.detach().contiguous()
        amp_c = amp_scale.detach().to(tank.dtype).contiguous()
        if recompute:
            out = tank.forward_raw(state_c, amp_c, params, t0, n)
            ctx.stored = None
        else:
            ctx.stored = tank.trajectory(state_c, amp_c, params, t0, n)
            out = ctx.stored[0][-1].clone()
        ctx.save_for_backward(state_c, amp_c, params)
        ctx.tank, ctx.t0, ctx.n = tank, t0, n
        ctx.episode = episode
        ctx.index = -1
        if episode is not None:
            ctx.index = episode.add((tank, state_c, amp_c, params, t0, n) if recompute
                                    else None)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        state, amp_c, params = ctx.saved_tensors
        stored, ep, bufs = ctx.stored, ctx.episode, None
        if ep is not None:
            if stored is None:
                got = ep.take(ctx.index)
                if got is not None:
                    traj, grids, bufs = got
                    stored = (traj, grids)
            ep.start(ctx.index - 1)    # the previous block's recomputation, beside this adjoint
        gin, gamp, gpar = ctx.tank.backward_raw(state, amp_c, params, ctx.t0, ctx.n, grad_out,
                                                stored=stored)
        if bufs is not None:
            ep.release(bufs)
        ctx.stored = None
        return (gin, gamp, gpar[:, 0], gpar[:, 1], gpar[:, 2], gpar[:, 4], None, None, None,
                None, None, None)
