"""jm_fwd_block_noalloc (the differentiable forward without device allocation) against
jm_fwd_block, through the C ABI.

Two calls in a row, the second with a smaller batch, each compared with jm_fwd_block on the
same inputs. In fp32, bit-exact where every instance is one workgroup of 256 particles: the
workgroup's P2G sums are exact fixed point and each grid node receives one global add per
substep, so both runs sum in the same order. In fp64 the workgroup's tile adds with fp64
LDS atomics from several waves, whose order varies from run to run: within 1e-10. For a
full reference tank, where several workgroups flush into shared nodes in a run-dependent
order, within fp32 atomic-order differences (1e-5 in x). A batch beyond the fixed scratch
is refused before any device work.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest
import torch

from jellyfih_amd import _lib
from jellyfih_amd.diff import DiffTank
from tests.cases import reference_tank, small_tank

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def block_tanks(M: int, seed: int):
    """M instances of exactly 256 live particles (a 16 x 16 lattice with a robot patch)."""
    pos = np.full((M, 256, 2), -1.0, np.float32)
    mat = np.full((M, 256), -1, np.int32)
    fib = np.zeros((M, 256, 2), np.float32)
    for m in range(M):
        x0 = 0.08 + 0.1 * (m % 8)
        y0 = 0.1 + 0.2 * (m // 8)
        p, mt, f = small_tank(seed + m, 256, region=(x0, x0 + 0.0624, y0, y0 + 0.0624))
        assert (mt >= 0).sum() == 256
        pos[m], mat[m], fib[m] = p, mt, f
    return pos, mat, fib


def noalloc_forward(tank: DiffTank, state, amp_scale, params, t0, n):
    cfg = _lib.JmConfig()
    _lib.call("jm_default_config", ctypes.byref(cfg))
    cfg.n_instances = tank.M
    cfg.n_grid_y = tank.ny if tank.ny != tank.nx else 0
    out = torch.empty_like(state)
    ins = tank._inputs(amp_scale, params)
    _lib.call("jm_fwd_block_noalloc", ctypes.byref(cfg), ctypes.c_int64(tank.P),
              tank.scalar_bits, ctypes.byref(ins),
              ctypes.c_void_p(state.data_ptr()), ctypes.c_double(t0), int(n),
              ctypes.c_void_p(out.data_ptr()),
              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return out


def inputs(tank, state):
    amp_scale = torch.ones(tank.P, dtype=tank.dtype, device=tank.device)
    params = torch.tensor([[500.0, 0.2, 1.0, 2.5, 700.0]] * tank.M, dtype=tank.dtype,
                          device=tank.device)
    params[:, 0] += torch.arange(tank.M, device=tank.device, dtype=tank.dtype) * 50.0
    # a moving start: some velocity, so that the state changes every substep
    g = torch.Generator(device="cpu").manual_seed(3)
    state = state.clone()
    state[:, 2:4] = (torch.randn(tank.P, 2, generator=g) * 0.3).to(state)
    return state, amp_scale, params


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_forward_matches_fwd_block(require_gpu, dtype):
    for M, n, seed in ((6, 60, 1), (3, 25, 20)):   # the second call is smaller
        say(f"noalloc vs fwd_block dtype={dtype} M={M} n={n}")
        tank, st = DiffTank.from_tanks(*block_tanks(M, seed), order="slot", dtype=dtype)
        state, amp_scale, params = inputs(tank, st)
        ref = tank.forward_raw(state, amp_scale, params, 0.05, n)
        got = noalloc_forward(tank, state, amp_scale, params, 0.05, n)
        torch.cuda.synchronize()
        assert torch.isfinite(ref).all()
        assert not torch.equal(ref[:, 0:2], state[:, 0:2])
        if dtype == torch.float32:
            assert torch.equal(got, ref), float((got - ref).abs().max())
        else:
            assert float((got - ref).abs().max()) < 1e-10
        tank.close()


def test_noalloc_forward_reference_tank(require_gpu):
    for seed, n in ((0, 100), (1, 40)):
        say(f"noalloc vs fwd_block reference tank seed={seed} n={n}")
        p, mt, f, _, _ = reference_tank(seed)
        tank, st = DiffTank.from_tanks(p[None], mt[None], f[None])
        state, amp_scale, params = inputs(tank, st)
        ref = tank.forward_raw(state, amp_scale, params, 0.1, n)
        got = noalloc_forward(tank, state, amp_scale, params, 0.1, n)
        torch.cuda.synchronize()
        assert torch.isfinite(got).all()
        assert float((got[:, 0:2] - ref[:, 0:2]).abs().max()) < 1e-5
        assert float((got[:, 2:4] - ref[:, 2:4]).abs().max()) < 1e-3
        tank.close()


def test_noalloc_rejects_a_batch_beyond_its_scratch(require_gpu):
    say("noalloc size limit")
    tank, st = DiffTank.from_tanks(*block_tanks(2, 1), order="slot", dtype=torch.float64)
    state, amp_scale, params = inputs(tank, st)
    cfg = _lib.JmConfig()
    _lib.call("jm_default_config", ctypes.byref(cfg))
    cfg.n_instances = 2
    out = torch.empty_like(state)
    ins = tank._inputs(amp_scale, params)
    too_many = _lib.JM_NOALLOC_SCRATCH_BYTES // (32 * 8) + 1   # particles, fp64
    with pytest.raises(_lib.JmError, match="scratch"):
        _lib.call("jm_fwd_block_noalloc", ctypes.byref(cfg), ctypes.c_int64(too_many), 64,
                  ctypes.byref(ins), ctypes.c_void_p(state.data_ptr()), ctypes.c_double(0.0),
                  10, ctypes.c_void_p(out.data_ptr()), None)
    tank.close()
