"""Host-side logic of the differentiable path (jellyfih_amd/diff.py) that runs without a
GPU: the per-instance reductions, the fp32 clock, the packing orders and the sizing of the
blocks a run keeps from its forward pass."""
from __future__ import annotations

import numpy as np
import torch

from jellyfih_amd.diff import DiffTank, clock


class _Packed:
    """The fields instance_sums reads, for particles in a given instance order."""

    def __init__(self, inst, M):
        inst = np.asarray(inst)
        self.P, self.M = len(inst), M
        order = np.argsort(inst, kind="stable")
        self._order = None if np.all(order == np.arange(self.P)) else torch.from_numpy(order)
        self._seg = torch.from_numpy(
            np.searchsorted(inst[order], np.arange(M + 1)).astype(np.int64))


def test_instance_sums_match_index_add():
    rng = np.random.default_rng(0)
    for inst, M in ((rng.integers(0, 5, 1000), 6), (np.repeat(np.arange(4), 50), 4),
                    (np.zeros(7, int), 1)):
        t = _Packed(inst, M)
        for shape in ((t.P,), (t.P, 4)):
            v = torch.from_numpy(rng.normal(size=shape))
            ref = torch.zeros((M,) + shape[1:], dtype=torch.float64).index_add_(
                0, torch.from_numpy(np.asarray(inst)), v)
            np.testing.assert_allclose(DiffTank.instance_sums(t, v).numpy(), ref.numpy(),
                                       rtol=0, atol=1e-12)


def test_instance_sums_are_differentiable():
    t = _Packed([1, 0, 1, 1, 0], 2)
    v = torch.arange(5, dtype=torch.float64, requires_grad=True)
    (DiffTank.instance_sums(t, v) * torch.tensor([2.0, 3.0], dtype=torch.float64)).sum().backward()
    np.testing.assert_array_equal(v.grad.numpy(), [3, 2, 3, 3, 2])


def test_clock_is_fp32_accumulation():
    """sim_time += dt in fp32 (mpm_sim.py:336-340): 150K substeps from 0 end at
    3.0033383 (not 3.0), and blocks compose."""
    assert np.float32(clock(0.0, 150000)) == np.float32(3.0033383)
    assert clock(clock(0.0, 700), 1300) == clock(0.0, 2000)


def test_packing_order_matches_engine_sort_order():
    """``order="cell"`` packs each instance like the engine's stable cell sort (strips of
    8 rows, the oracle's ``stable_cell_sort``); ``order="slot"`` keeps slot order."""
    import oracle
    from jellyfih_amd.diff import packing_order
    rng = np.random.default_rng(3)
    M, N = 3, 500
    pos = rng.uniform(0.0, 1.0, size=(M, N, 2)).astype(np.float32)
    mat = rng.integers(-1, 4, size=(M, N)).astype(np.int32)
    inst, slot = packing_order(pos, mat, "slot", 128, 128)
    assert np.all(mat[inst, slot] >= 0) and inst.size == int((mat >= 0).sum())
    assert np.all(np.diff(inst * N + slot) > 0)
    inst, slot = packing_order(pos, mat, "cell", 128, 128)
    c = oracle.reference_consts()
    for m in range(M):
        live = np.flatnonzero(mat[m] >= 0)
        expect = live[oracle.stable_cell_sort(c, pos[m, live], 8)]
        np.testing.assert_array_equal(slot[inst == m], expect)
    assert np.all(np.diff(inst) >= 0)


def test_packing_order_split_puts_the_water_last():
    """``order="split"``: every instance's body particles (material > 0) first, then every
    instance's water; within each part, instance by instance in the ``order="cell"``
    order of that instance's particles of the part."""
    import oracle
    from jellyfih_amd.diff import packing_order
    rng = np.random.default_rng(4)
    M, N = 3, 600
    pos = rng.uniform(0.0, 1.0, size=(M, N, 2)).astype(np.float32)
    mat = rng.choice([-1, 0, 0, 0, 1, 2, 3], size=(M, N)).astype(np.int32)
    inst, slot = packing_order(pos, mat, "split", 128, 128)
    m_all = mat[inst, slot]
    assert inst.size == int((mat >= 0).sum())
    nb = int((mat > 0).sum())
    assert np.all(m_all[:nb] > 0) and np.all(m_all[nb:] == 0)
    c = oracle.reference_consts()
    for part, sel in ((slice(0, nb), lambda mm: mm > 0), (slice(nb, None), lambda mm: mm == 0)):
        assert np.all(np.diff(inst[part]) >= 0)
        for m in range(M):
            live = np.flatnonzero(sel(mat[m]))
            expect = live[oracle.stable_cell_sort(c, pos[m, live], 8)]
            np.testing.assert_array_equal(slot[part][inst[part] == m], expect)


class _Sized:
    """The fields block_bytes and blocks_to_keep read."""

    def __init__(self, P, M, G, dtype=torch.float32, device="cuda"):
        self.P, self.M, self.G, self.dtype = P, M, G, dtype
        self.device = torch.device(device)

    block_bytes = DiffTank.block_bytes


def test_blocks_to_keep_arithmetic(monkeypatch):
    """The number of kept blocks: as many blocks of (checkpoint + 1) P 12 + checkpoint M 3 G
    scalars as fit in keep_memory times the free memory (plus the caching allocator's unused
    reserve) less 2 GiB, at most all blocks of the run; none for a zero fraction, an empty
    run or a tank off the GPU. The 32-tank tall batch at checkpoint 200 on a 288 GB device
    keeps 5 of its 20 blocks (41.4 GB each)."""
    GiB = 1 << 30
    tall = _Sized(P=4034949, M=32, G=128 * 256)
    assert tall.block_bytes(200) == (201 * 4034949 * 12 + 200 * 32 * 3 * 128 * 256) * 4
    assert _Sized(10, 1, 16, dtype=torch.float64).block_bytes(3) == (4 * 10 * 12 + 3 * 48) * 8
    state = {"free": 285 * 10**9, "reserved": 0, "allocated": 0}
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda dev=None: (state["free"], 288 * 10**9))
    monkeypatch.setattr(torch.cuda, "memory_reserved", lambda dev=None: state["reserved"])
    monkeypatch.setattr(torch.cuda, "memory_allocated", lambda dev=None: state["allocated"])
    keep = lambda obj, n, c, f: DiffTank.blocks_to_keep(obj, n, c, f)
    b = tall.block_bytes(200)
    assert keep(tall, 4000, 200, 0.8) == int((0.8 * 285e9 - 2 * GiB) // b) == 5
    assert keep(tall, 4000, 200, 0.0) == 0
    assert keep(tall, 0, 200, 0.8) == 0
    assert keep(tall, 400, 200, 0.8) == 2           # at most the run's 2 blocks
    assert keep(tall, 450, 200, 0.8) == 3           # a short last block counts as a block
    state.update(free=100 * 10**9, reserved=150 * 10**9, allocated=20 * 10**9)
    assert keep(tall, 4000, 200, 0.8) == int((0.8 * 230e9 - 2 * GiB) // b) == 4
    state.update(free=GiB, reserved=0, allocated=0)
    assert keep(tall, 4000, 200, 0.8) == 0          # less than 2 GiB of room: none
    assert keep(_Sized(4034949, 32, 128 * 256, device="cpu"), 4000, 200, 0.8) == 0
