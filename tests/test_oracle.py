"""Pins the oracle (oracle/liboracle.so, the CPU restatement of mpm_sim.py:153-340).

The reference ships no tests or golden vectors for this path, and its arithmetic runs in
Taichi 1.7.4 (absent here), so each part of the oracle is pinned by what the mathematics
fixes, chosen so that a plausible transcription mistake fails at least one check:

* actuation (mpm_sim.py:173-181): closed form and its special points;
* clock (mpm_sim.py:340): fp32 accumulation, 150K steps -> 3.0033383;
* SVD (taichi _svd2d): reconstruction, proper rotations, ordering, the polar closed form;
* P2G (mpm_sim.py:183-271): mass and momentum conservation (sum w = 1, sum w*dpos = 0),
  zero stress at rest;
* grid op (mpm_sim.py:273-309): node-by-node against an independent vectorised numpy pass;
* G2P (mpm_sim.py:311-333): APIC reproduces affine velocity fields exactly (v = a + Bx,
  C = B), which fixes the weights, the dpos convention and the 4*inv_dx scale;
* the whole substep: against oracle/brute.py (fp64 numpy with LAPACK SVD) on tiny inputs;
* fp32 vs fp64 restatement: error bound on a short run.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import brute


@pytest.fixture(scope="module")
def c():
    return oracle.reference_consts()


# --- actuation ---------------------------------------------------------------------------
def test_activation_closed_form(c):
    freq = 1.3
    cf = 0.25
    period = 1.0 / 1.3
    ts = np.linspace(0, 3, 2001).astype(np.float32)
    got = oracle.activation(c, ts, freq, cf, dtype=np.float64)
    phase = np.mod(ts.astype(np.float64), period) / period
    exp = np.where(phase < 0.25, 0.5 * (1 - np.cos(phase / 0.25 * 3.14159265)), 0.0)
    near_edge = np.abs(phase - 0.25) < 1e-6
    assert np.max(np.abs(got - exp)[~near_edge]) < 1e-9
    a = oracle.activation(c, [0.0, period * 0.125, period * 0.2499, period * 0.25, period * 0.6],
                          1.3, 0.25, dtype=np.float64)
    assert a[0] == 0.0
    assert abs(a[1] - 0.5) < 1e-7          # half-way through the contraction
    assert a[2] > 0.99                     # ramps up to 1 at the end of the contraction
    assert a[3] == 0.0 and a[4] == 0.0     # refractory: zero


def test_activation_fp32_matches_fp64(c):
    ts = np.linspace(0, 3, 501).astype(np.float32)
    a32 = oracle.activation(c, ts, 1.0, 0.2, dtype=np.float32)
    a64 = oracle.activation(c, ts.astype(np.float64), 1.0, 0.2, dtype=np.float64)
    ok = np.abs(np.mod(ts, 1.0) - 0.2) > 1e-5
    assert np.max(np.abs(a32 - a64)[ok]) < 1e-5


def test_clock_accumulates_in_fp32(c):
    s = oracle.OracleSim(1, 1, c)
    s.material[:] = -1
    s.run(150000, stages=0)
    assert s.t[0] == np.float32(3.0033383)
    t = np.float32(0)
    for _ in range(1000):
        t = np.float32(t + np.float32(2e-5))
    s2 = oracle.OracleSim(1, 1, c)
    s2.material[:] = -1
    s2.run(1000, stages=0)
    assert s2.t[0] == t


# --- SVD -----------------------------------------------------------------------------------
@pytest.mark.parametrize("dtype,tol", [(np.float32, 2e-6), (np.float64, 1e-13)])
def test_svd2d_reconstruction_and_rotations(dtype, tol):
    rng = np.random.default_rng(0)
    F = np.concatenate([
        np.eye(2) + rng.normal(size=(300, 2, 2)) * 0.1,
        rng.normal(size=(300, 2, 2)),                            # includes det < 0
        np.eye(2) * rng.uniform(0.5, 2, size=(50, 1, 1)),        # diagonal branch
        np.array([[[1.0, 3e-6], [0.0, 1.0]]]),                   # |S01| < 1e-5 branch
    ]).astype(dtype)
    U, sig, V, R = oracle.svd2d(F, dtype=dtype)
    rec = U @ (sig[:, :, None] * np.swapaxes(V, 1, 2))
    scale = np.maximum(1.0, np.abs(F).max(axis=(1, 2)))[:, None, None]
    branch = np.abs((np.swapaxes(R, 1, 2) @ F)[:, 0, 1]) >= 1e-5
    assert np.max((np.abs(rec - F) / scale)[branch]) < 50 * tol
    assert np.max(np.abs(np.linalg.det(U.astype(np.float64)) - 1)) < 10 * tol
    assert np.max(np.abs(np.linalg.det(V.astype(np.float64)) - 1)) < 10 * tol
    assert np.all(sig[:, 0] >= sig[:, 1])
    # U V^T is the polar rotation of F: (c, s) = normalize(F00 + F11, F10 - F01)
    x = F[:, 0, 0].astype(np.float64) + F[:, 1, 1]
    y = F[:, 1, 0].astype(np.float64) - F[:, 0, 1]
    r = np.hypot(x, y)
    Rc = np.stack([np.stack([x / r, -y / r], -1), np.stack([y / r, x / r], -1)], -2)
    assert np.max(np.abs(U @ np.swapaxes(V, 1, 2) - Rc)) < 50 * tol
    # singular values agree with LAPACK's up to sign when det F > 0
    pos = (np.linalg.det(F.astype(np.float64)) > 0.05) & branch   # exact (Jacobi) branch
    ref = np.linalg.svd(F.astype(np.float64), compute_uv=False)
    assert np.max(np.abs(np.sort(np.abs(sig[pos]), 1) - np.sort(ref[pos], 1))) < 100 * tol
    # det F < 0 gives a negative second singular value (rotation-variant convention)
    neg = np.linalg.det(F.astype(np.float64)) < -0.05
    assert np.all(sig[neg][:, 1] < 0)


def test_svd2d_identity():
    U, sig, V, R = oracle.svd2d(np.eye(2)[None], dtype=np.float64)
    np.testing.assert_array_equal(U[0], np.eye(2))
    np.testing.assert_array_equal(V[0], np.eye(2))
    np.testing.assert_array_equal(sig[0], [1, 1])


# --- P2G -----------------------------------------------------------------------------------
def _random_instance(rng, n, c, mats=(0, 1, 2, 3), lo=0.1, hi=0.85, sv=0.3, sc=1.0, sf=0.05,
                     dtype=np.float64):
    s = oracle.OracleSim(1, n, c, dtype=dtype)
    pos = rng.uniform(lo, hi, size=(n, 2))
    mat = rng.choice(mats, size=n).astype(np.int32)
    ang = rng.uniform(0, 2 * np.pi, n)
    fib = np.stack([np.cos(ang), np.sin(ang)], 1)
    s.load(0, pos, mat, fib)
    s.x[0] = pos                              # keep full precision in fp64
    s.fiber[0] = fib
    s.v[0] = rng.normal(size=(n, 2)) * sv
    s.C[0] = rng.normal(size=(n, 2, 2)) * sc
    s.F[0] = np.eye(2) + rng.normal(size=(n, 2, 2)) * sf
    s.t[0] = 0.05
    return s


def test_p2g_conserves_mass_and_momentum(c):
    rng = np.random.default_rng(1)
    s = _random_instance(rng, 400, c)
    mass = np.where(s.material[0] == 2, c.p_mass * s.pdens[0], c.p_mass)
    mom = (mass[:, None] * s.v[0]).sum(0)
    s.run(1, stages=oracle.ST_RESET | oracle.ST_P2G)
    assert abs(s.grid_m.sum() - mass.sum()) < 1e-12 * mass.sum()
    np.testing.assert_allclose(s.grid_v.reshape(-1, 2).sum(0), mom, rtol=1e-9, atol=1e-15)


def test_p2g_at_rest_has_no_stress(c):
    """F = I, C = 0, v = 0, no actuation: the grid receives mass only (zero stress)."""
    s = oracle.OracleSim(1, 3, c, dtype=np.float64)
    s.load(0, np.array([[0.3, 0.4], [0.5013, 0.61], [0.7, 0.2]]), np.array([0, 1, 2]))
    s.t[0] = 0.5   # refractory phase for the default contraction 0.2
    s.run(1, stages=oracle.ST_RESET | oracle.ST_P2G)
    assert np.all(s.grid_v == 0)
    exp = c.p_mass * (2 + 2.5)
    assert abs(s.grid_m.sum() - exp) < 1e-15


def test_water_stress_is_isotropic(c):
    """Water (mu = 0): the P2G force is a pure pressure; a sheared F still yields
    zero net torque on the grid about the particle."""
    s = oracle.OracleSim(1, 1, c, dtype=np.float64)
    s.load(0, np.array([[0.5004, 0.5007]]), np.array([0]))
    s.F[0, 0] = [[1.02, 0.05], [-0.03, 0.97]]
    s.run(1, stages=oracle.ST_RESET | oracle.ST_P2G)
    nodes = np.argwhere(s.grid_m[0] > 0)
    xs = nodes * c.dx - s.x[0, 0]
    f = s.grid_v[0][nodes[:, 0], nodes[:, 1]]
    torque = np.sum(xs[:, 0] * f[:, 1] - xs[:, 1] * f[:, 0])
    assert abs(torque) < 1e-18
    # and F is reset to sqrt(J) I (mpm_sim.py:221)
    J = np.linalg.det(np.array([[1.02, 0.05], [-0.03, 0.97]]))
    np.testing.assert_allclose(s.F[0, 0], np.eye(2) * np.sqrt(J), rtol=1e-12, atol=1e-15)


# --- grid op -------------------------------------------------------------------------------
@pytest.mark.parametrize("ny,H", [(128, 1.0), (256, 2.0)])
def test_grid_op_elementwise(ny, H):
    c = oracle.reference_consts(n_grid_y=ny, domain_height=H)
    rng = np.random.default_rng(2)
    s = oracle.OracleSim(1, 1, c, dtype=np.float64)
    s.material[:] = -1
    gm = rng.uniform(0, 1e-4, size=(128, ny))
    gm[rng.uniform(size=gm.shape) < 0.2] = 0
    gv = rng.normal(size=(128, ny, 2)) * 1e-5
    gv[gm == 0] = 0
    s.grid_m[0] = gm
    s.grid_v[0] = gv
    s.run(1, stages=oracle.ST_GRIDOP)
    # independent pass from the reference text (mpm_sim.py:274-309)
    exp = np.zeros_like(gv)
    for i in range(128):
        for j in range(ny):
            m = gm[i, j]
            if not m > 0:
                exp[i, j] = gv[i, j]
                continue
            v = gv[i, j] / m
            d = 1.0
            if i < 128 // 20:
                d *= 0.95 + 0.05 * i / (128 // 20)
            if i > 128 - 128 // 20:
                d *= 0.95 + 0.05 * (128 - i) / (128 // 20)
            if j < ny // 20:
                d *= 0.95 + 0.05 * j / (ny // 20)
            if j > ny - ny // 20:
                d *= 0.95 + 0.05 * (ny - j) / (ny // 20)
            v = v * d
            if i < 3 and v[0] < 0:
                v[0] = 0
            if i > 128 - 3 and v[0] > 0:
                v[0] = 0
            if j < 3 and v[1] < 0:
                v[1] = 0
            if j > ny - 3 and v[1] > 0:
                v[1] = 0
            exp[i, j] = v
    np.testing.assert_allclose(s.grid_v[0], exp, rtol=1e-14, atol=0)


# --- G2P -----------------------------------------------------------------------------------
def test_g2p_reproduces_affine_fields(c):
    rng = np.random.default_rng(3)
    n = 500
    s = oracle.OracleSim(1, n, c, dtype=np.float64)
    pos = rng.uniform(0.02, 0.97, size=(n, 2))
    s.load(0, pos, np.zeros(n, np.int32))
    s.x[0] = pos
    a = np.array([0.3, -0.2])
    B = np.array([[0.7, -1.1], [0.4, 0.25]])
    ii, jj = np.meshgrid(np.arange(128), np.arange(128), indexing="ij")
    xn = np.stack([ii, jj], -1) * c.dx
    s.grid_v[0] = a + xn @ B.T
    s.run(1, stages=oracle.ST_G2P)
    base = np.trunc(pos * 128 - 0.5)
    inr = np.all((base >= 0) & (base < 126), axis=1)
    v_exp = a + pos @ B.T
    v_exp[:, 1] -= c.dt * c.gravity
    np.testing.assert_allclose(s.v[0][inr], v_exp[inr], atol=1e-12)
    np.testing.assert_allclose(s.C[0][inr], np.broadcast_to(B, (inr.sum(), 2, 2)), atol=1e-10)
    x_exp = np.clip(pos + c.dt * s.v[0], 0.001, 0.999)
    np.testing.assert_allclose(s.x[0], x_exp, atol=1e-15)
    # out-of-range particles only get gravity and advection (mpm_sim.py:329-333)
    assert np.all(s.v[0][~inr][:, 0] == 0)


def test_trunc_cell_index(c):
    """`.cast(int)` truncates toward zero (mpm_sim.py:183): x*128-0.5 in (-1, 0) -> 0."""
    x = np.array([[0.001, 0.0039], [0.00390625, 0.5], [0.98828125, 0.988281]], np.float32)
    b = oracle.base_cells(c, x)
    np.testing.assert_array_equal(b, [[0, 0], [0, 63], [126, 125]])


# --- whole substep vs the independent fp64 model ---------------------------------------
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_substep_matches_brute_force(c, seed):
    rng = np.random.default_rng(10 + seed)
    n = [1, 7, 20, 50][seed]
    s = _random_instance(rng, n, c, sv=0.2, sc=0.5, sf=0.03, lo=0.2, hi=0.8)
    if seed == 2:
        s.x[0, :3, 0] = [0.9885, 0.995, 0.5]   # skip strip: no P2G/G2P, still advected
    s.t[0] = [0.05, 0.13, 0.71, 0.19][seed]
    state = {"x": s.x[0].copy(), "v": s.v[0].copy(), "C": s.C[0].copy(), "F": s.F[0].copy(),
             "Jp": s.Jp[0].copy(), "mat": s.material[0].copy(), "fiber": s.fiber[0].copy(),
             "t": float(s.t[0])}
    act = {"amp": 500.0, "contraction": 0.2, "freq": 1.0, "pdens": 2.5,
           "mu_j": float(s.mu_j[0]), "la_j": float(s.la_j[0])}
    ref = brute.substep(state, c, act)
    s.run(1)
    for k in ("x", "v", "C", "F", "Jp"):
        np.testing.assert_allclose(getattr(s, k)[0], ref[k], rtol=1e-9, atol=1e-12, err_msg=k)
    np.testing.assert_allclose(s.grid_m[0], ref["grid_m"], rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(s.grid_v[0], ref["grid_v"], rtol=1e-9, atol=1e-14)


def test_brute_force_detects_planted_mistake(c):
    """The brute-force comparison is sensitive: a sign flip in the affine term is caught."""
    rng = np.random.default_rng(20)
    s = _random_instance(rng, 10, c, lo=0.3, hi=0.7)
    state = {"x": s.x[0].copy(), "v": s.v[0].copy(), "C": -s.C[0].copy(), "F": s.F[0].copy(),
             "Jp": s.Jp[0].copy(), "mat": s.material[0].copy(), "fiber": s.fiber[0].copy(),
             "t": float(s.t[0])}
    act = {"amp": 500.0, "contraction": 0.2, "freq": 1.0, "pdens": 2.5,
           "mu_j": float(s.mu_j[0]), "la_j": float(s.la_j[0])}
    ref = brute.substep(state, c, act)
    s.run(1)
    assert not np.allclose(s.grid_v[0], ref["grid_v"], rtol=1e-6, atol=1e-14)


def test_fp32_restatement_error_bound(c):
    """fp32 vs fp64 oracle over 100 substeps of a small tank: the fp32 error stays small."""
    from tests.cases import small_tank
    pos, mat, fib = small_tank(7, 6000, region=(0.3, 0.6, 0.3, 0.5))
    out = {}
    for dt in (np.float32, np.float64):
        s = oracle.OracleSim(1, 6000, c, dtype=dt)
        s.load(0, pos, mat, fib)
        s.t[0] = 0.1
        s.run(100)
        out[dt] = s
    live = mat >= 0
    ex = np.max(np.abs(out[np.float32].x[0][live] - out[np.float64].x[0][live]))
    ev = np.max(np.abs(out[np.float32].v[0][live] - out[np.float64].v[0][live]))
    assert ex < 1e-6, ex
    assert ev < 1e-3, ev


# --- stats, headless, fitness, sort -------------------------------------------------------
def test_stats_and_headless_closed_forms(c):
    s = oracle.OracleSim(3, 4, c)
    s.load(0, [[0.4, 0.5], [0.6, 0.7], [0.1, 0.1], [0.2, 0.2]], [2, 2, 0, -1])
    s.load(1, [[0.3, 0.2], [0.5, 0.4], [0.1, 0.1], [0.2, 0.2]], [1, 3, 0, 0])   # no payload
    s.load(2, [[0.5, 0.005], [0.5, 0.007], [0.1, 0.1], [0.2, 0.2]], [2, 2, 0, 0])
    raw = s.raw_stats(False)
    np.testing.assert_allclose(raw[0], [1.2, 2, 1.0], rtol=1e-6)
    assert raw[1, 1] == 0
    st = s.get_stats_any()
    np.testing.assert_allclose(st[0], [0.6, 0.5, 2], rtol=1e-6)
    np.testing.assert_allclose(st[1], [0.3, 0.4, 2], rtol=1e-6)    # body CoM fallback
    res = oracle.assemble_results(st, st)
    assert res[0, 4] == 1 and res[1, 4] == 1
    assert res[2, 4] == 0                                          # payload below y = 0.01
    nan = st.copy()
    nan[0, 0] = np.nan
    assert oracle.assemble_results(st, nan)[0, 4] == 1             # NaN passes (H12)


def test_compute_fitness_closed_form():
    res = np.array([[0.40, 0.5, 0.55, 0.5, 1.0],
                    [0.40, 0.5, 0.99, 0.5, 1.0],    # capped at H - 0.07
                    [0.40, 0.5, 0.30, 0.5, 0.0],    # invalid
                    [0.40, 0.5, 0.50, 0.5, 1.0]])   # too few muscles
    muscles = [500, 300, 400, 150]
    genomes = [np.r_[np.zeros(9), 0.20, 1.0], np.r_[np.zeros(9), 0.9, 1.0], None, None]
    f = oracle.compute_fitness(res, muscles, genomes)
    assert f[0] == pytest.approx(-0.15 / np.sqrt(500 * 0.2 / 500))
    assert f[1] == pytest.approx(-(0.93 - 0.40) / np.sqrt(300 * 0.6 / 500))
    worst = max(f[0], f[1])
    assert f[2] == pytest.approx(worst + 1.0) and f[3] == pytest.approx(worst + 1.0)
    assert oracle.compute_fitness(res[2:], muscles[2:], None) == [100.0, 100.0]


def _walk_order(nx, ny, strip):
    """Rank of every cell (bx, by) in the documented sort order, built by walking it."""
    rank, n = {}, 0
    if strip <= 0:
        for by in range(ny):
            for bx in range(nx):
                rank[bx, by] = n
                n += 1
        return rank
    for s in range((ny + strip - 1) // strip):
        cols = range(nx) if s % 2 == 0 else range(nx - 1, -1, -1)
        for bx in cols:
            for by in range(s * strip, min((s + 1) * strip, ny)):
                rank[bx, by] = n
                n += 1
    return rank


@pytest.mark.parametrize("strip", [0, 8, 5])
def test_stable_cell_sort_is_a_stable_counting_sort(c, strip):
    rng = np.random.default_rng(4)
    x = rng.uniform(0.0, 1.0, size=(3000, 2)).astype(np.float32)
    x[::7] = x[3]                                      # many ties
    keys = oracle.cell_keys(c, x, strip)
    perm = oracle.stable_cell_sort(c, x, strip)
    brute_perm = sorted(range(len(x)), key=lambda i: (int(keys[i]), i))
    np.testing.assert_array_equal(perm, brute_perm)
    b = np.clip(np.trunc(x.astype(np.float64) * 128 - 0.5), 0, 127).astype(int)
    rank = _walk_order(128, 128, strip)
    order = np.array([rank[bx, by] for bx, by in b])
    if strip in (0, 8):   # 128 rows divide into strips: the keys are exactly the ranks
        np.testing.assert_array_equal(keys, order)
    else:                 # a short last strip leaves gaps in the keys, not in their order
        np.testing.assert_array_equal(np.argsort(keys, kind="stable"),
                                      np.argsort(order, kind="stable"))
