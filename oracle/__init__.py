"""TEST INFRASTRUCTURE — the oracle. Not part of the product path.

A CPU restatement of the reference's batched 2D MLS-MPM simulation core
(c-m-x-r/jellyfih ``mpm_sim.py``), used only as the reference results of ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg:

* ``liboracle.so`` (``mpm_oracle.cpp``): the substep in fp32 and fp64, following
  ``mpm_sim.py:153-340`` line by line; stats ``mpm_sim.py:740-762``.
* :class:`OracleSim` (this file): the reference module's host logic around it —
  instance load (``mpm_sim.py:715-736``), ``_get_stats_any`` (``:779-800``),
  ``run_batch_headless`` (``:802-834``) — and the evolve.py fitness closed form
  (``evolve.py:90-124``).
* ``brute.py``: an independent fp64 numpy model for tiny particle counts that uses
  a library SVD, to pin the C++ restatement.
* :func:`stable_cell_sort`: the stable per-instance counting-sort permutation by P2G
  base cell, which the GPU sort must reproduce bit-exactly.

Every function here cites the reference line it follows. Parity status of each part is
recorded in the tests that pin it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# Stage bits of oracle_run_* (one substep = RESET|P2G|GRIDOP|G2P [+GLOW]).
ST_RESET, ST_P2G, ST_GRIDOP, ST_G2P, ST_GLOW = 1, 2, 4, 8, 16
ST_SUBSTEP = ST_RESET | ST_P2G | ST_GRIDOP | ST_G2P


class OracleConsts(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("track_glow", ctypes.c_int32),
        ("axisym", ctypes.c_int32),
        ("dt", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("inv_dx", ctypes.c_double),
        ("p_vol", ctypes.c_double),
        ("p_mass", ctypes.c_double),
        ("gravity", ctypes.c_double),
        ("mu_base", ctypes.c_double),
        ("la_base", ctypes.c_double),
        ("mu_payload", ctypes.c_double),
        ("la_payload", ctypes.c_double),
        ("water_la", ctypes.c_double),
        ("act_freq", ctypes.c_double),
        ("domain_height", ctypes.c_double),
    ]


def reference_consts(n_grid_y: int | None = None, domain_height: float = 1.0,
                     quality: int = 1, track_glow: bool = False) -> OracleConsts:
    """Constants of ``mpm_sim.py:12-58`` evaluated in Python f64 as the reference does."""
    n_grid = 128 * quality                      # :14
    ny = n_grid if n_grid_y is None else int(n_grid_y)   # :15
    dx, inv_dx = 1 / n_grid, float(n_grid)      # :18
    dt = 2e-5 / quality                         # :19
    p_vol = (dx * 0.5) ** 2                     # :20
    p_mass = p_vol * 1                          # :21-22
    E_base, nu_base = 0.1e4, 0.2                # :25-26
    E_payload, nu_payload = 2.0e5, 0.2          # :43-44
    c = OracleConsts()
    c.nx, c.ny, c.track_glow = n_grid, ny, int(bool(track_glow))
    c.dt, c.dx, c.inv_dx, c.p_vol, c.p_mass = dt, dx, inv_dx, p_vol, p_mass
    c.gravity = 10.0                            # :49
    c.mu_base = E_base / (2 * (1 + nu_base))    # :27
    c.la_base = E_base * nu_base / ((1 + nu_base) * (1 - 2 * nu_base))   # :28
    c.mu_payload = E_payload / (2 * (1 + nu_payload))                    # :45
    c.la_payload = E_payload * nu_payload / ((1 + nu_payload) * (1 - 2 * nu_payload))  # :46
    c.water_la = 100000.0                       # :48
    c.act_freq = 1.0                            # :52
    c.domain_height = float(domain_height)      # :16
    return c


def jelly_lame(E: float | np.ndarray, nu: float = 0.3):
    """(mu, lambda) of the jelly/muscle material from Young's modulus (mpm_sim.py:31-34)."""
    E = np.asarray(E, dtype=np.float64)
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for suffix, ct in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            P = ctypes.POINTER(ct)
            fn = getattr(L, f"oracle_run_{suffix}")
            fn.restype = ctypes.c_int
            fn.argtypes = ([ctypes.POINTER(OracleConsts), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            P, P, P, P, ctypes.POINTER(ctypes.c_int32), P, P, P, P, P, P,
                            P, P, P, P, P, P, ctypes.c_int, ctypes.c_int])
            fn = getattr(L, f"oracle_run_axisym_{suffix}")
            fn.restype = ctypes.c_int
            fn.argtypes = ([ctypes.POINTER(OracleConsts), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            P, P, P, P, ctypes.POINTER(ctypes.c_int32), P, P, P, P, P, P,
                            P, P, P, P, P, P, P, P, P, ctypes.c_int, ctypes.c_int])
            fn = getattr(L, f"oracle_activation_{suffix}")
            fn.restype = None
            fn.argtypes = [ctypes.POINTER(OracleConsts), ctypes.c_int, P, P, P, P]
            fn = getattr(L, f"oracle_svd2d_{suffix}")
            fn.restype = None
            fn.argtypes = [ctypes.c_int, P, P, P, P, P]
            fn = getattr(L, f"oracle_stats_{suffix}")
            fn.restype = None
            fn.argtypes = [ctypes.c_int, ctypes.c_int, P, ctypes.POINTER(ctypes.c_int32),
                           ctypes.c_int, P]
        L.oracle_base_cell_f32.restype = None
        L.oracle_base_cell_f32.argtypes = [ctypes.POINTER(OracleConsts), ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_int32)]
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None, ct):
    if a is None:
        return ctypes.POINTER(ct)()
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _suffix(dtype) -> tuple[str, type]:
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f32", ctypes.c_float
    if dtype == np.float64:
        return "f64", ctypes.c_double
    raise TypeError(dtype)


def activation(consts: OracleConsts, t, freq, contraction, dtype=np.float32) -> np.ndarray:
    """Raised-cosine activation (mpm_sim.py:173-181), vectorised over instances."""
    sfx, ct = _suffix(dtype)
    t = np.ascontiguousarray(np.atleast_1d(t), dtype=dtype)
    n = t.size
    freq = np.ascontiguousarray(np.broadcast_to(freq, (n,)), dtype=dtype)
    contraction = np.ascontiguousarray(np.broadcast_to(contraction, (n,)), dtype=dtype)
    out = np.empty(n, dtype=dtype)
    getattr(lib(), f"oracle_activation_{sfx}")(ctypes.byref(consts), n, _ptr(t, ct),
                                               _ptr(freq, ct), _ptr(contraction, ct),
                                               _ptr(out, ct))
    return out


def svd2d(F: np.ndarray, dtype=np.float32):
    """taichi `_svd2d` restated: F (n,2,2) -> U (n,2,2), sig (n,2), V (n,2,2), R (n,2,2)."""
    sfx, ct = _suffix(dtype)
    F = np.ascontiguousarray(F, dtype=dtype).reshape(-1, 4)
    n = F.shape[0]
    U = np.empty((n, 4), dtype)
    V = np.empty((n, 4), dtype)
    R = np.empty((n, 4), dtype)
    sig = np.empty((n, 2), dtype)
    getattr(lib(), f"oracle_svd2d_{sfx}")(n, _ptr(F, ct), _ptr(U, ct), _ptr(sig, ct),
                                          _ptr(V, ct), _ptr(R, ct))
    return U.reshape(n, 2, 2), sig, V.reshape(n, 2, 2), R.reshape(n, 2, 2)


def base_cells(consts: OracleConsts, x: np.ndarray) -> np.ndarray:
    """P2G base cell per particle: trunc(x*inv_dx - 0.5) (mpm_sim.py:183), int32 (n,2)."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 2)
    out = np.empty((x.shape[0], 2), np.int32)
    lib().oracle_base_cell_f32(ctypes.byref(consts), x.shape[0], _ptr(x, ctypes.c_float),
                               _ptr(out, ctypes.c_int32))
    return out


def cell_keys(consts: OracleConsts, x: np.ndarray, strip: int = 8) -> np.ndarray:
    """Sort key of each particle's P2G base cell, clamped into the grid (the engine's sort
    order; the reference does not sort). ``strip`` = 0: row-major cell index. ``strip`` =
    S > 0: strips of S cell rows, one after the other upwards; columns left to right in
    even strips and right to left in odd ones; the S cells of a column upwards."""
    b = base_cells(consts, x)
    bx = np.clip(b[:, 0], 0, consts.nx - 1).astype(np.int64)
    by = np.clip(b[:, 1], 0, consts.ny - 1).astype(np.int64)
    if strip <= 0:
        return (by * consts.nx + bx).astype(np.int32)
    s, r = by // strip, by % strip
    xs = np.where(s % 2 == 1, consts.nx - 1 - bx, bx)
    return ((s * consts.nx + xs) * strip + r).astype(np.int32)


def stable_cell_sort(consts: OracleConsts, x: np.ndarray, strip: int = 8) -> np.ndarray:
    """Permutation that stably sorts one instance's particles by ``cell_keys``.

    ``perm[i]`` is the current index of the particle that goes to position ``i``; ties
    keep their current relative order (a stable counting sort)."""
    return np.argsort(cell_keys(consts, x, strip), kind="stable").astype(np.int32)


class OracleSim:
    """The reference module's state and host logic on the CPU (mpm_sim.py:86-149,715-834).

    Arrays use the reference layout: x, v, fiber (M,N,2); C, F (M,N,2,2); material, Jp (M,N);
    grid_v (M,nx,ny,2); grid_m (M,nx,ny)."""

    def __init__(self, M: int, N: int, consts: OracleConsts | None = None, dtype=np.float32,
                 axisym: bool = False):
        self.c = OracleConsts.from_buffer_copy(consts) if consts is not None \
            else reference_consts()
        self.c.axisym = int(bool(axisym))
        self.M, self.N = int(M), int(N)
        self.dtype = np.dtype(dtype)
        nx, ny = self.c.nx, self.c.ny
        d = self.dtype
        self.x = np.zeros((M, N, 2), d)
        self.v = np.zeros((M, N, 2), d)
        self.C = np.zeros((M, N, 2, 2), d)
        self.F = np.zeros((M, N, 2, 2), d)
        self.material = np.zeros((M, N), np.int32)
        self.Jp = np.zeros((M, N), d)
        self.fiber = np.zeros((M, N, 2), d)
        self.glow = np.zeros((M, N), d)
        self.grid_v = np.zeros((M, nx, ny, 2), d)
        self.grid_m = np.zeros((M, nx, ny), d)
        self.t = np.zeros(1, d)
        # axisymmetric variant (mpm_sim.py:127-133): spawn radius and the hoop-term grids
        self.axisym = bool(axisym)
        self.r_ref = np.zeros((M, N), d)
        self.grid_srr = np.zeros((M, nx, ny), d) if axisym else None
        self.grid_shoop = np.zeros((M, nx, ny), d) if axisym else None
        # mpm_sim.py:143-149 defaults
        self.amp = np.full(M, 500.0, d)
        self.contraction = np.full(M, 0.20, d)
        self.freq = np.full(M, 1.0, d)
        self.pdens = np.full(M, 2.5, d)
        self.set_e_jelly(np.full(M, 0.7e3))

    def set_e_jelly(self, E):
        mu, la = jelly_lame(np.broadcast_to(np.asarray(E, np.float64), (self.M,)))
        self.mu_j = mu.astype(self.dtype)
        self.la_j = la.astype(self.dtype)

    def load(self, m: int, pos, mat, fiber=None):
        """mpm_sim.py:715-736; the default fibre is (0, 1)."""
        N = self.N
        if fiber is None:
            fiber = np.zeros((N, 2), np.float32)
            fiber[:, 1] = 1.0
        self.x[m] = np.asarray(pos, np.float32)[:N].astype(self.dtype)
        self.material[m] = np.asarray(mat)[:N].astype(np.int32)
        self.v[m] = 0
        self.F[m] = np.eye(2, dtype=self.dtype)
        self.C[m] = 0
        self.Jp[m] = 1
        self.fiber[m] = np.asarray(fiber, np.float32)[:N].astype(self.dtype)
        self.r_ref[m] = self.x[m, :, 0]                      # mpm_sim.py:727

    def run(self, nsteps: int = 1, stages: int = ST_SUBSTEP, nthreads: int = 1) -> None:
        sfx, ct = _suffix(self.dtype)
        glow = self.glow if (stages & ST_GLOW) else None
        args = [ctypes.byref(self.c), stages, self.M, self.N, _ptr(self.x, ct), _ptr(self.v, ct),
                _ptr(self.C, ct), _ptr(self.F, ct), _ptr(self.material, ctypes.c_int32),
                _ptr(self.Jp, ct), _ptr(self.fiber, ct), _ptr(glow, ct), _ptr(self.grid_v, ct),
                _ptr(self.grid_m, ct), _ptr(self.t, ct), _ptr(self.amp, ct),
                _ptr(self.contraction, ct), _ptr(self.freq, ct), _ptr(self.pdens, ct),
                _ptr(self.mu_j, ct), _ptr(self.la_j, ct)]
        if self.axisym:
            fn = getattr(lib(), f"oracle_run_axisym_{sfx}")
            args += [_ptr(self.r_ref, ct), _ptr(self.grid_srr, ct), _ptr(self.grid_shoop, ct)]
        else:
            fn = getattr(lib(), f"oracle_run_{sfx}")
        rc = fn(*args, int(nsteps), int(nthreads))
        if rc != 0:
            raise RuntimeError(f"oracle_run_{sfx} failed ({rc})")

    def raw_stats(self, body: bool = False) -> np.ndarray:
        """fitness_buffer after clear + compute_{payload,body}_stats (mpm_sim.py:740-762)."""
        sfx, ct = _suffix(self.dtype)
        out = np.zeros((self.M, 3), self.dtype)
        getattr(lib(), f"oracle_stats_{sfx}")(self.M, self.N, _ptr(self.x, ct),
                                              _ptr(self.material, ctypes.c_int32),
                                              int(body), _ptr(out, ct))
        return out

    def get_stats_any(self) -> np.ndarray:
        """mpm_sim.py:779-800: payload CoM, body CoM where an instance has no payload."""
        raw = self.raw_stats(False)
        no_payload = raw[:, 1] == 0
        if no_payload.any():
            raw[no_payload] = self.raw_stats(True)[no_payload]
        return finish_stats(raw)

    def run_batch_headless(self, steps: int, nthreads: int = 1,
                           nan_invalid: bool = False) -> np.ndarray:
        """mpm_sim.py:802-834."""
        self.t[0] = 0.0
        initial = self.get_stats_any()
        self.run(steps, nthreads=nthreads)
        final = self.get_stats_any()
        return assemble_results(initial, final, nan_invalid=nan_invalid,
                                diverged=self.diverged())

    def diverged(self) -> np.ndarray:
        """Per instance: whether a live particle holds a non-finite v, C or F."""
        live = self.material >= 0
        bad = ~(np.isfinite(self.v).all(-1) & np.isfinite(self.C).all((-2, -1)) &
                np.isfinite(self.F).all((-2, -1)))
        return (bad & live).any(axis=1)


def finish_stats(raw: np.ndarray) -> np.ndarray:
    """Host division of the fitness buffer into [com_y, com_x, count] (mpm_sim.py:793-800).
    The reference's buffer is an fp32 array, so each quotient is rounded to the buffer's
    precision before it is stored in the float64 table."""
    M = raw.shape[0]
    stats = np.zeros((M, 3))
    for i in range(M):
        count = raw[i, 1]
        if count > 0:
            stats[i, 0] = raw[i, 0] / count
            stats[i, 1] = raw[i, 2] / count
            stats[i, 2] = count
    return stats


def assemble_results(initial: np.ndarray, final: np.ndarray, nan_invalid: bool = False,
                     diverged=None) -> np.ndarray:
    """(M,5) = [init_y, init_x, final_y, final_x, valid] (mpm_sim.py:821-834). The
    reference's gate ``final_y < 0.01`` passes a NaN CoM as valid; ``nan_invalid`` (the
    engine's opt-in flag) marks an instance invalid when its CoM is not finite or
    ``diverged[i]`` (a live particle with a non-finite v, C or F) holds."""
    M = initial.shape[0]
    results = np.zeros((M, 5))
    for i in range(M):
        if initial[i, 2] > 0 and final[i, 2] > 0:
            results[i, 0] = initial[i, 0]
            results[i, 1] = initial[i, 1]
            results[i, 2] = final[i, 0]
            results[i, 3] = final[i, 1]
            bad = nan_invalid and (not np.all(np.isfinite(results[i, :4])) or
                                   (diverged is not None and bool(diverged[i])))
            results[i, 4] = 0.0 if (final[i, 0] < 0.01 or bad) else 1.0
    return results


def compute_fitness(sim_results, muscle_counts, genomes=None, domain_height=1.0,
                    mode="efficiency"):
    """evolve.py:90-124 (constants evolve.py:66-68)."""
    raw = []
    for i in range(len(muscle_counts)):
        init_y, final_y, valid = sim_results[i, 0], sim_results[i, 2], sim_results[i, 4]
        if valid == 0 or muscle_counts[i] < 200:
            raw.append(None)
            continue
        displacement = min(final_y, domain_height - 0.07) - init_y
        if mode == "displacement":
            raw.append(-displacement)
        else:
            if genomes is not None and len(genomes[i]) > 9:
                active_frac = float(np.clip(genomes[i][9], 0.05, 0.60))
            else:
                active_frac = 0.20
            muscle_cost = (muscle_counts[i] * active_frac / 500) ** 0.5
            raw.append(-displacement / muscle_cost)
    valid_scores = [s for s in raw if s is not None]
    penalty = (max(valid_scores) + 1.0) if valid_scores else 100.0
    return [penalty if s is None else s for s in raw]
