"""TEST INFRASTRUCTURE — independent fp64 brute-force model of one reference substep.

Written from the reference text (c-m-x-r/jellyfih ``mpm_sim.py:153-340``) for a handful
of particles in one instance, with a different structure from the C++ restatement so a
shared transcription mistake is unlikely:

* the SVD is numpy's LAPACK ``np.linalg.svd`` (the polar rotation ``R = U Vᵀ`` and the
  clamped reconstruction ``U Σ' Vᵀ`` do not depend on the SVD's sign/order convention
  when ``det F > 0``), instead of the restated taichi ``_svd2d``;
* the B-spline weights and the grid node loop are vectorised with numpy;
* the grid op is a dense vectorised pass.

Used only to pin ``liboracle.so``'s fp64 substep on states with ``det F > 0``, in the planar
and the axisymmetric variant (``mpm_sim.py:127-133,163-165,236-304``).
"""
from __future__ import annotations

import numpy as np


def _weights(fx: np.ndarray) -> np.ndarray:
    """Quadratic B-spline weights for offsets 0,1,2 (mpm_sim.py:187); fx shape (2,)."""
    return np.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2])


def substep(state: dict, c, act_params: dict, axisym: bool = False) -> dict:
    """One substep of one instance, fp64. ``state`` keys: x (N,2) v (N,2) C (N,2,2) F (N,2,2)
    mat (N,) Jp (N,) fiber (N,2), t (float), and r_ref (N,) for the axisymmetric variant.
    ``c`` is an OracleConsts. ``act_params`` keys: amp, contraction, freq, pdens, mu_j, la_j.
    Returns the new state and the grid."""
    nx, ny = c.nx, c.ny
    dx, inv_dx, dt = c.dx, c.inv_dx, c.dt
    x = state["x"].astype(np.float64).copy()
    v = state["v"].astype(np.float64).copy()
    C = state["C"].astype(np.float64).copy()
    F = state["F"].astype(np.float64).copy()
    Jp = state["Jp"].astype(np.float64).copy()
    mat = state["mat"]
    fiber = state["fiber"].astype(np.float64)
    t = float(state["t"])

    period = 1.0 / (c.act_freq * act_params["freq"])
    phase = (t - period * np.floor(t / period)) / period
    cf = act_params["contraction"]
    a = 0.5 * (1.0 - np.cos(phase / cf * 3.14159265)) if phase < cf else 0.0

    gv = np.zeros((nx, ny, 2))
    gm = np.zeros((nx, ny))
    s_rr = np.zeros((nx, ny))       # axisymmetric: mass-weighted Kirchhoff sigma_rr
    s_tt = np.zeros((nx, ny))       # and sigma_theta_theta
    for p in range(x.shape[0]):
        if mat[p] < 0:
            continue
        base = np.trunc(x[p] * inv_dx - 0.5).astype(int)
        if base[0] < 0 or base[1] < 0 or base[0] >= nx - 2 or base[1] >= ny - 2:
            continue
        fx = x[p] * inv_dx - base
        w = _weights(fx)
        F[p] = (np.eye(2) + dt * C[p]) @ F[p]
        mu, la, mass = c.mu_base, c.la_base, c.p_mass
        if mat[p] in (1, 3):
            mu, la = act_params["mu_j"], act_params["la_j"]
        elif mat[p] == 2:
            mu, la, mass = c.mu_payload, c.la_payload, c.p_mass * act_params["pdens"]
        elif mat[p] == 0:
            mu, la = 0.0, c.water_la
        U, sig, Vt = np.linalg.svd(F[p])
        if np.linalg.det(U) < 0:      # make U, V proper rotations (det F > 0 assumed)
            U[:, 1] *= -1
            Vt[1, :] *= -1
        sig = np.maximum(sig, 1e-6)
        new_sig = np.clip(sig, 1 - 5e-3, 1 + 5e-3) if mat[p] == 2 else sig
        Jp[p] *= np.prod(sig / new_sig)
        J = float(np.prod(new_sig))
        if mat[p] == 0:
            F[p] = np.eye(2) * np.sqrt(J)
        elif mat[p] == 2:
            F[p] = U @ np.diag(new_sig) @ Vt
        R = U @ Vt
        stress = 2 * mu * (F[p] - R) @ F[p].T + np.eye(2) * la * J * (J - 1)
        if mat[p] == 3:
            stress += np.outer(fiber[p], fiber[p]) * act_params["amp"] * a * J
        ring = 1.0
        if axisym:   # the particle is a ring of radius r = x; hoop stretch r / r_ref
            ring = x[p, 0]
            lam = max(x[p, 0], 0.05 * dx) / max(state["r_ref"][p], 0.1 * dx)
            k_rr, k_tt = stress[0, 0], 2 * mu * (lam - 1) + la * J * (J - 1)
        stress = (-dt * c.p_vol * 4 * inv_dx * inv_dx) * stress
        affine = stress + mass * C[p]
        for i in range(3):
            for j in range(3):
                off = np.array([i, j])
                dpos = (off - fx) * dx
                wt = w[i, 0] * w[j, 1]
                node = (base[0] + i, base[1] + j)
                gv[node] += wt * (mass * ring * v[p] + affine @ dpos)
                gm[node] += wt * mass * ring
                if axisym:
                    s_rr[node] += wt * mass * k_rr
                    s_tt[node] += wt * mass * k_tt

    # grid op (mpm_sim.py:273-309), vectorised over nodes
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    has = gm > 0
    g = np.where(has[..., None], gv / np.where(has, gm, 1.0)[..., None], gv)
    bx, by = nx // 20, ny // 20
    damp = np.ones((nx, ny))
    damp = np.where(ii < bx, damp * (0.95 + 0.05 * ii / bx), damp)
    damp = np.where(ii > nx - bx, damp * (0.95 + 0.05 * (nx - ii) / bx), damp)
    damp = np.where(jj < by, damp * (0.95 + 0.05 * jj / by), damp)
    damp = np.where(jj > ny - by, damp * (0.95 + 0.05 * (ny - jj) / by), damp)
    g = np.where(has[..., None], g * damp[..., None], g)
    g0, g1 = g[..., 0], g[..., 1]
    if axisym:
        r_node = (ii + 0.5) * dx
        hoop = (s_rr - s_tt) / np.where(has, gm, 1.0)
        g0 = np.where(has & (r_node > dx), g0 + dt * hoop / r_node, g0)
        g0 = np.where(has & (ii == 0), 0.0, g0)
    else:
        g0 = np.where(has & (ii < 3) & (g0 < 0), 0.0, g0)
    g0 = np.where(has & (ii > nx - 3) & (g0 > 0), 0.0, g0)
    g1 = np.where(has & (jj < 3) & (g1 < 0), 0.0, g1)
    g1 = np.where(has & (jj > ny - 3) & (g1 > 0), 0.0, g1)
    g = np.stack([g0, g1], -1)

    # G2P (mpm_sim.py:311-333)
    for p in range(x.shape[0]):
        if mat[p] < 0:
            continue
        base = np.trunc(x[p] * inv_dx - 0.5).astype(int)
        if base[0] >= 0 and base[1] >= 0 and base[0] < nx - 2 and base[1] < ny - 2:
            fx = x[p] * inv_dx - base
            w = _weights(fx)
            nv = np.zeros(2)
            nC = np.zeros((2, 2))
            for i in range(3):
                for j in range(3):
                    dpos = np.array([i, j]) - fx
                    gvv = g[base[0] + i, base[1] + j]
                    wt = w[i, 0] * w[j, 1]
                    nv += wt * gvv
                    nC += 4 * inv_dx * wt * np.outer(gvv, dpos)
            v[p], C[p] = nv, nC
        v[p, 1] -= dt * c.gravity
        x[p] += dt * v[p]
        x[p, 0] = max(min(x[p, 0], 0.999), 0.001)
        x[p, 1] = max(min(x[p, 1], c.domain_height - 0.001), 0.001)

    return {"x": x, "v": v, "C": C, "F": F, "Jp": Jp, "mat": mat, "fiber": fiber,
            "t": t + dt, "grid_v": g, "grid_m": gm}
