"""Shared test helpers: seeded perturbed model states and field comparisons."""
from __future__ import annotations

import os

import numpy as np

import vvm_amd as V

PROG = ["zeta", "th", "qv", "qc", "qr"]
COMPARE_STEP = ["zeta", "zetam", "th", "thm", "qv", "qvm", "qc", "qcm", "qr", "qrm", "u", "w", "RKM", "RKH",
                "dth_advect", "dth_buoyancy", "dzeta_advect", "dqv_advect", "dqc_advect", "dqr_advect", "dqr_VT",
                "precip", "uxi", "thb", "qvb", "thvb", "evaporation", "accretion", "autoconversion",
                "condensation"]


def _bc_center(a):
    a[0, 1:-1] = a[-2, 1:-1]
    a[-1, 1:-1] = a[1, 1:-1]
    a[:, 0] = a[:, 1]
    a[:, -1] = a[:, -2]


def _bc_westdown(a):
    a[0, 1:-1] = a[-2, 1:-1]
    a[-1, 1:-1] = a[1, 1:-1]
    a[:, 0] = 0.
    a[:, 1] = 0.
    a[:, -1] = 0.


def perturbed_state(case: V.Case, seed: int = 7, step: int = 3, consistent_ghosts: bool = True,
                    negatives: bool = True) -> V.HostState:
    """A state that exercises every branch of the step: non-zero winds,
    cloud and rain (some negative), random AB2 histories and K fields."""
    st = V.init_state(case)
    rng = np.random.default_rng(seed)
    nx, nz = st.nx, st.nz
    sh = (nx, nz)
    a = st.a
    a["zeta"][:] = rng.normal(0, 2e-3, sh)
    a["u"][:] = rng.normal(0, 3.0, sh)
    a["w"][:] = rng.normal(0, 1.5, sh)
    a["th"][:] += rng.normal(0, 0.5, sh)
    a["qv"][:] *= 1 + rng.normal(0, 0.05, sh)
    a["qc"][:] = rng.normal(6e-4, 8e-4 if negatives else 3e-4, sh)
    # rain stays >= 0 (its fall speed is qr^0.1346); cloud may dip below 0
    a["qr"][:] = np.abs(rng.normal(4e-4, 6e-4, sh))
    a["RKM"][:] = rng.uniform(10, 300, sh)
    a["RKH"][:] = rng.uniform(10, 300, sh)
    for n in V.HISTS:
        a[n][:] = rng.normal(0, 1e-5, (nx, nz, 2))
    for f in PROG:
        a[f + "m"][:] = a[f] + rng.normal(0, 1e-6, sh)
        a[f + "p"][:] = rng.normal(0, 1, sh)  # overwritten by the step
    if consistent_ghosts:
        for f in ("th", "qv", "qc", "qr", "u", "RKM", "RKH"):
            _bc_center(a[f])
        _bc_westdown(a["zeta"])
        _bc_westdown(a["w"])
    st.s["ubarTop"] = 0.3
    st.s["ubarTopp"] = 0.31
    st.s["ubarTopm"] = 0.29
    st.dubarTop_advect = [1e-4, -2e-4]
    st.step = step
    return st


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    nb = float(np.linalg.norm(b))
    d = float(np.linalg.norm(np.asarray(a) - np.asarray(b)))
    if nb == 0.0:
        return 0.0 if d < 1e-300 else float("inf")
    return d / nb


# Diagnostics judged against the field they feed rather than their own norm:
# uxi (the mean-free x-derivative of the top-row G solution) enters u additively
# as u_top = uxi + ubarTopp (src/PoissonSolver.cpp:620-640).  It is formed from
# w one row below the lid, where w is small, so a round-off of the elliptic
# solve at the size of w in the interior is a large part of uxi when the flow
# near the lid is weak (a quiet start: ||uxi|| ~ 3e-4 ||u||).
SCALE_OF = {"uxi": "u"}


def assert_fields_close(got: V.HostState, want: V.HostState, names, tol: float, where: str = ""):
    bad = []
    for n in names:
        if n in SCALE_OF:
            ref = max(float(np.linalg.norm(want[n])), float(np.linalg.norm(want[SCALE_OF[n]])))
            d = float(np.linalg.norm(np.asarray(got[n]) - np.asarray(want[n])))
            e = d / ref if ref > 0 else (0.0 if d < 1e-300 else float("inf"))
        else:
            e = rel_err(got[n], want[n])
        if not (e <= tol):
            bad.append(f"{n}: rel {e:.3e}")
    assert not bad, f"{where} fields above {tol:g}: " + ", ".join(bad)


def gpu_run_op(dev: V.Device, st: V.HostState, op: str) -> V.HostState:
    out = st.copy()
    dev.upload(st)
    dev.op(op)
    dev.download(out)
    return out


SOUNDING = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "init_sounding.txt")


def branch_cases():
    """The step's branches that the benchmark cases leave off, each on a
    grid deep enough for the nudged layer (z > 12 km) where that matters:
      kick_fine   -- the per-step random theta kick (addforcingtime = 10 s,
                     the reference's own value, src/main.cpp:70) on a 50 m
                     grid with four kicked levels below 200 m;
      kick_c1     -- the kick on a c1-style 200 m grid (steps 1-5 kicked);
      nudge_qv    -- Nudge_qv with moisture_nudge_time = 3600 s
                     (src/NumericalProcess.cpp:87-96);
      case2       -- CASE 2, which skips Nudge_zeta (src/Iteration.cpp:416);
      refdefault  -- the reference's default run: Init1d from the sounding
                     file input/init.txt, Init2d's TROPICALFORCING branch,
                     dt = 3 s, dx = dz = 200 m, addforcingtime = 10 s.
    Returns name -> (cfg, initial HostState, steps)."""
    out = {}
    c = V.grid_case(24, 16, dx=50.0, dt=1.0, addforcingtime=10.0)
    out["kick_fine"] = (c.cfg, V.init_state(c), 12)
    c = V.grid_case(32, 64, dx=200.0, dt=2.0, addforcingtime=10.0)
    out["kick_c1"] = (c.cfg, V.init_state(c), 8)
    c = V.grid_case(32, 64, dx=200.0, dt=2.0, moisture_nudge_time=3600.0)
    out["nudge_qv"] = (c.cfg, perturbed_state(c, seed=31, step=0), 8)
    c = V.grid_case(32, 64, dx=200.0, dt=2.0, CASE=2)
    out["case2"] = (c.cfg, perturbed_state(c, seed=32, step=0), 8)
    cfg = V.make_config(32, 100, 200.0, 200.0, 3.0, CASE=1, addforcingtime=10.0, tropical_forcing=1)
    out["refdefault"] = (cfg, V.init_reference(cfg, SOUNDING), 10)
    return out


def input_checksums(st: V.HostState) -> np.ndarray:
    """Per-array sums of |value| of a state, in sorted name order: the
    tests regenerate each initial state and compare these first."""
    return np.array([float(np.abs(st.a[n]).sum()) for n in sorted(st.a)])


def single_case():
    """A perturbed state on a grid tall enough for the nudged layer (z > 12
    km), with p-fields near the now-fields (some cloud negative) and qv
    nudging on: every single operator changes something."""
    c = V.grid_case(24, 64, dx=200.0, dt=2.0, moisture_nudge_time=3600.0)
    st = perturbed_state(c, seed=77, step=5)
    rng = np.random.default_rng(78)
    for f in ("zeta", "th", "qv", "qc", "qr"):
        st[f + "p"][:] = st[f] * (1 + rng.normal(0, 1e-3, st[f].shape))
    st["qcp"][:] -= 2e-5
    st["qc"][:] -= 1e-4
    st["qvp"][3:7, 10:12] = -1e-5  # a few negatives for NegativeValueProcess
    return c, st
