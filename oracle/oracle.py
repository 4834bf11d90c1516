"""ctypes binding of the CPU oracle (oracle/vvm_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "vvm_ref")

GSOLVE_CG, GSOLVE_EXACT = 0, 1
_LIB = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        _LIB = C.CDLL(LIB_PATH)
    return _LIB


def _call(name, cfg, st, *extra):
    lib = load()
    v = st.view()
    rc = getattr(lib, name)(C.byref(cfg), C.byref(v), *extra)
    st.pull_scalars(v)
    if rc != 0:
        raise RuntimeError(f"{name} returned {rc}")


def step(cfg, st, nsteps: int = 1, gsolve: int = GSOLVE_EXACT) -> None:
    """nsteps TimeMarching loop bodies in place on a vvm_amd.HostState."""
    _call("vo_step", cfg, st, C.c_int(nsteps), C.c_int(gsolve))


def random_perturbation(cfg, st, t: int, lo: float = -0.001, hi: float = 0.001, sd: float = 1.0) -> None:
    """Init::RandomPerturbation into st['init_th_forcing'] (the oracle's own
    restatement, oracle/vvm_oracle_rng.cpp)."""
    _call("vo_random_perturbation", cfg, st, C.c_int(t), C.c_double(lo), C.c_double(hi), C.c_double(sd))


OPS = {
    "pzeta_pt": "vo_pzeta_pt", "pth_pt": "vo_pth_pt", "pqv_pt": "vo_pqv_pt", "pqc_pt": "vo_pqc_pt",
    "pqr_pt": "vo_pqr_pt", "forcing_bc": "vo_forcing_bc", "bc_all": "vo_bc_all",
    "pubarTop_pt": "vo_pubartop_pt", "cal_w": "vo_cal_w", "microphysics": "vo_microphysics",
    "turbulence": "vo_turbulence", "nudge": "vo_nudge", "updateMean": "vo_update_mean",
    "nextTimeStep": "vo_next_time_step",
}


def op(name: str, cfg, st, gsolve: int = GSOLVE_EXACT) -> None:
    if name == "cal_u":
        _call("vo_cal_u", cfg, st, C.c_int(gsolve))
    else:
        _call(OPS[name], cfg, st)


def op_single(op: int, cfg, st, field: int = 0) -> None:
    """One static of class vvm alone, by its VVM_OP_* id."""
    _call("vo_op_single", cfg, st, C.c_int(op), C.c_int(field))


def set_radiation(rate, cooling: bool = False) -> None:
    """AddForcing's radiation heating rate (K/day, (nx, nz), kept alive by
    the caller) and RADIATIONCOOLING for later op/step calls; None clears."""
    lib = load()
    P = C.POINTER(C.c_double)
    lib.vo_set_radiation.argtypes = [P, C.c_int]
    lib.vo_set_radiation.restype = None
    lib.vo_set_radiation(None if rate is None else rate.ctypes.data_as(P), int(cooling))


def poisson_exact(cfg, rhou, rhow, b):
    import numpy as np
    lib = load()
    x = np.zeros_like(b)
    P = C.POINTER(C.c_double)
    rc = lib.vo_poisson_exact(C.byref(cfg), rhou.ctypes.data_as(P), rhow.ctypes.data_as(P),
                              b.ctypes.data_as(P), x.ctypes.data_as(P))
    if rc != 0:
        raise RuntimeError("vo_poisson_exact failed")
    return x


def poisson_apply(cfg, rhou, rhow, x):
    import numpy as np
    lib = load()
    y = np.zeros_like(x)
    P = C.POINTER(C.c_double)
    lib.vo_poisson_apply(C.byref(cfg), rhou.ctypes.data_as(P), rhow.ctypes.data_as(P),
                         x.ctypes.data_as(P), y.ctypes.data_as(P))
    return y


def ref_available() -> bool:
    return os.path.exists(REF_BIN)


def run_ref(mode: str, arg: str, in_path: str, out_path: str, timeout: float = 600) -> None:
    """Run the compiled reference harness (oracle/_ref/vvm_ref), single thread.
    Its own prints go to stderr, so a caller's stdout carries only what the
    caller writes."""
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([REF_BIN, mode, str(arg), in_path, out_path], check=True, env=env, timeout=timeout,
                   stdout=sys.stderr)


# ------------------------------------------------------- guess network

def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def net_forward(params, inp, periodic: bool):
    """CPU forward of the guess network: inp [nb][3][h][w] -> [nb][h][w]."""
    import numpy as np
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(inp, dtype=np.float32)
    nb, _, h, w = x.shape
    out = np.zeros((nb, h, w), dtype=np.float32)
    rc = lib.vo_net_forward(nb, h, w, int(periodic), _fp(p), _fp(x), _fp(out))
    if rc != 0:
        raise RuntimeError(f"vo_net_forward returned {rc}")
    return out


def net_grad(params, inp, target, mask, periodic: bool, rounding: bool = True):
    """CPU loss and parameter gradient of the guess network."""
    import numpy as np
    lib = load()
    lib.vo_net_set_rounding(int(rounding))
    try:
        p = np.ascontiguousarray(params, dtype=np.float32)
        x = np.ascontiguousarray(inp, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.float32)
        nb, _, h, w = x.shape
        g = np.zeros_like(p)
        loss = C.c_double(0.)
        rc = lib.vo_net_grad(nb, h, w, int(periodic), _fp(p), _fp(x), _fp(t), _fp(m), _fp(g), C.byref(loss))
        if rc != 0:
            raise RuntimeError(f"vo_net_grad returned {rc}")
        return loss.value, g
    finally:
        lib.vo_net_set_rounding(1)


def net_param_count() -> int:
    return int(load().vo_net_param_count())
