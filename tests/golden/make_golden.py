"""Write the golden vectors in tests/golden/ by running the reference model,
compiled from its own sources (oracle/ref/build_ref.sh -> oracle/_ref/vvm_ref),
on seeded states.  Runs only where the reference sources exist; the vectors
are committed and the tests read them everywhere.

  python tests/golden/make_golden.py            # all fixtures
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "2d-vector-vorticity-model_amd"), os.path.join(ROOT, "tests")]

import vvm_amd as V  # noqa: E402
from helpers import branch_cases, input_checksums, perturbed_state, single_case  # noqa: E402
from oracle import oracle as O  # noqa: E402

REF_OPS = ["pzeta_pt", "pth_pt", "pqv_pt", "pqc_pt", "pqr_pt", "forcing_bc", "bc_all", "pubarTop_pt", "cal_w",
           "cal_u", "microphysics", "turbulence", "nudge", "updateMean", "nextTimeStep"]
SCAL = ["ubarTop", "ubarTopp", "ubarTopm", "dubarTop_advect0", "dubarTop_advect1", "step"]
STEP_FIELDS = ["zeta", "th", "qv", "qc", "qr", "u", "w", "RKM", "RKH", "uxi", "precip", "thb", "qvb", "thvb"]


def run_ref(mode, arg, st, cfg, tmp):
    pin, pout = os.path.join(tmp, "in.state"), os.path.join(tmp, "out.state")
    st.save(pin, cfg)
    O.run_ref(mode, arg, pin, pout, timeout=3600)
    return V.HostState.read_records(pout)


def state_arrays(st: V.HostState, prefix: str) -> dict:
    out = {prefix + n: a.copy() for n, a in st.a.items()}
    for n in ("ubarTop", "ubarTopp", "ubarTopm"):
        out[prefix + n] = np.array([st.s[n]])
    out[prefix + "dubarTop_advect0"] = np.array([st.dubarTop_advect[0]])
    out[prefix + "dubarTop_advect1"] = np.array([st.dubarTop_advect[1]])
    out[prefix + "step"] = np.array([float(st.step)])
    return out


def ops_fixture(tmp):
    c = V.case("tiny")
    data = {}
    for step in (3, 4):
        st = perturbed_state(c, seed=100 + step, step=step)
        rng = np.random.default_rng(step)
        for f in ("zeta", "th", "qv", "qc", "qr"):
            st[f + "p"][:] = st[f] * (1 + rng.normal(0, 1e-3, st[f].shape))
        st["qcp"][:] -= 2e-5
        data.update(state_arrays(st, f"s{step}/in/"))
        for op in REF_OPS:
            r = run_ref("op", op, st, c.cfg, tmp)
            for n, a in r.items():
                if n in st.a or n in SCAL:
                    ref_in = st.a.get(n)
                    if ref_in is not None and ref_in.size == a.size and np.array_equal(ref_in.reshape(a.shape), a):
                        continue
                    data[f"s{step}/{op}/{n}"] = a
    np.savez_compressed(os.path.join(HERE, "ops_tiny.npz"), **data)


# The statics of class vvm one at a time: reference harness op -> (VVM_OP_*
# id, field argument) as vvm_dev_op_field / vo_op_single take them.
SINGLE_OPS = {
    "advection_zeta": (20, 0), "bouyancy": (21, 0),
    "advection_thermo_th": (22, 1), "advection_thermo_qv": (22, 2), "advection_thermo_qc": (22, 3),
    "advection_thermo_qr": (22, 4), "advection_qrvt": (23, 0),
    "pth_pt": (24, 0), "pqv_pt": (25, 0), "pqc_pt": (26, 0), "pqr_pt": (27, 0), "updateMean": (28, 0),
    "autoconversion": (29, 0), "accretion": (30, 0), "evaporation": (31, 0), "condensation": (32, 0),
    "negval_qvp": (33, 2 * 8 + 2), "negval_qcp": (33, 2 * 8 + 3), "negval_qc": (33, 1 * 8 + 3),
    "mparam": (34, 0), "hparam_th": (35, 1), "hparam_qv": (35, 2), "hparam_qc": (35, 3), "hparam_qr": (35, 4),
    "nudge_theta": (36, 0), "nudge_zeta": (37, 0), "nudge_qv": (38, 0), "add_forcing": (39, 0),
}


def single_fixture(tmp):
    c, st = single_case()
    data = {"in_checksum": input_checksums(st)}
    for op in SINGLE_OPS:
        r = run_ref("op", op, st, c.cfg, tmp)
        for n, a in r.items():
            if n in st.a or n in SCAL:
                ref_in = st.a.get(n)
                if ref_in is not None and ref_in.size == a.size and np.array_equal(ref_in.reshape(a.shape), a):
                    continue
                data[f"{op}/{n}"] = a
    np.savez_compressed(os.path.join(HERE, "ops_single.npz"), **data)


def step_fixture(tmp):
    data = {}
    c = V.case("c0")
    st = V.init_state(c)
    for n in (1, 10):
        r = run_ref("step", n, st, c.cfg, tmp)
        for f in STEP_FIELDS:
            data[f"c0/{n}/{f}"] = r[f]
        data[f"c0/{n}/t_all"] = r["t_all"]
    t = V.case("tiny")
    st = perturbed_state(t, seed=21, step=0)
    data.update(state_arrays(st, "tiny/in/"))
    r = run_ref("step", 20, st, t.cfg, tmp)
    for f in STEP_FIELDS:
        data[f"tiny/20/{f}"] = r[f]
    np.savez_compressed(os.path.join(HERE, "steps.npz"), **data)


def misc_fixture(tmp):
    data = {}
    # RandomPerturbation: a grid fine enough to have levels below 200 m.
    c = V.grid_case(24, 16, dx=50.0, dt=1.0)
    st = V.init_state(c)
    for t in (1, 7):
        st.step = t
        r = run_ref("op", "random_perturbation", st, c.cfg, tmp)
        data[f"rp/{t}"] = r["init_th_forcing"]
    # cal_w on a seeded vorticity field (config-0 grid), the Eigen solve
    # at the reference tolerance and at 1e-10.
    c = V.case("c0")
    st = V.init_state(c)
    rng = np.random.default_rng(5)
    st["zetap"][:] = rng.normal(0, 1e-2, st["zetap"].shape)
    st["zetap"][:, :2] = 0.
    st["zetap"][:, -1] = 0.
    data["calw/zetap"] = st["zetap"].copy()
    r = run_ref("op", "cal_w", st, c.cfg, tmp)
    data["calw/w"] = r["w"]
    r = run_ref("solve", "1e-10", st, c.cfg, tmp)
    data["calw/w_1e-10"] = r["w"]
    data["calw/iters_1e-10"] = r["solve_iterations"]
    data["calw/relres_1e-10"] = r["solve_rel_residual"]
    np.savez_compressed(os.path.join(HERE, "misc.npz"), **data)


def long_fixture(tmp):
    """600 steps of config 0 and a twin started 1 ulp away in theta: the
    spread bounds the GPU-vs-reference difference of a chaotic run."""
    c = V.case("c0")
    st = V.init_state(c)
    r = run_ref("step", 600, st, c.cfg, tmp)
    twin = st.copy()
    i, k = 40, 5
    twin["th"][i, k] = np.nextafter(twin["th"][i, k], np.inf)
    twin["thm"][i, k] = twin["th"][i, k]
    r2 = run_ref("step", 600, twin, c.cfg, tmp)
    data = {}
    for f in ("zeta", "th", "qv", "qc", "qr", "u", "w"):
        data[f"c0/600/{f}"] = r[f]
        data[f"c0/600twin/{f}"] = r2[f]
    data["c0/600/t_all"] = r["t_all"]
    np.savez_compressed(os.path.join(HERE, "long_c0.npz"), **data)


INIT_VARIANTS = {
    # name: (binary, nx_int, nz_int, dx, CASE, addforcingtime, tropical)
    "analytic_trop": ("vvm_ref", 32, 40, 200.0, 1, 0.0, 1),
    "load_trop": ("vvm_ref_load", 32, 100, 200.0, 1, 0.0, 1),
    "notrop_case0": ("vvm_ref_notrop", 32, 40, 200.0, 0, 0.0, 0),
    "notrop_case1": ("vvm_ref_notrop", 48, 40, 200.0, 1, 0.0, 0),
    "notrop_case2": ("vvm_ref_notrop", 48, 40, 200.0, 2, 0.0, 0),
    "notrop_case1_kick": ("vvm_ref_notrop", 32, 40, 200.0, 1, 10.0, 0),
}
INIT_FIELDS = ["th", "thm", "qv", "qvm", "u", "zeta", "zetam", "init_th_forcing"]


def init_fixture(tmp):
    """The reference's own Init1d + Init2d (analytic sounding, sounding file
    input/init.txt, TROPICALFORCING on and off, CASE 0/1/2, the random kick)."""
    data = {}
    for name, (binary, nxi, nzi, dx, case_, aft, trop) in INIT_VARIANTS.items():
        cfg = V.make_config(nxi, nzi, dx, dx, 2.0, CASE=case_, addforcingtime=aft, tropical_forcing=trop)
        st = V.HostState(cfg.nx, cfg.nz)
        pin, pout = os.path.join(tmp, "init_in.state"), os.path.join(tmp, "init_out.state")
        st.save(pin, cfg)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", binary), "init", "-", pin, pout], check=True, env=env,
                       timeout=120)
        r = V.HostState.read_records(pout)
        for f in INIT_FIELDS:
            data[f"{name}/{f}"] = r[f]
        for f in V.PROFILES:
            if f in r:
                data[f"{name}/{f}"] = r[f]
        for f in ("ubarTop", "ubarTopm"):
            data[f"{name}/{f}"] = r[f]
    np.savez_compressed(os.path.join(HERE, "init_ref.npz"), **data)


BRANCH_FIELDS = ["zeta", "th", "qv", "qc", "qr", "u", "w", "RKM", "RKH", "init_th_forcing"]


def reference_init(binary: str, cfg, tmp) -> dict:
    """Init1d + Init2d as the compiled reference runs them on cfg's grid."""
    st = V.HostState(cfg.nx, cfg.nz)
    pin, pout = os.path.join(tmp, "init_in.state"), os.path.join(tmp, "init_out.state")
    st.save(pin, cfg)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([os.path.join(ROOT, "oracle", "_ref", binary), "init", "-", pin, pout], check=True, env=env,
                   timeout=120)
    return V.HostState.read_records(pout)


def branch_fixture(tmp):
    data = {}
    for name, (cfg, st, n) in branch_cases().items():
        if name == "refdefault":
            # the state stepped below is the reference's own initial state
            r = reference_init("vvm_ref_load", cfg, tmp)
            for f, a in r.items():
                if f in st.a and not np.array_equal(st.a[f], a.reshape(st.a[f].shape)):
                    sys.exit(f"refdefault: the host init differs from the reference's in {f}")
        data[f"{name}/in_checksum"] = input_checksums(st)
        r = run_ref("step", n, st, cfg, tmp)
        for f in BRANCH_FIELDS:
            data[f"{name}/{n}/{f}"] = r[f]
        data[f"{name}/steps"] = np.array([n])
    np.savez_compressed(os.path.join(HERE, "branches.npz"), **data)


def main():
    if not O.ref_available():
        subprocess.run(["bash", os.path.join(ROOT, "oracle", "ref", "build_ref.sh")], check=True)
    if not O.ref_available():
        sys.exit("reference binary unavailable")
    which = set(sys.argv[1:]) or {"ops", "steps", "misc", "long", "init", "branches", "single"}
    with tempfile.TemporaryDirectory(dir=os.path.join(ROOT, "oracle", "_ref")) as tmp:
        if "ops" in which:
            ops_fixture(tmp)
        if "steps" in which:
            step_fixture(tmp)
        if "misc" in which:
            misc_fixture(tmp)
        if "long" in which:
            long_fixture(tmp)
        if "init" in which:
            init_fixture(tmp)
        if "branches" in which:
            branch_fixture(tmp)
        if "single" in which:
            single_fixture(tmp)


if __name__ == "__main__":
    main()
