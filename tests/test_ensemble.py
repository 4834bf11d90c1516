"""The batched ensemble (include/vvm_ens.h): members of uneven grids
advanced together on one device.  Each member must give exactly what the
same configuration gives on its own handle, and the mixed-precision members
must stay within the benchmark bound of their FP64 twins."""
import numpy as np
import pytest

import vvm_amd as V

SHAPES = [(256, 64), (512, 64), (1024, 128), (256, 128), (512, 128), (2048, 64)]
FIELDS = ("zeta", "th", "qv", "u", "w")


def members(**over):
    rng = np.random.default_rng(11)
    out = []
    for k, (nx, nz) in enumerate(SHAPES):
        out.append(V.grid_case(nx, nz, seed=100 + k, sigma=float(rng.uniform(0.05, 0.2)),
                               heat=float(rng.uniform(50, 200)), **over))
    return out


def download(dev, case):
    st = V.init_state(case)
    dev.download(st)
    return st


@pytest.mark.gpu
def test_ensemble_members_match_single_handles():
    cases = members(stop_mode=V.VVM_STOP_REL, tol_rel=1e-10)
    steps = 12
    print(f"ensemble of {len(cases)} members, {steps} steps", flush=True)
    with V.Ensemble(cases, threads=3) as ens:
        s = ens.step(steps)
        assert s.members == len(cases) and s.not_converged == 0
        assert s.grid_point_steps == sum(nx * nz for nx, nz in SHAPES) * steps
        got = [download(ens.member(i), c) for i, c in enumerate(cases)]
    for i in (0, 2, 5):
        c = cases[i]
        with V.Device(c.cfg) as dev:
            dev.upload(V.init_state(c))
            dev.step(steps)
            want = download(dev, c)
        for f in FIELDS:
            np.testing.assert_array_equal(got[i][f], want[f], err_msg=f"member {i} {f}")


@pytest.mark.gpu
def test_ensemble_mixed_against_fp64():
    """Per-member error of the mixed-precision, tuned ensemble against the
    FP64 ensemble: relative L2 <= 1e-6 in every field (the benchmark bound)."""
    steps = 30
    fast = members(stop_mode=V.VVM_STOP_REL, tol_rel=1e-10, solver_mode=V.VVM_SOLVER_MIXED,
                   guess_mode=V.VVM_GUESS_EXTRAP)
    ref = members(stop_mode=V.VVM_STOP_PARITY)
    with V.Ensemble(fast) as ens:
        ens.tune(2)
        s = ens.step(steps)
        a = [download(ens.member(i), c) for i, c in enumerate(fast)]
    with V.Ensemble(ref) as ens:
        ens.step(steps)
        b = [download(ens.member(i), c) for i, c in enumerate(ref)]
    worst = 0.0
    for i in range(len(fast)):
        for f in FIELDS:
            e = np.linalg.norm(a[i][f] - b[i][f]) / max(np.linalg.norm(b[i][f]), 1e-300)
            worst = max(worst, e)
            assert e <= 1e-6, (i, f, e)
    print(f"mixed ensemble: {s.grid_point_steps / s.seconds:.4g} grid-point-steps/s, worst member field error "
          f"{worst:.2e}", flush=True)
