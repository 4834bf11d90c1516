"""The reference's own TimeMarching (src/Iteration.cpp compiled unchanged,
oracle/_ref/vvm_dropin, see oracle/ref/build_ref.sh) driving this build's
per-operator device statics through the drop-in C++ API: every operator
call of the reference's step lands on a device kernel, with the host arrays
synchronised around it.  The result must match the oracle's step to the
1e-10 bar of the timestep.  The binary is built where the reference sources
exist and travels with the built libraries; without it the test skips."""
import os
import subprocess

import pytest

import vvm_amd as V
from helpers import assert_fields_close, branch_cases, perturbed_state

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "vvm_dropin")
FIELDS = ["zeta", "th", "qv", "qc", "qr", "u", "w", "RKM", "RKH", "precip", "thb", "qvb", "init_th_forcing"]


def cases():
    t = V.case("tiny")
    out = {"tiny_perturbed": (t.cfg, perturbed_state(t, seed=9, step=0), 6)}
    b = branch_cases()
    for name in ("kick_fine", "nudge_qv", "case2"):
        cfg, st, _ = b[name]
        out[name] = (cfg, st, 4)
    return out


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="built only where the reference sources exist")
@pytest.mark.parametrize("name", ["tiny_perturbed", "kick_fine", "nudge_qv", "case2"])
def test_reference_timemarching_on_dropin_statics(oracle, tmp_path, name):
    cfg, st, n = cases()[name]
    print(f"reference TimeMarching on the drop-in statics: {name}, {cfg.nx - 2}x{cfg.nz - 2}, {n} steps",
          flush=True)
    pin, pout = str(tmp_path / "in.state"), str(tmp_path / "out.state")
    st.save(pin, cfg)
    subprocess.run([DROPIN, str(n), pin, pout], check=True, timeout=120, stdout=subprocess.DEVNULL)
    r = V.HostState.read_records(pout)
    got = st.copy()
    for f in got.a:
        if f in r:
            got.a[f][...] = r[f].reshape(got.a[f].shape)
    want = st.copy()
    want.step = 0
    oracle.step(cfg, want, n)
    assert int(r["step"][0]) == n
    assert_fields_close(got, want, FIELDS, 1e-10, where=f"{name} drop-in")
