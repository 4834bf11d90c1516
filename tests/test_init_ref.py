"""The reference's own initial state (vvm_init_reference) against the
reference compiled from its sources (tests/golden/init_ref.npz, written by
tests/golden/make_golden.py from Init1d + Init2d): the analytic sounding,
the sounding file input/init.txt, TROPICALFORCING on and off, CASE 0, 1, 2
and the random kick.  Host code, same libm: bit-exact."""
import os

import numpy as np
import pytest

import vvm_amd as V

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "init_ref.npz"))
SOUNDING = os.path.join(HERE, "golden", "init_sounding.txt")

VARIANTS = {
    "analytic_trop": (32, 40, 200.0, 1, 0.0, 1, False),
    "load_trop": (32, 100, 200.0, 1, 0.0, 1, True),
    "notrop_case0": (32, 40, 200.0, 0, 0.0, 0, False),
    "notrop_case1": (48, 40, 200.0, 1, 0.0, 0, False),
    "notrop_case2": (48, 40, 200.0, 2, 0.0, 0, False),
    "notrop_case1_kick": (32, 40, 200.0, 1, 10.0, 0, False),
}
FIELDS = ["th", "thm", "qv", "qvm", "u", "zeta", "zetam", "init_th_forcing"]


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_init_matches_reference(name):
    nxi, nzi, dx, case_, aft, trop, load = VARIANTS[name]
    cfg = V.make_config(nxi, nzi, dx, dx, 2.0, CASE=case_, addforcingtime=aft, tropical_forcing=trop)
    st = V.init_reference(cfg, SOUNDING if load else None)
    for f in FIELDS + V.PROFILES:
        key = f"{name}/{f}"
        if key not in GOLD.files:
            continue
        np.testing.assert_array_equal(st[f], GOLD[key].reshape(st[f].shape), err_msg=key)
    assert st.s["ubarTop"] == float(GOLD[f"{name}/ubarTop"][0])
    assert st.s["ubarTopm"] == float(GOLD[f"{name}/ubarTopm"][0])


def test_init_short_sounding_file_is_rejected(tmp_path):
    """Fewer levels in the file than the grid has: an error, not a read past
    the data."""
    lines = open(SOUNDING).read().splitlines()
    short = tmp_path / "short.txt"
    short.write_text("\n".join(lines[:20]) + "\n")
    cfg = V.make_config(32, 100, 200.0, 200.0, 2.0, tropical_forcing=1)
    with pytest.raises(V.VvmError):
        V.init_reference(cfg, str(short))


def test_case2_shear_profile():
    """Closed form of Init2d's CASE 2 wind: -8 m/s at the surface, linear to
    0 at 2 km, constant above (src/Init.cpp:213-223)."""
    cfg = V.make_config(16, 40, 200.0, 200.0, 2.0, CASE=2, tropical_forcing=0)
    st = V.init_reference(cfg)
    k = np.arange(1, 41)
    z_u = (k - 1) * 200.0
    want = np.where(z_u <= 2000.0, -8.0 + 8.0 / 2000.0 * z_u, 0.0)
    np.testing.assert_allclose(st["u"][5, 1:41], want, rtol=0, atol=1e-12)
