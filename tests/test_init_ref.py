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


def _write_nc(path, fields, ubar):
    """A per-step output as the reference's output_nc writes it, through an
    independent netCDF-3 writer (scipy)."""
    from scipy.io import netcdf_file
    nx, nz = fields["th"].shape
    with netcdf_file(str(path), "w") as nc:
        nc.createDimension("x", nx)
        nc.createDimension("z", nz)
        for f, a in fields.items():
            var = nc.createVariable(f, "d", ("x", "z"))
            var[:] = a
        var = nc.createVariable("ubarTop", "d", ())
        var.data[...] = ubar


@pytest.mark.parametrize("case_", [0, 1])
def test_restart_from_previous_outputs(tmp_path, case_):
    """LoadFromPreviousFile (src/Init.cpp:372-485) reads two outputs and the
    sounding; the fields, the bubble added back for CASE 1, the halos and the
    recomputed horizontal means follow the reference's statements."""
    nxi, nzi = 16, 100
    cfg = V.make_config(nxi, nzi, 200.0, 200.0, 2.0, CASE=case_, tropical_forcing=1)
    nx, nz = cfg.nx, cfg.nz
    rng = np.random.default_rng(case_)
    prev = {f: rng.normal(0, 1, (nx, nz)) for f in ("th", "zeta", "qv", "qc", "qr", "u", "w")}
    now = {f: rng.normal(0, 1, (nx, nz)) for f in ("th", "zeta", "qv", "qc", "qr", "u", "w")}
    now["th"] += 300.0
    prev["th"] += 300.0
    now["qv"] = np.abs(now["qv"]) * 1e-3
    _write_nc(tmp_path / "a.nc", prev, 0.5)
    _write_nc(tmp_path / "b.nc", now, -3.0)
    st = V.init_from_previous(cfg, SOUNDING, str(tmp_path / "a.nc"), str(tmp_path / "b.nc"))
    for f in ("zeta", "qv", "qc", "qr"):
        np.testing.assert_array_equal(st[f], now[f], err_msg=f)
        np.testing.assert_array_equal(st[f + "m"], prev[f], err_msg=f + "m")
    np.testing.assert_array_equal(st["u"], now["u"])
    np.testing.assert_array_equal(st["w"], now["w"])
    assert st.s["ubarTopm"] == 0.5
    bubble = np.zeros((nx, nz))
    if case_ == 1:
        i = np.arange(nx)[:, None]
        k = np.arange(nz)[None, :]
        rad = np.sqrt(((i - 0.5) * 200.0 - nx * 200.0 / 2) ** 2 / 4000.0 ** 2 +
                      ((k - 0.5) * 200.0 - 2500.0) ** 2 / 2000.0 ** 2)
        bubble = np.where(rad <= 1, 1.5 * (np.cos(np.pi * rad) + 1), 0.0)
    th = now["th"].copy()
    th[1:-1, 1:-1] += bubble[1:-1, 1:-1]
    th[0, 1:-1], th[-1, 1:-1] = th[-2, 1:-1], th[1, 1:-1]
    th[:, 0], th[:, -1] = th[:, 1], th[:, -2]
    np.testing.assert_allclose(st["th"], th, rtol=1e-15, atol=0)
    thb = th[1:-1, 1:-1].mean(axis=0)
    np.testing.assert_allclose(st["thb"][1:-1], thb, rtol=1e-14)
    np.testing.assert_allclose(st["qvb"][1:-1], now["qv"][1:-1, 1:-1].mean(axis=0), rtol=1e-12)
    sounding = np.loadtxt(SOUNDING, skiprows=2)[:nzi]
    np.testing.assert_array_equal(st["rhou"][1:-1], sounding[:, 9])
    np.testing.assert_array_equal(st["Q1LS"][1:-1], sounding[:, 7] * 6.0)


def test_case2_shear_profile():
    """Closed form of Init2d's CASE 2 wind: -8 m/s at the surface, linear to
    0 at 2 km, constant above (src/Init.cpp:213-223)."""
    cfg = V.make_config(16, 40, 200.0, 200.0, 2.0, CASE=2, tropical_forcing=0)
    st = V.init_reference(cfg)
    k = np.arange(1, 41)
    z_u = (k - 1) * 200.0
    want = np.where(z_u <= 2000.0, -8.0 + 8.0 / 2000.0 * z_u, 0.0)
    np.testing.assert_allclose(st["u"][5, 1:41], want, rtol=0, atol=1e-12)
