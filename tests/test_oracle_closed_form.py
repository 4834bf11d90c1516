"""Pins of the oracle that do not rely on the reference binary: closed
forms, invariants and the reference's own known-answer patterns."""
import ctypes as C

import numpy as np
import pytest

import vvm_amd as V
from helpers import perturbed_state, rel_err


def test_poisson_manufactured_solution(oracle):
    """The reference's commented Eigen check (src/PoissonSolver.cpp:815-848):
    x_ans = exp(sin(i+k)), b = A x_ans, solve, compare."""
    for name in ("tiny", "c0"):
        c = V.case(name)
        st = V.init_state(c)
        N, M = c.cfg.nx - 2, c.cfg.nz - 3
        kk, ii = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
        x_ans = np.exp(np.sin(ii + kk)).ravel()
        b = oracle.poisson_apply(c.cfg, st["rhou"], st["rhow"], x_ans)
        x = oracle.poisson_exact(c.cfg, st["rhou"], st["rhow"], b)
        assert rel_err(x, x_ans) <= 1e-12


def test_poisson_apply_matches_assembled_matrix(oracle):
    """poisson_apply against a dense assembly of InitPoissonMatrix's triplets
    (src/PoissonSolver.cpp:764-796) on a small grid."""
    c = V.case("tiny")
    st = V.init_state(c)
    N, M, nz = c.cfg.nx - 2, c.cfg.nz - 3, c.cfg.nz
    rhou, rhow = st["rhou"], st["rhow"]
    n = N * M
    A = np.zeros((n, n))
    k = 1
    for idx in range(n):
        if idx % N == 0:
            k += 1
        A[idx, idx] += -(2. + rhow[k] / rhou[k] + rhow[k] / rhou[k - 1])
        if idx % N != 0:
            A[idx, idx - 1] += 1.
        if (idx + 1) % N != 0:
            A[idx, idx + 1] += 1.
        if idx % N == 0:
            A[idx, idx + N - 1] += 1.
            A[idx + N - 1, idx] += 1.
    k = 1
    for idx in range(n):
        if idx % N == 0:
            k += 1
        if idx < N * (M - 1):
            A[idx, idx + N] += rhow[k] / rhou[k]
        if idx >= N:
            A[idx, idx - N] += rhow[k] / rhou[k - 1]
    x = np.random.default_rng(0).normal(size=n)
    assert rel_err(oracle.poisson_apply(c.cfg, rhou, rhow, x), A @ x) <= 1e-14
    b = np.random.default_rng(1).normal(size=n)
    assert rel_err(oracle.poisson_exact(c.cfg, rhou, rhow, b), np.linalg.solve(A, b)) <= 1e-12


def test_halos_idempotent_and_westdown_rows(oracle):
    c = V.case("tiny")
    st = perturbed_state(c, consistent_ghosts=False)
    oracle.op("bc_all", c.cfg, st)
    a = st.copy()
    oracle.op("bc_all", c.cfg, a)
    for f in ("zetap", "thp", "w", "u", "qvp", "qcp", "qrp"):
        assert np.array_equal(a[f], st[f]), f
    for f in ("zetap", "w"):
        assert not st[f][:, [0, 1, -1]].any()
        assert np.array_equal(st[f][0, 2:-1], st[f][-2, 2:-1])
    for f in ("thp", "u"):
        assert np.array_equal(st[f][:, 0], st[f][:, 1]) and np.array_equal(st[f][:, -1], st[f][:, -2])


def test_scalar_advection_conserves_mass(oracle):
    """sum_ik rhou[k] (phi_p - phi) = 0 on the first AB2 step: periodic x,
    zero flux through k = 0 and nz-1, w = 0 at k = 1."""
    c = V.case("tiny")
    st = perturbed_state(c, step=0)
    st["w"][:, 1] = 0.
    st["w"][:, 0] = 0.
    for h in V.HISTS:
        st[h][:] = 0.
    before = st.copy()
    oracle.op("pth_pt", c.cfg, st)
    d = (st["thp"] - before["th"])[1:-1, 1:-1] * st["rhou"][None, 1:-1]
    assert abs(d.sum()) <= 1e-12 * np.abs(d).sum()


def test_negative_value_process_preserves_sum(oracle):
    c = V.case("tiny")
    st = V.HostState(c.cfg.nx, c.cfg.nz)
    rng = np.random.default_rng(3)
    q = rng.normal(1e-3, 1e-3, st["qcp"].shape)
    st["qcp"][:] = q
    lib = oracle.load()
    lib.vo_negval(C.byref(c.cfg), st["qcp"].ctypes.data_as(C.POINTER(C.c_double)))
    inner_before = q[1:-1, 1:-1].sum()
    inner_after = st["qcp"][1:-1, 1:-1]
    assert (inner_after >= 0).all()
    assert inner_after.sum() == pytest.approx(inner_before, rel=1e-13)


def test_kessler_conserves_water_and_moist_static_energy(oracle):
    """autoconversion/accretion/evaporation/condensation move water between
    qv, qc and qr (their sum is unchanged per cell) and heat theta by
    Lv/(Cp pi) per unit condensed."""
    c = V.case("tiny")
    st = perturbed_state(c, negatives=False)
    for f in ("th", "qv", "qc", "qr"):
        st[f + "p"][:] = np.abs(st[f])
    st["qvp"][:] *= 1.3  # supersaturate some cells
    before = st.copy()
    oracle.op("microphysics", c.cfg, st)
    s = slice(1, -1)
    w0 = (before["qvp"] + before["qcp"] + before["qrp"])[s, s]
    w1 = (st["qvp"] + st["qcp"] + st["qrp"])[s, s]
    assert np.allclose(w1, w0, rtol=1e-13, atol=1e-18)
    pib = st["pib"][None, s]
    h0 = (before["thp"][s, s] - c.cfg.Lv / (c.cfg.Cp * pib) * (before["qcp"] + before["qrp"])[s, s])
    h1 = (st["thp"][s, s] - c.cfg.Lv / (c.cfg.Cp * pib) * (st["qcp"] + st["qrp"])[s, s])
    assert np.allclose(h1, h0, rtol=1e-12)
    assert (st["condensation"][s, s] != 0).any() and (st["evaporation"][s, s] >= 0).all()


def test_update_mean_closed_form(oracle):
    c = V.case("tiny")
    st = perturbed_state(c)
    oracle.op("updateMean", c.cfg, st)
    thb = st["thp"][1:-1, 1:-1].mean(axis=0)
    qvb = st["qvp"][1:-1, 1:-1].mean(axis=0)
    assert np.allclose(st["thb"][1:-1], thb, rtol=1e-14)
    assert np.allclose(st["thvb"][1:-1], thb * (1 + 0.608 * qvb), rtol=1e-14)
    assert np.allclose(st["thb_zeta"][2:-1], 0.5 * (st["thb"][1:-2] + st["thb"][2:-1]), rtol=1e-15)


def test_turbulence_branches(oracle):
    """RKM_RKH's closed form on a state covering Ri < 0, 0 < Ri < 0.25 and
    Ri >= 0.25, with the [10, 0.8 dx dz / dt] clamps (Turbulent.cpp:13-42)."""
    c = V.case("tiny")
    st = perturbed_state(c)
    st["thp"][:] = st["th"]
    oracle.op("turbulence", c.cfg, st)
    g, dx, dz, dt = c.cfg.GRAVITY, c.cfg.dx, c.cfg.dz, c.cfg.dt
    u, w, th, l2 = st["u"], st["w"], st["th"], st["lambda2"]
    nx, nz = c.cfg.nx, c.cfg.nz
    seen = set()
    for i in range(1, nx - 1):
        for k in range(1, nz - 1):
            Rz = 0.5 * ((w[i + 1, k + 1] + w[i + 1, k]) - (w[i - 1, k + 1] + w[i - 1, k])) / (2 * dx) + \
                0.5 * ((u[i + 1, k + 1] + u[i, k + 1]) - (u[i + 1, k - 1] + u[i, k - 1])) / (2 * dz)
            Ro = ((u[i + 1, k] - u[i, k]) / dx + (w[i, k + 1] - w[i, k]) / dz) ** 2
            Ri = (g / th[i, k] * (th[i, k + 1] - th[i, k - 1]) / (2 * dz)) / (Rz ** 2 + 2 * Ro)
            S = np.sqrt(Rz ** 2 + 2 * Ro)
            if Ri < 0:
                km, seen = l2[k] * S * np.sqrt(1 - 16 * Ri), seen | {"neg"}
            elif 0 < Ri < 0.25:
                km, seen = l2[k] * S * (1 - 4 * Ri) ** 4, seen | {"mid"}
            else:
                km, seen = 0.0, seen | {"stable"}
            km = min(max(km, 10.0), 0.8 * dx * dz / dt)
            assert st["RKM"][i, k] == pytest.approx(km, rel=1e-12)
    assert seen == {"neg", "mid", "stable"}


def test_init_base_state_closed_forms():
    c = V.case("c0")
    st = V.init_state(c)
    nz = c.cfg.nz
    z = (np.arange(nz) - 0.5) * c.cfg.dz
    k = slice(1, nz - 1)
    assert np.allclose(st["thb"][k], 300 * np.exp(0.01 ** 2 * z[k] / 9.80665), rtol=1e-15)
    rh = np.maximum(0.8 * (1 - z[k] / 12000.), 0.1)
    assert np.allclose(st["qvb"][k] / st["qvsb"][k], rh, rtol=1e-14)
    assert np.allclose(st["rhow"][2:nz - 1], 0.5 * (st["rhou"][2:nz - 1] + st["rhou"][1:nz - 2]), rtol=1e-15)
    assert np.allclose(st["thvb"], st["thb"] + 0.61 * st["qvb"], rtol=1e-15)
    # bubble: +2 K at its centre cell, noise <= 0.01 K below 1 km
    d = st["th"][1:-1, 1:-1] - st["thb"][None, 1:-1]
    assert 1.9 < d.max() <= 2.0 + 0.01
    assert np.abs(d[:, z[1:-1] > 4100]).max() == 0.0


@pytest.mark.parametrize("cooling", [False, True])
def test_radiation_heating_term(oracle, cooling):
    """AddForcing's radiation term and RADIATIONCOOLING (src/AddForcing.cpp:
    20-23): against the same forcing without them, thp at every interior
    point gains dt rate / 86400 (and dt (-2 / 86400)) and nothing else
    changes; a dropped 1/86400, a sign or an [i][k] / [k][i] mix-up fails."""
    c = V.case("tiny", tropical_forcing=1)
    st = perturbed_state(c, seed=41, step=3)
    nx, nz, dt = c.cfg.nx, c.cfg.nz, c.cfg.dt
    rate = np.random.default_rng(42).normal(0.0, 5.0, (nx, nz))
    rate[:, :nz // 2] += 3.0  # not symmetric under a transpose
    base, rad = st.copy(), st.copy()
    oracle.op("forcing_bc", c.cfg, base)
    try:
        oracle.set_radiation(rate, cooling)
        oracle.op("forcing_bc", c.cfg, rad)
    finally:
        oracle.set_radiation(None, False)
    want = dt * rate[1:-1, 1:-1] / 86400. + (dt * (-2. / 86400.) if cooling else 0.)
    got = rad["thp"][1:-1, 1:-1] - base["thp"][1:-1, 1:-1]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    for f in ("qvp", "qcp", "qrp", "zetap"):
        np.testing.assert_array_equal(rad[f], base[f])


def test_red_black_ilu0_formulas():
    """The red-black ILU0 the BiCGSTAB mode applies (k_bicgstab.hip: a black
    pass z_b = (r_b - sum a_bq r_q / d_q) / S_b, S_b = d_b - sum a_bq^2 / d_q,
    then a red pass z_r = (r_r - sum a_rq z_q) / d_r) against a textbook
    ILU(0) (IKJ elimination restricted to the pattern of A) of the same
    matrix in red-black order, on the tiny grid's operator."""
    c = V.case("tiny")
    st = V.init_state(c)
    N, M = c.cfg.nx - 2, c.cfg.nz - 3
    assert N % 2 == 0
    rhou, rhow = st["rhou"], st["rhow"]
    # the device operator L of the unknown grid: rows j = 1..M, periodic i
    cx = np.zeros(M + 2)
    cu = np.zeros(M + 1)
    cx[1:M + 1] = 1.0 / rhow[2:M + 2]
    cu[0:M + 1] = 1.0 / rhou[1:M + 2]
    idx = lambda i, j: (j - 1) * N + i  # noqa: E731
    n = N * M
    A = np.zeros((n, n))
    for j in range(1, M + 1):
        for i in range(N):
            p = idx(i, j)
            A[p, p] = 2 * cx[j] + cu[j] + cu[j - 1]
            A[p, idx((i - 1) % N, j)] -= cx[j]
            A[p, idx((i + 1) % N, j)] -= cx[j]
            if j < M:
                A[p, idx(i, j + 1)] -= cu[j]
            if j > 1:
                A[p, idx(i, j - 1)] -= cu[j - 1]
    red = [idx(i, j) for j in range(1, M + 1) for i in range(N) if (i + j) % 2 == 0]
    black = [idx(i, j) for j in range(1, M + 1) for i in range(N) if (i + j) % 2 == 1]
    order = red + black
    Ap = A[np.ix_(order, order)]
    # textbook ILU(0): IKJ Gaussian elimination, fill outside the pattern dropped
    LU = Ap.copy()
    pat = Ap != 0
    for i in range(1, n):
        for k in range(i):
            if not pat[i, k]:
                continue
            LU[i, k] /= LU[k, k]
            for jj in range(k + 1, n):
                if pat[i, jj]:
                    LU[i, jj] -= LU[i, k] * LU[k, jj]
    Lm = np.tril(LU, -1) + np.eye(n)
    Um = np.triu(LU)
    r = np.random.default_rng(7).normal(size=n)
    want = np.zeros(n)
    want[order] = np.linalg.solve(Um, np.linalg.solve(Lm, r[order]))
    # the kernels' two passes
    d = lambda j: 2 * cx[j] + cu[j] + cu[j - 1]  # noqa: E731
    z = np.zeros(n)
    for j in range(1, M + 1):
        for i in range(N):
            if (i + j) % 2 == 0:
                continue
            S = d(j) - 2 * cx[j] ** 2 / d(j)
            acc = -cx[j] * (r[idx((i - 1) % N, j)] + r[idx((i + 1) % N, j)]) / d(j)
            if j < M:
                S -= cu[j] ** 2 / d(j + 1)
                acc += -cu[j] * r[idx(i, j + 1)] / d(j + 1)
            if j > 1:
                S -= cu[j - 1] ** 2 / d(j - 1)
                acc += -cu[j - 1] * r[idx(i, j - 1)] / d(j - 1)
            z[idx(i, j)] = (r[idx(i, j)] - acc) / S
    for j in range(1, M + 1):
        for i in range(N):
            if (i + j) % 2 == 1:
                continue
            acc = -cx[j] * (z[idx((i - 1) % N, j)] + z[idx((i + 1) % N, j)])
            if j < M:
                acc += -cu[j] * z[idx(i, j + 1)]
            if j > 1:
                acc += -cu[j - 1] * z[idx(i, j - 1)]
            z[idx(i, j)] = (r[idx(i, j)] - acc) / d(j)
    np.testing.assert_allclose(z, want, rtol=1e-12, atol=1e-14)
