"""Failure handling of the elliptic solve on the GPU.

A non-finite residual is never a converged one: a NaN or Inf in the
right-hand side or the initial guess gives the not-converged warning
(VVM_WARN_NOT_CONVERGED), as the reference warns when its solver fails
(src/PoissonSolver.cpp:463-467).  A learned guess that starts worse than the
extrapolation it corrects, or is not finite, is dropped: the solve restarts
from the extrapolation, so the step equals the extrapolated path's bit for
bit and only the restart (and skip) counts tell the difference.
"""
import numpy as np
import pytest

import vvm_amd as V
from helpers import rel_err

pytestmark = pytest.mark.gpu

NET_L, NET_C, NET_T = 6, 32, 9
NET_LP = NET_C * NET_C * NET_T + NET_C


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("bad", ["guess_nan", "guess_inf", "rhs_nan"])
def test_poisson_nonfinite_is_not_converged(mode, bad):
    c = V.case("c0", solver_mode=V.VVM_SOLVER_MIXED if mode == "mixed" else V.VVM_SOLVER_FP64,
               stop_mode=V.VVM_STOP_REL, tol_rel=1e-10)
    st = V.init_state(c)
    n = (c.cfg.nx - 2) * (c.cfg.nz - 3)
    b = np.random.default_rng(2).normal(0, 1, n)
    x = np.zeros(n)
    if bad == "guess_nan":
        x[n // 3] = np.nan
    elif bad == "guess_inf":
        x[7] = np.inf
    else:
        b[n // 2] = np.nan
    print(f"poisson {mode} with {bad}", flush=True)
    with V.Device(c.cfg) as dev:
        dev.upload(st)
        s = dev.poisson(b, x)
        assert dev.last_rc == V.VVM_WARN_NOT_CONVERGED, dev.last_rc
        assert not np.isfinite(s.rel_residual)
        # the handle still solves a good system afterwards
        s2 = dev.poisson(np.random.default_rng(3).normal(0, 1, n), np.zeros(n))
        assert dev.last_rc == 0 and s2.rel_residual <= 1e-10


def _run(net_params, steps=10):
    """steps of a 512x64 case, mixed precision; the learned guess with the
    given parameters (None: the extrapolation the network corrects, linear
    on the third solve and quadratic from the fourth)."""
    guess = V.VVM_GUESS_EXTRAP2 if net_params is None else V.VVM_GUESS_LEARNED
    c = V.grid_case(512, 64, seed=3, guess_mode=guess, solver_mode=V.VVM_SOLVER_MIXED,
                    stop_mode=V.VVM_STOP_REL, tol_rel=1e-10, max_iter=400)
    st = V.init_state(c)
    net = None
    with V.Device(c.cfg) as dev:
        dev.upload(st)
        if net_params is not None:
            net = V.Net()
            net.set_params(net_params)
            dev.set_guess_net(net)
        s = dev.step(steps)
        out = st.copy()
        dev.download(out)
    if net is not None:
        net.close()
    return out, s


@pytest.mark.parametrize("kind", ["nan", "far"])
def test_bad_learned_guess_falls_back_to_extrapolation(kind):
    with V.Net() as net:
        net.init_params(1)
        p = net.params().copy()
    if kind == "nan":
        p[:] = np.nan  # every guess is NaN
    else:
        # the last layer's bias moves every guess 1e4 increments away: finite
        # and far worse than the extrapolation
        p[(NET_L - 1) * NET_LP + NET_C * NET_C * NET_T:] += 1e4
    print(f"learned guess with a {kind} network, 10 steps of 512x64", flush=True)
    want, s_ex = _run(None)
    got, s_le = _run(p)
    # the learned guess needs two previous solutions: solves 3..10 may use
    # it; after each restart the network sits out 1, 4, 16 .. solves
    # (solves 3, 5, 10 run it and are restarted; 4, 6-9 skip it)
    assert s_le.guess_fallbacks == 3, s_le.guess_fallbacks
    assert s_le.guess_skipped == 5, s_le.guess_skipped
    assert s_le.not_converged == 0 and s_le.rel_residual_last <= 1e-10
    assert s_le.solve_iters_total == s_ex.solve_iters_total
    for f in ("zeta", "th", "qv", "qc", "qr", "u", "w"):
        assert rel_err(got[f], want[f]) == 0.0, f
