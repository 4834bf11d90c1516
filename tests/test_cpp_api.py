"""The C++ state/config API (include/vvm.hpp) and the model driver
vvm_main: the reference's Config_VVM / vvm / vvm::Iteration::TimeMarching
surface (src/Declare.hpp:27-941, src/Iteration.cpp:208-438,
src/main.cpp:45-135) driving the device timestep through the C ABI."""
import os
import subprocess

import numpy as np
import pytest

import vvm_amd as V
from helpers import assert_fields_close

PKG = os.path.dirname(V.__file__)
MAIN = os.path.join(PKG, "vvm_main")
LIB = os.path.join(PKG, "libvvm.so")

# c0: 128x64 interior at 200 m, dt 2 s; the driver's keys follow the
# reference's vvm_config.txt
C0_KEYS = dict(VVM_XRANGE=128 * 200, VVM_ZRANGE=64 * 200, VVM_DT=2, VVM_DX=200, VVM_DZ=200, VVM_CASE=1,
               VVM_MOISTURE_NUDGE_TIME=0, VVM_ADDFORCINGTIME=0, VVM_TROPICAL_FORCING=1, VVM_BRUNT=0.01,
               VVM_BUBBLE_DT=2, VVM_BUBBLE_RADIUS=2000, VVM_BUBBLE_ZC=2000, VVM_NOISE="uniform",
               VVM_NOISE_AMP=0.01, VVM_NOISE_ZTOP=1000, VVM_SEED=1, VVM_PRINT_STEPS=0)


def write_config(path, **keys):
    with open(path, "w") as f:
        for k, v in keys.items():
            f.write(f"{k}={v}    # set by the test\n")


def built():
    if not (os.path.exists(MAIN) and os.path.exists(LIB)):
        pytest.skip("vvm_main / libvvm.so not built")


def test_driver_parses_reference_config(tmp_path):
    built()
    cfg = tmp_path / "vvm_config.txt"
    write_config(cfg, VVM_TIMEEND=20, VVM_OUTPUTSTEP=5, VVM_OUTPUTPATH=str(tmp_path / "out"), **C0_KEYS)
    r = subprocess.run([MAIN, str(cfg), "--dry-run"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    # XRANGE/ZRANGE gain one ghost cell each side (src/Declare.hpp:33,82-83)
    assert "nx=130 nz=66 dx=200 dz=200 dt=2 steps=10 solver=fp64" in r.stdout


def test_driver_rejects_bad_values(tmp_path):
    built()
    cfg = tmp_path / "vvm_config.txt"
    write_config(cfg, VVM_TIMEEND=20, VVM_SOLVER="fp16", **C0_KEYS)
    r = subprocess.run([MAIN, str(cfg), "--dry-run"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "VVM_SOLVER" in r.stderr
    write_config(cfg, VVM_DT=2)
    r = subprocess.run([MAIN, str(cfg), "--dry-run"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "missing key" in r.stderr


def test_library_exports_reference_names():
    built()
    r = subprocess.run(["nm", "-D", "-C", "--defined-only", LIB], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    for name in ("vvm::Iteration::TimeMarching(vvm&)", "vvm::Iteration::nextTimeStep(vvm&)",
                 "vvm::PoissonSolver::cal_w(vvm&, int, int, int)", "vvm::PoissonSolver::cal_u(vvm&)",
                 "vvm::PoissonSolver::pubarTop_pt(vvm&)", "vvm::PoissonSolver::InitPoissonMatrix(vvm&)",
                 "vvm::Turbulence::RKM_RKH(vvm&)", "vvm::BoundaryProcess2D_all(vvm&)",
                 "vvm::BoundaryProcess2D_center(double**, int, int)",
                 "vvm::BoundaryProcess2D_westdown(double**, int, int)", "vvm::vvm(Config_VVM const&)",
                 "Config_VVM::Config_VVM("):
        assert name in r.stdout, name


@pytest.mark.gpu
def test_time_marching_matches_oracle(tmp_path, oracle):
    """vvm_main runs 10 steps of c0 through TimeMarching, writing the state
    at the output steps 0 and 5; the oracle runs the same 10 steps from
    the step-0 file."""
    built()
    cfg = tmp_path / "vvm_config.txt"
    out = tmp_path / "out"
    final = tmp_path / "final.bin"
    write_config(cfg, VVM_TIMEEND=20, VVM_OUTPUTSTEP=5, VVM_OUTPUTPATH=str(out), VVM_FINAL_STATE=str(final),
                 **C0_KEYS)
    print(f"vvm_main c0 10 steps -> {final}", flush=True)
    r = subprocess.run([MAIN, str(cfg)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    s0 = V.HostState.load(str(out / "state" / "state_0.bin"))
    s5 = V.HostState.load(str(out / "state" / "state_5.bin"))
    got = V.HostState.load(str(final))
    assert s5.step == 5 and got.step == 10
    # the driver's initial state is the library's init layer, unchanged
    c = V.case("c0")
    ref0 = V.init_state(c)
    for n in V.FIELDS2D + V.PROFILES:
        np.testing.assert_array_equal(s0[n], ref0[n], err_msg=n)
    want = s0.copy()
    oracle.step(c.cfg, want, 10)
    assert_fields_close(got, want, ["zeta", "th", "qv", "qc", "qr", "u", "w"], 1e-9, where="TimeMarching c0")
