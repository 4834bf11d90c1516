"""The drop-in boundary, compiled: code written for the reference builds
against this include/ directory and links with libvvm.so.

Runs only where the reference sources exist (this build container); the
reference never travels.  Each test stages the one reference source it
compiles into a temporary directory (a quoted #include resolves next to the
including file first, so the file must not sit beside the reference's own
Declare.hpp), compiles it unchanged, and deletes it.
  - src/main.cpp, the reference's driver: compiles and links.
  - src/Iteration.cpp, the reference's TimeMarching and operator wrappers:
    compiles (oracle/ref/build_ref.sh links it into oracle/_ref/vvm_dropin,
    which tests/test_gpu_dropin.py runs on the GPU).
  - INTEGRATION.md section 1's Iteration_AMD.cpp, the stub a maintainer adds
    to the reference tree: compiles against the reference's own
    Declare.hpp with build_ref.sh's configuration edits.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "2d-vector-vorticity-model_amd")
INC = os.path.join(ROOT, "include")
REF = os.environ.get("VVM_REFERENCE", "/root/reference")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")),
                                reason="the reference sources exist only in the build container")


def _gxx(args, cwd):
    r = subprocess.run(["g++", "-std=c++17"] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]


def test_reference_main_links_against_dropin(tmp_path):
    if not os.path.exists(os.path.join(PKG, "libvvm.so")):
        pytest.skip("libvvm.so not built")
    shutil.copy(os.path.join(REF, "src", "main.cpp"), tmp_path / "main.cpp")
    _gxx([f"-I{INC}", "main.cpp", f"-L{PKG}", "-lvvm", "-lvvm_amd", f"-Wl,-rpath,{PKG}", "-o", "vvm_ref_main"],
         tmp_path)
    assert (tmp_path / "vvm_ref_main").exists()
    syms = subprocess.run(["nm", "-C", "-u", str(tmp_path / "vvm_ref_main")], capture_output=True, text=True).stdout
    # the reference's driver calls these, and they come from libvvm.so
    for name in ("vvm::Output::printInit", "vvm::Iteration::TimeMarching", "vvm::Init::Init1d",
                 "vvm::PoissonSolver::InitPoissonMatrix", "vvm_read_config"):
        assert name in syms, name


def test_reference_iteration_compiles_against_dropin(tmp_path):
    shutil.copy(os.path.join(REF, "src", "Iteration.cpp"), tmp_path / "Iteration.cpp")
    _gxx(["-c", f"-I{INC}", "-DTROPICALFORCING", "Iteration.cpp", "-o", "Iteration.o"], tmp_path)
    syms = subprocess.run(["nm", "-C", "-u", str(tmp_path / "Iteration.o")], capture_output=True, text=True).stdout
    # the per-operator statics the reference's step calls
    for name in ("vvm::Advection_zeta", "vvm::Bouyancy", "vvm::Advection_thermo", "vvm::Advection_qrVT",
                 "vvm::MicroPhysics::NegativeValueProcess", "vvm::MicroPhysics::autoconversion",
                 "vvm::MicroPhysics::condensation", "vvm::Turbulence::RKM_RKH", "vvm::NumericalProcess::Nudge_qv",
                 "vvm::NumericalProcess::Nudge_zeta", "vvm::PoissonSolver::cal_w", "vvm::AddForcing",
                 "vvm::Init::RandomPerturbation"):
        assert name in syms, name


def test_integration_stub_compiles_in_reference_tree(tmp_path):
    hdr = tmp_path / "ref_headers"
    subprocess.run(["bash", os.path.join(ROOT, "oracle", "ref", "build_ref.sh"), "--stage-headers", str(hdr)],
                   check=True, capture_output=True, timeout=120)
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```cpp\n(// src/Iteration_AMD.cpp.*?)```", text, re.S)
    assert m, "INTEGRATION.md has no Iteration_AMD.cpp block"
    (tmp_path / "Iteration_AMD.cpp").write_text(m.group(1))
    _gxx(["-c", "-DAMD_STEP", f"-I{hdr}", f"-I{os.path.join(REF, 'include')}", f"-I{INC}", "Iteration_AMD.cpp",
          "-o", "Iteration_AMD.o"], tmp_path)
