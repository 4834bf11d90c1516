"""Python host side of the MI355X VVM timestep.

Binds the C ABI of ``libvvm_amd.so`` (declared in ``include/vvm_dev.h``) with
ctypes, owns host copies of the model state in the reference's ``[i][k]``
layout (numpy arrays), reads and writes the binary state files the reference
harness uses, and defines the benchmark cases.

The compute path is the HIP library: nothing here computes a timestep, and
``Device`` raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
from dataclasses import dataclass, field

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libvvm_amd.so")

VVM_SOLVER_FP64, VVM_SOLVER_MIXED, VVM_SOLVER_BICGSTAB = 0, 1, 2
VVM_GUESS_ZERO, VVM_GUESS_PREV, VVM_GUESS_EXTRAP, VVM_GUESS_LEARNED, VVM_GUESS_EXTRAP2 = 0, 1, 2, 3, 4
VVM_STOP_PARITY, VVM_STOP_REL, VVM_STOP_ABS = 0, 1, 2
VVM_OK, VVM_WARN_NOT_CONVERGED = 0, 1
KT = {"none": 0, "pzeta": 1, "thermo4": 2, "scale": 3, "kessler": 4, "turb": 5, "diffusion": 6, "mean": 7,
      "calu": 8, "solve": 9}

OPS = {
    "pzeta_pt": 1, "pscalars": 2, "forcing_bc": 6, "pubarTop_pt": 7, "cal_w": 8, "cal_u": 9,
    "microphysics": 10, "turbulence": 11, "nudge": 12, "nextTimeStep": 13, "bc_all": 14,
}


class VvmDevConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("GRAVITY", C.c_double), ("Cp", C.c_double), ("Cv", C.c_double), ("Rd", C.c_double),
        ("Lv", C.c_double), ("P0", C.c_double), ("PSURF", C.c_double),
        ("addforcingtime", C.c_double), ("moisture_nudge_time", C.c_double), ("tolerance", C.c_double),
        ("CASE", C.c_int32), ("tropical_forcing", C.c_int32), ("solver_mode", C.c_int32),
        ("guess_mode", C.c_int32), ("stop_mode", C.c_int32), ("max_iter", C.c_int32),
        ("tol_rel", C.c_double),
    ]


FIELDS2D = ["zeta", "zetam", "zetap", "th", "thm", "thp", "qv", "qvm", "qvp", "qc", "qcm", "qcp",
            "qr", "qrm", "qrp", "u", "w", "RKM", "RKH", "init_th_forcing",
            "evaporation", "accretion", "autoconversion", "condensation"]
HISTS = ["dth_advect", "dth_buoyancy", "dzeta_advect", "dqv_advect", "dqc_advect", "dqr_advect", "dqr_VT"]
ROWS = ["precip", "xi", "uxi"]
PROFILES = ["rhou", "rhow", "pib", "pb", "thb", "thbm", "thb_init", "thb_zeta", "thvb", "thvbm",
            "qvb", "qvb0", "qvsb", "z", "z_zeta", "lambda2", "Q1LS", "Q2LS"]
SCALARS = ["ubarTop", "ubarTopp", "ubarTopm"]
PTR = C.POINTER(C.c_double)
FPTR = C.POINTER(C.c_float)


class VvmHostView(C.Structure):
    _fields_ = ([(n, PTR) for n in FIELDS2D + HISTS + ROWS + PROFILES]
                + [(n, C.c_double) for n in SCALARS]
                + [("dubarTop_advect", C.c_double * 2), ("step", C.c_int32),
                   ("status_for_adding_forcing", C.c_int32)])


# int fn(void *user, vvm_host_view *view): a host microphysics scheme
HOST_MP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(VvmHostView))


class VvmStepStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("solve_iters_total", C.c_int32), ("solve_iters_last", C.c_int32),
                ("refine_last", C.c_int32), ("not_converged", C.c_int32), ("rel_residual_last", C.c_double),
                ("t_advection", C.c_double), ("t_poisson", C.c_double), ("t_microphysics", C.c_double),
                ("t_diffusion", C.c_double), ("t_all", C.c_double), ("guess_fallbacks", C.c_int32),
                ("guess_skipped", C.c_int32)]


class VvmSolveStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("refine", C.c_int32), ("rel_residual", C.c_double),
                ("rel_residual0", C.c_double), ("seconds", C.c_double)]


class VvmCaseParams(C.Structure):
    _fields_ = [("sounding", C.c_int32), ("brunt", C.c_double), ("bubble_dT", C.c_double),
                ("bubble_radius", C.c_double), ("bubble_zc", C.c_double), ("noise_kind", C.c_int32),
                ("noise_amp", C.c_double), ("noise_ztop", C.c_double), ("seed", C.c_uint64),
                ("sfc_heat_flux", C.c_double), ("sfc_qv_ramp", C.c_double), ("sfc_qv_ramp_z", C.c_double)]


class VvmTuneResult(C.Structure):
    _fields_ = [("diffusion_strip_waves", C.c_int32), ("mg_coarse_lds_kib", C.c_int32), ("ms_per_step", C.c_double),
                ("ms_default", C.c_double)]


class VvmTrainOpts(C.Structure):
    _fields_ = [("steps", C.c_int32), ("batch", C.c_int32), ("crop_h", C.c_int32), ("crop_w", C.c_int32),
                ("margin", C.c_int32), ("lr", C.c_float), ("lr_final", C.c_float), ("seed", C.c_uint64),
                ("log_every", C.c_int32)]


class VvmTrainStats(C.Structure):
    _fields_ = [("loss_first", C.c_double), ("loss_last", C.c_double), ("seconds", C.c_double),
                ("steps", C.c_int32)]


_LIB = None


def load_lib(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.POINTER
    sig = {
        "vvm_dev_create": (C.c_int, [P(VvmDevConfig), C.c_int, P(C.c_void_p)]),
        "vvm_dev_destroy": (None, [C.c_void_p]),
        "vvm_dev_last_error": (C.c_char_p, [C.c_void_p]),
        "vvm_dev_upload": (C.c_int, [C.c_void_p, P(VvmHostView)]),
        "vvm_dev_download": (C.c_int, [C.c_void_p, P(VvmHostView)]),
        "vvm_dev_step": (C.c_int, [C.c_void_p, C.c_int, P(VvmStepStats)]),
        "vvm_dev_set_step_timers": (C.c_int, [C.c_void_p, PTR, C.c_int]),
        "vvm_dev_poisson": (C.c_int, [C.c_void_p, PTR, PTR, P(VvmSolveStats)]),
        "vvm_dev_poisson_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, P(VvmSolveStats)]),
        "vvm_dev_op": (C.c_int, [C.c_void_p, C.c_int]),
        "vvm_dev_kernel_timer": (C.c_int, [C.c_void_p, C.c_int]),
        "vvm_dev_kernel_time": (C.c_int, [C.c_void_p, P(C.c_double), P(C.c_int)]),
        "vvm_init_case": (C.c_int, [P(VvmDevConfig), P(VvmCaseParams), P(VvmHostView)]),
        "vvm_random_perturbation": (C.c_int, [P(VvmDevConfig), P(VvmHostView), C.c_int,
                                              C.c_double, C.c_double, C.c_double]),
        "vvm_init_reference": (C.c_int, [P(VvmDevConfig), C.c_char_p, P(VvmHostView)]),
        "vvm_init_from_previous": (C.c_int, [P(VvmDevConfig), C.c_char_p, C.c_char_p, C.c_char_p, P(VvmHostView)]),
        "vvm_field_len": (C.c_int64, [C.c_int, C.c_int]),
        "vvm_hist_len": (C.c_int64, [C.c_int, C.c_int]),
        # learned initial guess (include/vvm_net.h)
        "vvm_net_create": (C.c_int, [C.c_int, P(C.c_void_p)]),
        "vvm_net_destroy": (None, [C.c_void_p]),
        "vvm_net_last_error": (C.c_char_p, [C.c_void_p]),
        "vvm_net_param_count": (C.c_int, []),
        "vvm_net_set_params": (C.c_int, [C.c_void_p, FPTR]),
        "vvm_net_get_params": (C.c_int, [C.c_void_p, FPTR]),
        "vvm_net_get_grad": (C.c_int, [C.c_void_p, FPTR]),
        "vvm_net_init_params": (C.c_int, [C.c_void_p, C.c_uint64]),
        "vvm_net_forward_host": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, FPTR, FPTR]),
        "vvm_net_grad_host": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, FPTR, FPTR, FPTR, FPTR,
                                        P(C.c_double)]),
        "vvm_net_adam": (C.c_int, [C.c_void_p, C.c_float]),
        "vvm_net_set_conv_rows": (C.c_int, [C.c_int]),
        "vvm_net_set_layer_pairs": (C.c_int, [C.c_int]),
        "vvm_dev_collect": (C.c_int, [C.c_void_p, C.c_int]),
        "vvm_dev_collected": (C.c_int, [C.c_void_p]),
        "vvm_net_add_samples": (C.c_int, [C.c_void_p, C.c_void_p]),
        "vvm_net_pool_size": (C.c_int, [C.c_void_p]),
        "vvm_net_train": (C.c_int, [C.c_void_p, P(VvmTrainOpts), P(VvmTrainStats)]),
        "vvm_net_eval": (C.c_int, [C.c_void_p, C.c_void_p, P(C.c_double)]),
        "vvm_net_save": (C.c_int, [C.c_void_p, C.c_char_p]),
        "vvm_net_load": (C.c_int, [C.c_void_p, C.c_char_p]),
        "vvm_dev_set_guess_net": (C.c_int, [C.c_void_p, C.c_void_p]),
        "vvm_dev_tune": (C.c_int, [C.c_void_p, C.c_int, P(VvmTuneResult)]),
        # batched ensemble (include/vvm_ens.h)
        "vvm_ens_create": (C.c_int, [P(VvmEnsMember), C.c_int, C.c_int, C.c_int, P(C.c_void_p)]),
        "vvm_ens_destroy": (None, [C.c_void_p]),
        "vvm_ens_last_error": (C.c_char_p, [C.c_void_p]),
        "vvm_ens_size": (C.c_int, [C.c_void_p]),
        "vvm_ens_handle": (C.c_void_p, [C.c_void_p, C.c_int]),
        "vvm_ens_set_guess_net": (C.c_int, [C.c_void_p, C.c_void_p]),
        "vvm_ens_tune": (C.c_int, [C.c_void_p, C.c_int]),
        "vvm_ens_step": (C.c_int, [C.c_void_p, C.c_int, P(VvmEnsStats)]),
        "vvm_dev_set_shape": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "vvm_dev_set_radiation": (C.c_int, [C.c_void_p, PTR, C.c_int]),
        "vvm_dev_set_host_microphysics": (C.c_int, [C.c_void_p, HOST_MP_FN, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name, None)
        if fn is None:
            raise RuntimeError(f"{path} does not export {name}")
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


# ----------------------------------------------------------------- config

CONSTANTS = dict(GRAVITY=9.80665, Cp=1003.5, Cv=716.5, Rd=287.0, Lv=2.5e6, P0=1e5, PSURF=96500.0,
                 tolerance=1e-22)


def make_config(nx_int: int, nz_int: int, dx: float, dz: float, dt: float, *, CASE: int = 1,
                addforcingtime: float = 0.0, moisture_nudge_time: float = 0.0, tropical_forcing: int = 1,
                solver_mode: int = VVM_SOLVER_FP64, guess_mode: int = VVM_GUESS_ZERO,
                stop_mode: int = VVM_STOP_PARITY, max_iter: int = 200, tol_rel: float = 1e-10) -> VvmDevConfig:
    """Config with the constants main() passes to Config_VVM (src/main.cpp:68-72)."""
    return VvmDevConfig(nx=nx_int + 2, nz=nz_int + 2, dx=dx, dz=dz, dt=dt, addforcingtime=addforcingtime,
                        moisture_nudge_time=moisture_nudge_time, CASE=CASE, tropical_forcing=tropical_forcing,
                        solver_mode=solver_mode, guess_mode=guess_mode, stop_mode=stop_mode,
                        max_iter=max_iter, tol_rel=tol_rel, **CONSTANTS)


@dataclass
class Case:
    name: str
    cfg: VvmDevConfig
    params: VvmCaseParams
    steps: int


def case(name: str, seed: int = 1, **over) -> Case:
    """Benchmark cases.  c0: 128x64 warm bubble; c1: 1024x128 with surface
    fluxes; c2: 2048x256 with the c1 forcing; tiny: 16x12 for operator tests."""
    sfc = dict(sfc_heat_flux=100.0, sfc_qv_ramp=5e-3, sfc_qv_ramp_z=500.0)
    if name == "c0":
        cfg = make_config(128, 64, 200.0, 200.0, 2.0, **over)
        p = VvmCaseParams(sounding=0, brunt=0.01, bubble_dT=2.0, bubble_radius=2000.0, bubble_zc=2000.0,
                          noise_kind=1, noise_amp=0.01, noise_ztop=1000.0, seed=seed)
        return Case(name, cfg, p, 600)
    if name in ("c1", "c2"):
        nxi, nzi, steps = (1024, 128, 1000) if name == "c1" else (2048, 256, 500)
        cfg = make_config(nxi, nzi, 100.0, 100.0, 1.0, **over)
        p = VvmCaseParams(sounding=0, brunt=0.01, noise_kind=2, noise_amp=0.1, noise_ztop=500.0, seed=seed, **sfc)
        return Case(name, cfg, p, steps)
    if name == "tiny":
        cfg = make_config(16, 12, 200.0, 200.0, 2.0, **over)
        p = VvmCaseParams(sounding=0, brunt=0.01, bubble_dT=2.0, bubble_radius=800.0, bubble_zc=600.0,
                          noise_kind=1, noise_amp=0.05, noise_ztop=1000.0, seed=seed,
                          sfc_heat_flux=100.0, sfc_qv_ramp=5e-3, sfc_qv_ramp_z=500.0)
        return Case(name, cfg, p, 10)
    raise KeyError(name)


def grid_case(nx_int: int, nz_int: int, dx: float = 100.0, dt: float = 1.0, seed: int = 1,
              sigma: float = 0.1, heat: float = 100.0, **over) -> Case:
    """A c1-style case on any grid (ensemble members, parity at odd sizes)."""
    cfg = make_config(nx_int, nz_int, dx, dx, dt, **over)
    p = VvmCaseParams(sounding=0, brunt=0.01, noise_kind=2, noise_amp=sigma, noise_ztop=500.0, seed=seed,
                      sfc_heat_flux=heat, sfc_qv_ramp=5e-3, sfc_qv_ramp_z=500.0)
    return Case(f"{nx_int}x{nz_int}", cfg, p, 300)


# ------------------------------------------------------------- host state

class HostState:
    """Host copy of the model state: 2-D arrays shaped (nx, nz) ([i][k]),
    histories (nx, nz, 2), rows (nx,), profiles (nz,)."""

    def __init__(self, nx: int, nz: int):
        self.nx, self.nz = nx, nz
        self.a: dict[str, np.ndarray] = {}
        for n in FIELDS2D:
            self.a[n] = np.zeros((nx, nz))
        for n in HISTS:
            self.a[n] = np.zeros((nx, nz, 2))
        for n in ROWS:
            self.a[n] = np.zeros(nx)
        for n in PROFILES:
            self.a[n] = np.zeros(nz)
        self.s = {n: 0.0 for n in SCALARS}
        self.dubarTop_advect = [0.0, 0.0]
        self.step = 0
        self.status_for_adding_forcing = 0

    def __getitem__(self, name):
        return self.a[name]

    def copy(self) -> "HostState":
        h = HostState(self.nx, self.nz)
        for n, arr in self.a.items():
            h.a[n][...] = arr
        h.s = dict(self.s)
        h.dubarTop_advect = list(self.dubarTop_advect)
        h.step = self.step
        h.status_for_adding_forcing = self.status_for_adding_forcing
        return h

    def view(self) -> VvmHostView:
        v = VvmHostView()
        for n, arr in self.a.items():
            assert arr.flags.c_contiguous and arr.dtype == np.float64
            setattr(v, n, arr.ctypes.data_as(PTR))
        for n in SCALARS:
            setattr(v, n, self.s[n])
        v.dubarTop_advect[0], v.dubarTop_advect[1] = self.dubarTop_advect
        v.step = self.step
        v.status_for_adding_forcing = self.status_for_adding_forcing
        return v

    def pull_scalars(self, v: VvmHostView) -> None:
        for n in SCALARS:
            self.s[n] = getattr(v, n)
        self.dubarTop_advect = [v.dubarTop_advect[0], v.dubarTop_advect[1]]
        self.step = v.step
        self.status_for_adding_forcing = v.status_for_adding_forcing

    # -------- state files (format of oracle/ref/ref_harness.cpp)
    def save(self, path: str, cfg: VvmDevConfig) -> None:
        recs = []
        for n in ("nx", "nz", "dx", "dz", "dt", "GRAVITY", "Cp", "Cv", "Rd", "Lv", "P0", "PSURF",
                  "addforcingtime", "moisture_nudge_time", "tolerance", "CASE"):
            recs.append((f"cfg.{n}", np.array([float(getattr(cfg, n))])))
        for n, arr in self.a.items():
            recs.append((n, arr))
        for n in SCALARS:
            recs.append((n, np.array([self.s[n]])))
        recs.append(("dubarTop_advect0", np.array([self.dubarTop_advect[0]])))
        recs.append(("dubarTop_advect1", np.array([self.dubarTop_advect[1]])))
        recs.append(("step", np.array([float(self.step)])))
        recs.append(("status_for_adding_forcing", np.array([float(self.status_for_adding_forcing)])))
        with open(path, "wb") as f:
            f.write(b"VVMSTATE" + struct.pack("<ii", 1, len(recs)))
            for name, arr in recs:
                dims = list(arr.shape) + [1] * (3 - arr.ndim)
                f.write(name.encode().ljust(32, b"\0") + struct.pack("<iiii", 3, *dims))
                f.write(np.ascontiguousarray(arr, dtype="<f8").tobytes())

    @staticmethod
    def read_records(path: str) -> dict[str, np.ndarray]:
        out = {}
        with open(path, "rb") as f:
            data = f.read()
        if data[:8] != b"VVMSTATE":
            raise ValueError("bad state file")
        _, nrec = struct.unpack_from("<ii", data, 8)
        off = 16
        for _ in range(nrec):
            name = data[off:off + 32].split(b"\0")[0].decode()
            _, d0, d1, d2 = struct.unpack_from("<iiii", data, off + 32)
            off += 48
            n = d0 * d1 * d2
            arr = np.frombuffer(data, dtype="<f8", count=n, offset=off).copy()
            off += 8 * n
            shape = [d for d in (d0, d1, d2)]
            while len(shape) > 1 and shape[-1] == 1:
                shape.pop()
            out[name] = arr.reshape(shape)
        return out

    @classmethod
    def load(cls, path: str) -> "HostState":
        r = cls.read_records(path)
        nx, nz = r["zeta"].shape
        h = cls(nx, nz)
        for n in h.a:
            if n in r:
                h.a[n][...] = r[n].reshape(h.a[n].shape)
        for n in SCALARS:
            h.s[n] = float(r[n][0])
        h.dubarTop_advect = [float(r["dubarTop_advect0"][0]), float(r["dubarTop_advect1"][0])]
        h.step = int(r["step"][0])
        h.status_for_adding_forcing = int(r["status_for_adding_forcing"][0])
        return h


def init_reference(cfg: VvmDevConfig, sounding_path: str | None = None) -> HostState:
    """The reference's own initial state (Init1d + Init2d): the analytic
    sounding, or the 10-column sounding file when a path is given."""
    lib = load_lib()
    h = HostState(cfg.nx, cfg.nz)
    v = h.view()
    rc = lib.vvm_init_reference(C.byref(cfg), sounding_path.encode() if sounding_path else None, C.byref(v))
    if rc != 0:
        raise VvmError(f"vvm_init_reference failed: {rc}")
    h.pull_scalars(v)
    return h


def init_from_previous(cfg: VvmDevConfig, sounding_path: str, nc_prev: str, nc_now: str) -> HostState:
    """The reference's restart from two consecutive netCDF outputs."""
    lib = load_lib()
    h = HostState(cfg.nx, cfg.nz)
    v = h.view()
    rc = lib.vvm_init_from_previous(C.byref(cfg), sounding_path.encode(), nc_prev.encode(), nc_now.encode(),
                                    C.byref(v))
    if rc != 0:
        raise VvmError(f"vvm_init_from_previous failed: {rc}")
    h.pull_scalars(v)
    return h


def init_state(c: Case) -> HostState:
    """Initial state of a case, from the library's host init layer."""
    lib = load_lib()
    h = HostState(c.cfg.nx, c.cfg.nz)
    v = h.view()
    rc = lib.vvm_init_case(C.byref(c.cfg), C.byref(c.params), C.byref(v))
    if rc != 0:
        raise RuntimeError(f"vvm_init_case failed: {rc}")
    h.pull_scalars(v)
    return h


# ---------------------------------------------------------------- device

class VvmError(RuntimeError):
    pass


class Device:
    """One device handle (one GPU, one stream)."""

    def __init__(self, cfg: VvmDevConfig, device: int = 0, handle: int | None = None):
        self.lib = load_lib()
        self.cfg = cfg
        self.owned = handle is None
        self.h = C.c_void_p(handle) if handle is not None else C.c_void_p()
        if self.owned:
            rc = self.lib.vvm_dev_create(C.byref(cfg), device, C.byref(self.h))
            if rc != 0:
                raise VvmError(f"vvm_dev_create failed ({rc})")

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            raise VvmError(f"{what} failed ({rc}): {self.lib.vvm_dev_last_error(self.h).decode()}")
        return rc

    def upload(self, st: HostState) -> None:
        v = st.view()
        self._check(self.lib.vvm_dev_upload(self.h, C.byref(v)), "upload")

    def download(self, st: HostState) -> None:
        v = st.view()
        self._check(self.lib.vvm_dev_download(self.h, C.byref(v)), "download")
        st.pull_scalars(v)

    def step(self, n: int = 1) -> VvmStepStats:
        s = VvmStepStats()
        self._check(self.lib.vvm_dev_step(self.h, n, C.byref(s)), "step")
        return s

    def step_timed(self, n: int) -> tuple[VvmStepStats, np.ndarray]:
        """n steps with every step's phase times: an (n, 5) array of seconds,
        columns advection, poisson, microphysics, diffusion, all."""
        t = np.zeros((n, 5))
        self._check(self.lib.vvm_dev_set_step_timers(self.h, t.ctypes.data_as(PTR), n), "set_step_timers")
        try:
            s = self.step(n)
        finally:
            self.lib.vvm_dev_set_step_timers(self.h, None, 0)
        return s, t

    def kernel_timer(self, kernel: int) -> None:
        self._check(self.lib.vvm_dev_kernel_timer(self.h, kernel), "kernel_timer")

    def kernel_time(self) -> tuple[float, int]:
        avg, n = C.c_double(), C.c_int()
        self._check(self.lib.vvm_dev_kernel_time(self.h, C.byref(avg), C.byref(n)), "kernel_time")
        return avg.value, n.value

    def op(self, name: str) -> int:
        return self._check(self.lib.vvm_dev_op(self.h, OPS[name]), f"op {name}")

    def op_field(self, op: int, field: int = 0) -> int:
        """One static of class vvm alone (VVM_OP_ADVECTION_ZETA ..
        VVM_OP_NUDGE_QV), with the field argument it takes."""
        return self._check(self.lib.vvm_dev_op_field(self.h, op, field), f"op {op}")

    def tune(self, steps: int = 3) -> "VvmTuneResult":
        r = VvmTuneResult()
        self._check(self.lib.vvm_dev_tune(self.h, steps, C.byref(r)), "tune")
        return r

    def set_strip_waves(self, cw: int) -> None:
        self._check(self.lib.vvm_dev_set_shape(self.h, cw, 0), "set_shape")

    def set_shape(self, cw: int = 0, coarse_kib: int = 0) -> None:
        self._check(self.lib.vvm_dev_set_shape(self.h, cw, coarse_kib), "set_shape")

    def set_host_microphysics(self, fn) -> None:
        """Run fn(view) (view: a ctypes VvmHostView pointer into host arrays in
        the reference's [i][k] layout) in place of the device Kessler each
        step; it may change thp, qvp, qcp, qrp and the process rates and
        returns 0.  None restores the device Kessler."""
        if fn is None:
            self._mp = None
            self._check(self.lib.vvm_dev_set_host_microphysics(self.h, HOST_MP_FN(), None), "set_host_microphysics")
            return
        self._mp = HOST_MP_FN(lambda user, view: int(fn(view)))  # kept alive with the handle
        self._check(self.lib.vvm_dev_set_host_microphysics(self.h, self._mp, None), "set_host_microphysics")

    def set_radiation(self, heating_rate: np.ndarray | None, cooling: bool = False) -> None:
        """The radiation heating rate [K/day], (nx, nz) in the reference's
        layout, added to thp by every later step (None removes it); cooling
        adds RADIATIONCOOLING's -2 K/day."""
        if heating_rate is None:
            self._check(self.lib.vvm_dev_set_radiation(self.h, None, int(cooling)), "set_radiation")
            return
        r = np.ascontiguousarray(heating_rate, dtype=np.float64)
        assert r.shape == (self.cfg.nx, self.cfg.nz), r.shape
        self._check(self.lib.vvm_dev_set_radiation(self.h, r.ctypes.data_as(PTR), int(cooling)), "set_radiation")

    def collect(self, capacity: int) -> None:
        """Record training samples of the learned guess (after upload)."""
        self._check(self.lib.vvm_dev_collect(self.h, capacity), "collect")

    def collected(self) -> int:
        return int(self.lib.vvm_dev_collected(self.h))

    def set_guess_net(self, net: "Net | None") -> None:
        self._check(self.lib.vvm_dev_set_guess_net(self.h, net.h if net is not None else None), "set_guess_net")

    def poisson(self, b: np.ndarray, x: np.ndarray) -> VvmSolveStats:
        s = VvmSolveStats()
        assert b.dtype == np.float64 and x.dtype == np.float64 and b.flags.c_contiguous and x.flags.c_contiguous
        # VVM_WARN_NOT_CONVERGED (1) is kept in last_rc, as the reference
        # warns and goes on (src/PoissonSolver.cpp:463-467)
        self.last_rc = self._check(self.lib.vvm_dev_poisson(self.h, b.ctypes.data_as(PTR), x.ctypes.data_as(PTR),
                                                            C.byref(s)), "poisson")
        return s

    def poisson_device(self, d_b: int, d_x: int) -> VvmSolveStats:
        s = VvmSolveStats()
        self._check(self.lib.vvm_dev_poisson_device(self.h, C.c_void_p(d_b), C.c_void_p(d_x), C.byref(s)),
                    "poisson_device")
        return s

    def close(self) -> None:
        if self.h and self.owned:
            self.lib.vvm_dev_destroy(self.h)
        self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VvmEnsMember(C.Structure):
    _fields_ = [("cfg", VvmDevConfig), ("params", VvmCaseParams)]


class VvmEnsStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("grid_point_steps", C.c_double), ("solve_iters_total", C.c_int64),
                ("not_converged", C.c_int32), ("members", C.c_int32)]


class Ensemble:
    """Members of uneven grids advanced together on one device
    (include/vvm_ens.h)."""

    def __init__(self, cases: list, device: int = 0, threads: int = 0):
        self.lib = load_lib()
        arr = (VvmEnsMember * len(cases))()
        for i, c in enumerate(cases):
            arr[i].cfg = c.cfg
            arr[i].params = c.params
        self.cases = list(cases)
        self.h = C.c_void_p()
        rc = self.lib.vvm_ens_create(arr, len(cases), device, threads, C.byref(self.h))
        if rc != 0:
            raise VvmError(f"vvm_ens_create failed ({rc})")

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            raise VvmError(f"{what} failed ({rc}): {self.lib.vvm_ens_last_error(self.h).decode()}")
        return rc

    def member(self, i: int) -> Device:
        h = self.lib.vvm_ens_handle(self.h, i)
        return Device(self.cases[i].cfg, handle=h)

    def set_guess_net(self, net: "Net | None") -> None:
        self._check(self.lib.vvm_ens_set_guess_net(self.h, net.h if net is not None else None), "set_guess_net")

    def tune(self, steps: int = 2) -> None:
        self._check(self.lib.vvm_ens_tune(self.h, steps), "tune")

    def step(self, n: int) -> VvmEnsStats:
        s = VvmEnsStats()
        self._check(self.lib.vvm_ens_step(self.h, n, C.byref(s)), "step")
        return s

    def close(self) -> None:
        if self.h:
            self.lib.vvm_ens_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Net:
    """The learned initial-guess network on one device (include/vvm_net.h)."""

    def __init__(self, device: int = 0):
        self.lib = load_lib()
        self.h = C.c_void_p()
        rc = self.lib.vvm_net_create(device, C.byref(self.h))
        if rc != 0:
            raise VvmError(f"vvm_net_create failed ({rc})")
        self.n = int(self.lib.vvm_net_param_count())

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            raise VvmError(f"{what} failed ({rc}): {self.lib.vvm_net_last_error(self.h).decode()}")
        return rc

    def init_params(self, seed: int) -> None:
        self._check(self.lib.vvm_net_init_params(self.h, C.c_uint64(seed)), "init_params")

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n
        self._check(self.lib.vvm_net_set_params(self.h, p.ctypes.data_as(FPTR)), "set_params")

    def params(self) -> np.ndarray:
        p = np.zeros(self.n, dtype=np.float32)
        self._check(self.lib.vvm_net_get_params(self.h, p.ctypes.data_as(FPTR)), "get_params")
        return p

    def forward(self, inp: np.ndarray, periodic: bool) -> np.ndarray:
        x = np.ascontiguousarray(inp, dtype=np.float32)
        nb, c, h, w = x.shape
        assert c == 3
        out = np.zeros((nb, h, w), dtype=np.float32)
        self._check(self.lib.vvm_net_forward_host(self.h, nb, h, w, int(periodic), x.ctypes.data_as(FPTR),
                                                  out.ctypes.data_as(FPTR)), "forward")
        return out

    def grad(self, inp, target, mask, periodic: bool):
        x = np.ascontiguousarray(inp, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.float32)
        nb, _, h, w = x.shape
        g = np.zeros(self.n, dtype=np.float32)
        loss = C.c_double(0.)
        self._check(self.lib.vvm_net_grad_host(self.h, nb, h, w, int(periodic), x.ctypes.data_as(FPTR),
                                               t.ctypes.data_as(FPTR), m.ctypes.data_as(FPTR),
                                               g.ctypes.data_as(FPTR), C.byref(loss)), "grad")
        return loss.value, g

    def adam(self, lr: float) -> None:
        self._check(self.lib.vvm_net_adam(self.h, C.c_float(lr)), "adam")

    @staticmethod
    def set_conv_rows(rows: int) -> None:
        """Tile height (2, 4 or 8 rows) of every network's convolutions."""
        rc = load_lib().vvm_net_set_conv_rows(int(rows))
        if rc != 0:
            raise ValueError(f"vvm_net_set_conv_rows({rows}) returned {rc}")

    @staticmethod
    def set_layer_pairs(on: bool) -> None:
        """Whether inference runs hidden layers two at a time in one kernel."""
        rc = load_lib().vvm_net_set_layer_pairs(int(bool(on)))
        if rc != 0:
            raise ValueError(f"vvm_net_set_layer_pairs({on}) returned {rc}")

    def add_samples(self, dev: "Device") -> None:
        self._check(self.lib.vvm_net_add_samples(self.h, dev.h), "add_samples")

    def pool_size(self) -> int:
        return int(self.lib.vvm_net_pool_size(self.h))

    def train(self, steps: int, batch: int = 16, crop: tuple = (64, 64), margin: int = 6, lr: float = 2e-3,
              lr_final: float = 1e-4, seed: int = 1, log_every: int = 100) -> VvmTrainStats:
        o = VvmTrainOpts(steps=steps, batch=batch, crop_h=crop[0], crop_w=crop[1], margin=margin, lr=lr,
                         lr_final=lr_final, seed=seed, log_every=log_every)
        s = VvmTrainStats()
        self._check(self.lib.vvm_net_train(self.h, C.byref(o), C.byref(s)), "train")
        return s

    def eval(self, dev: "Device") -> float:
        e = C.c_double(0.)
        self._check(self.lib.vvm_net_eval(self.h, dev.h, C.byref(e)), "eval")
        return e.value

    def save(self, path: str) -> None:
        self._check(self.lib.vvm_net_save(self.h, path.encode()), "save")

    def load(self, path: str) -> None:
        self._check(self.lib.vvm_net_load(self.h, path.encode()), "load")

    def close(self) -> None:
        if self.h:
            self.lib.vvm_net_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
