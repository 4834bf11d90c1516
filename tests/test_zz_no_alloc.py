"""vvm_negative_value_noalloc (the entry point that allocates nothing)
against the handle's NegativeValueProcess operator
(vvm_dev_op_field(h, VVM_OP_NEGATIVE_VALUE, f); src/MicroPhysics.cpp:116-138),
through the C ABI, on two calls in a row (the second on a smaller field, so
it reuses the static scratch the first one left behind), plus the size
limit.  Both run the same kernels over the same 64x4 workgroups, so the
partial sums, the ratio and every output bit agree; the oracle's restatement
is checked beside them at 1e-14 (its serial sum order differs)."""
import ctypes as C

import numpy as np
import pytest

import vvm_amd as V
from helpers import perturbed_state

pytestmark = pytest.mark.gpu

OP_NEGATIVE_VALUE = 33
FIELD_QC_NOW = 1 * 8 + 3  # VVM_FIELD_AT(VVM_FIELD_QC, VVM_LEVEL_NOW)
VVM_ERR_SIZE = -3


def _noalloc_fn():
    lib = V.load_lib()
    fn = lib.vvm_negative_value_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    return fn


def _field(nx_int, nz_int, seed):
    c = V.grid_case(nx_int, nz_int, dx=200.0, dt=2.0)
    st = perturbed_state(c, seed=seed, step=0)
    # a cloud field with about a third of the points negative, outweighed
    # by the positives, so the rescale branch runs
    rng = np.random.default_rng(seed + 1)
    st["qc"][:] = rng.normal(4e-4, 5e-4, st["qc"].shape)
    return c, st


def _handle_result(c, st):
    with V.Device(c.cfg) as dev:
        dev.upload(st)
        dev.op_field(OP_NEGATIVE_VALUE, FIELD_QC_NOW)
        got = st.copy()
        dev.download(got)
    return got["qc"]


class _Hip:
    """The HIP runtime the library is linked against (same soname, so the
    same loaded copy): the test, as the caller, owns the device buffer."""

    def __init__(self):
        V.load_lib()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.chk(self.rt.hipSetDevice(0), "hipSetDevice")

    @staticmethod
    def chk(rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.chk(self.rt.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p

    def h2d(self, d, a):
        self.chk(self.rt.hipMemcpy(d, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")

    def d2h(self, a, d):
        self.chk(self.rt.hipMemcpy(a.ctypes.data, d, a.nbytes, 2), "hipMemcpy D2H")


def _noalloc_result(fn, a_ik, pad):
    hip = _Hip()
    nx, nz = a_ik.shape
    pitch = nx + pad
    host = np.zeros((nz, pitch))
    host[:, :nx] = a_ik.T
    d = hip.alloc(host.nbytes)
    try:
        hip.h2d(d, host)
        rc = fn(d, nx, nz, pitch, None)  # the null stream
        assert rc == 0, rc
        out = np.empty_like(host)
        hip.d2h(out, d)  # synchronous: waits for the null stream
    finally:
        hip.rt.hipFree(d)
    assert np.all(out[:, nx:] == 0.0), "wrote beyond the row"
    return np.ascontiguousarray(out[:, :nx].T)


def _oracle_negval(a_ik):
    """The reference's loop (src/MicroPhysics.cpp:116-138), k outer and i
    inner, serial sums."""
    a = a_ik.copy()
    nx, nz = a.shape
    inner = a[1:nx - 1, 1:nz - 1]
    neg = inner < 0.
    positive = float(np.sum(inner[~neg].T, dtype=np.float64))
    negative = float(np.sum(inner[neg].T, dtype=np.float64))
    inner[neg] = 0.
    if positive == 0. or abs(negative) > positive:
        return a
    ratio = 1. - abs(negative / positive)
    inner[inner > 0] *= ratio
    return a


def test_noalloc_matches_handle_operator_twice():
    fn = _noalloc_fn()
    for nx_int, nz_int, seed, pad in ((1024, 126, 5, 62), (126, 30, 9, 0)):
        c, st = _field(nx_int, nz_int, seed)
        print(f"noalloc NegativeValueProcess: {nx_int}x{nz_int}, seed {seed}", flush=True)
        a0 = st["qc"].copy()
        assert (a0[1:-1, 1:-1] < 0).any()
        want = _handle_result(c, st)
        got = _noalloc_result(fn, a0, pad)
        assert np.array_equal(got, want), float(np.max(np.abs(got - want)))
        ref = _oracle_negval(a0)
        e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert e <= 1e-14, e
        assert not np.array_equal(got, a0)


def test_noalloc_rejects_grids_beyond_its_scratch():
    """8192 workgroups of 64x4 points at most: a 4098 x 1026 field needs
    65 x 257 of them; the call returns VVM_ERR_SIZE before touching the
    GPU (the pointer is never read)."""
    fn = _noalloc_fn()
    hip = _Hip()
    d = hip.alloc(64)
    try:
        assert fn(d, 4098, 1026, 4098, None) == VVM_ERR_SIZE
        assert fn(d, 2050, 996, 2050, None) == VVM_ERR_SIZE
    finally:
        hip.rt.hipFree(d)
