"""mara_advance_noalloc -- one step whose working buffers come from the
library's static scratch array -- against mara_block_advance on the same
input: bit for bit (FP64 and mixed alike: the same kernels run on the same
data).  Calls follow one another with smaller inputs, so each must clear
what the previous one left in the shared array; one runs in place.  The
device buffers of the test are allocated through the HIP runtime the
library links."""
import ctypes

import numpy as np
import pytest

from mara2_amd import abi, problems
from mara2_amd.abi import MaraDesc, MaraReport
from mara2_amd.runtime import Block, load

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        load()  # maps the HIP runtime libmara_gpu.so links against
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def put(self, dst, a):
        assert self.h.hipMemcpy(dst, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a, src):
        assert self.h.hipMemcpy(a.ctypes.data, src, a.nbytes, D2H) == 0

    def free(self, p):
        assert self.h.hipFree(p) == 0


def _lib():
    lib = load()
    lib.mara_advance_noalloc.argtypes = [ctypes.POINTER(MaraDesc), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(MaraReport)]
    return lib


def _noalloc(hip, desc, P, dt, inplace=False):
    P = np.ascontiguousarray(P)
    din = hip.alloc(P.nbytes)
    dout = din if inplace else hip.alloc(P.nbytes)
    try:
        hip.put(din, P)
        rep = MaraReport()
        st = _lib().mara_advance_noalloc(ctypes.byref(desc), din, dout, 0.0, dt, None, ctypes.byref(rep))
        out = np.empty_like(P)
        hip.get(out, dout)
    finally:
        hip.free(din)
        if not inplace:
            hip.free(dout)
    return st, rep, out


def _block(desc, P, dt):
    with Block(desc) as blk:
        blk.upload(P)
        st, rep = blk.advance(dt, check=False)
        return st, rep, blk.download()


def _cases():
    """Largest first: each later call is smaller than the one before."""
    out = []
    d = problems.turbulence_desc(40)                       # fused 3-D hydro, RK2
    out.append(("turbulence 40^3 fp64", d, problems.turbulence_3d(d), 0.3, False))
    d = problems.turbulence_desc(24, "mixed")              # fused 3-D hydro, binary32 fluxes
    out.append(("turbulence 24^3 mixed", d, problems.turbulence_3d(problems.turbulence_desc(24)), 0.3, False))
    d = abi.make_desc(law="rmhd", gamma=4.0 / 3.0, recon="weno5", riemann="hlle", cells=(12, 12, 12), rk_order=3)
    out.append(("rmhd 12^3 flux arrays", d, problems.rmhd_3d(d, kmax=2, wmax=2.0), 0.3, False))
    d = problems.mhd_desc(48)                              # fused 2-D MHD + CT
    out.append(("mhd 48^2", d, problems.mhd_2d(d), 0.4, True))
    d = abi.make_desc(law="hydro", gamma=1.4, recon="plm", riemann="hlle", bc="outflow", cells=(256, 1, 1),
                      rk_order=3)                          # 1-D pencils, RK3
    out.append(("sod 256 rk3", d, problems.sod_1d(d), 0.4, False))
    return out


def test_noalloc_step_equals_block_step_on_shrinking_inputs():
    hip = _Hip()
    for name, desc, P0, cfl, inplace in _cases():
        with Block(desc) as blk:
            blk.upload(P0)
            dt = cfl * blk.courant_dt()
        st_b, rep_b, Pb = _block(desc, P0, dt)
        st_n, rep_n, Pn = _noalloc(hip, desc, P0, dt, inplace)
        print(f"{name}: status {st_n}, {rep_n.as_dict()}", flush=True)
        assert st_b == 0 and st_n == 0, (name, st_b, st_n)
        assert rep_n.as_dict() == rep_b.as_dict(), name
        assert np.array_equal(Pn, Pb), (name, int(np.sum(Pn != Pb)))


def test_noalloc_repeated_calls_are_independent():
    """The same call twice gives the same answer (nothing of the first call's
    scratch contents leaks into the second)."""
    hip = _Hip()
    desc = problems.mhd_desc(32)
    P0 = problems.mhd_2d(desc)
    a = _noalloc(hip, desc, P0, 1e-3)
    b = _noalloc(hip, desc, P0, 1e-3)
    assert a[0] == 0 and b[0] == 0
    assert np.array_equal(a[2], b[2])


def test_noalloc_refuses_what_does_not_fit():
    lib = _lib()
    desc = problems.turbulence_desc(256)
    rep = MaraReport()
    st = lib.mara_advance_noalloc(ctypes.byref(desc), ctypes.c_void_p(1), ctypes.c_void_p(1), 0.0, 1e-3, None,
                                  ctypes.byref(rep))
    assert st == abi.ERR_CAPACITY
    assert lib.mara_advance_noalloc(ctypes.byref(desc), None, None, 0.0, 1e-3, None, None) == abi.ERR_INVALID_ARGUMENT
