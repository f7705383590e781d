"""ctypes binding of libmara_gpu.so (the C ABI of include/mara_gpu.h).

The HIP library is the only compute path: if it is missing, loading raises
instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

from . import abi
from .abi import MaraDesc, MaraReport, MaraShape, array_shape

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmara_gpu.so")
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)


class MaraError(RuntimeError):
    """A non-zero status from the C ABI, with the text of mara_error_string."""

    def __init__(self, status, where, report=None):
        self.status = status
        self.report = report
        text = load().mara_error_string(status).decode()
        super().__init__(f"{where}: {text} (status {status})" + (f" {report}" if report else ""))


def load():
    """Load libmara_gpu.so once; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(make -C mara2_amd/csrc); there is no other compute path")
    L = ctypes.CDLL(LIB_PATH)
    blk = ctypes.c_void_p
    L.mara_version.restype = ctypes.c_char_p
    L.mara_error_string.restype = ctypes.c_char_p
    L.mara_error_string.argtypes = [ctypes.c_int]
    L.mara_num_fields.argtypes = [ctypes.c_int]
    L.mara_desc_num_fields.argtypes = [ctypes.POINTER(MaraDesc)]
    L.mara_stencil_size.argtypes = [ctypes.c_int]
    L.mara_desc_device_bytes.argtypes = [ctypes.POINTER(MaraDesc)]
    L.mara_desc_device_bytes.restype = ctypes.c_longlong
    L.mara_block_create.argtypes = [ctypes.POINTER(MaraDesc), ctypes.POINTER(blk)]
    L.mara_block_destroy.argtypes = [blk]
    L.mara_block_upload_prim.argtypes = [blk, _dp]
    L.mara_block_download_prim.argtypes = [blk, _dp]
    L.mara_block_advance.argtypes = [blk, ctypes.c_double, ctypes.c_double, ctypes.POINTER(MaraReport)]
    L.mara_block_courant_dt.argtypes = [blk, _dp]
    L.mara_block_apply_bc.argtypes = [blk]
    L.mara_block_monopole_max.argtypes = [blk, _dp]
    L.mara_block_run.argtypes = [blk, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp,
                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(MaraReport)]
    L.mara_block_set_timing.argtypes = [blk, ctypes.c_int]
    L.mara_block_set_graphs.argtypes = [blk, ctypes.c_int]
    L.mara_block_set_source.argtypes = [blk, ctypes.c_char_p, ctypes.c_char_p, _dp, ctypes.c_int]
    L.mara_block_set_source_aux.argtypes = [blk, _dp]
    L.mara_block_set_viscous_flux.argtypes = [blk, ctypes.c_char_p, ctypes.c_char_p, _dp, ctypes.c_int]
    shp = ctypes.POINTER(MaraShape)
    L.mara_block_get_shape.argtypes = [blk, shp]
    L.mara_block_set_shape.argtypes = [blk, shp]
    L.mara_block_autotune.argtypes = [blk, ctypes.c_int, shp, _dp]
    L.mara_tune_lookup.argtypes = [ctypes.POINTER(MaraDesc), shp]
    L.mara_tune_save.argtypes = [ctypes.c_char_p]
    L.mara_tune_load.argtypes = [ctypes.c_char_p]
    L.mara_block_stage_kernel_time.argtypes = [blk, _dp, ctypes.POINTER(ctypes.c_longlong),
                                               ctypes.POINTER(ctypes.c_char_p)]
    L.mara_block_device_bytes.argtypes = [blk]
    L.mara_block_device_bytes.restype = ctypes.c_longlong
    bat = ctypes.c_void_p
    L.mara_batch_create.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(bat)]
    L.mara_batch_set_graphs.argtypes = [bat, ctypes.c_int]
    L.mara_batch_get_shape.argtypes = [bat, ctypes.POINTER(MaraShape)]
    L.mara_batch_lane.argtypes = [bat, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.mara_batch_set_express.argtypes = [bat, ctypes.c_int]
    L.mara_batch_set_persistent.argtypes = [bat, ctypes.c_int]
    L.mara_batch_set_source.argtypes = [bat, ctypes.c_char_p, ctypes.c_char_p, _dp, ctypes.c_int]
    L.mara_batch_set_source_aux.argtypes = [bat, _dp]
    L.mara_batch_set_shape.argtypes = [bat, ctypes.POINTER(MaraShape)]
    L.mara_batch_destroy.argtypes = [bat]
    L.mara_batch_item_bytes.argtypes = [ctypes.POINTER(MaraDesc), ctypes.c_int]
    L.mara_batch_item_bytes.restype = ctypes.c_longlong
    L.mara_batch_free_bytes.argtypes = [bat]
    L.mara_batch_free_bytes.restype = ctypes.c_longlong
    L.mara_batch_admit.argtypes = [bat, ctypes.POINTER(MaraDesc), _dp, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.mara_batch_step.argtypes = [bat, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.mara_batch_query.argtypes = [bat, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(MaraReport)]
    L.mara_batch_retire.argtypes = [bat, ctypes.c_int, _dp]
    L.mara_batch_finished.argtypes = [bat, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    grp = ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    L.mara_partition.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip]
    L.mara_group_create.argtypes = [ctypes.POINTER(MaraDesc), ip, ctypes.POINTER(grp)]
    L.mara_group_destroy.argtypes = [grp]
    L.mara_group_num_blocks.argtypes = [grp]
    L.mara_group_block_region.argtypes = [grp, ctypes.c_int, ip, ip]
    L.mara_group_device_bytes.argtypes = [grp]
    L.mara_group_device_bytes.restype = ctypes.c_longlong
    L.mara_group_upload_prim.argtypes = [grp, _dp]
    L.mara_group_download_prim.argtypes = [grp, _dp]
    L.mara_group_apply_bc.argtypes = [grp]
    L.mara_group_advance.argtypes = [grp, ctypes.c_double, ctypes.c_double, ctypes.POINTER(MaraReport)]
    L.mara_group_courant_dt.argtypes = [grp, _dp]
    _LIB = L
    return L


def _check(status, where, report=None):
    if status != abi.OK:
        raise MaraError(status, where, report.as_dict() if report is not None else None)


def tune_lookup(desc):
    """Shape the tuning table holds for desc, or None."""
    s = MaraShape()
    r = load().mara_tune_lookup(ctypes.byref(desc), ctypes.byref(s))
    if r < 0:
        _check(r, "mara_tune_lookup")
    return s if r == 1 else None


def tune_save(path):
    _check(load().mara_tune_save(os.fsencode(path)), "mara_tune_save")


def tune_load(path):
    _check(load().mara_tune_load(os.fsencode(path)), "mara_tune_load")


class Block:
    """One device-resident mesh block (MeshData + MethodOfLinesTVD on the GPU)."""

    def __init__(self, desc: MaraDesc):
        self.lib = load()
        self.desc = desc
        self.shape = array_shape(desc)
        self._h = ctypes.c_void_p()
        _check(self.lib.mara_block_create(ctypes.byref(desc), ctypes.byref(self._h)), "mara_block_create")

    def close(self):
        if self._h:
            self.lib.mara_block_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def device_bytes(self):
        return self.lib.mara_block_device_bytes(self._h)

    def upload(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.shape == self.shape, (P.shape, self.shape)
        _check(self.lib.mara_block_upload_prim(self._h, P.ctypes.data_as(_dp)), "mara_block_upload_prim")

    def download(self):
        P = np.empty(self.shape)
        _check(self.lib.mara_block_download_prim(self._h, P.ctypes.data_as(_dp)), "mara_block_download_prim")
        return P

    def advance(self, dt, t0=0.0, check=True):
        rep = MaraReport()
        st = self.lib.mara_block_advance(self._h, t0, dt, ctypes.byref(rep))
        if check:
            _check(st, "mara_block_advance", rep)
        return st, rep

    def courant_dt(self):
        dt = ctypes.c_double(0)
        _check(self.lib.mara_block_courant_dt(self._h, ctypes.byref(dt)), "mara_block_courant_dt")
        return dt.value

    def apply_bc(self):
        _check(self.lib.mara_block_apply_bc(self._h), "mara_block_apply_bc")

    def monopole_max(self):
        m = ctypes.c_double(0)
        _check(self.lib.mara_block_monopole_max(self._h, ctypes.byref(m)), "mara_block_monopole_max")
        return m.value

    def run(self, cfl, t_final, max_steps, t=0.0, check=True):
        tt = ctypes.c_double(t)
        steps = ctypes.c_int(0)
        rep = MaraReport()
        st = self.lib.mara_block_run(self._h, cfl, t_final, max_steps, ctypes.byref(tt),
                                     ctypes.byref(steps), ctypes.byref(rep))
        if check:
            _check(st, "mara_block_run", rep)
        return tt.value, steps.value, rep

    def set_source(self, code_object, kernel=None, params=()):
        """Load a source-term code object (include/mara_source.h) and its
        kernel; S is added to L in every stage.  code_object=None removes it."""
        pr = np.ascontiguousarray(params, dtype=np.float64)
        _check(self.lib.mara_block_set_source(
            self._h, os.fsencode(code_object) if code_object else None, kernel.encode() if kernel else None,
            pr.ctypes.data_as(_dp) if len(pr) else None, len(pr)), "mara_block_set_source")

    def set_viscous_flux(self, code_object, kernel=None, params=()):
        """Load a viscous-flux code object (include/mara_source.h,
        MARA_VISCOUS_FLUX_KERNEL) and its kernel; it adds to the Godunov fluxes
        of every stage before the divergence.  code_object=None removes it."""
        pr = np.ascontiguousarray(params, dtype=np.float64)
        _check(self.lib.mara_block_set_viscous_flux(
            self._h, os.fsencode(code_object) if code_object else None, kernel.encode() if kernel else None,
            pr.ctypes.data_as(_dp) if len(pr) else None, len(pr)), "mara_block_set_viscous_flux")

    def set_source_aux(self, t):
        """The 8 auxiliary values passed to the source and viscous-flux
        functions."""
        t8 = np.zeros(8)
        t8[:len(t)] = t
        _check(self.lib.mara_block_set_source_aux(self._h, t8.ctypes.data_as(_dp)), "mara_block_set_source_aux")

    def set_graphs(self, on=True):
        """Replay run() steps from a captured HIP graph (default) or launch
        every kernel from the host."""
        _check(self.lib.mara_block_set_graphs(self._h, 1 if on else 0), "mara_block_set_graphs")

    def set_timing(self, on=True):
        _check(self.lib.mara_block_set_timing(self._h, 1 if on else 0), "mara_block_set_timing")

    def get_shape(self):
        s = MaraShape()
        _check(self.lib.mara_block_get_shape(self._h, ctypes.byref(s)), "mara_block_get_shape")
        return s

    def set_shape(self, shape):
        _check(self.lib.mara_block_set_shape(self._h, ctypes.byref(shape)), "mara_block_set_shape")

    def autotune(self, repeats=5):
        """Time the candidate launch shapes of the block's stage kernels, keep
        the fastest; returns (shape, mean stage ms)."""
        s = MaraShape()
        ms = ctypes.c_double(0)
        _check(self.lib.mara_block_autotune(self._h, repeats, ctypes.byref(s), ctypes.byref(ms)),
               "mara_block_autotune")
        return s, ms.value

    def stage_kernel_time(self):
        ms = ctypes.c_double(0)
        n = ctypes.c_longlong(0)
        name = ctypes.c_char_p()
        _check(self.lib.mara_block_stage_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(n),
                                                     ctypes.byref(name)), "mara_block_stage_kernel_time")
        return ms.value, n.value, name.value.decode()


class Batch:
    """Many independent solves resident in one device pool, stepped together
    (mara_batch_* of include/mara_gpu.h)."""

    def __init__(self, hbm_budget, rk_order=2, precision="fp64", max_solves=1024):
        self.lib = load()
        self.precision = abi.PRECISIONS[precision] if isinstance(precision, str) else precision
        self._h = ctypes.c_void_p()
        self._shapes = {}
        _check(self.lib.mara_batch_create(int(hbm_budget), rk_order, self.precision, max_solves,
                                          ctypes.byref(self._h)), "mara_batch_create")

    def close(self):
        if self._h:
            self.lib.mara_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def item_bytes(self, desc):
        return self.lib.mara_batch_item_bytes(ctypes.byref(desc), self.precision)

    @staticmethod
    def item_bytes_of(desc, precision):
        """Pool bytes of one solve (no device needed)."""
        prec = abi.PRECISIONS[precision] if isinstance(precision, str) else precision
        return load().mara_batch_item_bytes(ctypes.byref(desc), prec)

    @property
    def free_bytes(self):
        return self.lib.mara_batch_free_bytes(self._h)

    def set_graphs(self, on=True):
        """Replay each round of step() from a captured HIP graph (default) or
        launch every kernel from the host."""
        _check(self.lib.mara_batch_set_graphs(self._h, 1 if on else 0), "mara_batch_set_graphs")

    def get_shape(self):
        s = MaraShape()
        _check(self.lib.mara_batch_get_shape(self._h, ctypes.byref(s)), "mara_batch_get_shape")
        return s

    def set_express(self, max_solves):
        """The most long-chain solves the express lane may hold (0: none)."""
        _check(self.lib.mara_batch_set_express(self._h, max_solves), "mara_batch_set_express")

    def set_source(self, code_object, kernel=None, params=()):
        """Source terms for every solve (include/mara_source.h; the code
        object's <kernel>_batch kernel); before the first admission.
        code_object=None removes them."""
        pr = np.ascontiguousarray(params, dtype=np.float64)
        _check(self.lib.mara_batch_set_source(
            self._h, os.fsencode(code_object) if code_object else None, kernel.encode() if kernel else None,
            pr.ctypes.data_as(_dp) if len(pr) else None, len(pr)), "mara_batch_set_source")

    def set_source_aux(self, t):
        t8 = np.zeros(8)
        t8[:len(t)] = t
        _check(self.lib.mara_batch_set_source_aux(self._h, t8.ctypes.data_as(_dp)), "mara_batch_set_source_aux")

    def set_persistent(self, workgroups_per_cu):
        """Run the 2-D/3-D hydro + PLM tiles as a work queue over a fixed grid
        of workgroups_per_cu work-groups per CU (0: one work-group per tile)."""
        _check(self.lib.mara_batch_set_persistent(self._h, int(workgroups_per_cu)), "mara_batch_set_persistent")

    def lane(self, slot):
        """0 tile lane, 1 pencil lane, 2 express lane."""
        lane = ctypes.c_int(-1)
        _check(self.lib.mara_batch_lane(self._h, slot, ctypes.byref(lane)), "mara_batch_lane")
        return lane.value

    def set_shape(self, shape):
        _check(self.lib.mara_batch_set_shape(self._h, ctypes.byref(shape)), "mara_batch_set_shape")

    def admit(self, desc, P, cfl, t_final, max_steps):
        """Slot of the admitted solve, or None when the pool or the slot table
        is full (MARA_ERR_CAPACITY)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.shape == array_shape(desc), (P.shape, array_shape(desc))
        slot = ctypes.c_int(-1)
        st = self.lib.mara_batch_admit(self._h, ctypes.byref(desc), P.ctypes.data_as(_dp), cfl, t_final,
                                       max_steps, ctypes.byref(slot))
        if st == abi.ERR_CAPACITY:
            return None
        _check(st, "mara_batch_admit")
        self._shapes[slot.value] = array_shape(desc)
        return slot.value

    def step(self, n_steps):
        active = ctypes.c_int(0)
        _check(self.lib.mara_batch_step(self._h, n_steps, ctypes.byref(active)), "mara_batch_step")
        return active.value

    def query(self, slot):
        """(done, t, steps, status, report) of one solve."""
        done = ctypes.c_int(0)
        t = ctypes.c_double(0)
        steps = ctypes.c_int(0)
        rep = MaraReport()
        st = self.lib.mara_batch_query(self._h, slot, ctypes.byref(done), ctypes.byref(t), ctypes.byref(steps),
                                       ctypes.byref(rep))
        if st == abi.ERR_INVALID_ARGUMENT:
            _check(st, "mara_batch_query")
        return bool(done.value), t.value, steps.value, st, rep

    def finished(self):
        """Slots of solves that are done and not yet retired."""
        cap = max(len(self._shapes), 1)
        buf = (ctypes.c_int * cap)()
        n = ctypes.c_int(0)
        _check(self.lib.mara_batch_finished(self._h, buf, cap, ctypes.byref(n)), "mara_batch_finished")
        return list(buf[:n.value])

    def retire(self, slot, download=True):
        P = np.empty(self._shapes[slot]) if download else None
        _check(self.lib.mara_batch_retire(self._h, slot, P.ctypes.data_as(_dp) if download else None),
               "mara_batch_retire")
        del self._shapes[slot]
        return P


def desc_device_bytes(desc):
    return load().mara_desc_device_bytes(ctypes.byref(desc))


def partition(n, p, which):
    """(count, start) of partition `which` of n elements over p partitions
    (BlockDecomposition::partition / startIndex)."""
    c, s = ctypes.c_int(0), ctypes.c_int(0)
    _check(load().mara_partition(n, p, which, ctypes.byref(c), ctypes.byref(s)), "mara_partition")
    return c.value, s.value


class Group:
    """One mesh split into dims[0] x dims[1] x dims[2] blocks on one device
    (the single-device half of BlockDecomposition), with the guard exchange
    after every stage."""

    def __init__(self, desc: MaraDesc, dims):
        self.lib = load()
        self.desc = desc
        self.shape = array_shape(desc)
        self.dims = tuple(int(x) for x in dims)
        d3 = (ctypes.c_int * 3)(*self.dims)
        self._h = ctypes.c_void_p()
        _check(self.lib.mara_group_create(ctypes.byref(desc), d3, ctypes.byref(self._h)), "mara_group_create")

    def close(self):
        if self._h:
            self.lib.mara_group_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def num_blocks(self):
        return self.lib.mara_group_num_blocks(self._h)

    @property
    def device_bytes(self):
        return self.lib.mara_group_device_bytes(self._h)

    def block_region(self, index):
        s, c = (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
        _check(self.lib.mara_group_block_region(self._h, index, s, c), "mara_group_block_region")
        return tuple(s), tuple(c)

    def upload(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.shape == self.shape, (P.shape, self.shape)
        _check(self.lib.mara_group_upload_prim(self._h, P.ctypes.data_as(_dp)), "mara_group_upload_prim")

    def download(self):
        P = np.zeros(self.shape)
        _check(self.lib.mara_group_download_prim(self._h, P.ctypes.data_as(_dp)), "mara_group_download_prim")
        return P

    def apply_bc(self):
        _check(self.lib.mara_group_apply_bc(self._h), "mara_group_apply_bc")

    def advance(self, dt, t0=0.0, check=True):
        rep = MaraReport()
        st = self.lib.mara_group_advance(self._h, t0, dt, ctypes.byref(rep))
        if check:
            _check(st, "mara_group_advance", rep)
        return st, rep

    def courant_dt(self):
        dt = ctypes.c_double(0)
        _check(self.lib.mara_group_courant_dt(self._h, ctypes.byref(dt)), "mara_group_courant_dt")
        return dt.value
