"""Host-side parts of the batch and serving path that need no GPU: the
batch entry points' argument checks, pool sizes, and the seeded request
stream (mix, sizes, determinism, arrivals, work bound)."""
import ctypes

import numpy as np

from mara2_amd import abi, problems, serving
from mara2_amd.runtime import Batch, load


def test_batch_argument_checks_and_no_device():
    lib = load()
    h = ctypes.c_void_p()
    assert lib.mara_batch_create(0, 2, 0, 8, ctypes.byref(h)) == abi.ERR_INVALID_ARGUMENT
    assert lib.mara_batch_create(1 << 20, 4, 0, 8, ctypes.byref(h)) == abi.ERR_INVALID_ARGUMENT
    assert lib.mara_batch_create(1 << 20, 2, 0, 0, ctypes.byref(h)) == abi.ERR_INVALID_ARGUMENT
    st = lib.mara_batch_create(1 << 20, 2, 0, 8, ctypes.byref(h))
    assert st in (abi.OK, abi.ERR_NO_DEVICE)
    if st == abi.OK:
        lib.mara_batch_destroy(h)
    assert lib.mara_batch_step(None, 1, None) == abi.ERR_INVALID_ARGUMENT
    assert lib.mara_batch_retire(None, 0, None) == abi.ERR_INVALID_ARGUMENT
    assert lib.mara_batch_free_bytes(None) == -1


def test_item_bytes_cover_state_and_fluxes():
    d = problems.riemann_desc(dict(dim=2, n=100, law="mhd"), precision="mixed")
    cells = 104 * 104 * 8
    faces = 2 * (105 * 105 * 2) * 8  # x and y face arrays; face extent N3 + 1 = 2 on the flat axis
    got = Batch.item_bytes_of(d, "mixed")
    assert 3 * cells * 8 + faces * 4 <= got < 3 * cells * 8 + faces * 4 + 8 * 4096
    assert Batch.item_bytes_of(d, "fp64") > got


def test_stream_mix_and_sizes():
    reqs = serving.make_stream(2000, with_data=False)
    dims = np.array([r.params["dim"] for r in reqs])
    mhd = np.array([r.params["law"] == "mhd" for r in reqs])
    assert abs((dims == 1).mean() - 0.7) < 0.04
    assert abs(mhd.mean() - 0.5) < 0.04
    n1 = np.array([r.params["n"] for r in reqs if r.params["dim"] == 1])
    n2 = np.array([r.params["n"] for r in reqs if r.params["dim"] == 2])
    assert n1.min() >= 256 and n1.max() <= 16384
    assert n2.min() >= 128 and n2.max() <= 1024
    # log-uniform: the median sits near the geometric mean of the range
    assert abs(np.log(np.median(n1)) - 0.5 * np.log(256 * 16384)) < 0.3
    for r in reqs[:50]:
        p = r.params
        assert 0.1 <= p["rho"].min() and p["rho"].max() <= 10 and 0.1 <= p["p"].min() and p["p"].max() <= 10
        assert np.abs(p["v"]).max() <= 1
        assert r.desc.riemann == (abi.RIEMANN_HLLE if p["law"] == "mhd" else abi.RIEMANN_HLLC)
        assert r.desc.field_ct == (1 if (p["law"] == "mhd" and p["dim"] == 2) else 0)


def test_stream_is_deterministic_and_prefix_stable():
    a = serving.make_stream(40, with_data=False)
    b = serving.make_stream(80, with_data=False)
    for x, y in zip(a, b):
        assert x.params["n"] == y.params["n"] and x.params["law"] == y.params["law"]
        assert np.array_equal(x.params["rho"], y.params["rho"])


def test_mhd_transverse_jump_keeps_normal_field():
    rng = np.random.default_rng(3)
    for _ in range(200):
        p = problems.riemann_params(rng)
        if p["law"] != "mhd":
            continue
        n = np.array([np.cos(p["angle"]), np.sin(p["angle"]), 0.0])
        BL, BR = p["B"]
        assert abs(np.dot(BL - BR, n)) < 1e-12


def test_arrivals_are_poisson_at_the_rate():
    reqs = serving.make_stream(4000, with_data=False)
    serving.set_arrivals(reqs, rate=8.0)
    t = np.array([r.arrival for r in reqs])
    gaps = np.diff(t)
    assert t[0] == 0.0 and np.all(gaps >= 0)
    assert abs(gaps.mean() - 1 / 8.0) < 0.01
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.06  # exponential gaps


def test_work_bound():
    reqs = serving.make_stream(300, with_data=False)
    for r in reqs:
        m = serving.max_steps(r)
        assert m <= serving.MAX_STEPS and m * r.zones <= serving.MAX_ZONE_UPDATES
    # the regular extremes stay unbounded: 1-D 16384 zones at 1.6e5 steps,
    # 2-D 1024^2 at 7e3 steps
    assert 16384 * 160000 <= serving.MAX_ZONE_UPDATES and 160000 <= serving.MAX_STEPS
    assert 1024 * 1024 * 7000 <= serving.MAX_ZONE_UPDATES


class _FakeBatch:
    """Stands in for the device batch: admits a solve while its bytes fit the
    budget (item bytes = zones * 100), finishes every admitted solve on its
    first step."""

    def __init__(self, hbm_budget, rk_order=2, precision="mixed", max_solves=8):
        self.budget, self.used, self.slots, self.max = hbm_budget, 0, {}, max_solves

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        pass

    def set_graphs(self, on):
        pass

    def set_express(self, n):
        pass

    def set_shape(self, s):
        pass

    @property
    def free_bytes(self):
        return self.budget - self.used

    def admit(self, desc, P, cfl, t_final, max_steps):
        nbytes = int(np.prod(list(desc.cells))) * 100
        if self.used + nbytes > self.budget or len(self.slots) >= self.max:
            return None
        slot = min(set(range(self.max)) - set(self.slots))
        self.slots[slot] = nbytes
        self.used += nbytes
        return slot

    def step(self, n):
        return 0

    def finished(self):
        return sorted(self.slots)

    def query(self, slot):
        rep = abi.MaraReport()
        return 1, serving.T_FINAL, 10, 0, rep

    def retire(self, slot, download=True):
        self.used -= self.slots.pop(slot)
        return None


def test_serve_finishes_when_a_request_can_never_be_admitted():
    """A request larger than the whole budget is reported with
    MARA_ERR_CAPACITY instead of keeping the loop alive; the others are
    served."""
    reqs = serving.make_stream(6, with_data=False, resize=lambda dim, n: 300 if dim == 1 else 20)
    for r in reqs:
        r.P = np.zeros(abi.array_shape(r.desc))
    big = serving.make_stream(1, with_data=False, resize=lambda dim, n: 100000 if dim == 1 else 1000)[0]
    big.index = 99
    big.P = None
    reqs.append(big)
    serving.set_arrivals(reqs, rate=1e6)
    budget = 300 * 100 * 2
    wall, done = serving.serve(reqs, budget, max_solves=8, batch_factory=_FakeBatch)
    assert len(done) == len(reqs)
    by_index = {r.index: r for r in done}
    assert by_index[99].status == abi.ERR_CAPACITY
    assert all(r.status == 0 for r in done if r.index != 99)
