"""Checkpoint mirror: the reference's dataset names and shapes, and a round
trip that restores MeshData::P (interior from the file, guards from the
boundary condition)."""
import numpy as np

from mara2_amd import abi, checkpoint, problems


def test_names_shapes_and_points(tmp_path):
    desc = problems.mhd_desc(16)
    P = problems.mhd_2d(desc)
    path = tmp_path / "chkpt.0000.npz"
    checkpoint.write(path, desc, P, 0.25, 7)
    with np.load(path) as f:
        keys = set(f.files)
        assert {"primitive/density", "primitive/velocity1", "primitive/pressure", "primitive/magnetic3",
                "status/simulationTime", "status/simulationIter", "mesh/type", "mesh/points/x"} <= keys
        assert f["primitive/density"].shape == (16, 16, 1)
        assert str(f["mesh/type"]) == "cartesian"
        x = f["mesh/points/x"]
        assert x.shape == (17,) and x[0] == 0.0 and x[-1] == 1.0


def test_round_trip_restores_state(tmp_path):
    for desc, make in ((problems.sod_desc(64), problems.sod_1d), (problems.mhd_desc(16), problems.mhd_2d)):
        P = make(desc)
        path = tmp_path / "c.npz"
        checkpoint.write(path, desc, P, 0.5, 3)
        Q, t, it = checkpoint.read(path, desc)
        assert (t, it) == (0.5, 3)
        assert np.array_equal(P, Q)


def test_cooling_auxiliary_field_survives(tmp_path):
    """The cooled NewtonianMHD state has a ninth field, written as
    primitive/auxiliary1 (src/ConservationLaws.cpp:520-534) and restored."""
    desc = abi.make_desc(law="mhd", recon="plm", cells=(12, 10, 1), cooling_rate=0.5)
    rng = np.random.default_rng(3)
    P = problems.empty(desc)
    P[...] = rng.random(P.shape) + 0.5
    problems.apply_bc(desc, P)
    path = tmp_path / "cool.npz"
    checkpoint.write(path, desc, P, 1.25, 40)
    with np.load(path) as f:
        assert "primitive/auxiliary1" in f.files
        assert np.array_equal(f["primitive/auxiliary1"], P[2:14, 2:12, :, 8])
    Q, t, it = checkpoint.read(path, desc)
    assert (t, it) == (1.25, 40) and np.array_equal(P, Q)


def test_user_scheduler_and_particle_groups(tmp_path):
    """user/<name>, scheduler/<task>/{nextPhysTime, nextWallTime,
    nextIteration, repetition} and particles/data round-trip."""
    desc = problems.sod_desc(32)
    P = problems.sod_1d(desc)
    path = tmp_path / "g.npz"
    user = {"tfinal": 0.2, "problem": "sod", "cpi": 0.05}
    sched = {"checkpoint": {"nextPhysTime": 0.25, "nextWallTime": 0.0, "nextIteration": 0, "repetition": 5}}
    checkpoint.write(path, desc, P, 0.2, 11, user=user, scheduler=sched, particles=[0.1, 0.2, 0.3])
    groups = checkpoint.read_groups(path)
    assert groups["user"] == user
    assert groups["scheduler"] == sched
    assert np.array_equal(groups["particles"], [0.1, 0.2, 0.3])


def test_wrong_mesh_is_refused(tmp_path):
    desc = problems.sod_desc(32)
    path = tmp_path / "w.npz"
    checkpoint.write(path, desc, problems.sod_1d(desc), 0.0, 0)
    import pytest
    with pytest.raises(ValueError):
        checkpoint.read(path, problems.sod_desc(64))
