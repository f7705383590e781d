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
