"""Checkpoints with the reference's content and names, as NPZ archives.

CheckpointWriter::writeCheckpoint (src/Checkpoint.cpp:76-178) writes an HDF5
file with groups status, primitive, mesh (and scheduler, time_series, user,
diagnostic, particles, which the step does not touch).  HDF5 is not
available here, so the same datasets are stored in one .npz whose keys are
the HDF5 paths:

  run_name                    "Mara"
  status/simulationTime       SimulationStatus::pack (src/Mara.cpp:146-152)
  status/simulationIter
  mesh/type                   "cartesian"
  mesh/points/x, y, z         CartesianMeshGeometry::getPointCoordinates
                              (src/CartesianMeshGeometry.cpp:121-131): the
                              n + 1 face coordinates of each axis
  primitive/<name>            one dataset per primitive, interior zones only,
                              shape (n1, n2, n3), the law's primitive names
                              (src/ConservationLaws.cpp:308-319, 520-534)

Reading (CheckpointWriter::readCheckpoint, src/Checkpoint.cpp:178-215)
fills the interior of MeshData::P and then the guard zones from the boundary
condition.
"""
import numpy as np

from . import abi, problems

PRIMITIVE_NAMES = {
    abi.LAW_HYDRO: ["density", "velocity1", "velocity2", "velocity3", "pressure"],
    abi.LAW_MHD: ["density", "velocity1", "velocity2", "velocity3", "pressure",
                  "magnetic1", "magnetic2", "magnetic3"],
    abi.LAW_RMHD: ["density", "velocity1", "velocity2", "velocity3", "pressure",
                   "magnetic1", "magnetic2", "magnetic3"],
}


def point_coordinates(desc, axis):
    n = desc.cells[axis]
    dx = (desc.upper[axis] - desc.lower[axis]) / n
    # coordinateAtIndex(i - 0.5): lower + dx * (i - 0.5 + 0.5)
    return desc.lower[axis] + dx * ((np.arange(n + 1) - 0.5) + 0.5)


def write(path, desc, P, time, iteration):
    """Store MeshData::P (guards included, reference layout) with the clock."""
    g = desc.guards
    n = desc.cells
    inner = (slice(g[0], g[0] + n[0]), slice(g[1], g[1] + n[1]), slice(g[2], g[2] + n[2]))
    data = {"run_name": np.array("Mara"),
            "status/simulationTime": np.array(float(time)),
            "status/simulationIter": np.array(int(iteration)),
            "mesh/type": np.array("cartesian")}
    for a, name in enumerate("xyz"):
        data[f"mesh/points/{name}"] = point_coordinates(desc, a)
    for q, name in enumerate(PRIMITIVE_NAMES[desc.law]):
        data[f"primitive/{name}"] = np.ascontiguousarray(P[inner + (q,)])
    np.savez(path, **data)


def read(path, desc):
    """(P with guards filled from the boundary condition, time, iteration)."""
    with np.load(path, allow_pickle=False) as f:
        P = problems.empty(desc)
        g = desc.guards
        n = desc.cells
        inner = (slice(g[0], g[0] + n[0]), slice(g[1], g[1] + n[1]), slice(g[2], g[2] + n[2]))
        for q, name in enumerate(PRIMITIVE_NAMES[desc.law]):
            arr = f[f"primitive/{name}"]
            if arr.shape != tuple(n):
                raise ValueError(f"checkpoint field {name} has shape {arr.shape}, mesh has {tuple(n)}")
            P[inner + (q,)] = arr
        t = float(f["status/simulationTime"])
        it = int(f["status/simulationIter"])
    return problems.apply_bc(desc, P), t, it
