"""Checkpoints with the reference's content and names, as NPZ archives.

CheckpointWriter::writeCheckpoint (src/Checkpoint.cpp:76-178) writes an HDF5
file with groups status, scheduler, time_series, user, primitive,
diagnostic, mesh and particles.  HDF5 is not available here, so the same
datasets are stored in one .npz whose keys are the HDF5 paths:

  run_name                    "Mara"
  status/simulationTime       SimulationStatus::pack (src/Mara.cpp:146-152)
  status/simulationIter
  user/<name>                 CheckpointWriter::setUserParameters, one scalar
                              or string per named value (src/Checkpoint.cpp:156-161)
  scheduler/<task>/nextPhysTime, nextWallTime, nextIteration, repetition
                              TaskScheduler::write for each named task
                              (src/TaskScheduler.cpp:131-147)
  particles/data              MeshData::starParticles (src/Checkpoint.cpp:144-146)
  mesh/type                   "cartesian"
  mesh/points/x, y, z         CartesianMeshGeometry::getPointCoordinates
                              (src/CartesianMeshGeometry.cpp:121-131): the
                              n + 1 face coordinates of each axis
  primitive/<name>            one dataset per field of the law's state,
                              interior zones only, shape (n1, n2, n3), the
                              law's primitive names (src/ConservationLaws.cpp:
                              308-319, 520-534), "auxiliary1" for the cooled
                              NewtonianMHD's ninth field
  diagnostic/<name>           MeshData diagnostics, same shape

Reading (CheckpointWriter::readCheckpoint, src/Checkpoint.cpp:237-273)
fills the interior of MeshData::P and then the guard zones from the boundary
condition; read_groups returns the user, scheduler and particle data.
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

SCHEDULER_KEYS = ("nextPhysTime", "nextWallTime", "nextIteration", "repetition")


def primitive_names(desc):
    """NewtonianMHD::getPrimitiveName for every field of the descriptor's
    state: the law's names, then "auxiliary<n>" for field 7 + n."""
    names = list(PRIMITIVE_NAMES[desc.law])
    for q in range(len(names), abi.desc_num_fields(desc)):
        names.append("auxiliary" + str(q - 7))
    return names


def point_coordinates(desc, axis):
    n = desc.cells[axis]
    dx = (desc.upper[axis] - desc.lower[axis]) / n
    # coordinateAtIndex(i - 0.5): lower + dx * (i - 0.5 + 0.5)
    return desc.lower[axis] + dx * ((np.arange(n + 1) - 0.5) + 0.5)


def _interior(desc):
    g, n = desc.guards, desc.cells
    return (slice(g[0], g[0] + n[0]), slice(g[1], g[1] + n[1]), slice(g[2], g[2] + n[2]))


def write(path, desc, P, time, iteration, user=None, scheduler=None, particles=None, diagnostics=None):
    """Store MeshData::P (guards included, reference layout) with the clock.
    user: {name: number or str}; scheduler: {task: {nextPhysTime,
    nextWallTime, nextIteration, repetition}}; particles: sequence of
    doubles; diagnostics: {name: array shaped like P[..., 0] with guards}."""
    if P.shape != abi.array_shape(desc):
        raise ValueError(f"P has shape {P.shape}, descriptor wants {abi.array_shape(desc)}")
    inner = _interior(desc)
    data = {"run_name": np.array("Mara"),
            "status/simulationTime": np.array(float(time)),
            "status/simulationIter": np.array(int(iteration)),
            "mesh/type": np.array("cartesian"),
            "particles/data": np.asarray(particles if particles is not None else [], dtype=np.float64)}
    for name, value in (user or {}).items():
        data[f"user/{name}"] = np.array(value)
    for task, rec in (scheduler or {}).items():
        for key in SCHEDULER_KEYS:
            cast = int if key in ("nextIteration", "repetition") else float
            data[f"scheduler/{task}/{key}"] = np.array(cast(rec[key]))
    for a, name in enumerate("xyz"):
        data[f"mesh/points/{name}"] = point_coordinates(desc, a)
    for q, name in enumerate(primitive_names(desc)):
        data[f"primitive/{name}"] = np.ascontiguousarray(P[inner + (q,)])
    for name, arr in (diagnostics or {}).items():
        data[f"diagnostic/{name}"] = np.ascontiguousarray(np.asarray(arr)[inner])
    np.savez(path, **data)


def read(path, desc):
    """(P with guards filled from the boundary condition, time, iteration)."""
    with np.load(path, allow_pickle=False) as f:
        P = problems.empty(desc)
        inner = _interior(desc)
        for q, name in enumerate(primitive_names(desc)):
            key = f"primitive/{name}"
            if key not in f.files:
                raise KeyError(f"checkpoint has no dataset {key}")
            arr = f[key]
            if arr.shape != tuple(desc.cells):
                raise ValueError(f"checkpoint field {name} has shape {arr.shape}, mesh has {tuple(desc.cells)}")
            P[inner + (q,)] = arr
        t = float(f["status/simulationTime"])
        it = int(f["status/simulationIter"])
    return problems.apply_bc(desc, P), t, it


def read_groups(path):
    """The user parameters, the scheduler's task records and the star
    particle data of a checkpoint: {"user": {...}, "scheduler": {task: {...}},
    "particles": array}."""
    out = {"user": {}, "scheduler": {}, "particles": np.zeros(0)}
    with np.load(path, allow_pickle=False) as f:
        for key in f.files:
            parts = key.split("/")
            if parts[0] == "user":
                v = f[key]
                out["user"][parts[1]] = v.item() if v.dtype.kind != "U" else str(v)
            elif parts[0] == "scheduler":
                out["scheduler"].setdefault(parts[1], {})[parts[2]] = f[key].item()
            elif key == "particles/data":
                out["particles"] = np.array(f[key])
    return out
