"""MI355X (gfx950) Godunov finite-volume step of Mara2, behind a C ABI.

    from mara2_amd import abi, runtime, problems
    desc = abi.make_desc(law="hydro", cells=(256, 256, 256), riemann="hllc")
    with runtime.Block(desc) as blk:
        blk.upload(problems.turbulence_3d(desc))
        blk.run(cfl=0.3, t_final=1e30, max_steps=100)

Importing this package does not load the HIP library; runtime.load() does,
and raises if it was not built.
"""
from . import abi  # noqa: F401

__all__ = ["abi", "runtime", "problems"]
