"""underworld2_amd -- an MI355X (gfx950) particle-in-cell Stokes engine with the
``underworld`` (v2.16) API for its hot path:

    import underworld2_amd as uw
    mesh  = uw.mesh.FeMesh_Cartesian("Q1/dQ0", (64, 64), (0., 0.), (1., 1.))
    vel   = uw.mesh.MeshVariable(mesh, 2);  p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, 20, seed=0))
    stokes = uw.systems.Stokes(vel, p, fn_viscosity, fn_bodyforce, voronoi_swarm=swarm,
                               conditions=[uw.conditions.DirichletCondition(vel, (iw, jw))])
    uw.systems.Solver(stokes).solve()
    uw.systems.SwarmAdvector(vel, swarm, order=2).integrate(dt)

All numerical work runs in the HIP library ``libuwg.so`` (C ABI: include/uwg.h).
"""
from . import _lib
from . import function
from . import mesh
from . import swarm
from . import conditions
from . import systems

__version__ = "0.1.0"


def device_info():
    return _lib.device_info()
