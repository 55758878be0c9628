"""Structured meshes with the reference's generation rules
(python/meshes.py:47-241 and :273-493), triangles / Kuhn tetrahedra only."""
from ._native import N
from .comp import Mesh


def MakeStructured2DMesh(quads=False, nx=10, ny=10, **kwargs):
    if quads:
        raise NotImplementedError("only the triangular structured mesh (quads=False) is supported")
    for k in ("flip_triangles", "periodic_x", "periodic_y", "mapping", "secondorder", "boundarylayer"):
        if kwargs.get(k):
            raise NotImplementedError(f"MakeStructured2DMesh: option {k} is not supported")
    return Mesh(N.MakeStructured2DMesh(int(nx), int(ny)))


def MakeStructured3DMesh(hexes=False, nx=10, ny=None, nz=None, **kwargs):
    if hexes:
        raise NotImplementedError("only the tetrahedral structured mesh (hexes=False) is supported")
    for k in ("prism", "periodic_x", "periodic_y", "periodic_z", "mapping", "secondorder"):
        if kwargs.get(k):
            raise NotImplementedError(f"MakeStructured3DMesh: option {k} is not supported")
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return Mesh(N.MakeStructured3DMesh(int(nx), int(ny), int(nz)))
