"""Mesh topology and H1 dof tables built on the device (ngsamd_h1_create)
against the host setup (Mesh::BuildTopology, H1Space) and the oracle's
topology / h1_dofs / class_number: edge and face numbers, element and
boundary-facet edge / face numbers, element dof tables, free-dof masks and
element classes, all bit-exact; structured 2D and 3D meshes (in 3D also with
the boundary found on the device as the faces of one element), the
reference's square.vol.gz / cube.vol.gz, orders 1 to 5, Dirichlet on every,
one and no region; bad input refused."""
import gzip
import os

import numpy as np
import pytest

from oracle import fe_oracle as O
from helpers import N, say

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALL = (1 << 64) - 1


def _mesh(kind):
    if kind == "2d":
        return N.MakeStructured2DMesh(7, 5), O.structured_2d(7, 5)
    if kind == "3d":
        return N.MakeStructured3DMesh(3, 4, 2), O.structured_3d(3, 4, 2)
    text = gzip.open(os.path.join(GOLDEN, kind), "rt").read()
    return N.ReadVolMesh(text), O.read_vol(text)


def _device(m, p, mask, bnd=None, bnd_index=None):
    bnd = m.boundary if bnd is None else bnd
    bi = m.boundary_index if bnd_index is None else bnd_index
    return N.H1TablesDevice(m.dim, m.nv, np.ascontiguousarray(m.elements, dtype=np.int32),
                            np.ascontiguousarray(bnd, dtype=np.int32).reshape(-1, m.dim),
                            np.ascontiguousarray(bi, dtype=np.int32), p, mask)


def _check_against_host(D, m, fes):
    assert D["ndof"] == fes.ndof
    assert np.array_equal(D["edges"], m.edges)
    assert np.array_equal(D["element_edges"], m.element_edges)
    assert np.array_equal(D["boundary"], m.boundary)
    assert np.array_equal(D["boundary_index"], m.boundary_index)
    assert np.array_equal(D["boundary_edges"], m.boundary_edges)
    if m.dim == 3:
        assert np.array_equal(D["faces"], m.faces)
        assert np.array_equal(D["element_faces"], m.element_faces)
        assert np.array_equal(D["boundary_faces"], m.boundary_faces)
    assert np.array_equal(D["el_dofs"], fes.el_dofs)
    assert np.array_equal(D["freedofs"], fes.freedofs)
    assert np.array_equal(D["classnr"], fes.classnr)


MASKS = {"all": ALL, "region1": 1, "none": 0}


@pytest.mark.parametrize("kind", ["2d", "3d", "square.vol.gz", "cube.vol.gz"])
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("mask", list(MASKS))
def test_h1_tables_match_host_and_oracle(device, kind, p, mask):
    say(f"[h1 tables] {kind} p={p} dirichlet={mask}")
    m, om = _mesh(kind)
    bits = MASKS[mask]
    fes = N.H1Space(m, p, bits)
    D = _device(m, p, bits)
    _check_against_host(D, m, fes)
    # the oracle's own restatement (order-free quantities)
    topo = O.topology(om)
    assert np.array_equal(D["edges"], topo["edges"])
    assert np.array_equal(D["element_edges"], topo["el_edges"])
    if m.dim == 3:
        assert np.array_equal(D["faces"], topo["faces"])
        assert np.array_equal(D["element_faces"], topo["el_faces"])
    regions = {"all": None, "region1": [1], "none": []}[mask]
    dofs, ndof, free = O.h1_dofs(om, topo, p, regions)
    assert D["ndof"] == ndof
    assert np.array_equal(D["el_dofs"], dofs)
    assert np.array_equal(D["freedofs"], free)
    assert np.array_equal(D["classnr"], O.class_number(m.dim, om["els"]))


@pytest.mark.parametrize("kind", ["3d", "cube.vol.gz"])
def test_boundary_found_on_device(device, kind):
    """In 3D with no boundary facets given, the faces held by one element
    are the boundary (region 1, face order)."""
    say(f"[h1 tables] {kind} boundary from the faces of one element")
    m, om = _mesh(kind)
    D = _device(m, 3, ALL, bnd=np.zeros((0, 3), np.int32), bnd_index=np.zeros(0, np.int32))
    fv = np.sort(om["els"][:, [[3, 1, 2], [3, 2, 0], [3, 0, 1], [0, 2, 1]]].reshape(-1, 3), axis=1)
    uf, cnt = np.unique(fv, axis=0, return_counts=True)
    assert np.array_equal(D["boundary"], uf[cnt == 1])
    assert np.all(D["boundary_index"] == 1)
    assert np.array_equal(D["faces"][D["boundary_faces"]], uf[cnt == 1])
    # the same set of boundary facets as the mesh's own
    assert np.array_equal(np.unique(np.sort(m.boundary, axis=1), axis=0), uf[cnt == 1])
    if kind == "3d":  # the host setup finds it the same way
        _check_against_host(D, m, N.H1Space(m, 3, ALL))


def test_h1_tables_large_3d(device):
    """The 32^3 cube x 6 tets at p=4 (2,146,689 dofs), device against host."""
    say("[h1 tables] 32^3 p=4")
    m = N.MakeStructured3DMesh(32, 32, 32)
    fes = N.H1Space(m, 4, ALL)
    D = _device(m, 4, ALL, bnd=np.zeros((0, 3), np.int32), bnd_index=np.zeros(0, np.int32))
    assert D["ndof"] == 2146689
    _check_against_host(D, m, fes)


@pytest.mark.parametrize("nfan", [8, 40, 1000])
def test_fan_mesh_large_bucket(device, nfan):
    """A disc of nfan triangles around vertex 0: its bucket holds 2 nfan
    edge entries (heap-sorted above 64), device against host."""
    say(f"[h1 tables] fan of {nfan} triangles")
    i = np.arange(1, nfan + 1)
    els = np.stack([np.zeros(nfan, dtype=np.int32), i, i % nfan + 1], axis=1).astype(np.int32)
    bnd = np.stack([i, i % nfan + 1], axis=1).astype(np.int32)
    bi = np.ones(nfan, dtype=np.int32)
    m = N.BuildTopologyHost(2, nfan + 1, els, bnd, bi)
    for p in (1, 3):
        D = N.H1TablesDevice(2, nfan + 1, els, bnd, bi, p, ALL)
        _check_against_host(D, m, N.H1Space(m, p, ALL))


def test_bad_input_refused(device):
    say("[h1 tables] bad input")
    m, _ = _mesh("2d")
    els = np.ascontiguousarray(m.elements, dtype=np.int32).copy()
    els[3, 1] = m.nv
    with pytest.raises(RuntimeError, match="out of range"):
        N.H1TablesDevice(2, m.nv, els, np.ascontiguousarray(m.boundary, np.int32),
                         np.ascontiguousarray(m.boundary_index, np.int32), 2, ALL)
    bnd = np.ascontiguousarray(m.boundary, np.int32).copy()
    bnd[0] = [0, m.nv - 1]  # not an edge of the mesh
    with pytest.raises(RuntimeError, match="not a facet"):
        N.H1TablesDevice(2, m.nv, np.ascontiguousarray(m.elements, np.int32), bnd,
                         np.ascontiguousarray(m.boundary_index, np.int32), 2, ALL)
    with pytest.raises(RuntimeError, match="order"):
        _device(m, 0, ALL)
