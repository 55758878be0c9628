"""CPU oracle (test infrastructure only) for the H1 Laplace hot path.

This module restates, in numpy, the reference's algorithms for:
  structured meshes        python/meshes.py:137-195 (2D trigs), :351-410 (3D Kuhn tets)
  affine element map       comp/meshaccess.cpp:541-562
  H1 dof numbering         comp/h1hofespace.cpp:823-950, :1389-1413
  free dofs                comp/fespace.cpp:765-810
  element colouring        comp/fespace.cpp:518-760 (sequential order)
  shape functions          fem/h1hofe_impl.hpp:154-185, :510-570 with the
                           polynomials of fem/recursive_pol.hpp:1344-1373,
                           1641-1676, 1857-1940, 2820-2861, 2974-3005
  element classes          fem/elementtopology.hpp:999-1007, :1241-1252
  integration rules        fem/intrule.cpp:703-900, 2490-2571, 3143-3232
  element matrices         fem/symbolicintegrator.cpp:1549-1860 (grad.grad),
                           DiffOpId for mass
  CSR pattern + assembly   comp/bilinearform.cpp:552-825, linalg/sparsematrix.cpp:450-975,
                           linalg/sparsematrix_impl.hpp:150-189
  linear form              fem/symbolicintegrator.cpp:1036-1110
  Jacobi-PCG               python/krylovspace.py:263-291 (Jacobi or any
                           preconditioner)
  block Jacobi             linalg/blockjacobi.cpp:390, 404, 498, 596-632
  sparse direct inverse    linalg/sparsecholesky.cpp (a.mat.Inverse on free dofs)
  two-level preconditioner smoother + E C E^T (linalg/basematrix.hpp:413,592)
  multigrid V-cycle        multigrid/mgpre.cpp:205-318 with LinearProlongation
                           (multigrid/prolongation.cpp:77-144) on the
                           structured vertex hierarchy, Galerkin coarse levels
  lockstep multi-RHS CG    linalg/cg.cpp:132-250 (CGSolver::MultiMult)

It is deliberately written differently from the product code: shapes are
evaluated per element (vectorised over elements and points, orientation per
element, no class caching) with dual numbers, Gauss-Jacobi rules come from
the eigenvalues of the Jacobi matrix (numpy.linalg.eigh), element matrices
use F^{-T} per element and quadrature point, and the CSR pattern comes from
numpy.unique of all (row, col) pairs.

Global edge/face numbering: ascending lexicographic order of the sorted
vertex tuples (netgen's topology is not available; this is the documented
rule the product also follows).  Parity of that numbering against NGSolve
itself is unpinned.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, as the checker.
"""
import math

import numpy as np

# ----------------------------------------------------------------- meshes


def structured_2d(nx, ny=None):
    ny = nx if ny is None else ny
    i, j = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coords = np.stack([j.ravel() / nx, i.ravel() / ny], axis=1).astype(float)
    ii, jj = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    b = (ii * (nx + 1) + jj).ravel()
    t1 = np.stack([b, b + 1, b + nx + 1], axis=1)
    t2 = np.stack([b + 1, b + nx + 2, b + nx + 1], axis=1)
    els = np.empty((2 * len(b), 3), dtype=np.int64)
    els[0::2] = t1
    els[1::2] = t2
    bnd = []
    idx = []
    for k in range(nx):
        bnd.append((k, k + 1)); idx.append(1)
    for k in range(ny):
        bnd.append((k * (nx + 1) + nx, (k + 1) * (nx + 1) + nx)); idx.append(2)
    for k in range(nx):
        bnd.append((ny * (nx + 1) + k + 1, ny * (nx + 1) + k)); idx.append(3)
    for k in range(ny):
        bnd.append(((k + 1) * (nx + 1), k * (nx + 1))); idx.append(4)
    return dict(dim=2, coords=coords, els=els, bnd=np.array(bnd, dtype=np.int64),
                bnd_index=np.array(idx, dtype=np.int64))


KUHN = [[0, 4, 5, 6], [0, 6, 7, 4], [0, 3, 7, 6], [0, 1, 6, 5], [0, 1, 2, 6], [0, 3, 6, 2]]


def structured_3d(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    coords = np.stack([I.ravel() / nx, J.ravel() / ny, K.ravel() / nz], axis=1).astype(float)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = (ii * (ny + 1) * (nz + 1) + jj * (nz + 1) + kk).ravel()
    up = base + (ny + 1) * (nz + 1)
    p = np.stack([base, base + 1, base + nz + 2, base + nz + 1, up, up + 1, up + nz + 2, up + nz + 1], axis=1)
    els = np.stack([p[:, q] for q in KUHN], axis=1).reshape(-1, 4)
    # boundary: faces owned by a single element
    faces = np.sort(els[:, [[3, 1, 2], [3, 2, 0], [3, 0, 1], [0, 2, 1]]].reshape(-1, 3), axis=1)
    uf, cnt = np.unique(faces, axis=0, return_counts=True)
    bnd = uf[cnt == 1]
    return dict(dim=3, coords=coords, els=els.astype(np.int64), bnd=bnd.astype(np.int64),
                bnd_index=np.ones(len(bnd), dtype=np.int64))


def read_vol(text):
    """Netgen .vol text: 2D triangles in 'surfaceelements', boundary in
    'edgesegmentsgi2'; 3D tets in 'volumeelements', boundary in 'surfaceelements'."""
    lines = text.splitlines()
    dim, pts, surf, surf_bc, vol, segs, seg_bc = 3, [], [], [], [], [], []
    i = 0
    while i < len(lines):
        tok = lines[i].strip()
        if tok == "dimension":
            dim = int(lines[i + 1]); i += 2; continue
        if tok in ("surfaceelements", "surfaceelementsgi", "surfaceelementsuv"):
            n = int(lines[i + 1])
            for r in range(n):
                f = lines[i + 2 + r].split()
                np_ = int(f[4])
                surf.append([int(v) - 1 for v in f[5:5 + np_]])
                surf_bc.append(int(f[1]))
            i += 2 + n; continue
        if tok == "volumeelements":
            n = int(lines[i + 1])
            for r in range(n):
                f = lines[i + 2 + r].split()
                np_ = int(f[1])
                vol.append([int(v) - 1 for v in f[2:2 + np_]])
            i += 2 + n; continue
        if tok in ("edgesegmentsgi2", "edgesegmentsgi", "edgesegments"):
            n = int(lines[i + 1])
            for r in range(n):
                f = lines[i + 2 + r].split()
                segs.append([int(f[2]) - 1, int(f[3]) - 1])
                seg_bc.append(int(f[0]))
            i += 2 + n; continue
        if tok == "points":
            n = int(lines[i + 1])
            pts = [[float(v) for v in lines[i + 2 + r].split()] for r in range(n)]
            i += 2 + n; continue
        if tok == "endmesh":
            break
        i += 1
    coords = np.array(pts)[:, :dim]
    if dim == 2:
        return dict(dim=2, coords=coords, els=np.array(surf, dtype=np.int64),
                    bnd=np.array(segs, dtype=np.int64), bnd_index=np.array(seg_bc, dtype=np.int64))
    return dict(dim=3, coords=coords, els=np.array(vol, dtype=np.int64),
                bnd=np.array(surf, dtype=np.int64), bnd_index=np.array(surf_bc, dtype=np.int64))


# -------------------------------------------------------------- topology

TRIG_EDGES = [(2, 0), (1, 2), (0, 1)]
TET_EDGES = [(3, 0), (3, 1), (3, 2), (0, 1), (0, 2), (1, 2)]
TET_FACES = [(3, 1, 2), (3, 2, 0), (3, 0, 1), (0, 2, 1)]


def _row_lookup(table, keys):
    """Index of each row of keys in the lexicographically sorted unique table."""
    nv = int(max(table.max(), keys.max())) + 1
    if table.shape[1] == 2:
        tk = table[:, 0] * nv + table[:, 1]
        kk = keys[:, 0] * nv + keys[:, 1]
    else:
        tk = (table[:, 0] * nv + table[:, 1]) * nv + table[:, 2]
        kk = (keys[:, 0] * nv + keys[:, 1]) * nv + keys[:, 2]
    pos = np.searchsorted(tk, kk)
    assert np.all(tk[pos] == kk)
    return pos


def topology(mesh):
    els, dim = mesh["els"], mesh["dim"]
    le = TRIG_EDGES if dim == 2 else TET_EDGES
    ev = np.sort(np.stack([els[:, list(e)] for e in le], axis=1), axis=2)  # ne x nle x 2
    edges = np.unique(ev.reshape(-1, 2), axis=0)
    el_edges = _row_lookup(edges, ev.reshape(-1, 2)).reshape(len(els), len(le))
    topo = dict(edges=edges, el_edges=el_edges)
    bnd = mesh["bnd"]
    if dim == 2:
        topo["bnd_edges"] = _row_lookup(edges, np.sort(bnd, axis=1)).reshape(-1, 1)
    else:
        fv = np.sort(np.stack([els[:, list(f)] for f in TET_FACES], axis=1), axis=2)
        faces = np.unique(fv.reshape(-1, 3), axis=0)
        topo["faces"] = faces
        topo["el_faces"] = _row_lookup(faces, fv.reshape(-1, 3)).reshape(len(els), 4)
        sb = np.sort(bnd, axis=1)
        topo["bnd_faces"] = _row_lookup(faces, sb)
        be = np.stack([sb[:, [0, 1]], sb[:, [1, 2]], sb[:, [0, 2]]], axis=1).reshape(-1, 2)
        topo["bnd_edges"] = _row_lookup(edges, be).reshape(-1, 3)
    return topo


def h1_dofs(mesh, topo, p, dirichlet_regions=None):
    """Element dof table (reference GetDofNrs order), ndof and free-dof mask."""
    dim, els = mesh["dim"], mesh["els"]
    nv, ne, ned = len(mesh["coords"]), len(els), len(topo["edges"])
    pe = p - 1
    pf = (p - 1) * (p - 2) // 2 if (dim == 3 and p > 2) else 0
    pc = ((p - 1) * (p - 2) // 2 if p > 2 else 0) if dim == 2 else ((p - 1) * (p - 2) * (p - 3) // 6 if p > 3 else 0)
    nfa = len(topo["faces"]) if dim == 3 else 0
    e0 = nv
    f0 = e0 + ned * pe
    c0 = f0 + nfa * pf
    ndof = c0 + ne * pc
    parts = [els]
    if pe:
        parts.append((e0 + topo["el_edges"][:, :, None] * pe + np.arange(pe)).reshape(ne, -1))
    if dim == 3 and pf:
        parts.append((f0 + topo["el_faces"][:, :, None] * pf + np.arange(pf)).reshape(ne, -1))
    if pc:
        parts.append(c0 + np.arange(ne)[:, None] * pc + np.arange(pc))
    dofs = np.concatenate(parts, axis=1)
    free = np.ones(ndof, dtype=np.uint8)
    sel = np.ones(len(mesh["bnd"]), dtype=bool)
    if dirichlet_regions is not None:
        sel = np.isin(mesh["bnd_index"], list(dirichlet_regions))
    bnd = mesh["bnd"][sel]
    free[bnd.ravel()] = 0
    if pe:
        be = topo["bnd_edges"][sel].ravel()
        free[(e0 + be[:, None] * pe + np.arange(pe)).ravel()] = 0
    if dim == 3 and pf:
        bf = topo["bnd_faces"][sel]
        free[(f0 + bf[:, None] * pf + np.arange(pf)).ravel()] = 0
    return dofs, ndof, free


# ------------------------------------------------------------ dual numbers

class Dual:
    """Value and gradient, vectorised: val shape S, grad shape (D,)+S."""
    __slots__ = ("val", "grad")

    def __init__(self, val, grad):
        self.val = val
        self.grad = grad

    @staticmethod
    def const(c, like):
        c = np.broadcast_to(np.asarray(c, dtype=float), like.val.shape)
        return Dual(c.copy(), np.zeros_like(like.grad))

    def _lift(self, o):
        return o if isinstance(o, Dual) else Dual.const(o, self)

    def __add__(self, o):
        o = self._lift(o)
        return Dual(self.val + o.val, self.grad + o.grad)

    __radd__ = __add__

    def __sub__(self, o):
        o = self._lift(o)
        return Dual(self.val - o.val, self.grad - o.grad)

    def __rsub__(self, o):
        return self._lift(o) - self

    def __neg__(self):
        return Dual(-self.val, -self.grad)

    def __mul__(self, o):
        if not isinstance(o, Dual):
            return Dual(self.val * o, self.grad * o)
        return Dual(self.val * o.val, self.grad * o.val + self.val * o.grad)

    __rmul__ = __mul__


# ------------------------------------------------- one-dimensional families

def intleg_nobubble_scaled(n, x, y, c):
    """IntLegNoBubble, scaled: P0=-1/2, P1=-x/2,
    P_i = (2i+1)/(i+2) x P_{i-1} - (i-1)/(i+2) y^2 P_{i-2}."""
    out = [c * (-0.5)]
    if n >= 1:
        out.append(c * (x * -0.5))
    for i in range(2, n + 1):
        a = (2.0 * i + 1) / (i + 2.0)
        cc = -(i - 1.0) / (i + 2.0)
        out.append(x * out[-1] * a + y * y * out[-2] * cc)
    return out[: n + 1]


def jacobi11_scaled(n, x, y, c):
    """Scaled Jacobi P^(1,1): P0=1, P1=2x, P_j+1 from the standard recurrence."""
    out = [c * 1.0]
    if n >= 1:
        out.append(c * (x * 2.0))
    for i in range(2, n + 1):
        j = i - 1.0
        a = (2 * j + 2) * (2 * j + 3) * (2 * j + 4) / (2 * (j + 1) * (j + 3) * (2 * j + 2))
        cc = -2.0 * (j + 1) * (j + 1) * (2 * j + 4) / (2 * (j + 1) * (j + 3) * (2 * j + 2))
        out.append(x * out[-1] * a + y * y * out[-2] * cc)
    return out[: n + 1]


def _intjac_coef(i, al):
    den = (2.0 * i + 2) * (i + al) * (2.0 * i + al - 2)
    return ((2.0 * i + al - 1) * (2.0 * i + al - 2) * (2.0 * i + al) / den,
            (2.0 * i + al - 1) * al * (al - 2) / den,
            -2.0 * (i - 1) * (i + al - 2) * (2.0 * i + al) / den)


def intjac_alpha(n, al, x, c, y=None):
    """Integrated Jacobi (1/(1+x)) int_{-1}^x P^(al,0); scaled by y if given."""
    out = [c * 1.0]
    if n >= 1:
        if y is None:
            out.append(c * (x * (0.25 * (al + 2)) + 0.25 * (al - 2)))
        else:
            out.append(c * (x * (0.25 * (al + 2)) + y * (0.25 * (al - 2))))
    for i in range(2, n + 1):
        a, b, cc = _intjac_coef(i, al)
        if y is None:
            out.append((x * a + b) * out[-1] + out[-2] * cc)
        else:
            out.append((x * a + y * b) * out[-1] + y * y * out[-2] * cc)
    return out[: n + 1]


def dubiner_orthobub(n, x, y, c):
    res = []
    leg = jacobi11_scaled(n, y - (1.0 - x - y), 1.0 - x, c)
    for i in range(n + 1):
        res += intjac_alpha(n - i, 4 + 2 * i, 2.0 * x - 1.0, leg[i])
    return res


def dubiner_orthobub_scaled(n, x, y, t, c):
    res = []
    leg = jacobi11_scaled(n, y - (t - x - y), t - x, c)
    for i in range(n + 1):
        res += intjac_alpha(n - i, 4 + 2 * i, 2.0 * x - t, leg[i], y=t)
    return res


def dubiner3d_orthobub(n, x, y, z, c):
    res = []
    lam4 = 1.0 - x - y - z
    polz = jacobi11_scaled(n, z - lam4, z + lam4, c)
    for k in range(n + 1):
        polsy = intjac_alpha(n - k, 4 + 2 * k, y - z - lam4, polz[k], y=1.0 - x)
        for j in range(n - k + 1):
            res += intjac_alpha(n - k - j, 6 + 2 * k + 2 * j, 2.0 * x - 1.0, polsy[j])
    return res


# ------------------------------------------------------ element shapes

def _select(lams, idx):
    """lams: list of Dual with val shape (E, Q); idx: (E,) local index per element."""
    val = np.stack([l.val for l in lams], axis=0)
    grad = np.stack([l.grad for l in lams], axis=0)
    E = val.shape[1]
    ar = np.arange(E)
    return Dual(val[idx, ar], np.moveaxis(grad[idx, :, ar], 0, 1))


def element_shapes(dim, p, vnums, pts):
    """Shapes and reference gradients of every element at every point.
    vnums: (E, dim+1) global vertex numbers; pts: (Q, dim).
    Returns val (E, Q, nd) and grad (E, Q, nd, dim)."""
    E, Q = len(vnums), len(pts)
    coords = []
    for d in range(dim):
        g = np.zeros((dim, E, Q))
        g[d] = 1.0
        coords.append(Dual(np.broadcast_to(pts[:, d], (E, Q)).copy(), g))
    lam = coords + [1.0 - sum(coords[1:], coords[0])]
    shapes = list(lam)
    ar = np.arange(E)
    if p >= 2:
        for (a, b) in (TRIG_EDGES if dim == 2 else TET_EDGES):
            swap = vnums[:, a] > vnums[:, b]
            i0 = np.where(swap, b, a)
            i1 = np.where(swap, a, b)
            l0, l1 = _select(lam, i0), _select(lam, i1)
            shapes += intleg_nobubble_scaled(p - 2, l1 - l0, l0 + l1, l0 * l1)
    if dim == 2 and p >= 3:
        order = np.argsort(vnums, axis=1, kind="stable")
        f0, f1, f2 = (_select(lam, order[:, k]) for k in range(3))
        shapes += dubiner_orthobub(p - 3, f0, f1, f0 * f1 * f2)
    if dim == 3 and p >= 3:
        for face in TET_FACES:
            fv = vnums[:, list(face)]
            order = np.argsort(fv, axis=1, kind="stable")
            loc = np.array(face)[order]
            f0, f1, f2 = (_select(lam, loc[:, k]) for k in range(3))
            vop = _select(lam, 6 - loc.sum(axis=1))
            shapes += dubiner_orthobub_scaled(p - 3, f0, f1, 1.0 - vop, f0 * f1 * f2)
    if dim == 3 and p >= 4:
        shapes += dubiner3d_orthobub(p - 4, lam[0], lam[1], lam[2], lam[0] * lam[1] * lam[2] * lam[3])
    val = np.stack([s.val for s in shapes], axis=2)
    grad = np.stack([np.moveaxis(s.grad, 0, -1) for s in shapes], axis=2)
    del ar
    return val, grad


def class_number(dim, v):
    """Vertex-permutation class of each element (comparator networks of
    fem/elementtopology.hpp:999-1007 and :1241-1252)."""
    v = np.array(v, dtype=np.int64, copy=True)
    E = len(v)
    s = np.tile(np.arange(dim + 1), (E, 1))
    c = np.zeros(E, dtype=np.int64)
    ar = np.arange(E)
    net = [(0, 1, 1), (1, 2, 2), (0, 1, 2)] if dim == 2 else [(0, 1, 1), (2, 3, 2), (0, 2, 4), (1, 3, 8), (1, 2, 16)]
    for a, b, w in net:
        sw = v[ar, s[:, a]] > v[ar, s[:, b]]
        sa, sb = s[:, a].copy(), s[:, b].copy()
        s[sw, a], s[sw, b] = sb[sw], sa[sw]
        c += w * sw
    return c


# ------------------------------------------------------ integration rules

def gauss_jacobi01(n, alpha):
    """Nodes (ascending) and weights of int_0^1 (1-t)^alpha f(t) dt, Golub-Welsch."""
    k = np.arange(n, dtype=float)
    al = float(alpha)
    s = 2 * k + al
    a = np.where(k == 0, -al / (al + 2.0), -(al * al) / np.where(s * (s + 2) == 0, 1, s * (s + 2)))
    kk = np.arange(1, n, dtype=float)
    s1 = 2 * kk + al
    b = np.sqrt(4 * kk * (kk + al) * kk * (kk + al) / (s1 * s1 * (s1 + 1) * (s1 - 1)))
    J = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    x, V = np.linalg.eigh(J)
    mu0 = 2.0 ** (al + 1) / (al + 1)
    w = mu0 * V[0, :] ** 2
    return 0.5 * (x + 1), w * 0.5 ** (al + 1)


_TRIG4 = ([0.816847572980459, 0.091576213509771], [0.091576213509771, 0.816847572980459],
          [0.091576213509771, 0.091576213509771], [0.108103018168070, 0.445948490915965],
          [0.445948490915965, 0.108103018168070], [0.445948490915965, 0.445948490915965])
_TRIG4W = [0.054975871827661] * 3 + [0.111690794839005] * 3
_TRIG6 = ([0.873821971016996, 0.063089014491502], [0.063089014491502, 0.873821971016996],
          [0.063089014491502, 0.063089014491502], [0.501426509658179, 0.249286745170910],
          [0.249286745170910, 0.501426509658179], [0.249286745170910, 0.249286745170910],
          [0.636502499121399, 0.310352451033785], [0.310352451033785, 0.053145049844816],
          [0.053145049844816, 0.636502499121399], [0.636502499121399, 0.053145049844816],
          [0.310352451033785, 0.636502499121399], [0.053145049844816, 0.310352451033785])
_TRIG6W = [0.025422453185103] * 3 + [0.058393137863189] * 3 + [0.041425537809187] * 6


def trig_rule(order):
    if order <= 1:
        return np.array([[1 / 3, 1 / 3]]), np.array([0.5])
    if order == 2:
        return np.array([[0, 0.5], [0.5, 0], [0.5, 0.5]]), np.full(3, 1 / 6)
    if order <= 4:
        return np.array(_TRIG4), np.array(_TRIG4W)
    if order <= 6:
        return np.array(_TRIG6), np.array(_TRIG6W)
    n = order // 2 + 1
    xx, wx = gauss_jacobi01(n, 1)
    xy, wy = gauss_jacobi01(n, 0)
    xx, wx = xx[::-1], wx[::-1]
    X, Y = np.meshgrid(xx, xy, indexing="xy")    # rows: y index j, cols: x index i
    WX, WY = np.meshgrid(wx, wy, indexing="xy")
    pts = np.stack([X.ravel(), (Y * (1 - X)).ravel()], axis=1)
    return pts, (WX * WY).ravel()


def tet_rule(order):
    n = order // 2 + 1
    xx, wx = gauss_jacobi01(n, 2)
    xy, wy = gauss_jacobi01(n, 1)
    xz, wz = gauss_jacobi01(n, 0)
    xx, wx, xy, wy = xx[::-1], wx[::-1], xy[::-1], wy[::-1]
    pts, wts = [], []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                pts.append([xx[i], xy[j] * (1 - xx[i]), xz[k] * (1 - xx[i]) * (1 - xy[j])])
                wts.append(wx[i] * wy[j] * wz[k])
    return np.array(pts), np.array(wts)


def rule(dim, order):
    return trig_rule(order) if dim == 2 else tet_rule(order)


# ------------------------------------------------------------ geometry

def jacobians(mesh):
    """F (E, dim, dim) with columns v_j - v_last (comp/meshaccess.cpp:541-562)."""
    X = mesh["coords"][mesh["els"]]  # E x (dim+1) x dim
    dim = mesh["dim"]
    F = np.transpose(X[:, :dim, :] - X[:, dim:dim + 1, :], (0, 2, 1))
    return F


# ------------------------------------------------------------- FE problem

class Problem:
    """Everything the oracle knows about one discretisation."""

    def __init__(self, mesh, p, intorder=None, dirichlet_regions=None):
        self.mesh, self.p, self.dim = mesh, p, mesh["dim"]
        self.topo = topology(mesh)
        self.dofs, self.ndof, self.free = h1_dofs(mesh, self.topo, p, dirichlet_regions)
        self.nd = self.dofs.shape[1]
        self.intorder = 2 * p - 2 if intorder is None else intorder
        self.pts, self.wts = rule(self.dim, self.intorder)
        self.F = jacobians(mesh)
        self.detF = np.linalg.det(self.F)
        self.Finv = np.linalg.inv(self.F)
        self._shapes = None

    @property
    def ne(self):
        return len(self.mesh["els"])

    def shapes(self):
        if self._shapes is None:
            self._shapes = element_shapes(self.dim, self.p, self.mesh["els"], self.pts)
        return self._shapes

    def physical_gradients(self):
        """(E, Q, nd, dim) = F^{-T} grad_ref phi."""
        _, g = self.shapes()
        return np.einsum("eba,eqib->eqia", self.Finv, g)

    def element_matrices(self, kappa=None, stiffness=True, mass=False, rho=None):
        E = self.ne
        kap = np.ones(E) if kappa is None else np.asarray(kappa, dtype=float)
        rh = np.ones(E) if rho is None else np.asarray(rho, dtype=float)
        meas = np.abs(self.detF)
        K = np.zeros((E, self.nd, self.nd))
        if stiffness:
            G = self.physical_gradients()
            K += np.einsum("q,e,eqia,eqja->eij", self.wts, meas * kap, G, G)
        if mass:
            v, _ = self.shapes()
            K += np.einsum("q,e,eqi,eqj->eij", self.wts, meas * rh, v, v)
        return K

    def pattern(self):
        d = self.dofs
        rows = np.repeat(d, self.nd, axis=1).ravel()
        cols = np.tile(d, (1, self.nd)).ravel()
        key = np.unique(rows.astype(np.int64) * self.ndof + cols)
        r = key // self.ndof
        c = (key % self.ndof).astype(np.int32)
        rowptr = np.zeros(self.ndof + 1, dtype=np.int64)
        np.add.at(rowptr, r + 1, 1)
        return np.cumsum(rowptr), c

    def assemble(self, kappa=None, stiffness=True, mass=False, rho=None):
        """CSR (rowptr int64, cols int32, vals f64) of sum_e element matrices."""
        import scipy.sparse as sp
        K = self.element_matrices(kappa, stiffness, mass, rho)
        d = self.dofs
        rows = np.repeat(d, self.nd, axis=1).ravel()
        cols = np.tile(d, (1, self.nd)).ravel()
        A = sp.coo_matrix((K.ravel(), (rows, cols)), shape=(self.ndof, self.ndof)).tocsr()
        A.sum_duplicates()
        A.sort_indices()
        return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.copy()

    def linear_form(self, fvals):
        """f_i = sum_e sum_q w_q |det F| f(e,q) phi_i(q); fvals (E, Q)."""
        v, _ = self.shapes()
        fe = np.einsum("q,e,eq,eqi->ei", self.wts, np.abs(self.detF), fvals, v)
        out = np.zeros(self.ndof)
        np.add.at(out, self.dofs.ravel(), fe.ravel())
        return out

    def points(self):
        """Physical quadrature points (E, Q, dim)."""
        X0 = self.mesh["coords"][self.mesh["els"][:, self.dim]]
        return X0[:, None, :] + np.einsum("eab,qb->eqa", self.F, self.pts)

    def evaluate(self, u):
        v, _ = self.shapes()
        return np.einsum("eqi,ei->eq", v, u[self.dofs])


def jacobi_invdiag(rowptr, cols, vals, free):
    """JacobiPrecond (linalg/jacobi.cpp:41-66): the diagonal entry of each row
    (0 where the row stores none) and its inverse on the free dofs, 0 on the
    others (1 where a free row's diagonal is 0)."""
    rowptr = np.asarray(rowptr, dtype=np.int64)
    n = len(rowptr) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    on = np.asarray(cols, dtype=np.int64) == rows
    diag = np.zeros(n)
    diag[rows[on]] = np.asarray(vals)[on]
    inv = np.where(np.asarray(free) != 0, 1.0 / np.where(diag == 0, 1, diag), 0.0)
    return diag, inv


def block_jacobi(rowptr, cols, vals, blocks):
    """BlockJacobiPrecond (linalg/blockjacobi.cpp): each block's dofs sorted
    (:390), its diagonal block A[b][b] taken from the matrix, 0 where the
    pattern has no entry (:404), inverted (:498); MultAdd y += s sum_b
    P_b^T inv_b P_b x (:596-632).  Returns (inverses, apply) with
    apply(x) = sum_b P_b^T inv_b P_b x."""
    blocks = [np.sort(np.asarray(b, dtype=np.int64)) for b in blocks]
    invs = []
    for b in blocks:
        m = np.zeros((len(b), len(b)))
        for j, r in enumerate(b):
            s0, e0 = rowptr[r], rowptr[r + 1]
            k = np.searchsorted(cols[s0:e0], b)
            ok = (k < e0 - s0)
            ok[ok] = cols[s0:e0][k[ok]] == b[ok]
            m[j, ok] = vals[s0:e0][k[ok]]
        invs.append(np.linalg.inv(m) if len(b) else m)

    def apply(x):
        y = np.zeros_like(x)
        for b, inv in zip(blocks, invs):
            y[b] += inv @ x[b]
        return y

    return invs, apply


def element_coloring(els, nv):
    """Element colouring of FESpace::FinalizeUpdate (comp/fespace.cpp:518-760)
    for an H1 space, which takes its low-order (P1, vertex-dof) space's
    colouring: in rounds of 32 colours, each element not yet coloured gets
    the lowest colour of the round that no element of the round sharing one
    of its vertices has (a 32-bit mask per vertex), elements in order.  This
    is the sequential order of that loop (the reference runs it over threads,
    so its colours vary with the schedule; any such colouring is valid).
    Returns the colour of every element."""
    els = np.asarray(els, dtype=np.int64)
    col = np.full(len(els), -1, dtype=np.int64)
    mask = np.zeros(nv, dtype=np.int64)
    base, found = 0, 0
    while found < len(els):
        mask[:] = 0
        for e in range(len(els)):
            if col[e] >= 0:
                continue
            check = 0
            for v in els[e]:
                check |= int(mask[v])
            if check != 0xFFFFFFFF:
                c = 0
                while (check >> c) & 1:
                    c += 1
                col[e] = base + c
                found += 1
                for v in els[e]:
                    mask[v] |= 1 << c
        base += 32
    return col


def sparse_inverse(rowptr, cols, vals, free):
    """SparseCholesky behind a.mat.Inverse(freedofs, "sparsecholesky")
    (linalg/sparsecholesky.cpp): apply(x) = A_ff^-1 x on the free dofs, 0 on
    the others.  The oracle factorises A_ff with scipy's SuperLU (any exact
    factorisation gives the same inverse up to rounding); returns apply."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    rowptr = np.asarray(rowptr, dtype=np.int64)
    n = len(rowptr) - 1
    A = sp.csr_matrix((np.asarray(vals), np.asarray(cols), rowptr), shape=(n, n))
    f = np.flatnonzero(np.asarray(free) != 0)
    lu = spla.splu(A[f][:, f].tocsc())

    def apply(x):
        y = np.zeros(n)
        if len(f):
            y[f] = lu.solve(np.asarray(x, dtype=float)[f])
        return y

    return apply


def two_level(smoother, coarse, nc):
    """Additive two-level preconditioner y = S x + E C E^T x with E the
    embedding of the first nc dofs: the composite smoother + E @ C @ E.T of
    the reference's SumMatrix / ProductMatrix (linalg/basematrix.hpp:413,592).
    smoother, coarse: callables (full space, first nc dofs)."""
    def apply(x):
        y = np.array(smoother(x), dtype=float)
        y[:nc] += coarse(np.asarray(x)[:nc])
        return y

    return apply


def _grid_coords(dim, dims):
    """Grid coordinates of the vertices of the structured meshes above, in
    vertex order: 3D (i, j, k) with v = (i (ny+1) + j) (nz+1) + k; 2D
    (row, column) with v = row (nx+1) + column."""
    if dim == 3:
        nx, ny, nz = dims
        i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
        return np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    nx, ny = dims[0], dims[1]
    r, c = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    return np.stack([r.ravel(), c.ravel()], 1)


def _grid_number(dim, dims, X):
    if dim == 3:
        nx, ny, nz = dims
        return (X[:, 0] * (ny + 1) + X[:, 1]) * (nz + 1) + X[:, 2]
    return X[:, 0] * (dims[0] + 1) + X[:, 1]


def mg_prolongation(dim, dims):
    """Linear prolongation between the structured vertex grids dims/2 (coarse)
    and dims (fine) over all vertices: LinearProlongation
    (multigrid/prolongation.cpp:77-144), each vertex created by refinement
    the mean of the two end points of the coarse edge it halves.  The edge of
    a vertex with odd coordinates: in 3D from X//2 along the 0/1 vector X%2
    (the Kuhn cubes' edges); in 2D with both coordinates odd the square's
    diagonal from (row R, column C+1) to (row R+1, column C).  scipy CSR,
    shape (fine vertices, coarse vertices)."""
    import scipy.sparse as sp
    X = _grid_coords(dim, dims)
    cdims = [d // 2 for d in dims]
    lo, odd = X // 2, X % 2
    nf = len(X)
    rows, cols, vals = [], [], []
    plain = odd.sum(1) == 0
    rows.append(np.flatnonzero(plain))
    cols.append(_grid_number(dim, cdims, lo[plain]))
    vals.append(np.ones(plain.sum()))
    mid = ~plain
    if dim == 2:
        diag = odd.sum(1) == 2
        edge = mid & ~diag
        a, b = lo[edge], lo[edge] + odd[edge]
        da = lo[diag] + np.array([0, 1])
        db = lo[diag] + np.array([1, 0])
        for sel, P, Q in ((edge, a, b), (diag, da, db)):
            idx = np.flatnonzero(sel)
            rows += [idx, idx]
            cols += [_grid_number(dim, cdims, P), _grid_number(dim, cdims, Q)]
            vals += [np.full(len(idx), 0.5)] * 2
    else:
        idx = np.flatnonzero(mid)
        rows += [idx, idx]
        cols += [_grid_number(dim, cdims, lo[mid]), _grid_number(dim, cdims, lo[mid] + odd[mid])]
        vals += [np.full(len(idx), 0.5)] * 2
    ncv = int(np.prod([d + 1 for d in cdims]))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nf, ncv))


def mg_hierarchy(dim, dims, rowptr, cols, vals, free, max_levels=32, dense_max=1024):
    """Level hierarchy of the structured-mesh multigrid: the free block of the
    finest P1 matrix, then coarse levels while every grid size is even, the
    level holds more than dense_max free vertices and max_levels allows; a
    coarse vertex is free iff the fine vertex at the same place is; coarse
    matrices are the Galerkin products P^T A P over the free vertices (the
    reference assembles each level of its mesh hierarchy, mgpre.cpp; for a
    coefficient constant on coarse elements the same matrices).  Returns
    dict(free=[free vertex numbers of the finest level], A=[...], P=[...])."""
    import scipy.sparse as sp
    n = len(rowptr) - 1
    A = sp.csr_matrix((np.asarray(vals, dtype=float), np.asarray(cols), np.asarray(rowptr)), shape=(n, n))
    fr = np.flatnonzero(np.asarray(free) != 0)
    As, Ps = [A[fr][:, fr].tocsr()], []
    fmask = np.asarray(free) != 0
    d = list(dims)
    while len(As) < max_levels and As[-1].shape[0] > dense_max and all(x % 2 == 0 for x in d):
        cd = [x // 2 for x in d]
        Pall = mg_prolongation(dim, d)
        Xc = _grid_coords(dim, cd)
        cmask = fmask[_grid_number(dim, d, 2 * Xc)]
        if not cmask.any():
            break
        P = Pall[np.flatnonzero(fmask)][:, np.flatnonzero(cmask)].tocsr()
        Ps.append(P)
        As.append((P.T @ As[-1] @ P).tocsr())
        fmask, d = cmask, cd
    return dict(free=fr, A=As, P=Ps, n=n)


def mg_vcycle(H, r, nu=2, omega=2.0 / 3.0, level=0):
    """One V-cycle MGM (multigrid/mgpre.cpp:229-318) on level `level` for the
    right-hand side r: nu damped-Jacobi sweeps from 0, residual, restriction
    P^T, the cycle on the next level, prolongation, nu sweeps; the coarsest
    level solved exactly (dense solve)."""
    A = H["A"][level]
    if level == len(H["A"]) - 1:
        return np.linalg.solve(A.toarray(), r) if A.shape[0] else np.zeros(0)
    dinv = 1.0 / A.diagonal()
    x = np.zeros_like(r)
    for _ in range(nu):
        x = x + omega * dinv * (r - A @ x)
    P = H["P"][level]
    x = x + P @ mg_vcycle(H, P.T @ (r - A @ x), nu, omega, level + 1)
    for _ in range(nu):
        x = x + omega * dinv * (r - A @ x)
    return x


def mg_apply(H, x, nu=2, omega=2.0 / 3.0):
    """The multigrid preconditioner on a full vertex vector: the V-cycle on the
    free entries, 0 on the others."""
    y = np.zeros(H["n"])
    f = H["free"]
    y[f] = mg_vcycle(H, np.asarray(x, dtype=float)[f], nu, omega)
    return y


def pcg(matvec, invdiag, f, tol=1e-12, maxiter=100, atol=None):
    """NGSolve CGSolver recurrence (python/krylovspace.py:263-291), FP64,
    pure numpy.  invdiag: the Jacobi inverse diagonal, or a callable
    preconditioner w = C d.  Returns (u, iterations, residuals)."""
    pre = invdiag if callable(invdiag) else (lambda v: invdiag * v)
    u = np.zeros_like(f)
    d = f - matvec(u)
    w = pre(d)
    s = w.copy()
    wdn = float(w @ d)
    res = []
    final = [None]

    def check(r):
        res.append(r)
        if len(res) == 1:
            final[0] = atol if tol is None else (max(r * tol, atol) if atol is not None else r * tol)
        return len(res) >= maxiter or r <= final[0]

    if check(math.sqrt(abs(wdn))):
        return u, len(res), res
    while True:
        w = matvec(s)
        wd = wdn
        as_s = float(s @ w)
        if as_s == 0 or wd == 0:
            break
        alpha = wd / as_s
        u += alpha * s
        d -= alpha * w
        w = pre(d)
        wdn = float(w @ d)
        if check(math.sqrt(abs(wdn))):
            break
        beta = wdn / wd
        s = beta * s + w
    return u, len(res), res


def cg_multi(matvec, invdiag, F, tol=1e-8, maxiter=200, U0=None, absolute=False):
    """Lockstep CG for the K columns of F (n x K), CGSolver<double>::MultiMult
    (linalg/cg.cpp:132-250): per column alpha_j = wd_j / <s_j, A s_j> and
    beta_j = wdn_j / wd_j, one stopping test on the 2-norm of the vector of
    w.d values (loop while n++ < maxiter and ||wdn|| > tol^2 ||wdn_0||, or
    tol^2 if absolute; ||wdn_0|| = 0 -> wdn = 1), stop when ||kss|| = 0.  A
    column with wd_j or kss_j = 0 gets alpha_j = beta_j = 0 (the reference
    divides 0 by 0 there).  matvec maps an n x K array to A times it;
    invdiag: Jacobi inverse diagonal or None.  Returns (U, steps, history of
    sqrt(||wdn||))."""
    F = np.asarray(F, dtype=float)
    pre = (lambda V: V.copy()) if invdiag is None else (lambda V: V * np.asarray(invdiag)[:, None])

    def dots(A, B):
        return np.einsum("ij,ij->j", A, B)

    def safe(a, b):
        return np.where(b != 0, a / np.where(b != 0, b, 1.0), 0.0)

    if U0 is None:
        U = np.zeros_like(F)
        D = F.copy()
    else:
        U = np.array(U0, dtype=float)
        D = F - matvec(U)
    W = pre(D)
    S = W.copy()
    wdn = dots(W, D)
    hist = [math.sqrt(np.linalg.norm(wdn))]  # before the reset, as the reference prints it
    if np.linalg.norm(wdn) == 0.0:
        wdn = np.ones_like(wdn)
    err = tol * tol if absolute else tol * tol * np.linalg.norm(wdn)
    n = 0
    while True:
        go = n < maxiter
        n += 1
        if not (go and np.linalg.norm(wdn) > err):
            break
        W = matvec(S)
        wd = wdn
        kss = dots(S, W)
        if np.linalg.norm(kss) == 0.0:
            break
        al = safe(wd, kss)
        U += al * S
        D -= al * W
        W = pre(D)
        wdn = dots(W, D)
        be = safe(wdn, wd)
        S = S * be + W
        hist.append(math.sqrt(np.linalg.norm(wdn)))
    return U, n, np.array(hist)


def adjoint_gradient(pr, kappa, f, u_obs, M):
    """J(kappa) = 0.5 (u-u_obs)^T M (u-u_obs) with A(kappa) u = f on the free
    dofs (u = 0 on Dirichlet dofs), and dJ/dkappa_e = -lam_e^T K_e u_e / kappa_e
    with A lam = M (u - u_obs): the CPU FP64 adjoint gradient (direct solves).
    Returns (J, grad, u)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    rp, c, v = pr.assemble(kappa=kappa)
    A = sp.csr_matrix((v, c, rp), shape=(pr.ndof, pr.ndof))
    fr = np.where(pr.free)[0]
    lu = spl.splu(A[fr][:, fr].tocsc())
    u = np.zeros(pr.ndof)
    u[fr] = lu.solve(f[fr])
    r = u - u_obs
    Mr = M @ r
    J = 0.5 * float(r @ Mr)
    lam = np.zeros(pr.ndof)
    lam[fr] = lu.solve(Mr[fr])
    K = pr.element_matrices()
    g = -np.einsum("ei,eij,ej->e", lam[pr.dofs], K, u[pr.dofs])
    return J, g, u
