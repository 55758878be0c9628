"""Pins the CPU oracle against the reference's known answers and against what
the mathematics fixes (closed forms, exactness, orthogonality, invariants).
CPU only."""
import gzip
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl
from numpy.polynomial import legendre as L
from scipy.special import eval_jacobi

from oracle import fe_oracle as O
import oracle
from conftest import GOLDEN


def _vol(name):
    with gzip.open(os.path.join(GOLDEN, name), "rt") as f:
        return O.read_vol(f.read())


# ---------------------------------------------------------------- known answers
@pytest.mark.parametrize("name,ref", [("square.vol.gz", 0.004820065534967336),
                                      ("cube.vol.gz", 0.0002480226944690391)])
def test_mass_known_answer(name, ref):
    """tests/pytest/test_matrix.py:58-93: order-1 mass entry a.mat[1,1]."""
    m = _vol(name)
    pr = O.Problem(m, 1, intorder=2)
    rp, c, v = pr.assemble(stiffness=False, mass=True)
    k = np.searchsorted(c[rp[1]:rp[2]], 1)
    assert abs(v[rp[1] + k] - ref) < 1e-8
    # closed form: sum over elements at vertex 1 of |T|/6 (2D) or |T|/10 (3D)
    els = m["els"]
    vol = np.abs(pr.detF) / (2 if m["dim"] == 2 else 6)
    cf = vol[(els == 1).any(axis=1)].sum() / (6 if m["dim"] == 2 else 10)
    assert abs(v[rp[1] + k] - cf) < 1e-15


def test_element_volumes():
    """tests/catch/meshaccess.cpp:16-21 style: element volumes sum to the domain."""
    for m, dom in [(O.structured_2d(5), 1.0), (O.structured_3d(3), 1.0), (_vol("square.vol.gz"), 1.0),
                   (_vol("cube.vol.gz"), 1.0)]:
        pr = O.Problem(m, 1)
        fac = 2.0 if m["dim"] == 2 else 6.0
        assert abs(np.abs(pr.detF).sum() / fac - dom) < 1e-12


# ------------------------------------------------------------------ quadrature
@pytest.mark.parametrize("order", range(0, 11))
def test_trig_rule_exact(order):
    pts, w = O.trig_rule(order)
    for a in range(order + 1):
        for b in range(order + 1 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            tol = 1e-14 if order > 6 else 5e-14  # tabulated rules carry 15 digits
            assert abs(np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b) - exact) < tol


@pytest.mark.parametrize("order", range(0, 11))
def test_tet_rule_exact(order):
    pts, w = O.tet_rule(order)
    assert len(w) == (order // 2 + 1) ** 3
    for a in range(order + 1):
        for b in range(order + 1 - a):
            for c in range(order + 1 - a - b):
                exact = math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
                got = np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
                assert abs(got - exact) < 1e-15


def test_tet_order8_is_125_points():
    pts, w = O.tet_rule(8)
    assert len(w) == 125 and abs(w.sum() - 1 / 6) < 1e-15


# ------------------------------------------------------------- 1D polynomials
def test_intleg_nobubble_closed_form():
    """IntLegNoBubble_i(x) = L_{i+2}(x)/(1-x^2) = -P'_{i+1}(x)/((i+2)(i+1))."""
    x = np.linspace(-0.95, 0.95, 11)
    vals = O.intleg_nobubble_scaled(8, x, np.ones_like(x), 1.0)
    for i, v in enumerate(vals):
        c = np.zeros(i + 3)
        c[i + 1] = 1
        ref = -L.legval(x, L.legder(c)) / ((i + 2) * (i + 1))
        assert np.abs(v - ref).max() < 1e-14


def test_jacobi11_closed_form():
    x = np.linspace(-0.95, 0.95, 11)
    for i, v in enumerate(O.jacobi11_scaled(8, x, np.ones_like(x), 1.0)):
        assert np.abs(v - eval_jacobi(i, 1, 1, x)).max() < 1e-12


@pytest.mark.parametrize("alpha", [4, 6, 8, 10])
def test_integrated_jacobi_closed_form(alpha):
    """IntegratedJacobiAlpha_n(x) = (1/(1+x)) int_{-1}^x P_n^(alpha,0)(s) ds."""
    x = np.linspace(-0.9, 0.9, 7)
    gx, gw = np.polynomial.legendre.leggauss(30)
    for n, v in enumerate(O.intjac_alpha(6, alpha, x, 1.0)):
        ref = np.array([(xx + 1) / 2 * np.sum(gw * eval_jacobi(n, alpha, 0, (xx + 1) / 2 * gx + (xx - 1) / 2))
                        / (1 + xx) for xx in x])
        assert np.abs(v - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def test_scaled_polynomials_are_homogeneous():
    rng = np.random.default_rng(1)
    x, t = rng.uniform(-0.5, 0.5, 9), rng.uniform(0.5, 1.5, 9)
    for k, (a, b) in enumerate(zip(O.intleg_nobubble_scaled(6, x, t, 1.0),
                                   O.intleg_nobubble_scaled(6, x / t, np.ones_like(t), 1.0))):
        assert np.abs(a - t ** k * b).max() < 1e-14
    for k, (a, b) in enumerate(zip(O.intjac_alpha(5, 6, x, 1.0, y=t), O.intjac_alpha(5, 6, x / t, 1.0))):
        assert np.abs(a - t ** k * b).max() < 1e-13


# ------------------------------------------------------- 2D / 3D bubble bases
def test_trig_interior_bubble_orthogonal():
    """DubinerBasisOrthoBub is orthogonal w.r.t. the cubic bubble."""
    pts, w = O.trig_rule(16)
    lam = [pts[:, 0], pts[:, 1], 1 - pts[:, 0] - pts[:, 1]]
    D = np.array(O.dubiner_orthobub(5, lam[0], lam[1], np.ones(len(w))))
    G = (D * (w * lam[0] * lam[1] * lam[2])) @ D.T
    assert np.abs(G - np.diag(np.diag(G))).max() < 1e-13 * np.abs(np.diag(G)).max()


def test_tet_interior_bubble_orthogonal():
    pts, w = O.tet_rule(14)
    lam = [pts[:, 0], pts[:, 1], pts[:, 2], 1 - pts.sum(1)]
    D = np.array(O.dubiner3d_orthobub(3, lam[0], lam[1], lam[2], np.ones(len(w))))
    G = (D * (w * lam[0] * lam[1] * lam[2] * lam[3])) @ D.T
    assert np.abs(G - np.diag(np.diag(G))).max() < 1e-13 * np.abs(np.diag(G)).max()


def test_face_functions_scaling_identity():
    rng = np.random.default_rng(2)
    xx, yy, t = rng.uniform(0.1, 0.3, 5), rng.uniform(0.1, 0.3, 5), rng.uniform(0.7, 1.3, 5)
    n = 4
    sc = O.dubiner_orthobub_scaled(n, xx, yy, t, np.ones(5))
    un = O.dubiner_orthobub(n, xx / t, yy / t, np.ones(5))
    deg = [i + j for i in range(n + 1) for j in range(n - i + 1)]
    for s, u, d in zip(sc, un, deg):
        assert np.abs(s - t ** d * u).max() < 1e-14


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 3), (2, 5), (3, 1), (3, 2), (3, 4)])
def test_vertex_partition_of_unity_and_edge_traces(dim, p):
    pts, _ = O.rule(dim, 6)
    vn = np.array([[3, 1, 2] if dim == 2 else [3, 0, 2, 1]])
    val, grad = O.element_shapes(dim, p, vn, pts)
    nv = dim + 1
    assert np.abs(val[0, :, :nv].sum(axis=1) - 1).max() < 1e-15
    assert np.abs(grad[0, :, :nv, :].sum(axis=1)).max() < 1e-15
    # non-vertex functions vanish at the vertices
    verts = np.vstack([np.eye(dim), np.zeros(dim)])
    v2, _ = O.element_shapes(dim, p, vn, verts)
    if p > 1:
        assert np.abs(v2[0, :, nv:]).max() < 1e-15


# ------------------------------------------------------------ whole chain
def _poisson_l2_error(m, p, f_fn, exact_fn):
    pr = O.Problem(m, p, intorder=2 * p)
    rp, c, v = pr.assemble()
    A = sp.csr_matrix((v, c, rp), shape=(pr.ndof, pr.ndof))
    f = pr.linear_form(f_fn(pr.points()))
    fr = np.where(pr.free)[0]
    u = np.zeros(pr.ndof)
    u[fr] = spl.spsolve(A[fr][:, fr].tocsc(), f[fr])
    pr2 = O.Problem(m, p, intorder=2 * p + 4)
    e = pr2.evaluate(u) - exact_fn(pr2.points())
    return math.sqrt(np.sum(pr2.wts[None, :] * np.abs(pr2.detF)[:, None] * e * e))


@pytest.mark.parametrize("mesh", ["structured", "square.vol.gz"])
def test_p4_poisson_exact_2d(mesh):
    """tests/pytest/test_solvers.py:58-77: 'p4 should be exact'."""
    m = O.structured_2d(5) if mesh == "structured" else _vol(mesh)
    err = _poisson_l2_error(m, 4, lambda X: 32 * (X[..., 1] * (1 - X[..., 1]) + X[..., 0] * (1 - X[..., 0])),
                            lambda X: 16 * X[..., 0] * (1 - X[..., 0]) * X[..., 1] * (1 - X[..., 1]))
    assert err < 1e-12


def test_p6_poisson_exact_3d():
    def f(X):
        bx, by, bz = (X[..., k] * (1 - X[..., k]) for k in range(3))
        return 2 * (by * bz + bx * bz + bx * by)

    err = _poisson_l2_error(O.structured_3d(2), 6, f,
                            lambda X: np.prod([X[..., k] * (1 - X[..., k]) for k in range(3)], axis=0))
    assert err < 1e-12


@pytest.mark.parametrize("dim,p", [(2, 3), (3, 4)])
def test_stiffness_invariants(dim, p):
    m = O.structured_2d(4) if dim == 2 else O.structured_3d(2)
    pr = O.Problem(m, p)
    rp, c, v = pr.assemble()
    A = sp.csr_matrix((v, c, rp), shape=(pr.ndof, pr.ndof))
    assert abs(A - A.T).max() < 1e-14 * abs(A).max()
    one = np.zeros(pr.ndof)
    one[:len(m["coords"])] = 1.0            # interpolant of the constant 1
    assert np.abs(A @ one).max() < 1e-12
    xv = np.zeros(pr.ndof)
    xv[:len(m["coords"])] = m["coords"][:, 0]  # interpolant of u = x
    assert abs(xv @ (A @ xv) - 1.0) < 1e-12   # int |grad x|^2 over the unit domain


def test_class_numbers_of_structured_meshes():
    c2 = O.class_number(2, O.structured_2d(4)["els"])
    assert sorted(set(c2.tolist())) == [0, 2]
    c3 = O.class_number(3, O.structured_3d(3)["els"])
    vals, cnt = np.unique(c3, return_counts=True)
    assert vals.tolist() == [0, 2, 26] and cnt.tolist() == [81, 54, 27]


# -------------------------------------------------------------- C kernels
def test_oracle_spmv_and_mf_match_assembly():
    pr = O.Problem(O.structured_3d(2), 4, intorder=8)
    kappa = np.exp(np.random.default_rng(0).standard_normal(pr.ne))
    rp, c, v = pr.assemble(kappa=kappa)
    x = np.random.default_rng(1).uniform(-1, 1, pr.ndof)
    y_sp = sp.csr_matrix((v, c, rp), shape=(pr.ndof, pr.ndof)) @ x
    y_c = oracle.spmv(rp, c, v, x)
    assert np.abs(y_c - y_sp).max() <= 1e-13 * np.abs(y_sp).max()
    cls = O.class_number(3, pr.mesh["els"])
    uniq = np.unique(cls)
    _, g = pr.shapes()
    bref = np.stack([g[np.where(cls == k)[0][0]] for k in uniq])
    y_mf = oracle.mf_apply(3, bref, np.searchsorted(uniq, cls), pr.wts, pr.dofs, pr.Finv, pr.detF, kappa, x)
    assert np.abs(y_mf - y_sp).max() <= 1e-13 * np.abs(y_sp).max()


def test_cg_converges_in_k_steps_for_k_eigenvalues():
    """CG on an operator with k distinct eigenvalues stops after k steps: the
    reference's `iterations` then equals k+1 (the initial check counts)."""
    eig = np.repeat([1.0, 2.0, 5.0, 9.0], 25)
    n = len(eig)
    rp = np.arange(n + 1, dtype=np.int64)
    c = np.arange(n, dtype=np.int32)
    f = np.random.default_rng(3).uniform(-1, 1, n)
    u, its, res = oracle.pcg_csr(rp, c, eig, np.ones(n), f, tol=1e-12, maxiter=100)
    assert its == 5 and np.abs(u - f / eig).max() < 1e-12
    u2, its2, res2 = O.pcg(lambda v: eig * v, np.ones(n), f, tol=1e-12, maxiter=100)
    assert its2 == its and np.allclose(res, res2, rtol=1e-12)


def test_cg_maxiter_and_zero_rhs():
    eig = np.linspace(1, 100, 50)
    rp = np.arange(51, dtype=np.int64)
    c = np.arange(50, dtype=np.int32)
    # a zero preconditioner gives residual 0 at the first check: converged
    u, its, res = oracle.pcg_csr(rp, c, eig, np.zeros(50), np.ones(50), tol=1e-12, maxiter=3)
    assert its == 1 and res[0] == 0.0
    u, its, _ = oracle.pcg_csr(rp, c, eig, np.ones(50), np.zeros(50), tol=1e-12, maxiter=10)
    assert its == 1 and not u.any()
    u, its, res = oracle.pcg_csr(rp, c, eig, np.ones(50), np.ones(50), tol=1e-30, maxiter=7)
    assert its == 7 and len(res) == 7


def test_block_jacobi_special_cases():
    """Block Jacobi with singleton blocks is Jacobi; one block holding every
    dof is the exact inverse (A apply(x) = x, and PCG with it stops after one
    step)."""
    pr = O.Problem(O.structured_2d(4), 2)
    rp, c, v = pr.assemble()
    n = pr.ndof
    free = np.array(pr.free, dtype=np.uint8)
    _, inv = O.jacobi_invdiag(rp, c, v, free)
    fd = np.flatnonzero(free)
    _, apply1 = O.block_jacobi(rp, c, v, [[d] for d in fd])
    x = np.random.default_rng(0).uniform(-1, 1, n)
    assert np.allclose(apply1(x), inv * x, rtol=1e-14, atol=0)
    A = sp.csr_matrix((v, c, rp), shape=(n, n))
    # the free-free block, in shuffled order (the oracle sorts each block)
    _, applyall = O.block_jacobi(rp, c, v, [np.random.default_rng(1).permutation(fd)])
    z = applyall(x)
    assert np.allclose((A @ z)[fd], x[fd], rtol=0, atol=1e-12 * np.abs(x).max())
    assert np.all(z[free == 0] == 0)
    f = np.zeros(n)
    f[fd] = 1.0
    u, its, _ = O.pcg(lambda t: A @ t, applyall, f, tol=1e-12, maxiter=10)
    assert its == 2 and np.allclose((A @ u)[fd], f[fd], atol=1e-12)


def test_sparse_inverse_and_two_level():
    """The oracle's sparse inverse solves A_ff y = x and is 0 off the free
    dofs, equal to the dense inverse of the free block (block Jacobi with one
    block); the leading vertex block of a hierarchical p=3 matrix is the P1
    matrix of the same form (vertex shape functions are the barycentric
    coordinates), so the two-level preconditioner's coarse space is P1; with
    a zero smoother and the whole space as coarse space the two-level
    preconditioner is the exact inverse (PCG stops after one step)."""
    rng = np.random.default_rng(3)
    pr = O.Problem(O.structured_2d(6), 3)
    kappa = np.exp(rng.standard_normal(pr.mesh["els"].shape[0]))
    rp, c, v = pr.assemble(kappa=kappa)
    n = pr.ndof
    free = np.array(pr.free, dtype=np.uint8)
    fd = np.flatnonzero(free)
    A = sp.csr_matrix((v, c, rp), shape=(n, n))
    inv = O.sparse_inverse(rp, c, v, free)
    x = rng.uniform(-1, 1, n)
    y = inv(x)
    assert np.all(y[free == 0] == 0)
    assert np.abs((A @ y)[fd] - x[fd]).max() <= 1e-12 * np.abs(x).max()
    _, dense = O.block_jacobi(rp, c, v, [fd])
    assert np.allclose(y, dense(x), rtol=0, atol=1e-12 * np.abs(y).max())
    # leading vertex block = P1 matrix
    p1 = O.Problem(O.structured_2d(6), 1)
    rp1, c1, v1 = p1.assemble(kappa=kappa)
    nv = p1.ndof
    A1 = sp.csr_matrix((v1, c1, rp1), shape=(nv, nv))
    assert abs(A[:nv][:, :nv] - A1).max() <= 1e-13 * abs(A1).max()
    # two-level with the whole space as coarse space and no smoother: exact
    tl = O.two_level(lambda t: np.zeros_like(t), inv, n)
    f = np.where(free != 0, 1.0, 0.0)
    u, its, _ = O.pcg(lambda t: A @ t, tl, f, tol=1e-12, maxiter=10)
    assert its == 2 and np.allclose((A @ u)[fd], f[fd], atol=1e-12)
    # Jacobi + P1 coarse: the coarse part acts on the first nv entries only
    _, jinv = O.jacobi_invdiag(rp, c, v, free)
    cinv = O.sparse_inverse(rp1, c1, v1, free[:nv])
    tl2 = O.two_level(lambda t: jinv * t, cinv, nv)
    z = tl2(x)
    assert np.allclose(z[nv:], jinv[nv:] * x[nv:], rtol=0, atol=0)
    assert np.allclose(z[:nv], jinv[:nv] * x[:nv] + cinv(x[:nv]), rtol=1e-14, atol=0)


def test_multigrid_prolongation_galerkin_and_vcycle():
    """Multigrid oracle: the prolongation reproduces linear functions exactly
    (3D and 2D, both diagonals of the 2D rule used); the Galerkin product
    P^T A P of the P1 matrix (kappa = 1) equals the P1 matrix assembled on
    the coarse structured mesh (nested P1 spaces), on the free vertices; the
    V-cycle is a symmetric operator and, as a stationary iteration on the P1
    matrix with a lognormal coefficient, contracts the energy norm of the
    error by a factor below 0.45 on 8^3 and 16^3 alike (independent of h);
    with one level it is the exact inverse."""
    rng = np.random.default_rng(5)
    for dim, dims in ((3, [4, 6, 8]), (2, [8, 6])):
        P = O.mg_prolongation(dim, dims)
        Xf = O._grid_coords(dim, dims).astype(float)
        Xc = 2.0 * O._grid_coords(dim, [d // 2 for d in dims])
        a = rng.uniform(-1, 1, dim)
        assert np.abs(P @ (Xc @ a + 0.7) - (Xf @ a + 0.7)).max() <= 1e-14
        assert np.allclose(P.sum(1), 1.0)
    for dim, n in ((2, 16), (3, 8)):
        fine = O.Problem(O.structured_2d(n) if dim == 2 else O.structured_3d(n), 1)
        coarse = O.Problem(O.structured_2d(n // 2) if dim == 2 else O.structured_3d(n // 2), 1)
        rp, c, v = fine.assemble()
        H = O.mg_hierarchy(dim, [n] * dim, rp, c, v, np.array(fine.free, dtype=np.uint8), dense_max=0,
                           max_levels=2)
        rpc, cc, vc = coarse.assemble()
        nc = coarse.ndof
        Ac = sp.csr_matrix((vc, cc, rpc), shape=(nc, nc))
        fc = np.flatnonzero(np.array(coarse.free))
        G = H["A"][1]
        assert G.shape == (len(fc), len(fc))
        assert abs(G - Ac[fc][:, fc]).max() <= 1e-13 * abs(Ac).max()
    # V-cycle on 3D P1 with lognormal kappa: symmetric, contracting, exact on one level
    for n in (8, 16):
        pr = O.Problem(O.structured_3d(n), 1)
        kappa = np.exp(np.random.default_rng(n).standard_normal(pr.mesh["els"].shape[0]))
        rp, c, v = pr.assemble(kappa=kappa)
        free = np.array(pr.free, dtype=np.uint8)
        H = O.mg_hierarchy(3, [n] * 3, rp, c, v, free, dense_max=30)
        assert len(H["A"]) >= 2
        A = H["A"][0]
        x, y = rng.uniform(-1, 1, A.shape[0]), rng.uniform(-1, 1, A.shape[0])
        Mx, My = O.mg_vcycle(H, x), O.mg_vcycle(H, y)
        assert abs(y @ Mx - x @ My) <= 1e-12 * abs(y @ Mx)
        e = rng.uniform(-1, 1, A.shape[0])
        for _ in range(8):
            e_new = e - O.mg_vcycle(H, A @ e, nu=2, omega=0.85)
            rate = np.sqrt(e_new @ (A @ e_new) / (e @ (A @ e)))
            e = e_new
        assert rate < 0.45, (n, rate)  # 0.36 at 8^3, 0.39 at 16^3: independent of h
        H1 = O.mg_hierarchy(3, [n] * 3, rp, c, v, free, max_levels=1)
        b = rng.uniform(-1, 1, A.shape[0])
        assert np.abs(A @ O.mg_vcycle(H1, b) - b).max() <= 1e-10 * np.abs(b).max()
        z = O.mg_apply(H, rng.uniform(-1, 1, pr.ndof))
        assert np.all(z[free == 0] == 0)


def test_element_coloring_known_answers_and_first_fit():
    """Element colouring (comp/fespace.cpp:518-760): the 2 triangles of one
    square share an edge (colours 0, 1); the 6 Kuhn tetrahedra of one cube
    share the main diagonal (colours 0..5).  On larger meshes (fewer than 32
    colours) every element has the smallest colour that no earlier element
    sharing a vertex has (sequential first fit, checked against vertex
    neighbour sets), and no colour class shares a vertex."""
    assert O.element_coloring(O.structured_2d(1)["els"], 4).tolist() == [0, 1]
    assert O.element_coloring(O.structured_3d(1)["els"], 8).tolist() == [0, 1, 2, 3, 4, 5]
    for mesh in (O.structured_2d(6), O.structured_3d(3)):
        els = np.asarray(mesh["els"])
        col = O.element_coloring(els, int(els.max()) + 1)
        assert col.max() < 32
        by_vertex = {}
        for e, vs in enumerate(els):
            for v in vs:
                by_vertex.setdefault(int(v), []).append(e)
        for e, vs in enumerate(els):
            earlier = {int(col[f]) for v in vs for f in by_vertex[int(v)] if f < e}
            assert int(col[e]) == min(set(range(33)) - earlier)
            later = {int(col[f]) for v in vs for f in by_vertex[int(v)] if f != e}
            assert int(col[e]) not in later


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((n, n))
    return Q @ Q.T + n * np.eye(n)


def test_cg_multi_single_column_is_pcg():
    """One column: the lockstep recurrence and stopping rule are those of the
    single CG (same iterates, step count within the counting convention)."""
    A = _spd(40, 1)
    f = np.random.default_rng(2).standard_normal(40)
    inv = 1.0 / np.diag(A)
    U, steps, hist = O.cg_multi(lambda V: A @ V, inv, f[:, None], tol=1e-10, maxiter=500)
    u, its, res = O.pcg(lambda v: A @ v, inv, f, tol=1e-10, maxiter=500)
    assert abs(steps - its) <= 1
    assert np.abs(U[:, 0] - u).max() <= 1e-12 * np.abs(u).max()
    assert np.allclose(hist, res[:len(hist)], rtol=1e-10)


def test_cg_multi_columns():
    """Identical columns take the steps of one; every column solves its own
    system; k distinct eigenvalues give convergence within k iterations;
    a zero column stays zero; maxiter bounds the steps."""
    A = _spd(30, 3)
    rng = np.random.default_rng(4)
    f = rng.standard_normal(30)
    inv = 1.0 / np.diag(A)
    mv = lambda V: A @ V
    U1, s1, _ = O.cg_multi(mv, inv, f[:, None], tol=1e-10, maxiter=500)
    U3, s3, _ = O.cg_multi(mv, inv, np.stack([f, f, f], 1), tol=1e-10, maxiter=500)
    assert s3 == s1 and np.array_equal(U3[:, 0], U3[:, 2]) and np.allclose(U3[:, 1], U1[:, 0], rtol=0, atol=1e-14)
    F = rng.standard_normal((30, 4))
    U, steps, _ = O.cg_multi(mv, inv, F, tol=1e-11, maxiter=500)
    X = np.linalg.solve(A, F)
    assert np.abs(U - X).max() <= 1e-8 * np.abs(X).max()
    # diagonal matrix with 3 distinct eigenvalues, no preconditioner: 3 iterations (4 tests)
    Dg = np.diag(np.repeat([1.0, 4.0, 9.0], 5))
    _, sd, _ = O.cg_multi(lambda V: Dg @ V, None, rng.standard_normal((15, 3)), tol=1e-12, maxiter=50)
    assert sd == 4
    Fz = np.stack([f, np.zeros(30)], 1)
    Uz, _, _ = O.cg_multi(mv, inv, Fz, tol=1e-10, maxiter=500)
    assert np.all(Uz[:, 1] == 0) and np.all(np.isfinite(Uz))
    _, sm, hm = O.cg_multi(mv, inv, F, tol=1e-14, maxiter=3)
    assert sm == 4 and len(hm) == 4
