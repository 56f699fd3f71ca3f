"""Pins the oracle's cell and global assembly: reference element against
textbook quadrature, the literal per-(i,j,q) loop against the factorized
form, J = dR/dU and R = 1/2 dF/dU by finite differences, the constrained-row
convention, GMRES against a dense solve, and Newton convergence."""
import numpy as np
import pytest

import oracle_lib as O

RNG = np.random.default_rng(99)


@pytest.mark.parametrize("deg", [1, 2])
def test_reference_element(deg):
    N, dN, w = O.refel(deg)
    assert np.allclose(N.sum(1), 1.0, atol=1e-15)           # partition of unity
    assert np.allclose(dN.sum(1), 0.0, atol=1e-14)
    assert w.sum() == pytest.approx(1.0, abs=1e-15)
    # Gauss rule with k+1 points integrates x^(2k+1) exactly on [0,1]
    n1 = deg + 1
    q = np.arange(n1 ** 3)
    # recover 1D points from N of the first 1D node
    x1 = np.array([0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)]) if deg == 1 else \
        np.array([0.5 - 0.5 * np.sqrt(0.6), 0.5, 0.5 + 0.5 * np.sqrt(0.6)])
    X = x1[q % n1]
    p = 2 * deg + 1
    assert (w * X ** p).sum() == pytest.approx(1.0 / (p + 1), abs=1e-15)
    # interpolation of a linear function reproduces it
    n = n1 ** 3
    a = np.arange(n)
    nodes_x = (a % n1) / deg
    assert np.allclose(N @ nodes_x, X, atol=1e-15)
    assert np.allclose(dN[:, :, 0] @ nodes_x, 1.0, atol=1e-14)


def _cell_state(deg, scale=0.6):
    ns = (deg + 1) ** 3
    return RNG.standard_normal(ns * 18) * scale


@pytest.mark.parametrize("deg", [1, 2])
def test_faithful_equals_factorized(deg):
    co = O.coef(20.0, 0.5)
    h = np.array([1.3, 0.7, 2.1])
    U = _cell_state(deg)
    Jf, Rf = O.cell(deg, h, co, U, faithful=True)
    Jc, Rc = O.cell(deg, h, co, U, faithful=False)
    assert np.abs(Jf - Jc).max() <= 1e-13 * np.abs(Jf).max()
    assert np.abs(Rf - Rc).max() <= 1e-13 * np.abs(Rf).max()
    assert np.abs(Jc - Jc.T).max() <= 1e-14 * np.abs(Jc).max()


def _cell_energy(deg, h, co, U):
    ns = (deg + 1) ** 3
    m = O.Mesh(deg, np.arange(ns), h, [], ns, -np.ones(ns * 18), [0], [], [], np.r_[0, np.full(ns, ns)].cumsum()[:ns + 1] * 0, [])
    return O.free_energy(m, co, U)


@pytest.mark.parametrize("deg", [1, 2])
def test_cell_residual_and_jacobian_by_finite_differences(deg):
    co = O.coef(26.0, 0.7)
    h = np.array([0.9, 1.1, 0.8])
    U = _cell_state(deg)
    J, Rm = O.cell(deg, h, co, U)
    R = -Rm
    n = U.size
    eps = 1e-6
    idx = RNG.choice(n, 12, replace=False)
    for i in idx:
        e = np.zeros(n)
        e[i] = eps
        Fp = _cell_energy(deg, h, co, U + e)
        Fm = _cell_energy(deg, h, co, U - e)
        assert R[i] == pytest.approx(0.5 * (Fp - Fm) / (2 * eps), rel=1e-6, abs=1e-9)
        _, Rp = O.cell(deg, h, co, U + e)
        _, Rmm = O.cell(deg, h, co, U - e)
        col = -(Rp - Rmm) / (2 * eps)
        assert np.allclose(J[:, i], col, rtol=1e-6, atol=1e-7 * np.abs(J).max())


@pytest.mark.parametrize("bid,lf", [(2, 0), (3, 3), (4, 5)])
def test_robin_face_term(bid, lf):
    deg = 1
    bt = 2.0
    co = O.coef(20.0, 0.5, bt=bt)
    h = np.array([1.0, 1.2, 0.9])
    U = _cell_state(deg)
    J0, R0 = O.cell(deg, h, co, U)
    J1, R1 = O.cell(deg, h, co, U, face=[(lf, bid)])
    dJ, dR = J1 - J0, -(R1 - R0)
    # symmetric, positive semidefinite face mass matrix scaled by K1/bt, zero on the normal column
    assert np.allclose(dJ, dJ.T, atol=1e-15)
    assert np.linalg.eigvalsh(dJ).min() > -1e-14
    col = (bid - 2)
    zeroed = [a * 18 + c for a in range(8) for c in range(18) if (c % 9) % 3 == col]
    assert np.abs(dJ[zeroed]).max() == 0 and np.abs(dR[zeroed]).max() == 0
    # R_face = J_face U (the face energy is quadratic)
    assert np.allclose(dR, dJ @ U, atol=1e-13)
    # sum of the face mass over unit components = K1/bt * face area
    area = {2: h[1] * h[2], 3: h[0] * h[2], 4: h[0] * h[1]}[bid]
    one = np.zeros(144)
    one[[a * 18 + 1 + (col == 1) for a in range(8)]] = 1.0  # one tangential component everywhere
    assert one @ dJ @ one == pytest.approx(co[6] / bt * area, rel=1e-13)
    # bt >= 1e10 switches the term off (assemble.cc:293)
    co_off = O.coef(20.0, 0.5, bt=1e10)
    J2, R2 = O.cell(deg, h, co_off, U, face=[(lf, bid)])
    J3, R3 = O.cell(deg, h, co_off, U)
    assert np.array_equal(J2, J3) and np.array_equal(R2, R3)


def test_global_assembly_dirichlet_convention_and_symmetry():
    mesh = O.box_mesh(1, (2, 2, 2), (-1, -1, -1), (1, 1, 1))
    co = O.coef(20.0, 0.5)
    gap = O.matep(20.0, 0.5)["gap_B"]
    U = O.distribute(mesh, O.initial_guess(mesh.n_nodes, gap, seed=3))
    val, rhs = O.assemble(mesh, co, U)
    A = O.bsr_to_dense(mesh, val)
    cons = mesh.con_of >= 0
    assert np.abs(rhs[cons]).max() == 0
    off = A[cons][:, ~cons]
    assert np.abs(off).max() == 0 and np.abs(A[~cons][:, cons]).max() == 0
    d = np.diag(A)[cons]
    assert np.all(d > 0)
    assert np.abs(A - A.T).max() <= 1e-14 * np.abs(A).max()
    # unconstrained block equals the plain sum of cell matrices restricted there
    ns = 8
    Af = np.zeros_like(A)
    rf = np.zeros(mesh.n_dofs)
    for e in range(mesh.n_cells):
        cn = mesh.cell_nodes[e * ns:(e + 1) * ns]
        g = (cn[:, None] * 18 + np.arange(18)[None, :]).reshape(-1)
        J, R = O.cell(1, mesh.cell_h[3 * e:3 * e + 3], co, U[g])
        Af[np.ix_(g, g)] += J
        rf[g] += R
        # constrained diagonal: sum over cells of |local diagonal|
    assert np.abs(A[np.ix_(~cons, ~cons)] - Af[np.ix_(~cons, ~cons)]).max() <= 1e-14 * np.abs(A).max()
    assert np.abs(rhs[~cons] - rf[~cons]).max() <= 1e-14 * np.abs(rhs).max()
    assert np.allclose(np.diag(A)[cons], np.abs(np.diag(Af))[cons], rtol=1e-14)


def test_global_faithful_equals_factorized_two_cells():
    mesh = O.box_mesh(1, (2, 1, 2), (0, 0, 0), (2.0, 1.0, 2.0), dirichlet_faces=(4,))
    co = O.coef(22.0, 0.8)
    U = O.distribute(mesh, RNG.standard_normal(mesh.n_dofs) * 0.5)
    v1, r1 = O.assemble(mesh, co, U, faithful=True)
    v2, r2 = O.assemble(mesh, co, U, faithful=False)
    assert np.abs(v1 - v2).max() <= 1e-13 * np.abs(v1).max()
    assert np.abs(r1 - r2).max() <= 1e-13 * np.abs(r1).max()


def test_gmres_block_jacobi_against_dense_solve():
    mesh = O.box_mesh(1, (2, 2, 2), (-1, -1, -1), (1, 1, 1))
    co = O.coef(20.0, 0.5)
    gap = O.matep(20.0, 0.5)["gap_B"]
    U = O.distribute(mesh, O.initial_guess(mesh.n_nodes, gap, seed=3))
    val, rhs = O.assemble(mesh, co, U)
    x, its, res = O.gmres(mesh, val, rhs, rtol=1e-12, maxit=2000, restart=20)
    assert its > 0
    A = O.bsr_to_dense(mesh, val)
    xd = np.linalg.solve(A, rhs)
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()
    assert np.linalg.norm(A @ x - rhs) <= 1e-12 * np.linalg.norm(rhs) * 1.01


def test_gmres_reports_no_convergence():
    mesh = O.box_mesh(1, (2, 2, 2), (-1, -1, -1), (1, 1, 1))
    co = O.coef(20.0, 0.5)
    U = O.distribute(mesh, O.initial_guess(mesh.n_nodes, 1.0, seed=3))
    val, rhs = O.assemble(mesh, co, U)
    _, its, _ = O.gmres(mesh, val, rhs, rtol=1e-14, maxit=1, restart=20)
    assert its == -1


def test_newton_converges_and_lowers_energy():
    mesh = O.box_mesh(1, (3, 3, 3), (-3, -3, -3), (3, 3, 3))
    co = O.coef(20.0, 0.5)
    gap = O.matep(20.0, 0.5)["gap_B"]
    U0 = O.distribute(mesh, O.initial_guess(mesh.n_nodes, gap, amp=0.05, seed=0))
    U, rc, st = O.newton(mesh, co, U0, rtol=1e-10, lin_tol=1e-3)
    assert rc == 0
    assert st[3] <= 1e-10 * st[2]
    assert O.free_energy(mesh, co, U) < O.free_energy(mesh, co, U0)
    _, r = O.assemble(mesh, co, U, J=False)
    assert np.linalg.norm(r) <= 1e-10 * st[2]


@pytest.mark.parametrize("deg,sides,bt", [(1, [1, 1, 1, 1, 0, 0], 1.0e10), (2, [2, 2, 3, 0, 4, 1], 1.5),
                                          (2, [1, 1, 1, 1, 0, 0], 1.0e10)])
def test_cellwise_apply_equals_the_assembled_operator(deg, sides, bt):
    # orc_cells_apply (J x and -R without a global J, colour-parallel) is the
    # full-size check of the GPU path; here it must equal the BSR assembly of
    # orc_assemble (distribute_local_to_global semantics) on small meshes
    from product_util import glhip, oracle_mesh
    G = glhip()
    mesh = G.Mesh(deg, (3, 2, 2), (-3.0, -2.0, -2.0), (3.0, 2.0, 2.0), sides)
    om = oracle_mesh(mesh)
    co = O.coef(22.0, 0.8, bt=bt)
    rng = np.random.default_rng(deg)
    U = O.initial_guess(mesh.n_nodes, O.matep(22.0, 0.8)["gap_B"], amp=0.3, seed=5)
    J, r = O.assemble(om, co, U)
    X = rng.standard_normal((2, mesh.n_nodes * 18))
    r2, Y = O.cells_apply(deg, mesh.arrays(), mesh.n_nodes, co, U, X)
    assert np.abs(r2 - r).max() <= 1e-13 * np.abs(r).max()
    for x, y in zip(X, Y):
        ya = np.zeros_like(x)
        O.lib().orc_bsr_spmv(om.n_nodes, om.row_ptr, om.col, J.reshape(-1), x, ya)
        assert np.abs(y - ya).max() <= 1e-13 * np.abs(ya).max()
