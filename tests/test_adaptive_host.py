"""Adaptive Q1 meshes on the host (no GPU): forest refinement with 2:1
balance, hanging-node constraint lines against an independent derivation,
the condensed pattern, solution transfer, Kelly indicators and the
fixed-number refinement flags."""
import numpy as np
import pytest

import oracle_lib as O
from product_util import adaptive_forest, glhip, oracle_adaptive, oracle_from_library, side_masks

G = glhip()
LO, HI = (-20.0, -20.0, -6.0), (20.0, 20.0, 6.0)
SIDES = [1, 1, 1, 1, 0, 0]


def _continuous(mesh, U):
    """Hanging values from their masters (a conforming Q1 field)."""
    c = mesh.constraints()
    U = np.array(U, dtype=np.float64).reshape(-1, 18)
    hp = c["hang_ptr"]
    for v in range(mesh.n_nodes):
        if hp[v + 1] > hp[v]:
            U[v] = sum(c["hang_w"][k] * U[c["hang_m"][k]] for k in range(hp[v], hp[v + 1]))
    return U.reshape(-1)


def _geometry(mesh):
    a = mesh.arrays()
    xyz = a["node_xyz"].reshape(-1, 3)
    cn = a["cell_nodes"].reshape(-1, 8)
    h = a["cell_h"].reshape(-1, 3)
    return a, xyz, cn, xyz[cn[:, 0]], h


@pytest.mark.parametrize("steps,seed", [(1, 0), (2, 1), (3, 2)])
def test_forest_is_balanced_and_tiles_the_box(steps, seed):
    f = adaptive_forest(G, (4, 4, 2), LO, HI, SIDES, steps, seed)
    m = f.mesh()
    a, xyz, cn, lo, h = _geometry(m)
    vol = np.prod(h, 1)
    assert vol.sum() == pytest.approx(np.prod(np.subtract(HI, LO)), rel=1e-12)
    # 2:1 balance: cells whose closures touch differ by at most one level
    lvl = np.round(np.log2((HI[0] - LO[0]) / 4 / h[:, 0])).astype(int)
    hi = lo + h
    for e in range(m.n_cells):
        touch = np.all((lo <= hi[e] + 1e-9) & (hi >= lo[e] - 1e-9), 1)
        assert np.abs(lvl[touch] - lvl[e]).max() <= 1


@pytest.mark.parametrize("steps,seed", [(1, 0), (2, 1), (3, 2)])
def test_hanging_lines_match_independent_derivation(steps, seed):
    f = adaptive_forest(G, (4, 4, 2), LO, HI, SIDES, steps, seed)
    m = f.mesh()
    a, xyz, cn, lo, h = _geometry(m)
    ref = O.hanging_lines(a["cell_nodes"], xyz, lo, h)
    c = m.constraints()
    hp = c["hang_ptr"]
    got = {v: {int(c["hang_m"][k]): float(c["hang_w"][k]) for k in range(hp[v], hp[v + 1])}
           for v in range(m.n_nodes) if hp[v + 1] > hp[v]}
    assert set(got) == set(ref) and len(ref) == c["n_hanging"] > 0
    for v in ref:
        assert set(got[v]) == set(ref[v])
        for mm in ref[v]:
            assert got[v][mm] == pytest.approx(ref[v][mm], abs=1e-15)
        assert sum(got[v].values()) == pytest.approx(1.0, abs=1e-15)  # partition of unity
    # hanging nodes carry the flag bit and no Dirichlet bits
    mask = a["node_mask"]
    hang = np.array(sorted(ref))
    assert np.all(mask[hang] == np.uint32(1 << 31))


def test_condensed_pattern_holds_every_oracle_entry():
    f = adaptive_forest(G, (4, 4, 2), LO, HI, SIDES, 2, 3)
    m = f.mesh()
    a = m.arrays()
    om = oracle_adaptive(m, LO, HI, SIDES)
    prod = set()
    for r in range(m.n_nodes):
        for k in range(a["row_ptr"][r], a["row_ptr"][r + 1]):
            prod.add((r, int(a["col"][k])))
    for r in range(om.n_nodes):
        for k in range(om.row_ptr[r], om.row_ptr[r + 1]):
            assert (r, int(om.col[k])) in prod
    # hanging rows and columns hold only the diagonal block
    hp = m.constraints()["hang_ptr"]
    hang = set(np.nonzero(np.diff(hp))[0].tolist())
    for (r, s) in prod:
        if r in hang or s in hang:
            assert r == s


def test_transfer_reproduces_linear_fields_and_interpolates():
    f = adaptive_forest(G, (3, 3, 2), LO, HI, SIDES, 1, 4)
    m = f.mesh()
    rng = np.random.default_rng(0)
    A = rng.standard_normal((3, 18))
    b = rng.standard_normal(18)
    xyz = m.arrays()["node_xyz"].reshape(-1, 3)
    U = (xyz @ A + b).reshape(-1)
    f2 = f.refine((rng.random(f.n_cells) < 0.4).astype(np.uint8))
    m2 = f2.mesh()
    U2 = f.transfer(m, U, f2, m2)
    xyz2 = m2.arrays()["node_xyz"].reshape(-1, 3)
    assert np.abs(U2 - (xyz2 @ A + b).reshape(-1)).max() < 1e-12
    # a general conforming field: new nodes take the trilinear interpolant of
    # the old cell that contains them
    U = _continuous(m, rng.standard_normal(m.n_nodes * 18))
    U2 = f.transfer(m, U, f2, m2).reshape(-1, 18)
    a1, _, cn, lo, h = _geometry(m)
    for v in rng.choice(m2.n_nodes, 40, replace=False):
        x = xyz2[v]
        t = (x - lo) / h
        e = int(np.nonzero(np.all((t > -1e-12) & (t < 1 + 1e-12), 1))[0][0])
        tv = np.clip(t[e], 0, 1)
        val = np.zeros(18)
        for k in range(8):
            bits = (k & 1, (k >> 1) & 1, k >> 2)
            w = np.prod([tv[d] if bits[d] else 1 - tv[d] for d in range(3)])
            val += w * U.reshape(-1, 18)[cn[e][k]]
        assert np.abs(U2[v] - val).max() < 1e-12


def test_kelly_matches_independent_restatement():
    f = adaptive_forest(G, (3, 3, 2), LO, HI, SIDES, 2, 5, frac=0.3)
    m = f.mesh()
    a, xyz, cn, lo, h = _geometry(m)
    rng = np.random.default_rng(1)
    U = _continuous(m, rng.standard_normal(m.n_nodes * 18))
    eta = f.kelly(m, U)
    ref = O.kelly(a["cell_nodes"], lo, h, U)
    assert eta.dtype == np.float32
    assert np.abs(eta - ref).max() <= 1e-6 * ref.max()
    # a globally linear field has no jumps
    U_lin = (xyz @ rng.standard_normal((3, 18))).reshape(-1)
    assert np.abs(f.kelly(m, U_lin)).max() < 1e-10


def test_fixed_number_flags():
    eta = np.array([0.5, 3.0, 1.0, 3.0, 0.1, 2.0, 0.0, 7.0, 1.5, 0.2], np.float32)
    fl = G.flag_fixed_number(eta, 0.3)
    assert fl.sum() == 3 and set(np.nonzero(fl)[0]) == {7, 1, 3}
    assert G.flag_fixed_number(eta, 0.0).sum() == 0
    with pytest.raises(G.GLHError):
        G.flag_fixed_number(eta, 1.5)


@pytest.mark.parametrize("deg,axes", [(1, (0, 1)), (2, (0,)), (1, (0, 1, 2))])
def test_periodic_box_lines_match_independent_derivation(deg, axes):
    n = (3, 2, 2)
    lo, hi = (-4.0, -3.0, -2.0), (4.0, 3.0, 2.0)
    sides = [1] * 6
    for a in axes:
        sides[2 * a] = sides[2 * a + 1] = 5
    if 2 not in axes:
        sides[4] = sides[5] = 0
    m = G.Mesh(deg, n, lo, hi, sides)
    c = m.constraints()
    om = O.periodic_oracle_mesh(deg, n, lo, hi, side_masks(sides), axes)
    hp = c["hang_ptr"]
    # node-level lines of the product against the dof-level oracle lines
    for v in range(m.n_nodes):
        got = {int(c["hang_m"][k]): float(c["hang_w"][k]) for k in range(hp[v], hp[v + 1])}
        comp = []
        for cc in range(18):
            g = v * 18 + cc
            ci = om.con_of[g]
            if ci < 0:
                comp.append(None)
            else:
                comp.append({int(om.con_master[k]) // 18: float(om.con_weight[k])
                             for k in range(om.con_ptr[ci], om.con_ptr[ci + 1])})
        if got:
            # a periodic image: every component either follows the same image
            # or is Dirichlet there (its image is Dirichlet too, so the value is 0)
            for line in comp:
                assert line is not None and (line == got or line == {})
        else:
            assert all(line is None or line == {} for line in comp)
    assert c["n_hanging"] > 0


# ------------------------------------------------------------------ Q2 forests
def _q2(steps, seed, n=(3, 3, 2)):
    f = adaptive_forest(G, n, LO, HI, SIDES, steps, seed, degree=2)
    return f, f.mesh()


def _q2_geometry(mesh):
    a = mesh.arrays()
    xyz = a["node_xyz"].reshape(-1, 3)
    cn = a["cell_nodes"].reshape(-1, 27)
    return a, xyz, cn, xyz[cn[:, 0]], a["cell_h"].reshape(-1, 3)


def _poly(xyz, C):
    """sum_{i,j,k <= 2} C[i,j,k,:] x^i y^j z^k, scaled to the box (a Q2 field)."""
    x, y, z = (xyz[:, 0] / 20.0, xyz[:, 1] / 20.0, xyz[:, 2] / 6.0)
    out = np.zeros((xyz.shape[0], C.shape[-1]))
    for i in range(C.shape[0]):
        for j in range(C.shape[1]):
            for k in range(C.shape[2]):
                out += np.outer(x ** i * y ** j * z ** k, C[i, j, k])
    return out


@pytest.mark.parametrize("steps,seed", [(1, 0), (2, 1), (2, 5)])
def test_q2_hanging_lines_match_independent_derivation(steps, seed):
    f, m = _q2(steps, seed)
    assert m.deg == 2 and m.ns == 27
    a, xyz, cn, lo, h = _q2_geometry(m)
    ref = O.hanging_lines(a["cell_nodes"], xyz, lo, h, deg=2)
    c = m.constraints()
    hp = c["hang_ptr"]
    got = {v: {int(c["hang_m"][k]): c["hang_w"][k] for k in range(hp[v], hp[v + 1])}
           for v in range(m.n_nodes) if hp[v + 1] > hp[v]}
    assert len(got) > 0 and set(got) == set(ref)
    for v in got:
        assert set(got[v]) == set(ref[v]), v
        for mst, w in got[v].items():
            assert w == pytest.approx(ref[v][mst], abs=1e-14)
    # Q2 weights at a quarter of a coarse edge: 3/8, 3/4, -1/8 appear
    ws = np.round(np.concatenate([list(l.values()) for l in got.values()]), 12)
    assert any(np.isclose(ws, -1.0 / 8.0)) and any(np.isclose(ws, 3.0 / 8.0))


@pytest.mark.parametrize("steps,seed", [(2, 1), (3, 2)])
def test_q2_constraints_reproduce_triquadratics_and_only_those(steps, seed):
    f, m = _q2(steps, seed)
    a, xyz, cn, lo, h = _q2_geometry(m)
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((3, 3, 3, 18))
    P = _poly(xyz, C)
    U = _continuous(m, P.reshape(-1)).reshape(-1, 18)
    hang = m.constraints()["hang_ptr"]
    hv = np.nonzero(np.diff(hang) > 0)[0]
    assert hv.size > 0
    # a field of degree <= 2 in each direction is its own Q2 interpolant on
    # every cell, so the hanging values must equal it
    assert np.abs(U[hv] - P[hv]).max() <= 1e-12 * np.abs(P).max()
    # a cubic is not: the coarse face interpolant differs from it
    cub = np.outer((xyz[:, 0] / 20.0) ** 3 + (xyz[:, 1] / 20.0) ** 3, np.ones(18))
    Uc = _continuous(m, cub.reshape(-1)).reshape(-1, 18)
    assert np.abs(Uc[hv] - cub[hv]).max() > 1e-6


def test_q2_transfer_reproduces_triquadratics_and_interpolates():
    f, m = _q2(1, 4)
    a, xyz, cn, lo, h = _q2_geometry(m)
    rng = np.random.default_rng(0)
    C = rng.standard_normal((3, 3, 3, 18))
    U = _continuous(m, _poly(xyz, C).reshape(-1))
    f2 = f.refine((rng.random(f.n_cells) < 0.4).astype(np.uint8))
    m2 = f2.mesh()
    assert m2.deg == 2
    xyz2 = m2.arrays()["node_xyz"].reshape(-1, 3)
    U2 = f.transfer(m, U, f2, m2)
    assert np.abs(U2 - _poly(xyz2, C).reshape(-1)).max() <= 1e-12 * np.abs(U).max()
    # a general conforming field: new nodes take the Q2 interpolant of the
    # old cell that contains them
    U = _continuous(m, rng.standard_normal(m.n_nodes * 18)).reshape(-1, 18)
    U2 = f.transfer(m, U.reshape(-1), f2, m2).reshape(-1, 18)
    for v in rng.choice(m2.n_nodes, 40, replace=False):
        t = (xyz2[v] - lo) / h
        e = int(np.nonzero(np.all((t > -1e-12) & (t < 1 + 1e-12), 1))[0][0])
        tv = np.clip(t[e], 0, 1)
        val = np.zeros(18)
        for k in range(27):
            idx = (k % 3, (k // 3) % 3, k // 9)
            w = np.prod([O.lagrange_1d(2, idx[d], tv[d]) for d in range(3)])
            val += w * U[cn[e][k]]
        assert np.abs(U2[v] - val).max() < 1e-12 * max(1.0, np.abs(U).max())


def test_q2_kelly_matches_independent_restatement():
    f, m = _q2(2, 5)
    a, xyz, cn, lo, h = _q2_geometry(m)
    rng = np.random.default_rng(1)
    U = _continuous(m, rng.standard_normal(m.n_nodes * 18))
    eta = f.kelly(m, U)
    ref = O.kelly(a["cell_nodes"], lo, h, U, deg=2)
    assert eta.dtype == np.float32
    assert np.abs(eta - ref).max() <= 1e-6 * ref.max()
    # a global field of degree <= 2 per direction has no normal-derivative jumps
    Up = _continuous(m, _poly(xyz, rng.standard_normal((3, 3, 3, 18))).reshape(-1))
    assert np.abs(f.kelly(m, Up)).max() < 1e-8 * max(1.0, ref.max())


# ------------------------------------------------------------ periodic forests
PER = [5, 5, 5, 5, 0, 0]  # x and y periodic, A = 0 on the z faces


def _cells(mesh):
    a = mesh.arrays()
    ns = (mesh.deg + 1) ** 3
    xyz = a["node_xyz"].reshape(-1, 3)
    cn = a["cell_nodes"].reshape(-1, ns)
    return xyz, cn, xyz[cn[:, 0]], a["cell_h"].reshape(-1, 3)


def _value(deg, cn_e, lo_e, h_e, U, x):
    n1 = deg + 1
    t = (x - lo_e) / h_e
    out = np.zeros(18)
    for a in range(n1 ** 3):
        idx = (a % n1, (a // n1) % n1, a // (n1 * n1))
        out += np.prod([O.lagrange_1d(deg, idx[d], t[d]) for d in range(3)]) * U[cn_e[a]]
    return out


def _face_jumps(mesh, U, period, tol=1e-9):
    """Largest difference of the cell-wise Q_k fields of two face neighbours
    (across periodic faces too) at sample points of their common face."""
    deg = mesh.deg
    xyz, cn, lo, h = _cells(mesh)
    U = U.reshape(-1, 18)
    hi = lo + h
    worst, pairs = 0.0, 0
    rng = np.random.default_rng(0)
    for e in range(cn.shape[0]):
        for d in range(3):
            t1, t2 = (d + 1) % 3, (d + 2) % 3
            x_f = hi[e][d]
            for k in range(cn.shape[0]):
                wrap = 0.0
                if abs(lo[k][d] - x_f) > tol:
                    if period[d] <= 0 or abs(lo[k][d] + period[d] - x_f) > tol:
                        continue
                    wrap = period[d]
                a1, b1 = max(lo[e][t1], lo[k][t1]), min(hi[e][t1], hi[k][t1])
                a2, b2 = max(lo[e][t2], lo[k][t2]), min(hi[e][t2], hi[k][t2])
                if b1 - a1 <= tol or b2 - a2 <= tol:
                    continue
                pairs += 1
                for _ in range(3):
                    x = np.zeros(3)
                    x[d] = x_f
                    x[t1] = a1 + rng.random() * (b1 - a1)
                    x[t2] = a2 + rng.random() * (b2 - a2)
                    xk = x.copy()
                    xk[d] -= wrap
                    ve = _value(deg, cn[e], lo[e], h[e], U, x)
                    vk = _value(deg, cn[k], lo[k], h[k], U, xk)
                    worst = max(worst, np.abs(ve - vk).max())
    return worst, pairs


@pytest.mark.parametrize("deg", [1, 2])
def test_periodic_forest_uniform_node_count_and_identification(deg):
    f = G.Forest((4, 4, 2), LO, HI, PER, degree=deg)
    f = f.refine(np.ones(f.n_cells, np.uint8))
    m = f.mesh()
    assert m.n_nodes == (8 * deg) * (8 * deg) * (4 * deg + 1)
    assert m.constraints()["n_hanging"] == 0
    xyz, cn, lo, h = _cells(m)
    assert xyz[:, 0].max() < HI[0] - 1e-9 and xyz[:, 1].max() < HI[1] - 1e-9  # + faces are the - faces
    assert np.isclose(xyz[:, 2].max(), HI[2])


@pytest.mark.parametrize("deg,steps,seed", [(1, 2, 1), (1, 3, 2), (2, 2, 3)])
def test_periodic_forest_fields_are_continuous_across_every_face(deg, steps, seed):
    f = adaptive_forest(G, (3, 3, 2), LO, HI, PER, steps, seed, degree=deg)
    m = f.mesh()
    assert m.constraints()["n_hanging"] > 0
    rng = np.random.default_rng(seed)
    U = _continuous(m, rng.standard_normal(m.n_nodes * 18))
    L = (HI[0] - LO[0], HI[1] - LO[1], 0.0)
    worst, pairs = _face_jumps(m, U, L)
    # the periodic faces are among the checked pairs
    _, per_pairs = _face_jumps(m, U, (L[0], 0.0, 0.0))
    _, plain_pairs = _face_jumps(m, U, (0.0, 0.0, 0.0))
    assert pairs > per_pairs > plain_pairs > 0
    assert worst <= 1e-12 * np.abs(U).max()
    # without the lines the fields jump: the check sees a wrong constraint
    Ub = rng.standard_normal(m.n_nodes * 18)
    assert _face_jumps(m, Ub, L)[0] > 1e-3


def test_periodic_forest_is_balanced_across_the_periodic_faces():
    f = adaptive_forest(G, (3, 3, 2), LO, HI, PER, 3, 4)
    m = f.mesh()
    xyz, cn, lo, h = _cells(m)
    hi = lo + h
    checked = 0
    for d in (0, 1):
        t1, t2 = (d + 1) % 3, (d + 2) % 3
        top = np.nonzero(np.isclose(hi[:, d], HI[d]))[0]
        bot = np.nonzero(np.isclose(lo[:, d], LO[d]))[0]
        for e in top:
            for k in bot:
                ov1 = min(hi[e][t1], hi[k][t1]) - max(lo[e][t1], lo[k][t1])
                ov2 = min(hi[e][t2], hi[k][t2]) - max(lo[e][t2], lo[k][t2])
                if ov1 > 1e-9 and ov2 > 1e-9:
                    checked += 1
                    assert 0.5 - 1e-12 <= h[e][d] / h[k][d] <= 2.0 + 1e-12
    assert checked > 0


@pytest.mark.parametrize("deg", [1, 2])
def test_periodic_forest_kelly_matches_independent_restatement(deg):
    f = adaptive_forest(G, (3, 3, 2), LO, HI, PER, 2, 6, degree=deg)
    m = f.mesh()
    xyz, cn, lo, h = _cells(m)
    rng = np.random.default_rng(2)
    U = _continuous(m, rng.standard_normal(m.n_nodes * 18))
    eta = f.kelly(m, U)
    ref = O.kelly(m.arrays()["cell_nodes"], lo, h, U, deg=deg, period=(HI[0] - LO[0], HI[1] - LO[1], 0.0))
    assert np.abs(eta - ref).max() <= 1e-6 * ref.max()
    nonper = O.kelly(m.arrays()["cell_nodes"], lo, h, U, deg=deg)
    assert np.abs(eta - nonper).max() > 1e-3 * ref.max()  # the periodic faces count


def test_periodic_forest_rejects_one_sided_periodicity():
    with pytest.raises(G.GLHError):
        G.Forest((3, 3, 2), LO, HI, [5, 1, 1, 1, 0, 0])


@pytest.mark.parametrize("deg,steps,seed", [(1, 2, 1), (2, 2, 3)])
def test_periodic_forest_lines_match_independent_derivation(deg, steps, seed):
    f = adaptive_forest(G, (3, 3, 2), LO, HI, PER, steps, seed, degree=deg)
    m = f.mesh()
    xyz, cn, lo, h = _cells(m)
    ref = O.hanging_lines(m.arrays()["cell_nodes"], xyz, lo, h, deg=deg,
                          period=(HI[0] - LO[0], HI[1] - LO[1], 0.0))
    c = m.constraints()
    hp = c["hang_ptr"]
    got = {v: {int(c["hang_m"][k]): c["hang_w"][k] for k in range(hp[v], hp[v + 1])}
           for v in range(m.n_nodes) if hp[v + 1] > hp[v]}
    assert len(got) > 0 and set(got) == set(ref)
    for v in got:
        assert set(got[v]) == set(ref[v]), v
        for mst, w in got[v].items():
            assert w == pytest.approx(ref[v][mst], abs=1e-14)


# ----------------------------------------------------------------- coarsening
def _touching_levels_ok(mesh, tol=1e-9):
    """2:1 balance seen from geometry: cells whose closed boxes touch differ
    in size by at most a factor 2."""
    xyz, cn, lo, h = _cells(mesh)
    hi = lo + h
    for e in range(cn.shape[0]):
        touch = np.all((lo <= hi[e] + tol) & (hi >= lo[e] - tol), 1)
        r = h[touch, 0] / h[e, 0]
        if np.any(r > 2.0 + 1e-12) or np.any(r < 0.5 - 1e-12):
            return False
    return True


def test_refine_coarsen_flags():
    eta = np.array([0.5, 3.0, 1.0, 3.0, 0.1, 2.0, 0.0, 7.0, 1.5, 0.2], np.float32)
    fl = G.flag_refine_coarsen(eta, 0.3, 0.2)
    assert set(np.nonzero(fl == 1)[0]) == {7, 1, 3} and set(np.nonzero(fl == 2)[0]) == {6, 4}
    with pytest.raises(G.GLHError):
        G.flag_refine_coarsen(eta, 0.7, 0.5)


@pytest.mark.parametrize("deg", [1, 2])
def test_coarsening_a_uniform_refinement_restores_the_roots(deg):
    f0 = G.Forest((3, 2, 2), LO, HI, SIDES, degree=deg)
    m0 = f0.mesh()
    f1 = f0.refine(np.ones(f0.n_cells, np.uint8))
    f2 = f1.refine(np.full(f1.n_cells, 2, np.uint8))
    m2 = f2.mesh()
    assert f2.n_cells == f0.n_cells and m2.n_nodes == m0.n_nodes
    assert np.array_equal(m2.arrays()["node_xyz"], m0.arrays()["node_xyz"])
    # an incomplete family stays: one child not flagged
    fl = np.full(f1.n_cells, 2, np.uint8)
    fl[0] = 0
    assert f1.refine(fl).n_cells == f0.n_cells - 1 + 8


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_coarsening_keeps_the_balance_and_transfers_by_injection(seed):
    f = adaptive_forest(G, (3, 3, 2), LO, HI, SIDES, 3, seed, frac=0.35)
    m = f.mesh()
    rng = np.random.default_rng(seed)
    xyz0, cn0, lo0, h0 = _cells(m)
    centre = lo0 + 0.5 * h0
    # coarsen everything in x < 0 (whole families there collapse), refine
    # five cells in x > 0
    fl = np.where(centre[:, 0] < 0.0, 2, 0).astype(np.uint8)
    fl[rng.choice(np.nonzero(centre[:, 0] > 0.0)[0], 5, replace=False)] = 1
    f2 = f.refine(fl)
    m2 = f2.mesh()
    assert f2.n_cells < f.n_cells  # some families did coarsen
    assert _touching_levels_ok(m2)
    # coarse nodes sit on nodes of the fine mesh: the transfer is an injection
    U = _continuous(m, rng.standard_normal(m.n_nodes * 18)).reshape(-1, 18)
    U2 = f.transfer(m, U.reshape(-1), f2, m2).reshape(-1, 18)
    xyz, xyz2 = m.arrays()["node_xyz"].reshape(-1, 3), m2.arrays()["node_xyz"].reshape(-1, 3)
    key = {tuple(np.round(x, 9)): i for i, x in enumerate(xyz)}
    hang2 = np.diff(m2.constraints()["hang_ptr"]) if m2.constraints()["n_hanging"] else np.zeros(m2.n_nodes)
    checked = 0
    for v in range(m2.n_nodes):
        k = key.get(tuple(np.round(xyz2[v], 9)))
        if k is not None and not (len(hang2) and hang2[v] > 0):
            assert np.abs(U2[v] - U[k]).max() <= 1e-12 * np.abs(U).max()
            checked += 1
    assert checked > 0


@pytest.mark.parametrize("deg", [1, 2])
def test_a_family_next_to_a_refined_cell_keeps_its_children(deg):
    # coarsening root 0's family while its neighbour's child is refined would
    # put a level-0 cell next to level-2 cells: the family stays, and its
    # children keep their values
    f0 = G.Forest((3, 2, 2), LO, HI, SIDES, degree=deg)
    f1 = f0.refine(np.ones(f0.n_cells, np.uint8))
    m1 = f1.mesh()
    xyz1, cn1, lo1, h1 = _cells(m1)
    c = lo1 + 0.5 * h1
    w = (HI[0] - LO[0]) / 3.0
    fam = c[:, 0] < LO[0] + w  # the children of the roots in the first column
    fl = np.where(fam, 2, 0).astype(np.uint8)
    # a child of the next root on the shared face, in the corner of the domain
    # (so that it touches only one of the four families)
    cand = np.nonzero((c[:, 0] > LO[0] + w) & (c[:, 0] < LO[0] + 1.5 * w))[0]
    nb = cand[np.argmin(c[cand, 1] + c[cand, 2])]
    fl[nb] = 1
    f2 = f1.refine(fl)
    assert f2.n_cells == f1.n_cells + 7 - 7 * (4 - 1)  # 3 of the 4 families coarsen
    m2 = f2.mesh()
    assert _touching_levels_ok(m2)
    rng = np.random.default_rng(deg)
    U = _continuous(m1, rng.standard_normal(m1.n_nodes * 18)).reshape(-1, 18)
    U2 = f1.transfer(m1, U.reshape(-1), f2, m2).reshape(-1, 18)
    key = {tuple(np.round(x, 9)): i for i, x in enumerate(xyz1.reshape(-1, 3))}
    cons = m2.constraints()
    hang2 = np.diff(cons["hang_ptr"]) if cons["n_hanging"] else np.zeros(m2.n_nodes, int)
    xyz2 = m2.arrays()["node_xyz"].reshape(-1, 3)
    kept = 0
    for v in range(m2.n_nodes):
        k = key.get(tuple(np.round(xyz2[v], 9)))
        if k is not None and hang2[v] == 0:
            assert np.abs(U2[v] - U[k]).max() <= 1e-12 * np.abs(U).max()
            kept += 1
    assert kept > 0


@pytest.mark.parametrize("deg,sides,steps,seed", [(1, [1, 1, 1, 1, 0, 0], 2, 1), (2, [1, 1, 1, 1, 4, 4], 1, 2),
                                                  (1, [0, 0, 1, 1, 0, 0], 2, 3)])
def test_vectorised_oracle_lines_equal_the_independent_derivation(deg, sides, steps, seed):
    # the oracle mesh built from the library's own lines (used at full
    # size, where the independent derivation is too slow) distributes and
    # condenses exactly as the independent one on small forests
    lo, hi = (-6.0, -6.0, -4.0), (6.0, 6.0, 4.0)
    f = adaptive_forest(G, (3, 3, 2) if deg == 1 else (2, 2, 2), lo, hi, sides, steps, seed, degree=deg)
    m = f.mesh()
    om, ol = oracle_adaptive(m, lo, hi, sides), oracle_from_library(m)
    assert np.array_equal(om.con_of >= 0, ol.con_of >= 0)
    co = O.coef(22.0, 0.8)
    U = O.initial_guess(m.n_nodes, O.matep(22.0, 0.8)["gap_B"], amp=0.05, seed=seed)
    Ua, Ub = O.distribute(om, U), O.distribute(ol, U)
    assert np.array_equal(Ua, Ub)
    _, ra = O.assemble(om, co, Ua, J=False)
    _, rb = O.assemble(ol, co, Ub, J=False)
    assert np.abs(ra - rb).max() <= 1e-14 * np.abs(ra).max()
