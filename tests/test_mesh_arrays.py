"""glh_mesh_from_arrays (the per-cycle upload from a caller's own arrays) on
the CPU: a mesh rebuilt from the arrays a deal.II caller would hold -- its own
node numbering (a random permutation here), FE_Q's hierarchic local order,
periodic images as constraint lines -- has the same Dirichlet masks, boundary
faces, node-pair pattern and constraint lines as the library's own mesh, under
that permutation; colours never share a node; malformed input is refused with
a message."""
import numpy as np
import pytest

from product_util import adaptive_forest, caller_arrays, glhip

G = glhip()


def _blocks(m):
    a = m.arrays()
    rp, col = a["row_ptr"], a["col"]
    return {(r, int(col[b])) for r in range(m.n_nodes) for b in range(rp[r], rp[r + 1])}


def _check_colours(m):
    a = m.arrays()
    cn = a["cell_nodes"].reshape(-1, m.ns)
    cp, cc = a["color_ptr"], a["color_cells"]
    assert sorted(cc.tolist()) == list(range(m.n_cells))
    for k in range(len(cp) - 1):
        nodes = cn[cc[cp[k]:cp[k + 1]]].reshape(-1)
        assert len(np.unique(nodes)) == nodes.size  # no node twice in one colour


@pytest.mark.parametrize("deg", [1, 2])
@pytest.mark.parametrize("hier", [False, True])
def test_box_mesh_rebuilt_from_permuted_arrays(deg, hier):
    lo, hi, sides = (-3.0, -2.0, -2.0), (3.0, 2.0, 2.0), [0, 1, 2, 3, 4, 1]
    m = G.Mesh(deg, (3, 2, 2), lo, hi, sides)
    kw, perm, _ = caller_arrays(m, lo, hi, seed=deg + 2 * hier, hierarchic=hier)
    m2 = G.Mesh.from_arrays(**kw)
    assert (m2.deg, m2.n_cells, m2.n_nodes, m2.n_faces) == (m.deg, m.n_cells, m.n_nodes, m.n_faces)
    a, a2 = m.arrays(), m2.arrays()
    # masks derived from the faces' ids equal the box mesh's, node by node
    assert np.array_equal(a2["node_mask"][perm], a["node_mask"])
    assert np.array_equal(a2["node_xyz"].reshape(-1, 3)[perm], a["node_xyz"].reshape(-1, 3))
    # cell e keeps its nodes (in lexicographic order inside the library)
    assert np.array_equal(a2["cell_nodes"], perm[a["cell_nodes"]])
    assert np.array_equal(a2["faces"], a["faces"])
    assert _blocks(m2) == {(int(perm[r]), int(perm[c])) for r, c in _blocks(m)}
    _check_colours(m2)
    # the same mesh with the caller's masks instead of derived ones
    m3 = G.Mesh.from_arrays(node_mask=a2["node_mask"], **kw)
    assert np.array_equal(m3.arrays()["node_mask"], a2["node_mask"])
    # and without coordinates: the library places the nodes itself
    kw2 = dict(kw, node_xyz=None)
    m4 = G.Mesh.from_arrays(**kw2)
    assert np.allclose(m4.arrays()["node_xyz"], a2["node_xyz"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("deg,steps,seed", [(1, 2, 1), (2, 1, 2)])
def test_periodic_adaptive_forest_rebuilt_from_arrays(deg, steps, seed):
    lo, hi, sides = (-6.0, -6.0, -4.0), (6.0, 6.0, 4.0), [5, 5, 5, 5, 0, 0]
    f = adaptive_forest(G, (3, 3, 2) if deg == 1 else (2, 2, 2), lo, hi, sides, steps, seed, degree=deg)
    m = f.mesh()
    con = m.constraints()
    assert con["n_hanging"] > 0
    kw, perm, _ = caller_arrays(m, lo, hi, seed=7, hierarchic=True)
    n_img = len(perm) - m.n_nodes
    assert n_img > 0  # the + faces of the two periodic axes
    m2 = G.Mesh.from_arrays(**kw)
    assert m2.n_nodes == m.n_nodes + n_img
    c2 = m2.constraints()
    hp, hm, hw = con["hang_ptr"], con["hang_m"], con["hang_w"]
    hp2, hm2, hw2 = c2["hang_ptr"], c2["hang_m"], c2["hang_w"]

    def line2(u):
        return {int(x): w for x, w in zip(hm2[hp2[u]:hp2[u + 1]], hw2[hp2[u]:hp2[u + 1]])}

    # hanging lines are the library's own, renumbered
    for v in range(m.n_nodes):
        ref = {int(perm[x]): w for x, w in zip(hm[hp[v]:hp[v + 1]], hw[hp[v]:hp[v + 1]])}
        assert line2(int(perm[v])) == pytest.approx(ref)
    # every image is closed onto unconstrained masters: its - face node, or
    # that node's own line when it hangs
    lines_in = {ln[0]: ln for ln in kw["lines"]}
    for u in range(m.n_nodes, len(perm)):
        pu = int(perm[u])
        (tgt,) = lines_in[pu][1]
        ref = line2(tgt) if hp2[tgt + 1] > hp2[tgt] else {tgt: 1.0}
        assert line2(pu) == pytest.approx(ref)
        assert all(hp2[x + 1] == hp2[x] for x in line2(pu))
    # the library's Dirichlet masks on the unconstrained nodes
    mk, mk2 = m.arrays()["node_mask"], m2.arrays()["node_mask"]
    free = np.array([hp[v + 1] == hp[v] for v in range(m.n_nodes)])
    assert np.array_equal(mk2[perm[:m.n_nodes]][free], mk[free])
    _check_colours(m2)


def _box_kw():
    lo, hi = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    m = G.Mesh(2, (2, 1, 1), lo, hi, [1, 1, 1, 1, 0, 0])
    kw, _, _ = caller_arrays(m, lo, hi, seed=0)
    return kw


def test_wrong_local_order_is_refused():
    kw = _box_kw()
    with pytest.raises(G.GLHError, match="support point"):
        G.Mesh.from_arrays(**dict(kw, order=G.GLH_ORDER_HIERARCHIC))


def test_malformed_arrays_are_refused():
    kw = _box_kw()
    bad = kw["cell_nodes"].copy()
    bad[0, 0] = len(kw["node_xyz"])
    with pytest.raises(G.GLHError, match="out of range"):
        G.Mesh.from_arrays(**dict(kw, cell_nodes=bad))
    bad = kw["cell_nodes"].copy()
    bad[0, 1] = bad[0, 0]
    with pytest.raises(G.GLHError, match="repeated node"):
        G.Mesh.from_arrays(**dict(kw, cell_nodes=bad))
    n0, n1 = int(kw["cell_nodes"][0, 0]), int(kw["cell_nodes"][0, 1])
    with pytest.raises(G.GLHError, match="cyclic"):
        G.Mesh.from_arrays(**dict(kw, lines=[(n0, [n1], [1.0]), (n1, [n0], [1.0])]))
    with pytest.raises(G.GLHError, match="two constraint lines"):
        G.Mesh.from_arrays(**dict(kw, lines=[(n0, [n1], [1.0]), (n0, [n1], [0.5])]))
    h = np.array(kw["cell_h"], copy=True)
    h[1, 2] = 0.0
    with pytest.raises(G.GLHError, match="positive"):
        G.Mesh.from_arrays(**dict(kw, cell_h=h))
