"""CPU-side checks of the product library: it loads, exports every symbol of
include/glhip.h, its matep restatement agrees bit for bit with the oracle,
and its host mesh layer produces the expected structure (no GPU needed)."""
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from product_util import ROOT, cfg_mesh, glhip, oracle_mesh

G = glhip()


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "glhip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(glh_\w+)\(", hdr, re.M))
    assert declared, "no declarations found"
    assert declared == set(G.EXPORTS)
    L = G.lib()
    for name in declared:
        assert hasattr(L, name), name


@pytest.mark.parametrize("scc", [True, False])
@pytest.mark.parametrize("p,t", [(20.0, 0.5), (26.0, 0.7), (22.0, 0.8), (0.0, 0.3), (33.99, 0.95), (13.7, 0.41)])
def test_matep_bitwise_equal_to_oracle(p, t, scc):
    rc, m = G.matep(p, t, scc)
    assert rc == 0
    o = O.matep(p, t, scc)
    keys = ["alpha", "beta1", "beta2", "beta3", "beta4", "beta5", "beta_A", "beta_B", "gap_A", "gap_B", "f_A", "f_B"]
    for i, k in enumerate(keys):
        assert m[i] == o[k], k


def test_matep_range_error():
    assert G.matep(34.0, 0.5)[0] == -6
    assert G.matep(-1.0, 0.5)[0] == -6
    assert G.matep(40.0, 0.5, scc=False)[0] == 0


@pytest.mark.parametrize("deg,n", [(1, (3, 2, 4)), (2, (2, 3, 2))])
def test_box_mesh_matches_independent_construction(deg, n):
    lo, hi = (-1.0, -2.0, 0.0), (3.0, 1.0, 2.0)
    m = G.Mesh(deg, n, lo, hi, [1, 1, 1, 1, 0, 0])
    ref = O.box_mesh(deg, n, lo, hi, dirichlet_faces=(4, 5))
    a = m.arrays()
    assert m.n_nodes == ref.n_nodes and m.n_cells == ref.n_cells
    assert np.array_equal(a["cell_nodes"], ref.cell_nodes)
    assert np.allclose(a["cell_h"], ref.cell_h, rtol=1e-15)
    assert np.allclose(a["node_xyz"].reshape(-1, 3), ref.node_xyz, atol=1e-14)
    assert np.array_equal(a["row_ptr"], ref.row_ptr) and np.array_equal(a["col"], ref.col)
    om = oracle_mesh(m)
    assert np.array_equal(om.con_of >= 0, ref.con_of >= 0)


@pytest.mark.parametrize("deg", [1, 2])
def test_colouring_is_conflict_free_and_complete(deg):
    m = G.Mesh(deg, (5, 4, 3), (0, 0, 0), (1, 1, 1), [1] * 6)
    a = m.arrays()
    cells = a["cell_nodes"].reshape(m.n_cells, m.ns)
    seen = np.zeros(m.n_cells, int)
    for k in range(m.n_colors):
        ids = a["color_cells"][a["color_ptr"][k]:a["color_ptr"][k + 1]]
        seen[ids] += 1
        nodes = cells[ids].reshape(-1)
        assert np.unique(nodes).size == nodes.size  # no shared node inside a colour
    assert np.all(seen == 1)


def test_boundary_ids_and_masks():
    # AdGR z walls: only the z-normal orbital column (components 2,5,8,11,14,17)
    m = G.Mesh(1, (2, 2, 2), (0, 0, 0), (1, 1, 1), [1, 1, 1, 1, 4, 4])
    a = m.arrays()
    xyz = a["node_xyz"].reshape(-1, 3)
    zmask = sum(1 << c for c in (2, 5, 8, 11, 14, 17))
    on = (xyz[:, 2] == 0) | (xyz[:, 2] == 1)
    assert np.all(a["node_mask"][on] == zmask) and np.all(a["node_mask"][~on] == 0)
    f = a["faces"].reshape(-1, 3)
    assert len(f) == 8 and set(f[:, 2]) == {4} and set(f[:, 1]) == {4, 5}
    # workload sizes: Q1 ref 3 -> 729 nodes, Q2 ref 5 -> 274625 nodes
    m1 = cfg_mesh(G, "cube_q1_ref3")
    assert m1.n_nodes * 18 == 13122 and m1.nnzb * 324 == 5_062_500


def test_mesh_argument_errors():
    with pytest.raises(G.GLHError):
        G.Mesh(3, (2, 2, 2), (0, 0, 0), (1, 1, 1), [1] * 6)
    with pytest.raises(G.GLHError):
        G.Mesh(1, (0, 2, 2), (0, 0, 0), (1, 1, 1), [1] * 6)


def test_solver_option_argument_errors():
    # without a context the setters refuse before any device work (the
    # ratio's own range is checked on a live context in the GPU tests)
    L = G.lib()
    for r in (0.5, -0.1, float("nan"), float("inf")):
        assert L.glh_set_reorth(None, r) == -1
    assert L.glh_set_schedule(None, 100, 0, 0, 0, 8) == -1


@pytest.mark.parametrize("phase", [0, 1])
def test_initial_guess_bitwise_equal_to_oracle(phase):
    gap = O.matep(20.0, 0.5)["gap_B" if phase == 0 else "gap_A"]
    for seed in (0, 1, 12345):
        a = G.initial_guess(97, gap, phase=phase, amp=0.05, seed=seed)
        b = O.initial_guess(97, gap, phase=phase, amp=0.05, seed=seed)
        assert np.array_equal(a, b)


def test_bna_guess_bitwise_equal_to_oracle_and_phases():
    m = O.matep(22.0, 0.8)
    xyz = np.random.default_rng(3).uniform(-6, 6, (211, 3))
    a = G.initial_guess_bna(xyz, 0.0, m["gap_A"], m["gap_B"], amp=0.05, seed=1)
    b = O.initial_guess_bna(xyz, 0.0, m["gap_A"], m["gap_B"], amp=0.05, seed=1)
    assert np.array_equal(a, b)
    # without noise: B phase above the wall, A phase below (femgl/inc/BnA.h)
    u = G.initial_guess_bna(xyz, 0.0, m["gap_A"], m["gap_B"], amp=0.0).reshape(-1, 18)
    up = xyz[:, 2] >= 0
    assert np.allclose(u[up][:, [0, 4, 8]], m["gap_B"] / np.sqrt(3), rtol=1e-7)
    assert np.allclose(u[~up][:, [0, 10]], m["gap_A"] / np.sqrt(2), rtol=1e-7)
    assert np.count_nonzero(u[up]) == 3 * up.sum() and np.count_nonzero(u[~up]) == 2 * (~up).sum()


@pytest.mark.parametrize("deg", [1, 2])
def test_vtu_output_structure(tmp_path, deg):
    import xml.etree.ElementTree as ET
    m = G.Mesh(deg, (2, 2, 1), (-1, -1, -1), (1, 1, 1), [1, 1, 1, 1, 0, 0])
    U = np.arange(m.n_nodes * 18, dtype=np.float64)
    m.write_vtu(U, tmp_path, counter=3, dU=-U)
    root = ET.parse(tmp_path / "solution_03.0.vtu").getroot()
    piece = root.find("UnstructuredGrid/Piece")
    n_hex = 4 * (8 if deg == 2 else 1)
    assert int(piece.get("NumberOfPoints")) == m.n_nodes and int(piece.get("NumberOfCells")) == n_hex
    arrays = {a.get("Name"): a for a in piece.find("PointData")}
    names = ["%s_%d%d" % (p, i, j) for p in ("u", "v") for i in (1, 2, 3) for j in (1, 2, 3)]
    assert all(n in arrays for n in names) and all("d" + n in arrays for n in names)
    # component c of node v is U[v*18 + c]: u_12 is component 1, v_31 is 15
    assert np.array_equal(np.array(arrays["u_12"].text.split(), float), U.reshape(-1, 18)[:, 1])
    assert np.array_equal(np.array(arrays["v_31"].text.split(), float), U.reshape(-1, 18)[:, 15])
    # hexahedra are positively oriented (VTK order) and cover the box
    pts = np.array(piece.find("Points/DataArray").text.split(), float).reshape(-1, 3)
    con = np.array(piece.find("Cells/DataArray").text.split(), int).reshape(-1, 8)
    vol = 0.0
    for h in con:
        p = pts[h]
        vol += np.linalg.det(np.array([p[1] - p[0], p[3] - p[0], p[4] - p[0]]))
    assert vol == pytest.approx(8.0, rel=1e-12)
    assert (tmp_path / "solution_03.pvtu").exists()
