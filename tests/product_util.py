"""Helpers shared by the product tests: load the library binding, convert the
product's host mesh into the oracle's plain description, build inputs."""
import importlib.util
import os

import numpy as np

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "verkko-hem-repo_amd")


def glhip():
    spec = importlib.util.spec_from_file_location("glhip", os.path.join(PKG, "glhip.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def oracle_mesh(mesh):
    """glhip.Mesh -> oracle_lib.Mesh (Dirichlet masks become constraint lines
    without masters)."""
    a = mesh.arrays()
    mask = a["node_mask"]
    con_of = -np.ones(mesh.n_nodes * 18, np.int32)
    bits = ((mask[:, None] >> np.arange(18, dtype=np.uint32)[None, :]) & 1).astype(bool).reshape(-1)
    cdofs = np.nonzero(bits)[0]
    con_of[cdofs] = np.arange(cdofs.size, dtype=np.int32)
    con_ptr = np.zeros(cdofs.size + 1, np.int32)
    faces = a["faces"].reshape(-1, 3) if mesh.n_faces else np.zeros((0, 3), np.int32)
    return O.Mesh(mesh.deg, a["cell_nodes"], a["cell_h"], [tuple(f) for f in faces], mesh.n_nodes, con_of, con_ptr,
                  [], [], a["row_ptr"], a["col"], a["node_xyz"].reshape(-1, 3))


def cfg_mesh(G, name):
    """Meshes of the benchmark workloads (8^3 Q1 cube, 32^3 Q2 cube) and
    small stand-ins used by the parity tests."""
    Z = [1, 1, 1, 1, 0, 0]  # A = 0 on the two z faces, natural elsewhere
    if name == "cube_q1_ref3":      # [-10,10]^3, Q1, 8^3 cells
        return G.Mesh(1, (8, 8, 8), (-10, -10, -10), (10, 10, 10), Z)
    if name == "cube_q2_ref5":      # [-20,20]^3, Q2, 32^3 cells
        return G.Mesh(2, (32, 32, 32), (-20, -20, -20), (20, 20, 20), Z)
    if name == "cube_q2_small":     # same physics, 4^3 cells
        return G.Mesh(2, (4, 4, 4), (-20, -20, -20), (20, 20, 20), Z)
    if name == "cube_q1_small":
        return G.Mesh(1, (3, 3, 3), (-3, -3, -3), (3, 3, 3), Z)
    raise KeyError(name)


def initial_state(mesh_arrays_n_nodes, p, t, amp, seed, phase=0):
    m = O.matep(p, t)
    gap = m["gap_B"] if phase == 0 else m["gap_A"]
    return O.initial_guess(mesh_arrays_n_nodes, gap, phase=phase, amp=amp, seed=seed)


def side_masks(side_id):
    """18-bit Dirichlet component masks of the six box faces (femgl.h:294-301):
    id 0 all components, 2/3/4 the x/y/z-normal orbital column."""
    out = []
    for s in side_id:
        if s == 0:
            out.append((1 << 18) - 1)
        elif s in (2, 3, 4):
            out.append(sum(1 << c for c in range(18) if (c % 9) % 3 == s - 2))
        else:
            out.append(0)
    return out


def adaptive_forest(G, n, lo, hi, side_id, steps, seed=0, frac=0.25, degree=1):
    """Forest refined `steps` times, each time a seeded random fraction of its
    cells (with 2:1 balance applied by the library)."""
    rng = np.random.default_rng(seed)
    f = G.Forest(n, lo, hi, side_id, degree=degree)
    for _ in range(steps):
        fl = (rng.random(f.n_cells) < frac).astype(np.uint8)
        f = f.refine(fl)
    return f


def oracle_adaptive(mesh, lo, hi, side_id):
    period = tuple((hi[d] - lo[d]) if side_id[2 * d] == 5 and side_id[2 * d + 1] == 5 else 0.0 for d in range(3))
    """glhip adaptive Mesh -> oracle_lib.Mesh with independently derived
    hanging-node and Dirichlet constraints (deal.II order and closure)."""
    a = mesh.arrays()
    xyz = a["node_xyz"].reshape(-1, 3)
    cn = a["cell_nodes"].reshape(-1, (mesh.deg + 1) ** 3)
    cell_lo = xyz[cn[:, 0]]
    faces = a["faces"].reshape(-1, 3) if mesh.n_faces else np.zeros((0, 3), np.int32)
    return O.adaptive_oracle_mesh(a["cell_nodes"], xyz, cell_lo, a["cell_h"], [tuple(f) for f in faces], lo, hi,
                                  side_masks(side_id), deg=mesh.deg, period=period)
