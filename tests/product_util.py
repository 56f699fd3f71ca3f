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
    """Meshes of the benchmark configurations and small
    stand-ins used by the parity tests."""
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
