"""ctypes binding of the FP64 CPU oracle (oracle/libglorc.so).

Test infrastructure only: the oracle is the checker, never the product path.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "libglorc.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/libglorc.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.orc_matep.argtypes = [C.c_double, C.c_double, C.c_int, _dp]
        L.orc_point_hessian_literal.argtypes = [_dp, _dp, _dp]
        L.orc_point_rho_literal.argtypes = [_dp, _dp, _dp]
        L.orc_point_hessian.argtypes = [_dp, _dp, _dp]
        L.orc_point_rho.argtypes = [_dp, _dp, _dp]
        L.orc_point_energy.argtypes = [_dp, _dp, _dp]
        L.orc_point_energy.restype = C.c_double
        L.orc_lit_jac_point.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_lit_jac_point.restype = C.c_double
        L.orc_lit_res_point.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.orc_lit_res_point.restype = C.c_double
        L.orc_refel.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_cell_faithful.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_cell_factorized.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_face_faithful.argtypes = [C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_assemble.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, _ip, C.c_int, _ip, _ip, _ip, _dp,
                                   _dp, _dp, C.c_int, _ip, _ip, _dp, _dp]
        L.orc_free_energy.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, _ip, _dp, _dp]
        L.orc_free_energy.restype = C.c_double
        L.orc_bsr_spmv.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.orc_block_jacobi.argtypes = [C.c_int, _ip, _ip, _dp, _dp]
        L.orc_gmres.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int,
                                C.POINTER(C.c_double)]
        L.orc_distribute.argtypes = [C.c_int, _ip, _ip, _ip, _dp, _dp]
        L.orc_mt64.argtypes = [C.c_uint64, C.c_int, _up]
        L.orc_initial_guess.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_uint64, _dp]
        L.orc_initial_guess_bna.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_uint64, _dp]
        L.orc_newton.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, _ip, C.c_int, _ip, _ip, _ip, _dp, _dp,
                                 _dp, _ip, _ip, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                 C.c_double, C.c_int, _dp]
        _u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        L.orc_cells_apply.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, _ip, C.c_int, _u32p, _dp, _dp, C.c_int,
                                      _ip, _ip, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def matep(p, t, scc=True):
    out = np.zeros(12)
    rc = lib().orc_matep(float(p), float(t), int(scc), out)
    keys = ["alpha", "beta1", "beta2", "beta3", "beta4", "beta5", "beta_A", "beta_B", "gap_A", "gap_B",
            "f_A", "f_B"]
    d = dict(zip(keys, out.tolist()))
    d["rc"] = rc
    return d


K_CONST = 0.42072  # femgl/inc/femgl.h:320-322


def coef(p, t, scc=True, bt=1.0e10, K=(K_CONST, K_CONST, K_CONST)):
    m = matep(p, t, scc)
    return f64([m["alpha"], m["beta1"], m["beta2"], m["beta3"], m["beta4"], m["beta5"], K[0], K[1], K[2], bt])


def point_hessian(co, x, literal=False):
    H = np.zeros(324)
    (lib().orc_point_hessian_literal if literal else lib().orc_point_hessian)(f64(co), f64(x), H)
    return H.reshape(18, 18)


def point_rho(co, x, literal=False):
    r = np.zeros(18)
    (lib().orc_point_rho_literal if literal else lib().orc_point_rho)(f64(co), f64(x), r)
    return r


def point_energy(co, x, gx):
    return lib().orc_point_energy(f64(co), f64(x), f64(gx))


def refel(deg):
    n = (deg + 1) ** 3
    N = np.zeros(n * n)
    dN = np.zeros(n * n * 3)
    w = np.zeros(n)
    lib().orc_refel(deg, N, dN, w)
    return N.reshape(n, n), dN.reshape(n, n, 3), w


def cell(deg, h, co, Uloc, faithful=False, face=None):
    ns = (deg + 1) ** 3
    dpc = 18 * ns
    J = np.zeros(dpc * dpc)
    R = np.zeros(dpc)
    fn = lib().orc_cell_faithful if faithful else lib().orc_cell_factorized
    rc = fn(deg, f64(h), f64(co), f64(Uloc), J, R)
    assert rc == 0
    if face is not None:
        for (lf, bid) in face:
            lib().orc_face_faithful(deg, f64(h), lf, bid, f64(co), f64(Uloc), J, R)
    return J.reshape(dpc, dpc), R


class Mesh:
    """Plain description consumed by the oracle (and produced by the product's
    host mesh layer): lexicographic cell->node map, cell sizes, boundary
    faces (cell, local face, id), dof constraints and a BSR node pattern."""

    def __init__(self, deg, cell_nodes, cell_h, faces, n_nodes, con_of, con_ptr, con_master, con_weight,
                 row_ptr, col, node_xyz=None):
        self.deg = deg
        self.cell_nodes = i32(cell_nodes)
        self.cell_h = f64(cell_h)
        self.faces = i32(faces).reshape(-1) if len(faces) else np.zeros(0, np.int32)
        self.n_faces = len(self.faces) // 3
        self.n_cells = self.cell_nodes.size // ((deg + 1) ** 3)
        self.n_nodes = n_nodes
        self.con_of = i32(con_of)
        self.con_ptr = i32(con_ptr)
        self.con_master = i32(con_master) if len(con_master) else np.zeros(1, np.int32)
        self.con_weight = f64(con_weight) if len(con_weight) else np.zeros(1)
        self.row_ptr = i32(row_ptr)
        self.col = i32(col)
        self.node_xyz = node_xyz

    @property
    def n_dofs(self):
        return self.n_nodes * 18

    @property
    def nnzb(self):
        return int(self.row_ptr[-1])


def box_mesh(deg, n, lo, hi, dirichlet_faces=(4, 5), bids=None):
    """Independent (test-side) structured box mesh: n=(nx,ny,nz) cells,
    lexicographic node numbering, homogeneous Dirichlet on all 18 components
    of the nodes on the listed faces (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z)."""
    nx, ny, nz = n
    n1 = deg + 1
    Nx, Ny, Nz = deg * nx + 1, deg * ny + 1, deg * nz + 1
    h = [(hi[k] - lo[k]) / n[k] for k in range(3)]
    cells = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                cn = []
                for c in range(n1):
                    for b in range(n1):
                        for a in range(n1):
                            cn.append((deg * i + a) + Nx * ((deg * j + b) + Ny * (deg * k + c)))
                cells.append(cn)
    cells = np.array(cells, np.int32)
    n_nodes = Nx * Ny * Nz
    idx = np.arange(n_nodes)
    ix, iy, iz = idx % Nx, (idx // Nx) % Ny, idx // (Nx * Ny)
    xyz = np.stack([lo[0] + ix * h[0] / deg, lo[1] + iy * h[1] / deg, lo[2] + iz * h[2] / deg], 1)
    on = [ix == 0, ix == Nx - 1, iy == 0, iy == Ny - 1, iz == 0, iz == Nz - 1]
    cons = np.zeros(n_nodes, bool)
    for f in dirichlet_faces:
        cons |= on[f]
    con_of = -np.ones(n_nodes * 18, np.int32)
    cdofs = (np.nonzero(cons)[0][:, None] * 18 + np.arange(18)[None, :]).reshape(-1)
    con_of[cdofs] = np.arange(cdofs.size, dtype=np.int32)
    con_ptr = np.zeros(cdofs.size + 1, np.int32)
    # node pattern: nodes sharing a cell
    nb = [set() for _ in range(n_nodes)]
    for cn in cells:
        for a in cn:
            nb[a].update(cn.tolist())
    row_ptr = np.zeros(n_nodes + 1, np.int32)
    cols = []
    for r in range(n_nodes):
        s = sorted(nb[r])
        cols.extend(s)
        row_ptr[r + 1] = row_ptr[r] + len(s)
    faces = []
    if bids:
        for e in range(cells.shape[0]):
            i, j, k = e % nx, (e // nx) % ny, e // (nx * ny)
            for (lf, cond) in enumerate([i == 0, i == nx - 1, j == 0, j == ny - 1, k == 0, k == nz - 1]):
                if cond and lf in bids:
                    faces.append((e, lf, bids[lf]))
    return Mesh(deg, cells.reshape(-1), np.tile(h, (cells.shape[0], 1)).reshape(-1), faces, n_nodes, con_of,
                con_ptr, [], [], row_ptr, np.array(cols, np.int32), xyz)


def assemble(mesh, co, U, J=True, R=True, faithful=False):
    val = np.zeros(max(1, mesh.nnzb * 324)) if J else np.zeros(1)
    rhs = np.zeros(mesh.n_dofs)
    flags = (1 if J else 0) | (2 if R else 0) | (4 if faithful else 0)
    rc = lib().orc_assemble(mesh.deg, mesh.n_cells, mesh.cell_nodes, mesh.cell_h, mesh.n_faces,
                            mesh.faces if mesh.n_faces else np.zeros(3, np.int32), mesh.n_nodes, mesh.con_of,
                            mesh.con_ptr, mesh.con_master, mesh.con_weight, f64(co), f64(U), flags, mesh.row_ptr,
                            mesh.col, val, rhs)
    assert rc == 0, rc
    return (val.reshape(-1, 18, 18) if J else None), rhs


def cells_apply(deg, arrays, n_nodes, co, U, X=(), rhs=True):
    """system_rhs = -R and J x for each x in X cell by cell without a
    global J (orc_cells_apply): meshes with Dirichlet masks only.  arrays:
    the product mesh's arrays (cell_nodes, cell_h, faces, node_mask,
    color_ptr, color_cells).  Returns (rhs or None, [J x ...])."""
    a = arrays
    nd = n_nodes * 18
    Xs = np.ascontiguousarray(np.asarray(X, np.float64).reshape(-1, nd)) if len(X) else np.zeros((0, nd))
    Y = np.zeros_like(Xs)
    r = np.zeros(nd) if rhs else None
    faces = np.ascontiguousarray(a["faces"], np.int32)
    nc = len(a["cell_nodes"]) // (deg + 1) ** 3
    ptr = lambda v: None if v is None or v.size == 0 else v.ctypes.data_as(C.c_void_p)
    rc = lib().orc_cells_apply(deg, nc, i32(a["cell_nodes"]), f64(a["cell_h"]), len(faces) // 3,
                               faces if len(faces) else np.zeros(3, np.int32), n_nodes,
                               np.ascontiguousarray(a["node_mask"], np.uint32), f64(co), f64(U),
                               len(a["color_ptr"]) - 1, i32(a["color_ptr"]), i32(a["color_cells"]), len(Xs),
                               ptr(Xs), ptr(Y), ptr(r))
    assert rc == 0, rc
    return r, list(Y)


def free_energy(mesh, co, U):
    return lib().orc_free_energy(mesh.deg, mesh.n_cells, mesh.cell_nodes, mesh.cell_h, mesh.n_faces,
                                 mesh.faces if mesh.n_faces else np.zeros(3, np.int32), f64(co), f64(U))


def bsr_to_dense(mesh, val):
    n = mesh.n_dofs
    A = np.zeros((n, n))
    for r in range(mesh.n_nodes):
        for b in range(mesh.row_ptr[r], mesh.row_ptr[r + 1]):
            s = mesh.col[b]
            A[18 * r:18 * r + 18, 18 * s:18 * s + 18] = val[b]
    return A


def gmres(mesh, val, b, rtol=1e-10, maxit=1000, restart=30):
    x = np.zeros(mesh.n_dofs)
    res = C.c_double(0)
    its = lib().orc_gmres(mesh.n_nodes, mesh.row_ptr, mesh.col, f64(val.reshape(-1)), f64(b), x, rtol, maxit,
                          restart, C.byref(res))
    return x, its, res.value


def distribute(mesh, x):
    x = f64(x).copy()
    lib().orc_distribute(mesh.n_dofs, mesh.con_of, mesh.con_ptr, mesh.con_master, mesh.con_weight, x)
    return x


def mt64(seed, n):
    out = np.zeros(n, np.uint64)
    lib().orc_mt64(seed, n, out)
    return out


def initial_guess(n_nodes, gap, phase=0, amp=0.05, seed=0):
    U = np.zeros(n_nodes * 18)
    lib().orc_initial_guess(n_nodes, gap, phase, amp, seed, U)
    return U


def initial_guess_bna(xyz, z_a, gap_A, gap_B, amp=0.05, seed=1):
    xyz = f64(xyz).reshape(-1)
    n = xyz.size // 3
    U = np.zeros(n * 18)
    lib().orc_initial_guess_bna(n, xyz, z_a, gap_A, gap_B, amp, seed, U)
    return U


def newton(mesh, co, U, rtol=1e-10, max_newton=50, lin_tol=0.1, lin_maxit=5000, restart=50, ls_step=0.83,
           damped=True):
    U = f64(U).copy()
    stats = np.zeros(4)
    rc = lib().orc_newton(mesh.deg, mesh.n_cells, mesh.cell_nodes, mesh.cell_h, mesh.n_faces,
                          mesh.faces if mesh.n_faces else np.zeros(3, np.int32), mesh.n_nodes, mesh.con_of,
                          mesh.con_ptr, mesh.con_master, mesh.con_weight, f64(co), U, mesh.row_ptr, mesh.col,
                          rtol, max_newton, lin_tol, lin_maxit, restart, ls_step, int(damped), stats)
    return U, rc, stats


# ---------------------------------------------------------------- adaptive
def lagrange_1d(deg, i, t):
    """Lagrange basis function i of degree deg on the equidistant nodes of [0, 1]."""
    nodes = [k / deg for k in range(deg + 1)]
    out = 1.0
    for k in range(deg + 1):
        if k != i:
            out *= (t - nodes[k]) / (nodes[i] - nodes[k])
    return out


def hanging_lines(cell_nodes, xyz, cell_lo, cell_h, tol=1e-9, deg=1, period=(0.0, 0.0, 0.0)):
    """Independent (test-side) hanging-node detection for Q1/Q2 hex meshes: a
    node on the closure of a cell that the cell does not list follows that
    cell's Q_deg interpolant (the full tensor-product Lagrange basis evaluated
    at the node).  period[d] > 0: nodes on the - face of periodic axis d
    also stand for the + face (they are also tested one period up).  Lines
    are closed (masters that hang themselves are expanded).  Returns
    {node: {master: weight}}."""
    n1 = deg + 1
    ns = n1 ** 3
    cells = np.asarray(cell_nodes).reshape(-1, ns)
    lo = np.asarray(cell_lo).reshape(-1, 3)
    h = np.asarray(cell_h).reshape(-1, 3)
    xyz = np.asarray(xyz).reshape(-1, 3)
    raw = {}
    shifts = [np.array([sx, sy, sz]) for sx in {0.0, period[0]} for sy in {0.0, period[1]}
              for sz in {0.0, period[2]}]
    for e in range(cells.shape[0]):
        for sh in shifts:
            t = (xyz + sh - lo[e]) / h[e]              # node positions in cell coordinates
            inside = np.all((t > -tol) & (t < 1 + tol), 1)
            inside[cells[e]] = False
            for v in np.nonzero(inside)[0]:
                if v in raw:
                    continue
                raw[int(v)] = _interp_line(cells[e], np.clip(t[v], 0.0, 1.0), deg, ns, n1)
    return _close_lines(raw)


def _interp_line(cell, tv, deg, ns, n1):
    line = {}
    for a in range(ns):
        idx = (a % n1, (a // n1) % n1, a // (n1 * n1))
        w = 1.0
        for d in range(3):
            w *= lagrange_1d(deg, idx[d], tv[d])
        if abs(w) > 1e-14:
            line[int(cell[a])] = line.get(int(cell[a]), 0.0) + w
    return line


def _close_lines(raw):
    closed = {}
    for v in raw:
        line = dict(raw[v])
        for _ in range(32):
            nxt = {}
            changed = False
            for m, w in line.items():
                if m in raw:
                    changed = True
                    for m2, w2 in raw[m].items():
                        nxt[m2] = nxt.get(m2, 0.0) + w * w2
                else:
                    nxt[m] = nxt.get(m, 0.0) + w
            line = nxt
            if not changed:
                break
        closed[v] = line
    return closed


def adaptive_oracle_mesh(cell_nodes, xyz, cell_lo, cell_h, faces, box_lo, box_hi, side_masks, deg=1,
                         period=(0.0, 0.0, 0.0)):
    """Oracle description of an adaptive Q1 mesh with deal.II constraint
    semantics: hanging-node lines first (make_hanging_node_constraints), then
    homogeneous Dirichlet lines for boundary dofs not yet constrained
    (interpolate_boundary_values), then close (constrained masters replaced
    by their lines; Dirichlet masters drop out).  side_masks[6]: 18-bit
    component masks of the faces -x,+x,-y,+y,-z,+z."""
    xyz = np.asarray(xyz).reshape(-1, 3)
    n_nodes = xyz.shape[0]
    hang = hanging_lines(cell_nodes, xyz, cell_lo, cell_h, deg=deg, period=period)
    lines = {}                                      # dof -> {master dof: w}
    for v, line in hang.items():
        for c in range(18):
            lines[v * 18 + c] = {m * 18 + c: w for m, w in line.items()}
    tol = 1e-9 * max(abs(b - a) for a, b in zip(box_lo, box_hi))
    for v in range(n_nodes):
        for s in range(6):
            d = s // 2
            on = abs(xyz[v, d] - (box_lo[d] if s % 2 == 0 else box_hi[d])) < tol
            if not on:
                continue
            for c in range(18):
                if (side_masks[s] >> c) & 1 and v * 18 + c not in lines:
                    lines[v * 18 + c] = {}
    for g in list(lines):
        line = lines[g]
        for _ in range(32):
            nxt, changed = {}, False
            for m, w in line.items():
                if m in lines:
                    changed = True
                    for m2, w2 in lines[m].items():
                        nxt[m2] = nxt.get(m2, 0.0) + w * w2
                else:
                    nxt[m] = nxt.get(m, 0.0) + w
            line = nxt
            if not changed:
                break
        lines[g] = line
    n_dofs = n_nodes * 18
    con_of = -np.ones(n_dofs, np.int32)
    con_ptr, con_master, con_weight = [0], [], []
    for i, g in enumerate(sorted(lines)):
        con_of[g] = i
        for m, w in sorted(lines[g].items()):
            con_master.append(m)
            con_weight.append(w)
        con_ptr.append(len(con_master))
    # node pattern able to hold every condensed entry
    cells = np.asarray(cell_nodes).reshape(-1, (deg + 1) ** 3)
    tgt = {}
    for a in range(n_nodes):
        t = set()
        for c in range(18):
            g = a * 18 + c
            if g in lines:
                t.update(m // 18 for m in lines[g])
            else:
                t.add(a)
        tgt[a] = t
    nb = [set([r]) for r in range(n_nodes)]
    for cn in cells:
        for a in cn:
            for b in cn:
                for r in tgt[int(a)]:
                    nb[r].update(tgt[int(b)])
    row_ptr = np.zeros(n_nodes + 1, np.int32)
    cols = []
    for r in range(n_nodes):
        s = sorted(nb[r])
        cols.extend(s)
        row_ptr[r + 1] = row_ptr[r] + len(s)
    return Mesh(deg, np.asarray(cell_nodes).reshape(-1), np.asarray(cell_h).reshape(-1), faces, n_nodes, con_of,
                np.array(con_ptr, np.int32), con_master, con_weight, row_ptr, np.array(cols, np.int32), xyz)


def kelly(cell_nodes, cell_lo, cell_h, U, tol=1e-9, deg=1, period=(0.0, 0.0, 0.0)):
    """Test-side Kelly indicator (deal.II's documented formula, no Neumann
    data): eta_K^2 = sum_F h_K/24 int_F sum_c [du_c/dn]^2 over interior face
    pieces, integrated with (deg+1)^2 Gauss points (numpy's Gauss-Legendre)
    on the smaller face; gradients by central differences of the Q_deg
    interpolant (exact for polynomials of degree <= 2 per direction).
    period[d] > 0: axis d is periodic with that length (faces at the two ends
    of the box are neighbours)."""
    n1 = deg + 1
    ns = n1 ** 3
    cells = np.asarray(cell_nodes).reshape(-1, ns)
    lo = np.asarray(cell_lo).reshape(-1, 3)
    h = np.asarray(cell_h).reshape(-1, 3)
    U = np.asarray(U).reshape(-1, 18)
    n = cells.shape[0]
    e2 = np.zeros(n)
    gx, gwt = np.polynomial.legendre.leggauss(n1)
    g = [(0.5 * (x + 1.0), 0.5 * w) for x, w in zip(gx, gwt)]

    def value(e, t):
        out = np.zeros(18)
        for a in range(ns):
            idx = (a % n1, (a // n1) % n1, a // (n1 * n1))
            out += lagrange_1d(deg, idx[0], t[0]) * lagrange_1d(deg, idx[1], t[1]) * \
                lagrange_1d(deg, idx[2], t[2]) * U[cells[e][a]]
        return out

    def grad(e, x):
        t = (x - lo[e]) / h[e]
        out = np.zeros((3, 18))
        eps = 1e-3
        for d in range(3):
            tp, tm = t.copy(), t.copy()
            tp[d] += eps
            tm[d] -= eps
            out[d] = (value(e, tp) - value(e, tm)) / (2 * eps * h[e][d])
        return out

    hi = lo + h
    for e in range(n):
        for d in range(3):
            for sg in (0, 1):
                x_f = hi[e][d] if sg else lo[e][d]
                t1, t2 = (d + 1) % 3, (d + 2) % 3
                # neighbours sharing this face plane and overlapping it
                for k in range(n):
                    if k == e:
                        continue
                    x_k = lo[k][d] if sg else hi[k][d]
                    wrap = 0.0
                    if abs(x_k - x_f) > tol * max(1.0, abs(x_f)):
                        if period[d] <= 0:
                            continue
                        wrap = period[d] if sg else -period[d]
                        if abs(x_k + wrap - x_f) > tol * max(1.0, abs(x_f)):
                            continue
                    a1, b1 = max(lo[e][t1], lo[k][t1]), min(hi[e][t1], hi[k][t1])
                    a2, b2 = max(lo[e][t2], lo[k][t2]), min(hi[e][t2], hi[k][t2])
                    if b1 - a1 <= tol or b2 - a2 <= tol:
                        continue
                    # integrate each overlap piece once: from the smaller side,
                    # same-size faces from the lower cell
                    small_e = h[e][t1] * h[e][t2] < h[k][t1] * h[k][t2] - tol
                    same = abs(h[e][t1] * h[e][t2] - h[k][t1] * h[k][t2]) <= tol
                    if not (small_e or (same and sg == 1)):
                        continue
                    s = 0.0
                    for q1, w1 in g:
                        for q2, w2 in g:
                            x = np.zeros(3)
                            x[d] = x_f
                            x[t1] = a1 + q1 * (b1 - a1)
                            x[t2] = a2 + q2 * (b2 - a2)
                            xk = x.copy()
                            xk[d] -= wrap
                            jmp = grad(e, x)[d] - grad(k, xk)[d]
                            s += w1 * w2 * (b1 - a1) * (b2 - a2) * float(jmp @ jmp)
                    e2[e] += np.linalg.norm(h[e]) / 24.0 * s
                    e2[k] += np.linalg.norm(h[k]) / 24.0 * s
    return np.sqrt(e2)


def periodic_oracle_mesh(deg, n, lo, hi, side_masks, periodic_axes, faces=()):
    """Box mesh with periodic axes for the oracle, derived on the test side:
    Dirichlet lines first (masks of the non-periodic faces), then periodicity
    (a dof on a + face equals its image on the - face, skipped where the dof
    is already constrained), then closure (images that are constrained are
    replaced by their own lines)."""
    m = box_mesh(deg, n, lo, hi, dirichlet_faces=())
    xyz = m.node_xyz
    n_nodes = m.n_nodes
    tol = 1e-9 * max(abs(b - a) for a, b in zip(lo, hi))
    lines = {}
    for v in range(n_nodes):
        for s in range(6):
            d = s // 2
            if d in periodic_axes:
                continue
            if abs(xyz[v, d] - (lo[d] if s % 2 == 0 else hi[d])) < tol:
                for c in range(18):
                    if (side_masks[s] >> c) & 1:
                        lines[v * 18 + c] = {}
    key = {tuple(np.round(x / tol).astype(np.int64)): i for i, x in enumerate(xyz)}
    for v in range(n_nodes):
        img = xyz[v].copy()
        for d in periodic_axes:
            if abs(img[d] - hi[d]) < tol:
                img[d] = lo[d]
        w = key[tuple(np.round(img / tol).astype(np.int64))]
        if w == v:
            continue
        for c in range(18):
            if v * 18 + c not in lines:
                lines[v * 18 + c] = {w * 18 + c: 1.0}
    for g in list(lines):
        line = lines[g]
        for _ in range(8):
            nxt, changed = {}, False
            for mm, ww in line.items():
                if mm in lines:
                    changed = True
                    for m2, w2 in lines[mm].items():
                        nxt[m2] = nxt.get(m2, 0.0) + ww * w2
                else:
                    nxt[mm] = nxt.get(mm, 0.0) + ww
            line = nxt
            if not changed:
                break
        lines[g] = line
    con_of = -np.ones(n_nodes * 18, np.int32)
    con_ptr, con_master, con_weight = [0], [], []
    for i, g in enumerate(sorted(lines)):
        con_of[g] = i
        for mm, ww in sorted(lines[g].items()):
            con_master.append(mm)
            con_weight.append(ww)
        con_ptr.append(len(con_master))
    cells = m.cell_nodes.reshape(-1, (deg + 1) ** 3)
    tgt = {}
    for a in range(n_nodes):
        t = set()
        for c in range(18):
            g = a * 18 + c
            if g in lines:
                t.update(mm // 18 for mm in lines[g])
            else:
                t.add(a)
        tgt[a] = t
    nb = [set([r]) for r in range(n_nodes)]
    for cn in cells:
        for a in cn:
            for b in cn:
                for r in tgt[int(a)]:
                    nb[r].update(tgt[int(b)])
    row_ptr = np.zeros(n_nodes + 1, np.int32)
    cols = []
    for r in range(n_nodes):
        s = sorted(nb[r])
        cols.extend(s)
        row_ptr[r + 1] = row_ptr[r] + len(s)
    return Mesh(deg, m.cell_nodes, m.cell_h, list(faces), n_nodes, con_of, np.array(con_ptr, np.int32),
                con_master, con_weight, row_ptr, np.array(cols, np.int32), xyz)
