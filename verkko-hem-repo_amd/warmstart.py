"""Learned warm start for batched slab solves: a coordinate MLP (glhip.MLP)
trained on converged solutions predicts the initial order parameter of new
(T, p, D) instances; the seeded bulk-phase guess stays as the fallback where
the prediction's first residual is not lower.  Everything numeric runs in the
library (batched Newton-GMRES, BF16 MLP kernels); this module only arranges
the data.

Instances are slabs of lateral extent 40 xi0 and thickness D, A = 0 on the two
slab faces and natural on the sides, Q1 cells of 5 x 5 x (<= 3) xi0
(8 x 8 x ceil(D/3)), T/Tc ~ U[0.3, 0.95], p ~ U[0, 34) bar, D ~ U[6, 40] xi0."""
import numpy as np

LATERAL = 40.0
K_CONST = (0.42072, 0.42072, 0.42072)


class Instance:
    __slots__ = ("T", "p", "D", "mesh", "coef", "gap_B", "xyz")


def slab_instances(G, n, seed):
    rng = np.random.default_rng(seed)
    T = rng.uniform(0.3, 0.95, n)
    P = np.minimum(rng.uniform(0.0, 34.0, n), np.nextafter(34.0, 0.0))
    D = rng.uniform(6.0, 40.0, n)
    out = []
    h = LATERAL / 2
    for i in range(n):
        nz = int(np.ceil(D[i] / 3.0))
        m = G.Mesh(1, (8, 8, nz), (-h, -h, -D[i] / 2), (h, h, D[i] / 2), [1, 1, 1, 1, 0, 0])
        rc, mp = G.matep(P[i], T[i])
        it = Instance()
        it.T, it.p, it.D, it.mesh = float(T[i]), float(P[i]), float(D[i]), m
        it.coef = [mp[0], mp[1], mp[2], mp[3], mp[4], mp[5], K_CONST[0], K_CONST[1], K_CONST[2], 1.0e10]
        it.gap_B = float(mp[9])
        it.xyz = m.arrays()["node_xyz"].reshape(-1, 3)
        out.append(it)
    return out


def sweep_instances(G, n_inst=256, seed=2):
    """The batched phase-diagram sweep: n_inst slab instances, T/Tc ~
    U[0.3,0.95], p ~ U[0,34) bar, thickness D ~ U[6,40] xi0, lateral extent
    40 xi0, Q1 meshes of nx = ny in 8..16 and nz in 2..8 cells, A = 0 on the
    two slab faces and natural on the sides; B phase + 0.05 Delta_B noise
    (seed 1000 + i).  Returns (meshes, coefficient rows, guesses, params)."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(0.3, 0.95, n_inst)
    P = np.minimum(rng.uniform(0.0, 34.0, n_inst), np.nextafter(34.0, 0.0))
    D = rng.uniform(6.0, 40.0, n_inst)
    NX = rng.integers(8, 17, n_inst)
    NZ = rng.integers(2, 9, n_inst)
    meshes, coefs, U = [], [], []
    h = LATERAL / 2
    for i in range(n_inst):
        m = G.Mesh(1, (int(NX[i]), int(NX[i]), int(NZ[i])), (-h, -h, -D[i] / 2), (h, h, D[i] / 2),
                   [1, 1, 1, 1, 0, 0])
        rc, mp = G.matep(P[i], T[i])
        meshes.append(m)
        coefs.append([mp[0], mp[1], mp[2], mp[3], mp[4], mp[5], K_CONST[0], K_CONST[1], K_CONST[2], 1.0e10])
        U.append(G.initial_guess(m.n_nodes, mp[9], phase=0, amp=0.05, seed=1000 + i))
    return meshes, np.array(coefs), U, dict(T=T, p=P, D=D, nx=NX, nz=NZ)


def features(inst):
    """Network inputs of the nodes of an instance, each scaled to about [-1, 1]."""
    xyz = inst.xyz
    n = xyz.shape[0]
    X = np.empty((n, 6), np.float32)
    X[:, 0] = xyz[:, 0] / (LATERAL / 2)
    X[:, 1] = xyz[:, 1] / (LATERAL / 2)
    X[:, 2] = xyz[:, 2] / (inst.D / 2)
    X[:, 3] = (inst.T - 0.625) / 0.325
    X[:, 4] = (inst.p - 17.0) / 17.0
    X[:, 5] = (inst.D - 23.0) / 17.0
    return X


def plain_guess(G, inst, seed):
    return G.initial_guess(inst.mesh.n_nodes, inst.gap_B, phase=0, amp=0.05, seed=seed)


def solve_batch(G, insts, guesses, fp32=True, rtol=1e-10, lin_tol=0.1, lin_maxit=3000, restart=100,
                max_newton=60, verbose=0, ew=False, r_ref=None, matrix_free=False):
    """Batched Newton solve; returns (stats, per-instance stats, F, U list).
    ew: Eisenstat-Walker forcing; r_ref: per-instance stopping reference
    (||R|| <= rtol r_ref), e.g. the first residuals of the plain guesses, so
    that different guesses are held to the same absolute accuracy."""
    ctx = G.Context(0)
    try:
        ctx.set_verbose(verbose)
        batch = G.Mesh.concat([it.mesh for it in insts])
        ctx.upload_mesh(batch)
        ctx.set_materials(np.array([it.coef for it in insts]))
        ctx.set_forcing(ew)
        ctx.set_matrix_free(matrix_free)
        if r_ref is not None:
            ctx.set_residual_reference(r_ref)
        ctx.set_solution(np.concatenate(guesses))
        st = ctx.newton_solve(rtol=rtol, max_newton=max_newton, lin_tol=lin_tol, lin_maxit=lin_maxit,
                              restart=restart, fp32=fp32)
        ist = ctx.instance_stats()
        F = ctx.instance_energy()
        U = ctx.solution()
        no, _ = batch.instances()
        Us = [U[no[i] * 18:no[i + 1] * 18] for i in range(len(insts))]
        batch.close()
        return st, ist, F, Us
    finally:
        ctx.close()


def solve_from_candidates(G, insts, candidates, fp32=True, rtol=1e-10, lin_tol=0.1, lin_maxit=3000, restart=100,
                          max_newton=60, ew=True, verbose=0, matrix_free=False):
    """One context for the whole batch: the first residual of every candidate
    guess set (candidates[0] is the plain guess), per instance the candidate
    with the lowest residual (the plain guess wins ties: it is the fallback),
    then the batched Newton solve from the chosen guesses, stopped at
    ||R|| <= rtol ||R(plain guess)|| for every instance.
    Returns (stats, per-instance stats, F, first residuals [k][n], choice,
    final residuals recomputed over the whole batch)."""
    ctx = G.Context(0)
    try:
        ctx.set_verbose(verbose)
        batch = G.Mesh.concat([it.mesh for it in insts])
        ctx.upload_mesh(batch)
        ctx.set_materials(np.array([it.coef for it in insts]))
        res = []
        for cand in candidates:
            ctx.set_solution(np.concatenate(cand))
            res.append(ctx.instance_residuals())
        res = np.array(res)
        choice = np.argmin(res, axis=0)  # first minimum: the plain guess on ties
        guess = np.concatenate([candidates[c][i] for i, c in enumerate(choice)])
        ctx.set_forcing(ew)
        ctx.set_matrix_free(matrix_free)
        ctx.set_residual_reference(res[0])
        ctx.set_solution(guess)
        st = ctx.newton_solve(rtol=rtol, max_newton=max_newton, lin_tol=lin_tol, lin_maxit=lin_maxit,
                              restart=restart, fp32=fp32)
        ist = ctx.instance_stats()
        F = ctx.instance_energy()
        r_end = ctx.instance_residuals()
        batch.close()
        return st, ist, F, res, choice, r_end
    finally:
        ctx.close()


def first_residuals(G, insts, guesses):
    """||R|| of every instance at the given guesses (one batched assembly)."""
    ctx = G.Context(0)
    try:
        batch = G.Mesh.concat([it.mesh for it in insts])
        ctx.upload_mesh(batch)
        ctx.set_materials(np.array([it.coef for it in insts]))
        ctx.set_solution(np.concatenate(guesses))
        r = ctx.instance_residuals()
        batch.close()
        return r
    finally:
        ctx.close()


def training_set(insts, Us, converged):
    """(X, Y): node features and converged fields in units of the bulk B gap."""
    Xs, Ys = [], []
    for it, U, ok in zip(insts, Us, converged):
        if not ok:
            continue
        Xs.append(features(it))
        Ys.append((U.reshape(-1, 18) / it.gap_B).astype(np.float32))
    return np.concatenate(Xs), np.concatenate(Ys)


def train(net, steps, batch, lr=1e-3, lr_end=1e-5, segment=500, seed=3, log=None):
    """Adam on the GPU with a cosine learning-rate decay from lr to lr_end,
    held constant over segments of `segment` steps (the step counter, the
    moments and the batch sampler carry over between segments).  A constant
    rate leaves the final weights wherever the last fluctuation of the
    mini-batch loss put them; the decay settles them.  Returns the last
    mini-batch loss of every segment."""
    done, curve = 0, []
    while done < steps:
        k = min(segment, steps - done)
        rate = lr_end + 0.5 * (lr - lr_end) * (1.0 + np.cos(np.pi * (done + 0.5 * k) / steps))
        L = net.train(k, batch=batch, lr=rate, seed=seed, loss_every=k)
        done += k
        curve.append(float(L[-1]))
        if log:
            log(done, curve[-1], rate)
    return curve


def predict_guesses(net, insts):
    X = np.concatenate([features(it) for it in insts])
    Y = net.predict(X)
    out, k = [], 0
    for it in insts:
        n = it.xyz.shape[0]
        out.append((Y[k:k + n].astype(np.float64) * it.gap_B).reshape(-1))
        k += n
    return out


def choose(r_mlp, r_plain):
    """Per instance: the learned guess unless its first residual is not below
    the plain guess's (then the plain guess is the fallback)."""
    return np.asarray(r_mlp) < np.asarray(r_plain)
