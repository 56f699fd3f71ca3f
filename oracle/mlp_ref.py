"""CPU reference of the coordinate-MLP warm start (test infrastructure only:
imported by tests/ as the checker of the GPU kernels, never by the product).

The network has no counterpart in the reference code; this is the plain
restatement of its definition: Fourier features [sin, cos](2 pi x B),
`depth` SiLU layers y = x W + b, a linear output layer, mean-squared-error
loss, and the backward pass by the chain rule.  With emulate=True the points
where the GPU rounds to BF16 are reproduced (weights, layer inputs, the
output gradient and the hidden-layer gradients); the products accumulate in
float64.  With emulate=False everything is float64 (for the finite-difference
check of the backward pass itself).

Parameter layout (flat): for layer l, W_l (K_l x N_l row-major) then b_l."""
import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest even) and back to float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def layout(n_freq, width, depth, n_out):
    K = [2 * n_freq] + [width] * depth
    N = [width] * depth + [n_out]
    offs, o = [], 0
    for k, n in zip(K, N):
        offs.append((o, o + k * n))
        o += k * n + n
    return K, N, offs, o


def unpack(p, n_freq, width, depth, n_out):
    K, N, offs, _ = layout(n_freq, width, depth, n_out)
    Ws = [np.asarray(p[a:a + k * n], np.float64).reshape(k, n) for (a, b), k, n in zip(offs, K, N)]
    bs = [np.asarray(p[b:b + n], np.float64) for (a, b), n in zip(offs, N)]
    return Ws, bs


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def loss_grad(p, B, X, Y, n_freq, width, depth, n_out, emulate=True):
    """Mean-squared error and its gradient (flat, float64)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    Ws, bs = unpack(p, n_freq, width, depth, n_out)
    Wr = [np.asarray(r(W), np.float64) for W in Ws]
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    ang = 2.0 * np.pi * (X @ np.asarray(B, np.float64))
    H = [np.asarray(r(np.concatenate([np.sin(ang), np.cos(ang)], 1)), np.float64)]
    Z = []
    for l in range(depth):
        z = H[-1] @ Wr[l] + bs[l]
        Z.append(z)
        H.append(np.asarray(r(silu(z)), np.float64))
    out = H[-1] @ Wr[depth] + bs[depth]
    n = X.shape[0]
    d = out - Y
    loss = float(np.mean(d * d))
    dZ = np.asarray(r(2.0 * d / (n * n_out)), np.float64)
    gW, gb = [None] * (depth + 1), [None] * (depth + 1)
    for l in range(depth, -1, -1):
        gW[l] = H[l].T @ dZ
        gb[l] = dZ.sum(0)
        if l == 0:
            break
        dH = dZ @ Wr[l].T
        dZ = np.asarray(r(dH * dsilu(Z[l - 1])), np.float64)
    g = np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in zip(gW, gb)])
    return loss, g


def predict(p, B, X, n_freq, width, depth, n_out, emulate=True):
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    Ws, bs = unpack(p, n_freq, width, depth, n_out)
    ang = 2.0 * np.pi * (np.asarray(X, np.float64) @ np.asarray(B, np.float64))
    h = np.asarray(r(np.concatenate([np.sin(ang), np.cos(ang)], 1)), np.float64)
    for l in range(depth):
        h = np.asarray(r(silu(h @ np.asarray(r(Ws[l]), np.float64) + bs[l])), np.float64)
    return h @ np.asarray(r(Ws[depth]), np.float64) + bs[depth]
