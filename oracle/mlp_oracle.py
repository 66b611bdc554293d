"""FP64 CPU oracle of the learned PSOR warm start (test infrastructure only).

Nothing in the product imports this module: tests, the smoke check and the CPU
baseline leg of the benchmark use it to check the HIP kernels of
``rdk_mlp.hip``.  The learned start has no counterpart in the reference (whose
warm start is the previous-step impulse cache, src/physics/nsc_contact.cpp:
305-317); the network and features are this build's own design:

* MLP 12 -> 64 -> 64 -> 3, SiLU hidden activations, linear output.
* Parameters flattened FP32 row-major per layer: W1[12][64], b1[64],
  W2[64][64], b2[64], W3[64][3], b3[3]  (5187 values).
* Loss: mean squared error over all n*3 outputs.

Pins (tests/test_mlp_oracle.py):
  * ``grad`` against central finite differences of ``loss`` in FP64;
  * closed forms: with W3 = 0 the hidden-layer gradients vanish and
    d loss / d b3 = 2 * mean(y - t); SiLU'(z) against its difference quotient;
  * ``forward_bf16`` equals ``forward`` when every weight and input is exactly
    representable in BF16 and the hidden activations are too (checked on a
    hand-built case), and differs from it by at most the BF16 rounding bound
    otherwise;
  * ``features`` against an independent rotation-matrix restatement and
    hand-computed single-contact cases.
"""
import numpy as np

IN, H, OUT = 12, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * OUT
NPARAM = OFF_B3 + OUT


def unpack(p):
    p = np.asarray(p, np.float64)
    assert p.size == NPARAM
    W1 = p[OFF_W1:OFF_B1].reshape(IN, H)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(H, H)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3].reshape(H, OUT)
    b3 = p[OFF_B3:]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3]).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    h1 = silu(X @ W1 + b1)
    h2 = silu(h1 @ W2 + b2)
    return h2 @ W3 + b3


def loss(p, X, T):
    Y = forward(p, X)
    return float(np.mean((Y - np.asarray(T, np.float64)) ** 2))


def grad(p, X, T):
    """Analytic gradient of ``loss`` (reverse mode, FP64)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    z1 = X @ W1 + b1
    h1 = silu(z1)
    z2 = h1 @ W2 + b2
    h2 = silu(z2)
    Y = h2 @ W3 + b3
    dY = 2.0 * (Y - T) / Y.size
    gW3 = h2.T @ dY
    gb3 = dY.sum(0)
    dz2 = (dY @ W3.T) * dsilu(z2)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * dsilu(z1)
    gW1 = X.T @ dz1
    gb1 = dz1.sum(0)
    return pack(gW1, gb1, gW2, gb2, gW3, gb3)


def bf16(a):
    """Round FP32 values to the nearest BF16 (ties to even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(finite, r, u & 0xFFFF0000)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, X):
    """Forward pass as the matrix-core kernel computes it: BF16 weights, inputs
    and hidden activations, sums exact-ish (FP64 here, FP32 on the device),
    FP32 biases."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float32))
    b1, b2, b3 = (np.asarray(b, np.float32).astype(np.float64) for b in (b1, b2, b3))
    X = bf16(np.asarray(X, np.float32))
    h1 = bf16(silu(X @ bf16(W1) + b1).astype(np.float32))
    h2 = bf16(silu(h1 @ bf16(W2) + b2).astype(np.float32))
    return h2 @ bf16(W3) + b3


def quat_matrix(q):
    """Rotation matrices of unit quaternions q = (w, x, y, z), shape (n, 3, 3)."""
    w, x, y, z = (q[:, k] for k in range(4))
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def features(mf, a, b, state, r, mu_of_body, vs, prev=None, prev_wall=None):
    """Features of manifolds in natural order.

    mf: (42, m) manifold fields (normal 0-2, t1 3-5, t2 6-8, r_a 9-11,
    r_b 12-14, g_n 35, phi 38, v_n pre 39); a, b: body indices (b < 0 for a
    wall, slot = -1 - b); state: dict with x, q, v, w arrays; r: radii;
    mu_of_body: friction coefficient of each body's system; vs: velocity
    unit |g| dt; prev: dict (a, b) -> lambda_n of the previous step for rod
    pairs; prev_wall: dict (a, slot) -> lambda_n for walls.
    Returns X (m, 12) float64 and the impulse unit lscale (m,).
    """
    m = mf.shape[1]
    n = mf[0:3].T
    t1 = mf[3:6].T
    t2 = mf[6:9].T
    ra = mf[9:12].T
    rb = mf[12:15].T
    gn = mf[35]
    phi = mf[38]
    vnpre = mf[39]
    R = quat_matrix(state["q"])
    v, w = state["v"], state["w"]
    X = np.zeros((m, IN))
    ls = vs / gn
    for i in range(m):
        ai, bi = int(a[i]), int(b[i])
        va = v[ai] + np.cross(w[ai], ra[i])
        na = R[ai].T @ n[i]
        if bi >= 0:
            vb = v[bi] + np.cross(w[bi], rb[i])
            vrel = vb - va
            nb = R[bi].T @ n[i]
            rs = r[ai] + r[bi]
            lp = (prev or {}).get((ai, bi), 0.0)
        else:
            vrel = -va
            nb = np.zeros(3)
            rs = r[ai]
            lp = (prev_wall or {}).get((ai, -1 - bi), 0.0)
        X[i, 0] = phi[i] / rs
        X[i, 1:4] = na
        X[i, 4:7] = nb
        X[i, 7] = vnpre[i] / vs
        X[i, 8] = t1[i] @ vrel / vs
        X[i, 9] = t2[i] @ vrel / vs
        X[i, 10] = mu_of_body[ai]
        X[i, 11] = max(lp, 0.0) / ls[i]
    return X, ls


def apply_prediction(Y, ls, mu, lam_prev=None, all_manifolds=False):
    """Start impulses: manifolds with a previous normal impulse keep the
    previous-step impulses (unless all_manifolds); the others take the
    prediction (units of lscale) projected onto lambda_n >= 0 and the friction
    disc |lambda_t| <= mu lambda_n."""
    Y = np.asarray(Y, np.float64)
    base = np.zeros((Y.shape[0], 3)) if lam_prev is None else np.asarray(lam_prev, np.float64)
    ln = np.maximum(0.0, Y[:, 0] * ls)
    t = Y[:, 1:3] * ls[:, None]
    lim = mu * ln
    nt = np.linalg.norm(t, axis=1)
    sc = np.where(nt > lim, np.where(nt > 0, lim / np.where(nt > 0, nt, 1.0), 0.0), 1.0)
    pred = np.column_stack([ln, t * sc[:, None]])
    keep = (base[:, 0] > 0.0) & (not all_manifolds)
    return np.where(keep[:, None], base, pred)


def adamw_step(p, g, m, v, t, lr, wd, b1=0.9, b2=0.999, eps=1e-8):
    """One AdamW step (decoupled weight decay on the weight matrices only);
    t is the 1-based step count.  Returns (p, m, v)."""
    p = np.asarray(p, np.float64).copy()
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    wmask = np.zeros(NPARAM, bool)
    wmask[OFF_W1:OFF_B1] = True
    wmask[OFF_W2:OFF_B2] = True
    wmask[OFF_W3:OFF_B3] = True
    p[wmask] -= lr * wd * p[wmask]
    p -= lr * mh / (np.sqrt(vh) + eps)
    return p, m, v
