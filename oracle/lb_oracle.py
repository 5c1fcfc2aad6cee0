"""CPU restatement (numpy, FP64) of the learned thin absorbing boundary.

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by the product.  It
restates, from its definition, the model the library implements in
proxy-geos-hc_amd/csrc/mlp.hip (no reference counterpart: the reference's
boundaries are the sponge, src/fd/utilsFDTD/FDTDinit.hpp:319-396, and the PML,
FDTDkernels.hpp:36-104):
  * the layer geometry (faces 2*axis+side, compact index ((d*n_u)+u)*n_v+v);
  * the inputs of a layer point: (pn - h)/s and (pnm1 - h)/s at offsets -4..4
    along the inward normal (0 beyond the region's edge), sqrt(vp2 c_1);
  * the training target sum_{l>d} (c_l/c_1) (p(P + l n_out) - h)/s;
  * the network y = w3 . relu(W2 relu(W1 x + b1[d]) + b2) + b3 at bf16 weights;
  * the mean squared error and its analytic gradient;
  * the correction vp2 c_1 s y added at every layer point.
Parity: the analytic gradient is pinned by central finite differences
(test_lb_oracle.py); geometry and targets by closed forms on polynomial fields.
"""
import numpy as np

W = 4          # layer width
R = 4          # stencil half width
NIN = 19
H = 32
FACES = 6
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * NIN
OFF_W2 = OFF_B1 + W * H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NP = OFF_B3 + 1
NSTRIDE = 22
S_TINY = 1e-30


def bf16(a):
    """Round float32 values to bf16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, NIN), p[OFF_B1:OFF_W2].reshape(W, H),
            p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, X, depth):
    """y for samples X (N x 19) at layer depths `depth` with face parameters p."""
    W1, B1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    a1 = X @ W1.T + B1[depth]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    return h2 @ w3 + b3, (X, a1, h1, a2, h2)


def loss_grad(theta, samples):
    """Mean squared error over every sample of every face and its gradient with
    respect to theta (FACES x NP), both evaluated at bf16(theta).
    samples[f]: (N_f x 22) array (x[19], target, depth, index)."""
    th = bf16(np.asarray(theta, np.float32).reshape(FACES, NP))
    total = sum(len(s) for s in samples)
    loss = 0.0
    grad = np.zeros((FACES, NP))
    for f in range(FACES):
        S = np.asarray(samples[f], np.float64)
        if len(S) == 0:
            continue
        X, t, d = S[:, :NIN], S[:, NIN], S[:, NIN + 1].astype(int)
        y, (X, a1, h1, a2, h2) = forward(th[f], X, d)
        e = y - t
        loss += np.sum(e * e) / total
        dy = 2.0 * e / total
        W1, B1, W2, b2, w3, b3 = unpack(th[f])
        d2 = (dy[:, None] * w3[None, :]) * (a2 > 0)
        d1 = (d2 @ W2) * (a1 > 0)
        g = grad[f]
        g[OFF_W1:OFF_B1] = (d1.T @ X).reshape(-1)
        gb1 = np.zeros((W, H))
        np.add.at(gb1, d, d1)
        g[OFF_B1:OFF_W2] = gb1.reshape(-1)
        g[OFF_W2:OFF_B2] = (d2.T @ h1).reshape(-1)
        g[OFF_B2:OFF_W3] = d2.sum(0)
        g[OFF_W3:OFF_B3] = h2.T @ dy
        g[OFF_B3] = dy.sum()
    return loss, grad


def loss_at(theta64, samples):
    """Mean squared error at FP64 parameters (no rounding): for finite differences."""
    total = sum(len(s) for s in samples)
    loss = 0.0
    for f in range(FACES):
        S = np.asarray(samples[f], np.float64)
        if len(S) == 0:
            continue
        y, _ = forward(theta64[f], S[:, :NIN], S[:, NIN + 1].astype(int))
        loss += np.sum((y - S[:, NIN]) ** 2) / total
    return loss


def loss_grad64(theta64, samples):
    """Analytic gradient at FP64 parameters (the same formulas as loss_grad)."""
    total = sum(len(s) for s in samples)
    grad = np.zeros((FACES, NP))
    for f in range(FACES):
        S = np.asarray(samples[f], np.float64)
        if len(S) == 0:
            continue
        X, t, d = S[:, :NIN], S[:, NIN], S[:, NIN + 1].astype(int)
        y, (X, a1, h1, a2, h2) = forward(theta64[f], X, d)
        dy = 2.0 * (y - t) / total
        W1, B1, W2, b2, w3, b3 = unpack(theta64[f])
        d2 = (dy[:, None] * w3[None, :]) * (a2 > 0)
        d1 = (d2 @ W2) * (a1 > 0)
        g = grad[f]
        g[OFF_W1:OFF_B1] = (d1.T @ X).reshape(-1)
        gb1 = np.zeros((W, H))
        np.add.at(gb1, d, d1)
        g[OFF_B1:OFF_W2] = gb1.reshape(-1)
        g[OFF_W2:OFF_B2] = (d2.T @ h1).reshape(-1)
        g[OFF_B2:OFF_W3] = d2.sum(0)
        g[OFF_W3:OFF_B3] = h2.T @ dy
        g[OFF_B3] = dy.sum()
    return grad


# ------------------------------------------------------------------ geometry
def face_geom(n, f):
    a, side = f >> 1, f & 1
    nu = n[1] if a == 0 else n[0]
    nv = n[1] if a == 2 else n[2]
    return a, side, nu, nv


def face_points(n, f):
    """Grid coordinates (N x 3) and depths of face f's layer in compact order."""
    a, side, nu, nv = face_geom(n, f)
    d, u, v = np.meshgrid(np.arange(W), np.arange(nu), np.arange(nv), indexing="ij")
    d, u, v = d.reshape(-1), u.reshape(-1), v.reshape(-1)
    c = np.zeros((len(d), 3), int)
    ua = 1 if a == 0 else 0
    va = 1 if a == 2 else 2
    c[:, a] = np.where(side == 1, n[a] - 1 - d, d)
    c[:, ua] = u
    c[:, va] = v
    return c, d


def layer_inputs(pn, pnm1, vp, origin, n, f, h, c1, chat=None):
    """Inputs (N x 19), s, depths and (if chat is given) targets of face f's
    layer of the region origin..origin+n of a grid.  pn, pnm1: extended
    arrays (halo R on each side, as phc_fd_get_field returns them); vp:
    interior array (nx, ny, nz) of vp^2 dt^2."""
    c, d = face_points(n, f)
    a, side, _, _ = face_geom(n, f)
    dirv = np.zeros(3, int)
    dirv[a] = -1 if side else 1       # inward
    P = c + np.asarray(origin)[None, :]
    X = np.zeros((len(d), NIN))
    pn = np.asarray(pn, np.float64)
    pnm1 = np.asarray(pnm1, np.float64)
    for m in range(9):
        out = 4 - m
        Q = P + (m - 4) * dirv[None, :] + R
        a_ = pn[Q[:, 0], Q[:, 1], Q[:, 2]] - h
        b_ = pnm1[Q[:, 0], Q[:, 1], Q[:, 2]] - h
        keep = out <= d
        X[:, m] = np.where(keep, a_, 0.0)
        X[:, 9 + m] = np.where(keep, b_, 0.0)
    s = np.max(np.abs(X[:, :18]), axis=1)
    # amplitudes at or below S_TINY: a quiet point (the library's 1/s would overflow FP32)
    inv = np.where(s > S_TINY, 1.0 / np.where(s > S_TINY, s, 1.0), 0.0)
    X[:, :18] *= inv[:, None]
    v = np.asarray(vp, np.float64)[P[:, 0], P[:, 1], P[:, 2]]
    X[:, 18] = np.sqrt(v * c1)
    tgt = None
    if chat is not None:
        tgt = np.zeros(len(d))
        for l in range(1, R + 1):
            Q = P - l * dirv[None, :] + R
            val = pn[Q[:, 0], Q[:, 1], Q[:, 2]] - h
            tgt += np.where(l > d, chat[l] * val, 0.0)
        tgt *= inv
    return X, s, d, v, tgt


def ymax(chat, d):
    """Bound of y at layer depth d: sum_{l>d} |c_l / c_1|, summed in float32
    (the run-time clamp of the correction pass, csrc/lb.hpp lb_ymax)."""
    m = np.float32(0.0)
    for l in range(d + 1, R + 1):
        m = np.float32(m + np.float32(abs(np.float32(chat[l]))))
    return float(m)


def corrections(theta, pn, pnm1, vp, n, h, c1, chat):
    """Per face: vp2 c_1 s y at every layer point of a whole grid (compact
    order), y clamped to +-ymax(depth)."""
    th = bf16(np.asarray(theta, np.float32).reshape(FACES, NP))
    bound = np.array([ymax(chat, d) for d in range(W)])
    out = []
    for f in range(FACES):
        X, s, d, v, _ = layer_inputs(pn, pnm1, vp, (0, 0, 0), n, f, h, c1)
        y, _ = forward(th[f], X, d)
        yb = bound[d.astype(int)]
        out.append(v * c1 * s * np.clip(y, -yb, yb))
    return out


def apply_corrections(p_new, corr, n):
    """Add the per-face corrections to an extended field in place."""
    for f in range(FACES):
        c, d = face_points(n, f)
        Q = c + R
        np.add.at(p_new, (Q[:, 0], Q[:, 1], Q[:, 2]), corr[f])
    return p_new


def onpolicy_targets(tpn, tpm, tvp, wcur, origin, n, f, h, chat, c1, t_max=16.0):
    """On-policy samples of face f: inputs from the thin levels (extended
    arrays of the region grid), target sum_{l>d} chat_l (W(P + l n_out) - h) /
    s_thin from the wide level at the same step; mask of the kept points."""
    X, s, d, v, _ = layer_inputs(tpn, tpm, tvp, (0, 0, 0), n, f, h, c1)
    pts, _ = face_points(n, f)
    a, side, _, _ = face_geom(n, f)
    dirv = np.zeros(3, int)
    dirv[a] = -1 if side else 1
    Pw = pts + np.asarray(origin)[None, :]
    w = np.asarray(wcur, np.float64)
    t = np.zeros(len(pts))
    for l in range(1, R + 1):
        Q = Pw - l * dirv[None, :] + R
        t += np.where(l > d, chat[l] * (w[Q[:, 0], Q[:, 1], Q[:, 2]] - h), 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = t / s
    keep = (s > S_TINY) & (np.abs(t) <= t_max)
    return X, s, d, t, keep
