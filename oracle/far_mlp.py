"""far_mlp.py -- CPU restatement of the learned far-field network (numpy).

TEST INFRASTRUCTURE ONLY: the checker for the device's MLP kernels
(vvflow_amd/csrc/farfield.hip).  Only tests/, the smoke entry point and the
CPU-baseline leg of bench.py may import it.

The network and its flat parameter vector (include/vvhd_gpu.h):
    W1[64][12] b1[64] W2[64][64] b2[64] W3[2][64] b3[2] | mu[12] sd[12] ysd[2]
    y = W3 silu(W2 silu(W1 xs + b1) + b2) + b3,   xs = (x - mu) / sd.
With mirror=True the rounding points of the device are reproduced (round
to nearest even): in training (bf16) the standardised inputs, weights,
hidden activations fed to the next product and the deltas fed to the
backward products; in inference (forward_jvp) bf16 inputs and layer-1
weights, fp16 hidden values, tangents and layer-2/3 weights.  With
mirror=False everything is float64 (for finite-difference checks).
This network has no counterpart in the reference, so no reference line is
cited; the checks that pin it are the finite-difference gradient test and the
closed-form cases in tests/test_far_oracle.py.
"""
import numpy as np

N_IN, N_HID, N_OUT = 12, 64, 2
O_W1 = 0
O_B1 = O_W1 + N_HID * N_IN
O_W2 = O_B1 + N_HID
O_B2 = O_W2 + N_HID * N_HID
O_W3 = O_B2 + N_HID
O_B3 = O_W3 + N_OUT * N_HID
N_TRAIN = O_B3 + N_OUT
O_MU = N_TRAIN
O_SD = O_MU + N_IN
O_YSD = O_SD + N_IN
N_PARAMS = O_YSD + N_OUT


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(p):
    p = np.asarray(p)
    if p.dtype != np.float64:   # float64 parameters stay float64 (finite differences)
        p = p.astype(np.float32)
    W1 = p[O_W1:O_B1].reshape(N_HID, N_IN)
    b1 = p[O_B1:O_W2]
    W2 = p[O_W2:O_B2].reshape(N_HID, N_HID)
    b2 = p[O_B2:O_W3]
    W3 = p[O_W3:O_B3].reshape(N_OUT, N_HID)
    b3 = p[O_B3:N_TRAIN]
    return W1, b1, W2, b2, W3, b3, p[O_MU:O_SD], p[O_SD:O_YSD], p[O_YSD:N_PARAMS]


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s + z * s * (1.0 - s)


def init_params(rng, mu=None, sd=None, ysd=None):
    """He-style initialisation; mu/sd/ysd default to 0/1/1."""
    p = np.zeros(N_PARAMS, dtype=np.float32)
    p[O_W1:O_B1] = rng.normal(0, np.sqrt(2.0 / N_IN), N_HID * N_IN)
    p[O_W2:O_B2] = rng.normal(0, np.sqrt(2.0 / N_HID), N_HID * N_HID)
    p[O_W3:O_B3] = rng.normal(0, np.sqrt(1.0 / N_HID), N_OUT * N_HID)
    p[O_MU:O_SD] = 0 if mu is None else mu
    p[O_SD:O_YSD] = 1 if sd is None else sd
    p[O_YSD:N_PARAMS] = 1 if ysd is None else ysd
    return p


def forward(p, X, mirror=True):
    """Raw outputs (before ysd) and the cache for backward."""
    W1, b1, W2, b2, W3, b3, mu, sd, _ = unpack(p)
    r = bf16 if mirror else (lambda a: np.asarray(a, dtype=np.float64))
    X = np.asarray(X, dtype=np.float32 if mirror else np.float64)
    xs = r(((X - mu) / sd).astype(np.float32) if mirror else (X - mu) / sd)
    xs = xs.astype(np.float64)
    W1r, W2r, W3r = (r(W).astype(np.float64) for W in (W1, W2, W3))
    z1 = xs @ W1r.T + b1
    h1 = r(silu(z1).astype(np.float32) if mirror else silu(z1)).astype(np.float64)
    z2 = h1 @ W2r.T + b2
    h2 = r(silu(z2).astype(np.float32) if mirror else silu(z2)).astype(np.float64)
    y = h2 @ W3r.T + b3
    return y, dict(xs=xs, z1=z1, h1=h1, z2=z2, h2=h2, W=(W1r, W2r, W3r), mirror=mirror)


def loss_and_grad(p, X, Y, W, mirror=True):
    """Loss sum_j W_j |net(X_j) - Y_j/ysd|^2 / n and its gradient with
    respect to the N_TRAIN trainable parameters (the device's definition)."""
    n = len(X)
    y, c = forward(p, X, mirror)
    ysd = unpack(p)[8].astype(np.float64)
    W = np.asarray(W, dtype=np.float64)[:, None]
    res = y - np.asarray(Y, dtype=np.float64) / ysd
    loss = float(np.sum(W * res * res) / n)
    dy = 2.0 * W * res / n
    r = (lambda a: bf16(a.astype(np.float32)).astype(np.float64)) if mirror else (lambda a: a)
    W1r, W2r, W3r = c["W"]
    dyr = r(dy)
    dz2 = (dyr @ W3r) * dsilu(c["z2"])
    dz2r = r(dz2)
    dz1 = (dz2r @ W2r) * dsilu(c["z1"])
    dz1r = r(dz1)
    g = np.zeros(N_TRAIN)
    g[O_W3:O_B3] = (dyr.T @ c["h2"]).ravel()
    g[O_B3:N_TRAIN] = dy.sum(0)
    g[O_W2:O_B2] = (dz2r.T @ c["h1"]).ravel()
    g[O_B2:O_W3] = dz2.sum(0)
    g[O_W1:O_B1] = (dz1r.T @ c["xs"]).ravel()
    g[O_B1:O_W2] = dz1.sum(0)
    return loss, g


def correction(p, X, scale, ux, uy, mirror=True):
    """Velocity correction (u, v) of pairs from their features: the network
    output times ysd, scaled and rotated back to the target direction."""
    y, _ = forward(p, X, mirror=mirror)
    ysd = unpack(p)[8].astype(np.float64)
    a, b = y[:, 0] * ysd[0], y[:, 1] * ysd[1]
    return np.stack([(a * ux - b * uy) * scale, (a * uy + b * ux) * scale], 1)


def feature_tangents(X, ux, uy, rinv):
    """Derivatives of the features with respect to the target position,
    (dX/dpx, dX/dpy): with theta the angle of u, d theta/dp = (-uy, ux)/r;
    the rotated dipoles turn with theta (d(f2 + i f3)/d theta = f3 - i f2),
    the quadrupole twice as fast, d/r scales as 1/r, the rest is constant."""
    X = np.asarray(X, dtype=np.float64)
    thx, thy = -uy * rinv, ux * rinv
    rot = np.zeros_like(X)
    rot[:, 2], rot[:, 3] = X[:, 3], -X[:, 2]
    rot[:, 4], rot[:, 5] = X[:, 5], -X[:, 4]
    rot[:, 6], rot[:, 7] = 2 * X[:, 7], -2 * X[:, 6]
    TX, TY = rot * thx[:, None], rot * thy[:, None]
    TX[:, 10] = -X[:, 10] * ux * rinv
    TY[:, 10] = -X[:, 10] * uy * rinv
    return TX, TY


def feature_directions(X):
    """The two input directions the device propagates tangents along:
    d1, the rotated moments' derivative in theta (rot of feature_tangents),
    and d2, -d/r at input 10.  The feature tangents along x and y are
    thx d1 + (ux/r) d2 and thy d1 + (uy/r) d2 (thx, thy = -uy/r, ux/r), so
    by linearity the network's tangents along x and y are the same
    combinations of its tangents along d1 and d2."""
    X = np.asarray(X, dtype=np.float64)
    D1 = np.zeros_like(X)
    D1[:, 2], D1[:, 3] = X[:, 3], -X[:, 2]
    D1[:, 4], D1[:, 5] = X[:, 5], -X[:, 4]
    D1[:, 6], D1[:, 7] = 2 * X[:, 7], -2 * X[:, 6]
    D2 = np.zeros_like(X)
    D2[:, 10] = -X[:, 10]
    return D1, D2


LOG2E_F32 = np.float32(1.44269504088896341)
LN2_F32 = np.float32(0.693147180559945309)


def f16(a):
    """Round to fp16 (nearest even), returned as float64."""
    return np.asarray(a, dtype=np.float64).astype(np.float16).astype(np.float64)


def folded_inference_params(p):
    """The device's inference weights (farfield.hip, k_pack_model and
    k_far_learned): with L = log2 e, W1' = bf16((-L W1) / sd) applied to the
    centred features x - mu, b1' = -L b1, W2' = fp16(W2), b2' = -L b2,
    W3' = fp16(-W3 / L), b3; so layer 1 gives -L z, 2^(-L z) = e^-z and the
    hidden values carry the factor -L.  FP32 operations where the device
    uses FP32."""
    W1, b1, W2, b2, W3, b3, mu, sd, _ = unpack(np.asarray(p, dtype=np.float32))
    isd = (np.float32(1.0) / sd).astype(np.float32)
    W1f = bf16(((-LOG2E_F32) * W1).astype(np.float32) * isd[None, :]).astype(np.float64)
    b1f = ((-LOG2E_F32) * b1).astype(np.float32).astype(np.float64)
    b2f = ((-LOG2E_F32) * b2).astype(np.float32).astype(np.float64)
    W3f = f16((W3 * (-LN2_F32)).astype(np.float32))
    return W1f, b1f, f16(W2), b2f, W3f, b3.astype(np.float64)


def forward_jvp(p, X, TX, TY, mirror=True):
    """Raw outputs and their forward-mode tangents along TX and TY (inputs
    before standardisation).  mirror=True follows the device's inference
    path: the folded weights of folded_inference_params, bf16 centred
    features and bf16 raw tangents into layer 1, and fp16 hidden values and
    tangents (which carry the factor -L): per unit the matrix output z' and
    the sigmoid s = 1 / (1 + 2^z') (FP32) rounded to fp16, then
    h = fp16(z' s), d = fp16(h fp16(s ln2 - ln2) + s) (fused) and the
    tangents fp16(d fp16(t)), as the device's packed fp16 instructions do;
    mirror=False is the plain network in float64."""
    if mirror:
        r = lambda a: bf16(np.asarray(a, dtype=np.float32)).astype(np.float64)
        W1f, b1f, W2f, b2f, W3f, b3 = folded_inference_params(p)
        mu = unpack(np.asarray(p, dtype=np.float32))[6]
        xs, txs, tys = r((np.asarray(X, np.float32) - mu).astype(np.float32)), r(TX), r(TY)
        ln2 = f16(np.float32(LN2_F32))

        def layer(zp, tx, ty):
            z32 = np.asarray(zp, dtype=np.float32)
            s = f16((np.float32(1.0) / (np.float32(1.0) + np.exp2(z32))).astype(np.float32))
            h = f16(f16(z32) * s)
            d = f16(h * f16(s * ln2 - ln2) + s)
            return h, f16(d * f16(np.asarray(tx, np.float32))), f16(d * f16(np.asarray(ty, np.float32)))

        h1, t1x, t1y = layer(xs @ W1f.T + b1f, txs @ W1f.T, tys @ W1f.T)
        h2, t2x, t2y = layer(h1 @ W2f.T + b2f, t1x @ W2f.T, t1y @ W2f.T)
        return h2 @ W3f.T + b3, t2x @ W3f.T, t2y @ W3f.T
    W1, b1, W2, b2, W3, b3, mu, sd, _ = unpack(p)
    xs, txs, tys = (X - mu) / sd, TX / sd, TY / sd

    def layer(z, tx, ty):
        s = 1.0 / (1.0 + np.exp(-z))
        d = s * (1.0 + z * (1.0 - s))
        return z * s, d * tx, d * ty

    h1, t1x, t1y = layer(xs @ W1.T + b1, txs @ W1.T, tys @ W1.T)
    h2, t2x, t2y = layer(h1 @ W2.T + b2, t1x @ W2.T, t1y @ W2.T)
    return h2 @ W3.T + b3, t2x @ W3.T, t2y @ W3.T


def correction_jvp_dirs(p, X, D1, D2, scale, ux, uy, rinv, mirror=True):
    """correction_jvp with the network's tangents taken along the directions
    D1, D2 of feature_directions (as the device runs them) and combined into
    the tangents along x and y after the network."""
    y, j1, j2 = forward_jvp(p, X, D1, D2, mirror)
    thx, thy = -uy * rinv, ux * rinv
    yx = thx[:, None] * j1 + (ux * rinv)[:, None] * j2
    yy = thy[:, None] * j1 + (uy * rinv)[:, None] * j2
    return _expand(p, y, yx, yy, scale, ux, uy, rinv)


def correction_jvp(p, X, TX, TY, scale, ux, uy, rinv, mirror=True):
    """Correction (u, v) of pairs at their target points and its gradient
    with respect to the target position: columns u, v, du/dx, dv/dx, du/dy,
    dv/dy (network tangents, rotation back to the target frame and the 1/r^3
    of the scale)."""
    y, yx, yy = forward_jvp(p, X, TX, TY, mirror)
    return _expand(p, y, yx, yy, scale, ux, uy, rinv)


def _expand(p, y, yx, yy, scale, ux, uy, rinv):
    """Outputs and their tangents along x and y -> u, v and their gradient."""
    ysd = unpack(p)[8].astype(np.float64)
    a, b = y[:, 0] * ysd[0], y[:, 1] * ysd[1]
    U, V = (a * ux - b * uy) * scale, (a * uy + b * ux) * scale
    out = [U, V]
    for t, (ta, tb) in ((0, yx.T), (1, yy.T)):
        u_t = ux if t == 0 else uy
        th = -uy * rinv if t == 0 else ux * rinv
        k = -3 * u_t * rinv
        ga, gb = ta * ysd[0], tb * ysd[1]
        out.append(k * U - th * V + (ga * ux - gb * uy) * scale)
        out.append(k * V + th * U + (ga * uy + gb * ux) * scale)
    return np.stack(out, 1)
