"""CPU checks of the learned-far-field oracle (oracle/far_mlp.py and the
vvo_far_* functions): the network's gradient against finite differences,
bf16 rounding against torch, node moments against a direct computation, and
the labels against the exact field and the second-order multipole limit."""
import os
import sys

import numpy as np
import pytest

from cases import random_cluster_case
from oracle_bind import Oracle, direct_sum

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import far_mlp as M  # noqa: E402


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(size=4000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 4000),
                        np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 0.0, -0.0], np.float32)])
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(a).view(np.uint32), ref.view(np.uint32))


def test_mlp_gradient_finite_differences():
    """float64 backward (no rounding) against central differences."""
    rng = np.random.default_rng(1)
    p = M.init_params(rng, mu=rng.normal(size=12) * 0.1, sd=rng.uniform(0.5, 2, 12), ysd=[0.7, 1.3]).astype(np.float64)
    p[M.O_B1:M.O_W2] = rng.normal(size=64) * 0.1
    p[M.O_B2:M.O_W3] = rng.normal(size=64) * 0.1
    p[M.O_B3:M.N_TRAIN] = rng.normal(size=2) * 0.1
    X = rng.normal(size=(40, 12))
    Y = rng.normal(size=(40, 2))
    W = rng.uniform(0.5, 1.5, 40)

    def loss_of(q):
        return M.loss_and_grad(q, X, Y, W, mirror=False)[0]

    _, g = M.loss_and_grad(p, X, Y, W, mirror=False)
    idx = np.concatenate([rng.choice(M.N_TRAIN, 60, replace=False), [M.O_B3, M.O_B3 + 1, M.O_B1, M.O_B2]])
    for k in idx:
        h = 1e-6 * max(1.0, abs(p[k]))
        qp, qm = p.copy(), p.copy()
        qp[k] += h
        qm[k] -= h
        fd = (loss_of(qp) - loss_of(qm)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_mlp_mirror_close_to_exact():
    """The bf16 rounding points change outputs and gradients by bf16-level amounts only."""
    rng = np.random.default_rng(2)
    p = M.init_params(rng)
    X = rng.normal(size=(256, 12)).astype(np.float32)
    Y = rng.normal(size=(256, 2)).astype(np.float32)
    W = np.ones(256, np.float32)
    y0, _ = M.forward(p, X, mirror=False)
    y1, _ = M.forward(p, X, mirror=True)
    assert np.linalg.norm(y1 - y0) < 2e-2 * np.linalg.norm(y0)
    _, g0 = M.loss_and_grad(p, X, Y, W, mirror=False)
    _, g1 = M.loss_and_grad(p, X, Y, W, mirror=True)
    assert np.linalg.norm(g1 - g0) < 3e-2 * np.linalg.norm(g0)


def _case(n=3000, seed=5, fc=8):
    objs = random_cluster_case(n, seed, "mixed")
    o = Oracle()
    o.set_params(re=600.0, dt=0.005, inf=(0.0, 0.0))
    o.set_vortices(objs)
    o.tree_build(fc)
    o.epsilon(True)
    return o


def test_moments_direct():
    """Leaves: sums over their vortices.  Internal nodes (parallel-axis
    combination of the children) against the same sums taken directly."""
    o = _case()
    T = o.tree()
    v = o.vortices()
    mom = o.far_moments()
    ints = T["ints"]
    for nd in range(len(ints)):
        a, b = ints[nd, 0], ints[nd, 1]
        for s, sel in ((0, v[a:b, 2] > 0), (1, v[a:b, 2] < 0)):
            g = v[a:b, 2][sel]
            z = v[a:b, 0][sel] + 1j * v[a:b, 1][sel]
            eps2 = (1.0 / v[a:b, 5][sel]) ** 2
            G = g.sum()
            assert np.isclose(mom[s, nd], G, rtol=1e-12, atol=1e-18)
            if G == 0:
                assert np.array_equal(mom[[6 + 2 * s, 7 + 2 * s], nd], [0, 0])
                continue
            zc = np.sum(g * z) / G
            Q = np.sum(g * (z - zc) ** 2)
            assert abs(mom[2 + 2 * s, nd] + 1j * mom[3 + 2 * s, nd] - zc) <= 1e-12 * (1 + abs(zc))
            # relative bound plus a floor: centroid offsets carry ~1e-15 of the coordinates
            tolQ = 1e-9 * np.sum(np.abs(g) * np.abs(z - zc) ** 2) + np.sum(np.abs(g)) * (1e-15 * (1 + abs(zc))) ** 2
            assert abs(mom[6 + 2 * s, nd] + 1j * mom[7 + 2 * s, nd] - Q) <= tolQ
            assert np.isclose(mom[10 + s, nd], np.sum(g * eps2), rtol=1e-11)


def _pairs(o, every=7):
    T = o.tree()
    leaves = T["leaves"]
    foff, flst = T["far"]
    ints = T["ints"]
    tg, nd = [], []
    for l, L in enumerate(leaves):
        for i in range(ints[L, 0], ints[L, 1]):
            for F in flst[foff[l]:foff[l + 1]]:
                tg.append(i)
                nd.append(F)
    return np.array(tg[::every]), np.array(nd[::every]), T


def test_pair_exact_field_is_direct_sum_over_node():
    o = _case(2000, 6)
    tg, nd, T = _pairs(o, 101)
    v = o.vortices()
    X, Y, sc, base, exact = o.far_pairs(tg, nd)
    for p in range(0, len(tg), 13):
        a, b = T["ints"][nd[p], 0], T["ints"][nd[p], 1]
        src = v[a:b].copy()
        ref = direct_sum(np.concatenate([src, v[tg[p]:tg[p] + 1] * [1, 1, 0, 1, 1, 1]]),
                         targets=np.array([len(src)]), threads=1)[0]
        assert np.allclose(exact[p], ref, rtol=1e-12, atol=1e-18)


def test_pair_labels_second_order_limit():
    """Far from the node the residual (exact - monopoles) tends to the
    quadrupole and smoothing terms: in the model's frame the label tends to
    (f7, f6 - f8 - f9), with a remainder of order d/r."""
    o = _case(4000, 7, fc=16)
    tg, nd, _ = _pairs(o, 3)
    X, Y, sc, base, exact = o.far_pairs(tg, nd)
    ok = sc > 0
    X, Y = X[ok], Y[ok]
    pred = np.stack([X[:, 7], X[:, 6] - X[:, 8] - X[:, 9]], 1)
    d_over_r = X[:, 10]
    close = d_over_r < 0.08
    assert close.sum() > 200
    err = np.linalg.norm(Y[close] - pred[close]) / np.linalg.norm(pred[close])
    assert err < 0.15, err
    # and the leading term carries most of the residual everywhere
    assert np.linalg.norm(Y - pred) < 0.5 * np.linalg.norm(Y)


def test_single_vortex_node_residual():
    """A node holding one vortex: Q = 0 and the residual is the smoothing
    term alone, g rotl(D) (1/(r^2 + eps^2) - 1/r^2) / (2 pi)."""
    objs = random_cluster_case(1500, 8, "mixed")
    lone = np.array([[30.0, 30.0, 2e-3, 0, 0, 0], [-25.0, 4.0, -1e-3, 0, 0, 0]])
    o = Oracle()
    o.set_params(re=600.0, dt=0.005, inf=(0.0, 0.0))
    o.set_vortices(np.concatenate([objs, lone]))
    o.tree_build(8)
    o.epsilon(True)
    T = o.tree()
    ints = T["ints"]
    v = o.vortices()
    singles = [k for k in range(len(ints)) if ints[k, 1] - ints[k, 0] == 1 and v[ints[k, 0], 2] != 0]
    assert singles
    tg, nd = [], []
    for F in singles[:20]:
        tg.append(0 if ints[F, 0] != 0 else 1)
        nd.append(F)
    X, Y, sc, base, exact = o.far_pairs(np.array(tg), np.array(nd))
    for p, F in enumerate(nd):
        j = ints[F, 0]
        D = v[tg[p], :2] - v[j, :2]
        r2 = D @ D
        e2 = (1.0 / v[j, 5]) ** 2
        want = v[j, 2] * np.array([-D[1], D[0]]) * (1 / (r2 + e2) - 1 / r2) / (2 * np.pi)
        assert np.allclose(exact[p] - base[p], want, rtol=1e-6, atol=1e-16)
        assert X[p, 6] == 0 and X[p, 7] == 0


def test_point_features_match_pair_features():
    """Features at arbitrary points (the learned path's three points per
    leaf) are the pair features when the point is the vortex itself, and u
    is the unit vector from the node centre to the point."""
    o = _case(2000, 9)
    tg, nd, T = _pairs(o, 11)
    v = o.vortices()
    X, _, sc, _, _ = o.far_pairs(tg, nd)
    Xp, scp, u = o.far_point_features(v[tg, 0], v[tg, 1], nd)
    assert np.array_equal(X, Xp) and np.array_equal(sc, scp)
    D = v[tg, :2] - T["geom"][nd, :2]
    ok = sc > 0
    assert np.allclose(u[ok], D[ok] / np.hypot(D[ok, 0], D[ok, 1])[:, None], rtol=1e-14, atol=1e-15)
    assert np.all(u[~ok] == 0)


def test_feature_tangents_finite_differences():
    """Feature derivatives with respect to the target point (the learned
    path's expansion about the leaf centre) against central differences of
    the oracle's point features."""
    o = _case(2000, 10)
    tg, nd, T = _pairs(o, 9)
    v = o.vortices()
    X, sc, u = o.far_point_features(v[tg, 0], v[tg, 1], nd)
    ok = sc > 0
    D = v[tg, :2] - T["geom"][nd, :2]
    rinv = 1.0 / np.hypot(D[:, 0], D[:, 1])
    TX, TY = M.feature_tangents(X, u[:, 0], u[:, 1], rinv)
    h = 1e-5 / rinv
    for k, T_ in ((0, TX), (1, TY)):
        dp = np.zeros((len(tg), 2))
        dp[:, k] = h
        Xp = o.far_point_features(v[tg, 0] + dp[:, 0], v[tg, 1] + dp[:, 1], nd)[0]
        Xm = o.far_point_features(v[tg, 0] - dp[:, 0], v[tg, 1] - dp[:, 1], nd)[0]
        fd = (Xp - Xm) / (2 * h[:, None])
        scale = np.abs(X[ok]).max(0) * rinv[ok].max() + 1e-300
        assert np.all(np.abs(fd[ok] - T_[ok]) <= 1e-6 * scale + 1e-6 * np.abs(T_[ok]))


def test_network_jvp_finite_differences():
    """Forward-mode tangents of the network (float64) against central
    differences of its outputs along the tangent direction."""
    rng = np.random.default_rng(12)
    p = M.init_params(rng).astype(np.float64)
    p[M.O_B1:M.O_W2] = rng.normal(0, 0.3, M.N_HID)
    X = rng.normal(size=(50, 12))
    TX, TY = rng.normal(size=(50, 12)), rng.normal(size=(50, 12))
    y, yx, yy = M.forward_jvp(p, X, TX, TY, mirror=False)
    eps = 1e-6
    for T_, yt in ((TX, yx), (TY, yy)):
        fd = (M.forward(p, X + eps * T_, mirror=False)[0] - M.forward(p, X - eps * T_, mirror=False)[0]) / (2 * eps)
        assert np.allclose(fd, yt, rtol=1e-6, atol=1e-8)
    assert np.allclose(y, M.forward(p, X, mirror=False)[0], rtol=1e-13, atol=1e-15)


def test_correction_gradient_finite_differences():
    """The correction's gradient with respect to the target position
    (network tangents, rotation, 1/r^3 scale) against central differences of
    the correction at shifted points (float64 network)."""
    o = _case(2000, 11)
    tg, nd, T = _pairs(o, 13)
    v = o.vortices()
    rng = np.random.default_rng(5)
    X0, sc0, u0 = o.far_point_features(v[tg, 0], v[tg, 1], nd)
    ok = sc0 > 0
    tg, nd = tg[ok], nd[ok]
    p = M.init_params(rng, mu=X0[ok].mean(0), sd=X0[ok].std(0) + 1e-6).astype(np.float64)
    X, sc, u = o.far_point_features(v[tg, 0], v[tg, 1], nd)
    D = v[tg, :2] - T["geom"][nd, :2]
    rinv = 1.0 / np.hypot(D[:, 0], D[:, 1])
    TX, TY = M.feature_tangents(X, u[:, 0], u[:, 1], rinv)
    C = M.correction_jvp(p, X, TX, TY, sc, u[:, 0], u[:, 1], rinv, mirror=False)
    assert np.allclose(C[:, :2], M.correction(p, X, sc, u[:, 0], u[:, 1], mirror=False), rtol=1e-12, atol=0)
    h = 1e-6 / rinv
    for k in (0, 1):
        dp = np.zeros((len(tg), 2))
        dp[:, k] = h
        cp = M.correction(p, *_point(o, v[tg, :2] + dp, nd), mirror=False)
        cm = M.correction(p, *_point(o, v[tg, :2] - dp, nd), mirror=False)
        fd = (cp - cm) / (2 * h[:, None])
        ref = np.abs(C[:, 2 + 2 * k:4 + 2 * k]).max()
        assert np.max(np.abs(fd - C[:, 2 + 2 * k:4 + 2 * k])) < 1e-5 * ref
    # the device's form: tangents along the two fixed directions, combined
    # into x and y after the network, give the same gradient
    Cd = M.correction_jvp_dirs(p, X, *M.feature_directions(X), sc, u[:, 0], u[:, 1], rinv, mirror=False)
    assert np.max(np.abs(Cd - C)) < 1e-10 * np.abs(C).max()


def _point(o, P, nd):
    X, sc, u = o.far_point_features(P[:, 0], P[:, 1], nd)
    return X, sc, u[:, 0], u[:, 1]


def test_folded_inference_is_the_same_network():
    """The device's inference weights fold log2 e and the features' 1/sd
    into the layers (folded_inference_params); through them,
    with the device's bf16 rounding points, the outputs and both tangents
    stay within bf16 precision of the plain float64 network (a wrong sign or
    factor in the fold is off by order one)."""
    rng = np.random.default_rng(21)
    mu = rng.normal(size=M.N_IN).astype(np.float32)
    sd = rng.uniform(0.05, 2.0, M.N_IN).astype(np.float32)
    p = M.init_params(rng, mu=mu, sd=sd, ysd=np.float32([0.03, 0.07]))
    X = (mu + sd * rng.normal(size=(3000, M.N_IN))).astype(np.float32)
    TX, TY = rng.normal(size=(2, 3000, M.N_IN)) * sd
    got = M.forward_jvp(p, X, TX, TY, mirror=True)
    want = M.forward_jvp(p, X.astype(np.float64), TX, TY, mirror=False)
    for g, w in zip(got, want):
        assert np.linalg.norm(g - w) < 2e-2 * np.linalg.norm(w)
    W1f, b1f, W2f, b2f, W3f, b3 = M.folded_inference_params(p)
    W1, b1 = M.unpack(p)[:2]
    # layer 1 of the folded network is -log2(e) times the standardised one
    z = ((X.astype(np.float64) - mu) / sd) @ W1.T.astype(np.float64) + b1
    zf = (X.astype(np.float64) - mu) @ W1f.T + b1f
    assert np.linalg.norm(zf + np.log2(np.e) * z) < 1e-2 * np.linalg.norm(z)
