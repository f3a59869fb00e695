"""The FP64 residual-energy model of the gradient-tuned form (oracle_round_energy),
pinned without a GPU:

* against an independent pure-Python enumeration that walks the reference's
  block and target loops (compress.tpp:51-88, 179-208) on a small field;
* by the algebra the model must obey: E(t) is a quadratic whose derivative is
  the returned gradient (central differences, h = 1e-3), and whose Hessian
  2 sum A^2 follows from two gradients;
* by fields on which a predictor is exact: a cubic polynomial along the round
  axis leaves the interior 1D4R residual zero at t = 5/16 (cubic Lagrange,
  s1_compress.cpp:96-97) and a quintic polynomial leaves the 1D6R residual
  zero at the reference weights a = 3/256, b = -25/256 (asz_param.hpp:72-74),
  so those are the minimisers.
"""
import numpy as np
import pytest

from oracle import api

PLAN_AXIS = [0, 2, 1, 2, 0, 1]


def _strides(dims, axis, bpr, round_):
    ps, bs, bn = api.geometry(dims, axis, bpr)
    s = [int(x) for x in ps]
    for r in range(round_):
        s[axis[r]] //= 2
    return s, [int(x) for x in bs], [int(x) for x in bn]


def brute_energy(state, axis, bpr, round_, quintic, theta):
    """Pure-Python restatement of the reference's target loops for one round."""
    dims = state.shape
    s, B, nb = _strides(dims, axis, bpr, round_)
    d = axis[round_]
    h = s[d] // 2
    reach = 5 if quintic else 3
    E = np.zeros(nb[0] * nb[1] * nb[2])
    G = np.zeros(theta.size)
    x = state.astype(np.float64)
    for b0 in range(nb[0]):
        for b1 in range(nb[1]):
            for b2 in range(nb[2]):
                bid = (b0 * nb[1] + b1) * nb[2] + b2
                bb = (b0, b1, b2)
                lo = [bb[a] * B[a] for a in range(3)]
                hi = [min(lo[a] + B[a] + 1, dims[a]) for a in range(3)]  # exclusive (compress.tpp:183-188)
                rng = []
                for a in range(3):
                    if a == d:
                        rng.append(range(lo[a] + h, min(lo[a] + B[a], dims[a]), 2 * h))
                    else:
                        start = lo[a] if lo[a] == 0 else lo[a] + s[a]
                        rng.append(range(start, hi[a], s[a]))
                for i0 in rng[0]:
                    for i1 in rng[1]:
                        for i2 in rng[2]:
                            i = (i0, i1, i2)[d]
                            right = hi[d]
                            if i < reach * h or i + reach * h >= right:
                                continue
                            def nbv(k):
                                idx = [i0, i1, i2]
                                idx[d] += k * h
                                return x[tuple(idx)]
                            x0 = x[i0, i1, i2]
                            if not quintic:
                                A = (nbv(-1) + nbv(1)) - (nbv(-3) + nbv(3))
                                Bv = 0.25 * (nbv(-1) + nbv(1) + nbv(-3) + nbv(3)) - x0
                                r = A * theta[bid] + Bv
                                G[bid] += 2 * r * A
                            else:
                                S1, S3, S5 = nbv(-1) + nbv(1), nbv(-3) + nbv(3), nbv(-5) + nbv(5)
                                a_, b_ = theta[2 * bid], theta[2 * bid + 1]
                                r = a_ * (S5 - S1) + b_ * (S3 - S1) + 0.5 * S1 - x0
                                G[2 * bid] += 2 * r * (S5 - S1)
                                G[2 * bid + 1] += 2 * r * (S3 - S1)
                            E[bid] += r * r
    return E, G


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    g = [np.linspace(0, 3, n) for n in shape]
    f = (np.sin(g[0])[:, None, None] * np.cos(1.3 * g[1])[None, :, None] + np.sin(0.7 * g[2])[None, None, :])
    return (f + 0.05 * rng.standard_normal(shape)).astype(np.float32)


@pytest.mark.parametrize("round_", [2, 4, 5])
@pytest.mark.parametrize("quintic", [False, True])
def test_oracle_energy_matches_reference_loops(round_, quintic):
    state = _field((23, 19, 26), 3)
    bpr = 2
    _, _, nb = _strides(state.shape, PLAN_AXIS, bpr, round_)
    nblk = int(np.prod(nb))
    rng = np.random.default_rng(round_)
    theta = rng.uniform(0.05, 0.35, nblk) if not quintic else np.stack(
        [rng.uniform(-0.02, 0.04, nblk), rng.uniform(-0.15, 0.0, nblk)], 1).ravel()
    E, G = api.round_energy(state, PLAN_AXIS, bpr, round_, quintic, theta)
    Eb, Gb = brute_energy(state, PLAN_AXIS, bpr, round_, quintic, theta)
    assert np.allclose(E, Eb, rtol=1e-12, atol=1e-12)
    assert np.allclose(G, Gb, rtol=1e-10, atol=1e-10)
    if not quintic or round_ >= 4:  # coarse rounds have no 1D6R targets in this small field
        assert E.sum() > 0


def test_energy_is_quadratic_with_the_returned_gradient():
    state = _field((40, 36, 44), 5)
    axis, bpr, r = PLAN_AXIS, 3, 5
    _, _, nb = _strides(state.shape, axis, bpr, r)
    nblk = int(np.prod(nb))
    t = np.full(nblk, 0.25)
    E0, g0 = api.round_energy(state, axis, bpr, r, False, t)
    h = 1e-3
    Ep, _ = api.round_energy(state, axis, bpr, r, False, t + h)
    Em, _ = api.round_energy(state, axis, bpr, r, False, t - h)
    fd = (Ep - Em) / (2 * h)
    scale = np.maximum(np.abs(g0), 1e-12 * E0.max())
    assert np.all(np.abs(fd - g0) <= 1e-7 * scale + 1e-9 * E0)
    # Hessian from two gradients, and the quadratic identity at t = 0.4
    _, g1 = api.round_energy(state, axis, bpr, r, False, t + 0.05)
    H = (g1 - g0) / 0.05
    assert np.all(H >= 0)
    E4, _ = api.round_energy(state, axis, bpr, r, False, np.full(nblk, 0.4))
    pred = E0 + g0 * 0.15 + 0.5 * H * 0.15 ** 2
    assert np.allclose(E4, pred, rtol=1e-9, atol=1e-9 * E0.max())
    # the unconstrained minimiser -sum AB / sum A^2 beats every grid point
    tstar = t - g0 / np.where(H > 0, H, 1)
    Es, _ = api.round_energy(state, axis, bpr, r, False, tstar)
    for ti in range(2, 13):
        Eg, _ = api.round_energy(state, axis, bpr, r, False, np.full(nblk, ti / 32))
        assert np.all(Es <= Eg + 1e-9 * E0.max())


def _poly_field(shape, axis_, degree):
    i = np.arange(shape[axis_], dtype=np.float64) / shape[axis_]
    p = sum((k + 1) * 0.3 * i ** k for k in range(degree + 1))
    sh = [1, 1, 1]
    sh[axis_] = shape[axis_]
    return np.broadcast_to(p.reshape(sh), shape).astype(np.float32).copy()


def test_cubic_lagrange_is_the_minimiser_on_cubic_fields():
    axis = [0, 1, 2, 2, 1, 0]
    r = 4  # a round along axis 1
    state = _poly_field((30, 64, 30), 1, 3)
    _, _, nb = _strides(state.shape, axis, 2, r)
    nblk = int(np.prod(nb))
    E_lag, g_lag = api.round_energy(state, axis, 2, r, False, np.full(nblk, 5 / 16))
    E_lin, g_lin = api.round_energy(state, axis, 2, r, False, np.full(nblk, 4 / 16))
    assert E_lin.sum() > 0
    # the field is float32: its rounding (~1e-7 relative) is the floor of the residual
    assert E_lag.sum() <= 1e-7 * E_lin.sum()
    # Newton from 1/4 lands on 5/16
    H = (g_lag - g_lin) / (1 / 16)
    live = H > 1e-6 * H.max()
    tstar = 0.25 - g_lin[live] / H[live]
    assert np.allclose(tstar, 5 / 16, atol=1e-3)


def test_quintic_weights_are_the_minimiser_on_quintic_fields():
    axis = [1, 2, 2, 2, 0, 2]
    r = 5  # a round along axis 2 with stride 2 (h = 1)
    state = _poly_field((24, 20, 96), 2, 5)
    _, _, nb = _strides(state.shape, axis, 8, r)
    nblk = int(np.prod(nb))
    ref = np.tile([3 / 256, -25 / 256], nblk)
    E_ref, _ = api.round_energy(state, axis, 8, r, True, ref)
    E_off, _ = api.round_energy(state, axis, 8, r, True, np.tile([0.0, -0.1], nblk))
    assert E_off.sum() > 0
    assert E_ref.sum() <= 1e-7 * E_off.sum()
