"""GPU parity of the learned-band kernels against the CPU oracle
(oracle/mlp_oracle.py, pinned by tests/test_mlp_oracle.py).

Tolerances (floating point): features f[3:] bit-exact (integer counts and
correctly rounded fp32 divisions), f[:3] within 1e-6 (device log2f);
network output within 2e-3 + 1e-2 |y|, gradients within a relative L2 error
of 2e-2 per parameter block -- the device accumulates in a different order
(MFMA), which can move a bf16 rounding of an activation by one unit.
"""
import numpy as np
import pytest

import mauvehip
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _jobs(rng, shapes):
    A, B, rows = [], [], []
    ao = bo = 0
    for la, lb in shapes:
        a = rng.integers(0, 4, la, dtype=np.uint8)
        if la == lb:
            b = a.copy()
            hit = rng.random(la) < 0.1
            b[hit] = rng.integers(0, 4, int(hit.sum()), dtype=np.uint8)
        else:
            b = rng.integers(0, 4, lb, dtype=np.uint8)
        A.append(a)
        B.append(b)
        rows.append((ao, la, bo, lb))
        ao += la
        bo += lb
    r = np.array(rows, dtype=np.int64)
    return np.concatenate(A), np.concatenate(B), mauvehip.make_jobs(r[:, 0], r[:, 1], r[:, 2], r[:, 3])


def test_features_match_oracle(ctx):
    rng = np.random.default_rng(21)
    shapes = [(1, 1), (3, 5), (7, 9), (8, 8), (9, 200), (64, 64), (500, 500), (3000, 2900), (20000, 20000),
              (1500, 40), (40, 1500)] + [(int(x), int(x)) for x in rng.integers(10, 4000, 40)]
    A, B, jobs = _jobs(rng, shapes)
    dA, dB = ctx.alloc(len(A) + 16), ctx.alloc(len(B) + 16)
    dF = ctx.alloc(4 * 272 * len(jobs))
    ctx.h2d(dA, A)
    ctx.h2d(dB, B)
    mauvehip.gap_features_device(ctx, dA, dB, jobs, dF)
    F = np.zeros((len(jobs), 272), dtype=np.float32)
    ctx.d2h(F, dF)
    for k, j in enumerate(jobs):
        a = A[j["a_off"]:j["a_off"] + j["a_len"]]
        b = B[j["b_off"]:j["b_off"] + j["b_len"]]
        want = mo.features(a, b)
        assert np.array_equal(F[k, 3:], want[3:]), k
        assert np.allclose(F[k, :3], want[:3], rtol=0, atol=1e-6), k
    for p in (dA, dB, dF):
        ctx.free(p)


def _net_and_batch(rng, n):
    p = rng.normal(0, 0.06, mo.NPARAM).astype(np.float32)
    p[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.05, mo.NH1)
    x = rng.normal(0, 1, (n, mo.NIN)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    return p, x, t


@pytest.mark.parametrize("n", [1, 37, 1024, 5000])
def test_forward_matches_oracle(ctx, n):
    rng = np.random.default_rng(n)
    p, x, _ = _net_and_batch(rng, n)
    net = mauvehip.Mlp(ctx, params=p)
    dX, dY = ctx.alloc(x.nbytes), ctx.alloc(4 * n + 16)
    ctx.h2d(dX, x)
    net.forward_device(dX, n, dY)
    y = np.zeros(n, dtype=np.float32)
    ctx.d2h(y, dY)
    want, _ = mo.forward(p, x)
    assert np.all(np.abs(y - want) <= 2e-3 + 1e-2 * np.abs(want))
    assert np.array_equal(net.params(), p)
    for q in (dX, dY):
        ctx.free(q)
    net.free()


@pytest.mark.parametrize("n", [16, 300, 2048])
def test_gradients_match_oracle(ctx, n):
    rng = np.random.default_rng(100 + n)
    p, x, t = _net_and_batch(rng, n)
    net = mauvehip.Mlp(ctx, params=p)
    dX, dT = ctx.alloc(x.nbytes), ctx.alloc(t.nbytes)
    ctx.h2d(dX, x)
    ctx.h2d(dT, t)
    loss, g = net.grad_device(dX, dT, n)
    wl, wg = mo.grad(p, x, t)
    assert abs(loss - wl) <= 1e-3 * max(1.0, abs(wl))
    blocks = {"W1": (mo.OFF_W1, mo.OFF_B1), "b1": (mo.OFF_B1, mo.OFF_W2), "W2": (mo.OFF_W2, mo.OFF_B2),
              "b2": (mo.OFF_B2, mo.OFF_W3), "w3": (mo.OFF_W3, mo.OFF_B3), "b3": (mo.OFF_B3, mo.NPARAM)}
    for name, (lo, hi) in blocks.items():
        err = np.linalg.norm(g[lo:hi] - wg[lo:hi]) / max(np.linalg.norm(wg[lo:hi]), 1e-12)
        assert err <= 2e-2, (name, err)
    for q in (dX, dT):
        ctx.free(q)
    net.free()


def test_training_fits_a_smooth_target(ctx):
    """Adam on the device drives the loss of a learnable target down."""
    rng = np.random.default_rng(7)
    n = 16384
    x = rng.normal(0, 1, (n, mo.NIN)).astype(np.float32)
    w = rng.normal(0, 1, mo.NIN).astype(np.float32) / np.sqrt(mo.NIN)
    t = (np.tanh(x @ w) * 2 + 0.5 * np.maximum(x[:, 0], 0)).astype(np.float32)
    dX, dT = ctx.alloc(x.nbytes), ctx.alloc(t.nbytes)
    ctx.h2d(dX, x)
    ctx.h2d(dT, t)
    net = mauvehip.Mlp(ctx, seed=3)
    base = float(((t - t.mean()) ** 2).mean() / 2)
    curve = net.train_device(dX, dT, n, epochs=8, batch=512, lr=2e-3, seed=3)
    assert curve[-1] < 0.2 * base, (curve, base)
    assert curve[-1] < curve[0]
    for q in (dX, dT):
        ctx.free(q)
    net.free()
