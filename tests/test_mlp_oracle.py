"""Pins of the learned-band oracle (oracle/mlp_oracle.py), CPU only.

The backward pass is checked against central differences of the forward pass
in float64 with bf16 rounding switched off (a dropped term, a wrong sign or a
transposed operand fails it); bf16 rounding against known answers; the
features against brute-force recounts written with plain loops in the
forward coordinates of both ends.
"""
import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 0.0, 3.0e-3], dtype=np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 0.0], dtype=np.float32)
    got = mo.bf16(x)
    assert np.array_equal(got[:6], want)
    assert abs(got[6] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_backward_is_the_derivative_of_forward():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.08, mo.NPARAM)
    x = rng.normal(0, 1, (5, mo.NIN))
    t = rng.normal(0, 1, 5)
    ident = lambda v: v
    loss, g = mo.grad(p, x, t, rnd=ident, dtype=np.float64)

    def L(q):
        y, _ = mo.forward(q, x, rnd=ident, dtype=np.float64)
        return float(((y - t) ** 2).sum() / (2 * len(t)))

    assert abs(L(p) - loss) < 1e-12
    idx = np.concatenate([rng.choice(mo.NPARAM, 60, replace=False),
                          [mo.OFF_B3, mo.OFF_W3, mo.OFF_B2, mo.OFF_B1, mo.OFF_W2, mo.OFF_W1]])
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        up = L(q)
        q[i] -= 2 * eps
        dn = L(q)
        fd = (up - dn) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def _brute_features(a, b):
    la, lb = len(a), len(b)
    ca = np.zeros(256, dtype=np.int64)
    cb = np.zeros(256, dtype=np.int64)
    for s, c in ((a, ca), (b, cb)):
        for p in range(len(s) - 3):
            c[s[p] * 64 + s[p + 1] * 16 + s[p + 2] * 4 + s[p + 3]] += 1
    dens = np.zeros(12)
    if la >= 8 and lb >= 8:
        for side in range(2):
            for sc in range(6):
                w = 16 << sc
                hits = valid = 0
                for p in range(min(w, 512)):
                    if p + 8 > la:
                        continue
                    pa = p if side == 0 else la - 8 - p
                    qc = p * lb // la
                    hit = False
                    for q in range(qc - 8, qc + 9):
                        if q < 0 or q + 8 > lb:
                            continue
                        pb = q if side == 0 else lb - 8 - q
                        hit |= list(a[pa:pa + 8]) == list(b[pb:pb + 8])
                    valid += 1
                    hits += hit
                dens[6 * side + sc] = hits / valid if valid else 0.0
    return ca, cb, dens


def test_features_brute_force():
    rng = np.random.default_rng(4)
    for la, lb in [(1, 1), (3, 9), (7, 7), (8, 8), (9, 30), (40, 33), (100, 150), (600, 580), (700, 1400)]:
        a = rng.integers(0, 4, la, dtype=np.uint8)
        b = a.copy() if la == lb else rng.integers(0, 4, lb, dtype=np.uint8)
        if la == lb and la > 10:
            b[::7] = (b[::7] + 1) % 4
        f = mo.features(a, b)
        ca, cb, dens = _brute_features(a, b)
        tot = max(0, la - 3) + max(0, lb - 3)
        assert abs(f[0] - np.log2(la) / 16) < 1e-6 and abs(f[1] - np.log2(lb) / 16) < 1e-6
        assert abs(f[2] - np.log2(1 + abs(la - lb)) / 16) < 1e-6
        assert abs(f[3] - np.abs(ca - cb).sum() / max(1, tot)) < 1e-6
        want = np.clip((ca - cb) * 512 / max(tot, 512), -8, 8) / 8
        assert np.allclose(f[4:260], want, atol=1e-6)
        assert np.allclose(f[260:], dens, atol=1e-6)
    # identical sequences: every flank 8-mer is found on the diagonal
    a = rng.integers(0, 4, 2000, dtype=np.uint8)
    f = mo.features(a, a)
    assert np.all(f[260:] == 1.0) and f[3] == 0 and np.all(f[4:260] == 0)
