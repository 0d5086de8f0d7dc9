"""CPU oracle of the learned band: gap features and the bf16 MLP.

TEST INFRASTRUCTURE ONLY: used by tests/ (and nothing in the product) as the
checker of the library's features kernel, forward pass and gradients.  There
is no counterpart in the reference (its NWSmall always fills the whole
matrix), so this restates the library's own documented definitions
(include/mauvehip.h, mauve-py_amd/csrc/mlp.hip header comment):

  features  f[0..2] log2 lengths / length difference over 16, f[3] 4-mer L1
            distance, f[4..259] clipped normalised 4-mer count differences,
            f[260..271] flank seed densities
  network   272 -> 256 -> 128 -> 1, ReLU, bf16 weights and activations
            (round to nearest even), fp32 accumulation, loss sum (y-t)^2/(2n)

Parity pins: the backward pass below is the derivative of the forward pass
(checked by central differences in float64 with rounding switched off,
tests/test_mlp_oracle.py); features are pinned by brute-force recounts on
small strings in the same test file.
"""
import numpy as np

NIN, NH1, NH2 = 272, 256, 128
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH1 * NIN
OFF_W2 = OFF_B1 + NH1
OFF_B2 = OFF_W2 + NH2 * NH1
OFF_W3 = OFF_B2 + NH2
OFF_B3 = OFF_W3 + NH2
NPARAM = OFF_B3 + 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(NH1, NIN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH2, NH1),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3:OFF_B3 + 1])


def forward(p, x, rnd=bf16, dtype=np.float32):
    W1, b1, W2, b2, w3, b3 = (a.astype(dtype) for a in unpack(p))
    xb = rnd(np.asarray(x, dtype=dtype))
    z1 = xb @ rnd(W1).T + b1
    h1 = rnd(np.where(z1 > 0, z1, 0).astype(dtype))
    z2 = h1 @ rnd(W2).T + b2
    h2 = rnd(np.where(z2 > 0, z2, 0).astype(dtype))
    y = h2 @ rnd(w3) + b3[0]
    return y, (xb, h1, h2)


def grad(p, x, t, rnd=bf16, dtype=np.float32):
    """(loss, flat gradient) of sum (y - t)^2 / (2n) for one batch."""
    W1, b1, W2, b2, w3, b3 = (a.astype(dtype) for a in unpack(p))
    y, (xb, h1, h2) = forward(p, x, rnd, dtype)
    n = len(y)
    d = y - np.asarray(t, dtype=dtype)
    loss = float((d * d).sum() / (2 * n))
    g = rnd((d / dtype(n)).astype(dtype))
    out = np.zeros(NPARAM, dtype=dtype)
    out[OFF_W3:OFF_B3] = h2.T @ g
    out[OFF_B3] = g.sum()
    dz2 = rnd((g[:, None] * rnd(w3)[None, :]).astype(dtype)) * (h2 > 0)
    out[OFF_W2:OFF_B2] = (dz2.T @ h1).ravel()
    out[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = rnd((dz2 @ rnd(W2)).astype(dtype)) * (h1 > 0)
    out[OFF_W1:OFF_B1] = (dz1.T @ xb).ravel()
    out[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, out


def _kmers(s, k):
    s = np.asarray(s, dtype=np.int64) & 3
    n = len(s) - k + 1
    if n <= 0:
        return np.zeros(0, dtype=np.int64)
    v = np.zeros(n, dtype=np.int64)
    for q in range(k):
        v = (v << 2) | s[q:q + n]
    return v


def features(a, b):
    """The 272 features of one gap (A, B letter codes)."""
    a = np.asarray(a, dtype=np.uint8)
    b = np.asarray(b, dtype=np.uint8)
    la, lb = len(a), len(b)
    f = np.zeros(NIN, dtype=np.float32)
    f32 = np.float32
    f[0] = np.log2(f32(la)) / f32(16)
    f[1] = np.log2(f32(lb)) / f32(16)
    f[2] = np.log2(f32(1) + f32(abs(la - lb))) / f32(16)
    ca = np.bincount(_kmers(a, 4), minlength=256)
    cb = np.bincount(_kmers(b, 4), minlength=256)
    na, nb = max(0, la - 3), max(0, lb - 3)
    tot = na + nb
    f[3] = f32(np.abs(ca - cb).sum()) / f32(max(1, tot))
    r = (ca - cb).astype(np.float32) * f32(512) / f32(max(tot, 512))
    f[4:260] = np.clip(r, f32(-8), f32(8)) / f32(8)
    if la >= 8 and lb >= 8:
        for side in range(2):
            aa, bb = (a, b) if side == 0 else (a[::-1], b[::-1])
            # an 8-mer read from the right end is the reversed string's 8-mer,
            # reversed; equality is unaffected by reversing both sides
            ka, kb = _kmers(aa, 8), _kmers(bb, 8)
            hits = np.zeros(512, dtype=bool)
            pv = np.arange(min(512, la - 7))
            for p in pv:
                qc = p * lb // la
                lo, hi = max(0, qc - 8), min(lb - 8, qc + 8)
                hits[p] = lo <= hi and bool((kb[lo:hi + 1] == ka[p]).any())
            for s in range(6):
                w = 16 << s
                v = min(w, len(pv))
                f[260 + 6 * side + s] = f32(hits[:v].sum()) / f32(v) if v else f32(0)
    return f
