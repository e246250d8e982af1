"""The proposal-network oracle pinned without the GPU: its float64 backward
against central differences of its float64 forward (argmax routing through
the pool included), and its bf16 rounding against exact cases."""
import numpy as np

from oracle import proposal_oracle as PO


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    shapes = {"W1": (64, 3), "b1": (64,), "W2": (128, 64), "b2": (128,), "W3": (256, 128), "b3": (256,),
              "W4": (128, 256), "b4": (128,), "W5": (192, 128), "b5": (192,)}
    return {k: rng.normal(size=s) * scale / np.sqrt(s[-1] if len(s) > 1 else 4) for k, s in shapes.items()}


def test_bf16_rounding_exact_cases():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, 1.0 + 3 * 2 ** -8, -0.0, np.inf, 3.0e38],
                 dtype=np.float32)
    got = PO.bf16_round(x)
    # bf16 spacing at 1 is 2^-7; ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6;
    # below the tie rounds down: 1 + 2^-7 + 2^-9 -> 1 + 2^-7
    assert got[0] == 1.0 and got[1] == 1.0
    assert got[2] == np.float32(1.0 + 2 ** -7)
    assert got[3] == np.float32(1.0 + 2 ** -6)
    assert np.signbit(got[4]) and got[5] == np.inf
    assert np.isnan(PO.bf16_round(np.array([np.nan], dtype=np.float32)))[0]
    # every bf16 value is a fixed point
    vals = (np.arange(0, 1 << 16, 7, dtype=np.uint32) << 16).view(np.float32)
    vals = vals[np.isfinite(vals)]
    assert np.array_equal(PO.bf16_round(vals).view(np.uint32), vals.view(np.uint32))


def test_backward_matches_central_differences():
    rng = np.random.default_rng(1)
    prm = _params(2)
    pts = rng.normal(size=(2, 40, 3))
    dout = rng.normal(size=(2, 192))
    o = PO.Oracle(prm, bf16=False)
    out, cache = o.forward(pts)
    G = o.backward(dout, cache)
    h = 1e-6
    for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4", "W5", "b5"):
        flat = prm[k].reshape(-1)
        picks = rng.choice(flat.size, size=min(6, flat.size), replace=False)
        for i in picks:
            pp = {kk: v.copy() for kk, v in prm.items()}
            pp[k].reshape(-1)[i] += h
            lp = (PO.Oracle(pp, bf16=False).forward(pts)[0] * dout).sum()
            pp[k].reshape(-1)[i] -= 2 * h
            lm = (PO.Oracle(pp, bf16=False).forward(pts)[0] * dout).sum()
            fd = (lp - lm) / (2 * h)
            an = G[k].reshape(-1)[i]
            assert abs(fd - an) <= 1e-5 * (1 + abs(fd)), (k, i, fd, an)


def test_adam_formula():
    p = np.array([1.0, -2.0], dtype=np.float32)
    g = np.array([0.5, -0.25], dtype=np.float32)
    p1, m1, v1 = PO.adam(p, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    # first step: m_hat = g, v_hat = g^2, so the update is lr * sign(g) (up to eps)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-6)
