"""The CPU restatement of the learned surrogate (oracle/net_oracle.py) pinned
before it is used to check the GPU kernels: BF16 rounding against known bit
patterns, and the backward pass against central differences of its own loss
(plain float64, no rounding)."""
import numpy as np

from oracle import net_oracle as no


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 3.140625, 1e-3], dtype=np.float32)
    r = no.bf16(x)
    # 1 + 2^-8 is a tie: rounds to even (1.0); 1 + 3 * 2^-9 rounds up to 1 + 2^-7
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == 1.0
    assert r[4] == -2.5 and r[5] == 3.140625
    assert abs(r[6] - 1e-3) <= 2 ** -9 * 1e-3


def test_backward_matches_finite_differences():
    M, C, r, L, B = 6, 4, 3, 2, 3
    off, n = no.layout(M, C, r, L)
    rng = np.random.default_rng(0)
    p = rng.standard_normal(n) * 0.4
    net = no.NetOracle(M, C, r, L, p)
    X = rng.standard_normal((M, B, 4))
    T = rng.standard_normal((M, B, 2))
    loss, g = net.loss_grad(X, T, rnd=False)
    idx = rng.choice(n, 40, replace=False)
    for i in list(idx) + [off["ob"][0], off["pb"][0], off["b10"][0], off["U1"][0]]:
        h = 1e-6
        pp = p.copy(); pp[i] += h
        lp, _ = no.NetOracle(M, C, r, L, pp).loss_grad(X, T, rnd=False)
        pp[i] -= 2 * h
        lm, _ = no.NetOracle(M, C, r, L, pp).loss_grad(X, T, rnd=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
