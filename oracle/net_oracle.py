"""CPU restatement of the learned collision surrogate (test infrastructure only).

The network of spectralbte_amd/csrc/spbte_net.c written out in numpy, so the GPU
forward and backward kernels (BF16 MFMA GEMMs with FP32 sums, elementwise
kernels) can be checked against it.  There is no reference counterpart: the
reference has no learned component.  GEMM operands are rounded to BF16 (round to
nearest even) exactly where the GPU rounds them (every operand of every
product); everything else is float64 here.  With ``rnd=False`` the restatement is
plain float64 and its gradient is checked against central differences of its
own loss (tests/test_net_oracle.py), which pins the backward formulas.

Layout (as on the GPU): activations [m][b][c] (mode-major); parameters one flat
vector with blocks P (4 x C), pb (C), E (M x C), per layer U (M x r), V (M x r),
W1 (C x 2C), b1 (2C), W2 (2C x C), b2 (C), then O (C x 2), ob (2).
"""
import numpy as np
from scipy.special import erf


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def _gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def _gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def layout(M, C, r, L):
    """Offsets and shapes of the parameter blocks (same order as spbte_net_layout)."""
    shapes = [("P", (4, C)), ("pb", (C,)), ("E", (M, C))]
    for l in range(L):
        shapes += [(f"U{l}", (M, r)), (f"V{l}", (M, r)), (f"W1{l}", (C, 2 * C)), (f"b1{l}", (2 * C,)),
                   (f"W2{l}", (2 * C, C)), (f"b2{l}", (C,))]
    shapes += [("O", (C, 2)), ("ob", (2,))]
    off, o = {}, 0
    for name, sh in shapes:
        off[name] = (o, sh)
        o += int(np.prod(sh))
    return off, o


class NetOracle:
    def __init__(self, M, C, r, L, params):
        self.M, self.C, self.r, self.L = M, C, r, L
        self.off, self.n = layout(M, C, r, L)
        self.p = np.asarray(params, dtype=np.float64).copy()
        assert self.p.size == self.n

    def w(self, name):
        o, sh = self.off[name]
        return self.p[o:o + int(np.prod(sh))].reshape(sh)

    def _mm(self, a, b, rnd):
        """Product of two operands as the GPU forms it (BF16 operands, wide sums)."""
        if rnd:
            a, b = bf16(a), bf16(b)
        return a @ b

    def _f32(self, x, rnd):
        # the GPU keeps activations in FP32 between kernels
        return x.astype(np.float32).astype(np.float64) if rnd else x

    def forward(self, X, rnd=True):
        """X: (M, B, 4) features -> Y (M, B, 2) and the cache of the backward pass."""
        M, C, L = self.M, self.C, self.L
        B = X.shape[1]
        X = np.asarray(X, dtype=np.float64)
        H = X @ self.w("P") + self.w("E")[:, None, :] + self.w("pb")
        H = self._f32(H, rnd)
        cache = {"X": X, "H": [H], "Hm": [], "S": [], "Z": [], "G": []}
        for l in range(L):
            Hf = H.reshape(M, B * C)
            S = self._f32(self._mm(self.w(f"V{l}").T, Hf, rnd), rnd)                 # (r, BC)
            Hm = self._f32(Hf + self._mm(self.w(f"U{l}"), S, rnd), rnd).reshape(M * B, C)
            Z = self._f32(self._mm(Hm, self.w(f"W1{l}"), rnd) + self.w(f"b1{l}"), rnd)
            G = self._f32(_gelu(Z), rnd)
            Hn = self._f32(Hm + self._mm(G, self.w(f"W2{l}"), rnd) + self.w(f"b2{l}"), rnd)
            cache["S"].append(S); cache["Hm"].append(Hm); cache["Z"].append(Z); cache["G"].append(G)
            H = Hn.reshape(M, B, C)
            cache["H"].append(H)
        Y = H.reshape(M * B, C) @ self.w("O") + self.w("ob")
        return Y.reshape(M, B, 2), cache

    def loss_grad(self, X, T, rnd=True):
        """Mean over cells of |Y - T|^2 / |T|^2 and its gradient (flat, float64)."""
        M, C, L = self.M, self.C, self.L
        Y, c = self.forward(X, rnd)
        B = Y.shape[1]
        T = np.asarray(T, dtype=np.float64)
        num = ((Y - T) ** 2).sum(axis=(0, 2))
        den = (T ** 2).sum(axis=(0, 2)) + 1e-30
        loss = float(np.mean(num / den))
        dY = 2.0 * (Y - T) / (den[None, :, None] * B)
        g = np.zeros(self.n)

        def put(name, val):
            o, sh = self.off[name]
            g[o:o + int(np.prod(sh))] = np.asarray(val).reshape(-1)

        HL = c["H"][L].reshape(M * B, C)
        dYf = dY.reshape(M * B, 2)
        put("O", HL.T @ dYf)
        put("ob", dYf.sum(axis=0))
        dH = self._f32(dYf @ self.w("O").T, rnd)                                  # (MB, C)
        for l in range(L - 1, -1, -1):
            G, Z, Hm, S = c["G"][l], c["Z"][l], c["Hm"][l], c["S"][l]
            dG = self._f32(self._mm(dH, self.w(f"W2{l}").T, rnd), rnd)
            put(f"W2{l}", self._mm(G.T, dH, rnd))
            put(f"b2{l}", dH.sum(axis=0))
            dZ = self._f32(dG * _gelu_grad(Z), rnd)
            put(f"W1{l}", self._mm(Hm.T, dZ, rnd))
            put(f"b1{l}", dZ.sum(axis=0))
            dHm = self._f32(dH + self._mm(dZ, self.w(f"W1{l}").T, rnd), rnd).reshape(M, B * C)
            put(f"U{l}", self._mm(dHm, S.T, rnd))
            dS = self._f32(self._mm(self.w(f"U{l}").T, dHm, rnd), rnd)
            Hin = c["H"][l].reshape(M, B * C)
            put(f"V{l}", self._mm(Hin, dS.T, rnd))
            dH = self._f32(dHm + self._mm(self.w(f"V{l}"), dS, rnd), rnd).reshape(M * B, C)
        dH3 = dH.reshape(M, B, C)
        X = c["X"]
        put("E", dH3.sum(axis=1))
        put("pb", dH3.sum(axis=(0, 1)))
        put("P", np.einsum("mbq,mbc->qc", X, dH3))
        return loss, g
