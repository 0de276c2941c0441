"""FP64 NumPy restatement of the spectral neural operator surrogate
(test infrastructure only: the GPU kernels are checked against it).

The network and the flat parameter layout are those of the engine's
gnlse_sno_* entry points:

  H0 = x Wl + bl
  layer l: V_k = sum_n H[n] exp(-2 pi i k n / N), k < K
           W_k = V_k (Rr_k + i Ri_k)
           U[n] = (1/N) sum_k c_k Re(W_k exp(+2 pi i k n / N)), c_0 = 1, c_k = 2
           Z = H Wp + bp + U;  H = gelu(Z) (no activation after the last layer)
  Y1 = gelu(H Wq1 + bq1);  y = x[..., :2] + Y1 Wq2 + bq2
  loss = mean_b |y_b - t_b|^2 / |t_b|^2

The backward pass is written out by hand; it is pinned by central finite
differences (tests/test_sno_oracle.py).
"""

import numpy as np
from scipy.special import erf

SQRT1_2 = 1.0 / np.sqrt(2.0)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * SQRT1_2)) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def layout(N, C, K, L, CI, CH):
    """name -> (offset, shape) in the flat parameter vector."""
    out, o = {}, 0

    def add(name, shape):
        nonlocal o
        out[name] = (o, shape)
        o += int(np.prod(shape))

    add("Wl", (CI, C))
    add("bl", (C,))
    for l in range(L):
        add(f"Rr{l}", (K, C, C))
        add(f"Ri{l}", (K, C, C))
        add(f"Wp{l}", (C, C))
        add(f"bp{l}", (C,))
    add("Wq1", (C, CH))
    add("bq1", (CH,))
    add("Wq2", (CH, 2))
    add("bq2", (2,))
    out["_total"] = o
    return out


def unpack(flat, lay):
    return {k: flat[v[0]:v[0] + int(np.prod(v[1]))].reshape(v[1]) for k, v in lay.items() if k != "_total"}


def dft_mats(N, K):
    """Exact truncated DFT matrices: Ft (2K x N) rows cos, -sin; G (N x 2K)."""
    n = np.arange(N)
    k = np.arange(K)
    ang = 2.0 * np.pi * ((k[:, None] * n[None, :]) % N) / N
    Ft = np.empty((2 * K, N))
    Ft[0::2] = np.cos(ang)
    Ft[1::2] = -np.sin(ang)
    ck = np.where(k == 0, 1.0, 2.0)[:, None] / N
    G = np.empty((N, 2 * K))
    G[:, 0::2] = (ck * np.cos(ang)).T
    G[:, 1::2] = (-ck * np.sin(ang)).T
    return Ft, G


class SNO:
    def __init__(self, N, C=64, K=128, L=4, CI=3, CH=128):
        self.N, self.C, self.K, self.L, self.CI, self.CH = N, C, K, L, CI, CH
        self.lay = layout(N, C, K, L, CI, CH)
        self.Ft, self.G = dft_mats(N, K)

    @property
    def size(self):
        return self.lay["_total"]

    def forward(self, flat, x, keep=False):
        p = unpack(np.asarray(flat, dtype=np.float64), self.lay)
        x = np.asarray(x, dtype=np.float64)
        C, K = self.C, self.K
        H = x @ p["Wl"] + p["bl"]
        cache = {"x": x, "H": [H], "Z": [], "V": []}
        for l in range(self.L):
            V = np.einsum("kn,bnc->bkc", self.Ft, H, optimize=True)  # (B, 2K, C)
            Vc = V[:, 0::2, :] + 1j * V[:, 1::2, :]
            R = p[f"Rr{l}"] + 1j * p[f"Ri{l}"]
            Wc = np.einsum("bki,kio->bko", Vc, R, optimize=True)
            Ws = np.empty_like(V)
            Ws[:, 0::2, :] = Wc.real
            Ws[:, 1::2, :] = Wc.imag
            U = np.einsum("nk,bko->bno", self.G, Ws, optimize=True)
            Z = H @ p[f"Wp{l}"] + p[f"bp{l}"] + U
            H = Z if l == self.L - 1 else gelu(Z)
            cache["V"].append(Vc)
            cache["Z"].append(Z)
            cache["H"].append(H)
        Zq = H @ p["Wq1"] + p["bq1"]
        Y1 = gelu(Zq)
        y = x[..., :2] + Y1 @ p["Wq2"] + p["bq2"]
        cache.update(Zq=Zq, Y1=Y1)
        return (y, cache) if keep else y

    def loss_grad(self, flat, x, t):
        """(loss, per-field relative error, flat gradient)."""
        p = unpack(np.asarray(flat, dtype=np.float64), self.lay)
        y, c = self.forward(flat, x, keep=True)
        t = np.asarray(t, dtype=np.float64)
        B = y.shape[0]
        d = (y - t).reshape(B, -1)
        den = np.sum(t.reshape(B, -1) ** 2, axis=1)
        num = np.sum(d ** 2, axis=1)
        loss = float(np.mean(num / den))
        rel = np.sqrt(num / den)
        dy = (2.0 / B) * (y - t) / den[:, None, None]
        g = {}
        g["Wq2"] = np.einsum("bnh,bnj->hj", c["Y1"], dy, optimize=True)
        g["bq2"] = dy.sum(axis=(0, 1))
        dZq = (dy @ p["Wq2"].T) * gelu_grad(c["Zq"])
        g["Wq1"] = np.einsum("bnc,bnh->ch", c["H"][self.L], dZq, optimize=True)
        g["bq1"] = dZq.sum(axis=(0, 1))
        dH = dZq @ p["Wq1"].T
        for l in reversed(range(self.L)):
            Z = c["Z"][l]
            dZ = dH if l == self.L - 1 else dH * gelu_grad(Z)
            Hl = c["H"][l]
            g[f"bp{l}"] = dZ.sum(axis=(0, 1))
            g[f"Wp{l}"] = np.einsum("bnc,bno->co", Hl, dZ, optimize=True)
            dH = dZ @ p[f"Wp{l}"].T
            dWs = np.einsum("nk,bno->bko", self.G, dZ, optimize=True)  # (B, 2K, C)
            dWc = dWs[:, 0::2, :] + 1j * dWs[:, 1::2, :]  # d/dRe + i d/dIm
            Vc = c["V"][l]
            R = p[f"Rr{l}"] + 1j * p[f"Ri{l}"]
            # W = V R (complex, linear in (Re, Im) parts); gradients of real
            # parameters from the real-valued chain rule
            dR = np.einsum("bki,bko->kio", np.conj(Vc), dWc, optimize=True)
            g[f"Rr{l}"] = dR.real
            g[f"Ri{l}"] = dR.imag
            dVc = np.einsum("bko,kio->bki", dWc, np.conj(R), optimize=True)
            dV = np.empty((dZ.shape[0], 2 * self.K, self.C))
            dV[:, 0::2, :] = dVc.real
            dV[:, 1::2, :] = dVc.imag
            dH = dH + np.einsum("kn,bkc->bnc", self.Ft, dV, optimize=True)
        g["Wl"] = np.einsum("bni,bnc->ic", c["x"], dH, optimize=True)
        g["bl"] = dH.sum(axis=(0, 1))
        flat_g = np.zeros(self.size)
        for k, v in self.lay.items():
            if k != "_total":
                flat_g[v[0]:v[0] + int(np.prod(v[1]))] = g[k].ravel()
        return loss, rel, flat_g


def init_params(net, seed=0):
    """Uniform initialisation scaled by fan-in (spectral weights by
    1/(C*C), as in the published Fourier neural operator)."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(net.size, dtype=np.float64)
    for k, v in net.lay.items():
        if k == "_total":
            continue
        o, s = v
        n = int(np.prod(s))
        if k.startswith("R"):
            v = rng.uniform(0.0, 1.0, n) / (net.C * net.C)
        elif k.startswith("b"):
            v = np.zeros(n)
        else:
            v = rng.uniform(-1.0, 1.0, n) / np.sqrt(s[0])
        flat[o:o + n] = v
    return flat
