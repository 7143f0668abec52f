"""CPU oracle (numpy, float64) of the learned warm start's network and
features.  TEST INFRASTRUCTURE ONLY: the checker the HIP kernels are compared
against; no product code imports it.

The warm start has no counterpart in the reference (OpenADMIXTURE.jl starts
from a random point plus FRAPPE EM, src/algorithms_outer.jl:12-45), so this is
a restatement of the project's own definitions: the genotype-projection
features, the MLP 64 -> hidden -> hidden -> out with ReLU and a softmax head,
the soft-target cross entropy and Adam (Kingma & Ba 2015).  The gradient is
pinned by central finite differences of the loss in the tests.
"""
import numpy as np


def unpack(w, in_dim, hid, out):
    """Flat parameters [W1 (hid x in), b1, W2 (hid x hid), b2, W3 (out x hid), b3]."""
    w = np.asarray(w, np.float64)
    o = 0
    shapes = [(hid, in_dim), (hid,), (hid, hid), (hid,), (out, hid), (out,)]
    parts = []
    for s in shapes:
        n = int(np.prod(s))
        parts.append(w[o:o + n].reshape(s))
        o += n
    assert o == w.size
    return parts


def forward(w, X, in_dim, hid, out):
    W1, b1, W2, b2, W3, b3 = unpack(w, in_dim, hid, out)
    X = np.asarray(X, np.float64)
    A1 = X @ W1.T + b1
    H1 = np.maximum(A1, 0.0)
    A2 = H1 @ W2.T + b2
    H2 = np.maximum(A2, 0.0)
    Z = H2 @ W3.T + b3
    Zs = Z - Z.max(1, keepdims=True)
    E = np.exp(Zs)
    Y = E / E.sum(1, keepdims=True)
    return dict(X=X, H1=H1, H2=H2, Z=Z, Y=Y, logY=Zs - np.log(E.sum(1, keepdims=True)))


def loss(w, X, T, in_dim, hid, out):
    f = forward(w, X, in_dim, hid, out)
    T = np.asarray(T, np.float64)
    return float(-(T * f["logY"]).sum() / X.shape[0])


def loss_grad(w, X, T, in_dim, hid, out):
    """Mean soft-target cross entropy and its gradient (flat, float64)."""
    W1, b1, W2, b2, W3, b3 = unpack(w, in_dim, hid, out)
    f = forward(w, X, in_dim, hid, out)
    T = np.asarray(T, np.float64)
    B = X.shape[0]
    dZ = (f["Y"] - T) / B
    gW3 = dZ.T @ f["H2"]
    gb3 = dZ.sum(0)
    dH2 = (dZ @ W3) * (f["H2"] > 0)
    gW2 = dH2.T @ f["H1"]
    gb2 = dH2.sum(0)
    dH1 = (dH2 @ W2) * (f["H1"] > 0)
    gW1 = dH1.T @ f["X"]
    gb1 = dH1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return float(-(T * f["logY"]).sum() / B), g


def adam_step(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v


def features(g, R, mu, isd):
    """g: I x Jp int8 genotypes (-1 missing); features (1/sqrt(Jp)) sum_j z_ij R[j]."""
    g = np.asarray(g, np.float64)
    Jp = R.shape[0]
    Z = np.where(g[:, :Jp] < 0, 0.0, (g[:, :Jp] - mu[None, :]) * isd[None, :])
    return Z @ np.asarray(R, np.float64) / np.sqrt(Jp)
