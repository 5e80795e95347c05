"""FP64 numpy restatement of the coupling-predictor network (TEST
INFRASTRUCTURE ONLY: imported by tests/, never by the product package).

Same network and parameter layout as fribois2d_amd/csrc/fribo_mlp.hip:
ni -> 64 tanh -> 64 tanh -> no linear, parameters [W1 (64 x ni), b1, W2
(64 x 64), b2, W3 (no x 64), b3]; loss = mean over samples and outputs of
(Y - T)^2; Adam (Kingma and Ba, ICLR 2015) with beta1 = 0.9, beta2 = 0.999,
eps = 1e-8 and bias-corrected moments.
"""
import numpy as np

H = 64


def layout(ni, no):
    w1 = 0
    b1 = w1 + H * ni
    w2 = b1 + H
    b2 = w2 + H * H
    w3 = b2 + H
    b3 = w3 + no * H
    return dict(w1=w1, b1=b1, w2=w2, b2=b2, w3=w3, b3=b3, np=b3 + no)


def unpack(p, ni, no):
    L = layout(ni, no)
    p = np.asarray(p, dtype=np.float64)
    W1 = p[L["w1"]:L["b1"]].reshape(H, ni)
    b1 = p[L["b1"]:L["w2"]]
    W2 = p[L["w2"]:L["b2"]].reshape(H, H)
    b2 = p[L["b2"]:L["w3"]]
    W3 = p[L["w3"]:L["b3"]].reshape(no, H)
    b3 = p[L["b3"]:L["np"]]
    return W1, b1, W2, b2, W3, b3


def forward(p, X, no):
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, X.shape[1], no)
    H1 = np.tanh(X @ W1.T + b1)
    H2 = np.tanh(H1 @ W2.T + b2)
    return H1, H2, H2 @ W3.T + b3


def loss(p, X, T):
    T = np.asarray(T, dtype=np.float64)
    Y = forward(p, X, T.shape[1])[2]
    return float(np.mean((Y - T) ** 2))


def grad(p, X, T):
    """Analytic gradient of loss() (back-propagation)."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    no = T.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(p, X.shape[1], no)
    H1, H2, Y = forward(p, X, no)
    D3 = 2.0 * (Y - T) / Y.size
    D2 = (D3 @ W3) * (1.0 - H2 ** 2)
    D1 = (D2 @ W2) * (1.0 - H1 ** 2)
    return np.concatenate([(D1.T @ X).ravel(), D1.sum(0), (D2.T @ H1).ravel(), D2.sum(0),
                           (D3.T @ H2).ravel(), D3.sum(0)])


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step t (1-based); returns new (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16(x):
    """Round to bfloat16 (nearest even), as float64."""
    a = np.asarray(x, dtype=np.float32).copy()
    u = a.view(np.uint32)
    u += np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))
    u &= np.uint32(0xFFFF0000)
    return a.astype(np.float64)


def forward_bf16(p, X, no):
    """Inference with inputs, parameters and hidden activations rounded to
    bfloat16 and exact accumulation."""
    X = bf16(X)
    W1, b1, W2, b2, W3, b3 = (bf16(a) for a in unpack(p, np.shape(X)[1], no))
    H1 = bf16(np.tanh(X @ W1.T + b1))
    H2 = bf16(np.tanh(H1 @ W2.T + b2))
    return H2 @ W3.T + b3
