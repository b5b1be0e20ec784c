"""CPU oracle of the learned curvature closure: a 30 -> 64 -> 64 -> 1 MLP.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ to check the GPU forward and backward
kernels; never by the product path.

The network has no counterpart in the reference.  It stands in for get_curvature
(interface.f90:20-119) on under-resolved interface cells, and its definition is fixed
by the C ABI (include/icls.h):

  input f (32, the last two zero):
      f[0..26]  = phi(i+a, j+b, k+c) / dz in reshape order (a fastest, then b, then c,
                  each -1..1; the gather order of interface.f90:59)
      f[27..29] = the unit normal grad(phi)/|grad(phi)| from the least-squares rows 1..3
  h1 = tanh(W1 f + b1)        W1: 64 x 32
  h2 = tanh(W2 h1 + b2)       W2: 64 x 64
  y  = w3 . h2 + b3           w3: 64          (y = kappa * dz)
  loss = mean over the batch of (y - t)^2

Arithmetic: double throughout.  With bf16=True the operands the GPU feeds to its
BF16 matrix instructions (f, W1, W2, h1, and for the backward pass the outgoing
gradients) are rounded to bfloat16 first, as the GPU does; sums stay double.  The
rounding is round-to-nearest-even on the float32 bit pattern, the conversion the GPU
kernels use.
"""
import numpy as np

NIN, NPAD, NH = 30, 32, 64


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    x = np.asarray(a, dtype=np.float32).copy()
    u = x.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(x)
    if nan.any():
        out[nan] = np.nan
    return out


def to_bf16_bits(a):
    """bfloat16 bit patterns (uint16) of the nearest-even rounding of a."""
    x = np.ascontiguousarray(a, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b):
    """float64 values of bfloat16 bit patterns."""
    u = np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def flatten(p):
    """The flat parameter vector of the C ABI: W1, b1, W2, b2, w3, b3."""
    return np.concatenate([p["W1"].ravel(), p["b1"], p["W2"].ravel(), p["b2"], p["w3"], p["b3"]])


def unflatten(v):
    v = np.asarray(v, dtype=np.float64)
    o = 0
    p = {}
    for k, shape in (("W1", (NH, NPAD)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("w3", (NH,)), ("b3", (1,))):
        n = int(np.prod(shape))
        p[k] = v[o:o + n].reshape(shape).copy()
        o += n
    return p


def init_params(seed=1234):
    """Deterministic initial weights (uniform, fan-in scaled)."""
    rng = np.random.default_rng(seed)
    p = {
        "W1": rng.uniform(-1, 1, (NH, NPAD)) / np.sqrt(NIN),
        "b1": np.zeros(NH),
        "W2": rng.uniform(-1, 1, (NH, NH)) / np.sqrt(NH),
        "b2": np.zeros(NH),
        "w3": rng.uniform(-1, 1, NH) / np.sqrt(NH),
        "b3": np.zeros(1),
    }
    p["W1"][:, NIN:] = 0.0
    return p


def forward(p, X, bf=False):
    """X: (B, 32).  Returns y (B,) and the cache for backward."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    Xr = r(X)
    z1 = Xr @ r(p["W1"]).T + p["b1"]
    h1 = np.tanh(z1)
    h1r = r(h1)
    z2 = h1r @ r(p["W2"]).T + p["b2"]
    h2 = np.tanh(z2)
    y = h2 @ p["w3"] + p["b3"][0]
    return y, {"X": Xr, "h1": h1, "h1r": h1r, "h2": h2}


def loss_and_grads(p, X, t, bf=False):
    """Mean squared error and its gradient with respect to every parameter."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    y, c = forward(p, X, bf)
    B = X.shape[0]
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / B                                   # (B,)
    g = {"w3": c["h2"].T @ dy, "b3": np.array([dy.sum()])}
    dz2 = np.outer(dy, p["w3"]) * (1.0 - c["h2"] ** 2)  # (B, 64)
    dz2r = r(dz2)
    g["W2"] = dz2r.T @ c["h1r"]
    g["b2"] = dz2.sum(axis=0)
    dh1 = dz2r @ r(p["W2"])                            # (B, 64)
    dz1 = dh1 * (1.0 - c["h1"] ** 2)
    dz1r = r(dz1)
    g["W1"] = dz1r.T @ c["X"]
    g["b1"] = dz1.sum(axis=0)
    return loss, g


def features_from_stencil(phi27, normal, dz):
    """Feature rows from (B, 27) stencils and (B, 3) unit normals."""
    B = phi27.shape[0]
    X = np.zeros((B, NPAD))
    X[:, :27] = phi27 / dz
    X[:, 27:30] = normal
    return X
