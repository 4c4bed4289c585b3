"""FP64 numpy restatement of the learned coarse-space correction.

TEST INFRASTRUCTURE ONLY: the checker of ``within_amd/csrc/coarse.hip`` and
the coarse-correction host code in ``within_amd/csrc/runtime.cpp``;
``within_amd`` never imports it.  The reference has no learned component (it
would extend the CG loop of ``schwarz-precond/src/solve/cg.rs:45-133`` beside
``FePreconditioner``, ``crates/within/src/operator/preconditioner.rs:47-154``),
so this module is pinned by mathematics, not by reference vectors:

* the loss gradient against central finite differences (``mimic=False``: every
  operation in FP64);
* the coarse-space loss against its meaning (-1 when span(P) holds every
  sample, 0 when P is G-orthogonal to them);
* the two-level apply against dense linear algebra (it is the identity plus
  projections: z = G^+ r when M = G^+, and Q r = P (P'GP)^+ P' r).

With ``mimic=True`` the network rounds its matrix-core operands to bfloat16
(nearest even) where the device does, so device and oracle agree to FP32
accumulation error.

Network (internal DOF numbering, m rows, width 32):
    Z_l = H_l Ws_l + (A H_l) Wa_l + b_l,  A = (1/Q) D^-1/2 G D^-1/2,
    H_{l+1} = ReLU(Z_l) for l < L-1,  P = Z_{L-1}.
Loss on error samples E with residuals R = G E, nu = sum_s e_s' r_s:
    L = -<U, (P'GP + eps tr(P'GP)/32 I)^-1 U> / nu,  U = P'R  (eps: the ridge, 1e-4 by default).
"""

from __future__ import annotations

import numpy as np

K = 32
LAYERS = 3
MAT = K * K
LAYER_PARAMS = 2 * MAT + K
N_PARAMS = LAYERS * LAYER_PARAMS


def bf16(x):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _splitmix(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def randn(seed, idx):
    """N(0, 1) from counters idx: Box-Muller on two 53-bit uniforms of splitmix64."""
    idx = np.asarray(idx, dtype=np.uint64)
    s = np.uint64(seed)
    with np.errstate(over="ignore"):
        h1 = _splitmix(s ^ _splitmix(np.uint64(2) * idx))
        h2 = _splitmix(s ^ _splitmix(np.uint64(2) * idx + np.uint64(1)))
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 0.5) / 9007199254740992.0
    u2 = ((h2 >> np.uint64(11)).astype(np.float64) + 0.5) / 9007199254740992.0
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def features(diag, factor, user_dof, seed):
    """Input features of each internal DOF: [1, z, one-hot(factor mod 4), z^2 - 1,
    25 N(0,1) keyed by the user DOF], z = standardised log of the Gramian diagonal."""
    z = np.log(np.maximum(np.asarray(diag, dtype=np.float64), 1e-300))
    z = (z - z.mean()) / max(z.std(), 1e-12)
    m = len(z)
    F = np.zeros((m, K))
    F[:, 0] = 1.0
    F[:, 1] = z
    F[np.arange(m), 2 + (np.asarray(factor) & 3)] = 1.0
    F[:, 6] = z * z - 1.0
    fseed = int(_splitmix(np.uint64(seed + 1)))
    cols = np.arange(7, K, dtype=np.uint64)
    idx = np.asarray(user_dof, dtype=np.uint64)[:, None] * np.uint64(K) + cols[None, :]
    F[:, 7:] = randn(fseed, idx)
    return F.astype(np.float32).astype(np.float64)


def _layer_params(theta, l):
    t = np.asarray(theta, dtype=np.float64)[l * LAYER_PARAMS:(l + 1) * LAYER_PARAMS]
    return t[:MAT].reshape(K, K), t[MAT:2 * MAT].reshape(K, K), t[2 * MAT:]


def forward(theta, X0, G, dinv, Q, mimic=True):
    """Returns P and the activations the gradient needs."""
    r = bf16 if mimic else (lambda a: np.asarray(a, dtype=np.float64))
    c = dinv / Q
    H = [np.asarray(X0, dtype=np.float64)]
    T = []
    Z = None
    for l in range(LAYERS):
        Ws, Wa, b = _layer_params(theta, l)
        S = dinv[:, None] * H[l]
        T.append(G @ S)
        Z = r(H[l]) @ r(Ws) + r(c[:, None] * T[l]) @ r(Wa) + b
        if l < LAYERS - 1:
            H.append(np.maximum(Z, 0.0))
    return Z, {"H": H, "T": T, "c": c}


def coarse_solve(Ac, U, nu, eps):
    """V = (Ac + eps tr(Ac)/K I)^-1 U by the same Cholesky as the device; loss."""
    A = np.array(Ac, dtype=np.float64)
    reg = eps * max(np.trace(A), 0.0) / K + 1e-300
    A[np.diag_indices(K)] += reg
    L = np.zeros_like(A)
    for j in range(K):
        d = A[j, j] - L[j, :j] @ L[j, :j]
        d = np.sqrt(d) if d > reg else np.sqrt(reg)
        L[j, j] = d
        L[j + 1:, j] = (A[j + 1:, j] - L[j + 1:, :j] @ L[j, :j]) / d
    y = np.linalg.solve(L, U)
    V = np.linalg.solve(L.T, y)
    return V, -float(np.sum(U * V)) / nu


def coarse_loss(P, GP, R, nu, eps=1e-4):
    """The loss at the coarse space P (GP = G P) and its gradient dL/dP."""
    V, loss = coarse_solve(P.T @ GP, P.T @ R, nu, eps)
    return loss, -(2.0 / nu) * (R - GP @ V) @ V.T


def backward(theta, cache, dz, G, dinv, mimic=True, scale=None):
    """Gradient with respect to theta from dL/dP = dz through the layers of
    ``forward`` (its cache).  With a list ``scale``, also appends each entry's
    sum of absolute terms (the bound that floating-point summation errors of
    the same sums scale with)."""
    r = bf16 if mimic else (lambda a: np.asarray(a, dtype=np.float64))
    g = np.zeros(N_PARAMS)
    H, T, c = cache["H"], cache["T"], cache["c"]
    for l in range(LAYERS - 1, -1, -1):
        Ws, Wa, _ = _layer_params(theta, l)
        o = l * LAYER_PARAMS
        g[o:o + MAT] = (r(H[l]).T @ r(dz)).ravel()
        g[o + MAT:o + 2 * MAT] = (r(c[:, None] * T[l]).T @ r(dz)).ravel()
        g[o + 2 * MAT:o + LAYER_PARAMS] = dz.sum(axis=0)
        if scale is not None:
            if not scale:
                scale.append(np.zeros(N_PARAMS))
            a = np.abs(r(dz))
            scale[0][o:o + MAT] = (np.abs(r(H[l])).T @ a).ravel()
            scale[0][o + MAT:o + 2 * MAT] = (np.abs(r(c[:, None] * T[l])).T @ a).ravel()
            scale[0][o + 2 * MAT:o + LAYER_PARAMS] = np.abs(dz).sum(axis=0)
        if l == 0:
            break
        ys = dinv[:, None] * (r(dz) @ r(Wa.T))
        dh = r(dz) @ r(Ws.T) + c[:, None] * (G @ ys)
        dz = dh * (H[l] > 0.0)
    return g


def loss_grad(theta, X0, G, dinv, Q, E, R, nu, eps=1e-4, mimic=True, grad=True, scale=None):
    """The training loss and its gradient with respect to theta."""
    P, cache = forward(theta, X0, G, dinv, Q, mimic)
    loss, dP = coarse_loss(P, G @ P, R, nu, eps)
    if not grad:
        return loss, None
    return loss, backward(theta, cache, dP, G, dinv, mimic, scale)


def sym_pinv(A, rtol=1e-10):
    """Pseudo-inverse of a symmetric matrix (eigenvalues below rtol of the largest dropped)."""
    A = 0.5 * (A + A.T)
    w, V = np.linalg.eigh(A)
    keep = w > rtol * np.max(np.abs(w))
    return (V[:, keep] / w[keep]) @ V[:, keep].T


def two_level_apply(P, GP, Ainv, M, r):
    """Balancing two-level preconditioner z = Q r + (I - Q G) M (I - G Q) r with
    Q = P Ainv P' and GP standing for G P; M applies the one-level preconditioner."""
    c1 = Ainv @ (P.T @ r)
    w = M(r - GP @ c1)
    return w + P @ (c1 - Ainv @ (GP.T @ w))
