"""FP64 numpy restatement of the surrogate MLP and its training gradients.

TEST INFRASTRUCTURE ONLY (tests/ use it to check libtgx's MLP kernels).

Network: z1 = W1 q + b1, h1 = silu(z1), z2 = W2 h1 + b2, h2 = silu(z2),
E = w3 . h2 + b3, g = dE/dq. Loss of a batch of B rows:
    L = w_E mean_b (E_b - e_b)^2 + w_G mean_b |u_b|^2,  u_b = P_b (g_b - g_ref_b),
    P_b = 1 - q_b q_b^T.
The gradient of the second term is the reverse-mode derivative of u . g(q)
with u fixed (forward mode along u, then backward) -- written here from the
chain rule independently of the HIP kernels. force_loss_and_grads replaces the
gradient term by the force error of whole structures, which enters the same
way with u_i = -scale sum_p J_p^T dF (the force error carried back to the
descriptor of centre i). `bf16` rounds the operands the
kernels feed to BF16 matrix products (round to nearest even).
"""
import numpy as np


def bf16(x):
    x = np.asarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    rounded = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sig(z)


def silu1(z):
    s = sig(z)
    return s * (1 + z * (1 - s))


def silu2(z):
    s = sig(z)
    return s * (1 - s) * (2 + z * (1 - 2 * s))


def forward(w, q, emulate_bf16=False):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, w3, b3 = (np.asarray(w[k], np.float64) for k in ("W1", "b1", "W2", "b2", "w3", "b3"))
    z1 = r(q) @ r(W1).T + b1
    h1 = silu(z1)
    z2 = r(h1) @ r(W2).T + b2
    h2 = silu(z2)
    E = h2 @ w3 + b3[0]
    a2 = w3 * silu1(z2)
    v1 = r(a2) @ r(W2)
    a1 = v1 * silu1(z1)
    g = r(a1) @ r(W1)
    return dict(z1=z1, h1=h1, z2=z2, h2=h2, E=E, a2=a2, v1=v1, a1=a1, g=g)


def loss_and_grads(w, q, e_ref, g_ref, w_e=1.0, w_g=1.0, emulate_bf16=False):
    q = np.asarray(q, np.float64)
    B = q.shape[0]
    f = forward(w, q, emulate_bf16)
    res = f["E"] - e_ref
    d = f["g"] - g_ref
    u = d - q * np.sum(q * d, axis=1, keepdims=True)
    loss_e = np.mean(res ** 2)
    loss_g = np.mean(np.sum(u * u, axis=1))
    return loss_e, loss_g, _param_grads(w, q, f, 2 * w_e * res / B, u, 2 * w_g / B, emulate_bf16)


def force_loss_and_grads(blocks, n_atoms, w_e=1.0, w_f=1.0, emulate_bf16=False):
    """Energy plus force loss of whole structures. blocks: one dict per network
    with its weights w, normalisation (shift, scale), and its centres' q (n, D),
    e_ref (n), g_ref = dE/dq of the reference (n, D), the list entries' dq/dr
    J (P, D, 3) with site_beg (n + 1) and the 0-based atom nbr (P) of each
    entry. The network gives E_hat = shift + scale E and g_hat = scale g; the
    force error dF_j = -sum over entries p -> j of (g_hat - g_ref)_i(p) . J_p
    sums over every block, and
        L = w_e sum_blocks mean_i r_i^2 + w_f / n_atoms sum_j |dF_j|^2,
    r_i = E_i - (e_ref_i - shift) / scale. Returns (loss_e, loss_f, [grads])
    with loss_e the sum over blocks of the mean r^2."""
    dF = np.zeros((n_atoms, 3))
    fw = []
    for b in blocks:
        f = forward(b["w"], b["q"], emulate_bf16)
        dg = b["scale"] * f["g"] - b["g_ref"]
        sb = b["site_beg"]
        for i in range(len(sb) - 1):
            for p in range(sb[i], sb[i + 1]):
                dF[b["nbr"][p]] -= b["J"][p].T @ dg[i]
        fw.append(f)
    loss_f = np.sum(dF * dF) / n_atoms
    out, le = [], 0.0
    for b, f in zip(blocks, fw):
        n = b["q"].shape[0]
        res = f["E"] - (b["e_ref"] - b["shift"]) / b["scale"]
        le += np.mean(res ** 2)
        G = np.zeros_like(b["q"])
        sb = b["site_beg"]
        for i in range(n):
            for p in range(sb[i], sb[i + 1]):
                G[i] += b["J"][p] @ dF[b["nbr"][p]]
        # dL/dg_i = d/dg_i (w_f / n_atoms |dF|^2) = cF u_i, cF = 2 w_f / n_atoms
        u = -b["scale"] * G
        out.append(_param_grads(b["w"], b["q"], f, 2 * w_e * res / n, u, 2 * w_f / n_atoms, emulate_bf16))
    return le, loss_f, out


def _param_grads(w, q, f, dE, u, cG, emulate_bf16=False):
    """Weight gradients of sum_b dE_b E_b + cG u_b . g_b with dE and u fixed
    (forward mode along u, then reverse)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    q = np.asarray(q, np.float64)
    W1, W2, w3 = (np.asarray(w[k], np.float64) for k in ("W1", "W2", "w3"))
    s1, s1pp = silu1(f["z1"]), silu2(f["z1"])
    s2, s2pp = silu1(f["z2"]), silu2(f["z2"])
    # forward mode along u
    t0 = r(u) @ r(W1).T
    t1 = s1 * t0
    t2 = r(t1) @ r(W2).T
    gam2 = s2 * t2
    zet2 = w3 * t2 * s2pp
    wz = r(zet2) @ r(W2)
    zet1 = f["v1"] * t0 * s1pp + wz * s1
    a1, a2, h1, h2 = f["a1"], f["a2"], f["h1"], f["h2"]
    # the weight gradients are products too: BF16 operands under emulation
    g = {
        "W1": r(a1).T @ r(dE[:, None] * q + cG * u) + r(cG * zet1).T @ r(q),
        "b1": (dE[:, None] * a1 + cG * zet1).sum(0),
        "W2": r(a2).T @ r(dE[:, None] * h1 + cG * t1) + r(cG * zet2).T @ r(h1),
        "b2": (dE[:, None] * a2 + cG * zet2).sum(0),
        "w3": (dE[:, None] * h2 + cG * gam2).sum(0),
        "b3": np.array([dE.sum()]),
    }
    return g
