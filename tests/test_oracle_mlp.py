"""CPU pins of the MLP oracle (oracle/mlp_oracle.py): its input gradient and its
weight gradients of the energy + projected-gradient loss against central
finite differences in float64."""
import numpy as np

from oracle import mlp_oracle as M


def _weights(rng, D, H):
    return {"W1": rng.normal(size=(H, D)), "b1": rng.normal(size=H) * 0.1, "W2": rng.normal(size=(H, H)) / np.sqrt(H),
            "b2": rng.normal(size=H) * 0.1, "w3": rng.normal(size=H) / np.sqrt(H), "b3": np.array([0.05])}


def test_input_gradient_fd():
    rng = np.random.default_rng(0)
    D, H = 9, 7
    w = _weights(rng, D, H)
    q = rng.normal(size=(3, D))
    g = M.forward(w, q)["g"]
    h = 1e-6
    for k in range(D):
        e = np.zeros(D)
        e[k] = h
        fd = (M.forward(w, q + e)["E"] - M.forward(w, q - e)["E"]) / (2 * h)
        np.testing.assert_allclose(g[:, k], fd, rtol=1e-7, atol=1e-9)


def test_weight_gradients_fd():
    rng = np.random.default_rng(1)
    D, H, B = 8, 6, 5
    w = _weights(rng, D, H)
    q = rng.normal(size=(B, D))
    q /= np.linalg.norm(q, axis=1)[:, None]
    e_ref = rng.normal(size=B)
    g_ref = rng.normal(size=(B, D))
    _, _, g = M.loss_and_grads(w, q, e_ref, g_ref, 0.7, 1.3)

    def L(ww):
        le, lg, _ = M.loss_and_grads(ww, q, e_ref, g_ref, 0.7, 1.3)
        return 0.7 * le + 1.3 * lg

    h = 1e-6
    for name in ("W1", "b1", "W2", "b2", "w3", "b3"):
        flat = w[name].reshape(-1)
        for idx in range(0, flat.size, max(1, flat.size // 7)):
            wp = {k: v.copy() for k, v in w.items()}
            wm = {k: v.copy() for k, v in w.items()}
            wp[name].reshape(-1)[idx] += h
            wm[name].reshape(-1)[idx] -= h
            fd = (L(wp) - L(wm)) / (2 * h)
            assert abs(g[name].reshape(-1)[idx] - fd) < 1e-6 * max(1.0, abs(fd)), (name, idx)


def test_bf16_rounding():
    # bf16 keeps 7 mantissa bits: ulp(1) = 2^-7; ties go to the even mantissa
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9, -3.14159265, 1e-3], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2 * 2 ** -7 and r[3] == 1.0
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def _force_blocks(rng, D, H, n_atoms):
    """Two networks, each with a few centres whose list entries point at atoms
    of one structure (the first entry of a centre is its own atom); dq/dr of
    the entries random."""
    blocks = []
    atom = 0
    for k in range(2):
        n = 3
        sizes = rng.integers(2, 5, size=n)
        sb = np.concatenate([[0], np.cumsum(sizes)])
        nbr = []
        for i in range(n):
            nbr.append(atom)
            nbr += list(rng.integers(0, n_atoms, size=sizes[i] - 1))
            atom += 1
        q = rng.normal(size=(n, D))
        q /= np.linalg.norm(q, axis=1)[:, None]
        blocks.append({"w": _weights(rng, D, H), "q": q, "e_ref": rng.normal(size=n), "g_ref": rng.normal(size=(n, D)),
                       "J": rng.normal(size=(sb[-1], D, 3)), "site_beg": sb, "nbr": np.array(nbr), "shift": 0.3 * k,
                       "scale": 1.5 + k})
    return blocks


def test_force_loss_gradients_fd():
    """Weight gradients of the energy + force loss of whole structures (both
    networks, their force errors summed over the blocks) against central
    finite differences of the loss itself."""
    rng = np.random.default_rng(4)
    D, H, n_atoms = 7, 5, 9
    blocks = _force_blocks(rng, D, H, n_atoms)
    le, lf, grads = M.force_loss_and_grads(blocks, n_atoms, 0.6, 1.7)

    def L(bl):
        a, b, _ = M.force_loss_and_grads(bl, n_atoms, 0.6, 1.7)
        return 0.6 * a + 1.7 * b

    h = 1e-6
    for k in range(2):
        for name in ("W1", "b1", "W2", "b2", "w3", "b3"):
            flat = blocks[k]["w"][name].reshape(-1)
            for idx in range(0, flat.size, max(1, flat.size // 5)):
                bp = [dict(b, w={kk: v.copy() for kk, v in b["w"].items()}) for b in blocks]
                bm = [dict(b, w={kk: v.copy() for kk, v in b["w"].items()}) for b in blocks]
                bp[k]["w"][name].reshape(-1)[idx] += h
                bm[k]["w"][name].reshape(-1)[idx] -= h
                fd = (L(bp) - L(bm)) / (2 * h)
                assert abs(grads[k][name].reshape(-1)[idx] - fd) < 1e-6 * max(1.0, abs(fd)), (k, name, idx)


def test_force_loss_is_the_force_error():
    """The force loss is the mean squared error of the forces the networks
    imply: F_hat_j - F_j = -sum over entries p -> j of (g_hat - g)_i(p) . J_p,
    checked against forces assembled entry by entry from each network's own
    forces and the reference's."""
    rng = np.random.default_rng(5)
    D, H, n_atoms = 6, 4, 8
    blocks = _force_blocks(rng, D, H, n_atoms)
    _, lf, _ = M.force_loss_and_grads(blocks, n_atoms, 0.0, 1.0)
    F_hat = np.zeros((n_atoms, 3))
    F_ref = np.zeros((n_atoms, 3))
    for b in blocks:
        g_hat = b["scale"] * M.forward(b["w"], b["q"])["g"]
        for i in range(b["q"].shape[0]):
            for p in range(b["site_beg"][i], b["site_beg"][i + 1]):
                F_hat[b["nbr"][p]] += -(g_hat[i] @ b["J"][p])
                F_ref[b["nbr"][p]] += -(b["g_ref"][i] @ b["J"][p])
    assert abs(lf - np.sum((F_hat - F_ref) ** 2) / n_atoms) <= 1e-12 * max(1.0, lf)
