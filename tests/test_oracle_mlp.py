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
