"""MLP surrogate of the per-site GAP energy (host side).

``MLP`` wraps one network in libtgx (tgx_mlp_*): its dataset, weights and
optimiser state live on the device; training, evaluation and inference run on
the library's own kernels. The energy targets are the GAP local energies E_i and
the gradient targets the descriptor-space gradients dE_i/dq_i (= -Qss_i of
src/gap.f90:184), both taken from FP64 tgx_compute runs; the network is then
installed as an ensemble member of a block's surrogate (TGX_PATH_SURROGATE).

Two losses train a network. On environments taken one by one (MLP.train):
energy plus the descriptor-gradient term |P (g - g_ref)|^2, which matches
dE_i/dq_i in the directions a force can see (the tangent space of the
normalised descriptor). On whole structures (force_gradients, train_forces):
energy plus the force error itself, F_j = -sum_i (dE_i/dq_i) . (dq_i/dr_j)
summed over every block, formed from the FP64 plain path's dq/dr of the
structure.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L


class MLP:
    def __init__(self, ctx, dim: int, hidden: int = 256, max_batch: int = 4096, seed: int = 1):
        self.ctx = ctx
        self.dim, self.hidden, self.max_batch = dim, hidden, max_batch
        h = C.c_void_p()
        L.check(L.load().tgx_mlp_create(ctx.h, dim, hidden, max_batch, C.c_uint64(seed), C.byref(h)), "mlp_create")
        self.h = h
        ctx._children.append(self)  # the context destroys its networks before itself

    def close(self):
        if self.h is not None and self.h.value:
            L.load().tgx_mlp_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_samples(self, model_id: int, stride: int = 1, offset: int = 0) -> int:
        n = C.c_int()
        L.check(L.load().tgx_mlp_add_samples(self.h, model_id, stride, offset, C.byref(n)), "mlp_add_samples")
        return n.value

    def dataset_stats(self, n_use: int):
        mu, sd = C.c_double(), C.c_double()
        L.check(L.load().tgx_mlp_dataset_stats(self.h, n_use, C.byref(mu), C.byref(sd)), "mlp_dataset_stats")
        return mu.value, sd.value

    def set_normalization(self, shift: float, scale: float):
        L.check(L.load().tgx_mlp_set_normalization(self.h, float(shift), float(scale)), "mlp_set_normalization")

    def train(self, n_train: int, epochs: int, batch: int = 4096, lr: float = 1e-3, w_energy: float = 1.0,
              w_grad: float = 1.0, seed: int = 7):
        le, lg = C.c_double(), C.c_double()
        L.check(L.load().tgx_mlp_train(self.h, n_train, epochs, batch, C.c_float(lr), C.c_float(w_energy),
                                       C.c_float(w_grad), C.c_uint64(seed), C.byref(le), C.byref(lg)), "mlp_train")
        return le.value, lg.value

    def evaluate(self, begin: int, end: int):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        L.check(L.load().tgx_mlp_evaluate(self.h, begin, end, C.byref(a), C.byref(b), C.byref(c)), "mlp_evaluate")
        return {"e_mae": a.value, "e_rmse": b.value, "g_rmse": c.value}

    def forward(self, q: np.ndarray):
        q = np.ascontiguousarray(q, dtype=np.float32)
        B = q.shape[0]
        e = np.empty(B, np.float32)
        g = np.empty((B, self.dim), np.float32)
        L.check(L.load().tgx_mlp_forward_host(self.h, B, L.fptr(q), L.fptr(e), L.fptr(g)), "mlp_forward_host")
        return e, g

    def gradients(self, q, e_ref, g_ref, w_energy=1.0, w_grad=1.0):
        q = np.ascontiguousarray(q, dtype=np.float32)
        e_ref = np.ascontiguousarray(e_ref, dtype=np.float32)
        g_ref = np.ascontiguousarray(g_ref, dtype=np.float32)
        H, D = self.hidden, self.dim
        out = {"W1": np.empty((H, D), np.float32), "b1": np.empty(H, np.float32), "W2": np.empty((H, H), np.float32),
               "b2": np.empty(H, np.float32), "w3": np.empty(H, np.float32), "b3": np.empty(1, np.float32)}
        le, lg = C.c_double(), C.c_double()
        L.check(L.load().tgx_mlp_gradients_host(self.h, q.shape[0], L.fptr(q), L.fptr(e_ref), L.fptr(g_ref),
                                                C.c_float(w_energy), C.c_float(w_grad), L.fptr(out["W1"]),
                                                L.fptr(out["b1"]), L.fptr(out["W2"]), L.fptr(out["b2"]),
                                                L.fptr(out["w3"]), L.fptr(out["b3"]), C.byref(le), C.byref(lg)),
                "mlp_gradients_host")
        return out, le.value, lg.value

    def get_gradients(self):
        """The weight gradients of the last backward pass (layout of get_weights)."""
        H, D = self.hidden, self.dim
        g = {"W1": np.empty((H, D), np.float32), "b1": np.empty(H, np.float32), "W2": np.empty((H, H), np.float32),
             "b2": np.empty(H, np.float32), "w3": np.empty(H, np.float32), "b3": np.empty(1, np.float32)}
        L.check(L.load().tgx_mlp_get_gradients(self.h, *[L.fptr(g[k]) for k in ("W1", "b1", "W2", "b2", "w3", "b3")]),
                "mlp_get_gradients")
        return g

    def adam_step(self, lr: float = 1e-3):
        L.check(L.load().tgx_mlp_adam_step(self.h, C.c_float(lr)), "mlp_adam_step")

    def get_weights(self):
        H, D = self.hidden, self.dim
        w = {"W1": np.empty((H, D), np.float32), "b1": np.empty(H, np.float32), "W2": np.empty((H, H), np.float32),
             "b2": np.empty(H, np.float32), "w3": np.empty(H, np.float32), "b3": np.empty(1, np.float32)}
        L.check(L.load().tgx_mlp_get_weights(self.h, *[L.fptr(w[k]) for k in ("W1", "b1", "W2", "b2", "w3", "b3")]),
                "mlp_get_weights")
        return w

    def set_weights(self, w):
        arr = [np.ascontiguousarray(w[k], dtype=np.float32) for k in ("W1", "b1", "W2", "b2", "w3", "b3")]
        L.check(L.load().tgx_mlp_set_weights(self.h, *[L.fptr(a) for a in arr]), "mlp_set_weights")

    def install(self, model_id: int, member: int = 0):
        L.check(L.load().tgx_surrogate_install(self.ctx.h, model_id, member, self.h), "surrogate_install")


def set_gate(ctx, n_members: int, threshold: float):
    L.check(L.load().tgx_surrogate_gate(ctx.h, n_members, float(threshold)), "surrogate_gate")


def gate_stats(ctx):
    f, m = C.c_double(), C.c_double()
    L.check(L.load().tgx_surrogate_stats(ctx.h, C.byref(f), C.byref(m)), "surrogate_stats")
    return f.value, m.value


def md_dataset(ctx, model_ids, system, mlps, temperatures=(300.0, 1000.0, 2000.0, 3000.0), steps=50, dt=1.0,
               tau=100.0, seed=17, path="fp64_plain", log=None):
    """Environments for training: for each temperature, NVT Berendsen MD from the
    initial structure with seeded Maxwell velocities, one snapshot after `steps`
    steps (dt in fs). The snapshot's centres (q, E_i, dE_i/dq_i from the FP64
    path) of block k are appended to mlps[k][...] (every network of the block)."""
    from .md import MD, maxwell_velocities
    from . import structures as S
    mass = S.masses_for(system.species)
    for it, T in enumerate(temperatures):
        ctx.set_positions(system.positions)
        md = MD(ctx, mass, maxwell_velocities(mass, T, seed=seed + it), 3.7, path)
        for _ in range(steps):
            E, ek, Tinst = md.step(dt, T0=T, tau=tau)
        md.forces(path)  # the snapshot: FP64 results at the final positions
        for k, mid in enumerate(model_ids):
            for net in mlps[k]:
                net.add_samples(mid, 1, 0)
        if log:
            log(f"md dataset: T={T:.0f} K, {steps} steps, E={md.energy:.6f} eV, T_inst={Tinst:.1f} K")
    ctx.set_positions(system.positions)


def train_ensemble(mlps, n_samples, epochs=30, batch=4096, lr=1e-3, w_energy=1.0, w_grad=1.0, log=None):
    """Train every network on samples [0, n_samples[k]) of its block with the
    energy mean/std of those samples as normalisation; member i shuffles with
    seed 1000 k + i."""
    import time
    for k, nets in enumerate(mlps):
        for i, net in enumerate(nets):
            mu, sd = net.dataset_stats(n_samples[k])
            net.set_normalization(mu, sd if sd > 0 else 1.0)
            t0 = time.perf_counter()
            le, lg = net.train(n_samples[k], epochs, batch, lr, w_energy, w_grad, seed=1000 * k + i)
            if log:
                log(f"train block {k} member {i}: {n_samples[k]} samples, {epochs} epochs, "
                    f"loss_E={le:.4g} loss_g={lg:.4g} ({time.perf_counter() - t0:.1f} s)")


def force_gradients(ctx, nets, model_ids, w_energy=1.0, w_force=1.0):
    """Energy plus force loss of the context's current structure for one network
    per block (nets[b] for model_ids[b]) after ctx.compute("fp64_plain"): the
    force error of all blocks together. Leaves each network's weight gradients;
    returns (loss_energy, loss_force) with the force loss in (eV/A)^2 per atom."""
    n = len(nets)
    arr = (C.c_void_p * n)(*[net.h.value for net in nets])
    ids = np.ascontiguousarray(model_ids, dtype=np.int32)
    le, lf = C.c_double(), C.c_double()
    L.check(L.load().tgx_mlp_force_gradients(ctx.h, n, arr, L.iptr(ids), float(w_energy), float(w_force), C.byref(le),
                                             C.byref(lf)), "mlp_force_gradients")
    return le.value, lf.value


def md_structures(ctx, system, temperatures=(300.0, 1000.0, 2000.0, 3000.0), steps=200, every=50, dt=1.0,
                  tau=100.0, seed=23, rcut=3.7, path="fp64_plain"):
    """Whole-structure snapshots for the force loss: for each temperature, NVT
    Berendsen MD of `system` from its initial structure with seeded Maxwell
    velocities, a snapshot (positions) every `every` steps up to `steps`."""
    from .md import MD, maxwell_velocities
    from . import structures as S
    mass = S.masses_for(system.species)
    snaps = []
    for it, T in enumerate(temperatures):
        ctx.set_positions(system.positions)
        md = MD(ctx, mass, maxwell_velocities(mass, T, seed=seed + it), rcut, path)
        for k in range(steps):
            md.step(dt, T0=T, tau=tau)
            if (k + 1) % every == 0:
                snaps.append(md.state()[0].copy())
    ctx.set_positions(system.positions)
    return snaps


def train_forces(ctx, mlps, model_ids, structures, epochs=5, lr=1e-3, w_energy=1.0, w_force=1.0, rcut=3.7,
                 log=None):
    """Energy plus force loss over whole structures (positions of the context's
    system): each epoch visits every structure once, evaluates it with the FP64
    plain path and takes one Adam step per network, member i of every block
    trained together (mlps[b][i]). Returns the last epoch's mean losses."""
    import time
    t0 = time.perf_counter()
    n_mem = len(mlps[0])
    le_m = lf_m = 0.0
    for ep in range(epochs):
        le_s = lf_s = 0.0
        for pos in structures:
            ctx.set_positions(pos)
            ctx.build_neighbors_list(rcut, fetch=False)
            ctx.compute("fp64_plain")
            for i in range(n_mem):
                nets = [mlps[b][i] for b in range(len(mlps))]
                le, lf = force_gradients(ctx, nets, model_ids, w_energy, w_force)
                for net in nets:
                    net.adam_step(lr)
                le_s += le / n_mem
                lf_s += lf / n_mem
        le_m, lf_m = le_s / len(structures), lf_s / len(structures)
        if log:
            log(f"force training epoch {ep + 1}: {len(structures)} structures, loss_E={le_m:.4g} "
                f"loss_F={lf_m:.4g} (eV/A)^2 ({time.perf_counter() - t0:.1f} s)")
    return le_m, lf_m
