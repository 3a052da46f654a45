"""The learned warm start: a 4-level BF16 convolutional encoder-decoder, trained on the
GPU by the library's own kernels, whose prediction of the next time level
starts the FP32 defect-correction APT (include/cdiff.h, cdiff_net_*).

The APT still iterates to the same FP64 residual tolerance, so the network
only changes how far the iteration has to go; a step falls back to T^n
whenever the predicted start has the larger FP64 residual.  This part has no
counterpart in the reference.

Workflow::

    net = WarmStartNet(seed=0)
    pairs = generate_pairs(net, n=1024, nseeds=..., steps=...)   # APT step pairs in HBM
    net.train(pairs, epochs=..., lr=...)
    net.attach(solver)            # FP32_DC steps of `solver` now warm-start
    net.save(path) / WarmStartNet.load(path)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from . import solver as S
from .mesh import CurvilinearGrid2D, wavy_map_2d

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warm_start_4096.npz")


class NetError(RuntimeError):
    pass


class WarmStartNet:
    """A cdiff_net handle: the 4-level encoder-decoder of include/cdiff.h (32
    channels per level, 3x3 convolutions, 56,644 parameters)."""

    def __init__(self, seed: int = 0, gate: float = 1e-5):
        lib = L.load()
        h = C.c_void_p()
        rc = lib.cdiff_net_create(int(seed), C.byref(h))
        self._h = h
        self._check(rc)
        self.capacity = 0
        self.shape = None
        self.count = 0
        self.gate = gate

    @property
    def gate(self) -> float:
        return self._gate

    @gate.setter
    def gate(self, g: float):
        self._check(L.load().cdiff_net_set_gate(self._h, float(g)))
        self._gate = float(g)

    # -- lifetime ---------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.load().cdiff_net_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != L.OK:
            msg = L.load().cdiff_net_last_error(self._h).decode() if self._h and self._h.value else ""
            raise NetError(f"{L.load().cdiff_status_string(rc).decode()}: {msg}")

    # -- parameters -------------------------------------------------------------
    @staticmethod
    def num_params() -> int:
        return int(L.load().cdiff_net_num_params())

    @property
    def params(self) -> np.ndarray:
        n = self.num_params()
        out = np.zeros(n, dtype=np.float32)
        self._check(L.load().cdiff_net_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), n))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(L.load().cdiff_net_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size))

    def save(self, path: str, **meta):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.params, gate=np.float32(self.gate), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str = WEIGHTS) -> "WarmStartNet":
        with np.load(path, allow_pickle=False) as f:
            p = f["params"]
            gate = float(f["gate"]) if "gate" in f else 1e-5
        net = cls(gate=gate)
        net.params = p
        return net

    # -- training ---------------------------------------------------------------
    LOSSES = {"mse": 0, "residual": 1}

    def set_loss(self, kind: str):
        """Training loss: "mse" (mean squared error of the normalised
        increment) or "residual" (squared relative residual of the predicted
        start, to first order)."""
        self._check(L.load().cdiff_net_set_loss(self._h, self.LOSSES[kind]))

    def dataset(self, height: int, width: int, capacity: int):
        self._check(L.load().cdiff_net_dataset(self._h, int(height), int(width), int(capacity)))
        self.capacity, self.shape, self.count = capacity, (height, width), 0

    def capture(self, solver: S.PseudoTransientSolver, slot: int, stage: int, dt: float):
        self._check(L.load().cdiff_net_capture(self._h, solver._h, int(slot), int(stage), float(dt)))

    def train(self, epochs: int, lr: float = 1e-3, seed: int = 0, nsamples: int | None = None) -> float:
        n = self.count if nsamples is None else nsamples
        loss = C.c_double()
        self._check(L.load().cdiff_net_train(self._h, int(n), int(epochs), float(lr), int(seed), C.byref(loss)))
        return loss.value

    def evaluate(self, slot: int):
        h, w = self.shape
        delta = np.zeros(h * w, dtype=np.float32)
        loss = C.c_double()
        self._check(L.load().cdiff_net_eval(self._h, int(slot), delta.ctypes.data_as(C.POINTER(C.c_float)),
                                            C.byref(loss)))
        return delta.reshape(h, w), loss.value

    def loss_grad(self, features_bf16: np.ndarray, target: np.ndarray):
        """Loss, gradient and prediction for host features ((H+2, W+2, 8) bf16
        bit patterns) and target (H, W): the backward kernels' test hook."""
        H, W = target.shape
        f = np.ascontiguousarray(features_bf16, dtype=np.uint16)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.num_params(), dtype=np.float32)
        d = np.zeros(H * W, dtype=np.float32)
        loss = C.c_double()
        fp = C.POINTER(C.c_float)
        self._check(L.load().cdiff_net_loss_grad(self._h, H, W, f.ctypes.data_as(C.POINTER(C.c_uint16)),
                                                 t.ctypes.data_as(fp), g.ctypes.data_as(fp), d.ctypes.data_as(fp),
                                                 C.byref(loss)))
        return loss.value, g, d.reshape(H, W)

    def predict(self, features_bf16: np.ndarray, height: int, width: int) -> np.ndarray:
        """Prediction for host features ((H+2, W+2, 8) bf16 bit patterns) through
        the inference kernels the warm start runs."""
        f = np.ascontiguousarray(features_bf16, dtype=np.uint16)
        d = np.zeros(height * width, dtype=np.float32)
        self._check(L.load().cdiff_net_predict(self._h, int(height), int(width), f.ctypes.data_as(C.POINTER(C.c_uint16)),
                                               d.ctypes.data_as(C.POINTER(C.c_float))))
        return d.reshape(height, width)

    # -- use --------------------------------------------------------------------
    def attach(self, solver: S.PseudoTransientSolver):
        rc = L.load().cdiff_set_warm_start(solver._h, self._h)
        S._check(rc, solver._h)
        solver._warm_net = self  # keep the network alive while attached

    @staticmethod
    def detach(solver: S.PseudoTransientSolver):
        S._check(L.load().cdiff_set_warm_start(solver._h, None), solver._h)
        solver._warm_net = None


# ---------------------------------------------------------------------------
# synthetic problems: kappa = T^n, n ~ U[1, 3], 1-16 Gaussian blobs


def blob_temperature(mesh, rng, nblobs: int, background: float = 0.01, width=(0.02, 0.05)):
    """Background + Gaussian blobs (amplitude U[0.5, 1], width U[width] in the
    unit square, centres inside the mesh's extent), scaled so that the peak
    stays at or below 1: with kappa = T^n, n <= 3, the conductivity stays
    <= 1 and the FP64 residual of a 4096-cell-wide grid can reach 1e-8
    (overlapping blobs of height 2 put its rounding floor near 2e-8)."""
    x, y = mesh.centroid_coordinates.x, mesh.centroid_coordinates.y
    x0, x1, y0, y1 = x.min(), x.max(), y.min(), y.max()
    B = np.zeros(mesh.dims_full)
    for _ in range(nblobs):
        a = rng.uniform(0.5, 1.0)
        w = rng.uniform(*width)
        cx, cy = rng.uniform(x0, x1), rng.uniform(y0, y1)
        B = B + a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / w**2)
    peak = B.max()
    if peak > 1.0 - background:
        B *= (1.0 - background) / peak
    return np.asfortranarray(background + B)


def crop_mesh(n: int, full: int, offset=(0.0, 0.0), amplitude: float = 0.05):
    """An n x n window of the full x full wavy grid (same cell size)."""
    a0, b0 = offset
    ext = ((a0, a0 + n / full), (b0, b0 + n / full))
    return CurvilinearGrid2D.from_mapping(wavy_map_2d(amplitude), (n, n), 1, extent=ext)


def generate_pairs(net: WarmStartNet, *, n: int, full: int, nseeds: int, steps: int, dt0: float, seed0: int = 0,
                   atol: float = 1e-8, first_slot: int = 0, log=None) -> int:
    """Run the FP32 defect-correction APT on `nseeds` blob problems (n x n
    windows of the full x full grid, adaptive dt from dt0) and store
    (features, target) of every step after the first in the network's
    training store.  Returns the number of pairs stored."""
    slot = first_slot
    for s in range(nseeds):
        rng = np.random.default_rng(seed0 + s)
        off = tuple(rng.uniform(0.0, 1.0 - n / full, size=2))
        mesh = crop_mesh(n, full, off)
        T0 = blob_temperature(mesh, rng, int(rng.integers(1, 17)))
        law = S.PowerLawConductivity(1.0, float(rng.uniform(1.0, 3.0)), 0.0)
        sol = S.PseudoTransientSolver(mesh, [S.NeumannBC()] * 4)
        S.update_conductivity(sol, mesh, T0, 1.0, 1.0, law)
        sol.set_temperature(T0)
        dt = dt0
        kw = dict(rtol=0.0, atol=atol, max_iter=10**7, precision="fp32_dc")
        for k in range(steps):
            capture = k > 0 and slot < net.capacity
            if capture:
                net.capture(sol, slot, 0, dt)
            st, dt_next = sol.step_resident(dt, None, 1.0, law, **kw)
            if capture:
                net.capture(sol, slot, 1, dt)
                slot += 1
            dt = dt_next
        sol.close()
        if log:
            log(f"seed {seed0 + s}: n={law.t_exp:.2f}, {slot - first_slot} pairs so far")
        if slot >= net.capacity:
            break
    net.count = max(net.count, slot)
    return slot - first_slot


# ---------------------------------------------------------------------------
# command line: training on APT step pairs, and evaluation against the FP64 APT


def _log(*a):
    import sys

    print(*a, flush=True, file=sys.stderr)


def train_main(args):
    import time

    net = WarmStartNet(seed=args.seed, gate=args.gate)
    per_seed = args.steps - 1
    nseeds = (args.pairs + per_seed - 1) // per_seed
    net.dataset(args.crop, args.crop, args.pairs)
    t0 = time.time()
    got = generate_pairs(net, n=args.crop, full=args.full, nseeds=nseeds, steps=args.steps, dt0=args.dt0,
                         seed0=args.seed * 100000, log=_log if args.verbose else None)
    _log(f"{got} step pairs ({args.crop}^2 windows of the {args.full}^2 grid) in {time.time() - t0:.1f} s")
    # mean squared error first, then (optionally) the residual of the
    # predicted start at a lower rate
    plan = [("mse", args.lr)] * args.epochs
    if args.epochs > 1:
        plan[-1] = ("mse", args.lr / 10)
    plan += [("residual", args.finetune_lr)] * args.finetune_epochs
    for ep, (kind, lr) in enumerate(plan):
        t1 = time.time()
        net.set_loss(kind)
        loss = net.train(epochs=1, lr=lr, seed=args.seed + ep)
        _log(f"epoch {ep}: {kind} loss, lr {lr:.1e}: {loss:.4e}, {time.time() - t1:.1f} s")
    net.save(args.out, pairs=got, crop=args.crop, full=args.full, epochs=args.epochs, lr=args.lr,
             finetune_epochs=args.finetune_epochs, finetune_lr=args.finetune_lr, final_loss=loss,
             final_loss_kind=plan[-1][0])
    _log(f"weights -> {args.out}")


def evaluate(net, problem: str, n: int, steps: int, dt0: float, seed: int = 0, atol: float = 1e-8):
    """FP32_DC steps with and without the warm start, and the FP64 APT, on one
    problem: 'disk' (the hot-disk workload) or 'blobs' (kappa = T^n, n in
    [1, 3], 1-16 blobs).  Returns per-variant iterations, times and the final
    difference from the FP64 APT."""
    import time

    if problem == "disk":
        from .mesh import CurvilinearGrid2D as G  # noqa: F401

        mesh = crop_mesh(n, n, (0.0, 0.0))
        rng = np.random.default_rng(seed)
        cx, cy = rng.uniform(0.3, 0.7, size=2)
        x, y = mesh.centroid_coordinates.x, mesh.centroid_coordinates.y
        T0 = np.asfortranarray(np.where((x - cx) ** 2 + (y - cy) ** 2 <= 0.05**2, 1.0, 0.01))
        law = S.PowerLawConductivity(1.0, 2.5, 0.0)
        bcs = [S.DirichletBC(1.0), S.NeumannBC(), S.NeumannBC(), S.NeumannBC()]
    else:
        rng = np.random.default_rng(10**6 + seed)  # disjoint from the training seeds
        mesh = crop_mesh(n, n, (0.0, 0.0))
        T0 = blob_temperature(mesh, rng, int(rng.integers(1, 17)))
        law = S.PowerLawConductivity(1.0, float(rng.uniform(1.0, 3.0)), 0.0)
        bcs = [S.NeumannBC()] * 4
    out = {}
    # "+learned": consulted where a saving above the consultation's cost is
    # expected (the default); "+learned(always)": at every step from the second
    for name, prec, warm in (("fp64", "fp64", False), ("fp64+learned", "fp64", True), ("fp32_dc", "fp32_dc", False),
                             ("fp32_dc+learned", "fp32_dc", True), ("fp32_dc+learned(always)", "fp32_dc", True)):
        sol = S.PseudoTransientSolver(mesh, bcs, tuning={"warm_policy": 1} if name.endswith("(always)") else None)
        if warm:
            net.attach(sol)
        S.update_conductivity(sol, mesh, T0, 1.0, 1.0, law)
        sol.set_temperature(T0)
        dt, its, used, cons, tt = dt0, [], [], [], 0.0
        for _ in range(steps):
            t0 = time.perf_counter()
            st, dt = sol.step_resident(dt, None, 1.0, law, rtol=0.0, atol=atol, max_iter=10**7, precision=prec)
            tt += time.perf_counter() - t0
            its.append(st.niter)
            used.append(int(st.warm_start))
            cons.append(int(st.warm_consulted))
        out[name] = dict(iters=its, warm_used=used, warm_consulted=cons, s_per_step=tt / steps,
                         T=sol.get_temperature())
        sol.close()
    ref = out["fp64"]["T"]
    scale = float(np.max(np.abs(ref)))
    for name in out:
        out[name]["max_rel_diff_vs_fp64"] = float(np.max(np.abs(out[name].pop("T") - ref))) / scale
    out["problem"] = dict(kind=problem, n=n, steps=steps, dt0=dt0, t_exp=law.t_exp)
    return out


def evaluate_main(args):
    import json

    net = WarmStartNet.load(args.weights)
    for problem in args.problems.split(","):
        res = evaluate(net, problem, args.n, args.steps, args.dt0, seed=args.seed)
        print(json.dumps(res), flush=True)


def main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(prog="python -m curvilineardiffusionmanuscript_amd.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="generate APT step pairs on the GPU and train the network on them")
    t.add_argument("--pairs", type=int, default=2000)
    t.add_argument("--crop", type=int, default=1024)
    t.add_argument("--full", type=int, default=4096)
    t.add_argument("--steps", type=int, default=8, help="time steps per problem (pairs per problem = steps - 1)")
    t.add_argument("--dt0", type=float, default=1e-6)
    t.add_argument("--epochs", type=int, default=8)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--finetune-epochs", type=int, default=16,
                   help="epochs on the residual of the predicted start after the mean-squared-error epochs")
    t.add_argument("--finetune-lr", type=float, default=3e-5)
    t.add_argument("--gate", type=float, default=1e-5)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--verbose", action="store_true")
    e = sub.add_parser("evaluate", help="warm-started FP32_DC steps against FP32_DC and FP64 APT")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--problems", default="blobs,disk")
    e.add_argument("--n", type=int, default=4096)
    e.add_argument("--steps", type=int, default=10)
    e.add_argument("--dt0", type=float, default=1e-6)
    e.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    (train_main if args.cmd == "train" else evaluate_main)(args)


if __name__ == "__main__":
    main()
