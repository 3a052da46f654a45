"""The learned warm start on the GPU: the network's MFMA forward/backward
kernels against the CPU oracle (oracle/net_oracle.py), training on captured
APT step pairs, and FP32 defect-correction steps started from the prediction.

Tolerances: the kernels store activations and pre-activation gradients in
bfloat16 (8 significant bits) and accumulate in FP32; the oracle rounds the
same quantities to bfloat16 and accumulates in float64.  A bfloat16 rounding
that lands on the other side of a tie, or an FP32 sum that differs in its
last bit, changes a stored value by one bf16 ulp (2^-8 relative), so the
comparison is norm-wise: 2e-2 relative per parameter block, 1e-2 on the
prediction."""
import numpy as np
import pytest

import net_oracle as N
from cases import hot_disk_2d
from curvilineardiffusionmanuscript_amd import learned as LW
from curvilineardiffusionmanuscript_amd import solver as S

pytestmark = pytest.mark.gpu


def _blocks(g):
    offs = np.cumsum([0] + [n for _, n in N.SIZES])
    return {name: g[offs[k] : offs[k + 1]] for k, (name, _) in enumerate(N.SIZES)}


@pytest.mark.parametrize("H,W,kind", [(37, 70, "mse"), (8, 16, "mse"), (64, 64, "mse"), (37, 70, "residual"),
                                      (64, 64, "residual")])
def test_loss_gradient_matches_oracle(H, W, kind):
    p = N.init_params(11)
    X = N.random_features(H, W, 12)
    X[1 : H // 2 + 1, 1 : W // 3 + 1, 5] = 0.0  # an inactive region: prediction and gradient 0 there
    X[..., 3:5] = np.abs(X[..., 3:5])  # face diffusion numbers expm1(f) >= 0
    Xb = N.bf16_round(X)
    t = np.random.default_rng(13).normal(0, 0.5, (H, W)).astype(np.float32)
    net = LW.WarmStartNet()
    net.params = p
    net.set_loss(kind)
    loss, g, d = net.loss_grad(N.bf16_bits(X), t)
    lo, go, do = N.loss_grad(p.astype(np.float64), Xb.astype(np.float64), t.astype(np.float64), bf16=True,
                             kind=LW.WarmStartNet.LOSSES[kind])
    print(f"{H}x{W} {kind}: loss gpu {loss:.6e} oracle {lo:.6e}")
    assert abs(loss - lo) <= 1e-2 * lo
    assert np.linalg.norm(d - do) <= 1e-2 * np.linalg.norm(do)
    gb, ob = _blocks(g), _blocks(go)
    for name, _ in N.SIZES:
        err = np.linalg.norm(gb[name] - ob[name]) / max(np.linalg.norm(ob[name]), 1e-30)
        print(f"  {name}: rel err {err:.2e}")
        assert err <= 2e-2, name
    net.close()


@pytest.mark.parametrize("H,W", [(37, 70), (64, 64)])
def test_inference_prediction_matches_oracle_and_training_head(H, W):
    # the warm start's inference kernels (MFMA head) against the oracle's
    # prediction, and against the training path's prediction (same conv layers,
    # vector-ALU head): those two differ only in the FP32 summation order
    p = N.init_params(21)
    X = N.random_features(H, W, 22)
    X[H // 2 :, : W // 4 + 1, 5] = 0.0  # an inactive region: prediction 0 there
    Xb = N.bf16_round(X)
    net = LW.WarmStartNet()
    net.params = p
    pred = net.predict(N.bf16_bits(X), H, W)
    _, _, d_train = net.loss_grad(N.bf16_bits(X), np.zeros((H, W), np.float32))
    _, _, do = N.loss_grad(p.astype(np.float64), Xb.astype(np.float64), np.zeros((H, W)), bf16=True)
    err_o = np.linalg.norm(pred - do) / np.linalg.norm(do)
    err_t = np.max(np.abs(pred - d_train)) / np.max(np.abs(d_train))
    print(f"{H}x{W}: |pred - oracle| {err_o:.2e}, max|pred - training head| {err_t:.2e}")
    assert err_o <= 1e-2
    assert err_t <= 1e-5
    assert np.all(pred[d_train == 0.0] == 0.0)
    net.close()


def test_training_on_captured_pairs_reduces_loss():
    n, full = 48, 512
    net = LW.WarmStartNet(seed=1)
    net.dataset(n, n, 12)
    got = LW.generate_pairs(net, n=n, full=full, nseeds=3, steps=5, dt0=2e-5)
    assert got == 12
    _, l0 = net.evaluate(0)
    loss = net.train(epochs=40, lr=2e-3, seed=0)
    _, l1 = net.evaluate(0)
    print(f"loss before {l0:.3e}, last epoch {loss:.3e}, sample 0 after {l1:.3e}")
    # the loss is the squared relative residual of the predicted start: a
    # third of it is a predicted start with 58 % of the residual of T^n
    assert l1 < 0.5 * l0
    net.close()


@pytest.mark.parametrize("precision", ["fp32_dc", "fp64"])
def test_warm_started_step_reaches_the_same_tolerance(precision):
    # train briefly, then step a held-out problem with and without the warm
    # start (mixed-precision and FP64 steps)
    n, full = 64, 512
    net = LW.WarmStartNet(seed=2)
    net.dataset(n, n, 24)
    LW.generate_pairs(net, n=n, full=full, nseeds=4, steps=7, dt0=2e-5, seed0=100)
    net.train(epochs=60, lr=2e-3, seed=1)
    mesh, T0 = hot_disk_2d(n, seed=5)
    bcs = [S.DirichletBC(1.0), S.NeumannBC(), S.NeumannBC(), S.NeumannBC()]
    law = S.PowerLawConductivity(1.0, 2.5, 0.0)
    out = {}
    for warm in (False, True):
        # consulted at every step (warm_policy 1): the cost rule would keep a
        # network this small out of steps this cheap
        sol = S.PseudoTransientSolver(mesh, bcs, tuning={"warm_policy": 1} if warm else None)
        if warm:
            net.attach(sol)
        S.update_conductivity(sol, mesh, T0, 1.0, 1.0, law)
        sol.set_temperature(T0)
        dt, its, used, errs = 2e-5, [], [], []
        for _ in range(6):
            st, dt = sol.step_resident(dt, None, 1.0, law, rtol=0.0, atol=1e-8, max_iter=10**6, precision=precision)
            its.append(st.niter)
            used.append(st.warm_start)
            errs.append(st.abs_err)
        out[warm] = (its, used, errs, sol.get_temperature())
        sol.close()
    print("cold", out[False][:2], "warm", out[True][:2])
    assert not any(out[False][1])
    assert any(out[True][1])
    assert all(e <= 1e-8 for e in out[True][2])
    # both trajectories satisfy the same FP64 residual tolerance every step
    T_cold, T_warm = out[False][3], out[True][3]
    assert np.max(np.abs(T_warm - T_cold)) <= 6 * 2e-5 * np.sqrt(T0.size) * 1e-8 * 10
    net.close()


def test_network_sits_out_where_it_cannot_pay():
    # the mixed-precision hot disk: a decade of residual reduction costs the
    # step ~13-30 correction sweeps, so a prediction would have to take 2-5
    # decades off the start to repay the ~62 sweeps a consultation costs;
    # the network is never consulted and every step is the plain
    # mixed-precision step, bit for bit.  Consulted anyway (warm_policy 1),
    # the steps still end at the tolerance.
    net = LW.WarmStartNet.load()
    mesh, T0 = hot_disk_2d(512, seed=0)
    bcs = [S.DirichletBC(1.0), S.NeumannBC(), S.NeumannBC(), S.NeumannBC()]
    law = S.PowerLawConductivity(1.0, 2.5, 0.0)
    out = {}
    for name, warm, tuning in (("plain", False, None), ("auto", True, None), ("always", True, {"warm_policy": 1})):
        sol = S.PseudoTransientSolver(mesh, bcs, tuning=tuning)
        if warm:
            net.attach(sol)
        S.update_conductivity(sol, mesh, T0, 1.0, 1.0, law)
        sol.set_temperature(T0)
        dt, rec = 4e-6, []
        for _ in range(12):
            st, dt = sol.step_resident(dt, None, 1.0, law, rtol=0.0, atol=1e-8, max_iter=10**6, precision="fp32_dc")
            rec.append((st.niter, st.outer, int(st.warm_start), int(st.warm_consulted), st.abs_err))
            assert st.abs_err <= 1e-8
        out[name] = (rec, sol.get_temperature())
        sol.close()
    net.close()
    for name in ("auto", "always"):
        print(name, "consulted at steps", [k for k, r in enumerate(out[name][0]) if r[3]],
              "used at", [k for k, r in enumerate(out[name][0]) if r[2]])
    assert not any(r[3] for r in out["auto"][0])
    assert [r[:2] for r in out["auto"][0]] == [r[:2] for r in out["plain"][0]]
    np.testing.assert_array_equal(out["auto"][1], out["plain"][1])
    # from the second step on (the network reads the previous increment)
    assert [r[3] for r in out["always"][0]] == [0] + [1] * 11
