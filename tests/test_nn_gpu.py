"""Warm-start network kernels (mflbm_nn_*) against the CPU oracle (oracle/nn_oracle.py, pinned in
tests/test_nn_oracle.py): the matrix-core fragment layout, the convolution and its weight gradient, and the
whole encoder-decoder's loss and gradients; a short training run lowers the loss.

Tolerances: activations are BF16 on both sides with the rounding at the same points; the device accumulates in
a different order (matrix cores, FP32 atomics), so a value may land one BF16 step (2^-8 relative) away, and
the differences propagate through the layers: outputs to 3 % of their scale, gradients to 5 % in norm."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import nn_oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


def _t():
    import torch
    return torch


def _bf(a):
    torch = _t()
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda", torch.bfloat16)


def test_mfma_fragment_layout():
    import ctypes as C
    torch = _t()
    from mflbm import native
    lib = native.load()
    rng = np.random.default_rng(0)
    a = O.to_bf16(rng.standard_normal((16, 32)).astype(np.float32))
    b = O.to_bf16(rng.standard_normal((16, 32)).astype(np.float32))   # 16 columns of 32
    d = torch.zeros((16, 16), device="cuda")
    ga, gb = _bf(a), _bf(b)
    assert lib.mflbm_nn_mfma_probe(C.c_void_p(ga.data_ptr()), C.c_void_p(gb.data_ptr()), C.c_void_p(d.data_ptr()), None) == 0
    torch.cuda.synchronize()
    np.testing.assert_allclose(d.cpu().numpy(), a @ b.T, rtol=1e-5, atol=1e-5)


def _conv_case(cin, cout, shape, seed):
    rng = np.random.default_rng(seed)
    x = O.to_bf16(rng.standard_normal(shape + (cin,)).astype(np.float32))
    w = rng.standard_normal((cout, 27, cin)).astype(np.float32) * np.float32(np.sqrt(2.0 / (27 * cin)))
    b = rng.standard_normal(cout).astype(np.float32) * 0.1
    return x, w, b


@pytest.mark.parametrize("cin,cout,relu", [(8, 32, True), (32, 8, False), (16, 48, True)])
def test_conv3d_forward_and_weight_gradient_match_the_oracle(cin, cout, relu):
    import ctypes as C
    torch = _t()
    from mflbm import native
    from mflbm.warmstart import _kpad
    lib = native.load()
    shape = (2, 6, 5, 16)
    x, w, b = _conv_case(cin, cout, shape, cin + cout)
    gx, gw_t = _bf(x), torch.from_numpy(w.reshape(cout, -1)).cuda()
    wf = torch.zeros((cout, _kpad(cin)), dtype=torch.bfloat16, device="cuda")
    wb = torch.zeros((cin, _kpad(cout)), dtype=torch.bfloat16, device="cuda")
    P = lambda t: C.c_void_p(t.data_ptr())
    assert lib.mflbm_nn_weights_to_bf16(P(gw_t), P(wf), P(wb), cin, cout, None) == 0
    gb_t = torch.from_numpy(b).cuda()
    y = torch.empty(shape + (cout,), dtype=torch.bfloat16, device="cuda")
    assert lib.mflbm_nn_conv3d_fwd(P(gx), P(wf), P(gb_t), P(y), *shape, cin, cout, int(relu), None) == 0
    ref = O.to_bf16(O.conv3d(x, O.to_bf16(w), b, relu))
    got = y.float().cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 2 ** -7 * scale, np.abs(got - ref).max() / scale
    # weight and bias gradients for an output gradient g, and the input gradient through the flipped weights
    rng = np.random.default_rng(7)
    g = O.to_bf16(rng.standard_normal(shape + (cout,)).astype(np.float32))
    gg = _bf(g)
    dw = torch.zeros((cout, 27 * cin), device="cuda")
    db = torch.zeros(cout, device="cuda")
    assert lib.mflbm_nn_conv3d_dw(P(gx), P(gg), P(dw), P(db), *shape, cin, cout, None) == 0
    dx = torch.empty(shape + (cin,), dtype=torch.bfloat16, device="cuda")
    if cout % 8 == 0:
        assert lib.mflbm_nn_conv3d_fwd(P(gg), P(wb), None, P(dx), *shape, cout, cin, 0, None) == 0
    torch.cuda.synchronize()
    rgx, rgw, rgb = O.conv3d_grads(x, O.to_bf16(w), g)
    np.testing.assert_allclose(dw.cpu().numpy().reshape(cout, 27, cin), rgw, rtol=2e-3, atol=2e-3 * np.abs(rgw).max())
    np.testing.assert_allclose(db.cpu().numpy(), rgb, rtol=2e-3, atol=2e-3 * np.abs(rgb).max())
    if cout % 8 == 0:
        rgx = O.to_bf16(rgx)
        assert np.abs(dx.float().cpu().numpy() - rgx).max() <= 2 ** -7 * np.abs(rgx).max()


@pytest.mark.parametrize("cin,cout,shape", [(8, 32, (2, 6, 5, 16)), (32, 8, (2, 6, 5, 16)), (16, 48, (2, 6, 5, 16)),
                                            (16, 48, (2, 12, 16, 48))])
def test_ordered_weight_gradient_is_repeatable_and_matches_the_oracle(cin, cout, shape):
    """The fixed-order gradient sums: the oracle's gradients to the same tolerance as the atomic form, identical
    bits on a repeated call, and accumulation onto what dw / dbias already hold.  The last case has 18432
    positions: 5 weight chunks and 3 bias chunks, so the chunk sums are really summed."""
    import ctypes as C
    torch = _t()
    from mflbm import native
    lib = native.load()
    P = lambda t: C.c_void_p(t.data_ptr())
    x, w, _ = _conv_case(cin, cout, shape, 3 * cin + cout)
    g = O.to_bf16(np.random.default_rng(11).standard_normal(shape + (cout,)).astype(np.float32))
    gx, gg = _bf(x), _bf(g)
    need = lib.mflbm_nn_conv3d_dw_work(*shape, cin, cout)
    M = int(np.prod(shape))
    assert need == -(-M // 4096) * cout * 27 * cin + -(-M // 8192) * cout
    print(f"[nn dw ordered] shape {shape}: {-(-M // 4096)} weight chunks, {-(-M // 8192)} bias chunks", flush=True)
    work = torch.empty(need, device="cuda")
    runs = []
    for _ in range(2):
        dw = torch.zeros((cout, 27 * cin), device="cuda")
        db = torch.zeros(cout, device="cuda")
        assert lib.mflbm_nn_conv3d_dw_ordered(P(gx), P(gg), P(dw), P(db), P(work), need, *shape, cin, cout, None) == 0
        runs.append((dw.clone(), db.clone()))
    # a workspace one float short is refused
    assert lib.mflbm_nn_conv3d_dw_ordered(P(gx), P(gg), P(dw), P(db), P(work), need - 1, *shape, cin, cout, None) != 0
    # accumulation: a second call onto the first call's result doubles it exactly (x + x is exact in FP32)
    assert lib.mflbm_nn_conv3d_dw_ordered(P(gx), P(gg), P(dw), P(db), P(work), need, *shape, cin, cout, None) == 0
    torch.cuda.synchronize()
    assert torch.equal(runs[0][0], runs[1][0]) and torch.equal(runs[0][1], runs[1][1])
    assert torch.equal(dw, 2 * runs[0][0]) and torch.equal(db, 2 * runs[0][1])
    _, rgw, rgb = O.conv3d_grads(x, O.to_bf16(w), g)
    got_w = runs[0][0].cpu().numpy().reshape(cout, 27, cin)
    np.testing.assert_allclose(got_w, rgw, rtol=2e-3, atol=2e-3 * np.abs(rgw).max())
    np.testing.assert_allclose(runs[0][1].cpu().numpy(), rgb, rtol=2e-3, atol=2e-3 * np.abs(rgb).max())
    # the atomic form agrees with the ordered one to FP32 summation-order differences
    dwa = torch.zeros((cout, 27 * cin), device="cuda")
    dba = torch.zeros(cout, device="cuda")
    assert lib.mflbm_nn_conv3d_dw(P(gx), P(gg), P(dwa), P(dba), *shape, cin, cout, None) == 0
    torch.cuda.synchronize()
    scale = float(runs[0][0].abs().max())
    assert float((dwa - runs[0][0]).abs().max()) <= 1e-5 * scale * np.sqrt(M)
    assert float((dba - runs[0][1]).abs().max()) <= 1e-5 * float(runs[0][1].abs().max()) * np.sqrt(M)


def test_repeated_training_steps_give_identical_weights():
    """Three loss/gradient/Adam steps from the same start, twice: the weights agree bit for bit (the weight and
    bias gradients are summed in a fixed order).  16384 positions at the first level: 4 weight chunks and 2
    bias chunks, so the fixed-order sums of several chunks are what is repeated."""
    torch = _t()
    from mflbm.warmstart import USED, WarmStartNet
    rng = np.random.default_rng(21)
    x = np.zeros((2, 16, 16, 32, 8), np.float32)
    x[..., 0] = rng.random((2, 16, 16, 32)) > 0.35
    t = np.zeros((2, 16, 16, 32, 8), np.float32)
    t[..., :USED] = O.to_bf16(rng.standard_normal((2, 16, 16, 32, USED)).astype(np.float32))
    finals = []
    for _ in range(2):
        net = WarmStartNet((8, 16, 32), seed=3)
        for _ in range(3):
            net.loss_and_backward(_bf(x), _bf(t))
            net.adam(1e-3)
        torch.cuda.synchronize()
        finals.append(net.get_params())
    for name in finals[0]:
        for a, b in zip(finals[0][name], finals[1][name]):
            assert np.array_equal(a, b), name


def test_network_loss_and_gradients_match_the_oracle():
    torch = _t()
    from mflbm.warmstart import USED, WarmStartNet
    levels = (8, 16)
    net = WarmStartNet(levels, seed=5)
    params = net.get_params()
    rng = np.random.default_rng(9)
    for n in params:   # non-zero biases so that every ReLU sees both signs
        params[n] = (params[n][0], (rng.standard_normal(params[n][1].shape) * 0.05).astype(np.float32))
    net.set_params(params)
    x = np.zeros((2, 8, 8, 16, 8), np.float32)
    x[..., 0] = rng.random((2, 8, 8, 16)) > 0.35
    t = np.zeros((2, 8, 8, 16, 8), np.float32)
    t[..., :USED] = O.to_bf16(rng.standard_normal((2, 8, 8, 16, USED)).astype(np.float32))
    loss = net.loss_and_backward(_bf(x), _bf(t))
    orc = O.UNetOracle(params, list(levels), 8, 8)
    rloss, ry, rg = orc.loss_and_grads(x, t, USED)
    print(f"[nn gpu vs oracle] loss {loss:.6g} oracle {rloss:.6g}", flush=True)
    assert abs(loss - rloss) <= 1e-2 * rloss
    for name, L in net.layers.items():
        gw = L.gw.cpu().numpy().reshape(L.cout, 27, L.cin)
        gb = L.gb.cpu().numpy()
        rw, rb = rg[name]
        ew = np.linalg.norm(gw - rw) / (np.linalg.norm(rw) + 1e-30)
        eb = np.linalg.norm(gb - rb) / (np.linalg.norm(rb) + 1e-30)
        print(f"[nn gpu vs oracle] {name}: |dgw|/|gw| = {ew:.3e}  |dgb|/|gb| = {eb:.3e}", flush=True)
        assert ew < 5e-2 and eb < 5e-2, name


def test_training_lowers_the_loss():
    torch = _t()
    from mflbm.warmstart import Dataset, WarmStartNet, train
    rng = np.random.default_rng(4)
    n, s = 8, 16
    x = np.zeros((n, s, s, s, 8), np.float32)
    x[..., 0] = rng.random((n, s, s, s)) > 0.3
    t = np.zeros((n, s, s, s, 8), np.float32)
    # a target the network can represent: a smoothed copy of the indicator in channel 2
    k = x[..., 0]
    t[..., 2] = (k + np.roll(k, 1, 1) + np.roll(k, -1, 1) + np.roll(k, 1, 2) + np.roll(k, -1, 2)) / 5.0 * k
    data = Dataset(_bf(x), _bf(t), np.zeros(n, np.int64), 0.0)
    net = WarmStartNet((8, 16), seed=1)
    hist = train(net, data, epochs=30, batch_size=4, lr=3e-3, lr_final=3e-3, mirror=False)
    print(f"[nn training] loss {hist[0]:.4g} -> {hist[-1]:.4g}", flush=True)
    assert hist[-1] < 0.25 * hist[0]


def test_trained_network_predicts_the_steady_velocity_on_unseen_packs():
    """What training buys, and what it does not: on packs it has not seen, the network's velocity is closer
    to the steady field than the rest state (median relative L2 error below 0.8; the rest state's is 1), and runs
    restarted from its (u, v, w, rho) meet the reference's 1e-7 flow-rate criterion at the permeability of runs
    started from rest at the collision's equilibrium (the method without the network) within 0.5 %.  The network
    does not shorten the runs measurably (the criterion's last decades are set by the density field, which it
    predicts to tens of percent): the step counts of both starts are printed, not asserted."""
    _t()
    from mflbm import batch
    from mflbm.singlephase import SPCase, SPSimulation, permeability
    from mflbm.warmstart import T_PROBE, WarmStartNet, make_dataset, probe_fields, train, warm_state
    la_nu, force, side = 0.1, 1e-6, 32
    data = make_dataset(192, side=side, seed=3, la_nu=la_nu, force=force, per_batch=192)
    net = WarmStartNet((32, 64, 128), seed=2)
    hist = train(net, data, epochs=40, batch_size=4, lr=1e-3)
    print(f"[nn warm start] training loss {hist[0]:.4g} -> {hist[-1]:.4g}, target rms {data.target_rms}", flush=True)
    steps = {"cold": 0, "warm": 0}
    errs = []
    for k, s in enumerate(batch.rev_samples(3, seed=777, sides=(side,))):
        nx, ny, nz = s.shape
        ctl = {"lattice_dimensions": (nx, ny, nz), "periodic_indicator": (0, s.jper, 1),
               "domain_wall_status_y": (0, 0) if s.jper else (1, 1), "domain_wall_status_z": (0, 0),
               "MPI_async_layers_num": (0, 4 if s.jper else 0, 4), "fluid_viscosity": la_nu,
               "body_force_0": force, "MRT_collision_parameter_preset": 1, "steady_state_option": 1,
               "convergence_criteria": 1e-7}
        case = SPCase(ctl, s.walls, "fp64")
        res = {}
        for mode in ("cold", "warm"):
            sim = SPSimulation(case, case.collision_equilibrium_state())
            n0 = 1
            if mode == "warm":   # T_PROBE steps from rest, then restarted from the network's fields
                probe = probe_fields(sim, case, la_nu, force)
                sim.upload(warm_state(net, case, la_nu, force, probe))
                u0 = sim.macro()
                n0 = T_PROBE + 1
            n, status, hist_q = sim.run_steady(n0, 200000, 100)
            assert status == "converged", (mode, status)
            res[mode] = (n, permeability(case, hist_q[-1][1]), sim.macro())
            steps[mode] += n
            sim.close()
        ref = res["cold"][2]
        fl = np.pad(case.walls_global == 0, 1)
        num = sum(((u0[c] - ref[c])[fl] ** 2).sum() for c in ("u", "v", "w"))
        den = sum((ref[c][fl] ** 2).sum() for c in ("u", "v", "w"))
        err = float(np.sqrt(num / den))
        errs.append(err)
        kc, kw = res["cold"][1], res["warm"][1]
        print(f"[nn warm start] pack {k}: velocity error {err:.3f}, steps cold {res['cold'][0]} warm {res['warm'][0]}, "
              f"k {kc:.6e} vs {kw:.6e}", flush=True)
        assert abs(kw / kc - 1.0) < 5e-3
    print(f"[nn warm start] steps cold {steps['cold']} warm {steps['warm']}, velocity errors {np.round(errs, 3)}",
          flush=True)
    assert np.median(errs) < 0.8             # closer to the steady field than the rest state (error 1)


@pytest.mark.parametrize("periodic", [(True, False, False), (True, True, False), (False, False, False)])
def test_distance_features_on_the_device_equal_the_host_transform(periodic):
    """mflbm_nn_distance_features (separable capped squared-distance minimum) against the exact Euclidean
    distance transform of scipy on the host (features): equal after the BF16 rounding both go through."""
    torch = _t()
    from mflbm.warmstart import CIN, NFEAT, device_features, features
    rng = np.random.default_rng(21)
    fluid = rng.random((24, 40, 32)) > 0.08          # sparse solids: distances up to the cap
    fluid[:, :, :2] = False                          # and a wall slab
    x = torch.zeros((24, 40, 32, CIN), dtype=torch.bfloat16, device="cuda")
    device_features(x, fluid, periodic)
    ref = torch.from_numpy(features(fluid, periodic)).to(torch.bfloat16)
    got = x[..., :NFEAT].cpu()
    assert torch.equal(got, ref)
    assert float(x[..., NFEAT:].float().abs().max()) == 0.0
