"""The surrogate network and its hand-off to the solver, on the GPU.

- Whole-network gradients: the U-Net on the MFMA kernels against the same
  network in float64 on the CPU (torch conv2d, same weights and inputs),
  rounded to bf16 at the same places.  Without that rounding, ReLU masks
  flipping near zero make the deepest gradients differ by ~10 %; with it,
  the two agree to ~1e-2, while a wrong tap, transpose or missing term
  would decorrelate them.
- fd_init_fields: a running flow restarted from its own fields continues
  close to the original, far closer than a restart at rest.
- The device-array entry points give bit for bit what the host-array ones
  give.
- The accelerated loop with every jump rejected (interface error forced
  above the threshold) is the plain LBM, bit for bit; with none rejected it
  takes every jump.
- Training on LBM trajectories lowers the loss below persistence.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


class _Round(torch.autograd.Function):
    """Rounds to bf16 where the MFMA network stores bf16: the value on the way
    forward (fwd=True) and the gradient on the way back."""

    @staticmethod
    def forward(ctx, v, fwd):
        return v.to(torch.bfloat16).double() if fwd else v.clone()

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).double(), None


def _ref_forward(model, x):
    """The U-Net in float64 on the CPU with the weights of model (rounded to
    bf16), rounding to bf16 where the GPU network stores bf16: every
    convolution's output and input gradient, the pooled tensors and the
    output gradient entering each convolution.  The ReLU masks of the two
    networks then agree, and what is left is FP32 accumulation and the
    bf16 sums of the upsampling and skip gradients."""
    F = torch.nn.functional
    R = _Round.apply

    def c(name, v, relu=True):
        w = getattr(model, "w_" + name).detach().to(torch.bfloat16).double().cpu().permute(0, 3, 1, 2)
        w.requires_grad_(True)
        b = getattr(model, "b_" + name).detach().double().cpu().requires_grad_(True)
        ws[name] = (w, b)
        y = F.conv2d(R(v, False), w, b, padding=1)
        return R(F.relu(y), True) if relu else R(y, False)

    ws = {}
    pool = lambda v: R(F.avg_pool2d(v, 2), True)
    up = lambda v: v.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3)
    v = x.double().cpu().permute(0, 3, 1, 2)
    s0 = c("e0b", c("e0a", v))
    s1 = c("e1b", c("e1a", pool(s0)))
    s2 = c("e2b", c("e2a", pool(s1)))
    b = c("e3b", c("e3a", pool(s2)))
    d = c("d2b", c("d2a", torch.cat([up(b), s2], dim=1)))
    d = c("d1b", c("d1a", torch.cat([up(d), s1], dim=1)))
    d = c("d0b", c("d0a", torch.cat([up(d), s0], dim=1)))
    return c("head", d, relu=False)[:, :3].permute(0, 2, 3, 1), ws


def test_unet_gradients_match_float64_reference():
    from fingering_dynamics_amd.surrogate import UNet
    torch.manual_seed(0)
    model = UNet(channels=(16, 16, 32, 32), seed=3).cuda()
    x = (torch.randn(2, 16, 32, 8) * 0.5).to(torch.bfloat16).cuda()
    target = torch.randn(2, 16, 32, 3).cuda()
    out = model(x)
    loss = ((out - target) ** 2).mean()
    loss.backward()
    ref_out, ws = _ref_forward(model, x)
    ref_loss = ((ref_out - target.double().cpu()) ** 2).mean()
    ref_loss.backward()
    rel_out = float((out.detach().double().cpu() - ref_out.detach()).norm() / ref_out.detach().norm())
    print("output rel %.2e" % rel_out)
    assert rel_out < 3e-3, rel_out
    for name in model.names:
        g = getattr(model, "w_" + name).grad.double().cpu()
        r = ws[name][0].grad.permute(0, 2, 3, 1)
        rel = float((g - r).norm() / r.norm())
        cos = float((g * r).sum() / (g.norm() * r.norm()))
        gb = getattr(model, "b_" + name).grad.double().cpu()
        rb = ws[name][1].grad
        relb = float((gb - rb).norm() / rb.norm())
        print("%-5s dW rel %.2e cos %.6f  db rel %.2e" % (name, rel, cos, relb))
        assert rel < 2e-2 and cos > 0.9998 and relb < 2e-2, (name, rel, cos, relb)


def test_restart_from_fields_continues_the_flow():
    from fingering_dynamics_amd.lbm import Compute
    from fingering_dynamics_amd.ensemble import domain_spec, build_domain
    s = domain_spec(1)
    s.update(H=128, W=256)
    m, l, p = build_domain(s)
    a = Compute(m, l, p)
    a.step(3000)
    psi, rho, ux, uy = [v.astype(np.float64) for v in a.fields()]
    b = Compute(m, l, p)
    b.init_fields(psi, rho, ux, uy)
    c = Compute(m, l, p)
    c.init(psi, rho)
    for s_ in (a, b, c):
        s_.step(200)
    ua = a.fields()[2][m]
    eb = float(np.abs(b.fields()[2][m] - ua).max())
    ec = float(np.abs(c.fields()[2][m] - ua).max())
    pb = float(np.abs(b.fields()[0] - a.fields()[0]).max())
    print("restart |du| with velocity %.2e, at rest %.2e; |dpsi| %.2e" % (eb, ec, pb))
    assert eb < 0.25 * ec
    assert pb < 0.2
    for s_ in (a, b, c):
        s_.close()


def test_fields_right_after_a_restart_are_the_given_fields():
    """Before the first iteration after fd_init_fields(_device), the fields
    read back are the ones handed over (velocity included), and a plain
    fd_init afterwards reads back at rest again."""
    from fingering_dynamics_amd.lbm import Compute
    from fingering_dynamics_amd.ensemble import domain_spec, build_domain
    s = domain_spec(3)
    s.update(H=64, W=192)
    m, l, p = build_domain(s)
    a = Compute(m, l, p)
    a.step(500)
    psi, rho, ux, uy = [v.astype(np.float64) for v in a.fields()]
    b = Compute(m, l, p)
    b.init_fields(psi, rho, ux, uy)
    got = b.fields()
    for x, y in zip(got, (psi, rho, ux, uy)):
        assert np.array_equal(x[m], y[m].astype(np.float32))
    assert np.abs(ux[m]).max() > 0
    c = Compute(m, l, p)
    with torch.cuda.stream(c.torch_stream()):
        c.init_fields_device(*[torch.as_tensor(v, dtype=torch.float32).cuda() for v in (psi, rho, ux, uy)])
    assert np.array_equal(c.fields()[2][m], ux[m].astype(np.float32))
    b.init(psi, rho)
    assert not b.fields()[2].any()
    for s_ in (a, b, c):
        s_.close()


def test_device_coupling_matches_host_path():
    """fd_fields_device / fd_init_fields_device / device state blobs give
    bit-for-bit what the host-array entry points give."""
    from fingering_dynamics_amd.lbm import Compute
    from fingering_dynamics_amd.ensemble import domain_spec, build_domain
    s = domain_spec(2)
    s.update(H=96, W=200)
    m, l, p = build_domain(s)
    a = Compute(m, l, p)
    a.step(700)
    with torch.cuda.stream(a.torch_stream()):
        dev = a.fields_device()
        host_dev = dev.cpu().numpy()
    for i, h in enumerate(a.fields()):
        assert np.array_equal(host_dev[i], h), i
    psi, rho, ux, uy = [v.astype(np.float64) for v in a.fields()]
    psi = np.clip(psi + 0.05 * np.sin(np.arange(p.W) / 7.0)[None, :], -1, 1)
    b = Compute(m, l, p)
    b.init_fields(psi, rho, ux, uy)
    c = Compute(m, l, p)
    with torch.cuda.stream(c.torch_stream()):
        t = [torch.as_tensor(v, dtype=torch.float32).cuda() for v in (psi, rho, ux, uy)]
        c.init_fields_device(*t)
    b.step(300)
    c.step(300)
    for x, y in zip(b.fields(), c.fields()):
        assert np.array_equal(x, y)
    # device state blob: save, run on, restore, run again -> identical
    with torch.cuda.stream(a.torch_stream()):
        st = a.save_state_device()
    a.step(250)
    first = [v.copy() for v in a.fields()]
    assert a.t == 950
    a.load_state_device(st)
    assert a.t == 700
    a.step(250)
    for x, y in zip(first, a.fields()):
        assert np.array_equal(x, y)
    with pytest.raises(ValueError):
        a.init_fields_device(t[0].double(), *t[1:])
    # host memory handed to a device entry point is an argument error, not a fault
    host = np.zeros((p.H, p.W), dtype=np.float32)
    from fingering_dynamics_amd import _native
    assert a._lib.fd_fields_device(a._h, _native.ptr(host), None, None, None) == -1
    assert a._lib.fd_init_fields_device(a._h, *[_native.ptr(host)] * 4) == -1
    for s_ in (a, b, c):
        s_.close()


def test_training_beats_persistence():
    from fingering_dynamics_amd import surrogate as sg
    data = sg.generate(n_domains=10, H=128, W=256, f=2, snap=500, max_steps=12000, log=lambda *a: None)
    assert data["psi"].shape[0] >= 4
    idx = list(range(data["psi"].shape[0]))
    model = sg.UNet(seed=1).cuda()
    before = sg.evaluate(model, data, idx, 2000, 500)
    hist, _ = sg.train(model, data, idx, 2000, 500, iters=300, bs=16, lr=2e-3, log=lambda *a: None)
    after = sg.evaluate(model, data, idx, 2000, 500)
    print("mse untrained %.3e trained %.3e persistence %.3e" % (before["mse"], after["mse"], after["mse_persist"]))
    assert hist[-1][1] < hist[0][1]
    assert after["mse"] < after["mse_persist"]


def test_forced_fallback_is_the_plain_lbm():
    from fingering_dynamics_amd import surrogate as sg
    from fingering_dynamics_amd.lbm import Compute
    spec = sg.flow_spec(1, H=128, W=256)
    m, l, p = sg.build_flow(spec)
    model = sg.UNet(seed=0).cuda().eval()
    a = Compute(m, l, p)
    tb, _, rec_p = sg.plain_run(a, 8, max_steps=11200, chunk=200, record_every=2200, log=print)
    assert tb is None and a.t == 11200
    b = Compute(m, l, p)
    tb2, _, st, rec_a = sg.accelerated_run(b, model, spec["M"], 180.0 - spec["theta"], 8, jump=2000, relax=200,
                                           threshold=-1.0, max_steps=11200, log=print)
    print(st)
    assert tb2 is None and st["jumps"] == 0 and st["fallbacks"] == 5 and b.t == 11200
    assert sorted(rec_a) == [200, 2400, 4600, 6800, 9000]
    for t in rec_a:
        assert torch.equal(rec_a[t], rec_p[t]), t
    for x, y in zip(a.fields(), b.fields()):
        assert np.array_equal(x, y)
    c = Compute(m, l, p)
    _, _, st2, _ = sg.accelerated_run(c, model, spec["M"], 180.0 - spec["theta"], 8, jump=2000, relax=200,
                                      threshold=2.0, max_steps=11200, log=print)
    assert st2["fallbacks"] == 0 and st2["jumps"] == 5
    assert c.mass()[2] == 0
    for s_ in (a, b, c):
        s_.close()


def test_coupled_pairs_start_from_the_plain_state():
    """coupled_pairs: one pair per live flow and cycle; the first input is
    the plain LBM's coarse state after `relax` iterations, bit for bit, and
    the first target that after relax + jump."""
    from fingering_dynamics_amd import surrogate as sg
    from fingering_dynamics_amd.lbm import Compute
    specs = [sg.flow_spec(k, H=64, W=256, regime="target") for k in range(2)]
    flows = [sg.build_flow(s) for s in specs]
    model = sg.UNet(seed=2).cuda().eval()
    cp = sg.coupled_pairs(model, flows, specs, 8, jump=400, relax=100, cycles=3, band=2, log=print)
    assert cp["n"] == 6 and cp["psi0"].shape == (6, 8, 32)
    assert sorted(cp["flow"].tolist()) == [0, 0, 0, 1, 1, 1]
    for v in ("psi0", "psi1", "ux0", "ux1"):
        assert bool(torch.isfinite(cp[v]).all())
    m, l, p = flows[1]
    c = Compute(m, l, p)
    fl = torch.as_tensor(m, dtype=torch.float32).cuda()
    c.step(100)
    first = sg.coarsen(torch.as_tensor(c.fields()[0]).cuda(), fl, 8)
    c.step(400)
    teach = sg.coarsen(torch.as_tensor(c.fields()[0]).cuda(), fl, 8)
    k = cp["flow"].tolist().index(1)
    assert torch.equal(cp["psi0"][k], first)
    assert torch.equal(cp["psi1"][k], teach)
    assert not torch.equal(cp["psi0"][k], cp["psi1"][k])
    c.close()
