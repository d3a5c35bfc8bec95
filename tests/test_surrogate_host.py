"""Host-side pieces of the surrogate (CPU tensors): coarsening and its
inverse, the interface error, the flow generators, the network's parameter
and FLOP counts, the weight file round trip, the network input layout, the
jump hand-off arithmetic, the breakthrough probe, the training pair index,
and the many-domain ensemble's generators."""
import math

import numpy as np
import torch

from fingering_dynamics_amd import surrogate as sg


def _fluid(H, W, seed):
    rng = np.random.default_rng(seed)
    m = rng.random((H, W)) > 0.3
    m[:8, :8] = False            # one all-solid 8 x 8 block
    return torch.as_tensor(m, dtype=torch.float32)


def test_coarsen_averages_fluid_nodes_and_conserves_the_sum():
    H, W, f = 32, 48, 8
    fl = _fluid(H, W, 1)
    field = torch.as_tensor(np.random.default_rng(2).standard_normal((H, W)), dtype=torch.float32)
    c = sg.coarsen(field, fl, f)
    assert c.shape == (H // f, W // f)
    assert float(c[0, 0]) == 0.0
    # brute force on one block
    blk = (slice(8, 16), slice(16, 24))
    want = float((field[blk] * fl[blk]).sum() / fl[blk].sum())
    assert abs(float(c[1, 2]) - want) < 1e-6
    n = fl.view(H // f, f, W // f, f).sum(dim=(1, 3))
    assert abs(float((c * n).sum()) - float((field * fl).sum())) < 1e-3
    # batched leading dimensions give the same per-field result
    cb = sg.coarsen(torch.stack([field, 2 * field]), fl, f)
    assert torch.allclose(cb[0], c) and torch.allclose(cb[1], 2 * c)
    sf = sg.solid_fraction(fl, f)
    assert float(sf[0, 0]) == 1.0 and torch.allclose(sf, 1.0 - n / 64)


def test_refine_inverts_coarsen_on_block_constant_fields():
    H, W, f = 16, 32, 8
    fl = _fluid(H, W, 3)
    coarse = torch.arange(8, dtype=torch.float32).view(2, 4) - 3.5
    fine = sg.refine(coarse, f)
    assert fine.shape == (H, W)
    c = sg.coarsen(fine, fl, f)
    n = fl.view(2, f, 4, f).sum(dim=(1, 3))
    assert torch.equal(c[n > 0], coarse[n > 0])


def test_interface_error_counts_fluid_cells_of_other_sign():
    a = torch.tensor([[1.0, -1.0], [0.5, -0.2]])
    b = torch.tensor([[0.9, 0.3], [-0.5, -0.9]])
    fc = torch.tensor([[1.0, 1.0], [1.0, 0.0]])
    assert abs(float(sg.interface_error(a, b, fc)) - 2.0 / 3.0) < 1e-7
    assert float(sg.interface_error(a, a, fc)) == 0.0


def test_near_interface_is_the_dilated_interface():
    """Brute force: a cell is near when some interface cell (|psi| < 0.9)
    lies within `band` rows and columns of it."""
    rng = np.random.default_rng(7)
    pc = torch.as_tensor(np.sign(rng.standard_normal((2, 9, 13))) * (rng.random((2, 9, 13)) > 0.1),
                         dtype=torch.float32)
    iface = pc.abs() < 0.9
    for band in (1, 2):
        got = sg.near_interface(pc, band)
        assert got.shape == pc.shape and torch.equal(sg.near_interface(pc[1], band), got[1])
        for n in range(2):
            for i in range(9):
                for j in range(13):
                    w = iface[n, max(0, i - band):i + band + 1, max(0, j - band):j + band + 1]
                    assert float(got[n, i, j]) == float(w.any())


def test_flow_specs():
    s1, s2 = sg.flow_spec(5), sg.flow_spec(5)
    assert s1 == s2
    for k in range(50):
        s = sg.flow_spec(k)
        assert 0.6 <= s["porosity"] <= 0.8 and 1.0 <= s["M"] <= 20.0 and 30.0 <= s["theta"] <= 90.0
    b = sg.breakthrough_spec()
    assert (b["H"], b["W"], b["porosity"], b["M"], b["theta"]) == (1024, 2048, 0.7, 15.0, 45.0)
    m, l, p = sg.build_flow(dict(sg.flow_spec(3), H=64, W=320))
    assert m.shape == (64, 320) and l.shape == (64, 320)
    bed = m[:, 64:256]
    assert abs(bed.mean() - (1 - (1 - sg.flow_spec(3)["porosity"]) * 320 / 192)) < 0.08
    assert m[:, :64].all() and m[:, -64:].all()          # inlet and outlet kept clear
    assert m[:2].all() and m[-2:].all()                  # obstacles stay inside the rows (no wrap)
    assert p.u0 == sg.U_IN and p.n_inv_cols == sg.INV_COLS and p.y_bc == 0
    assert sg.probe_column(2048) == 2047 - 8


def test_unet_counts():
    net = sg.UNet(channels=(8, 16, 16, 16), seed=0)
    c0, c1, c2, c3 = 8, 16, 16, 16
    plan = [(8, c0), (c0, c0), (c0, c1), (c1, c1), (c1, c2), (c2, c2), (c2, c3), (c3, c3), (c3 + c2, c2), (c2, c2),
            (c2 + c1, c1), (c1, c1), (c1 + c0, c0), (c0, c0), (c0, 8)]
    assert net.n_params() == sum(9 * a * b + b for a, b in plan)
    H, W = 16, 32
    res = [1, 1, 2, 2, 4, 4, 8, 8, 4, 4, 2, 2, 1, 1, 1]
    want = sum(2 * 9 * a * b * (H // r) * (W // r) for (a, b), r in zip(plan, res))
    assert net.flops_per_sample(H, W) == want
    full = sg.UNet()
    assert 1.0e6 < full.n_params() < 1.5e6


def test_model_file_round_trip(tmp_path):
    from fingering_dynamics_amd.surrogate_run import save_model, load_model
    a = sg.UNet(channels=(8, 8, 16, 16), seed=1)
    b = sg.UNet(channels=(8, 8, 16, 16), seed=2)
    path = str(tmp_path / "w.npz")
    save_model(a, path)
    load_model(b, path)
    for (ka, va), (kb, vb) in zip(a.state_dict().items(), b.state_dict().items()):
        assert ka == kb and torch.equal(va, vb)


def test_pipeline_rejects_extra_seeds_with_a_loaded_model(tmp_path):
    """--extra-seeds trains on generated data, so it cannot follow
    --load-model; the parser refuses before anything touches the GPU."""
    import pytest
    from fingering_dynamics_amd import surrogate_run
    with pytest.raises(SystemExit):
        surrogate_run.main(["--out", str(tmp_path), "--extra-seeds", "1", "--load-model", "w.npz"])


def test_jump_change_clamps_bands_and_scales_to_the_inflow():
    """The coarse change of one jump: the prediction clamped so psi stays in
    [-1, 1], zero on solid cells and beyond `band` cells of the interface,
    and scaled so that sum(change * n_c) equals the inflow, with the scale
    held to [0.5, 2]."""
    pc = torch.tensor([[1.0, 1.0, 0.2, -1.0, -1.0, -1.0, -1.0]] * 2)
    dpsi = torch.tensor([[0.5, 0.0, 0.3, 0.4, 0.2, 0.1, 0.1]] * 2)
    fluid = torch.ones_like(pc)
    fluid[1, 3] = 0.0
    n_c = torch.full_like(pc, 64.0)
    d, s = sg.jump_change(pc, dpsi, fluid, n_c, band=0)
    assert float(s) == 1.0
    want = torch.tensor([[0.0, 0.0, 0.3, 0.4, 0.2, 0.1, 0.1], [0.0, 0.0, 0.3, 0.0, 0.2, 0.1, 0.1]])
    assert torch.allclose(d, want)                           # 1 + 0.5 clamps to 1: no change
    d1, _ = sg.jump_change(pc, dpsi, fluid, n_c, band=1)    # interface cell: column 2 only
    assert torch.allclose(d1, want * (torch.arange(7) <= 3).float())
    gain = float((d1 * n_c).sum())
    d2, s2 = sg.jump_change(pc, dpsi, fluid, n_c, band=1, inflow=torch.tensor(1.5 * gain))
    assert abs(float(s2) - 1.5) < 1e-6 and abs(float((d2 * n_c).sum()) - 1.5 * gain) < 1e-3
    _, s3 = sg.jump_change(pc, dpsi, fluid, n_c, band=1, inflow=torch.tensor(10.0 * gain))
    _, s4 = sg.jump_change(pc, dpsi, fluid, n_c, band=1, inflow=torch.tensor(0.1 * gain))
    assert float(s3) == 2.0 and float(s4) == 0.5
    _, s5 = sg.jump_change(pc, -dpsi.abs() - 2.0, fluid, n_c, band=1, inflow=torch.tensor(5.0))
    assert float(s5) == 1.0                                  # no gain: nothing to scale


def test_inputs_layout_and_scaling():
    """Channels: psi, ux / U_REF, uy / U_REF, solid fraction, log M / log 20,
    theta / 180, then two zero channels; bf16, contiguous, one M and theta
    per sample."""
    N, H, W = 2, 3, 4
    g = torch.Generator().manual_seed(0)
    psi, ux, uy, sol = (torch.rand((N, H, W), generator=g) for _ in range(4))
    x = sg.inputs(psi, ux * sg.U_REF, uy * sg.U_REF, sol, [20.0, 1.0], [90.0, 135.0])
    assert x.shape == (N, H, W, sg.IN_CH) and x.dtype == torch.bfloat16 and x.is_contiguous()
    for c, want in enumerate((psi, ux, uy, sol)):
        assert torch.equal(x[..., c], want.to(torch.bfloat16))
    assert torch.all(x[0, ..., 4] == 1.0) and torch.all(x[1, ..., 4] == 0.0)
    assert torch.all(x[0, ..., 5] == 0.5) and torch.all(x[1, ..., 5] == 0.75)
    assert torch.all(x[..., 6:] == 0)


def test_breakthrough_probe_matches_the_output_module():
    """sg.reached and sg.status read the probe column as output.outlet_hits
    does: fluid nodes only, psi strictly above the threshold."""
    from fingering_dynamics_amd import output
    rng = np.random.default_rng(4)
    H, W, col = 6, 10, 7
    fluid = torch.as_tensor(rng.random((H, W)) > 0.3, dtype=torch.float32)
    for k in range(20):
        fields = torch.as_tensor(rng.uniform(-1.0, 0.2, (4, H, W)), dtype=torch.float32)
        for thr in (0.0, 0.1):
            want = output.outlet_hits(fields[0][:, col].numpy(), fluid[:, col].numpy(), thr).size > 0
            assert bool(sg.reached(fields, fluid, col, thr)) == want
        assert sg.status(fields, fluid, col) == (bool(sg.reached(fields, fluid, col)), False)
    fields = -torch.ones((4, H, W))
    solid_row = int(torch.nonzero(fluid[:, col] == 0)[0])
    fields[0, solid_row, col] = 1.0                    # a solid node above the threshold is not a breakthrough
    assert sg.status(fields, fluid, col) == (False, False)
    r, c = [int(v) for v in torch.nonzero(fluid)[0]]
    fields[1, r, c] = float("nan")                     # non-finite rho on a fluid node
    assert sg.status(fields, fluid, col) == (False, True)


def test_pairs_keep_t_plus_jump_among_the_usable_snapshots():
    data = {"psi": torch.zeros((3, 10, 1, 1)), "valid": [10, 4, 1]}
    items, d = sg.pairs(data, [0, 1, 2], jump=2000, snap=500)
    assert d == 4
    assert items == [(0, t) for t in range(6)]        # domain 1 has 4 usable snapshots: t + 4 never fits
    items, _ = sg.pairs(data, [1], jump=1000, snap=500)
    assert items == [(1, 0), (1, 1)]


def test_ensemble_domain_specs_and_square_beds():
    """The many-domain workload's specs stay in their stated ranges and are
    reproducible; square beds keep a one-node gap between squares, stay
    inside their column and row bounds and never exceed the target solid
    fraction; a built domain carries the spec's flow parameters."""
    from fingering_dynamics_amd import ensemble, wetting
    for k in range(40):
        s = ensemble.domain_spec(k)
        assert s == ensemble.domain_spec(k) and s != ensemble.domain_spec(k, seed=4)
        assert s["W"] % 32 == 0 and 256 <= s["W"] <= 1024 and s["H"] % 32 == 0 and 128 <= s["H"] <= 512
        assert 0.6 <= s["porosity"] <= 0.85 and 1.0 <= s["M"] <= 20.0 and 31.0 <= s["theta"] <= 90.0
    H, W, size = 64, 200, 6
    bed = ensemble._square_bed(H, W, 0.8, seed=9, x_lo=24, x_hi=W - 25, size=size)
    assert 0 < bed.sum() <= 0.2 * H * W and bed.sum() % (size * size) == 0
    ys, xs = np.nonzero(bed)
    assert xs.min() >= 24 and xs.max() < W - 25 and ys.min() >= 2 and ys.max() <= H - 3
    # squares: flood-label the bed and check each component is one size x size block
    from scipy import ndimage
    lab, n = ndimage.label(bed, structure=np.ones((3, 3)))   # 8-connectivity: touching corners merge
    assert n * size * size == bed.sum()
    spec = next(ensemble.domain_spec(k) for k in range(40) if not ensemble.domain_spec(k)["circles"]
                and ensemble.domain_spec(k)["H"] * ensemble.domain_spec(k)["W"] <= 256 * 512)
    mask, links, p = ensemble.build_domain(spec)
    assert mask.shape == (spec["H"], spec["W"]) and links.shape == mask.shape
    assert p.H == spec["H"] and p.W == spec["W"] and p.M == spec["M"] and p.u0 == 0.001
    assert p.y_bc == 0 and p.psi_visc_bound == 1.0      # periodic y, |psi| bound in tau_mix
    assert p.psi_wall == wetting.psi_wall_for_angle(180.0 - spec["theta"])


def test_volume_term_of_the_loss():
    """The volume term adds LAMBDA_V times the squared error of each
    sample's total psi change, relative to the batch's mean true total: zero
    when every sample's predicted total equals the true one (whatever the
    per-cell errors), and quadratic in a uniform under-prediction."""

    class Const(torch.nn.Module):
        def __init__(self, out):
            super().__init__()
            self.out = out

        def forward(self, x):
            return self.out

    B, H, W = 2, 4, 6
    x = torch.zeros(B, H, W, sg.IN_CH)
    p0 = torch.zeros(B, H, W)
    p1 = torch.full((B, H, W), 0.02)
    v = torch.zeros(B, H, W, 2)
    wf = torch.ones(B, H, W)
    old = sg.LAMBDA_V
    try:
        # same totals, different cells: the volume term is zero
        out = torch.zeros(B, H, W, 3)
        out[:, 0, :, 0] = 0.02 * H
        sg.LAMBDA_V = 0.0
        base = float(sg.loss_fn(Const(out), x, p0, p1, v, v, wf))
        sg.LAMBDA_V = 0.5
        assert math.isclose(float(sg.loss_fn(Const(out), x, p0, p1, v, v, wf)), base, rel_tol=1e-6)
        # a uniform under-prediction by a fraction e adds LAMBDA_V * e^2
        for e in (0.25, 0.5):
            out = torch.zeros(B, H, W, 3)
            out[..., 0] = 0.02 * (1 - e)
            sg.LAMBDA_V = 0.0
            base = float(sg.loss_fn(Const(out), x, p0, p1, v, v, wf))
            sg.LAMBDA_V = 0.5
            assert math.isclose(float(sg.loss_fn(Const(out), x, p0, p1, v, v, wf)) - base, 0.5 * e * e, rel_tol=1e-5)
    finally:
        sg.LAMBDA_V = old


def test_channel_slice_layout_of_the_pooled_skip_gradient():
    """The layouts pool_skip_bwd reads in place (contiguous, or a channel
    slice at a multiple of 8) and those it does not (any other stride or
    offset, or not bf16): the backward copies the latter first."""
    import torch
    from fingering_dynamics_amd import surrogate_ops as so
    wide = torch.zeros(2, 4, 6, 40, dtype=torch.bfloat16)
    assert so.channel_slice_ld(wide) == 40
    assert so.channel_slice_ld(wide[..., 8:24]) == 40
    assert so.channel_slice_ld(wide[..., 4:20]) is None          # offset not a multiple of 8
    assert so.channel_slice_ld(wide[:, :, ::2, :16]) == 80       # every other pixel: still uniform
    assert so.channel_slice_ld(wide[:, ::2, :, :16]) is None     # rows skipped: not one pixel stride
    assert so.channel_slice_ld(wide.permute(0, 2, 1, 3)) is None
    assert so.channel_slice_ld(wide.float()) is None
    assert so.channel_slice_ld(torch.zeros(2, 4, 6, 12, dtype=torch.bfloat16)) is None   # pixels 12 apart
