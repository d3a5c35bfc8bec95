"""The surrogate pipeline end to end on one GPU, as one job:

1. trajectories of `--domains` flows (surrogate.flow_spec, H x W), coarsened
   by f every `--snap` iterations (batched LBM, one launch per iteration);
2. training of the U-Net on pairs (t, t + jump) of all but `--holdout`
   flows (AdamW, batch `--bs`);
3. evaluation on the held-out flows against persistence;
4. the MFMA throughput of a training step at batch 64 on the 256 x 128
   coarse grid of the 2048 x 1024 workload;
5. the 2048 x 1024 breakthrough workload (surrogate.breakthrough_spec) run
   to breakthrough by the plain LBM and by the surrogate-accelerated loop,
   with wall times, breakthrough iterations and the coarse interface error
   of the accelerated run against the plain one at equal iterations.

    python -m fingering_dynamics_amd.surrogate_run --out DIR [options]

Writes DIR/surrogate_run.json (every figure above) and DIR/unet.npz (the
trained weights); --load-model NPZ skips steps 1-3.  `--mfma-only` runs
step 4 alone (for profiling the convolution kernels).  Further options:
--variants (more accelerated runs: relax, velocity scale, band),
--dagger-cycles (fine-tuning on the coupled loop's own states, labelled
by the solver), --phase-weight (the phase cross-entropy term of the loss),
--teacher (jumps to the plain run's own coarse future: the coupling with
a perfect coarse predictor), --workload-seed (another obstacle bed).
"""
import argparse
import json
import os
import sys

import numpy as np
import torch

from . import surrogate as sg
from .lbm import Compute

BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16 MFMA


def log(*a):
    print(*a, flush=True)


def save_model(model, path):
    np.savez(path, **{k: v.detach().cpu().numpy() for k, v in model.state_dict().items()})


def load_model(model, path):
    with np.load(path, allow_pickle=False) as z:
        model.load_state_dict({k: torch.as_tensor(z[k]) for k in z.files})
    return model


class MeanNet(torch.nn.Module):
    """The mean of several networks' outputs (an ensemble surrogate)."""

    def __init__(self, nets):
        super().__init__()
        self.nets = torch.nn.ModuleList(nets)

    def forward(self, x):
        return torch.stack([n(x) for n in self.nets]).mean(dim=0)


def mfma_throughput(model, H=128, W=256, bs=64, iters=20):
    """Forward + backward of the U-Net on a random batch: TFLOP/s of the
    convolutions (3x the forward FLOPs) against the bf16 dense peak."""
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn(bs, H, W, sg.IN_CH, device="cuda", generator=g).to(torch.bfloat16)
    tgt = torch.randn(bs, H, W, 3, device="cuda", generator=g)
    for _ in range(3):
        ((model(x) - tgt) ** 2).mean().backward()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        model.zero_grad(set_to_none=True)
        ((model(x) - tgt) ** 2).mean().backward()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / iters
    flops = 3.0 * model.flops_per_sample(H, W) * bs
    tf = flops / (ms * 1e-3) / 1e12
    return {"batch": bs, "H": H, "W": W, "ms_per_step": ms, "tflops": tf, "peak_tflops": BF16_PEAK_TFLOPS,
            "frac": tf / BF16_PEAK_TFLOPS, "gflop_per_step": flops / 1e9}


def select_network(a, nets, seeds, log=print):
    """Model selection on a validation bed: the breakthrough workload's
    flow with obstacle seed a.select_workload_seed (a different bed from the
    target's), run plainly and accelerated by each network with the main
    jump, relax, band and threshold; the network whose breakthrough comes
    closest to the plain run's is selected."""
    spec = dict(sg.breakthrough_spec(), obstacle_seed=a.select_workload_seed)
    mask, links, params = sg.build_flow(spec)
    cm = Compute(mask, links, params)
    t_p, wall_p, _ = sg.plain_run(cm, a.f, a.max_steps, chunk=a.relax, record_every=a.jump + a.relax, log=log)
    cm.close()
    out = {"workload_seed": a.select_workload_seed, "plain": {"breakthrough_t": t_p, "wall_s": wall_p}, "runs": []}
    best, best_err = 0, None
    for k, (net, sd) in enumerate(zip(nets, seeds)):
        cm = Compute(mask, links, params)
        t_a, wall_a, stats, _ = sg.accelerated_run(cm, net, spec["M"], 180.0 - spec["theta"], a.f, a.jump, a.relax,
                                                   a.threshold, a.max_steps, mass_fix=not a.no_mass_fix,
                                                   velocity=a.velocity, band=a.band, log=log)
        cm.close()
        err = (t_a - t_p) / t_p if (t_a and t_p) else None
        out["runs"].append({"seed": sd, "breakthrough_t": t_a, "wall_s": wall_a, "rel_error": err,
                            "speedup_wall": wall_p / wall_a if wall_a > 0 else None})
        log("selection: seed %d breakthrough %s (plain %s)" % (sd, t_a, t_p))
        if err is not None and (best_err is None or abs(err) < best_err):
            best, best_err = k, abs(err)
    out["selected_index"], out["selected_seed"] = best, seeds[best]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--domains", type=int, default=48)
    ap.add_argument("--H", type=int, default=256)
    ap.add_argument("--W", type=int, default=2048)
    ap.add_argument("--f", type=int, default=8)
    ap.add_argument("--snap", type=int, default=500)
    ap.add_argument("--data-steps", type=int, default=320000)
    ap.add_argument("--holdout", type=int, default=8)
    ap.add_argument("--jump", type=int, default=2000)
    ap.add_argument("--relax", type=int, default=200)
    ap.add_argument("--threshold", type=float, default=0.02)
    ap.add_argument("--iters", type=int, default=8000)
    ap.add_argument("--bs", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--max-steps", type=int, default=3000000)
    ap.add_argument("--no-mass-fix", action="store_true")
    ap.add_argument("--load-model")
    ap.add_argument("--mfma-only", action="store_true")
    ap.add_argument("--skip-plain", action="store_true", help="accelerated run only")
    ap.add_argument("--variants", default="",
                    help="further accelerated runs as relax[:velocity-scale[:band[:threshold]]], "
                         "e.g. 500:0:2,1000:0.5:2,200:0:2:0.005")
    ap.add_argument("--band", type=int, default=2, help="coarse cells around the interface a jump may change")
    ap.add_argument("--dagger-cycles", type=int, default=0,
                    help="coupled-loop cycles on the training flows whose states, labelled by the solver, "
                         "fine-tune the network (0: none)")
    ap.add_argument("--dagger-iters", type=int, default=4000)
    ap.add_argument("--teacher-velocity", action="store_true",
                    help="with --teacher: the jumps' velocity change is the plain run's too (use --velocity 1)")
    ap.add_argument("--teacher", action="store_true",
                    help="diagnostic: jumps to the plain run's own coarse psi (perfect coarse predictor)")
    ap.add_argument("--extra-seeds", default="",
                    help="further networks (initialisation and batch order seeds, e.g. 1,2,3) trained on the "
                         "same data and run through the same accelerated runs: the spread over training")
    ap.add_argument("--select-workload-seed", type=int, default=None,
                    help="choose, among the networks trained (the first and --extra-seeds), the one whose "
                         "breakthrough on the workload with this obstacle seed is closest to that workload's "
                         "plain run (a validation bed, not the target), and report its target run as 'selected'")
    ap.add_argument("--ensemble-eval", action="store_true",
                    help="with --extra-seeds: also run the accelerated runs with the mean of all the networks")
    ap.add_argument("--workload-seed", type=int, default=None,
                    help="obstacle seed of the 2048 x 1024 workload (default 5)")
    ap.add_argument("--phase-weight", type=float, default=sg.LAMBDA_S,
                    help="weight of the phase cross-entropy term of the training loss")
    ap.add_argument("--volume-weight", type=float, default=sg.LAMBDA_V,
                    help="weight of each sample's total psi-change error in the training loss")
    ap.add_argument("--loss-band", type=int, default=sg.LOSS_BAND,
                    help="train on cells within this many coarse cells of the interface only (0: all)")
    ap.add_argument("--regime", default="target", choices=sorted(sg.REGIMES))
    ap.add_argument("--velocity", type=float, default=0.0,
                    help="scale of the predicted velocity change applied at each jump (0: keep the solver's)")
    a = ap.parse_args(argv)
    if a.extra_seeds and a.load_model:
        ap.error("--extra-seeds trains on generated data; it cannot be combined with --load-model")
    sg.LAMBDA_S = a.phase_weight
    sg.LOSS_BAND = a.loss_band
    sg.LAMBDA_V = a.volume_weight
    os.makedirs(a.out, exist_ok=True)
    res = {"args": vars(a)}
    model = sg.UNet(seed=0).cuda()
    res["n_params"] = model.n_params()

    if a.mfma_only:
        res["mfma"] = mfma_throughput(model)
        log(json.dumps(res["mfma"]))
        return 0

    if a.load_model:
        load_model(model, a.load_model)
    else:
        log("generating %d flows %dx%d, %d iterations" % (a.domains, a.W, a.H, a.data_steps))
        data = sg.generate(a.domains, a.H, a.W, a.f, a.snap, a.data_steps, regime=a.regime, log=log)
        D = data["psi"].shape[0]
        res["data"] = {"domains": D, "dropped": data["dropped"], "snapshots": len(data["steps"]),
                       "non_finite": data["non_finite"], "broke_through": data["broke_through"],
                       "usable_snapshots": int(sum(data["valid"])),
                       "seconds": data["seconds"], "mlups": data["lattice_updates"] / data["seconds"] / 1e6}
        log(json.dumps(res["data"]))
        hold = min(a.holdout, max(D // 4, 1))
        tr, te = list(range(D - hold)), list(range(D - hold, D))
        res["eval_untrained"] = sg.evaluate(model, data, te, a.jump, a.snap)
        hist, secs = sg.train(model, data, tr, a.jump, a.snap, iters=a.iters, bs=a.bs, lr=a.lr, log=log)
        res["train"] = {"iters": a.iters, "seconds": secs, "loss_first": hist[0][1], "loss_last": hist[-1][1]}
        res["eval_train"] = sg.evaluate(model, data, tr[:hold], a.jump, a.snap)
        res["eval_heldout"] = sg.evaluate(model, data, te, a.jump, a.snap)
        log("held-out: %s" % json.dumps(res["eval_heldout"]))
        if a.dagger_cycles > 0:
            cp = sg.coupled_pairs(model, [data["flows"][i] for i in tr], [data["specs"][i] for i in tr], a.f,
                                  a.jump, a.relax, a.dagger_cycles, a.band, log=log)
            hist2, secs2 = sg.train(model, data, tr, a.jump, a.snap, iters=a.dagger_iters, bs=a.bs, lr=a.lr * 0.5,
                                    seed=1, extra=cp, log=log)
            res["dagger"] = {"pairs": cp["n"], "seconds": cp["seconds"], "train_seconds": secs2,
                             "loss_last": hist2[-1][1]}
            res["eval_heldout_after_dagger"] = sg.evaluate(model, data, te, a.jump, a.snap)
            log("held-out after coupled pairs: %s" % json.dumps(res["eval_heldout_after_dagger"]))
            del cp
        save_model(model, os.path.join(a.out, "unet.npz"))
        if not a.extra_seeds:
            del data
            torch.cuda.empty_cache()
    res["mfma"] = mfma_throughput(model)
    log("mfma: %s" % json.dumps(res["mfma"]))

    spec = sg.breakthrough_spec()
    if a.workload_seed is not None:
        spec["obstacle_seed"] = a.workload_seed
    mask, links, params = sg.build_flow(spec)
    res["workload"] = dict(spec, porosity_reached=float(mask.mean()), psi_wall=params.psi_wall, u0=params.u0,
                           probe_column=sg.probe_column(spec["W"]))
    cycle = a.jump + a.relax
    rec_p, rec_u = {}, ({} if a.teacher_velocity else None)
    if not a.skip_plain:
        cm = Compute(mask, links, params)
        t_b, wall, rec_p = sg.plain_run(cm, a.f, a.max_steps, chunk=a.relax,
                                        record_every=a.relax if a.teacher else cycle, velocity_record=rec_u,
                                        log=log)
        res["plain"] = {"breakthrough_t": t_b, "wall_s": wall, "mlups": cm.t * mask.size / wall / 1e6,
                        "non_finite": cm.mass()[2]}
        cm.close()
        log("plain: %s" % json.dumps(res["plain"]))
    model.eval()
    fluid = torch.as_tensor(mask, dtype=torch.float32).cuda()
    fluid_c = (fluid.view(spec["H"] // a.f, a.f, spec["W"] // a.f, a.f).sum(dim=(1, 3)) > 0).float()
    runs = [(a.relax, a.velocity, a.band, a.threshold)]
    for v in a.variants.split(","):
        if v:
            f = v.split(":")
            f += [str(a.velocity), str(a.band), str(a.threshold)][len(f) - 1:]
            runs.append((int(f[0]), float(f[1]), int(f[2]), float(f[3])))

    def run_accelerated(net):
        out = []
        for relax, vel, band, threshold in runs:
            # the plain run's coarse snapshots sit at chunk + k * (jump + relax)
            # only for the main relax length; other variants compare at breakthrough
            cm = Compute(mask, links, params)
            t_a, wall_a, stats, rec_a = sg.accelerated_run(cm, net, spec["M"], 180.0 - spec["theta"], a.f, a.jump,
                                                           relax, threshold, a.max_steps,
                                                           mass_fix=not a.no_mass_fix, velocity=vel, band=band,
                                                           teacher=rec_p if a.teacher else None,
                                                           teacher_velocity=rec_u, log=log)
            acc = dict(stats, relax=relax, velocity=vel, band=band, threshold=threshold, breakthrough_t=t_a,
                       wall_s=wall_a,
                       non_finite=cm.mass()[2])
            cm.close()
            if "plain" in res:
                p = res["plain"]
                acc["speedup_wall"] = p["wall_s"] / wall_a if wall_a > 0 else None
                acc["breakthrough_t_rel_error"] = ((t_a - p["breakthrough_t"]) / p["breakthrough_t"]
                                                   if t_a and p["breakthrough_t"] else None)
                common = sorted(set(rec_p) & set(rec_a))
                errs = [float(sg.interface_error(rec_a[t], rec_p[t], fluid_c)) for t in common]
                acc["interface_error_vs_plain_mean"] = float(np.mean(errs)) if errs else None
                acc["interface_error_vs_plain_last"] = errs[-1] if errs else None
                acc["compared_at"] = len(common)
            log("accelerated: %s" % json.dumps(acc))
            out.append(acc)
        return out

    res["accelerated"] = run_accelerated(model)
    nets = [model]
    for sd in [int(v) for v in a.extra_seeds.split(",") if v]:
        net = sg.UNet(seed=sd).cuda()
        hist_s, secs_s = sg.train(net, data, tr, a.jump, a.snap, iters=a.iters, bs=a.bs, lr=a.lr, seed=sd, log=log)
        ev = sg.evaluate(net, data, te, a.jump, a.snap)
        net.eval()
        nets.append(net)
        res.setdefault("seeds", []).append({"seed": sd, "train_seconds": secs_s, "eval_heldout": ev,
                                            "accelerated": run_accelerated(net)})
    if a.select_workload_seed is not None:
        res["selection"] = select_network(a, nets, [0] + [int(v) for v in a.extra_seeds.split(",") if v], log)
        k = res["selection"]["selected_index"]
        res["selected"] = dict(seed=res["selection"]["selected_seed"],
                               accelerated=(res["accelerated"] if k == 0 else res["seeds"][k - 1]["accelerated"])[0])
        log("selected: %s" % json.dumps(res["selected"], default=float))
    if a.ensemble_eval and len(nets) > 1:
        mean = MeanNet(nets).eval()
        res["ensemble"] = {"networks": len(nets), "eval_heldout": sg.evaluate(mean, data, te, a.jump, a.snap),
                           "accelerated": run_accelerated(mean)}
    log(json.dumps(res, indent=1, default=float))
    with open(os.path.join(a.out, "surrogate_run.json"), "w") as fh:
        json.dump(res, fh, indent=1, default=float)
    return 0


if __name__ == "__main__":
    sys.exit(main())
