"""Throughput of the fused two-phase D2Q9 step on one MI355X.

Workload (the single-domain configuration): 4096 x 2048 nodes (W x H),
Zou-He velocity inlet u = 0.001, pressure outlet rho = 1, half-way walls on
the bottom and top rows, seeded circular obstacles of radius 8..24 to
porosity 0.75 (seed 2), viscosity ratio M = 10, FP32 populations.  A step is
one iteration of the hot loop of lattice_boltzmann/fingering_periodic.py
(collide, stream, bounce-back, walls, Zou-He, moments) over the whole
domain; the populations stay resident in HBM.

Prints ONE JSON line.  ``value`` is MLUPS over all H*W nodes.  ``roofline``
prices the dominant kernel at 144 algorithmic bytes per fluid-node update
(9 f + 9 g FP32 read and written) against 8 TB/s, using the kernel's average
launch time from HIP events on the solver's stream.  By default one launch
advances two iterations (the intermediate state stays in registers), so a
launch carries 2 x 144 algorithmic bytes per fluid node while moving the
populations through HBM once.  ``cpu_baseline`` times
the FP64 numpy restatement (the oracle, one core) on a few iterations of the
same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps S]
"""
import argparse
import json
import os
import platform
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

H, W = 2048, 4096
BYTES_PER_FLUID_UPDATE = 144          # 9 f + 9 g, FP32, read + write
HBM_PEAK_GBS = 8000.0                 # MI355X HBM3E
CONTACT_ANGLE = 60.0                  # degrees, through the defending (psi = -1) fluid
U_IN = 0.001                          # inlet velocity (see workload())


def wall_value():
    """psi_wall for CONTACT_ANGLE from the droplet calibration
    (fingering_dynamics_amd/data/wetting_table.json).  The table gives the
    angle through the psi = +1 (invading) phase, so 60 degrees through the
    defending fluid is 120 degrees there: psi_wall = -0.349 (the reference's
    own -0.5 is 47 degrees by the same measure).  Measured through the
    invading fluid instead (psi_wall = +0.355), the invading phase wets every
    obstacle against the defending bulk and the run diverges within 2,500
    iterations at the reference's interface constants."""
    from fingering_dynamics_amd import wetting
    return wetting.psi_wall_for_angle(180.0 - CONTACT_ANGLE)


def workload(h=H, w=W):
    """Masks, link bits and parameters of the 4096 x 2048 channel (or of an
    h x w copy built by the same generator, for the parity tests).

    Zou-He inlet with a constant velocity, pressure outlet rho = 1, half-way
    walls on rows 0 and H-1, circles of radius 8..24 to porosity 0.75 kept
    64 columns clear of the inlet and outlet, M = 10, contact angle 60
    degrees (wall_value()), psi = +1 in the first 40 columns.  Bounce-back uses the complete half-way link set (every solid
    -> fluid link), so solid nodes are inert.  The inlet velocity is 0.001:
    at 0.01 the Darcy pressure drop across 4096 columns of this bed
    (about 3 nu L u / k in density units, ~3) is far beyond what the weakly
    compressible scheme carries, and every variant tried diverged within
    2000-7000 iterations; the step's arithmetic and memory traffic do not
    depend on these values."""
    from fingering_dynamics_amd import geometry
    from fingering_dynamics_amd.params import LbmParams, Y_WALLS, PROFILE_CONST, OUTLET_PRESSURE

    circles = geometry.random_circles(h, w, 8, 24, 0.75, seed=2, margin=2, x_lo=64, x_hi=w - 65)
    bpa, _, _, _ = geometry.Createblock(h, w).setCirleblock(circles)
    mask = geometry.fluid_mask(bpa)
    bits = geometry.link_bits_complete(mask)
    params = LbmParams.periodic(
        h, w, M=10.0, psi_wall=wall_value(), u0=U_IN, profile=PROFILE_CONST,
        outlet=OUTLET_PRESSURE, rho_out=1.0, y_bc=Y_WALLS, wall_lo=0, wall_hi=h - 1,
        n_inv_cols=40)
    return mask, bits, params, circles


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE), as recorded in profiles/; None if not
    recorded."""
    path = os.path.join(ROOT, "profiles", "%s_4096x2048_hbm_traffic.txt" % kernel)
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def copy_probe(reps=10):
    """Best 1 GiB device-to-device streaming copy (read + write bytes / time)."""
    import ctypes
    from fingering_dynamics_amd import _native
    lib = _native.load()
    ms = ctypes.c_double()
    nbytes = 1 << 30
    _native.check(lib.fd_copy_probe(nbytes, reps, ctypes.byref(ms)))
    return ms.value, 2 * nbytes / (ms.value * 1e-3) / 1e9


CPU_WINDOW = (512, 1024)              # rows, columns of the CPU-baseline sample


def cpu_baseline(mask, bits, params, steps):
    """The FP64 oracle (numpy, one core) on `steps` iterations of the
    bottom-left 1024 x 512 window of the same workload (same obstacles,
    inlet, walls and parameters; the full domain costs about a minute per
    iteration in numpy).  Its bounce-back masks are the unpacked link bits
    (the geometry itself is not what is timed)."""
    from oracle import lbm64
    from fingering_dynamics_amd.geometry import unpack_links
    h, w = CPU_WINDOW
    m = np.ascontiguousarray(mask[:h, :w])
    b = np.ascontiguousarray(bits[:h, :w])
    p = params.with_(H=h, W=w, wall_lo=0, wall_hi=h - 1)
    o = lbm64.Lbm64.from_params(p, m, unpack_links(b))
    o.step(1)                                    # first iteration pays page faults
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return {
        "value": round(h * w * steps / dt / 1e6, 4),
        "unit": "MLUPS",
        "cores": 1,
        "kind": "port",
        "sample": "%d iterations of the %dx%d bottom-left window of the workload, FP64 numpy "
                  "restatement of fingering_periodic.py:454-479, %.1f s (host %s, %d cpus visible, "
                  "numpy single-threaded)" % (steps, w, h, dt, platform.processor() or platform.machine(),
                                              os.cpu_count() or 0),
    }


def ensemble_bench(n, steps, warmup):
    """Aggregate MLUPS of n independent domains (the seeded many-domain
    workload of fingering_dynamics_amd.ensemble.domain_spec) advanced by
    graph-replayed batched launches of two iterations each, against one
    launch per domain per launch step on the same stream, and against
    batched launches of one iteration.  Prints one JSON line."""
    from fingering_dynamics_amd.ensemble import Ensemble, domain_spec, build_domain
    from fingering_dynamics_amd.lbm import Compute
    t0 = time.perf_counter()
    cs = [Compute(*build_domain(domain_spec(k))) for k in range(n)]
    setup_s = time.perf_counter() - t0
    ens = Ensemble(cs)
    tune = ens.autotune(steps=10)
    ens.step(warmup)
    ens.synchronize()
    nodes = ens.nodes
    res = {}
    fusion, depth = ens.fusion, ens.depth
    for mode in ("graph", "eager", "sequential", "one_iteration_per_launch"):
        k = steps if mode != "sequential" else max(steps // 8, 10)
        k += k & 1
        ens.use_graph(mode != "eager")
        if mode == "one_iteration_per_launch":
            ens.set_fusion(1)
            ens.step(4)
        t1 = time.perf_counter()
        (ens.step_sequential if mode == "sequential" else ens.step)(k)
        ens.synchronize()
        wall = time.perf_counter() - t1
        res[mode] = {"steps": k, "ms_per_step": round(ens.last_step_ms(), 4),
                     "wall_ms_per_step": round(wall * 1e3 / k, 4),
                     "mlups": round(nodes * k / wall / 1e6, 1)}
    bad = [c.mass()[2] for c in cs]
    out = {
        "metric": "aggregate MLUPS of %d independent domains, batched launches vs one launch per domain" % n,
        "value": res["graph"]["mlups"], "unit": "MLUPS", "n_gpus": 1, "higher_is_better": True,
        "dtype": "fp32", "data": "synthetic: seeded squares/circles, sizes 256..1024 x 128..512",
        "config": {"workload": "many-domain ensemble", "domains": n, "nodes": nodes},
        "batched_graph": res["graph"], "batched_eager": res["eager"], "sequential": res["sequential"],
        "batched_one_iteration_per_launch": res["one_iteration_per_launch"],
        "speedup_vs_sequential": round(res["graph"]["mlups"] / res["sequential"]["mlups"], 2),
        "iterations_per_launch": depth,
        "autotune_ms_per_step": {"%dx%d" % t: round(v, 4) for t, v in tune.items()},
        "shape": "%dx%d" % fusion[1:], "tile_one_iteration": "%dx%d" % ens.tile, "setup_s": round(setup_s, 1),
        "domains_non_finite": int(sum(1 for v in bad if v)), "iterations_run": int(cs[0].t),
    }
    ens.close()
    for c in cs:
        c.close()
    print(json.dumps(out), flush=True)


def cfg1_bench(steps):
    """The 400 x 100 channel of the reference-faithful configuration
    (periodic y, Gaussian Zou-He inlet/outlet, 56 seeded 6x6 squares, M = 5,
    psi_wall = 0): `steps` iterations, one launch per iteration or two, eager
    or replayed from a captured graph, each at the library's default launch
    shape for this size (one row per wave).  Its state (6 MB) stays in the
    caches, so this measures launch overhead and latency, not HBM.  Prints
    one JSON line."""
    from fingering_dynamics_amd.lbm import Compute
    from fingering_dynamics_amd.params import LbmParams
    from fingering_dynamics_amd import geometry
    h, w = 100, 400
    rects = geometry.random_squares(h, w, 56, size=6, seed=1, x_min=41)
    bpa, corners = geometry.Createblock(h, w).setblock(rects)
    mask = geometry.fluid_mask(bpa)
    bits = geometry.Bounce_back(h, w).link_bits_rec(corners)
    params = LbmParams.periodic(h, w, M=5.0, psi_wall=0.0, n_inv_cols=40)
    res = {}
    for depth in (2, 1):
        for graph in (True, False):
            cm = Compute(mask, bits, params)
            if depth == 1:
                cm.set_fusion(1)
            cm.use_graph(graph)
            cm.step(100)
            cm.synchronize()
            t0 = time.perf_counter()
            cm.step(steps)
            cm.synchronize()
            dt = time.perf_counter() - t0
            res["%s_%s" % ("two_per_launch" if depth == 2 else "one_per_launch", "graph" if graph else "eager")] = {
                "us_per_iteration": round(dt * 1e6 / steps, 3), "device_us_per_iteration": round(cm.last_step_ms() * 1e3, 3),
                "mlups": round(h * w * steps / dt / 1e6, 1)}
            bad = cm.mass()[2]
            cm.close()
    from oracle import lbm64
    o = lbm64.Lbm64.from_params(params, mask, geometry.unpack_links(bits))
    o.step(1)
    n_cpu = 200
    t0 = time.perf_counter()
    o.step(n_cpu)
    cpu_dt = time.perf_counter() - t0
    best = max(res, key=lambda k: res[k]["mlups"])
    print(json.dumps({
        "metric": "MLUPS of the 400x100 channel (%d iterations), by launch mode" % steps,
        "value": res[best]["mlups"], "unit": "MLUPS", "best_mode": best, "modes": res,
        "graph_speedup_two_per_launch": round(res["two_per_launch_graph"]["mlups"] / res["two_per_launch_eager"]["mlups"], 2),
        "non_finite_after": bad,
        "cpu_baseline": {"value": round(h * w * n_cpu / cpu_dt / 1e6, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
                         "sample": "%d iterations of the same channel, FP64 numpy restatement, %.1f s" % (n_cpu, cpu_dt)},
        "config": {"workload": "400x100 channel, periodic y, Gaussian Zou-He, 56 squares 6x6, M=5, psi_wall=0"},
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="iterations of the CPU baseline sample (0 skips it)")
    ap.add_argument("--tile", default="", help="launch shape WX,ROWS of the one-iteration kernel")
    ap.add_argument("--fusion", type=int, default=0, choices=(0, 1, 2, 3),
                    help="iterations per launch (1, 2 or 3; default 0: the library's choice for the domain)")
    ap.add_argument("--ftile", default="", help="launch shape WX,ROWS of the two-iteration kernel")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only run the plain-copy HBM probe (1 GiB copies) and print its bandwidth")
    ap.add_argument("--cfg1", action="store_true",
                    help="instead: the 400x100 channel, eager vs graph-replayed launches (--steps, default 5000)")
    ap.add_argument("--ensemble", type=int, default=0, metavar="N",
                    help="instead: the many-domain workload with N domains, batched vs sequential launches")
    args = ap.parse_args()
    if args.hbm_probe:
        ms, gbps = copy_probe()
        print(json.dumps({"probe": "streaming copy (fd_copy_probe)", "bytes_read": 1 << 30,
                          "bytes_written": 1 << 30, "ms_per_copy": ms, "GBps": gbps}), flush=True)
        return
    if args.cfg1:
        cfg1_bench(args.steps if args.steps != 20000 else 5000)
        return
    if args.ensemble:
        ensemble_bench(args.ensemble, args.steps if args.steps != 20000 else 2000, args.warmup if args.warmup != 500 else 20)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path "
              "(x-slab sharding with halo exchange is a design, not built)", file=sys.stderr)
        sys.exit(2)
    # single process: any launcher-provided rank other than 0 has nothing to do
    if int(os.environ.get("RANK", "0")) != 0:
        sys.exit(0)

    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ.setdefault(k, "1")

    from fingering_dynamics_amd.lbm import Compute
    mask, bits, params, _ = workload()
    n_fluid = int(mask.sum())
    cm = Compute(mask, bits, params)
    if args.tile:
        wx, rows = (int(v) for v in args.tile.split(","))
        cm.set_tile(wx, rows)
    fwx, frows = (int(v) for v in args.ftile.split(",")) if args.ftile else (0, 0)
    if args.fusion or args.ftile:
        cm.set_fusion(args.fusion or cm.fusion[0], fwx, frows)
    fusion = cm.fusion
    cm.step(args.warmup)
    cm.synchronize()
    t0 = time.perf_counter()
    cm.step(args.steps)
    cm.synchronize()
    dt = time.perf_counter() - t0
    kernel_ms = cm.last_step_ms()
    s_rho, s_psi, bad = cm.mass()
    _, tau_margin = cm.watchdog()
    cm.close()
    _, copy_gbps = copy_probe()

    ms_per_step = dt * 1e3 / args.steps
    mlups = H * W * args.steps / dt / 1e6
    # the dominant kernel: fused3_kernel (three iterations per launch),
    # fused_kernel (two) or step_kernel (one); kernel_ms is the device time
    # per iteration
    per_launch = fusion[0]
    kernel = {3: "fused3_kernel", 2: "fused_kernel", 1: "step_kernel"}[per_launch]
    algo_bytes = per_launch * BYTES_PER_FLUID_UPDATE * n_fluid
    launch_ms = kernel_ms * per_launch
    achieved = algo_bytes / (launch_ms * 1e-3) / 1e9 if kernel_ms > 0 else None
    traffic = traffic_from_profiles(kernel)
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
        "traffic": traffic,
        "kernel": kernel,
        "iterations_per_launch": per_launch,
        "kernel_ms": round(launch_ms, 5),
        "kernel_ms_per_iteration": round(kernel_ms, 5),
        "algorithmic_bytes_per_launch": algo_bytes,
        "fluid_nodes": n_fluid,
        # measured HBM bytes (rocprofv3, profiles/) over this run's kernel time
        "hbm_rate_GBps": round(traffic / (launch_ms * 1e-3) / 1e9, 1) if traffic and kernel_ms > 0 else None,
        # the best streaming copy this device reached in the same process
        "copy_GBps": round(copy_gbps, 1),
    }
    cpu = cpu_baseline(mask, bits, params, args.cpu_steps) if args.cpu_steps > 0 else None
    out = {
        "metric": "MLUPS (all H*W nodes) of the fused two-phase D2Q9 step; % of HBM roofline",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded circular obstacles (seed 2), initial psi=+1 in the first 40 columns",
        "config": {
            "workload": "single domain 4096x2048 (W x H), Zou-He velocity inlet (u=%g), pressure outlet, "
                        "half-way walls, seeded circles r 8-24 to porosity 0.75, M=10, contact angle %g deg "
                        "through the defending fluid (psi_wall=%.4f)" % (U_IN, CONTACT_ANGLE, params.psi_wall),
            "H": H, "W": W, "porosity": round(n_fluid / (H * W), 4), "M": params.M,
            "psi_wall": params.psi_wall,
        },
        "mflups": round(n_fluid * args.steps / dt / 1e6, 1),
        "mass_check": {"sum_rho": s_rho, "sum_psi": s_psi, "non_finite": bad, "min_tau_minus_half": tau_margin},
        "launch": {"iterations_per_launch": fusion[0], "waves_per_workgroup": fusion[1], "rows_per_wave": fusion[2]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
