"""Benchmark: grid points per second at standard error <= 1e-3 on one MI355X.

Workload (one "step" = one full solve: classify, walk, merge):
  3D screened Poisson on the unit icosphere at subdivision 4 (2562 vertices,
  5120 triangles), alpha = 10, constant source f = 1 sampled uniformly in the
  ball, g = 0, 32^3 grid over [-1, 1]^3, 4096 walks per inside point,
  epsilon = 1e-4, seed 7.
Forms: the FP32 persistent walker (the value) and the FP64 plain walker.

Grid pts/s at SE <= 1e-3 = #inside / sum_p n_p * t_walk,p with
n_p = ceil(var_p / 1e-6) from the FP64 run (shared by every form) and
t_walk,p = (T / total steps) * mean_steps_p of the form being measured.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-adjoint]
       python bench.py --fp32-only --steps 1 --warmup 0   (walker launches only,
                                                          for counter runs)
       python bench.py --workload stars | plate | welding  (batched stars, plate,
       welding transients)
       python bench.py --copy-calibration 1024            (one 1024 MiB fp32
                                                          read+write kernel, the
                                                          scale for HBM counters)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "walk-on-spheres_amd"))

METRIC = "grid pts/s at SE≤1e-3, FP32 persistent vs FP64 plain; adjoint walks/s"
WORKLOAD = ("3D screened Poisson, unit icosphere s4 (2562 v / 5120 t), alpha=10, f=1 uniform "
            "ball sampling, g=0, 32^3 grid, N=4096 walks/point, eps=1e-4, seed 7")
HBM_PEAK_GBS = 8000.0
SE_TARGET = 1e-3


def pts_per_s_at_se(n_inside, var_ref, msteps, seconds, total_steps):
    n_p = np.maximum(1.0, np.ceil(var_ref / (SE_TARGET ** 2)))
    t_step = seconds / max(1, total_steps)
    return n_inside / float(np.sum(n_p * t_step * msteps))


def cpu_baseline(V, F, eq, grid, inside_idx, var_ref, eps, seed, budget_s=12.0):
    """The reference core (oracle/_ref) on the host cores, on a bounded sample
    of the inside points; falls back to the C oracle if it is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cores = min(os.cpu_count() or 1, 16)
    pts_all = pyoracle.grid_points(pyoracle.Oracle().mesh(V, F).bbox(), *grid, 3)
    try:
        lib = pyoracle.Reference()
        kind = "reference"
    except FileNotFoundError:
        lib = pyoracle.Oracle()
        kind = "port"
    mesh = lib.mesh(V, F)
    # calibrate on a small sample, then size the sample to ~budget_s
    stride = max(1, len(inside_idx) // 256)
    pick = inside_idx[::stride]
    walks = 64

    def run(pick, walks):
        t0 = time.perf_counter()
        if kind == "reference":
            r = mesh.solve_points(eq, pts_all[pick], walks, eps, seed=seed, n_threads=cores,
                                  seed_index=pick)
        else:
            r = mesh.solve_points(eq, pts_all[pick], walks, eps, seed=seed, rng_mode=1,
                                  n_threads=cores, seed_index=pick)
        dt = time.perf_counter() - t0
        steps = float(np.nansum(r["mean_steps"]) * walks)
        return dt, steps, r

    dt, steps, r = run(pick, walks)
    scale = max(1.0, budget_s / max(dt, 1e-3))
    walks = int(min(4096, walks * scale))
    dt, steps, r = run(pick, walks)
    t_step = dt / steps
    ms = r["mean_steps"]
    # cost model over the sampled points, scaled to the full inside set
    n_p = np.maximum(1.0, np.ceil(var_ref[np.searchsorted(inside_idx, pick)] / SE_TARGET ** 2))
    pts_s = len(pick) / float(np.sum(n_p * t_step * ms))
    return dict(value=pts_s, unit="grid pts/s", cores=cores, kind=kind,
                steps_per_s=steps / dt, walks_per_s=len(pick) * walks / dt,
                sample=f"{len(pick)} of {len(inside_idx)} inside points (every {stride}th), "
                       f"{walks} walks each, {dt:.1f} s on {cores} threads")


def copy_calibration(mib):
    """One elementwise fp32 kernel that reads and writes `mib` MiB each: the
    known-size transfer the FETCH_SIZE / WRITE_SIZE counters are scaled by."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 2.0, out=dst)
    torch.cuda.synchronize()
    assert float(dst[-1]) == 2.0
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}))


# FP32 walker records read per query unit (bytes): a grid cell (GridCellF),
# a cell-list entry (GridEntry), a primitive (TriF 64 / SegF + 1/|v|^2 20) and
# an internal node (NodeF3 64 / NodeF2 48); per walk: its start record (120),
# value + steps (12), the FP64 terminal primitive (TriD 128 / SegD 32) and
# its slot -> primitive id (4).
# walk: FP32 start copy + the value and step count written;
# terminal: FP64 primitive record + slot -> primitive id (read unless g = 0)
RECORDS = {3: dict(cell=16, entry=8, prim=48, node=64, walk=32 + 12, terminal=128 + 4),
           2: dict(cell=16, entry=8, prim=16, node=48, walk=32 + 12, terminal=32 + 4)}


def walker_bytes(s, dim, terminal=True):
    """Algorithmic bytes of one FP32 walker launch from its query-unit counts:
    every query reads its grid cell; a list scan reads each entry and its
    primitive; a tree query reads its start primitive, each opened node and
    each leaf primitive; every walk reads its start and writes its result, and
    reads its terminal primitive when the boundary data is not zero."""
    r = RECORDS[dim]
    tree_q = s["q_queries"] - s["q_grid_bound"] - s["q_list_scans"]
    return (s["q_queries"] * r["cell"] + s["q_list_entries"] * r["entry"]
            + s["q_list_prims"] * r["prim"]
            + s["q_nodes"] * r["node"] + (s["q_leaf_prims"] + tree_q) * r["prim"]
            + s["total_walks"] * (r["walk"] + (r["terminal"] if terminal else 0)))


def torus_probes(n, seed):
    """n points uniform inside the torus R = 1, r = 0.35 (rejection in the box)."""
    rng = np.random.default_rng(seed)
    out = []
    while sum(len(o) for o in out) < n:
        p = rng.uniform([-1.35, -1.35, -0.35], [1.35, 1.35, 0.35], size=(4 * n, 3))
        q = (np.hypot(p[:, 0], p[:, 1]) - 1.0) ** 2 + p[:, 2] ** 2 < 0.35 ** 2
        out.append(p[q])
    return np.concatenate(out)[:n]


def adjoint_bench(Context, meshgen, problems, precision, walks=2048):
    """Differentiable 3D Laplace on the torus R = 1, r = 0.35 (200 x 100 quads,
    40 000 triangles), per-vertex values v ~ N(0, 1) (seed 3), 10 000 probes,
    loss 0.5 sum (u - t)^2 with t from v + 0.1 N(0, 1): forward solve and
    adjoint replay (dLoss/dv) at equal walks per point."""
    V, F, _ = meshgen.torus(1.0, 0.35, 200, 100)
    rng = np.random.default_rng(3)
    v = rng.normal(size=len(V))
    vt = v + 0.1 * rng.normal(size=len(V))
    pts = torus_probes(10000, 3)
    ctx = Context(0)
    ctx.set_mesh(V, F)
    p = Context.params(walks, 1e-3, 3, precision=precision)
    target = ctx.solve_points(pts, problems.laplace_vertex_values(vt), p)["mean"]
    eq = problems.laplace_vertex_values(v)
    ctx.solve_points(pts, eq, p)  # warm-up
    t0 = time.perf_counter()
    fwd = ctx.solve_points(pts, eq, p)
    t_fwd = time.perf_counter() - t0
    dl = np.nan_to_num(fwd["mean"] - target)
    t0 = time.perf_counter()
    out, gv, _ = ctx.adjoint_points(pts, eq, p, dl)
    t_adj = time.perf_counter() - t0
    n_walks = out["summary"]["total_walks"]
    return dict(workload="torus 40000 tris, 10000 probes, %d walks/point, eps 1e-3, seed 3, "
                         "per-vertex Dirichlet values" % walks,
                precision="fp32" if precision else "fp64",
                loss=float(0.5 * np.sum(dl ** 2)),
                forward_ms=t_fwd * 1e3, adjoint_ms=t_adj * 1e3,
                adjoint_over_forward=t_adj / t_fwd,
                walks_per_s=n_walks / out["summary"]["walk_kernel_seconds"],
                walks_per_s_end_to_end=n_walks / t_adj,
                grad_norm=float(np.linalg.norm(gv)))


def workload_stars(Context, meshgen, problems, caps):
    """Batched uneven 2D Poisson, Green-function source sampling: 16 star
    domains r = 1 + a cos(k theta) (a ~ U(0.1, 0.4), k ~ U{3..9}, 512..16384
    segments, meshgen seed 2024), f = exp(-|x|^2/0.1), g = sin 3x cos 2y,
    256^2 grid each, walks per point adaptive from 64 to 16384 until
    SE <= 1e-3, FP32, eps 1e-3, seed 11: every round is one walker launch over
    all domains.  Points reaching SE <= 1e-3 per second, capped points apart."""
    fam = meshgen.star_family(16, seed=2024, n_min=512, n_max=16384)
    ctx = Context(0)
    ctx.set_domains([f["mesh"] for f in fam])
    eq = problems.star_poisson_gaussian(1)
    p = Context.params(64, 1e-3, 11, precision=caps.WOS_FP32_PERSISTENT, target_se=SE_TARGET,
                       max_walks=16384)
    ctx.solve_grid_batch(16, 16, 1, eq, p)  # builds the distance grids
    t0 = time.perf_counter()
    res, tot = ctx.solve_grid_batch(256, 256, 1, eq, p)
    dt = time.perf_counter() - t0
    inside = reached = capped = 0
    per = []
    for f, r in zip(fam, res):
        ins = r["location"] == 1
        ok = ins & (r["standard_error"] <= SE_TARGET)
        cap = ins & ~ok
        inside += int(ins.sum()); reached += int(ok.sum()); capped += int(cap.sum())
        per.append(dict(segments=f["n"], a=round(f["a"], 4), k=f["k"], inside=int(ins.sum()),
                        reached=int(ok.sum()), capped=int(cap.sum()),
                        mean_walks=float(r["walks"][ins].mean())))
    return dict(workload="stars", description=workload_stars.__doc__.split(":")[0].strip(),
                metric="points reaching SE<=1e-3 per s (FP32, batched, adaptive)",
                value=reached / dt, unit="pts/s", seconds=dt, inside=inside, reached=reached,
                capped=capped, rounds=tot["rounds"], total_walks=tot["total_walks"],
                total_steps=tot["total_steps"], steps_per_s=tot["total_steps"] / dt,
                walk_kernel_s=tot["walk_kernel_seconds"], domains=per)


def workload_plate(Context, meshgen, problems, caps):
    """Welding-style stationary screened Poisson: box plate 120 x 60 x 10 mm,
    200 316 triangles, alpha = 0.05, f = Q exp(-|x - x0|^2 / sigma^2) with
    Q = 1, sigma = 3 mm, x0 = (40, 30, 10), uniform ball sampling, g = 0,
    64 x 32 x 16 grid, 8192 walks per point, eps 0.01, seed 5: FP32 and FP64
    grid pts/s at SE <= 1e-3 and the FP32 adjoint of L = sum u (gradients
    with respect to Q, sigma, x0, alpha)."""
    V, F, _ = meshgen.plate_200k()
    ctx = Context(0)
    ctx.set_mesh(V, F)
    eq = problems.plate_gaussian(alpha=0.05, q=1.0, sigma=3.0, x0=(40.0, 30.0, 10.0))
    grid, nw, eps, seed = (64, 32, 16), 8192, 0.01, 5
    p32 = Context.params(nw, eps, seed, precision=caps.WOS_FP32_PERSISTENT)
    p64 = Context.params(nw, eps, seed, precision=caps.WOS_FP64_PLAIN)
    ctx.solve_grid(*grid, eq, p32)
    t0 = time.perf_counter()
    r32 = ctx.solve_grid(*grid, eq, p32)
    t32 = time.perf_counter() - t0
    t0 = time.perf_counter()
    r64 = ctx.solve_grid(*grid, eq, p64)
    t64 = time.perf_counter() - t0
    ins = r64["location"] == 1
    var = r64["variance"][ins]
    v32 = pts_per_s_at_se(ins.sum(), var, r32["mean_steps"][ins], t32, r32["summary"]["total_steps"])
    v64 = pts_per_s_at_se(ins.sum(), var, r64["mean_steps"][ins], t64, r64["summary"]["total_steps"])
    dl = np.where(ins, 1.0, 0.0)
    t0 = time.perf_counter()
    _, _, gp = ctx.adjoint_grid(*grid, eq, p32, dl)
    tadj = time.perf_counter() - t0
    # a floor for points where the source underflows FP32 (f ~ 1e-170 far
    # from x0: both estimates are 0 to 1e-160 and their SEs vanish)
    floor = 1e-9 * np.abs(r64["mean"][ins]).max()
    z = np.abs(r32["mean"][ins] - r64["mean"][ins]) / np.maximum(
        np.sqrt(r32["standard_error"][ins] ** 2 + r64["standard_error"][ins] ** 2), floor)
    return dict(workload="plate", description=workload_plate.__doc__.split(":")[0].strip(),
                tuning=ctx.get_tuning(),
                metric="grid pts/s at SE<=1e-3 (FP32 vs FP64), adjoint time",
                value=v32, unit="grid pts/s", fp64_value=v64, fp32_over_fp64=v32 / v64,
                inside=int(ins.sum()), fp32_ms=t32 * 1e3, fp64_ms=t64 * 1e3,
                fp32_steps_per_s=r32["summary"]["total_steps"] / t32,
                fp32_vs_fp64_max_z=float(z.max()), fp32_vs_fp64_mean_z2=float(np.mean(z ** 2)),
                adjoint_ms=tadj * 1e3, adjoint_over_forward=tadj / t32,
                grad={"Q": gp[caps.WOS_GRAD_SRC_VALUE], "sigma": gp[caps.WOS_GRAD_SRC_RADIUS],
                      "x0": list(gp[caps.WOS_GRAD_SRC_CENTER:caps.WOS_GRAD_SRC_CENTER + 3]),
                      "alpha": gp[caps.WOS_GRAD_ALPHA]})


def workload_welding(Context, caps):
    """Welding transients, one screened WoS solve per implicit-Euler step:
    level 2 (plate 120 x 80 mm, 49 x 33 grid, 500 walks, 100 steps of 0.1 s,
    cubic history, stationary Gaussian for 3 s, seed 20260819) in FP32 and
    FP64, level 3 (97 x 65 grid, 500 walks, 150 steps, source moving for 5 s)
    in FP32; the compiled reference on the host cores times step 21 of level 2
    on a sample of its points (the reference's own seeds), scaled to the run:
    whole-run seconds"""
    from wos_amd import welding as W
    ctx = Context(0)
    out = {}
    for label, prec in (("fp32", caps.WOS_FP32_PERSISTENT), ("fp64", caps.WOS_FP64_PLAIN)):
        W.run_level2(ctx, W.Level2Config(), prec, steps=2)   # warm-up
        t0 = time.perf_counter()
        rec, _ = W.run_level2(ctx, W.Level2Config(), prec)
        out[f"level2_{label}_s"] = time.perf_counter() - t0
        out[f"level2_{label}_ms_per_step"] = 1e3 * out[f"level2_{label}_s"] / len(rec)
        out[f"level2_{label}_peak_rise"] = max(r["peak_rise"] for r in rec)
    t0 = time.perf_counter()
    rec3, _, _ = W.run_level3(ctx, W.Level3Config(), caps.WOS_FP32_PERSISTENT)
    out["level3_fp32_s"] = time.perf_counter() - t0
    out["level3_fp32_ms_per_step"] = 1e3 * out["level3_fp32_s"] / len(rec3)
    out["level3_fp32_peak_rise"] = max(r["peak_rise"] for r in rec3)
    # the reference on the host: step 21 of level 2 from the GPU's step-20 field
    c = W.Level2Config()
    _, prev = W.run_level2(ctx, c, caps.WOS_FP32_PERSISTENT, steps=20)
    cpu = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        from wos_amd import meshgen
        ref = pyoracle.Reference()
        x0, x1, y0, y1 = -0.06, 0.06, -0.04, 0.04
        kappa = c.conductivity / (c.density * c.specific_heat)
        alpha = 1.0 / np.sqrt(kappa * c.dt)
        ix = W.gaussian_integral(x0, x1, 0.0, c.sigma_x)
        iy = W.gaussian_integral(y0, y1, 0.0, c.sigma_y)
        term = (0.0, 0.0, c.sigma_x, c.sigma_y, c.electrical_power * c.efficiency,
                c.thickness * ix * iy)
        frac = W.heating_fraction(2.0, 2.1, c.heat_on_time)
        fs = W.field_source(prev, (c.nx, c.ny, x0, x1, y0, y1), W.WOS_INTERP_CUBIC, True,
                            alpha * alpha, frac, c.conductivity, [term])
        eq = W.field_equation(alpha, fs)
        idx = np.array([i * c.ny + j for i in range(1, c.nx - 1) for j in range(1, c.ny - 1)])
        pick = idx[::4]
        pts = np.array([(W.grid_coordinate(x0, x1, k // c.ny, c.nx),
                         W.grid_coordinate(y0, y1, k % c.ny, c.ny)) for k in pick])
        cores = min(os.cpu_count() or 1, 16)
        rm = ref.mesh(*meshgen.rectangle(x0, x1, y0, y1))
        t0 = time.perf_counter()
        rm.solve_points(eq, pts, c.walks, c.epsilon, seed=W.step_seed(c.seed, 21),
                        n_threads=cores, seed_index=pick)
        dt = time.perf_counter() - t0
        run_s = dt * len(idx) / len(pick) * 100
        cpu = dict(value=run_s, unit="s (level-2 run, estimated)", cores=cores, kind="reference",
                   sample=f"step 21 of level 2, {len(pick)} of {len(idx)} interior points, "
                          f"{c.walks} walks each, {dt:.2f} s on {cores} threads, x 100 steps")
        out["level2_fp32_speedup_vs_cpu"] = run_s / out["level2_fp32_s"]
    except (FileNotFoundError, OSError):
        pass
    return dict(workload="welding", description=workload_welding.__doc__.split(":")[0].strip(),
                metric="whole-run seconds of the level-2 and level-3 transients",
                value=out["level2_fp32_s"], unit="s", higher_is_better=False, **out,
                cpu_baseline=cpu)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--walks", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--fp32-only", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    ap.add_argument("--workload", default="icosphere",
                    choices=("icosphere", "stars", "plate", "welding"),
                    help="icosphere: the headline line (default); stars: 16 batched star "
                         "domains, adaptive; plate: the 200k-triangle plate with gradients; "
                         "welding: the level-2/3 transient time loops")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the single-GPU kernel at full shape is the measured unit).", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    from wos_amd import Context, meshgen, problems
    from wos_amd import capi
    from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN
    if args.workload == "stars":
        print(json.dumps(workload_stars(Context, meshgen, problems, capi)))
        return
    if args.workload == "plate":
        print(json.dumps(workload_plate(Context, meshgen, problems, capi)))
        return
    if args.workload == "welding":
        print(json.dumps(workload_welding(Context, capi)))
        return

    V, F, _ = meshgen.icosphere(4)
    eq = problems.sphere_screened_constant(alpha=10.0, f=1.0)
    grid = (32, 32, 32)
    eps, seed, nw = 1e-4, 7, args.walks
    ctx = Context(0)
    ctx.set_mesh(V, F)

    p32 = Context.params(nw, eps, seed, precision=WOS_FP32_PERSISTENT)
    p64 = Context.params(nw, eps, seed, precision=WOS_FP64_PLAIN)

    for _ in range(args.warmup):
        ctx.solve_grid(*grid, eq, p32)
    walk_k = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r32 = ctx.solve_grid(*grid, eq, p32)
        walk_k.append(r32["summary"]["walk_kernel_seconds"])
    t32 = (time.perf_counter() - t0) / args.steps
    s32 = r32["summary"]
    if args.fp32_only:
        print(json.dumps({"fp32_ms_per_solve": t32 * 1e3, "walk_kernel_ms": np.mean(walk_k) * 1e3,
                          "total_steps": s32["total_steps"], "total_walks": s32["total_walks"],
                          "launch_blocks": s32["launch_blocks"],
                          "launch_threads": s32["launch_threads"]}))
        return
    # the same solve as a captured graph (captured by the first call)
    pg = Context.params(nw, eps, seed, precision=WOS_FP32_PERSISTENT, use_graph=1)
    rg = ctx.solve_grid(*grid, eq, pg)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rg = ctx.solve_grid(*grid, eq, pg)
    tg = (time.perf_counter() - t0) / args.steps
    graph_same = bool(np.array_equal(rg["mean"], r32["mean"], equal_nan=True))
    # query-unit counts of the same solve (an extra, untimed launch)
    pc = Context.params(nw, eps, seed, precision=WOS_FP32_PERSISTENT, count_units=1)
    sc = ctx.solve_grid(*grid, eq, pc)["summary"]
    assert sc["total_steps"] == s32["total_steps"]

    t0 = time.perf_counter()
    r64 = ctx.solve_grid(*grid, eq, p64)
    t64 = time.perf_counter() - t0
    s64 = r64["summary"]

    ins = r64["location"] == 1
    assert np.array_equal(r64["location"], r32["location"])
    var_ref = r64["variance"][ins]
    v32 = pts_per_s_at_se(ins.sum(), var_ref, r32["mean_steps"][ins], t32, s32["total_steps"])
    v64 = pts_per_s_at_se(ins.sum(), var_ref, r64["mean_steps"][ins], t64, s64["total_steps"])

    kern_s = float(np.mean(walk_k))
    launch_bytes = walker_bytes(sc, 3, terminal=False)  # g = 0 on the sphere
    achieved = launch_bytes / kern_s / 1e9
    with open(os.path.join(ROOT, "profiles", "algorithmic_counts.json")) as f:
        ref_counts = json.load(f)["icosphere4_screened_32cubed"]
    ref_achieved = s32["total_steps"] * ref_counts["B_step_f32"] / kern_s / 1e9
    traffic, counters = None, {}
    tpath = os.path.join(ROOT, "profiles", "walk_fp32_icosphere4_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f)
        traffic = prof.get("hbm_bytes_per_launch")
        counters = {k: prof[k] for k in ("valu_insts_per_cu_cycle", "valu_lane_utilisation",
                                         "td_busy", "l2_hit_rate")
                    if k in prof}
    steps = s32["total_steps"]

    z = np.abs(r32["mean"][ins] - r64["mean"][ins]) / np.maximum(
        np.sqrt(r32["standard_error"][ins] ** 2 + r64["standard_error"][ins] ** 2), 1e-300)
    d = np.abs(r32["mean"][ins] - r64["mean"][ins])
    line = {
        "metric": METRIC,
        "value": v32,
        "unit": "grid pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "inside_points": int(ins.sum()), "walks_per_point": nw,
                   "accumulation": "f64 (Welford, terminal projection, boundary value)",
                   "tuning": ctx.get_tuning(), "parallelism": "single GPU"},
        "fp64_plain": {"value": v64, "ms_per_solve": t64 * 1e3,
                       "walk_kernel_ms": s64["walk_kernel_seconds"] * 1e3,
                       "steps_per_s": s64["total_steps"] / t64},
        "fp32_over_fp64": v32 / v64,
        "fp32_graph": {"ms_per_solve": tg * 1e3, "results_equal_eager": graph_same},
        "walks_per_s": s32["total_walks"] / t32,
        "steps_per_s": s32["total_steps"] / t32,
        "fp32_vs_fp64": {"max_abs_diff": float(d.max()), "mean_abs_diff": float(d.mean()),
                         "max_z": float(z.max()), "mean_z2": float(np.mean(z ** 2)),
                         "mean_steps_ratio": float(np.mean(r32["mean_steps"][ins]) /
                                                   np.mean(r64["mean_steps"][ins]))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_walk_fp32<3,1,true,false>",
                     "kernel_ms": kern_s * 1e3,
                     "bytes_per_launch": launch_bytes,
                     "bytes_per_step": launch_bytes / steps,
                     "units_per_step": {k: sc[k] / steps for k in
                                        ("q_queries", "q_grid_bound", "q_list_scans",
                                         "q_list_entries", "q_list_prims", "q_nodes",
                                         "q_leaf_prims")},
                     "reference_layout_bytes_per_step": ref_counts["B_step_f32"],
                     "reference_layout_achieved": ref_achieved,
                     "profiled_counters": counters},
    }
    if not args.no_adjoint:
        line["adjoint"] = adjoint_bench(Context, meshgen, problems, WOS_FP32_PERSISTENT)
    if not args.no_cpu:
        inside_idx = np.nonzero(ins)[0]
        line["cpu_baseline"] = cpu_baseline(V, F, eq, grid, inside_idx, var_ref, eps, seed)
        line["vs_cpu"] = v32 / line["cpu_baseline"]["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
