"""Throughput benchmark of the FV residual + SSP-RK3 loop on one MI355X.

Workload (the large shock problem): 2-D four-quadrant Riemann problem on
[0,1]^2 split at (0.8,0.8), 1025 x 1025 jittered-triangle mesh (seed 1,
1,050,625 vertices), transmissive (outlet) sides, KEPES-TeCNO Rusanov flux,
minmod reconstruction, SSP-RK3, CFL 0.5, FP64.  A step is one SSP-RK3 time
step (dt pass + 3 stages of gradient, edge flux, gather + update).  The metric
is FP64 vertex-stage updates per second = n_vertex * 3 * steps / wall time of
the timed steps, inputs resident in HBM.

Prints one JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def kernel_bytes(nv, nf, nc, tiles):
    """Algorithmic HBM bytes per launch of each kernel: FP64 8 B, int32 4 B,
    every array counted once per launch, halo re-reads not counted."""
    fl, halo = tiles["faces"], tiles["halo"]
    return {
        # CSR kernels
        "dt": 76 * nv + 24 * nf,
        "gradient": 124 * nv + 24 * nc,
        "flux": 112 * nv + 56 * nf,
        "update": 148 * nv + 40 * nf,
        # tiled kernels: W + coords + 1/(6 mcarea) + cell entries (3 per cell,
        # 4 B) + CSR offsets + halo ids, dE write
        "grad_tile": (32 + 16 + 8 + 4 + 64) * nv + 12 * nc + 4 * halo,
        # W, dE, coords, U_old source W^n, dcarea, weight, offsets, W write;
        # local faces (slot pair 4 B + normal 16 B), 16-bit gather entries
        "fu_tile": (32 + 64 + 16 + 32 + 8 + 8 + 4 + 32) * nv + 20 * fl + 4 * nf + 4 * halo,
        # fused two-ring stage kernel: W, coords, 1/(6 mcarea), U_old source
        # W^n, dcarea, weight, gather offsets, W write; gradient cell entries
        # of the owned and first-ring slots (4 B + 4 B offsets per slot);
        # local faces, gather entries, halo ids of both rings
        "stage_fused": (32 + 16 + 8 + 32 + 8 + 8 + 4 + 32) * nv + 4 * tiles.get("grad_cells", 0)
                       + 4 * (nv + halo) + 20 * fl + 4 * nf + 4 * (halo + tiles.get("halo2", 0)),
    }


PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                       "riemann1025_tiled_kernels.md")


def profile_figure(what, kernel):
    """A per-launch figure of `kernel` from the committed rocprofv3 summary:
    the line '<what> <kernel>: <n> B' ('traffic': HBM bytes from FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by the copy-kernel calibration) or
    '<what> <kernel>: <n> wave-instructions' ('valu': SQ_INSTS_VALU per
    dispatch).  None if absent."""
    if not os.path.exists(PROFILE):
        return None
    with open(PROFILE) as f:
        for line in f:
            m = re.match(r"^(\w+) (\w+): (\d+) ", line.strip())
            if m and m.group(1) == what and m.group(2) == kernel:
                return int(m.group(3))
    return None


def measured_traffic(kernel):
    return profile_figure("traffic", kernel)


# FP64 vector issue peak of MI355X in wave64 instructions per second: one per
# CU per clock (256 CUs x 2.4 GHz; 78.6 TFLOP/s = that x 64 lanes x 2 flops).
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 2.4e9


# Whole-step algorithmic bytes per vertex-stage of an SSP-RK3 step at the
# 1025^2 size (3 (256 V + 24 E + 12 T) + (40 V + 24 E) divided by 3 V).
STEP_BYTES_PER_VERTEX_STAGE = 389.2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=334)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1025)
    ap.add_argument("--workload", default="riemann", choices=["riemann", "vortex", "tgv"],
                    help="riemann: the headline 1025^2 case; vortex: periodic isentropic vortex "
                         "(KEPES-TeCNO Roe, CFL 0.8); tgv: Navier-Stokes Taylor-Green vortex "
                         "(KEPEC + viscous, Re 1600); --n sets the mesh")
    ap.add_argument("--tile", type=int, default=0, help="owned vertices per tile (0: 128)")
    ap.add_argument("--path", default="auto", choices=["auto", "csr"])
    ap.add_argument("--fused-gradient", action="store_true",
                    help="gradients inside the stage kernel (two halo rings) instead of a pass "
                         "of their own")
    ap.add_argument("--graph", type=int, default=0, help="steps per captured HIP graph (0: none)")
    ap.add_argument("--block", type=int, default=0, help="threads per tile of the stage kernel")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--cpu-ranks", type=int, default=min(16, os.cpu_count() or 1),
                    help="MPI ranks of the reference's multi-rank CPU baseline (1: one rank only)")
    ap.add_argument("--mc", action="store_true",
                    help="second metric: surrogate control-variate MC against plain MC on the "
                         "uncertain isentropic vortex (prints its own JSON line)")
    ap.add_argument("--mc-fine", type=int, default=None,
                    help="fine mesh (vortex 129, riemann 1025)")
    ap.add_argument("--mc-coarse", type=int, default=None,
                    help="mesh of the surrogate's training samples (vortex 129, riemann 257)")
    ap.add_argument("--mc-time", type=float, default=None, help="t_final (vortex 10, riemann 0.05)")
    ap.add_argument("--mc-train", type=int, default=None, help="training samples (512 / 2048)")
    ap.add_argument("--mc-pilot", type=int, default=None,
                    help="fine samples fitting the control coefficients (64 / 0: fitted on the "
                         "training samples)")
    ap.add_argument("--mc-pair", type=int, default=None, help="paired fine samples (1024 / 64)")
    ap.add_argument("--mc-ref", type=int, default=None,
                    help="plain-MC reference samples on the fine mesh (4096 / 512)")
    ap.add_argument("--mc-fine-batch", type=int, default=None,
                    help="fine samples per solver context (riemann: 32)")
    ap.add_argument("--mc-ref-chunk", default=None, metavar="START:COUNT",
                    help="solve only samples [START, START+COUNT) of the plain-MC reference and "
                         "save their QoIs to --mc-ref-out (a reference longer than one run)")
    ap.add_argument("--mc-ref-out", default=None, help=".npz file of --mc-ref-chunk")
    ap.add_argument("--mc-ref-files", default=None,
                    help="comma-separated --mc-ref-chunk files that together hold samples "
                         "[0, --mc-ref) of the reference, used instead of solving it")
    ap.add_argument("--mlmc", action="store_true",
                    help="multilevel MC of the uncertain vortex against plain MC at equal "
                         "GPU-seconds (prints its own JSON line)")
    ap.add_argument("--mlmc-levels", default="129,257")
    ap.add_argument("--mlmc-samples", default="512,64")
    ap.add_argument("--mlmc-throughput", action="store_true",
                    help="vertex-stage updates/s of the multilevel batch (levels --mlmc-levels with "
                         "--mlmc-samples samples, uncertain vortex) stepped by packed launches "
                         "against per-level launches interleaved on streams (prints its own JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 1 GiB (calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": x.numel() * 8, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    if args.mc:
        mc_bench(args)
        return
    if args.mlmc:
        mlmc_bench(args)
        return
    if args.mlmc_throughput:
        mlmc_throughput(args)
        return

    import numpy as np

    import tensum_amd as T
    from cases import riemann, riemann_mesh

    n = args.n
    common = dict(time_scheme="ssprk3", kernel_path=args.path, tile_vertices=args.tile,
                  fused_gradient=args.fused_gradient, graph_steps=args.graph, block_fu=args.block)
    if args.workload == "riemann":
        mesh = riemann_mesh(n, seed=1)
        prm = T.Params(flux="kepes_tecno_rusanov", reconstruct="minmod", cfl=0.5, final_time=0.8,
                       **common)
        w0 = riemann(mesh.x, mesh.y)
        workload = "2D Riemann, KEPES-TeCNO Rusanov, minmod, SSP-RK3, CFL 0.5"
        mesh_name = f"{n}x{n} jittered triangles, seed 1"
    elif args.workload == "vortex":
        from cases import vortex, vortex_mesh
        mesh = vortex_mesh(n, seed=0)
        prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=1e9,
                       **common)
        w0 = vortex(mesh.x, mesh.y)
        workload = "2D isentropic vortex, KEPES-TeCNO Roe, minmod, SSP-RK3, CFL 0.8"
        mesh_name = f"{n}x{n} jittered triangles on periodic [-5,5]^2, seed 0"
    else:
        from cases import taylor_green, tgv_mesh
        mesh = tgv_mesh(n, seed=2)
        gam, mach = 1.4, 0.1
        prm = T.Params(flux="kepec", reconstruct="first", model="ns", cfl=0.5, final_time=1e9,
                       mu_ref=1.0 / 1600.0, prandtl=0.72, **common)
        w0 = taylor_green(mesh.x, mesh.y, mach=mach, gam=gam)
        workload = "2D Taylor-Green vortex, Navier-Stokes, KEPEC + viscous, Re 1600, SSP-RK3"
        mesh_name = f"{n}x{n} jittered triangles on periodic [0,2pi]^2, seed 2"
    print(f"bench: {args.workload} mesh {n}x{n} ({mesh.n_vertex} vertices), building context", file=sys.stderr,
          flush=True)
    s = T.Solver(mesh, prm)
    s.set_state(w0)
    geo = s.geometry()
    nv, nf, nc = mesh.n_vertex, geo["face_v"].shape[0], mesh.cell_v.size // 3
    # warm-up, per-kernel durations (HIP events on the library's stream, 20
    # steps plus 20 more without events), then the timed steps
    s.run(args.warmup)
    ku = s.bench(20)
    # timed region: K whole steps, synchronised on both sides (tsm_run
    # synchronises the stream before returning)
    t_before, it_before = s.time()
    t0 = time.perf_counter()
    done = s.run(args.steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    assert int(done[0]) == args.steps, f"only {int(done[0])} steps ran (final time reached?)"
    value = nv * 3 * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps
    print(f"bench: {args.steps} steps in {wall:.3f} s, {ms_per_step:.3f} ms/step", file=sys.stderr,
          flush=True)

    tiles = T.host_tiling(mesh, prm.tile_vertices or 128, two_rings=args.fused_gradient)
    kb = kernel_bytes(nv, nf, nc, tiles)
    if ku["dt_us"] == 0.0 and tiles["two_rings"]:  # fused stage kernel
        per_launch = {"stage_fused": ku["flux_us"]}
    elif ku["dt_us"] == 0.0:  # tiled kernels with a gradient pass
        per_launch = {"grad_tile": ku["gradient_us"], "fu_tile": ku["flux_us"]}
    else:
        per_launch = {"dt": ku["dt_us"], "gradient": ku["gradient_us"], "flux": ku["flux_us"],
                      "update": ku["update_us"]}
    per_step = {k: v * (1 if k == "dt" else 3) for k, v in per_launch.items()}
    dom = max(per_step, key=per_step.get)
    dur_us = per_launch[dom]
    achieved = kb[dom] / (dur_us * 1e-6) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": (measured_traffic(dom) if args.workload == "riemann" and n == 1025
                            else None),
                "algorithmic_bytes_per_launch": kb[dom],
                "valu_issue_frac": valu_issue_frac(dom, dur_us, args),
                "kernel_us": {k: round(v, 2) for k, v in ku.items()},
                "step_frac_at_389B_per_vertex_stage":
                    round(value * STEP_BYTES_PER_VERTEX_STAGE / (HBM_PEAK_GBS * 1e9), 4)}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = reference_cpu_baseline(args, mesh, prm, w0) if args.workload == "riemann" else None
        if cpu is None:
            import oracle_binding as O
            print("bench: timing the CPU oracle on a bounded sample", file=sys.stderr, flush=True)
            o = O.Oracle(mesh, prm)
            o.set_state(w0)
            c0 = time.perf_counter()
            k = o.run(args.cpu_steps)
            c1 = time.perf_counter()
            cpu = {"value": nv * 3 * k / (c1 - c0), "unit": "vertex-stage updates/s", "cores": 1,
                   "kind": "port",
                   "sample": f"{k} SSP-RK3 steps of the same {n}x{n} {args.workload} workload, "
                             f"oracle (C++ restatement of the reference), single thread"}

    out = {
        "metric": "FP64 vertex-stage updates/s at 1M verts",
        "value": value,
        "unit": "vertex-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (seeded jittered-triangle mesh, {args.workload} initial state)",
        "config": {"workload": workload, "mesh": mesh_name, "n_vertex": nv, "n_face": nf,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def valu_issue_frac(kernel, dur_us, args):
    """VALU wave-instructions per launch of `kernel` (the committed profile's
    SQ_INSTS_VALU count, 1025^2 Riemann only) over its measured duration,
    against the FP64 issue peak: the fraction of the vector issue rate the
    kernel uses."""
    if args.workload != "riemann" or args.n != 1025:
        return None
    n = profile_figure("valu", kernel)
    return round(n / (dur_us * 1e-6) / VALU_PEAK_WAVE_INSTR_PER_S, 4) if n else None


def _reference_run(mesh, t_k, ranks, d, say):
    """One run of the reference binary on `ranks` MPI ranks in directory d up
    to time t_k: (steps, SAMPLE_SOLVE seconds, wall seconds) or None.  More
    than one rank: the mesh goes through the reference's own partitioner
    (oracle/_ref/grid_part) as a gmsh file split into x-strips."""
    import subprocess

    from cases import write_msh_strips, write_part_file
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from ref_cases import param_text
    exe = os.path.join(ROOT, "oracle", "_ref", "tensum")
    os.makedirs(d, exist_ok=True)
    cmd = [exe, "-i", "param.in", "-npart", str(ranks), "-L", "."]
    if ranks == 1:
        write_part_file(mesh, d)
    else:
        part = os.path.join(ROOT, "oracle", "_ref", "grid_part")
        mpirun = "/opt/conda/bin/mpirun"
        if not (os.access(part, os.X_OK) and os.access(mpirun, os.X_OK)):
            return None
        say(f"partitioning the mesh into {ranks} strips (the reference's grid_part)")
        write_msh_strips(mesh, os.path.join(d, "mesh.msh"), ranks)
        r = subprocess.run([part, "-D", "2", "-I", os.path.join(d, "mesh.msh"), "-P", str(ranks),
                            "-L", d], capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            return None
        os.remove(os.path.join(d, "mesh.msh"))
        cmd = [mpirun, "-np", str(ranks)] + cmd
    text = param_text("riemann33", t_k, 10 ** 8).replace(
        "   restart               yes", "   restart               no")
    with open(os.path.join(d, "param.in"), "w") as f:
        f.write(text)
    c0 = time.perf_counter()
    r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=900)
    wall = time.perf_counter() - c0
    sd = os.path.join(d, "SAMPLE_0")
    try:
        steps = len(open(os.path.join(sd, "residual.dat")).read().split("\n")) - 1
        solve = None
        for line in open(os.path.join(sd, "time_log.log")):
            if line.startswith("SAMPLE_SOLVE"):
                solve = 60.0 * float(line.split()[2])  # T_MAX over the ranks, minutes
    except OSError:
        return None
    if r.returncode != 0 or not solve or steps < 1:
        return None
    return steps, solve, wall


def reference_cpu_baseline(args, mesh, prm, w0):
    """The reference itself (oracle/_ref/tensum, built from its sources) on the
    first steps of the same Riemann workload, on the host's cores: the mesh
    partitioned into cpu_ranks x-strips by the reference's own partitioner and
    run under MPICH's mpirun, and on one rank for comparison.  The parameter
    file is the reference's own grammar with final_time at the time of the
    sample's last step (from this library's dt, which is the reference's);
    rate = V x 3 x steps / SAMPLE_SOLVE (the reference's own loop timer, the
    slowest rank's).  None when the binary is absent or fails."""
    import tempfile

    import tensum_amd as T
    say = lambda m: print("bench: " + m, file=sys.stderr, flush=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "tensum")
    if not os.access(exe, os.X_OK):
        return None
    k1 = max(1, args.cpu_steps)
    ranks = max(1, args.cpu_ranks)
    kn = k1 * max(1, ranks // 2)  # a sample of similar length on many ranks
    g = T.Solver(mesh, prm)
    g.set_state(w0)
    _, log = g.run(max(k1, kn), want_log=True)
    g.close()
    rate = lambda st, sv: mesh.n_vertex * 3 * st / sv
    with tempfile.TemporaryDirectory() as d:
        say(f"timing the reference binary on one rank, {k1} steps")
        one = _reference_run(mesh, float(log[k1 - 1, 2]), 1, os.path.join(d, "one"), say)
        many = None
        if ranks > 1:
            say(f"timing the reference binary on {ranks} ranks, {kn} steps")
            many = _reference_run(mesh, float(log[kn - 1, 2]), ranks, os.path.join(d, "many"), say)
    if one is None:
        say("reference run failed; timing the oracle instead")
        return None
    s1, v1, w1 = one
    single = {"value": rate(s1, v1), "cores": 1,
              "sample": f"{s1} steps, SAMPLE_SOLVE {v1:.1f} s ({w1:.1f} s with start-up)"}
    if many is None:
        return {"value": single["value"], "unit": "vertex-stage updates/s", "cores": 1,
                "kind": "reference",
                "sample": f"{s1} SSP-RK3 steps from t = 0 of the same {args.n}x{args.n} Riemann "
                          f"workload, the reference built from its sources (one MPI rank), timed by "
                          f"its SAMPLE_SOLVE loop timer ({v1:.1f} s; {w1:.1f} s with start-up "
                          f"and preprocessing)"}
    sn, vn, wn = many
    return {"value": rate(sn, vn), "unit": "vertex-stage updates/s", "cores": ranks,
            "kind": "reference",
            "sample": f"{sn} SSP-RK3 steps from t = 0 of the same {args.n}x{args.n} Riemann workload, "
                      f"the reference built from its sources on {ranks} MPI ranks (MPICH, one per "
                      f"core; the mesh in {ranks} x-strips through the reference's own partitioner), "
                      f"timed by its SAMPLE_SOLVE loop timer, slowest rank ({vn:.1f} s; {wn:.1f} s "
                      f"with start-up and preprocessing)",
            "one_rank": single}


MC_DEFAULTS = {  # fine, coarse, t_final, train, pilot, pair, ref, fine batch
    "vortex": dict(fine=129, coarse=129, time=10.0, train=512, pilot=64, pair=1024, ref=4096,
                   fine_batch=None),
    "riemann": dict(fine=1025, coarse=257, time=0.05, train=2048, pilot=0, pair=64, ref=1024,
                    fine_batch=32),
}


def mc_bench(args):
    """Surrogate control variate vs plain MC.  vortex: uncertain strength,
    centre, gamma and Mach; riemann: the four-quadrant Riemann problem with
    uncertain split (U(0.75,0.85)^2) and quadrant pressures (x U(0.9,1.1)),
    surrogate trained on 257^2 solves and paired with 1025^2 solves.  QoIs: 64
    probe densities, kinetic energy, entropy integral.  value = GPU-seconds
    plain MC needs for the control variate's standard error divided by the
    control variate's GPU-seconds (training samples, fit, fine samples and
    surrogate evaluations all charged)."""
    from tensum_amd import uq
    d = dict(MC_DEFAULTS[args.workload if args.workload in MC_DEFAULTS else "vortex"])
    for k, a in (("fine", args.mc_fine), ("coarse", args.mc_coarse), ("time", args.mc_time),
                 ("train", args.mc_train), ("pilot", args.mc_pilot), ("pair", args.mc_pair),
                 ("ref", args.mc_ref), ("fine_batch", args.mc_fine_batch)):
        if a is not None:
            d[k] = a
    riemann = args.workload == "riemann"
    # the Riemann study runs first-order reconstruction: with perturbed quadrant
    # pressures the second-order scheme leaves the physical states on part of
    # the samples within a few hundred steps (see uq.riemann_problem)
    p = (uq.riemann_problem(final_time=d["time"], reconstruct="first") if riemann
         else uq.vortex_problem(final_time=d["time"]))
    say = lambda m: print("bench: " + m, file=sys.stderr, flush=True)
    if args.mc_ref_chunk:
        import numpy as np
        a, b = (int(x) for x in args.mc_ref_chunk.split(":"))
        Q, sec = uq.reference_chunk(p, d["fine"], a, b, log=say, fine_batch=d["fine_batch"])
        np.savez(args.mc_ref_out, start=a, count=b, seconds=sec, qoi=Q)
        print(json.dumps({"reference_chunk": [a, b], "seconds": sec, "file": args.mc_ref_out}),
              flush=True)
        return
    ref = None
    if args.mc_ref_files:
        import numpy as np
        parts = sorted((np.load(f) for f in args.mc_ref_files.split(",")), key=lambda z: int(z["start"]))
        nxt = 0
        for z in parts:
            if int(z["start"]) != nxt:
                raise SystemExit(f"reference chunks leave a gap at sample {nxt}")
            nxt += int(z["count"])
        if nxt != d["ref"]:
            raise SystemExit(f"reference chunks hold {nxt} samples, --mc-ref is {d['ref']}")
        ref = (np.vstack([z["qoi"] for z in parts]), sum(float(z["seconds"]) for z in parts))
        say(f"reference: {nxt} samples from {len(parts)} files")
    r = uq.cv_experiment(p, d["fine"], d["coarse"], d["train"], d["pilot"], d["pair"], 65536,
                         d["ref"], log=say, fine_batch=d["fine_batch"], ref=ref)
    nin = len(p.bounds)
    out = {
        "metric": "MC moment error per GPU-second (control variate vs plain MC, equal error)",
        "value": r["speedup_at_equal_error"],
        "unit": "x fewer GPU-seconds than plain batched MC",
        "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 solver, bf16/fp32 surrogate",
        "data": "synthetic (seeded jittered meshes, scrambled Sobol inputs)",
        "config": {"workload": ("2D Riemann UQ (split ~U(0.75,0.85)^2, quadrant pressures "
                                "x U(0.9,1.1)), KEPES-TeCNO Rusanov, first order, CFL 0.5"
                                if riemann else "isentropic vortex UQ") + f", t_final {d['time']}",
                   "fine_mesh": d["fine"], "coarse_mesh": d["coarse"], "train_samples": d["train"],
                   "paired_samples": d["pair"], "pilot_samples": d["pilot"],
                   "surrogate_evaluations": 65536, "reference_samples": d["ref"],
                   "surrogate": f"MLP {nin}-256-256-256-66, GELU, BF16 MFMA, FP32 accumulation, Adam"},
        "detail": r,
    }
    print(json.dumps(out), flush=True)


def mlmc_bench(args):
    from tensum_amd import uq
    p = uq.vortex_problem(final_time=args.mc_time if args.mc_time is not None else 10.0)
    levels = tuple(int(x) for x in args.mlmc_levels.split(","))
    samples = tuple(int(x) for x in args.mlmc_samples.split(","))
    say = lambda m: print("bench: " + m, file=sys.stderr, flush=True)
    r = uq.mlmc_experiment(p, levels, samples, args.mc_ref if args.mc_ref is not None else 4096,
                           log=say, batch=args.mc_fine_batch)
    out = {"metric": "MLMC moment error at the GPU-seconds of plain MC (uncertain vortex)",
           "value": r["plain_mc"]["mean_error"] / max(r["mlmc"]["mean_error"], 1e-300),
           "unit": "plain-MC mean error / MLMC mean error at equal GPU-seconds",
           "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (seeded jittered meshes, scrambled Sobol inputs)",
           "config": {"workload": f"isentropic vortex UQ, t_final {p.params.final_time}",
                      "levels": levels, "samples": samples},
           "detail": r}
    print(json.dumps(out), flush=True)


def mlmc_throughput(args):
    """Throughput of one multilevel batch: every level's samples in one
    context, advanced by packed launches spanning all levels (tsm_run_packed)
    and, for comparison, by each level's own launches interleaved on per-level
    streams (tsm_run_group).  Both runs start from the same states; the packed
    states are checked equal to the interleaved ones bit for bit."""
    import numpy as np

    import tensum_amd as T
    from tensum_amd import uq
    p = uq.vortex_problem(final_time=1e9)
    levels = tuple(int(x) for x in args.mlmc_levels.split(","))
    samples = tuple(int(x) for x in args.mlmc_samples.split(","))
    say = lambda m: print("bench: " + m, file=sys.stderr, flush=True)
    sets = {}
    for mode in ("packed", "group"):
        sets[mode] = []
        for n, k in zip(levels, samples):
            inp = uq.sample_inputs(k, p.bounds, seed=n)
            sets[mode].append(uq.make_batch(p, p.mesh(n), inp))
        say(f"{mode}: {len(levels)} levels built")
    vs = sum(s.nv * s.n_samples for s in sets["packed"])
    res = {}
    for mode, fn in (("packed", T.run_packed), ("group", T.run_group)):
        fn(sets[mode], args.warmup)
        t0 = time.perf_counter()
        done = fn(sets[mode], args.steps)
        wall = time.perf_counter() - t0
        assert all(int(d) == args.steps * s.n_samples for d, s in zip(done, sets[mode]))
        res[mode] = {"vertex_stage_per_s": vs * 3 * args.steps / wall, "ms_per_step": 1e3 * wall / args.steps}
        say(f"{mode}: {res[mode]['ms_per_step']:.2f} ms per step of all levels")
    same = all(np.array_equal(a.get_state(k), b.get_state(k)) for a, b in zip(sets["packed"], sets["group"])
               for k in range(a.n_samples))
    out = {"metric": "FP64 vertex-stage updates/s of a multilevel batch (packed launches)",
           "value": res["packed"]["vertex_stage_per_s"], "unit": "vertex-stage updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded jittered meshes, scrambled Sobol inputs)",
           "config": {"workload": "uncertain isentropic vortex, KEPES-TeCNO Roe, minmod, SSP-RK3, CFL 0.8",
                      "levels": levels, "samples": samples, "vertex_samples": vs},
           "per_level_streams": res["group"], "packed": res["packed"],
           "packed_over_streams": res["packed"]["vertex_stage_per_s"] / res["group"]["vertex_stage_per_s"],
           "states_identical": bool(same)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
