"""Throughput of the FRiBoIS2D fluid step on one MI355X.

One "step" is one call of the drop-in fluid_tsch (CFL/dt control + 3 RK
substeps of HDNL/INVTR, MLS forcing, ADI Helmholtz, divergence, eigen-transform
Poisson solve and projection) on the 2048x1536-cell rectangle workload
(chord 1, thickness 0.25, Re=200, 1000 Lagrangian markers pitching about the
centre, CFL 0.3), fields resident in HBM.  Marker kinematics for each step are
produced on the host (as rigid() does) and passed through the C ABI inside the
timed region.

Prints one JSON line: Mcell-steps/s, the roofline of the dominant kernel
(the FP64 MFMA eigen-transform GEMM) measured with HIP events on the library's
stream over the timed steps, and the CPU oracle timed on one step of the same
workload on the host.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA_TFLOPS = 78.6
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def pmc_traffic(name, mirror=None):
    """HBM bytes per dispatch of `name` from the committed rocprofv3 --pmc
    summary (for the transform GEMM only when the summary's run took the same
    form of the transform, mirror or dense)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        e = json.load(open(path)).get(name, {})
        if mirror is not None and e.get("mirror_form", False) != mirror:
            return None
        return e.get("bytes_per_dispatch")
    except Exception:
        return None


def run_gemm_reference(reps=50):
    """The library-independent yardstick of the transform GEMM: torch's FP64
    matmul (hipBLASLt / rocBLAS) on the y transform of the 2048x1536 grid,
    C (1536 x 2048) = A (1536 x 1536) * B (1536 x 2048) in the column-major
    layout k_gemm_dma reads (A modes-contiguous, B k-contiguous), i.e. the
    row-major product B^T (2048 x 1536) @ A^T (1536 x 1536)."""
    import torch
    M, N, K = 1536, 2048, 1536
    g = torch.Generator(device="cuda").manual_seed(5)
    bt = torch.randn(N, K, dtype=torch.float64, device="cuda", generator=g)
    at = torch.randn(K, M, dtype=torch.float64, device="cuda", generator=g)
    for _ in range(10):
        ct = bt @ at
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ct = bt @ at
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    tf = 2.0 * M * N * K / (us * 1e-6) / 1e12
    print(json.dumps({"kernel": "torch.matmul float64 (vendor BLAS)", "M": M, "N": N, "K": K,
                      "us_per_gemm": round(us, 2), "TFLOP/s": round(tf, 2),
                      "frac_of_fp64_mfma_peak": round(tf / PEAK_FP64_MFMA_TFLOPS, 4), "checksum": float(ct[0, 0])}))


def ensemble_members(n, seed=20261202):
    """The batched ensemble: n cylinders, n3m uniform in 256..1024 (multiples
    of 8), n2m = 3/4 n3m (multiple of 8), Re uniform in [60, 200], reduced
    velocity Ur uniform in [3, 10] giving a heave of 0.1 D at frequency 1/Ur."""
    import numpy as np
    from fribois2d_amd import cases
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        n3m = 8 * int(rng.integers(32, 129))
        n2m = 8 * int(round(0.75 * n3m / 8))
        re = float(rng.uniform(60.0, 200.0))
        ur = float(rng.uniform(3.0, 10.0))
        out.append(cases.ensemble_case(i, n3m, n2m, re, amp=0.1, freq=1.0 / ur))
    return out


class BodyMarkers:
    """Marker kinematics of a body in prescribed rigid motion (Case.motion)
    at time t, as fribo_set_markers takes them, from one small matrix
    product: the 4 x 3 rigid map (positions, velocities) times the rows
    [1, ry, rz] of the rest positions relative to the centre.  The same
    motion as Case.markers, whose ten array operations cost the host about
    27 us per step against about 5 us here (time in which the GPU, past the
    next step's CFL kernel, waits); rounded in another order."""

    def __init__(self, c):
        import numpy as np
        c.markers(0.0)
        _, lmA, ry, rz = c._rest
        self.c = c
        self.R = np.stack([np.ones_like(ry), ry, rz])
        self.T = np.zeros((4, 3))
        self.buf = np.empty((4, ry.size))
        self.lmA = lmA.copy()

    def at(self, t):
        import numpy as np
        if self.c.motion is None:
            return self.c.markers(t)
        dy, dz, th, vy, vz, om = self.c.motion(t)
        yc, zc = self.c.center
        cs, sn = math.cos(th), math.sin(th)
        T = self.T
        T[0] = (yc + dy, cs, -sn)          # y  = yc + dy + c ry - s rz
        T[1] = (zc + dz, sn, cs)           # z  = zc + dz + s ry + c rz
        T[2] = (vy, -om * sn, -om * cs)    # vy - om (s ry + c rz)
        T[3] = (vz, om * cs, -om * sn)     # vz + om (c ry - s rz)
        np.dot(T, self.R, out=self.buf)
        return self.buf[0:2].T, self.buf[2:4].T, self.lmA


class EnsembleMarkers:
    """Marker kinematics of every heaving member at time t[i], laid out as
    fribo_batch_set_markers takes them, with the arithmetic of
    Case.markers (so each member sees exactly its single-instance input)."""

    def __init__(self, cs):
        import numpy as np
        self.cs = cs
        rest = [c.markers(0.0) for c in cs]
        self.nl = np.array([r[2].size for r in rest])
        self.off = np.concatenate([[0], np.cumsum(5 * self.nl)])
        self.buf = np.zeros(int(self.off[-1]))
        self.iy, self.iz, self.ivy, self.ivz = [], [], [], []
        ry, rz, yc, zc, inst = [], [], [], [], []
        for i, (c, r) in enumerate(zip(cs, rest)):
            o, n = int(self.off[i]), int(self.nl[i])
            self.buf[o + 4 * n:o + 5 * n] = r[2]
            self.iy.append(np.arange(o, o + n))
            self.iz.append(np.arange(o + n, o + 2 * n))
            self.ivy.append(np.arange(o + 2 * n, o + 3 * n))
            self.ivz.append(np.arange(o + 3 * n, o + 4 * n))
            l0 = c._rest
            ry.append(l0[2])
            rz.append(l0[3])
            yc.append(np.full(n, c.center[0]))
            zc.append(np.full(n, c.center[1]))
            inst.append(np.full(n, i))
        cat = np.concatenate
        self.iy, self.iz, self.ivy, self.ivz = cat(self.iy), cat(self.iz), cat(self.ivy), cat(self.ivz)
        self.ry, self.rz, self.yc, self.zc, self.inst = cat(ry), cat(rz), cat(yc), cat(zc), cat(inst)
        self.buf[self.iz] = (self.zc + 0.0) + self.rz
        self.buf[self.ivz] = 0.0
        self.amp = np.array([c.motion.amp for c in cs])
        self.w = np.array([c.motion.w for c in cs])

    def at(self, t):
        import numpy as np
        wt = self.w * t
        dy = self.amp * np.sin(wt)
        vy = self.amp * self.w * np.cos(wt)
        self.buf[self.iy] = (self.yc + dy[self.inst]) + self.ry
        self.buf[self.ivy] = vy[self.inst]
        return self.buf


def run_ensemble(args):
    """Batched ensemble of uneven cylinders: one launch per stage for the
    whole batch, against the same instances stepped one at a time."""
    import numpy as np
    from fribois2d_amd import cases
    from fribois2d_amd.gpu import FluidBatch, FluidStep
    n = int(args.workload[len("ensemble"):] or 256)
    cs = ensemble_members(n)
    t0 = time.perf_counter()
    gs = []
    for c in cs:
        g = FluidStep(c.params, c.rc, c.zz)
        g.set_state(cases.initial_state(c.params, c.seed))
        g.set_markers(*c.markers(0.0))
        gs.append(g)
    t_init = time.perf_counter() - t0
    cells = np.array([(c.params.n2 - 1) * (c.params.n3 - 1) for c in cs], dtype=np.float64)
    mirror = [g.ymirror for g in gs]
    mk = EnsembleMarkers(cs)
    b = FluidBatch(gs)
    best, tune_ms, tune_total = b.autotune(5)
    b.set_graph(args.launch == "graph")
    t = np.zeros(n)
    dt = np.array([c.params.dt0 for c in cs])
    ns = 0

    def bstep():
        nonlocal t, dt, ns
        b.set_markers_flat(mk.at(t))
        dtn, _, err = b.step(ns, dt, dt)
        if err.any():
            raise RuntimeError(f"instance errors {np.nonzero(err)[0][:8]} at step {ns}")
        t = t + dtn
        dt = dtn
        ns += 1

    for _ in range(args.warmup):
        bstep()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        bstep()
    el = time.perf_counter() - t1
    value = float(cells.sum()) * args.steps / el / 1e6
    b.close()
    # the same instances, one fribo_fluid_tsch at a time (their own streams again)
    nst = max(1, min(3, args.steps))
    t2 = time.perf_counter()
    for i, (c, g) in enumerate(zip(cs, gs)):
        ti, di = float(t[i]), float(dt[i])
        for k in range(nst):
            g.set_markers(*c.markers(ti))
            dn, _ = g.fluid_tsch(ns + k, di, di)
            ti += dn
            di = dn
    g.sync()
    el1 = time.perf_counter() - t2
    single = float(cells.sum()) * nst / el1 / 1e6
    for g in gs:
        g.close()
    # transform GEMM of the batch (forward, the autotuned shape): flops of
    # every member's 2*n3m*n2m^2 (along y) per grouped launch, half that for
    # the members in the mirror form (two n2m/2 blocks)
    flops = float(sum((1.0 if m else 2.0) * (c.params.n3 - 1) * (c.params.n2 - 1) ** 2
                      for c, m in zip(cs, mirror)))
    gemm_ms = float(tune_ms[best])
    achieved = flops / (gemm_ms * 1e-3) / 1e12
    shapes = {0: "32x32", 1: "64x64", 2: "96x128", 3: "direct-to-LDS 96x128", 4: "direct-to-LDS 96x64"}
    out = {
        "metric": "Mcell-steps/s (batched ensemble, plain FP64)",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated grids and markers, seeded 1e-3 velocity perturbation per member)",
        "config": {"workload": f"ensemble{n}: {n} heaving cylinders, grids 256x192..1024x768 (uneven, "
                               f"{int(cells.sum())} cells in all), Re 60-200, Ur 3-10, CFL 0.3",
                   "precision": "plain FP64", "launch": f"batched, one launch per stage ({args.launch})"},
        "unbatched": {"value": round(single, 3), "unit": "Mcell-steps/s",
                      "how": f"the same {n} instances, {nst} fribo_fluid_tsch calls each, one after another"},
        "batched_over_unbatched": round(value / single, 2),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_MFMA_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_MFMA_TFLOPS, 4), "traffic": None,
                     "kernel": "kb_gemm (grouped forward transform)", "avg_ms": round(gemm_ms, 4),
                     "members_in_mirror_form": int(sum(mirror))},
        "autotune": {"shape": shapes.get(best, best),
                     "ms_per_launch": {shapes[k]: round(float(v), 4) for k, v in enumerate(tune_ms) if v > 0},
                     "ms_with_unfused_chunk_pass": {shapes[k]: round(float(v), 4)
                                                    for k, v in enumerate(tune_total) if v > 0}},
        "init_s": round(t_init, 1),
    }
    print(json.dumps(out))


def run_coupled(args):
    """Strongly coupled cylinder at m* = 1 (512x384, Re 150, Ur 4-8, fixed
    dt, coupling tolerance 1e-8): the learned first iterate against the
    plain one (the previous step's loads) and linear extrapolation."""
    import numpy as np
    from fribois2d_amd import cases, predictor as P
    from fribois2d_amd.gpu import FluidStep
    n3m, n2m = 512, 384
    ntrain, nsteps, neval = args.train_runs, args.coupled_steps, 6
    c, p, _ = P.coupled_case(0, n3m, n2m)
    g = FluidStep(p, c.rc, c.zz)
    t0 = time.perf_counter()
    X, T, its0 = [], [], []
    gen_stats = {"skipped_runs": 0}
    for run in range(ntrain):
        c, p, body = P.coupled_case(run, n3m, n2m)
        g.set_state(cases.initial_state(p, c.seed))
        g.set_markers(*body.markers())
        st = {}
        hist, its = P.run_coupled(g, body, nsteps, p.dt0, stats=st)
        gen_stats["failed"] = gen_stats.get("failed", 0) + st.get("failed", 0)
        if st.get("failed"):
            gen_stats["skipped_runs"] += 1
            continue   # a step stopped at a support switch: not a clean training history
        x, t = P.samples(hist)
        X.append(x)
        T.append(t)
        its0.append(its)
        if (run + 1) % 20 == 0:
            print(f"training run {run + 1}/{ntrain}", file=sys.stderr, flush=True)
    t_gen = time.perf_counter() - t0
    X, T = np.concatenate(X), np.concatenate(T)
    t1 = time.perf_counter()
    pred = P.Predictor.train(X, T, epochs=args.epochs)
    t_train = time.perf_counter() - t1
    res = {}
    hists = {}
    eval_runs = list(range(10000, 10000 + neval))
    for name, guess in (("plain", None), ("extrapolated", P.extrapolate), ("learned", pred)):
        its_all, hs = [], []
        el = 0.0
        st = {}
        for run in eval_runs:
            c, p, body = P.coupled_case(run, n3m, n2m)
            g.set_state(cases.initial_state(p, c.seed))
            g.set_markers(*body.markers())
            g.sync()
            t2 = time.perf_counter()
            h, its = P.run_coupled(g, body, nsteps, p.dt0, guess=guess, stats=st)
            el += time.perf_counter() - t2
            its_all.append(its[P.HIST:])
            hs.append(np.array(h))
        hists[name] = hs
        res[name] = {"subiterations_per_step": round(float(np.mean(np.concatenate(its_all))), 3),
                     "Mcell_steps_per_s": round(n3m * n2m * nsteps * neval / el / 1e6, 3),
                     "steps_stopped_at_a_support_switch": st.get("failed", 0)}
    diff = max(float(np.max(np.abs(a - b) / np.max(np.abs(b), axis=0)))
               for a, b in zip(hists["learned"], hists["plain"]))
    saved = 1.0 - res["learned"]["subiterations_per_step"] / res["plain"]["subiterations_per_step"]
    out = {
        "metric": "Mcell-steps/s (strongly coupled, learned first iterate)",
        "value": res["learned"]["Mcell_steps_per_s"],
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": nsteps * neval,
        "warmup": 0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 fluid, fp32 training / bf16 inference of the predictor",
        "data": f"synthetic: {ntrain} training runs x {nsteps} steps generated by the solver, {neval} held-out runs",
        "config": {"workload": f"coupled{n3m}: cylinder on a transverse spring, m*=1, Re=150, Ur 4-8, "
                               f"{n3m}x{n2m} cells, fixed dt {p.dt0}, coupling tolerance 1e-8",
                   "network": f"{P.NI} -> 64 tanh -> 64 tanh -> {P.NO}"},
        "first_iterate": res,
        "saved_subiterations": round(saved, 4),
        "learned_vs_plain_history_max_rel_diff": diff,
        "training": {"samples": int(X.shape[0]), "epochs": args.epochs, "final_loss": pred.loss,
                     "steps_stopped_at_a_support_switch": gen_stats.get("failed", 0),
                     "runs_left_out": gen_stats["skipped_runs"],
                     "generation_s": round(t_gen, 1), "training_s": round(t_train, 2)},
        "plain_subiterations_training_runs": round(float(np.mean(np.concatenate([i[P.HIST:] for i in its0]))), 3),
    }
    g.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=30,
                    help="untimed steps first (the engine clock takes some tens of ms to settle)")
    ap.add_argument("--workload", default="rect2048")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--train-runs", type=int, default=200, help="coupled512: solver runs the predictor learns from")
    ap.add_argument("--coupled-steps", type=int, default=40, help="coupled512: steps per run")
    ap.add_argument("--epochs", type=int, default=300, help="coupled512: training epochs")
    ap.add_argument("--transform", default="auto", choices=["auto", "y", "z"],
                    help="axis of the Poisson eigen-transform (auto: the shorter one)")
    ap.add_argument("--dense-y", action="store_true",
                    help="transform along y in the dense form even where the grid has the mirror form")
    ap.add_argument("--precision", type=int, default=-1,
                    help="Poisson transforms: -1 FP64; n >= 0 FP32 MFMA + n FP64-residual refinement sweeps")
    ap.add_argument("--launch", default="stream", choices=["stream", "graph"],
                    help="kernel launches of a step: one by one on the stream, or one captured HIP graph "
                         "(steps with GEMM timing events always launch on the stream)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the library's plain 256 MiB copy kernel 3+1 times (HBM counter scale)")
    ap.add_argument("--gemm-reference", action="store_true",
                    help="only time the vendor FP64 GEMM (torch.matmul: hipBLASLt/rocBLAS) on the y transform "
                         "of the 2048x1536 grid, the yardstick of k_gemm_dma")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        from fribois2d_amd.gpu import plain_copy
        nbytes = 256 << 20
        ms = plain_copy(nbytes, 3)
        print(json.dumps({"kernel": "k_plain_copy", "bytes_read": nbytes, "bytes_written": nbytes,
                          "dispatches": 4, "ms_per_copy": round(ms, 4),
                          "GB/s": round(2 * nbytes / ms / 1e6, 1)}))
        return

    if args.gemm_reference:
        run_gemm_reference()
        return

    if args.workload.startswith("ensemble"):
        run_ensemble(args)
        return
    if args.workload == "coupled512":
        run_coupled(args)
        return

    import numpy as np
    from fribois2d_amd import cases
    from fribois2d_amd.gpu import FluidStep

    c = cases.make_case(args.workload)
    p = c.params
    n2m, n3m = p.n2 - 1, p.n3 - 1
    cells = n2m * n3m
    zmm, zmmt, wr = cases.phini(c.zz, p.alx3)
    st0 = cases.initial_state(p, c.seed)
    g = FluidStep(p, c.rc, c.zz, zmm, zmmt, wr)
    g.set_transform(args.transform)
    if args.dense_y and g.ymirror:
        g.set_ymirror(False)
    if args.precision >= 0:
        g.set_precision(args.precision)
    g.set_graph(args.launch == "graph")
    g.set_state(st0)

    t, dt = 0.0, p.dt0
    ns = 0
    body = BodyMarkers(c)

    def step():
        nonlocal t, dt, ns
        g.set_markers(*body.at(t))
        dtn, _ = g.fluid_tsch(ns, dt, dt)
        t += dtn
        dt = dtn
        ns += 1

    for _ in range(args.warmup):
        step()
    g.sync()
    # timed region: the transform GEMMs (the roofline kernel) of every tenth
    # step record their start and end in HIP events on the library's stream;
    # each timed launch costs the stream about 10 us of event handling, so the
    # other steps run with no events, back to back as in production
    g.reset_profile()
    t0 = time.perf_counter()
    for i in range(args.steps):
        g.set_profiling(i % 10 == 9 or (args.steps < 10 and i == args.steps - 1), gemm_only=True)
        step()
    g.sync()
    el = time.perf_counter() - t0
    g.set_profiling(False)
    prof = g.profile()
    ms_step = el * 1e3 / args.steps
    value = cells * args.steps / el / 1e6
    # per-kernel-class breakdown from a separate, untimed pass with every
    # launch bracketed
    g.set_profiling(True)
    g.reset_profile()
    nb = max(1, min(args.steps, 10))
    for _ in range(nb):
        step()
    g.sync()
    prof_all = g.profile()
    g.set_profiling(False)

    gemm_ms, gemm_n = prof["gemm"]
    gemm_avg_s = gemm_ms / gemm_n / 1e3
    # flops of one transform GEMM: along y (Yinv*dph, Y*V) 2*n3m*n2m^2, along
    # z (dph*zmmt^T, V*zmm^T, the reference's order) 2*n2m*n3m^2
    axis = g.transform
    # the mirror form (y grids symmetric about the centre) multiplies by two
    # n2m/2 x n2m/2 blocks: n3m*n2m^2
    mirror = axis == "y" and g.ymirror and args.precision < 0
    gemm_flops = (1.0 if mirror else 2.0) * n3m * n2m * n2m if axis == "y" else 2.0 * n2m * n3m * n3m
    achieved = gemm_flops / gemm_avg_s / 1e12
    # the transforms of this workload take the direct-to-LDS kernel (at least
    # 192 tiles of 96x128 in the library's size test, operands with padded
    # leading dimensions)
    gemm_kernel = "k_gemm_dma" if -(-n2m // 96) * -(-n3m // 128) >= 192 else "k_gemm_nt"
    peak = PEAK_FP64_MFMA_TFLOPS
    if args.precision >= 0:
        gemm_kernel, peak = "k_gemm32", PEAK_FP32_MFMA_TFLOPS
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": pmc_traffic(gemm_kernel, mirror),
            "kernel": gemm_kernel, "avg_ms": round(gemm_avg_s * 1e3, 4), "transform_axis": axis,
            "mirror_form": mirror, "flops_per_launch": gemm_flops}
    kernels = {k: {"ms_per_step": round(v[0] / nb, 4), "launches_per_step": v[1] / nb}
               for k, v in prof_all.items() if v[1]}
    # HBM fraction of the memory-bound classes: compulsory FP64 field values
    # per cell per launch (reads + writes), at the 8 TB/s MI355X peak
    # (the per-mode Poisson lines: along z one read and one write of hmn; along
    # y the two chunk passes read it twice and write it once, unless the first
    # pass runs in the forward transform's epilogue: equal chunks of at most
    # 33 rows dividing its 64-column tiles, which the direct-to-LDS kernel
    # takes at >= 192 tiles of 96x128)
    nch = max(1, min(n3m // 2, 64))
    csz = n3m // nch
    fused = (axis == "y" and n3m % nch == 0 and 1 < csz <= 33 and 64 % csz == 0 and n3m % 64 == 0
             and -(-n2m // 96) * -(-n3m // 128) >= 192)
    values = {"rhs": 9, "jsolve_helm": 4, "ksolve": 6, "div": 3,
              "jsolve_poisson": (2 if fused else 3) if axis == "y" else 2,
              "upd": 7, "cfl": 2}
    for k, nv in values.items():
        if k in kernels and kernels[k]["ms_per_step"] > 0:
            byts = nv * 8.0 * cells * kernels[k]["launches_per_step"]
            gbs = byts / (kernels[k]["ms_per_step"] * 1e-3) / 1e9
            kernels[k]["alg_GBs"] = round(gbs, 1)
            kernels[k]["hbm_frac"] = round(gbs / PEAK_HBM_GBS, 4)

    # the drop-in fluid_tsch body per step (markers, step, checkbcs, loads,
    # ttime's diagnostics on the device), and the same with q2, q3, pr, ru2,
    # ru3 copied back every step as a host-side ttime would need them
    nor, acc = c.surface(t)

    def dropin(copy_state, n=10):
        g.sync()
        t2 = time.perf_counter()
        for _ in range(n):
            step()
            g.checkbcs()
            g.loads_transfer(1.0, nor, acc)
            if copy_state:
                g.get_state()
            else:
                g.diagnostics()
        g.sync()
        return (time.perf_counter() - t2) * 1e3 / n

    drop_in = {"ms_per_step_device_diagnostics": round(dropin(False), 4),
               "ms_per_step_full_state_copied_back": round(dropin(True), 4)}

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        o = Oracle(p, c.rc, c.zz, zmm, zmmt, wr)
        o.set_state(st0)
        o.set_markers(*c.markers(0.0))
        ncpu = 2
        t1 = time.perf_counter()
        tt, dtc = 0.0, p.dt0
        for k in range(ncpu):
            o.set_markers(*c.markers(tt))
            rc, dtc, _ = o.fluid_tsch(k, dtc, dtc)
            tt += dtc
        cel = time.perf_counter() - t1
        model = ""
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": round(cells * ncpu / cel / 1e6, 4), "unit": "Mcell-steps/s", "cores": 1, "kind": "port",
               "sample": f"{ncpu} time steps (3 substeps each) of the same {args.workload} workload by the "
                         f"single-threaded C oracle, {cel:.1f} s, on one core of the GPU host "
                         f"({os.cpu_count()} cores visible, {model})",
               "host_cores": os.cpu_count()}

    out = {
        "metric": "Mcell-steps/s (plain FP64 GPU fluid step)",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated grid and markers, seeded 1e-3 velocity perturbation)",
        "config": {"workload": f"{args.workload}: {n3m}x{n2m} cells, {c.nlm} markers, Re={1.0 / p.visc:g}, CFL {p.cflmax}",
                   "precision": "plain FP64" if args.precision < 0 else
                   f"FP32 MFMA transforms + {args.precision} FP64-residual refinement sweeps",
                   "poisson_transform": f"along {axis}" + (" (mirror form)" if mirror else ""),
                   "launch": args.launch},
        "roofline": roof,
        "cpu_baseline": cpu,
        "drop_in": drop_in,
        "chosen_form": "plain FP64 transforms: FP32 MFMA transforms with FP64-residual refinement measured "
                       "slower on MI355X (bench.py --precision n)",
        "kernels": kernels,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
