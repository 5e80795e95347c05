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
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def pmc_traffic(name):
    """HBM bytes per dispatch of `name` from the committed rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return d.get(name, {}).get("bytes_per_dispatch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=30,
                    help="untimed steps first (the engine clock takes some tens of ms to settle)")
    ap.add_argument("--workload", default="rect2048")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--transform", default="auto", choices=["auto", "y", "z"],
                    help="axis of the Poisson eigen-transform (auto: the shorter one)")
    ap.add_argument("--launch", default="stream", choices=["stream", "graph"],
                    help="kernel launches of a step: one by one on the stream, or one captured HIP graph "
                         "(steps with GEMM timing events always launch on the stream)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the library's plain 256 MiB copy kernel 3+1 times (HBM counter scale)")
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
    g.set_graph(args.launch == "graph")
    g.set_state(st0)

    t, dt = 0.0, p.dt0
    ns = 0

    def step():
        nonlocal t, dt, ns
        g.set_markers(*c.markers(t))
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
    gemm_flops = 2.0 * n3m * n2m * n2m if axis == "y" else 2.0 * n2m * n3m * n3m
    achieved = gemm_flops / gemm_avg_s / 1e12
    # the transforms of this workload take the direct-to-LDS kernel (at least
    # 192 tiles of 96x128 in the library's size test, operands with padded
    # leading dimensions; the kernel itself runs 96x64 tiles)
    gemm_kernel = "k_gemm_dma" if -(-n2m // 96) * -(-n3m // 128) >= 192 else "k_gemm_nt"
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_MFMA_TFLOPS, 4), "traffic": pmc_traffic(gemm_kernel),
            "kernel": gemm_kernel, "avg_ms": round(gemm_avg_s * 1e3, 4), "transform_axis": axis,
            "flops_per_launch": gemm_flops}
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

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        o = Oracle(p, c.rc, c.zz, zmm, zmmt, wr)
        o.set_state(st0)
        o.set_markers(*c.markers(0.0))
        t1 = time.perf_counter()
        rc, _, _ = o.fluid_tsch(0, p.dt0, p.dt0)
        cel = time.perf_counter() - t1
        cpu = {"value": round(cells / cel / 1e6, 4), "unit": "Mcell-steps/s", "cores": 1, "kind": "port",
               "sample": f"1 time step (3 substeps) of the same {args.workload} workload, single-threaded C "
                         f"oracle, {cel:.1f} s"}

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
                   "precision": "plain FP64", "poisson_transform": f"along {axis}", "launch": args.launch},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
