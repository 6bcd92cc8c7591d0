"""Benchmark: unknowns/s of the implicit diffusion solve to an FP64 relative residual of 1e-10.

Workload (one step = one whole solve; --workload picks another): 8192 x 4096 cells, periodic in
x, closed pole rows, K log-normal (ln K ~ N(0,1) smoothed by a Gaussian of 8 cells, renormalised), dt mean(K)/dx^2 =
50, h0 isotropic k^-2 spectral noise (mean 0, std 1000 m); seeded, synthetic.  Each step copies
h0 into the working array and solves it on the device (operator and levels rebuilt, the
splitting's first guess, f64 CG preconditioned by the f32 V-cycle, mass put-back, the true f64
residual checked) with the inputs already resident in HBM.

Also reported: the smoother's roofline (one f32 damped-Jacobi sweep of the finest level timed
with HIP events on the library's stream; compulsory bytes per launch / duration vs 8 TB/s), and
the CPU oracle (the reference algorithm restated in C, OpenMP over strips) timed on the host on
the same inputs, with the relative L2 difference of the two answers.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "deep-time_amd"))

METRIC = "unknowns/s to FP64 rel. residual 1e-10 at 8192×4096; smoother HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 matrix-core peak


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_of(kernel, w, h):
    """HBM bytes per launch of `kernel` on the finest level, from the committed counter summary
    (profiles/traffic_<w>x<h>.json: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors from a
    float4 copy of known size), or None when no summary covers this grid."""
    path = os.path.join(REPO, "profiles", f"traffic_{w}x{h}.json")
    try:
        with open(path) as f:
            t = json.load(f)["kernels"][kernel]
        return t["bytes"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # The first solve captures the iteration graphs and the second is their first replay; the
    # timed steps start from the third.
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="lognormal-8192x4096",
                    help="constant-512x256 | lognormal-8192x4096 | constant-16384x8192 | "
                         "belt-8192x4096 (the Jacobi step's mostly empty grid) | "
                         "regional-512 (many grids, one batched solve)")
    ap.add_argument("--size", default=None, help="WxH: the same workload at another size")
    ap.add_argument("--solver", choices=["multigrid", "jacobi"], default="multigrid",
                    help="jacobi: the Jacobi-preconditioned step (its own 1e-6 m tolerance)")
    ap.add_argument("--rel", type=float, default=1e-10)
    ap.add_argument("--cycle", choices=["f32", "f64", "bf16"], default="f32",
                    help="bf16: the f32 cycle with its two finest levels' records stored in bf16")
    ap.add_argument("--guess", choices=["lod", "rhs"], default="lod")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg")
    ap.add_argument("--stage", choices=["solve", "failure", "cube", "learned"], default="solve",
                    help="failure: the submarine slope failure after the solve (8192x4096 world "
                         "with sediment everywhere, 40 iterations); cube: the cube-grid solve "
                         "(--cube-n cells a face side); learned: the solve from the learned "
                         "first guess beside the plain solve")
    ap.add_argument("--cube-n", type=int, default=512)
    ap.add_argument("--kernels-only", action="store_true",
                    help="one solve, then only the per-kernel timings (for counter runs)")
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="timed oracle solves after one untimed one (the best is reported)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the built-in band heights (no dtd_tune before the warm-up)")
    ap.add_argument("--no-extras", action="store_true",
                    help="the default workload only: skip the other configurations' block")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); sharding over several "
              "GPUs is a design here, not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import dtdiff
    from dtdiff import fields as F
    from dtdiff.hipmem import DeviceArray, synchronize

    if args.stage == "failure":
        return bench_failure(args)
    if args.stage == "cube":
        return bench_cube(args)
    gen, kw = F.WORKLOADS[args.workload]
    if args.stage == "learned":
        return bench_learned(args, gen, kw)
    if gen is F.case_regional:
        return bench_batch(args, gen, dict(kw))
    if args.solver == "jacobi":
        return bench_jacobi(args, gen, dict(kw))
    kw = dict(kw)
    if args.size:
        kw["w"], kw["h"] = (int(v) for v in args.size.split("x"))
    w, h = kw["w"], kw["h"]
    t = time.time()
    d, h0, kd, dt = gen(**kw)
    log(f"inputs {args.workload} at {w}x{h} generated in {time.time() - t:.1f} s")
    n = d.cells()

    t = time.time()
    h0d = DeviceArray.from_numpy(h0)
    kdd = DeviceArray.from_numpy(kd)
    synchronize()
    upload_s = time.time() - t
    hd = DeviceArray(n)

    ws = dtdiff.Workspace(device=0,
                          cycle_precision={"f32": dtdiff.CYCLE_F32, "f64": dtdiff.CYCLE_F64,
                                           "bf16": dtdiff.CYCLE_BF16}[args.cycle],
                          guess=dtdiff.GUESS_LOD if args.guess == "lod" else dtdiff.GUESS_RHS)
    def step():
        hd.copy_from(h0d)  # synchronous device-to-device copy
        return dtdiff.solve_rel_device(ws, d, hd, kdd, dt, args.rel)

    tuned = None
    if not args.no_tune and not args.kernels_only:
        # The launch plan for this shape, searched once (kept in the workspace, like the graphs).
        t = time.time()
        ms_rule, ms_tuned, plan = dtdiff.tune(ws, d, h0d, kdd, dt, reps=5)
        tuned = {"tune_s": time.time() - t, "ms_rule": ms_rule, "ms_tuned": ms_tuned, "plan": plan}
        log(f"tuned in {tuned['tune_s']:.1f} s: whole solve {ms_rule:.3f} ms with the built-in "
            f"band heights, {ms_tuned:.3f} ms with the plan kept")

    if args.kernels_only:
        st = step()
        log(f"solve: {st.iterations} its, rel {st.rel_residual:.2e}, {st.ms_total:.2f} ms")
        for name, kid in (("copy_f32", dtdiff.KERNEL_COPY), ("stream5_f32", dtdiff.KERNEL_STREAM5),
                          ("bandcopy_f32", dtdiff.KERNEL_BANDCOPY),
                          ("pre_smooth_f32", dtdiff.KERNEL_PRE),
                          ("post_smooth_f32", dtdiff.KERNEL_POST),
                          ("direction_f64", dtdiff.KERNEL_DIRECTION),
                          ("update_f64", dtdiff.KERNEL_UPDATE)):
            ms, b = dtdiff.time_kernel(ws, kid, 5)
            log(f"kernel {name}: {ms * 1e3:.1f} us, {b / 1e6:.1f} MB, {b / ms / 1e6:.0f} GB/s")
        return

    for k in range(args.warmup):
        st = step()
        log(f"warmup {k}: {st.iterations} its, rel {st.rel_residual:.2e}, {st.ms_total:.2f} ms")
    synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(args.steps):
        stats.append(step())
    synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    last = stats[-1]
    x_gpu = hd.numpy()
    log(f"timed {args.steps} steps: {elapsed * 1e3 / args.steps:.3f} ms/step, "
        f"{last.iterations} its, {last.restarts} restarts, rel {last.rel_residual:.3e}; "
        f"setup {last.ms_setup:.2f} guess {last.ms_guess:.2f} cg {last.ms_cg:.2f} ms")

    kernels = {}
    roof = None
    if args.cycle in ("f32", "bf16"):
        for name, kid in (("sweep_f32", dtdiff.KERNEL_SWEEP),
                          ("direction_f64", dtdiff.KERNEL_DIRECTION),
                          ("update_f64", dtdiff.KERNEL_UPDATE),
                          ("restrict_f32", dtdiff.KERNEL_RESTRICT),
                          ("pre_smooth_f32", dtdiff.KERNEL_PRE),
                          ("post_smooth_f32", dtdiff.KERNEL_POST),
                          ("copy_f32", dtdiff.KERNEL_COPY),
                          ("stream5_f32", dtdiff.KERNEL_STREAM5)):
            ms, b = dtdiff.time_kernel(ws, kid, 20)
            kernels[name] = {"ms": ms, "bytes": b, "GB/s": b / (ms * 1e-3) / 1e9}
            log(f"kernel {name}: {ms * 1e3:.1f} us, {b / 1e6:.1f} MB, "
                f"{kernels[name]['GB/s']:.0f} GB/s")
        roof = smoother_roofline(kernels, args.cycle, w, h)

    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O

        tol_m = args.rel * float(np.linalg.norm(h0)) / np.sqrt(n)
        xo, its_o, cpu_s = oracle_best(O, d, h0, kd, dt, tol_m, args.cpu_reps)
        rel_o = O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, xo)
        parity = float(np.linalg.norm(x_gpu - xo) / np.linalg.norm(xo))
        cpu = {"value": n / cpu_s, "unit": "unknowns/s", "cores": O.threads(),
               "kind": "port",
               "sample": f"full {w}x{h} solves (reference algorithm restated in C, OpenMP over "
                         f"8-row strips), best of {max(args.cpu_reps, 1)} after one untimed: "
                         f"{cpu_s:.2f} s, {its_o} CG iterations, rel residual {rel_o:.2e}"}
        log(f"cpu oracle: {cpu_s:.2f} s, {its_o} its, rel {rel_o:.2e}; GPU vs oracle rel L2 "
            f"{parity:.3e}")

    ms_step = elapsed * 1e3 / args.steps
    out = {
        "metric": METRIC,
        "value": n * args.steps / elapsed,
        "unit": "unknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f64 (CG, residual, update) + f32 (V-cycle)",
                  "bf16": "f64 (CG, residual, update) + f32 (V-cycle; bf16 records on the two "
                          "finest levels)",
                  "f64": "f64"}[args.cycle],
        "data": "synthetic",
        "config": {"workload": f"diffusion solve {w}x{h}, " + (
                       "log-normal K (sigma 8 cells), dt*mean(K)/dx^2=50"
                       if gen is F.case_lognormal else "constant K, dt*K/dx^2=10")
                   + f", k^-2 h0, to FP64 rel. residual {args.rel:g}",
                   "grid": [w, h], "cycle": args.cycle, "guess": args.guess},
        "roofline": roof,
        "cpu_baseline": cpu,
        "iterations": int(last.iterations),
        "restarts": int(last.restarts),
        "rel_residual": last.rel_residual,
        "parity_rel_l2_vs_oracle": parity,
        "ms_breakdown": {"setup": last.ms_setup, "guess": last.ms_guess, "cg": last.ms_cg,
                         "device_total": last.ms_total},
        "upload_s_pcie": upload_s,
        "kernels": kernels,
        "tuning": tuned,
    }
    default_run = (args.workload == "lognormal-8192x4096" and not args.size
                   and args.cycle == "f32" and args.guess == "lod" and args.rel == 1e-10)
    if default_run and not args.no_extras:
        # The plain FP64 cycle and the other configurations, each measured the same way.
        ws.close()
        del hd
        out["extra_configs"] = extra_configs(args, d, h0, kd, dt, h0d, kdd, ms_step,
                                             None if args.no_cpu else xo)
    print(json.dumps(out), flush=True)


def kd_records(cycle):
    """Whether the f32 cycle's finest smoothers form their records from kd (the library's
    default; DTD_NO_KD=1 at workspace creation reads stored records instead).  The bf16-record
    cycle always reads its stored records."""
    v = os.environ.get("DTD_NO_KD", "")
    return cycle == "f32" and not (v and v != "0")


def smoother_roofline(kernels, cycle, w, h):
    """The finest level's fused smoothing, both kernels as the solve runs them: the
    pre-smoothing (the residual's update r -= alpha A p with r.r, two damped-Jacobi sweeps, the
    residual restricted 2x2) and the post-smoothing (the corrected sweep, the last sweep, r.z).
    Compulsory bytes of the two launches over their summed HIP-event durations."""
    pre, post = kernels["pre_smooth_f32"], kernels["post_smooth_f32"]
    if kd_records(cycle):
        # kRUpdateK = 4, kR64K = 3: the records (and the pre-smoothing's A p) formed from kd.
        kp, kq = "k_spre<float, 2, 4, 0>", "k_spost<float, 2, 3, 0>"
        what = ("pre: r64, p and kd in (A p and the faces formed from them), r64, z and the "
                "coarse r out, 37 B/cell; post: r64, z, kd and the coarse z in, z out, 25 B/cell")
    else:
        cb = int(cycle == "bf16")
        rec = 8 if cb else 12
        kp, kq = f"k_spre<float, 2, 2, {cb}>", f"k_spost<float, 2, 1, {cb}>"
        what = (f"pre: r64, A p and the {rec}-byte record in, r64, z and the coarse r out, "
                f"{28 + rec + 1} B/cell; post: r64, z, the record and the coarse z in, z out, "
                f"{17 + rec} B/cell")
    achieved = (pre["bytes"] + post["bytes"]) / ((pre["ms"] + post["ms"]) * 1e-3) / 1e9
    tp, tq = traffic_of(kp, w, h), traffic_of(kq, w, h)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": tp + tq if tp is not None and tq is not None else None,
            "kernel": f"{kp} + {kq}, the finest level's fused smoothing ({what}), "
                      f"{pre['ms'] * 1e3:.1f} + {post['ms'] * 1e3:.1f} us",
            "per_kernel": {kp: {"us": pre["ms"] * 1e3, "GB/s": pre["GB/s"],
                                "frac": pre["GB/s"] / HBM_PEAK_GBS},
                           kq: {"us": post["ms"] * 1e3, "GB/s": post["GB/s"],
                                "frac": post["GB/s"] / HBM_PEAK_GBS}}}


def oracle_best(O, d, h0, kd, dt, tol_m, reps):
    """The oracle's solve: one untimed, then the best of reps timed (reps < 1: one timed)."""
    if reps < 1:
        t = time.perf_counter()
        xo, its = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol_m)
        return xo, its, time.perf_counter() - t
    xo, its = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol_m)
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol_m)
        best = min(best, time.perf_counter() - t)
    return xo, its, best


def extra_configs(args, d2, h02, kd2, dt2, h0d2, kdd2, ms_f32, xo2):
    """Every other configuration of the benchmark, each as a whole solve to the true FP64
    relative residual 1e-10 on resident inputs (2 untimed solves, then the mean of 5), with its
    iterations, the true residual recomputed on the host, and the relative L2 difference from the
    oracle's answer where the oracle's solve takes seconds: the 512x256 constant case, the plain
    FP64 cycle, the bf16-record cycle and the f32 cycle with bilinear transfers or the exact f64
    coarsest solve at 8192x4096, the 16384x8192 constant case with the f32,
    bf16-record and plain FP64 cycles, the 512 regional grids in one batched solve, and the learned
    first guess at 8192x4096.  vs_plain_f64 is the plain FP64 cycle's time at the same size over
    this line's."""
    import numpy as np

    import dtdiff
    from dtdiff import fields as F
    from dtdiff.hipmem import DeviceArray, synchronize

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    warm, steps = 2, 5
    cycles = {"f32": dtdiff.CYCLE_F32, "bf16": dtdiff.CYCLE_BF16, "f64": dtdiff.CYCLE_F64}

    def timed(step):
        for _ in range(warm):
            step()
        synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = step()
        synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, st

    def single(d, h0, kd, dt, h0d, kdd, cycle, xo=None, **opts):
        ws = dtdiff.Workspace(device=0, cycle_precision=cycles[cycle], **opts)
        hd = DeviceArray(d.cells())
        try:
            def step():
                hd.copy_from(h0d)
                return dtdiff.solve_rel_device(ws, d, hd, kdd, dt, 1e-10)
            ms, st = timed(step)
            x = hd.numpy()
        finally:
            ws.close()
            hd.free()
        r = {"grid": [d.w, d.h], "cycle": cycle, **opts, "ms": ms,
             "unknowns_per_s": d.cells() / ms * 1e3,
             "iterations": int(st.iterations), "restarts": int(st.restarts),
             "rel_residual_true": O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x)}
        if xo is not None:
            r["parity_rel_l2_vs_oracle"] = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
        log(f"extra {d.w}x{d.h} {cycle}: {ms:.2f} ms, {st.iterations} its, true rel "
            f"{r['rel_residual_true']:.2e}")
        return r

    out = {}
    # 512 x 256, constant K (the reference's CPU-runnable case).
    d, h0, kd, dt = F.case_constant()
    h0d, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
    tol = 1e-10 * float(np.linalg.norm(h0)) / np.sqrt(d.cells())
    xo, its_o, cpu_s = oracle_best(O, d, h0, kd, dt, tol, 5)
    r = single(d, h0, kd, dt, h0d, kdd, "f32", xo)
    r["cpu_oracle"] = {"s": cpu_s, "unknowns_per_s": d.cells() / cpu_s, "iterations": its_o,
                       "cores": O.threads(), "sample": "best of 5 after one untimed"}
    out["constant-512x256"] = r
    # 8192 x 4096 log-normal: the plain FP64 cycle and the bf16-record cycle on the headline inputs.
    f64 = single(d2, h02, kd2, dt2, h0d2, kdd2, "f64", xo2)
    f64["vs_plain_f64"] = 1.0
    out["lognormal-8192x4096-f64"] = f64
    out["lognormal-8192x4096-f32"] = {"ms": ms_f32, "vs_plain_f64": f64["ms"] / ms_f32}
    bf = single(d2, h02, kd2, dt2, h0d2, kdd2, "bf16", xo2)
    bf["vs_plain_f64"] = f64["ms"] / bf["ms"]
    out["lognormal-8192x4096-bf16"] = bf
    # The cycle's options (no reference counterpart): bilinear transfers, the exact f64 coarsest
    # solve; the f32 cycle otherwise as the headline's.
    for name, opts in (("bilinear", {"transfer": dtdiff.TRANSFER_BILINEAR}),
                       ("exact-coarsest", {"coarsest": dtdiff.COARSEST_EXACT})):
        r = single(d2, h02, kd2, dt2, h0d2, kdd2, "f32", xo2, **opts)
        r["vs_plain_f64"] = f64["ms"] / r["ms"]
        out[f"lognormal-8192x4096-{name}"] = r
    # The reference's summation order on the headline inputs, the reference's call (an RMS
    # tolerance): the answer and the iteration count against the oracle's, bit for bit.
    out["lognormal-8192x4096-strips"] = strips_line(d2, h02, kd2, dt2, h0d2, kdd2, f64["ms"],
                                                    timed, xo2)
    # The learned first guess on the headline inputs.
    out["lognormal-8192x4096-learned"] = learned_line(d2, h02, kd2, dt2, h0d2, kdd2, f64["ms"],
                                                      timed, xo2)
    out["lognormal-8192x4096-learned-coarse"] = learned_line(d2, h02, kd2, dt2, h0d2, kdd2,
                                                             f64["ms"], timed, xo2, coarse=True)
    # 16384 x 8192 constant K.
    gen, kw = F.WORKLOADS["constant-16384x8192"]
    d, h0, kd, dt = gen(**kw)
    h0d, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
    xo = None
    if xo2 is not None:
        tol = 1e-10 * float(np.linalg.norm(h0)) / np.sqrt(d.cells())
        xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol)
    c3 = {c: single(d, h0, kd, dt, h0d, kdd, c, xo) for c in ("f64", "f32", "bf16")}
    for c in c3:
        c3[c]["vs_plain_f64"] = c3["f64"]["ms"] / c3[c]["ms"]
        out[f"constant-16384x8192-{c}"] = c3[c]
    del h0d, kdd, xo
    # The 512 regional grids, one batched solve a step.
    out["regional-512"] = batch_line(timed, xo2 is not None)
    return out


def strips_line(d, h0, kd, dt, h0d, kdd, ms_f64, timed, xo):
    """The solve with the reference's summation order (DTD_SUMS_STRIPS) to the RMS tolerance
    tol_m = 1e-10 ||h0|| / sqrt(n), the oracle's call on the CPU leg: whether h1 and the
    iteration count equal the oracle's exactly."""
    import numpy as np

    import dtdiff
    from dtdiff.hipmem import DeviceArray

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    tol_m = 1e-10 * float(np.linalg.norm(h0)) / np.sqrt(d.cells())
    ws = dtdiff.Workspace(device=0, sums=dtdiff.SUMS_STRIPS)
    hd = DeviceArray(d.cells())
    try:
        def step():
            hd.copy_from(h0d)
            return dtdiff.solve_device(ws, d, hd, kdd, dt, tol_m)

        ms, st = timed(step)
        x = hd.numpy()
    finally:
        ws.close()
        hd.free()
    r = {"grid": [d.w, d.h], "cycle": "f32", "sums": "strips (the reference's order)",
         "tol_m": tol_m, "ms": ms, "unknowns_per_s": d.cells() / ms * 1e3,
         "iterations": int(st.iterations),
         "rel_residual_true": O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x),
         "vs_plain_f64": ms_f64 / ms}
    if xo is not None:
        its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol_m)[1]
        r["iterations_oracle"] = int(its_o)
        r["bits_equal_oracle"] = bool(np.array_equal(x.view(np.uint64), xo.view(np.uint64)))
        r["parity_rel_l2_vs_oracle"] = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
    log(f"extra strips: {ms:.2f} ms, {st.iterations} its, bits equal the oracle's: "
        f"{r.get('bits_equal_oracle')}")
    return r


def learned_line(d, h0, kd, dt, h0d, kdd, ms_f64, timed, xo, coarse=False):
    """The committed network's guess (splitting + U-Net correction) and the solve from it; coarse:
    the learned coarse correction (the network on 8 x 8 block means, interpolated back)."""
    import numpy as np

    import dtdiff
    from dtdiff import learned as LN
    from dtdiff.hipmem import DeviceArray

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    ws = dtdiff.Workspace(device=0)
    hd, gd = DeviceArray(d.cells()), DeviceArray(d.cells())
    L = dtdiff.load()
    try:
        net = (LN.CoarseGuess(ws, LN.load_params(LN.WEIGHTS_COARSE)) if coarse
               else LN.UNet(ws, LN.load_params()))

        def step():
            hd.copy_from(h0d)
            net.guess(d, hd, kdd, dt, gd)
            return dtdiff.solve_rel_guess_device(ws, d, hd, kdd, dt, 1e-10, gd)

        ms, st = timed(step)
        x = hd.numpy()
        ws.check(L.dtd_sync(ws._ctx))
    finally:
        ws.close()
        hd.free()
        gd.free()
    r = {"grid": [d.w, d.h], "cycle": "f32",
         "guess": ("splitting + U-Net coarse correction (1/8 resolution)" if coarse
                   else "splitting + U-Net"), "ms": ms,
         "unknowns_per_s": d.cells() / ms * 1e3, "iterations": int(st.iterations),
         "rel_residual_true": O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x),
         "vs_plain_f64": ms_f64 / ms}
    if xo is not None:
        r["parity_rel_l2_vs_oracle"] = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
    log(f"extra learned{' coarse' if coarse else ''}: {ms:.2f} ms, {st.iterations} its")
    return r


def batch_line(timed, with_cpu):
    """The 512 regional grids in one dtd_solve_batch a step; the oracle on a fixed sample of 32
    grids (every 16th) for parity and its rate."""
    import numpy as np

    import dtdiff
    from dtdiff import fields as F
    from dtdiff.hipmem import DeviceArray

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    gen, kw = F.WORKLOADS["regional-512"]
    grids = gen(**kw)
    offs = np.cumsum([0] + [d.cells() for d, *_ in grids])
    total = int(offs[-1])
    h0_all = DeviceArray.from_numpy(np.concatenate([g[1] for g in grids]))
    kd_all = DeviceArray.from_numpy(np.concatenate([g[2] for g in grids]))
    h_all = DeviceArray(total)
    items = dtdiff.BatchGrids([(d, h_all.data_ptr() + 8 * int(o), kd_all.data_ptr() + 8 * int(o),
                                dt, 0.0) for (d, _, _, dt), o in zip(grids, offs[:-1])])
    ws = dtdiff.Workspace(device=0)
    try:
        def step():
            h_all.copy_from(h0_all)
            return dtdiff.solve_batch(ws, items, 1e-10)

        ms, (its, rels, st) = timed(step)
        x_all = h_all.numpy()
    finally:
        ws.close()
    worst = 0.0
    for (d, h0, kd, dt), o in zip(grids, offs[:-1]):
        x = x_all[int(o):int(o) + d.cells()]
        worst = max(worst, O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x))
    r = {"grids": len(grids), "cells": total, "ms": ms, "unknowns_per_s": total / ms * 1e3,
         "iterations_max": int(st.iterations_max), "restarts": int(st.restarts),
         "rel_residual_true_max": worst}
    if with_cpu:
        l2, cells, t = 0.0, 0, 0.0
        for k in range(0, len(grids), 16):
            d, h0, kd, dt = grids[k]
            tol = 1e-10 * float(np.linalg.norm(h0)) / np.sqrt(d.cells())
            t0 = time.perf_counter()
            xo, _ = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol)
            t += time.perf_counter() - t0
            cells += d.cells()
            x = x_all[int(offs[k]):int(offs[k]) + d.cells()]
            l2 = max(l2, float(np.linalg.norm(x - xo) / np.linalg.norm(xo)))
        r["parity_rel_l2_vs_oracle_max"] = l2
        r["cpu_oracle"] = {"unknowns_per_s": cells / t, "cores": O.threads(),
                           "sample": f"32 of the grids (every 16th, {cells} cells), one solve each"}
    log(f"extra regional-512: {ms:.2f} ms, its max {st.iterations_max}, restarts {st.restarts}, "
        f"true rel max {worst:.2e}")
    return r


def bench_jacobi(args, gen, kw):
    """The Jacobi-preconditioned step (`step_in`) on the device at the workload's size: its own
    tolerance of 1e-6 m RMS, its sums in the reference's order.  The CPU leg is the oracle's
    (serial) restatement on the same generator at 1/16 of the cells."""
    import numpy as np

    import dtdiff
    from dtdiff.hipmem import DeviceArray, synchronize

    if args.size:
        kw["w"], kw["h"] = (int(v) for v in args.size.split("x"))
    w, h = kw["w"], kw["h"]
    d, h0, kd, dt = gen(**kw)
    n = d.cells()
    h0d, kdd, hd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd), DeviceArray(n)
    ws = dtdiff.Workspace(device=0)

    def step():
        hd.copy_from(h0d)
        return dtdiff.step_device(ws, d, hd, kdd, dt)

    for k in range(args.warmup):
        st = step()
        log(f"warmup {k}: {st.iterations} its, {st.ms_total:.2f} ms")
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
    synchronize()
    elapsed = time.perf_counter() - t0
    log(f"timed {args.steps} steps: {elapsed * 1e3 / args.steps:.3f} ms/step, {st.iterations} "
        f"its, {st.rows_visited} of {h} rows visited; setup {st.ms_setup:.2f} guess "
        f"{st.ms_guess:.2f} cg {st.ms_cg:.2f} ms")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O

        ks = dict(kw, w=max(w // 4, 1), h=max(h // 4, 1))
        ds, h0s, kds, dts = gen(**ks)
        t = time.perf_counter()
        _, its_o = O.step_in(ds.w, ds.h, ds.dx, ds.wrap_x, h0s, kds, dts)
        cpu_s = time.perf_counter() - t
        cpu = {"value": ds.cells() / cpu_s, "unit": "unknowns/s", "cores": 1, "kind": "port",
               "sample": f"the same generator at {ds.w}x{ds.h} (1/16 of the cells), one step by "
                         f"the oracle's serial restatement: {cpu_s:.2f} s, {its_o} iterations"}
        log(f"cpu oracle sample {ds.w}x{ds.h}: {cpu_s:.2f} s, {its_o} its")
    out = {
        "metric": f"unknowns/s of the Jacobi-preconditioned step (1e-6 m RMS) at {w}x{h}",
        "value": n * args.steps / elapsed,
        "unit": "unknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Jacobi step {w}x{h} ({args.workload} generator)"},
        "roofline": None,
        "cpu_baseline": cpu,
        "iterations": int(st.iterations),
        "rows_visited": int(st.rows_visited),
        "ms_breakdown": {"setup": st.ms_setup, "guess": st.ms_guess, "cg": st.ms_cg,
                         "device_total": st.ms_total},
    }
    print(json.dumps(out), flush=True)


def bench_failure(args):
    """The submarine slope failure (`marine::slope_failure`) on resident arrays: the reference
    tests' world generator (`failure_world`, sediment everywhere, so every iteration moves) at
    8192x4096 (or --size), the stage's 40 iterations.  A step restores the inputs (a device copy,
    outside the timed region) and runs the call, which returns once the device is done (host clock
    around it).  Roofline: k_fail_sweep's compulsory bytes, h and avail in and out (32 B) and the
    sea flag (1 B) a cell an iteration, over the whole call's time (its step kernels and launches
    included, so a lower bound on the sweep kernel's own rate)."""
    import numpy as np

    import dtdiff
    from dtdiff.hipmem import DeviceArray, synchronize

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    w, h = (8192, 4096) if not args.size else (int(v) for v in args.size.split("x"))
    t = time.time()
    h0, a0, sea = O.failure_world(w, h, 7, 0)
    log(f"failure world {w}x{h} generated in {time.time() - t:.1f} s")
    d = dtdiff.Dims(w, h, 4800.0, True)
    h0d, a0d, sd = (DeviceArray.from_numpy(v) for v in (h0, a0, sea))
    hd, ad = DeviceArray(w * h), DeviceArray(w * h)
    ws = dtdiff.Workspace(device=0)

    def step():
        hd.copy_from(h0d)
        ad.copy_from(a0d)
        synchronize()
        t0 = time.perf_counter()
        r = dtdiff.slope_failure_device(ws, d, sd, 0.0, hd, ad, O.CRITICAL_SLOPE)
        return r, (time.perf_counter() - t0) * 1e3

    for _ in range(args.warmup):
        step()
    total_ms, its = 0.0, 0
    for _ in range(args.steps):
        (moved, n), ms = step()
        total_ms += ms
        its += n
    ms_step = total_ms / args.steps
    sweep_bytes = 33.0 * w * h
    achieved = sweep_bytes * its / (total_ms * 1e-3) / 1e9
    log(f"timed {args.steps} calls: {ms_step:.3f} ms/call, {n} iterations, moved {moved:.6e}, "
        f"{achieved:.0f} GB/s over the call")
    cpu = None
    if not args.no_cpu:
        ws_, hs_ = 1024, 512
        hc, ac, sc = O.failure_world(ws_, hs_, 7, 0)
        t = time.perf_counter()
        _, _, mo = O.slope_failure(ws_, hs_, 4800.0, True, sc, 0.0, O.CRITICAL_SLOPE,
                                   O.FAILURE_RATE, O.FAILURE_ITERATIONS, hc, ac)
        cpu_s = time.perf_counter() - t
        cpu = {"value": ws_ * hs_ * O.FAILURE_ITERATIONS / cpu_s, "unit": "cell-iterations/s",
               "cores": O.threads(), "kind": "port",
               "sample": f"the same generator at {ws_}x{hs_}, 40 iterations of the oracle's full "
                         f"sweep (OpenMP over cells): {cpu_s:.2f} s"}
        log(f"cpu oracle sample {ws_}x{hs_}: {cpu_s:.2f} s")
    out = {
        "metric": f"cell-iterations/s of the submarine slope failure at {w}x{h}",
        "value": w * h * its / (total_ms * 1e-3),
        "unit": "cell-iterations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"slope failure {w}x{h}, failure_world seed 7 with sediment "
                               f"everywhere, S_c = tan 2.5 deg, rate 0.05, 40 iterations"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_fail_sweep (+ its one-block step, over the whole call): h and "
                               "avail in and out, sea flag in: 33 B/cell an iteration"},
        "cpu_baseline": cpu,
        "iterations": int(n),
    }
    print(json.dumps(out), flush=True)


def bench_cube(args):
    """The cube-grid solve (`surface::cube::diffusion::solve`, multigrid) on resident arrays: a
    cube of --cube-n cells a face side (6 n^2 cells) from `dtdiff.cubegrid`, a log-normal kd, k^-2
    spectral heights mapped face by face, dt mean(kd) / mean area = 50, to an RMS residual of
    1e-6 m.  A step restores h (a device copy, outside the timed region) and runs the call, which
    returns when the device is done (host clock around it).  The CPU leg is the oracle's solve at
    n / 4 on the same generator."""
    import numpy as np

    import dtdiff
    from dtdiff import cubegrid as CG
    from dtdiff import fields as F
    from dtdiff.hipmem import DeviceArray, synchronize

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    def inputs(n):
        g = CG.cube_grid(n)
        h = np.concatenate([F.spectral_noise(n, n, 11 + f) for f in range(6)])
        lnk = F.splitmix_normal(23, g.cells)
        kd = np.exp(0.5 * lnk)
        kd *= 1.0e3 / kd.mean()
        dt = 50.0 * g.mean_area / 1.0e3
        return g, h, kd, dt

    n = args.cube_n
    t = time.time()
    g, h0, kd, dt = inputs(n)
    log(f"cube n {n} ({g.cells} cells) generated in {time.time() - t:.1f} s")
    ws = dtdiff.Workspace(device=0)
    cube = dtdiff.Cube(ws, g)
    h0d, kdd, hd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd), DeviceArray(g.cells)

    def step():
        hd.copy_from(h0d)
        synchronize()
        t0 = time.perf_counter()
        its = cube.solve_device(hd, kdd, dt, dtdiff.DiffusionSolver.MULTIGRID, 1e-6)
        return its, time.perf_counter() - t0

    for _ in range(args.warmup):
        step()
    total, its = 0.0, 0
    for _ in range(args.steps):
        its, el = step()
        total += el
    ms = total * 1e3 / args.steps
    log(f"timed {args.steps} solves: {ms:.3f} ms/solve, {its} iterations")
    cpu = None
    if not args.no_cpu:
        ns = max(n // 4, 2)
        gs, hs, kds, dts = inputs(ns)
        t = time.perf_counter()
        _, its_o = O.cube_solve(gs, hs, kds, dts, 0, 1e-6)
        cpu_s = time.perf_counter() - t
        cpu = {"value": gs.cells / cpu_s, "unit": "unknowns/s", "cores": O.threads(),
               "kind": "port",
               "sample": f"the same generator at n {ns} ({gs.cells} cells), one multigrid solve "
                         f"by the oracle (OpenMP over cells): {cpu_s:.2f} s, {its_o} iterations"}
        log(f"cpu oracle sample n {ns}: {cpu_s:.2f} s")
    cube.close()
    out = {
        "metric": f"unknowns/s of the cube-grid multigrid solve at n {n} (6 n^2 cells)",
        "value": g.cells * args.steps / total,
        "unit": "unknowns/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"cube grid n {n}: {g.cells} cells, log-normal kd, "
                               f"dt mean(kd)/mean area = 50, RMS residual 1e-6 m"},
        "roofline": None,
        "cpu_baseline": cpu,
        "iterations": int(its),
    }
    print(json.dumps(out), flush=True)


def bench_learned(args, gen, kw):
    """The learned first guess at the workload's size (8192x4096 log-normal by default): the plain
    solve (the splitting's guess) and the solve started from the network's guess, each to the
    same true FP64 relative residual, on resident arrays.  A learned step is the guess (splitting +
    network) plus the solve, host clock around both with the device synchronised.  The network's
    weights are the committed ones (trained by `python -m dtdiff.learned train` on 1024x512 solves
    of the same generators)."""
    import numpy as np

    import dtdiff
    from dtdiff import learned as LN
    from dtdiff.hipmem import DeviceArray, synchronize

    kw = dict(kw)
    if args.size:
        kw["w"], kw["h"] = (int(v) for v in args.size.split("x"))
    w, h = kw["w"], kw["h"]
    d, h0, kd, dt = gen(**kw)
    n = d.cells()
    h0d, kdd, hd, gd = (DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd), DeviceArray(n),
                        DeviceArray(n))
    ws = dtdiff.Workspace(device=0)
    net = LN.UNet(ws, LN.load_params())
    cnet = LN.CoarseGuess(ws, LN.load_params(LN.WEIGHTS_COARSE))
    L = dtdiff.load()

    def plain():
        hd.copy_from(h0d)
        synchronize()
        t0 = time.perf_counter()
        st = dtdiff.solve_rel_device(ws, d, hd, kdd, dt, args.rel)
        return st, time.perf_counter() - t0

    def learned():
        hd.copy_from(h0d)
        synchronize()
        t0 = time.perf_counter()
        net.guess(d, hd, kdd, dt, gd)
        ws.check(L.dtd_sync(ws._ctx))
        t1 = time.perf_counter()
        st = dtdiff.solve_rel_guess_device(ws, d, hd, kdd, dt, args.rel, gd)
        return st, time.perf_counter() - t0, t1 - t0

    def coarse():
        hd.copy_from(h0d)
        synchronize()
        t0 = time.perf_counter()
        cnet.guess(d, hd, kdd, dt, gd)
        ws.check(L.dtd_sync(ws._ctx))
        t1 = time.perf_counter()
        st = dtdiff.solve_rel_guess_device(ws, d, hd, kdd, dt, args.rel, gd)
        return st, time.perf_counter() - t0, t1 - t0

    for _ in range(args.warmup):
        plain()
        learned()
        coarse()
    tp, tl, tg, tc, tcg = 0.0, 0.0, 0.0, 0.0, 0.0
    for _ in range(args.steps):
        sp, e = plain()
        tp += e
        sl, e, g = learned()
        tl += e
        tg += g
        sc, e, g = coarse()
        tc += e
        tcg += g
    x_ref = None
    # Accuracy of the two guesses against the converged answer.
    plain()
    synchronize()
    x_ref = hd.numpy()
    dtdiff.lod_device(ws, d, h0d, kdd, dt, gd)
    ws.check(L.dtd_sync(ws._ctx))
    err_lod = float(np.sqrt(np.mean((gd.numpy() - x_ref) ** 2)))
    net.guess(d, h0d, kdd, dt, gd)
    ws.check(L.dtd_sync(ws._ctx))
    err_net = float(np.sqrt(np.mean((gd.numpy() - x_ref) ** 2)))
    cnet.guess(d, h0d, kdd, dt, gd)
    ws.check(L.dtd_sync(ws._ctx))
    err_coarse = float(np.sqrt(np.mean((gd.numpy() - x_ref) ** 2)))
    # The network's forward alone as the guess runs it (the convolutions on the matrix cores, the
    # decoder reading its skip concatenations in place), and its flops.
    net.set_input(d, h0d, kdd, gd, dt)
    ws.check(L.dtd_sync(ws._ctx))
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        net.forward(train=False)
    ws.check(L.dtd_sync(ws._ctx))
    fwd_ms = (time.perf_counter() - t0) * 1e3 / reps
    flops = 0.0
    for name, ci, co, _ in LN.LAYERS:
        lvl = 0 if name == "out" else int(name[3])
        flops += 2.0 * (n >> (2 * lvl)) * co * 9 * ci
    tflops = flops / (fwd_ms * 1e-3) / 1e12
    log(f"plain {tp * 1e3 / args.steps:.2f} ms ({sp.iterations} its), learned "
        f"{tl * 1e3 / args.steps:.2f} ms ({sl.iterations} its, guess {tg * 1e3 / args.steps:.2f} "
        f"ms), coarse correction {tc * 1e3 / args.steps:.2f} ms ({sc.iterations} its, guess "
        f"{tcg * 1e3 / args.steps:.2f} ms), network forward {fwd_ms:.1f} ms = {tflops:.0f} "
        f"TFLOP/s; guess RMS error: splitting {err_lod:.3f} m, network {err_net:.3f} m, coarse "
        f"correction {err_coarse:.3f} m")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O

        ds, hs, kds, dts = gen(**dict(kw, w=1024, h=512))
        tol_m = args.rel * float(np.linalg.norm(hs)) / np.sqrt(ds.cells())
        t = time.perf_counter()
        _, its_o = O.solve(ds.w, ds.h, ds.dx, ds.wrap_x, hs, kds, dts, tol_m)
        cpu_s = time.perf_counter() - t
        cpu = {"value": ds.cells() / cpu_s, "unit": "unknowns/s", "cores": O.threads(),
               "kind": "port",
               "sample": f"the plain solve of the same generator at 1024x512 by the oracle "
                         f"(reference algorithm restated in C, OpenMP): {cpu_s:.2f} s, "
                         f"{its_o} iterations"}
    out = {
        "metric": f"unknowns/s to FP64 rel. residual {args.rel:g} at {w}x{h} with the learned "
                  f"first guess",
        "value": n * args.steps / tl,
        "unit": "unknowns/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": tl * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 network (f32 accumulation) + f64 CG with the f32 V-cycle",
        "data": "synthetic",
        "config": {"workload": f"diffusion solve {w}x{h}, log-normal K (sigma 8 cells), "
                               f"dt*mean(K)/dx^2=50, k^-2 h0, first guess = splitting + U-Net "
                               f"correction (32/64/128/256 channels)"},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": tflops / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
                     "kernel": f"the network's forward (15 3x3 convolutions, {flops / 1e12:.2f} "
                               f"TFLOP, bf16 MFMA with f32 accumulation), {fwd_ms:.1f} ms"},
        "cpu_baseline": cpu,
        "plain_unknowns_per_s": n * args.steps / tp,
        "plain_ms": tp * 1e3 / args.steps,
        "iterations_plain": int(sp.iterations),
        "iterations_learned": int(sl.iterations),
        "guess_ms": tg * 1e3 / args.steps,
        "network_forward_ms": fwd_ms,
        "guess_rms_error_m": {"splitting": err_lod, "network": err_net,
                              "coarse_correction": err_coarse},
        "rel_residual": {"plain": sp.rel_residual, "learned": sl.rel_residual,
                         "coarse_correction": sc.rel_residual},
        "coarse_correction": {"ms": tc * 1e3 / args.steps, "guess_ms": tcg * 1e3 / args.steps,
                              "iterations": int(sc.iterations),
                              "unknowns_per_s": n * args.steps / tc},
    }
    print(json.dumps(out), flush=True)


def bench_batch(args, gen, kw):
    """The regional workload: 512 grids (w 64..1024, h 64..512, multiples of 32) solved by one
    dtd_solve_batch call a step, every grid to the relative target.  The grids' h0 and kd lie in
    two contiguous device arrays; a step copies h0 into the working array and solves."""
    import numpy as np

    import dtdiff
    from dtdiff.hipmem import DeviceArray, synchronize

    t = time.time()
    grids = gen(**kw)
    log(f"inputs {args.workload}: {len(grids)} grids generated in {time.time() - t:.1f} s")
    offs = np.cumsum([0] + [d.cells() for d, *_ in grids])
    total = int(offs[-1])
    h0_all = DeviceArray.from_numpy(np.concatenate([g[1] for g in grids]))
    kd_all = DeviceArray.from_numpy(np.concatenate([g[2] for g in grids]))
    h_all = DeviceArray(total)
    synchronize()
    items = dtdiff.BatchGrids([(d, h_all.data_ptr() + 8 * int(o), kd_all.data_ptr() + 8 * int(o),
                                dt, 0.0) for (d, _, _, dt), o in zip(grids, offs[:-1])])
    ws = dtdiff.Workspace(device=0)

    def step():
        h_all.copy_from(h0_all)
        return dtdiff.solve_batch(ws, items, args.rel)

    for k in range(args.warmup):
        its, rels, st = step()
        log(f"warmup {k}: its max {st.iterations_max} sum {st.iterations_sum}, rel max "
            f"{st.rel_residual_max:.2e}, {st.ms_total:.2f} ms")
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        its, rels, st = step()
    synchronize()
    elapsed = time.perf_counter() - t0
    x_all = h_all.numpy()
    log(f"timed {args.steps} steps: {elapsed * 1e3 / args.steps:.3f} ms/step; its max "
        f"{st.iterations_max} mean {st.iterations_sum / st.grids:.2f}, restarts {st.restarts}, "
        f"rel max {st.rel_residual_max:.3e}; setup {st.ms_setup:.2f} guess {st.ms_guess:.2f} "
        f"cg {st.ms_cg:.2f} ms")
    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O

        worst = 0.0
        its_o = []
        t = time.perf_counter()
        for (d, h0, kd, dt), o in zip(grids, offs[:-1]):
            tol_m = args.rel * float(np.linalg.norm(h0)) / np.sqrt(d.cells())
            xo, k = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol_m)
            its_o.append(k)
            x = x_all[int(o):int(o) + d.cells()]
            worst = max(worst, float(np.linalg.norm(x - xo) / np.linalg.norm(xo)))
        cpu_s = time.perf_counter() - t
        parity = worst
        cpu = {"value": total / cpu_s, "unit": "unknowns/s", "cores": O.threads(), "kind": "port",
               "sample": f"all {len(grids)} grids solved one after another by the reference "
                         f"algorithm restated in C (OpenMP over 8-row strips): {cpu_s:.2f} s, "
                         f"{sum(its_o)} CG iterations in all"}
        log(f"cpu oracle: {cpu_s:.2f} s; worst per-grid GPU vs oracle rel L2 {worst:.3e}")
    out = {
        "metric": "unknowns/s of a batch of regional grids to FP64 rel. residual "
                  f"{args.rel:g} (one batched solve)",
        "value": total * args.steps / elapsed,
        "unit": "unknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (CG, residual, update) + f32 (V-cycle)",
        "data": "synthetic",
        "config": {"workload": f"{len(grids)} grids, w 64..1024 x h 64..512 (multiples of 32), "
                               f"periodic in x, log-normal K (sigma 8 cells), "
                               f"dt*mean(K)/dx^2 in [5, 200], k^-2 h0, {total} cells"},
        "roofline": None,
        "cpu_baseline": cpu,
        "iterations_max": int(st.iterations_max),
        "iterations_mean": st.iterations_sum / st.grids,
        "restarts": int(st.restarts),
        "rel_residual_max": st.rel_residual_max,
        "parity_rel_l2_vs_oracle_max": parity,
        "ms_breakdown": {"setup": st.ms_setup, "guess": st.ms_guess, "cg": st.ms_cg,
                         "device_total": st.ms_total},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
