#!/usr/bin/env python3
"""Benchmark of the MI355X projector path on one GPU.

Workload (N=1): plain FP32 OS-SART on a 512x512x256 volume from 121 projections
of 512x512 (-60..+60 deg, 1 deg step), 10 subsets (13 projections per update,
TomoJ's update = 13), relaxation 1.0, seeded 40-ellipsoid phantom with analytic
projections + 1% Gaussian noise generated on the device.  One step = one full
OS-SART iteration (every projection forward-projected, differenced and
backprojected once, 10 volume updates).  Inputs are resident in HBM before the
timed region.  Extras on the same line: voxel-ray GUPS, the per-kernel HBM
roofline from HIP events, the literal 512^3 SIRT iteration rate, the weighted
backprojection pass, the precision form (16-bit storage modes against the FP64
twin, at 512^3), the batched form (64 uneven subtomograms in one graph launch),
the serving form (a Poisson stream, p50/p99 latency), and TomoJ's CPU
algorithm (the C restatement in oracle/) timed on the host cores for one full
iteration of the same workload.

Roofline figures: `roofline` is the dominant kernel (algorithmic bytes per
launch over its HIP-event time, with its counter-measured traffic and its VALU
issue fraction); `hbm_iteration` is the counter-measured HBM bytes of a whole
iteration (profiles/traffic.json) over the measured iteration time; the
`iteration_model` entry is the per-projection streaming model (one volume
read per projection), a model figure that counts more volume reads than the
kernels make.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-forms]
       python bench.py --workload sirt_512x512x512_121 ...   (another timed workload)
       python bench.py --precision-study W H D N   (one precision study only)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SIRT iters/s + voxel-ray GUPS at 512³ (HBM %roofline); serve p50/p99 latency"
HBM_PEAK_GBS = 8000.0


# MI355X vector issue rate: 256 CUs x 4 SIMDs x 16 lanes per clock x 2.4 GHz
# (FP32 and FP64 adds and multiplies both issue at this rate; packed FP32 is
# not used).
VALU_PEAK = 256 * 64 * 2.4e9
def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(w):
    """Algorithmic HBM bytes per iteration of each timed kernel group (FP32
    storage), each the least any implementation of the step must move.  fp:
    the forward projection fused with the residual, the volume read once per
    subset launch (every projection of the subset samples all of it) plus the
    measured images read and the differences written (the per-projection
    streaming model, one volume read per projection, is reported apart as
    `iteration_model`: it counts reads no implementation needs to make).  res:
    the fold of the per-block error partials and row flags.  bpprep: the
    backprojection's row table and FP64 pair images (difference read, 16-byte
    pair written per pixel).  bp: the backprojection kernel, the volume read
    and written once per subset plus the differences (4 bytes per pixel; the
    kernel stages the pair images)."""
    nvox, npx, n, s = w.nvox, w.npx, w.n, w.subsets()
    ntile = (w.W + 255) // 256
    nrb = (w.H + 3) // 4
    return {
        "fp": s * nvox * 4 + n * npx * 8,
        "res": n * (nrb * ntile * 8 + w.H * ntile + w.H + 8),
        "bpprep": n * npx * (4 + 16) + s * w.H * 32,
        "bp": s * 2 * nvox * 4 + n * npx * 4,
    }


# Kernels of one plain OS-SART/SIRT iteration on the LDS path, as rocprofv3
# names them, grouped as tj_timing groups them.  The backprojection group
# launches both forms; the row table picks one on the device and the other
# returns at once.
ITERATION_KERNELS = {"fp": "k_fp_lds", "res": "k_err_fold", "bpprep": "k_bp_prep",
                     "bp": ("k_bp_fast", "k_bp_lds")}


def working_kernel(group, tim):
    """The kernel of a timed group that did the group's work: the fast
    backprojection when its launches ran (tj_timing counts them), else the
    general one."""
    if group == "bp":
        return "k_bp_fast" if tim.get("fast_bp_launches", 0) > 0 else "k_bp_lds"
    names = ITERATION_KERNELS[group]
    return names if isinstance(names, str) else names[-1]


def load_traffic():
    """profiles/traffic.json: HBM bytes per launch of each kernel of a workload
    from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs, scaled by the copy-kernel
    calibration.  None when the file is absent."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return None


def _counter_record(traffic, workload, kernel):
    try:
        ks = traffic["workloads"][workload]["kernels"]
    except (TypeError, KeyError):
        return None
    for name, rec in ks.items():
        if name == kernel or name.startswith(kernel + "<"):
            return rec
    return None


def measured_traffic(traffic, workload, kernel):
    """Calibrated HBM bytes per launch of `kernel` (a rocprofv3 name prefix) on
    `workload`, or None."""
    rec = _counter_record(traffic, workload, kernel)
    return rec["hbm_bytes_per_launch"] if rec else None


def valu_lane_ops_per_update(traffic, w, kernel):
    """VALU lane operations per voxel-ray update of `kernel` on workload `w`:
    SQ_INSTS_VALU (wave instructions per launch, rocprofv3) x 64 lanes over the
    updates of an average launch (N x Nvox / subsets)."""
    rec = _counter_record(traffic, w.name, kernel)
    if not rec or not rec.get("valu_wave_instr"):
        return None
    return rec["valu_wave_instr"] * 64 / (w.n * w.nvox / w.subsets())


def iteration_traffic(traffic, w, launches_per_iteration):
    """Counter-measured HBM bytes of one iteration: the calibrated bytes per
    launch of every kernel of the iteration times its launches per iteration."""
    total = 0.0
    parts = {}
    for group, names in ITERATION_KERNELS.items():
        for kname in (names if isinstance(names, tuple) else (names,)):
            b = measured_traffic(traffic, w.name, kname)
            if b is None:
                return None, None
            parts[kname] = b * launches_per_iteration[group]
            total += parts[kname]
    return total, parts


def kernel_table(w, tim, steps):
    """Per timed group: launches, average launch time (HIP events on the
    library's stream) and algorithmic bytes per launch."""
    kb = kernel_bytes(w)
    out = {}
    for k in ("fp", "res", "bpprep", "bp"):
        launches = tim[f"{k}_launches"]
        ms_k = tim[f"{k}_ms"]
        total_bytes = kb[k] * steps
        out[k] = {
            "launches": launches,
            "avg_ms": ms_k / max(launches, 1),
            "bytes_per_launch": total_bytes / max(launches, 1),
            "achieved_GBs": total_bytes / (ms_k * 1e-3) / 1e9 if ms_k > 0 else None,
        }
    return out


def roofline_of(w, kernels, tim, steps, traffic):
    """The roofline object of the dominant kernel (the longest of the forward
    and backprojection launches): algorithmic bytes per launch over its live
    launch time against the 8 TB/s HBM peak, its counter-measured traffic, and
    the same launch against the vector issue rate."""
    dom = max(("fp", "bp"), key=lambda k: tim[f"{k}_ms"])
    kname = working_kernel(dom, tim)
    achieved = kernels[dom]["achieved_GBs"]
    r = {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(traffic, w.name, kname),
        "kernel": kname,
        "bytes_per_launch": kernels[dom]["bytes_per_launch"],
        "avg_launch_ms": kernels[dom]["avg_ms"],
        "limiter": "VALU issue and LDS: FP64 positions and ImageJ's FP64 bilinear weights per voxel update",
    }
    if r["traffic"]:
        r["traffic_over_algorithmic"] = r["traffic"] / r["bytes_per_launch"]
    per_update = valu_lane_ops_per_update(traffic, w, kname)
    if per_update:
        upd = w.voxel_ray_updates_per_iteration() / 2 * steps / max(tim[f"{dom}_launches"], 1)
        valu = per_update * upd / (kernels[dom]["avg_ms"] * 1e-3)
        r["valu"] = {"lane_ops_per_update": per_update, "achieved": valu / 1e12, "peak": VALU_PEAK / 1e12,
                     "unit": "T lane-ops/s", "frac": valu / VALU_PEAK}
    return r


def hbm_iteration(w, tim, steps, ms_per_step, traffic):
    """Counter-measured HBM bytes of one whole iteration over the measured
    iteration time, against the 8 TB/s peak."""
    lpi = {k: tim[f"{k}_launches"] / steps for k in ITERATION_KERNELS}
    total, parts = iteration_traffic(traffic, w, lpi)
    if total is None:
        return None
    gbs = total / (ms_per_step * 1e-3) / 1e9
    return {"bytes_per_iteration": total, "achieved_GBs": gbs, "peak": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS,
            "by_kernel": parts,
            "source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, copy-calibrated (profiles/traffic.json)"}


def run_gpu(w, steps, warmup, tj):
    from tomoj_amd.workloads import eulers_for
    E = eulers_for(w)
    P = tj.VoxelProjector3DHIP((w.W, w.H, w.D))
    ell = tj.phantom_ellipsoids(w.seed, w.ellipsoids, w.W, w.H, w.D)
    P.synth_tilt_series(ell, E, w.noise_rel, w.noise_seed)
    if warmup > 0:
        P.run_ossart(warmup, w.relaxation, w.update, first_index=0)
    P.reset_volume()
    P.enable_timing(True)
    t0 = time.perf_counter()
    err = P.run_ossart(steps, w.relaxation, w.update, first_index=0)
    t1 = time.perf_counter()
    tim = P.timing()
    return P, E, err, (t1 - t0), tim


def host_cpu():
    """Host CPU description: the model name, the logical CPUs the OS reports,
    the CPUs this process may run on, and OMP_NUM_THREADS."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(w, P, E, oracle):
    """TomoJ's CPU algorithm (C restatement, TomoJ's 2-row task decomposition with a
    join per projection) for one full OS-SART iteration of the same workload, on
    the host CPUs this process is given: OMP_NUM_THREADS when it is set (the
    share of the host a GPU job gets), else every CPU in the affinity mask."""
    cpu = host_cpu()
    cores = int(cpu["omp_num_threads"] or 0) or cpu["affinity"] or 1
    cores = max(1, min(cores, cpu["affinity"] or cores))
    exps = np.stack([P.get_loaded_projection(k) for k in range(w.n)])
    g = oracle.make_geom(w.W, w.H, w.D, nthreads=cores)
    vol = np.zeros((w.D, w.H, w.W), np.float32)
    t0 = time.perf_counter()
    oracle.ossart(g, vol, exps, E, 1, w.relaxation, w.update, first_index=0)
    dt = time.perf_counter() - t0
    return {
        "value": 1.0 / dt,
        "unit": "OS-SART iterations/s",
        "cores": cores,
        "kind": "port",
        "sample": f"one full OS-SART iteration ({w.subsets()} subsets, {w.n} projections) of "
                  f"{w.W}x{w.H}x{w.D}, C restatement of VoxelProjector3D/TomoReconstruction2 "
                  f"(TomoJ CPU algorithm, not the JVM), {cores} OpenMP threads, {dt:.1f} s",
        "gups": w.voxel_ray_updates_per_iteration() / dt / 1e9,
        "host": cpu,
    }


def regularisation_extras(w, P, tj, iters=3):
    """SIRT, TVM (TomoJ's defaults: 5 TV steps, theta 25, g 1, lambda 0.1),
    Bayesian and BgART (k 1, dark background, cross median) iterations on the
    workload's resident data, a TV step alone, and the FSC of two resident
    volumes."""
    import ctypes as C
    res = {}
    for key, setup in (("sirt", lambda: None), ("tvm", lambda: P.setTvm(5, 25.0, 1.0, 0.1)),
                       ("bayesian", lambda: P.setBayesian(True, 1.0, 0.01)),
                       ("bgart", lambda: P.setBgart(True, 1.0, True, True))):
        P.setTvm(0)
        P.setBayesian(False)
        P.setBgart(False)
        setup()
        P.reset_volume()
        P.run_ossart(1, 1.0, 0)
        t0 = time.perf_counter()
        P.run_ossart(iters, 1.0, 0)
        res[f"{key}_it_s"] = iters / (time.perf_counter() - t0)
    P.setTvm(0)
    P.setBayesian(False)
    P.setBgart(False)
    P.tv_denoise(25.0, 1.0, 0.1, 1)
    t0 = time.perf_counter()
    P.tv_denoise(25.0, 1.0, 0.1, 10)
    t1 = time.perf_counter()
    P.tv_denoise(25.0, 1.0, 0.1, 2)
    t2 = time.perf_counter()
    step = ((t1 - t0) - (t2 - t1)) / 8          # per step, set-up and flag read-back removed
    nvox = w.W * w.H * w.D
    res["tv_step_ms"] = step * 1e3
    res["tv_step_GBs"] = 48.0 * nvox / step / 1e9   # 28 + 20 algorithmic bytes per voxel
    Q = tj.VoxelProjector3DHIP((w.W, w.H, w.D))
    try:
        Q.set_volume(P.get_volume() * np.float32(0.9))
        out = np.zeros((4, w.W // 2 + 1))
        tj._lib.call("tj_fsc", P._ctx, Q._ctx, out.ctypes.data_as(C.c_void_p))
        t0 = time.perf_counter()
        tj._lib.call("tj_fsc", P._ctx, Q._ctx, out.ctypes.data_as(C.c_void_p))
        res["fsc_ms"] = (time.perf_counter() - t0) * 1e3
    finally:
        Q.close()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="ossart_512x512x256_121",
                    choices=["ossart_512x512x256_121", "sirt_512x512x512_121", "sirt_256x256x128_61"],
                    help="the timed workload (default: the production OS-SART configuration)")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="also run a plain device copy of BYTES (HBM counter calibration)")
    ap.add_argument("--no-forms", action="store_true", help="skip the precision/batched/serving forms")
    ap.add_argument("--no-serving", action="store_true", help="forms: skip the serving runs")
    ap.add_argument("--precision-study", type=int, nargs=4, metavar=("W", "H", "D", "N"),
                    help="run only the precision study (30 + 2 iterations, 80 ellipsoids, 2%% noise)")
    ap.add_argument("--modes", nargs="+", default=None,
                    help="precision study: only these storage modes (e.g. bf16+master+p32)")
    ap.add_argument("--no-reference", action="store_true", help="precision study: skip the FP64 twin")
    ap.add_argument("--no-tune", action="store_true", help="precision study: default tile shapes, no autotuning")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import tomoj_amd as tj
    from tomoj_amd import workloads as WL

    if args.precision_study:
        from tomoj_amd import forms
        W, H, D, N = args.precision_study
        modes = forms.PRECISION_MODES
        if args.modes:
            modes = [m for m in modes if m[0] in args.modes]
            if not modes:
                ap.error("--modes: none of " + ", ".join(m[0] for m in forms.PRECISION_MODES))
        res = forms.precision_study(W, H, D, N, iterations=30, refine=2, ellipsoids=80, noise=0.02, modes=modes,
                                    reference=not args.no_reference, tune=not args.no_tune)
        print(json.dumps({"precision_study": res}), flush=True)
        return

    w = WL.ALL[args.workload]
    traffic = load_traffic()
    log(f"[bench] workload {w.name}: {w.W}x{w.H}x{w.D}, {w.n} projections, update {w.update}")
    P, E, err, dt, tim = run_gpu(w, args.steps, args.warmup, tj)
    ms = dt / args.steps * 1e3
    iters = args.steps / dt
    gups = w.voxel_ray_updates_per_iteration() * iters / 1e9
    log(f"[bench] {args.steps} iterations in {dt:.3f} s: {iters:.2f} it/s, {gups:.1f} GUPS; "
        f"fp {tim['fp_ms']:.1f} ms, res {tim['res_ms']:.1f} ms, bp prep {tim['bpprep_ms']:.1f} ms, "
        f"bp {tim['bp_ms']:.1f} ms, other {tim['other_ms']:.1f} ms; fused subsets {tim['fused_subsets']}")

    kernels = kernel_table(w, tim, args.steps)
    roofline = roofline_of(w, kernels, tim, args.steps, traffic)
    alg = WL.algorithmic_bytes_per_iteration(w)
    algo_name = "OS-SART" if w.update else "SIRT"
    out = {
        "metric": METRIC,
        "value": iters,
        "unit": f"{algo_name} iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, analytic FP64 projections + 1% noise, generated on device)",
        "config": {"workload": w.name, "volume": [w.W, w.H, w.D], "projections": w.n,
                   "detector": [w.W, w.H], "tilts_deg": [w.tilts[0], w.tilts[-1], w.tilts[1] - w.tilts[0]],
                   "algorithm": algo_name, "update": w.update, "subsets": w.subsets(),
                   "relaxation": w.relaxation, "first_index": 0, "ellipsoids": w.ellipsoids},
        "gups": gups,
        "roofline": roofline,
        # counter-measured HBM bytes of the whole iteration over its time: the
        # honest bandwidth figure of the path (it is bound by VALU issue and
        # LDS, not by HBM)
        "hbm_iteration": hbm_iteration(w, tim, args.steps, ms, traffic),
        # the per-projection slab-streaming model: it counts one volume
        # read per projection, which the LDS kernels never make (they read the
        # volume about once per subset), so it is a model figure, above the
        # bytes that cross HBM, not a measurement of HBM use; it carries no
        # fraction of the peak (at 512^3 the modelled rate exceeds the peak)
        "iteration_model": {"model_bytes": alg, "model_GBs": alg * iters / 1e9,
                            "compulsory_bytes": WL.compulsory_bytes_per_iteration(w),
                            "note": "model, not measured traffic: counts one volume read per projection"},
        "kernels": kernels,
        "fused_subsets": tim["fused_subsets"],
        "final_residual": float(err[-1, 0]),
    }
    log(f"[bench] dominant kernel {roofline['kernel']}: {roofline['avg_launch_ms']:.4f} ms/launch, "
        f"{roofline['achieved']:.0f} GB/s algorithmic ({roofline['frac']:.3f} of HBM peak)" +
        (f", VALU {roofline['valu']['frac']:.3f}" if "valu" in roofline else "") +
        (f"; iteration HBM (counters) {out['hbm_iteration']['achieved_GBs']:.0f} GB/s = "
         f"{out['hbm_iteration']['frac']:.3f}" if out["hbm_iteration"] else ""))

    if not args.no_extras:
        # weighted backprojection pass (Radermacher weighting, diameter 0.5*W) on the same data
        P.reset_volume()
        P.setWeightingFilter(0.5 * w.W, w.tilts)
        P.run_wbp()
        t0 = time.perf_counter()
        P.run_wbp()
        out["wbp_radermacher_ms"] = (time.perf_counter() - t0) * 1e3
        P.setRampFilter(hamming=True, cutoff=1.0)
        t0 = time.perf_counter()
        P.run_wbp()
        out["wbp_ramp_hamming_ms"] = (time.perf_counter() - t0) * 1e3
        log(f"[bench] WBP: radermacher {out['wbp_radermacher_ms']:.1f} ms, "
            f"ramp+hamming {out['wbp_ramp_hamming_ms']:.1f} ms")
        # the drop-in boundary's own cost: TomoReconstruction2.OSSART's loop made
        # call by call through the Projector interface (per subset 13
        # addProjection, projectAndDifference, backproject and
        # clearAllProjections over the C ABI, each synchronous) against the
        # library's own loop above
        from tomoj_amd import reconstruction as REC
        P.reset_volume()
        REC.ossart(P, E, 1, w.relaxation, w.update)
        P.reset_volume()
        k = 3
        t0 = time.perf_counter()
        REC.ossart(P, E, k, w.relaxation, w.update)
        dts = time.perf_counter() - t0
        out["spi_ossart"] = {"iters_per_s": k / dts, "ms_per_iteration": dts / k * 1e3, "vs_library_loop": (k / dts) / iters,
                             "calls": "per subset: addProjection x update, projectAndDifference, backproject, "
                                      "clearAllProjections (C ABI, synchronous)"}
        log(f"[bench] OS-SART through the Projector interface: {k / dts:.1f} it/s "
            f"({out['spi_ossart']['vs_library_loop']:.2f} x the library loop)")
        # literal 512^3 SIRT, with its own kernel times and roofline
        c = WL.SIRT_512_CUBE
        if c.name != w.name:
            Q, _, _, dtc, timc = run_gpu(c, 3, 1, tj)
            kc = kernel_table(c, timc, 3)
            out["sirt_512cube"] = {"iters_per_s": 3 / dtc,
                                   "ms_per_iteration": dtc / 3 * 1e3,
                                   "gups": c.voxel_ray_updates_per_iteration() * 3 / dtc / 1e9,
                                   "roofline": roofline_of(c, kc, timc, 3, traffic),
                                   "hbm_iteration": hbm_iteration(c, timc, 3, dtc / 3 * 1e3, traffic),
                                   "iteration_model_GBs": WL.algorithmic_bytes_per_iteration(c) * 3 / dtc / 1e9,
                                   "kernels": kc}
            log(f"[bench] SIRT 512^3: {3 / dtc:.2f} it/s")
            Q.close()
        out["regularisation"] = regularisation_extras(w, P, tj)
        r = out["regularisation"]
        log(f"[bench] {w.W}x{w.H}x{w.D}: SIRT {r['sirt_it_s']:.1f} it/s, TVM (5 TV steps) "
            f"{r['tvm_it_s']:.1f} it/s, Bayesian {r['bayesian_it_s']:.1f} it/s, BgART {r['bgart_it_s']:.1f} it/s, TV step {r['tv_step_ms']:.3f} ms "
            f"({r['tv_step_GBs']:.0f} GB/s), FSC {r['fsc_ms']:.1f} ms")

    if not args.no_forms:
        from tomoj_amd import forms
        t0 = time.perf_counter()
        out["precision_512cube"] = forms.precision_study(512, 512, 512, 121, iterations=10, refine=2,
                                                         ellipsoids=40, noise=0.01)
        modes = out["precision_512cube"]["modes"]
        log(f"[bench] precision 512^3 ({time.perf_counter() - t0:.1f} s): plain "
            f"{out['precision_512cube']['plain_f32_it_s']:.1f} it/s; " +
            ", ".join(f"{k} {v['it_s']:.1f} it/s x{v['vs_plain']:.2f} rel-L2 {v.get('rel_l2_vs_f64', float('nan')):.1e}"
                      for k, v in modes.items()))
        out["batched"] = {}
        for st, master in (("f32", False), ("bf16", False), ("bf16", True)):
            key = st + ("+master" if master else "")
            out["batched"][key] = forms.batched_run(64, 15, st, master)
            log(f"[bench] batched 64 jobs {key}: {out['batched'][key]['ms']:.1f} ms for 15 iterations, "
                f"{out['batched'][key]['gups']:.0f} GUPS")
    if not args.no_forms and not args.no_serving:
        # the serving configuration: 200 requests of a seeded Poisson stream at
        # 20/s (more work than one GPU does, so latency grows with queueing),
        # then 120 requests at 0.8 x the capacity that run measured (jobs per
        # second of device-busy time): the steady state below saturation
        out["serving"] = forms.serving_run(200, 20.0)
        sv = out["serving"]
        log(f"[bench] serving 200 requests at 20/s: p50 {sv['p50_s']:.2f} s, p99 {sv['p99_s']:.2f} s, "
            f"{sv['gups']:.0f} GUPS, capacity {sv['capacity_jobs_per_s']:.2f} jobs/s; by width " +
            ", ".join(f"{k}: {v['p50_s']:.2f}/{v['p99_s']:.2f} s" for k, v in sv["by_width"].items()))
        rate = 0.8 * sv["capacity_jobs_per_s"]
        out["serving_sustained"] = forms.serving_run(120, rate)
        ss = out["serving_sustained"]
        log(f"[bench] serving 120 requests at {rate:.2f}/s (0.8 x capacity): p50 {ss['p50_s']:.3f} s, "
            f"p99 {ss['p99_s']:.2f} s; by width " +
            ", ".join(f"{k}: {v['p50_s']:.3f}/{v['p99_s']:.2f} s" for k, v in ss["by_width"].items()))

    if args.copy_probe > 0:
        cms = tj._lib.copy_probe(args.copy_probe, repeats=3)
        out["copy_probe"] = {"bytes": args.copy_probe, "ms": cms,
                             "read_plus_write_GBs": 2 * args.copy_probe / (cms * 1e-3) / 1e9}
        log(f"[bench] copy probe {args.copy_probe} B: {cms:.3f} ms "
            f"({out['copy_probe']['read_plus_write_GBs']:.0f} GB/s read+write)")

    if not args.no_cpu_baseline:
        import oracle
        oracle.build()
        log("[bench] timing TomoJ's CPU algorithm (one iteration) ...")
        out["cpu_baseline"] = cpu_baseline(w, P, E, oracle)
        log(f"[bench] CPU: {out['cpu_baseline']['value']:.4f} it/s on {out['cpu_baseline']['cores']} cores")
    P.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
