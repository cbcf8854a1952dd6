#!/usr/bin/env python3
"""Throughput of the SPH particle loops on one MI355X.

A step is one SphBase::runOneStep of the solver over the whole particle set
(neighbour search, density/alpha, non-pressure forces, the density and
divergence solver loops, integration).  Inputs are resident in HBM before the
timed region.  Prints ONE JSON line on stdout; progress goes to stderr.

Default workload: DFSPH, two facing blocks of 126^3 fluid particles
(4,000,752) at 0.01 m in a closed 6x3x3 m box, CFL dt (0.4, dt_max 1e-3 s),
density tolerance 0.01 % average, divergence tolerance 1 kg m^-3 s^-1
average, with the learned warm start trained online during the untimed
steps; the plain solver and the previous-stiffness warm start run as A/B
arms over the same steps.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from graphics_amd import scenes, sphgpu as sg  # noqa: E402

METRIC = "particle-steps/s at <=0.1% density error; solver iters/step learned vs plain"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per fluid particle of each pass (FP32 SoA, vec3 = 12 B,
# scalar = 4 B, each array counted once per pass; neighbour lists not counted)
ALG_BYTES = {
    "df_pass_velocity": 44, "df_pass_density": 36, "df_pass_divergence": 36,
    "df_density_alpha": 20, "df_aadv": 40, "wc_pressure": 20, "wc_force": 56,
    "ii_sumdij": 32, "ii_pressure": 56, "ii_aadv_dii": 64, "ii_aii": 48, "ii_force": 56,
    "pci_aext": 56, "pci_density": 44, "pci_pressure": 64, "pci_integrate": 84,
}
# the solver-loop passes with an LDS-window build and a gathers-only build
WIN_BIT = {"df_pass_density": sg.WIN_DENSITY, "df_pass_divergence": sg.WIN_DIVERGENCE,
           "df_pass_velocity": sg.WIN_VELOCITY, "ii_sumdij": sg.WIN_SUMDIJ, "ii_pressure": sg.WIN_PRESSURE}
DOMINANT = {sg.DFSPH: "df_pass_divergence", sg.IISPH: "ii_pressure", sg.WCSPH: "wc_force",
            sg.PCISPH: "pci_pressure"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


COPY_PROBE_BYTES = 1 << 30


def traffic_file(workload: str) -> Path:
    return ROOT / "profiles" / f"{workload}_traffic.txt"


def read_traffic(workload: str, kernel: str, window: bool | None = None):
    """HBM bytes per launch of `kernel` measured by the FETCH_SIZE/WRITE_SIZE
    passes recorded in profiles/<workload>_traffic.txt (None if absent); for
    a solver-loop pass, of its LDS-window build (window) or its gathers-only
    build."""
    path = traffic_file(workload)
    if not path.exists():
        return None
    names = {"k_" + kernel}
    if window is not None:
        names.add(f"k_{kernel}<{'true' if window else 'false'}>")
    for line in path.read_text().splitlines():
        if line.startswith("{"):
            rec = json.loads(line)
            if rec.get("kernel") in names:
                return rec
    return None


def kernel_name(raw: str) -> str:
    """Short kernel name from a demangled or mangled (Itanium) symbol."""
    import re
    m = re.match(r"_ZN(?:\d+\w+?)*?\d+_GLOBAL__N_1(\d+)", raw)
    if m:
        k = int(m.group(1))
        start = m.end()
        return raw[start:start + k]
    name = raw.replace("sphgpu::", "").replace("(anonymous namespace)::", "").split("(")[0]
    return name[5:] if name.startswith("void ") else name


def summarize_trace(stats_csv: str, trace_csv: str, workload: str, command: str, timed: str | None = None) -> int:
    """Kernel table of a rocprofv3 --kernel-trace --stats run: all dispatches,
    and the dispatches that processed particles (solver iterations launched
    past convergence return at once; launches under 10 us at this size are
    those); writes profiles/<workload>_kernel_stats.txt.  `timed` =
    "KERNEL:N" adds the mean of that kernel's last N such dispatches (the
    timed steps of the profiled bench run, whose line reports N launches)."""
    import csv
    import collections

    per = collections.defaultdict(list)
    for r in csv.DictReader(open(trace_csv)):
        name = kernel_name(r["Kernel_Name"])
        per[name].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    total = sum(sum(v) for v in per.values())
    lines = [f"# kernel durations, {workload}", f"# command: {command}",
             f"# {'kernel':28s} {'calls':>6s} {'avg_us':>9s} {'pct':>6s} {'calls>=10us':>11s} {'avg_us>=10us':>12s}"]
    for name, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        real = [x for x in v if x >= 10.0]
        ra = sum(real) / len(real) if real else 0.0
        lines.append(f"  {name[:28]:28s} {len(v):6d} {sum(v) / len(v):9.2f} {100 * sum(v) / total:6.2f} "
                     f"{len(real):11d} {ra:12.2f}")
    if timed:
        kname, n = timed.rsplit(":", 1)
        allv = per.get(kname, [])
        if allv:
            # iterations launched past convergence return early; at this
            # size they take well under a quarter of a working launch
            thr = max(10.0, 0.25 * sorted(allv)[len(allv) // 2])
            real = [x for x in allv if x >= thr][-int(n):]
            lines.append(f"# timed steps: last {len(real)} dispatches of {kname} of {thr:.0f} us or more "
                         f"(a quarter of its median): mean {sum(real) / len(real):.2f} us")
    out = ROOT / "profiles" / f"{workload}_kernel_stats.txt"
    out.write_text("\n".join(lines) + "\n")
    print(out)
    return 0


def summarize_pmc(fetch_csv: str, write_csv: str, workload: str, command: str) -> int:
    """Per-kernel median FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of two
    rocprofv3 --pmc runs (or of several pairs: comma-separated CSV lists, e.g.
    runs with and without LDS windows), scaled by what the counters read for
    the copy probe (k_copy_probe, COPY_PROBE_BYTES read and written); writes
    profiles/<workload>_traffic.txt."""
    import csv
    import collections

    def load(paths):
        d = collections.defaultdict(list)
        for path in paths.split(","):
            for r in csv.DictReader(open(path)):
                name = kernel_name(r["Kernel_Name"])
                dur = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
                d[name].append((float(r["Counter_Value"]), dur, int(r["Grid_Size"])))
        return d

    f, w = load(fetch_csv), load(write_csv)
    kib = COPY_PROBE_BYTES / 1024.0
    probe_f = max(v[0] for v in f["k_copy_probe"])
    probe_w = max(v[0] for v in w["k_copy_probe"])
    cf, cw = kib / probe_f, kib / probe_w
    lines = [f"# HBM traffic per dispatch, {workload}",
             f"# command: {command}",
             f"# copy probe {COPY_PROBE_BYTES} B: FETCH_SIZE {probe_f:.1f} KiB, WRITE_SIZE {probe_w:.1f} KiB;"
             f" scale factors fetch {cf:.4f}, write {cw:.4f}",
             "# dispatches that processed particles only (launches shorter than 10 us are skipped"
             " iterations past convergence); medians over dispatches"]
    for name in sorted(f, key=lambda k: -sum(v[0] for v in f[k])):
        if name not in w or name == "k_copy_probe":
            continue
        fv = sorted(v[0] for v in f[name] if v[1] >= 10000) or sorted(v[0] for v in f[name])
        wv = sorted(v[0] for v in w[name] if v[1] >= 10000) or sorted(v[0] for v in w[name])
        med = lambda a: a[len(a) // 2]
        rec = {"kernel": name, "dispatches": len(fv), "fetch_kib": round(med(fv), 1),
               "write_kib": round(med(wv), 1),
               "traffic_bytes": round((med(fv) * cf + med(wv) * cw) * 1024.0)}
        lines.append(json.dumps(rec))
    traffic_file(workload).write_text("\n".join(lines) + "\n")
    print(traffic_file(workload))
    return 0


def build_scene(name: str):
    return scenes.WORKLOADS[name]()


def run_gpu(sc, warm_start: int, steps: int, warmup: int, profile: str | None, graph: int = 0,
            stats_every: int = 1, capture_at: int | None = None, label: str = "", train: int = 0,
            lds_windows: str = "auto"):
    """One arm: `train` untimed steps in which a learned warm start trains its
    network (every arm runs them, so that all arms time the same steps of the
    simulation), `warmup` untimed steps, then `steps` timed steps.  Returns
    per-step statistics of the untimed and the timed steps; with capture_at
    the fluid state after that many untimed steps (for the CPU baseline)."""
    sim = scenes.make_simulation(sc, warm_start=warm_start, use_graph=graph)
    if lds_windows != "auto":
        sim.set_lds_windows(lds_windows == "on")
    if sc.rigids:
        # the host rigid-body step (Bullet's place in the reference) inside
        # every step, at the updateSolids point; part of the timed work
        world = scenes.rigid_world(sc)

        def update_solids(s):
            world.step(s.step_dt(), s.solid_force_torque)
            world.apply(s.set_solid_pose)

        sim.set_update_solids(update_solids)
    t0 = time.perf_counter()
    sim.init()
    log(f"[bench] {label} init {time.perf_counter() - t0:.2f}s  fluid={sc.num_fluid} boundary={sc.num_boundary}")
    warm_times, warm_stats, captured = [], [], None
    t_warm = time.perf_counter()
    train_elapsed = 0.0
    for k in range(train + warmup):
        t = time.perf_counter()
        warm_stats.append(sim.step())
        warm_times.append(time.perf_counter() - t)
        if k + 1 == train:
            train_elapsed = time.perf_counter() - t_warm
        if capture_at is not None and k + 1 == capture_at:
            captured = dict(x=sim.get(sg.F_POS), v=sim.get(sg.F_VEL), ids=sim.ids(),
                            phase=sim.get(sg.F_PHASE).astype(np.int64), step=k + 1)
        if k % 25 == 0 or k == train + warmup - 1:
            s = warm_stats[-1]
            what = "train" if k < train else "warmup"
            log(f"[bench] {label} {what} {k}: {warm_times[-1] * 1e3:.2f} ms it={s['iter_density']}/{s['iter_divergence']}")
    warm_elapsed = time.perf_counter() - t_warm
    if profile:
        sim.profile_kernel(profile)
    sim.synchronize()
    t0 = time.perf_counter()
    stats = []
    for k in range(steps):
        s = sim.step(stats=(k + 1) % stats_every == 0 or k == steps - 1)
        if s is not None:
            stats.append(s)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    timing = sim.kernel_timing() if profile else (0, 0.0, 0.0)
    n_live = sim.num_fluid()
    tuning = sim.tuning()
    tuning["lists"] = sim.list_form()
    tuning["lds_windows"] = sim.lds_windows()
    sim.close()
    return dict(elapsed=elapsed, stats=stats, warm_times=warm_times, warm_stats=warm_stats,
                warm_elapsed=warm_elapsed, train_elapsed=train_elapsed, timing=timing, n_live=n_live,
                tuning=tuning, captured=captured)


def arm_summary(run: dict, nf: int, steps: int, tol: float | None) -> dict:
    """Throughput and solver statistics of one arm over its timed steps."""
    st = run["stats"]
    mean = lambda k: float(np.mean([s[k] for s in st]))
    mx = lambda k: float(np.max([s[k] for s in st]))
    out = {"particle_steps_per_s": round(nf * steps / run["elapsed"], 2),
           "ms_per_step": round(run["elapsed"] / steps * 1e3, 4),
           "iter_density": mean("iter_density"), "iter_divergence": mean("iter_divergence"),
           "max_iter_hit_rate": mean("max_iter_hit"),
           "density_err_mean": mean("density_err"), "density_err_max": mx("density_err"),
           "divergence_err_mean": mean("divergence_err"), "divergence_err_max": mx("divergence_err"),
           "rho_mean_err": mean("rho_mean_err"),
           "too_many_neigb": int(sum(s["too_many_neigb"] for s in st)),
           "list_overflow": int(sum(s["list_overflow"] for s in st)),
           "zero_dis": int(sum(s["zero_dis"] for s in st)), "removed": int(sum(s["removed"] for s in st)),
           "untimed_s": round(run["warm_elapsed"], 3), "train_s": round(run["train_elapsed"], 3)}
    if tol is not None:
        out["density_err_within_tol"] = bool(out["density_err_max"] <= tol)
    return out


def cpu_threads() -> int:
    """Host threads the CPU baseline may use: the cores this process may run
    on (omp_get_num_procs() of the reference, SphBase.cpp:242-243), within
    OMP_NUM_THREADS when the environment sets a per-GPU share."""
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(ncpu, int(omp))) if omp.isdigit() and int(omp) > 0 else ncpu


def cpu_baseline(sc, state, budget_s: float):
    """The FP32 oracle (the reference's loop structure and algorithm, OpenMP)
    built for this host's CPU, timed on the same workload: from the GPU's
    state after `state['step']` steps where one is given, else from the
    initial state; as many steps as fit in budget_s (at least one)."""
    import tempfile
    from oracle import oracle as orc
    isa = "native"
    try:
        orc.build_native(tempfile.mkdtemp(prefix="oracle_native_"))
    except (OSError, subprocess.CalledProcessError, RuntimeError) as e:
        isa = f"x86-64-v3 (native build failed: {e})"
    threads = cpu_threads()
    orc.set_threads(threads)
    # the reference algorithm: no warm start
    if state is not None:
        o = orc.from_state(sc, state["x"], state["v"], state["ids"], state["phase"], 32, warm_start=0)
        first = state["step"]
    else:
        o = orc.from_scene(sc, 32, warm_start=0)
        first = 0
    per_step, iters = [], []
    t0 = time.perf_counter()
    while True:
        t = time.perf_counter()
        s = o.step()
        per_step.append(time.perf_counter() - t)
        iters.append((s["iter_density"], s["iter_divergence"]))
        log(f"[bench] cpu step {first + len(per_step) - 1}: {per_step[-1]:.2f}s it={iters[-1]}")
        if time.perf_counter() - t0 >= budget_s or len(per_step) >= 50:
            break
    elapsed = time.perf_counter() - t0
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    o.close()
    return dict(steps=len(per_step), elapsed=elapsed, threads=threads, model=model, per_step=per_step,
                first=first, isa=isa, iters=iters)


def surface_bench(args) -> int:
    """Surface reconstruction of one frame of the 1M-particle DFSPH dam (after
    `--warmup` simulation steps) with the reference surfacer's parameters
    (re.c: r_h = 2 r_p, node spacing r_p / 4, iso 0.5): a "step" is the
    anisotropic field plus marching cubes of that frame; value = particles
    surfaced per second."""
    sc = scenes.dfsph_c2()
    sim = scenes.make_simulation(sc, warm_start=sg.WARM_PREVIOUS)
    for k in range(args.warmup):
        sim.step(stats=(k + 1) % 50 == 0)
    x = sim.get(sg.F_POS)
    sim.close()
    s = sc.params["spacing_r"]
    mn, mx = x.min(0), x.max(0)
    sf = sg.Surfacer(mn, mx, s, 2 * s, s / 4)
    log(f"[bench] frame after {args.warmup} steps: {len(x)} particles, field grid {sf.grid_num}")
    sf.aniso(x)                         # warm-up run (allocations, code load)
    times, ntri = [], 0
    for _ in range(args.steps):
        t0 = time.perf_counter()
        sf.aniso(x)
        tris = sf.marching_cubes(0.5)
        times.append(time.perf_counter() - t0)
        ntri = len(tris)
    sf.close()
    ms = float(np.mean(times)) * 1e3
    value = len(x) / (ms * 1e-3)
    log(f"[bench] surface: {ms:.1f} ms per frame, {ntri} triangles")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        # serial FP64 restatement (the reference's loop is serial too) on a
        # sample: the particles of a sub-box holding about 3000 of them
        lo = np.quantile(x, 0.45, axis=0)
        sel = x[np.all((x >= lo) & (x <= lo + 0.15), axis=1)][:3000]
        t0 = time.perf_counter()
        orc.surface_field(sel.min(0), sel.max(0), s, 2 * s, s / 4, sel, True)
        el = time.perf_counter() - t0
        cpu = {"value": round(len(sel) / el, 2), "unit": "particles/s", "cores": 1, "kind": "port",
               "sample": f"anisotropic field of {len(sel)} particles of the same frame, FP64 oracle, serial"}
    out = {"metric": "particles surfaced per second (anisotropic field + marching cubes)", "value": round(value, 1),
           "unit": "particles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"surface_dam_{len(x)}", "grid": list(map(int, sf.grid_num)),
                      "triangles": ntri, "radius_p": s, "radius_h": 2 * s, "cell": s / 4, "iso": 0.5},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps before the timed ones (after the training steps of a learned "
                         "warm start)")
    ap.add_argument("--workload", default="dfsph_c3", choices=sorted(scenes.WORKLOADS) + ["surface_dam"])
    ap.add_argument("--warm-start", default="learned", choices=["none", "previous", "learned"])
    ap.add_argument("--no-ab", action="store_true", help="skip the plain-solver arm")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--cpu-from-step", type=int, default=10,
                    help="the CPU baseline starts from the GPU plain arm's state after this many steps")
    ap.add_argument("--profile-kernel", default=None)
    ap.add_argument("--lds-windows", default="auto", choices=["auto", "on", "off"],
                    help="LDS windows in every solver-loop pass (on), in none (off), or per pass as the autotuner times them")
    ap.add_argument("--stats-every", type=int, default=1,
                    help="read statistics back every N timed steps (the others do not wait for the device)")
    ap.add_argument("--graph", action="store_true",
                    help="replay captured HIP graphs of whole steps (no per-kernel timing)")
    ap.add_argument("--summarize-pmc", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "COMMAND"),
                    help="write profiles/<workload>_traffic.txt from two rocprofv3 --pmc runs")
    ap.add_argument("--timed", default=None, metavar="KERNEL:N",
                    help="with --summarize-trace: mean of the kernel's last N dispatches (the timed steps)")
    ap.add_argument("--summarize-trace", nargs=2, metavar=("TRACE_CSV", "COMMAND"),
                    help="write profiles/<workload>_kernel_stats.txt from a rocprofv3 --kernel-trace run")
    args = ap.parse_args()
    wname = (lambda: "surface_dam_1000000") if args.workload == "surface_dam" else \
        (lambda: build_scene(args.workload).name)
    if args.summarize_pmc:
        return summarize_pmc(*args.summarize_pmc[:2], wname(), args.summarize_pmc[2])
    if args.summarize_trace:
        return summarize_trace(None, args.summarize_trace[0], wname(), args.summarize_trace[1], args.timed)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the path runs on one "
              "GPU's share; sharding across GPUs is a design, not built", file=sys.stderr)
        return 2
    if args.workload == "surface_dam":
        if not sg.device_available():
            print("surface_dam needs one gfx9 device", file=sys.stderr)
            return 3
        if args.warmup is None:
            args.warmup = 300
        return surface_bench(args)
    if not sg.device_available():
        print("no gfx9 HIP device visible", file=sys.stderr)
        return 3

    # plain copy of 1 GiB: counter calibration and the copy bandwidth this
    # device reaches
    probe_ms = sg.copy_probe(COPY_PROBE_BYTES)
    copy_gbs = 2.0 * COPY_PROBE_BYTES / (probe_ms * 1e-3) / 1e9
    log(f"[bench] copy probe: {COPY_PROBE_BYTES >> 20} MiB in {probe_ms:.3f} ms = {copy_gbs:.0f} GB/s (read+write)")

    sc = build_scene(args.workload)
    solver = sc.params["solver"]
    warm = {"none": sg.WARM_NONE, "previous": sg.WARM_PREVIOUS, "learned": sg.WARM_LEARNED}[args.warm_start]
    if warm == sg.WARM_LEARNED and solver not in (sg.DFSPH, sg.IISPH):
        warm = sg.WARM_PREVIOUS
    learn_steps = int(sc.params.get("learn_steps", 40))
    # the learned arm trains its network online during its first learn_steps
    # steps; these run untimed before the warm-up, in every arm
    train = learn_steps if warm == sg.WARM_LEARNED else 0
    prof = args.profile_kernel or DOMINANT[solver]
    tol = sc.params.get("density_tol")
    labels = {sg.WARM_NONE: "plain", sg.WARM_PREVIOUS: "previous", sg.WARM_LEARNED: "learned"}
    # the CPU baseline starts from the plain arm's state (the reference algorithm)
    want_cpu = not args.no_cpu and sc.num_scenes == 1
    cap_main = args.cpu_from_step if (want_cpu and warm == sg.WARM_NONE and not sc.rigids
                                      and args.cpu_from_step <= train + args.warmup) else None
    run = run_gpu(sc, warm, args.steps, args.warmup, None if args.graph else prof, int(args.graph),
                  max(1, args.stats_every), capture_at=cap_main, label=labels[warm], train=train,
                  lds_windows=args.lds_windows)
    nf = sc.num_fluid
    value = nf * args.steps / run["elapsed"]
    ms = run["elapsed"] / args.steps * 1e3
    main_sum = arm_summary(run, nf, args.steps, tol)
    log(f"[bench] {labels[warm]}: {ms:.3f} ms/step, {value:.4g} particle-steps/s, "
        f"iters {main_sum['iter_density']:.2f}/{main_sum['iter_divergence']:.2f}, "
        f"max density err {main_sum['density_err_max']:.3g}")

    launches, tot_ms, tot_bytes = run["timing"]
    roofline = None
    if launches:
        avg_s = tot_ms / launches * 1e-3
        alg = ALG_BYTES[prof] * nf
        achieved = alg / avg_s / 1e9
        bit = WIN_BIT.get(prof)
        lw = (run["tuning"] or {}).get("lds_windows")
        window = bool(lw["mask"] & bit) if (bit and lw) else None
        tr = read_traffic(sc.name, prof, window)
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                    "traffic": tr["traffic_bytes"] if tr else None,
                    "kernel": prof, "lds_window": window, "avg_us": round(avg_s * 1e6, 3), "launches": launches,
                    "alg_bytes_per_launch": alg,
                    "traffic_achieved_GBs": round(tr["traffic_bytes"] / avg_s / 1e9, 1) if tr else None,
                    "copy_probe_GBs": round(copy_gbs, 1)}

    arms = {labels[warm]: main_sum}
    captured = run["captured"]
    if not args.no_ab and warm != sg.WARM_NONE:
        # same workload, same steps: the plain solver (the reference) and, for
        # the learned guess, the previous-stiffness guess it competes with
        ke = np.array([s["kinetic_energy"] for s in run["warm_stats"] + run["stats"]])
        # (IISPH has no previous-stiffness guess of its own: its plain start
        # is the reference's 0.5 p(t-dt))
        others = [sg.WARM_NONE] + ([sg.WARM_PREVIOUS] if warm == sg.WARM_LEARNED and solver == sg.DFSPH else [])
        for w in others:
            cap = args.cpu_from_step if (w == sg.WARM_NONE and want_cpu and not sc.rigids
                                         and args.cpu_from_step <= train + args.warmup) else None
            arm = run_gpu(sc, w, args.steps, args.warmup, None, capture_at=cap, label=labels[w], train=train,
                          lds_windows=args.lds_windows)
            if cap is not None:
                captured = arm["captured"]
            a = arm_summary(arm, nf, args.steps, tol)
            ke_a = np.array([s["kinetic_energy"] for s in arm["warm_stats"] + arm["stats"]])
            a["max_kinetic_energy_dev"] = float(np.abs(ke - ke_a).max() / max(ke_a.max(), 1e-30))
            arms[labels[w]] = a
            log(f"[bench] {labels[w]} arm: {a['ms_per_step']:.3f} ms/step, iters "
                f"{a['iter_density']:.2f}/{a['iter_divergence']:.2f}, cap hit {a['max_iter_hit_rate']:.2f}")
    iters = {"warm_start": labels[warm], "density": main_sum["iter_density"],
             "divergence": main_sum["iter_divergence"]}
    if "plain" in arms and warm != sg.WARM_NONE:
        tot = lambda a: a["iter_density"] + a["iter_divergence"]
        iters["reduction_vs_plain"] = round(1.0 - tot(main_sum) / tot(arms["plain"]), 4)
        pv = arms["plain"]["iter_divergence"]
        iters["divergence_reduction_vs_plain"] = round(1.0 - main_sum["iter_divergence"] / pv, 4) if pv else None
        if "previous" in arms and warm == sg.WARM_LEARNED:
            iters["reduction_vs_previous"] = round(1.0 - tot(main_sum) / tot(arms["previous"]), 4)

    cpu = None
    if want_cpu:
        c = cpu_baseline(sc, captured, args.cpu_seconds)
        cpu_rate = nf * c["steps"] / c["elapsed"]
        start = (f"the GPU plain arm's state after step {c['first']}" if c["first"] else "the initial state")
        cpu = {"value": round(cpu_rate, 2), "unit": "particle-steps/s", "cores": c["threads"],
               "kind": "port",
               "sample": (f"steps {c['first']}-{c['first'] + c['steps'] - 1} of the same workload from {start}, "
                          f"FP32 oracle (the reference's loop structure and algorithm, no warm start; OpenMP, "
                          f"-march={c['isa']}) on {c['model']}"),
               "iters_per_step": [list(t) for t in c["iters"]],
               "speedup": round(value / cpu_rate, 1)}

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": sc.name, "solver": ["WCSPH", "DFSPH", "IISPH", "PCISPH"][solver],
                   "fluid": nf, "boundary": sc.num_boundary, "scenes": sc.num_scenes,
                   "density_tol": tol, "divergence_tol": sc.params.get("divergence_tol"),
                   "max_iter_density": sc.params.get("max_iter_density"),
                   "max_iter_div": sc.params.get("max_iter_div"),
                   "learn_steps": learn_steps if warm == sg.WARM_LEARNED else None,
                   "untimed_training_steps": train,
                   "training_s": round(run["train_elapsed"], 3),
                   "density_err_within_tol": main_sum.get("density_err_within_tol"),
                   "live_after": run["n_live"], "hip_graph": bool(args.graph),
                   "stats_every": max(1, args.stats_every)},
        "iters_per_step": iters,
        "arms": arms,
        "autotune": run["tuning"],
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if args.lds_windows != "auto":
        out["config"]["lds_windows"] = args.lds_windows
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
