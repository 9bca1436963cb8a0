#!/usr/bin/env python3
"""Benchmark of the PD force/update hot path on one MI355X.

Workload (default): the pre-notched PMB brittle-fracture plate under edge
impact -- 200 x 400 x 18 = 1.44M particles, dx = 5e-4, delta = 3.01 dx,
E = 191 GPa, nu = 0.25, rho = 8000, G0 = 42408 J/m^2, two 0.05 m notches,
16.5 m/s impact, dt = 7.5e-8 (0.85 x the reference's critical step estimate).

A "step" is one velocity-Verlet time step (force pass with breaks + fused
integrator).  The metric is bond evaluations per second: directed bonds
(broken ones included) x steps / wall time, of the mixed-precision form,
reported beside the plain FP64 form on the same specimen; the accuracy of the
mixed form against FP64 (relative displacement error, broken-bond set) is
part of the line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cabanapd_amd import configs, engine  # noqa: E402
from cabanapd_amd._abi import KCLS_FORCE  # noqa: E402

HBM_PEAK_GBS = 8000.0
DEFAULT_CELLS = (200, 400, 18)

PROFILES = os.path.join(ROOT, "profiles")


def measured_counters(kernel, cells):
    """HBM traffic per launch and the VALU-issue share of a kernel, as the
    last rocprofv3 counter passes on this build measured them: the lines
    "traffic <kernel> <cells> <bytes>" and "valu_busy <kernel> <cells> <frac>"
    of the counter summaries under profiles/ (FETCH_SIZE x 2.0 + WRITE_SIZE x
    1.0 per launch, the factors calibrated on a plain device copy).  None when
    no summary covers this kernel on this workload."""
    import re
    tag = "%dx%dx%d" % tuple(cells)
    out = {"traffic": None, "valu_busy": None, "sca_busy": None, "source": None}
    if not os.path.isdir(PROFILES):
        return out
    for name in sorted(os.listdir(PROFILES)):
        if not name.endswith(".md"):
            continue
        for ln in open(os.path.join(PROFILES, name)):
            m = re.match(r"^(traffic|valu_busy|sca_busy) (.+) (\d+x\d+x\d+) ([0-9.eE+-]+)\s*$", ln.strip())
            if m and m.group(2) == kernel and m.group(3) == tag:
                out[m.group(1)] = float(m.group(4))
                out["source"] = "profiles/" + name
    return out


def roofline_note(ctr, avg_ms):
    """What the counter passes say bounds the kernel, in words."""
    base = "frac is the algorithmic-bytes yardstick against HBM peak"
    if ctr["traffic"] is None or ctr["valu_busy"] is None:
        return base + "; no counter summary covers this kernel on this workload"
    hbm = ctr["traffic"] / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    parts = ["VALU issue %.2f" % ctr["valu_busy"]]
    if ctr["sca_busy"] is not None:
        parts.append("scalar issue %.2f" % ctr["sca_busy"])
    parts.append("measured HBM traffic %.2f of peak" % hbm)
    top = max(ctr["valu_busy"], ctr["sca_busy"] or 0.0, hbm)
    verdict = ("bound by VALU issue" if top == ctr["valu_busy"] and top >= 0.8 else
               "bound by HBM" if top == hbm and top >= 0.8 else
               "no unit saturated: bound by issue and load latency")
    return "%s; counters (share of launch cycles): %s -- %s" % (base, ", ".join(parts), verdict)


def host_cores():
    """Physical cores of the host (lscpu: sockets x cores per socket) and the
    CPUs this process may run on."""
    import subprocess
    phys = None
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict((a.strip(), b.strip()) for a, b in (ln.split(":", 1) for ln in txt.splitlines() if ":" in ln))
        phys = int(kv["Socket(s)"]) * int(kv["Core(s) per socket"])
    except Exception:
        pass
    return phys, len(os.sched_getaffinity(0))


def algorithmic_bytes_per_step(n_particles, bonds, passes=1, fracture=True, lps=False):
    """B_alg = P B (4 + f/8) + N (24 + 48 + 48 + 16 L): int32 neighbor index
    and one broken bit per bond per pass; x read, u and v read + write per
    particle (and m, theta for LPS)."""
    f = 1.0 if fracture else 0.0
    return passes * bonds * (4.0 + f / 8.0) + n_particles * (24 + 48 + 48 + (16 if lps else 0))


def make_spec(ctx, w, precision, bond_layout=0):
    d = type(w["desc"]).from_buffer_copy(w["desc"])
    d.precision = {"fp64": 0, "mixed": 1}[precision]
    d.bond_layout = bond_layout
    t0 = time.perf_counter()
    s = engine.Spec(ctx, d, w["x"].shape[0])
    configs.upload(s, w)
    configs.setup(s, w)
    s.sync()
    s.setup_s = time.perf_counter() - t0
    return s


def timed_run(spec, warmup, steps):
    spec.step(warmup) if warmup else None
    spec.sync()
    spec.timing(True)
    t0 = time.perf_counter()
    spec.step(steps)
    spec.sync()
    t1 = time.perf_counter()
    force_ms, launches = spec.kernel_time(KCLS_FORCE)
    spec.timing(False)
    return t1 - t0, force_ms, launches


def cpu_baseline_child(cells, budget_s):
    """Runs in a child process whose OpenMP environment is set before the
    oracle loads: the oracle on the first steps of the same specimen."""
    from oracle import binding as orc
    threads = int(os.environ["OMP_NUM_THREADS"])
    orc.set_threads(threads)
    w = configs.kw_plate(num_cells=cells, steps=1)
    d = type(w["desc"]).from_buffer_copy(w["desc"])
    d.precision = 0
    o = orc.OracleSpec(d, w["x"].shape[0], w["x"], w["vol"], u=w["u"], v=w["v"], rho=w["rho"])
    configs.setup(o, w)
    bonds = o.bond_stats()[0]
    t0 = time.perf_counter()
    o.step(1)
    dt1 = time.perf_counter() - t0
    more = max(0, min(20, int(budget_s / max(dt1, 1e-3)) - 1))
    t0 = time.perf_counter()
    if more:
        o.step(more)
    dtm = time.perf_counter() - t0
    o.close()
    print(json.dumps({"bonds": bonds, "steps": 1 + more, "seconds": dt1 + dtm}))


def cpu_baseline(cells, budget_s=12.0):
    """The oracle (a port of the reference's CPU path: one OpenMP thread per
    particle chunk, serial neighbor loop) timed on this host in a child
    process with OMP_PROC_BIND=close and OMP_PLACES=cores.  Threads: the CPUs
    this process may use, capped by OMP_NUM_THREADS when the environment sets
    it (a GPU box allots 16 host CPUs per GPU)."""
    import subprocess
    phys, avail = host_cores()
    threads = avail
    if os.environ.get("OMP_NUM_THREADS"):
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    p = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", ",".join(map(str, cells)),
                        "--cpu-budget", str(budget_s)], env=env, capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        return {"error": p.stderr.strip()[-500:]}
    r = json.loads(p.stdout.strip().splitlines()[-1])
    return {"value": r["bonds"] * r["steps"] / r["seconds"], "unit": "bond-evals/s", "cores": threads,
            "kind": "port", "threads": threads, "host_physical_cores": phys, "cpus_available": avail,
            "binding": "OMP_PROC_BIND=close OMP_PLACES=cores",
            "sample": "first %d steps of the same specimen (oracle, OpenMP, %d threads)" % (r["steps"], threads)}


def serving(num_requests=24, rate=50.0, seed=20261202, max_particles=2e6):
    """Serving stream: Poisson arrivals, specimens of uneven size.  Specimen
    setup (upload, bond build, prenotch, init) runs on two host worker threads
    as requests arrive.  The mixed form runs through the continuous-batching
    engine (specimens admitted first come, first served as slots free up, one
    packed launch per step for all running specimens); the plain FP64 form
    serves the same stream one specimen at a time.  Latency is arrival ->
    result, setup included."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    # unit-rate arrival times, scaled by each stream's rate
    arrivals_unit = np.cumsum(rng.exponential(1.0, size=num_requests))
    reqs = [configs.serving_request(rng, n_range=(2e4, max_particles)) for _ in range(num_requests)]

    def make(w, precision, ctx):
        d = type(w["desc"]).from_buffer_copy(w["desc"])
        d.precision = {"fp64": 0, "mixed": 1}[precision]
        s = engine.Spec(ctx, d, w["x"].shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        return s

    def summarize(lat, wall, bond_evals):
        lat = np.sort(np.asarray(lat))
        return {"solves_per_s": len(lat) / wall, "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "wall_s": wall, "bond_evals_per_s": bond_evals / wall}

    def stream(precision, batched, rate):
        arrivals = arrivals_unit / rate
        ctx = engine.Context(0)
        pool = ThreadPoolExecutor(max_workers=2)
        t0 = time.perf_counter()
        setups = {}
        arrived = 0
        started = 0
        running = {}
        lat = [0.0] * num_requests
        finished = 0
        bond_evals = 0.0
        while finished < num_requests:
            now = time.perf_counter() - t0
            while arrived < num_requests and arrivals[arrived] <= now:
                setups[arrived] = pool.submit(make, reqs[arrived], precision, ctx)
                arrived += 1
            # start specimens in arrival order once their setup is done
            while started < arrived and (setups[started].done() or not batched):
                s = setups.pop(started).result()
                bond_evals += s.bond_stats()[0] * reqs[started]["steps"]
                if batched:
                    running[started] = (s, ctx.submit(s, reqs[started]["steps"]))
                else:
                    s.step(reqs[started]["steps"])  # one specimen at a time
                    lat[started] = time.perf_counter() - t0 - arrivals[started]
                    s.close()
                    finished += 1
                    started += 1
                    break  # take in the arrivals of that solve first
                started += 1
            if running:
                ctx.progress(16)
                for k in list(running):
                    s, t = running[k]
                    if ctx.poll(t):
                        lat[k] = time.perf_counter() - t0 - arrivals[k]
                        s.close()
                        del running[k]
                        finished += 1
            elif finished < num_requests:
                nxt = arrivals[arrived] if arrived < num_requests else now + 1e-3
                time.sleep(min(1e-3, max(0.0, nxt - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        pool.shutdown()
        ctx.close()
        return summarize(lat, wall, bond_evals)

    def small_batch(graph_steps, n_range, count=16, steps=400):
        """count small specimens of n_range particles stepped together in the
        batch engine; packed-step time with the step replayed as a captured
        graph for up to graph_steps steps (0: eager launches)."""
        brng = np.random.default_rng(seed + 1)
        ws = [configs.serving_request(brng, n_range=n_range, steps_range=(steps, steps)) for _ in range(count)]
        ctx = engine.Context(0)
        ctx.set_graph(graph_steps)
        specs = [make(w, "mixed", ctx) for w in ws]
        # warm up: a few packed steps, then the timed run of `steps` steps
        def drain(tickets):
            ctx.progress(0)
            while not all(ctx.poll(t) for t in tickets):
                time.sleep(1e-4)

        drain([ctx.submit(s, 8) for s in specs])
        t0 = time.perf_counter()
        drain([ctx.submit(s, steps) for s in specs])
        dt = time.perf_counter() - t0
        npart = sum(w["x"].shape[0] for w in ws)
        for s in specs:
            s.close()
        ctx.close()
        return {"specimens": count, "particles": int(npart), "steps": steps, "graph_steps": graph_steps,
                "ms_per_step": 1e3 * dt / steps}

    mixed = stream("mixed", True, rate)
    fp64 = stream("fp64", False, rate)
    # the same requests offered below the batched capacity: latencies of a
    # sustained stream rather than of a queue draining
    sustained_rate = 5.0
    sustained = stream("mixed", True, sustained_rate)
    sustained["arrival_rate_per_s"] = sustained_rate
    sizes = [r["x"].shape[0] for r in reqs]
    bond_evals = mixed["bond_evals_per_s"] * mixed["wall_s"] / num_requests
    return {"requests": num_requests, "arrival_rate_per_s": rate, "seed": seed,
            "particles_min": int(min(sizes)), "particles_max": int(max(sizes)),
            "steps_total": int(sum(r["steps"] for r in reqs)),
            "bond_evals_per_request": bond_evals, "offered_bond_evals_per_s": bond_evals * rate,
            "mixed_batched": mixed, "fp64_per_specimen": fp64, "mixed_batched_sustained": sustained,
            "small_batch_graph": small_batch(64, (2e4, 4e4)), "small_batch_eager": small_batch(0, (2e4, 4e4)),
            "tiny_batch_graph": small_batch(64, (1e3, 2e3)), "tiny_batch_eager": small_batch(0, (1e3, 2e3))}


def copy_kernel(mib, repeats=5):
    """A plain device copy of a known size: rocprofv3's FETCH_SIZE / WRITE_SIZE
    for this kernel calibrate the counters (bytes moved / counter reading)."""
    import torch
    n = int(mib * (1 << 20)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(repeats):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel_bytes_read": n * 4, "copy_kernel_bytes_written": n * 4, "repeats": repeats}))


def rows_leg(args):
    """The rows beyond the benchmark path, each on the benchmark plate's
    geometry (200 x 400 x 18 by default): linearized PMB / LPS (no fracture),
    the stress and energy output passes, a device-resident region series (the
    crack-extent max over every particle), and the FP64 gather forms of two
    materials, ElasticPerfectlyPlastic, TemperatureDependent and heat
    transfer.  Per row: kernel time from HIP events on the specimen's stream,
    rates, the kernel's algorithmic HBM bytes against 8 TB/s (the same
    yardstick as the main line), and the oracle on the host cores over two
    steps (or calls) of the same workload."""
    from cabanapd_amd import _abi
    from cabanapd_amd._abi import KCLS_ENERGY, KCLS_LPS, KCLS_OTHER, make_desc
    from oracle import binding as orc
    from oracle import regions as oregions
    threads = min(int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1, 16)
    orc.set_threads(threads)
    cells = tuple(int(c) for c in args.cells.split(","))
    steps = min(args.steps, 200)
    base = configs.kw_plate(num_cells=cells, steps=steps)
    x = base["x"]
    n = x.shape[0]
    bd = base["desc"]
    ctx = engine.Context(0)
    out = {"workload": "rows on the plate %dx%dx%d (%d particles), %d steps each" % (cells + (n, steps)),
           "cpu_threads": threads, "rows": {}}

    def workload(model="PMB", fracture=True, **kw):
        w = dict(base)
        w["desc"] = make_desc(model, fracture=fracture, horizon=bd.horizon, K=bd.K, G=kw.pop("G", 0.0),
                              G0=bd.G0, dt=bd.dt, **kw)
        if not fracture:
            w["notches"] = None
        w["prescribed"] = []
        return w

    def frac_of(b_alg, ms):
        gbs = b_alg / (ms * 1e-3) / 1e9
        return {"achieved_GBs": gbs, "frac_of_8TBs": gbs / HBM_PEAK_GBS, "alg_bytes": b_alg}

    def oracle_rate(w, bonds, nsteps=2):
        d = type(w["desc"]).from_buffer_copy(w["desc"])
        d.precision = 0
        o = orc.OracleSpec(d, n, w["x"], w["vol"], u=w["u"], v=w["v"], rho=w["rho"])
        configs.setup(o, w)
        t0 = time.perf_counter()
        o.step(nsteps)
        dt_ = (time.perf_counter() - t0) / nsteps
        return o, {"ms_per_step": dt_ * 1e3, "bond_evals_per_s": bonds / dt_}

    def force_row(name, w, precisions=("fp64", "mixed"), lps=False, extra_cls=None):
        row = {}
        bonds = 0
        for prec in precisions:
            s = make_spec(ctx, w, prec)
            bonds, _ = s.bond_stats()
            s.step(5)
            s.sync()
            s.timing(True)
            t0 = time.perf_counter()
            s.step(steps)
            s.sync()
            wall = time.perf_counter() - t0
            fms, fl = s.kernel_time(KCLS_FORCE)
            r = {"ms_per_step": wall / steps * 1e3, "bond_evals_per_s": bonds * steps / wall,
                 "force_kernel": s.kernel_name(KCLS_FORCE), "force_kernel_ms": fms / max(fl, 1)}
            r.update(frac_of(algorithmic_bytes_per_step(n, bonds, fracture=w["desc"].fracture != 0, lps=lps),
                             r["force_kernel_ms"]))
            if lps:
                lms, ll = s.kernel_time(KCLS_LPS)
                r["state_kernel_ms"] = lms / max(ll, 1)
            if extra_cls is not None:
                oms, ol = s.kernel_time(extra_cls[0])
                r[extra_cls[1] + "_ms_total"] = oms
                r[extra_cls[1] + "_launches"] = ol
            s.timing(False)
            s.close()
            row[prec] = r
        _, row["cpu_oracle"] = oracle_rate(w, bonds)
        row["bonds"] = bonds
        out["rows"][name] = row
        return bonds

    # rank 1: linearized models (NoFracture only)
    force_row("linear_pmb", workload("LINEAR_PMB", fracture=False))
    force_row("linear_lps", workload("LINEAR_LPS", fracture=False, G=bd.K * 0.6), lps=True)

    # ranks 2-3: output passes and a region series on the fracturing plate (mixed spec)
    w = workload()
    s = make_spec(ctx, w, "mixed")
    bonds, _ = s.bond_stats()
    s.step(50)
    s.sync()
    s.timing(True)
    for _ in range(20):
        s.compute_stress()
    sms, sl = s.kernel_time(KCLS_OTHER)
    for _ in range(20):
        s.compute_energy_async()
    s.sync()
    ems, el = s.kernel_time(KCLS_ENERGY)
    s.timing(False)
    box = _abi.box_region(x.min(axis=0), x.max(axis=0))
    ser = s.series(box, "x", component=1, reducer="max", capacity=256, gate="phi", threshold=0.3)
    ser.update(0.0)
    s.sync()
    t0 = time.perf_counter()
    for k in range(200):
        ser.update(float(k))
    s.sync()
    upd = (time.perf_counter() - t0) / 200
    ser.close()
    s.close()
    o, _ = oracle_rate(w, bonds, nsteps=1)
    t0 = time.perf_counter()
    o.compute_stress()
    cpu_stress = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.compute_energy()
    cpu_energy = time.perf_counter() - t0
    ox, ophi = o.get("x"), o.get("phi")
    t0 = time.perf_counter()
    mem = oregions.members(ox, box)
    oregions.reduce(np.where(ophi[mem] > 0.3, ox[mem, 1], 0.0), "max")
    cpu_series = time.perf_counter() - t0
    stress_ms = sms / max(sl, 1)
    energy_ms = ems / max(el, 1)
    out["rows"]["stress"] = dict(kernel_ms=stress_ms, bond_evals_per_s=bonds / (stress_ms * 1e-3),
                                 cpu_oracle_ms=cpu_stress * 1e3,
                                 **frac_of(bonds * (4 + 1 / 8) + n * (24 + 24 + 72), stress_ms))
    out["rows"]["energy"] = dict(kernel_ms=energy_ms, bond_evals_per_s=bonds / (energy_ms * 1e-3),
                                 cpu_oracle_ms=cpu_energy * 1e3,
                                 **frac_of(bonds * (4 + 1 / 8) + n * (24 + 24 + 16), energy_ms))
    out["rows"]["series_update"] = dict(members=n, us_per_update=upd * 1e6, cpu_host_ms=cpu_series * 1e3,
                                        **frac_of(n * (8 + 8 + 8 + 8), upd * 1e3))

    # rank 4: gather forms (multi-material FP64 only)
    w = workload()
    mtype = (x[:, 0] >= 0.0).astype(np.int32)
    w["materials"] = ([(bd.K, 0.0, bd.G0), (2.0 * bd.K, 0.0, 2.0 * bd.G0)], mtype)
    force_row("multi_material", w)
    force_row("epp", workload(mechanics="epp", sigma_y=3.0 * bd.K * 1e-3))
    w = workload(thermal=True, alpha=1e-5, temp0=0.0)
    w["prescribed"] = [("temperature", np.arange(n), np.zeros(n), 1e7 * (x[:, 0] - x[:, 0].min()) / 0.1)]
    force_row("thermal", w)
    w = workload(thermal="dynamic", alpha=1e-5, temp0=0.0, kappa=50.0, cp=500.0, thermal_subcycle=10)
    w["temperature"] = 100.0 * (x[:, 0] - x[:, 0].min()) / 0.1
    hb = force_row("heat_transfer", w, extra_cls=(KCLS_OTHER, "heat_and_half_step"))
    r = out["rows"]["heat_transfer"]["fp64"]
    passes = steps // 10
    heat_ms = r["heat_and_half_step_ms_total"] / max(passes, 1)
    r["heat_pass_ms"] = heat_ms
    r["heat_pass"] = frac_of(hb * (4 + 1 / 8) + n * (24 + 8 + 8 + 8 + 8), heat_ms)
    return out


def lps_leg(args, n_side=128):
    """LPS elastic-brittle cube 128^3 (2.1M particles): weighted volume,
    dilatation and force passes every step, both precision forms."""
    from cabanapd_amd._abi import KCLS_LPS
    w = configs.lps_cube(n_side=n_side, steps=args.steps, fracture=True)
    n = w["x"].shape[0]
    ctx = engine.Context(0)
    res, specs = {}, {}
    for prec in ("fp64", "mixed"):
        s = make_spec(ctx, w, prec)
        bonds, _ = s.bond_stats()
        unroll = s.autotune(3) if args.unroll < 0 else (s.set_tuning(args.unroll) or args.unroll)
        s.step(args.warmup) if args.warmup else None
        s.sync()
        s.timing(True)
        t0 = time.perf_counter()
        s.step(args.steps)
        s.sync()
        wall = time.perf_counter() - t0
        fms, fl = s.kernel_time(KCLS_FORCE)
        lms, ll = s.kernel_time(KCLS_LPS)
        s.timing(False)
        res[prec] = dict(value=bonds * args.steps / wall, ms_per_step=wall / args.steps * 1e3,
                         force_kernel=s.kernel_name(KCLS_FORCE), force_avg_ms=fms / max(fl, 1),
                         state_kernel=s.kernel_name(KCLS_LPS), state_avg_ms=lms / max(ll, 1), unroll=unroll)
        specs[prec] = s
    u64, u32 = specs["fp64"].get("u"), specs["mixed"].get("u")
    den = float(np.max(np.abs(u64)))
    _, _, i64 = specs["fp64"].bonds()
    _, _, i32 = specs["mixed"].bonds()
    for s in specs.values():
        s.close()
    # The specimen's own sensitivity: plain FP64 from the initial displacement
    # rounded to FP32 once (a relative change <= 6e-8), same steps (untimed)
    w32 = dict(w)
    w32["u"] = w["u"].astype(np.float32).astype(np.float64)
    f = make_spec(ctx, w32, "fp64")
    f.step(args.warmup + args.steps)
    uf = f.get("u")
    _, _, i_f = f.bonds()
    f.close()
    ctx.close()
    b_alg = algorithmic_bytes_per_step(n, bonds, passes=3, fracture=True, lps=True)
    return {"workload": "LPS elastic-brittle cube %d^3 (%d particles, %d bonds), seeded perturbation, "
                        "+-10 MPa y-face traction, %d steps" % (n_side, n, bonds, args.steps),
            "metric": "bond-evals/s", "mixed": res["mixed"], "fp64": res["fp64"],
            "speedup": res["mixed"]["value"] / res["fp64"]["value"],
            "u_rel_err": float(np.max(np.abs(u32 - u64)) / den) if den > 0 else 0.0,
            "broken_bonds_fp64": int((i64 == 0).sum()), "broken_bond_mismatches": int((i64 != i32).sum()),
            "fp32_input_floor": {"u_rel_err": float(np.max(np.abs(uf - u64)) / den) if den > 0 else 0.0,
                                 "broken_bond_mismatches": int((i64 != i_f).sum()),
                                 "note": "plain FP64 started from u0 rounded to FP32, against plain FP64"},
            "alg_bytes_per_step": b_alg,
            "achieved_GBs_mixed": b_alg / (res["mixed"]["ms_per_step"] * 1e-3) / 1e9}


def cube_leg(args):
    """PMB elastic cube 41^3 (68,921 particles, 7,720,394 bonds), Gaussian
    radial displacement, dt 1e-7: both precision forms, with the captured
    step graphs and with every kernel launched on its own."""
    w = configs.elastic_cube(n_side=41, steps=args.steps)
    ctx = engine.Context(0)
    out = {}
    u = {}
    for prec in ("fp64", "mixed"):
        for graph in (True, False):
            s = make_spec(ctx, w, prec)
            bonds, _ = s.bond_stats()
            s.set_graph(graph)
            s.step(args.warmup) if args.warmup else None
            s.sync()
            t0 = time.perf_counter()
            s.step(args.steps)
            s.sync()
            wall = time.perf_counter() - t0
            out["%s_%s" % (prec, "graph" if graph else "eager")] = {
                "value": bonds * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
                "kernel": s.kernel_name(KCLS_FORCE)}
            if graph:
                u[prec] = s.get("u")
            s.close()
    ctx.close()
    den = float(np.max(np.abs(u["fp64"])))
    return {"workload": "PMB elastic cube 41^3 (68921 particles, %d bonds), %d steps" % (bonds, args.steps),
            "metric": "bond-evals/s", "runs": out,
            "u_rel_err": float(np.max(np.abs(u["mixed"] - u["fp64"])) / den)}


def dem_leg(args, n=1000000):
    """The granular box: 1e6 spheres, radii log-normal (sigma 0.1) clipped to
    +-30 %, compacted to a seeded random non-overlapping packing at solid
    fraction 0.55 (configs.dem_random_packing, untimed setup) inside a
    container of frozen wall spheres, Hertz-JKR (restitution 0.5, surface
    energy 0.05 J/m^2), gravity, contact list rebuilt every 20 steps.  The
    contact pass's algorithmic bytes per step: 4 B per listed pair (neighbor
    id) + 120 B per particle, walls included (x and u read, v read and
    written, f written)."""
    ctx = engine.Context(0)
    t0 = time.perf_counter()
    w = configs.dem_box(ctx, n=n, steps=args.steps, log=lambda m: print(m, file=sys.stderr, flush=True))
    pack_s = time.perf_counter() - t0
    s = engine.Spec(ctx, w["desc"], w["x"].shape[0])
    configs.upload(s, w)
    configs.setup(s, w)
    s.step(args.warmup) if args.warmup else None
    s.sync()
    s.timing(True)
    t0 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    fms, fl = s.kernel_time(KCLS_FORCE)
    bms, bl = s.kernel_time(3)
    oms, ol = s.kernel_time(4)
    kname = s.kernel_name(KCLS_FORCE)
    s.timing(False)
    pairs, maxn = s.bond_stats()
    frac = w["solid_fraction"]
    s.close()
    ctx.close()
    favg = fms / max(fl, 1)
    npart = w["x"].shape[0]
    b_alg = pairs * 4.0 + npart * 120.0
    achieved = b_alg / (favg * 1e-3) / 1e9
    return {"workload": "Hertz-JKR granular box, %d polydisperse spheres (radii log-normal sigma 0.1 clipped to "
                        "+-30 %%), random packing at solid fraction %.4f in a container of %d frozen wall spheres, "
                        "gravity, list rebuilt every 20 steps, %d steps" % (n, frac, w["wall_spheres"], args.steps),
            "particles": npart,
            "metric": "contact-pair evals/s", "value": pairs * args.steps / wall,
            "steps_per_s": args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
            "list_pairs": pairs, "max_neighbors": maxn, "solid_fraction": frac,
            "packing_stages": w["packing_stages"], "packing_setup_s": pack_s,
            "force_kernel": kname, "force_avg_ms": favg, "build_ms_total": bms, "builds": bl,
            "other_ms_total": oms,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "alg_bytes_per_launch": b_alg,
                         "traffic": measured_counters(kname, (npart, 1, 1))["traffic"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the workload's own count -- plate 1000, cube 500, lps 500, "
                         "dem 2000, rows 200)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cells", type=str, default="200,400,18")
    ap.add_argument("--unroll", type=int, default=-1,
                    help="force-pass variant: 0 row windows in LDS, 1/2/4/8 gathers in slot groups, -1 autotune")
    ap.add_argument("--requests", type=int, default=24, help="serving stream length (0 = skip serving)")
    ap.add_argument("--max-particles", type=float, default=2e6)
    ap.add_argument("--bond-layout", choices=("auto", "compact"), default="auto",
                    help="auto: stencil slots on full lattices; compact: stored neighbor ids")
    ap.add_argument("--workload", choices=("plate", "cube", "lps", "dem", "rows"), default="plate",
                    help="plate: the PMB fracture plate (the benchmark line); cube: PMB elastic cube 41^3; "
                         "lps: LPS fracture cube 128^3; dem: Hertz-JKR contact box of 1e6 spheres; "
                         "rows: linearized models, output passes, region series, multi-material, EPP, "
                         "thermal and heat transfer on the plate geometry")
    ap.add_argument("--cpu-baseline-child", type=str, default="", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help=argparse.SUPPRESS)
    ap.add_argument("--copy-kernel", type=float, default=0.0,
                    help="only run a device copy of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"plate": 1000, "cube": 500, "lps": 500, "dem": 2000, "rows": 200}[args.workload]
    if args.cpu_baseline_child:
        cpu_baseline_child(tuple(int(c) for c in args.cpu_baseline_child.split(",")), args.cpu_budget)
        return
    if args.copy_kernel > 0:
        copy_kernel(args.copy_kernel)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; sharding across GPUs is a design, "
                                   "not implemented", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "lps":
        print(json.dumps(lps_leg(args)))
        return
    if args.workload == "cube":
        print(json.dumps(cube_leg(args)))
        return
    if args.workload == "dem":
        print(json.dumps(dem_leg(args)))
        return
    if args.workload == "rows":
        print(json.dumps(rows_leg(args)))
        return
    cells = tuple(int(c) for c in args.cells.split(","))
    w = configs.kw_plate(num_cells=cells, steps=args.steps)
    n = w["x"].shape[0]
    ctx = engine.Context(0)
    res = {}
    specs = {}
    for prec in ("fp64", "mixed"):
        s = make_spec(ctx, w, prec, {"auto": 0, "compact": 1}[args.bond_layout])
        bonds, _ = s.bond_stats()
        if args.unroll >= 0:
            s.set_tuning(args.unroll)
            unroll = args.unroll
        else:
            unroll = s.autotune(3)
        wall, force_ms, launches = timed_run(s, args.warmup, args.steps)
        res[prec] = dict(wall=wall, force_ms=force_ms, launches=launches, bonds=bonds,
                         kernel=s.kernel_name(KCLS_FORCE), unroll=unroll, layout=s.layout(), setup_s=s.setup_s)
        specs[prec] = s
    bonds = res["mixed"]["bonds"]
    u64 = specs["fp64"].get("u")
    u32 = specs["mixed"].get("u")
    den = float(np.max(np.abs(u64)))
    u_err = float(np.max(np.abs(u32 - u64)) / den) if den > 0 else 0.0
    _, _, i64 = specs["fp64"].bonds()
    _, _, i32 = specs["mixed"].bonds()
    broken64 = int((i64 == 0).sum())
    broken_diff = int((i64 != i32).sum())
    for s in specs.values():
        s.close()

    m = res["mixed"]
    value = bonds * args.steps / m["wall"]
    fp64_value = bonds * args.steps / res["fp64"]["wall"]
    avg_ms = m["force_ms"] / max(m["launches"], 1)
    b_alg = algorithmic_bytes_per_step(n, bonds, passes=1, fracture=True)
    achieved = b_alg / (avg_ms * 1e-3) / 1e9
    ctr = measured_counters(m["kernel"], cells)
    line = {
        "metric": "bond evals/s vs plain FP64 at fixed tol; specimen solves/s + p50/p99 latency",
        "value": value,
        "unit": "bond-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 bond math / fp64 accumulate+integrate",
        "data": "synthetic",
        "config": {"workload": "PMB brittle fracture pre-notched plate %dx%dx%d (%d particles, %d bonds), "
                               "edge impact 16.5 m/s, dt 7.5e-8" % (cells + (n, bonds)),
                   "particles": n, "bonds": bonds, "precision": "mixed", "layout": m["layout"],
                   "unroll": m["unroll"], "setup_s": m["setup_s"]},
        "fp64": {"value": fp64_value, "unit": "bond-evals/s", "ms_per_step": res["fp64"]["wall"] / args.steps * 1e3,
                 "kernel": res["fp64"]["kernel"],
                 "force_kernel_avg_ms": res["fp64"]["force_ms"] / max(res["fp64"]["launches"], 1)},
        "mixed_vs_fp64": {"speedup": value / fp64_value, "u_rel_err": u_err,
                          "broken_bonds_fp64": broken64, "broken_bond_mismatches": broken_diff},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": ctr["traffic"],
                     "kernel": m["kernel"], "kernel_avg_ms": avg_ms, "alg_bytes_per_launch": b_alg,
                     "valu_busy": ctr["valu_busy"], "sca_busy": ctr["sca_busy"], "counters_source": ctr["source"],
                     "note": roofline_note(ctr, avg_ms)},
        "cpu_baseline": None,
    }
    if args.requests > 0:
        line["serving"] = serving(args.requests, max_particles=args.max_particles)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cells)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
