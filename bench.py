"""Benchmark of the electron-phonon vertex path on one MI355X.

Workload: the 3C-SiC-like synthetic material (fcc a = 4.36 A, 2 atoms, 8
Wannier bands, 6 branches, Wigner-Seitz sets of an 8^3 coarse grid, 617 cells)
on a 24^3 k = q mesh with a window keeping ~4 of 8 bands, driven through
calculate_eph_interaction_ibzk('g').  One step = the whole job: every in-window
IBZ state (all 413 IBZ k) against every k' of the window and every branch,
inputs resident in HBM, |g|^2 written to HBM.  Reported: |g|^2 vertices/s, the roofline of the
dominant kernel (fused phase-contraction + rotation + |g|^2 kernel) from HIP
events on its stream, and the CPU reference method (the oracle) timed on a
bounded sample on this host.

Serving (north-star pattern 3): a Poisson stream of |g|^2 requests at arbitrary
k- and q-lists over resident materials of the multi-material generator,
served with continuous batching by the library's server in mixed precision
and in plain FP64; p50/p99 latency and sustained vertices/s in "serving".

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-serving] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "|g|² vertices/sec, 1 MI355X, mixed vs FP64 at ≤1e-5 rel err; p50/p99 latency"
FP64_PEAK_TFLOPS = 78.6
INT8_PEAK_TOPS = 5000.0
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def log_mem(tag):
    from elphbolt_amd import _lib
    f, t = _lib.device_meminfo(0)
    log(f"[bench] device memory {tag}: {f / 2**30:.1f} GiB free of {t / 2**30:.1f} GiB")


def pair_rows(cells):
    """Rows of the cell-pair form of the Rp contraction: the cells (R, -R) are
    folded into one cos row and one sin row (a cell without its opposite, R = 0,
    still takes two), so the phase operand is real."""
    cells = [tuple(c) for c in np.asarray(cells, np.int64)]
    present, used, npair = set(cells), set(), 0
    for c in cells:
        if c in used:
            continue
        used.add(c)
        neg = tuple(-x for x in c)
        if neg in present and neg not in used:
            used.add(neg)
        npair += 1
    return 2 * npair


def algorithmic_flops(nst, nwv, nstates_inwindow, nb, nbr, k2):
    """Flops of one pair-kernel launch (nst IBZ states x nwv k' rows): the Rp
    phase contraction of the band-rotated g(k,Rp) in cell-pair form (k2 real
    phase rows times complex columns: 4 k2 nb nbr per state and row), the
    final-band and mode rotations (8 nb nbr + 8 nbr^2 per vertex column block)
    and |g|^2/omega (4 per vertex)."""
    gemm = 4.0 * k2 * nb * nbr * nwv * nst
    epi = nst * nstates_inwindow * (8.0 * nbr * nb + 8.0 * nbr * nbr + 4.0 * nbr)
    return gemm + epi


def survey_flops(nst, nwv, nstates_inwindow, nb, nbr, nwsg, nm):
    """The same launch counted with the unrotated per-pair form (Rp contraction
    over all nb^2 nbr columns of g(k,Rp), then rotations), shared over the nm
    in-window initial bands of the IBZ point."""
    pairs = nst * nwv / max(nm, 1)
    rot = 8.0 * (nb * nb * nbr * nm) + 8.0 * nb * nm * nbr * (nstates_inwindow / nwv) \
        + 8.0 * nm * (nstates_inwindow / nwv) * nbr * nbr
    return pairs * (8.0 * nwsg * nb * nb * nbr + rot) + 4.0 * nst * nstates_inwindow * nbr


def host_cpu_info():
    """The host's CPU as this process sees it: logical CPUs (nproc), the CPUs
    this process may run on, OMP_NUM_THREADS and the model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity": aff, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "model": model}


def cpu_baseline(mat, el, ph, states, target_s=15.0, threads=None):
    """The oracle (literal C restatement of the reference's per-vertex g2 in the
    ibzk loop order, OpenMP over states) on a bounded sample: the given IBZ
    states, each over the first `ikp_limit` final k' of the window.  Threads:
    the CPUs this process may use, capped by OMP_NUM_THREADS when set."""
    from oracle import oracle as O
    info = host_cpu_info()
    if threads is None:
        threads = info["affinity"]
        if info["omp_num_threads"]:
            threads = min(threads, int(info["omp_num_threads"]))
    o = O.Wannier(mat)
    first, last = states
    # two pilot sizes: the difference removes the per-state g(k,Rp) setup cost
    t = []
    pilots = (10, 80)
    for lim in pilots:
        t0 = time.perf_counter()
        c, _ = O.ibzk_g2(o, el, ph, 1, first, last, nthreads=threads, ikp_limit=lim)
        t.append(time.perf_counter() - t0)
    nact = int((c >= 0).sum())
    per_row = max((t[1] - t[0]) / (pilots[1] - pilots[0]), 1e-6)
    ikp_limit = int(max(10, min(el.nwv, (target_s - t[0]) / per_row)))
    t0 = time.perf_counter()
    c, g = O.ibzk_g2(o, el, ph, 1, first, last, nthreads=threads, ikp_limit=ikp_limit)
    dt = time.perf_counter() - t0
    nvert = int(c[c >= 0].sum())
    return {"value": nvert / dt, "unit": "vertices/s", "cores": threads, "kind": "port",
            "host": info,
            "sample": f"ibzk states {first}..{last} ({nact} in window), first {ikp_limit} of {el.nwv} k' each, "
                      f"{nvert} vertices in {dt:.1f} s on {threads} OpenMP threads"}


SERVING_SEED = 0xE1B04000


def serving_requests(nreq, rate, sigma, nmat, seed=SERVING_SEED):
    """Seed-fixed request stream: Poisson arrivals at `rate`/s, material id
    uniform, k- and q-list lengths independent lognormal (median 500, log-sd
    sigma, clamped to [1, 20000]), coordinates uniform in [0, 1)^3."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, nreq))
    mid = rng.integers(0, nmat, nreq)
    nk = np.clip(np.rint(np.exp(np.log(500.0) + sigma * rng.standard_normal(nreq))), 1, 20000).astype(int)
    nq = np.clip(np.rint(np.exp(np.log(500.0) + sigma * rng.standard_normal(nreq))), 1, 20000).astype(int)
    return [dict(t=float(t[i]), mat=int(mid[i]), k=rng.random((nk[i], 3)), q=rng.random((nq[i], 3)))
            for i in range(nreq)]


def serve_stream(srv, mids, reqs):
    """Submits `reqs` on their arrival schedule from a thread; waits for each
    in order and releases it (the result is consumed in HBM).  Returns the
    latencies (arrival to completion) of completed requests and the number
    refused at admission (larger than the pool)."""
    import threading
    from elphbolt_amd._lib import ElphError
    ids = [None] * len(reqs)
    ready = [threading.Event() for _ in reqs]

    def submitter():
        t0 = time.perf_counter()
        for i, r in enumerate(reqs):
            dt = t0 + r["t"] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            try:
                ids[i] = srv.submit(mids[r["mat"]], r["k"], r["q"])
            except ElphError:
                ids[i] = -1
            ready[i].set()

    th = threading.Thread(target=submitter)
    th.start()
    lat, refused, failed = [], 0, 0
    for i in range(len(reqs)):
        ready[i].wait()
        if ids[i] < 0:
            refused += 1
            continue
        st, ms = srv.wait(ids[i])
        srv.release(ids[i])
        if st == 0:
            lat.append(ms)
        else:
            failed += 1
    th.join()
    return lat, refused, failed


def serving_bench(args):
    """Serving form: the eight resident materials of the serving set (the
    first eight draws of the multi-material generator with a 12 GB vertex cap
    and a vertex of at least 2 GB: 38.8 GB in HBM), continuous batching in the
    library's server, in mixed precision and in plain FP64.  Two request
    streams per server: the Poisson stream at 200/s (far above one GPU's
    capacity, so its latencies are queue-drain times), then, from the
    capacity that stream measured, a Poisson stream at 80% of it, whose
    p50/p99 are latencies outside overload."""
    from elphbolt_amd import synth
    from elphbolt_amd.serving import Server
    from elphbolt_amd.wannier import Wannier
    t0 = time.perf_counter()
    idx = synth.serving_indices()
    mats = [synth.multi_material(i, budget_bytes=synth.SERVING_BUDGET, window=False) for i in idx]
    ws = [Wannier(m) for m in mats]
    gb = sum(m.gwann_bytes for m in mats) / 1e9
    log(f"[bench] serving: {len(mats)} materials, {gb:.1f} GB of Wannier vertex in HBM "
        f"({time.perf_counter() - t0:.1f} s)")
    reqs = serving_requests(args.serving_requests, args.serving_rate, args.serving_sigma, len(mats))
    probe = [(m, np.random.default_rng(7 + m).random((8, 3)), np.random.default_rng(9 + m).random((8, 3)))
             for m in range(min(3, len(mats)))]
    out = {"workload": f"{len(mats)} resident materials of the multi-material generator (seeds "
                       f"{synth.MULTI_SEED_BASE:#x}+i for i in {idx}: the first eight draws with a "
                       f"{synth.SERVING_BUDGET / 1e9:g} GB vertex cap and at least "
                       f"{synth.SERVING_MIN_BYTES / 1e9:g} GB of vertex): " + ", ".join(m.name for m in mats),
           "resident_vertex_gb": round(gb, 2),
           "list_lengths": f"k and q independent lognormal, median 500, log-sd {args.serving_sigma:g}, "
                           f"clamped [1, 20000]; every (k, q) pair, all m, n, s; |g|^2 FP32",
           "pool_gb": args.serving_pool_gb, "slots": args.serving_slots}
    results = {}

    def run_stream(srv, mids, rq, tag, prec):
        base = srv.stats()
        t1 = time.perf_counter()
        lat, refused, failed = serve_stream(srv, mids, rq)
        wall = time.perf_counter() - t1
        st = srv.stats()
        vert = st["vertices"] - base["vertices"]
        lat = np.asarray(lat)
        res = {"requests": len(rq), "offered_rate_per_s": round(len(rq) / max(rq[-1]["t"], 1e-9), 2),
               "completed": int(lat.size), "refused_larger_than_pool": refused, "failed": failed,
               "p50_latency_ms": float(np.percentile(lat, 50)) if lat.size else None,
               "p99_latency_ms": float(np.percentile(lat, 99)) if lat.size else None,
               "mean_latency_ms": float(lat.mean()) if lat.size else None,
               "sustained_vertices_per_s": vert / wall, "wall_s": round(wall, 2), "vertices": vert,
               "achieved_requests_per_s": lat.size / wall, "offered_span_s": round(rq[-1]["t"], 3),
               "max_slots_busy": st["max_slots_busy"], "units": st["units"] - base["units"]}
        log(f"[bench] serving {prec} {tag}: {res['completed']} done at offered {res['offered_rate_per_s']}/s, "
            f"p50 {res['p50_latency_ms']:.0f} ms p99 {res['p99_latency_ms']:.0f} ms, "
            f"{res['sustained_vertices_per_s']:.3e} vertices/s, {res['achieved_requests_per_s']:.2f} req/s, "
            f"wall {wall:.1f} s")
        return res

    for prec in ("mixed", "fp64"):
        log_mem(f"serving {prec}: before server")
        with Server(prec, pool_bytes=int(args.serving_pool_gb * (1 << 30)), nslots=args.serving_slots) as srv:
            mids = [srv.add_material(w) for w in ws]
            log_mem(f"serving {prec}: server ready")
            pr = []
            for m, k, q in probe:   # warm-up and accuracy probe
                rid = srv.submit(mids[m], k, q)
                srv.wait(rid)
                pr.append(srv.fetch(rid, release=True))
            over = run_stream(srv, mids, reqs, f"Poisson {args.serving_rate:g}/s", prec)
            cap = over["achieved_requests_per_s"]
            rate_b = args.serving_load * cap
            reqs_b = serving_requests(args.serving_requests_steady, rate_b, args.serving_sigma, len(mats),
                                      seed=SERVING_SEED + 1)
            steady = run_stream(srv, mids, reqs_b, f"Poisson at {args.serving_load:g} x capacity", prec)
            log_mem(f"serving {prec}: streams done")
        results[prec] = ({"poisson_200_per_s": over, "capacity_requests_per_s": round(cap, 3),
                          f"poisson_{args.serving_load:g}_capacity": steady}, pr)
        log_mem(f"after serving {prec}")
    rel = []
    for a, b in zip(results["mixed"][1], results["fp64"][1]):
        a, b = a.astype(np.float64).ravel(), b.astype(np.float64).ravel()
        sel = b > 1e-6 * b.max()
        rel.append(np.abs(a[sel] - b[sel]) / b[sel])
    rel = np.concatenate(rel)
    out["mixed"] = results["mixed"][0]
    out["fp64"] = results["fp64"][0]
    out["mixed_vs_fp64_sustained"] = round(
        out["mixed"]["poisson_200_per_s"]["sustained_vertices_per_s"]
        / out["fp64"]["poisson_200_per_s"]["sustained_vertices_per_s"], 3)
    out["error_mixed_vs_fp64"] = {"max_rel": float(rel.max()), "median_rel": float(np.median(rel)),
                                  "entries": int(rel.size), "floor": "1e-6*max|g|^2",
                                  "probe": f"{len(probe)} requests of 8 k x 8 q", "bound_met": bool(rel.max() <= 1e-5)}
    return out


BATCH_SEED = 0xE1B05000


def transport_bench(args, w, sym, el, plans, g2, nst):
    """The consumers of one step's g2 rows (all IBZ states, device-resident):
    calculate_eph_interaction_ibzk('X') with the tetrahedron delta, the e-ph
    RTA rates, then the RTA transport coefficients with the field terms
    multiplied by the symmetrizers and the tensors symmetrized over the
    Cartesian operations, as the reference's dragless RTA step does
    (src/bte.f90:236-277).  The chain runs on the FP64 rows (timed) and on the
    mixed rows of the same states; sigma, sigma*S and S of both are reported
    with their relative differences."""
    from elphbolt_amd import interactions
    from elphbolt_amd.bte import ElectronRTA
    T, mu = 300.0, float(el.enref)
    plan = plans["fp64"]
    t0 = time.perf_counter()
    tet = interactions.Tetrahedra(w, el.wvmesh, el.ens, indexlist=el.indexlist)
    t_tet = time.perf_counter() - t0
    bufs = plan.run_X(tet, T, mu, 1, nst, g2)          # allocates the outputs; untimed
    log(f"[bench] X: tetrahedra {t_tet:.2f} s, first pass {plan.timing()['dominant_ms']:.2f} ms")
    kms, walls = [], []
    for i in range(args.transport_steps):
        t = time.perf_counter()
        plan.run_X(tet, T, mu, 1, nst, g2, buffers=bufs)
        walls.append(time.perf_counter() - t)
        kms.append(plan.timing()["dominant_ms"])
        log(f"[bench] X step {i}: kernel {kms[-1]:.3f} ms, call {walls[-1] * 1e3:.2f} ms")
    s = np.arange(nst)
    act = np.abs(el.ens_irred[s // el.nb, s % el.nb] - el.enref) <= el.fsthick
    procs = int(act.sum()) * plan.nprocs
    k_s = float(np.mean(kms)) * 1e-3
    algo = procs * (8 + 16 + 16)    # |g|^2 read, X+ and X- written, the two state maps written
    ti = el.transport_inputs(w, sym)
    vol = abs(float(np.linalg.det(np.asarray(w.mat.lattvecs))))
    tr = ElectronRTA(w, el.ens, ti["vels"], el.indexlist, ti["nequiv"], ti["ibz2fbz_map"], vol, el.wvmesh, 2,
                     symmetrizers=ti["symmetrizers"], crotations=ti["crotations"])
    tr.rta(bufs["rates_device"], T, mu)
    t = time.perf_counter()
    coef = {"fp64": tr.rta(bufs["rates_device"], T, mu)}
    t_rta = time.perf_counter() - t
    # the same chain on the mixed-precision rows of the same states
    g2_mx = w.device_buffer((nst, plan.nprocs))
    plans["mixed"].run(1, nst, out=g2_mx)
    plans["mixed"].run_X(tet, T, mu, 1, nst, g2_mx, buffers=bufs)
    coef["mixed"] = tr.rta(bufs["rates_device"], T, mu)
    g2_mx.free()
    vals = {}
    for prec, c in coef.items():
        sig, sigS = ElectronRTA.scalar(c["sigma"]), ElectronRTA.scalar(c["sigmaS"])
        vals[prec] = {"sigma": sig, "sigmaS": sigS, "seebeck_V_per_K": sigS / sig if sig else None}
    d = {k: abs(vals["mixed"][k] / vals["fp64"][k] - 1.0) for k in ("sigma", "sigmaS", "seebeck_V_per_K")}
    log(f"[bench] RTA transport {t_rta * 1e3:.2f} ms: sigma {vals['fp64']['sigma']:.6e} 1/Ohm/m (mixed rel diff "
        f"{d['sigma']:.2e}), sigmaS rel diff {d['sigmaS']:.2e}, S rel diff {d['seebeck_V_per_K']:.2e}")
    ach = algo / k_s / 1e9
    return {"T_K": T, "chempot_eV": mu, "states": nst, "processes": procs,
            "X_kernel_ms": round(k_s * 1e3, 3), "X_call_ms": round(float(np.mean(walls)) * 1e3, 3),
            "processes_per_s": procs / k_s,
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "kernel": "k_ibzk_X",
                         "bytes_per_launch": algo},
            "rta_transport_ms": round(t_rta * 1e3, 3), "tetrahedra_s": round(t_tet, 3),
            "symmetrized": "field terms x symmetrizers, tensors over the Cartesian operations",
            "fp64": vals["fp64"], "mixed": vals["mixed"], "mixed_vs_fp64_rel_diff": d,
            "sigma": vals["fp64"]["sigma"], "sigmaS": vals["fp64"]["sigmaS"],
            "seebeck_V_per_K": vals["fp64"]["seebeck_V_per_K"]}


def ibzq_bench(args, w, el, ph, mat):
    """calculate_eph_interaction_ibzq('g') over all IBZ phonon states of the
    same workload (FP64), output rows in HBM."""
    from elphbolt_amd import interactions
    plan = interactions.IbzqPlan(w, el, ph, precision="fp64")
    nst = ph.nwv_irred * mat.nbr
    out = w.device_buffer((nst, plan.stride))
    counts, _ = plan.run(1, nst, out=out)            # warm-up
    dom, tot, walls = [], [], []
    for i in range(args.ibzq_steps):
        w.synchronize()
        t = time.perf_counter()
        counts, _ = plan.run(1, nst, out=out)
        w.synchronize()
        walls.append(time.perf_counter() - t)
        tm = plan.timing()
        dom.append(tm["dominant_ms"])
        tot.append(tm["all_ms"])
        log(f"[bench] ibzq step {i}: {walls[-1] * 1e3:.1f} ms wall, pair kernels {tm['dominant_ms']:.1f} ms "
            f"in {tm['launches']} launches")
    vertices = int(counts.sum())
    k2 = -(-pair_rows(mat.rcells_k) // 32) * 32    # cos/sin rows of the Re cell pairs, padded
    flops = 4.0 * k2 * mat.nb ** 2 * el.nwv * nst   # real phase x complex h, per (state, k row)
    ach = flops / (np.mean(dom) * 1e-3) / 1e12
    out.free()
    plan.close()
    return {"workload": f"all {nst} IBZ phonon states (q, s), {el.nwv} k rows each, FP64",
            "vertices_per_step": vertices, "value": vertices / float(np.mean(walls)), "unit": "vertices/s",
            "ms_per_step": round(float(np.mean(walls)) * 1e3, 2),
            "roofline": {"bound": "mfma", "achieved": round(ach, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(ach / FP64_PEAK_TFLOPS, 4), "kernel": "k_pair_q_fp64",
                         "pair_ms_per_step": round(float(np.mean(dom)), 2), "flops_per_step": flops}}


def _batched_material(i, budget):
    from elphbolt_amd import synth
    return synth.multi_material(i, budget_bytes=budget, bands_in_window=4.0, pair_floor=True)


def start_batched_generation(args):
    """Generates the batched materials in worker processes started with the
    spawn method before this process touches the GPU (they overlap the rest
    of the benchmark); returns (executor, futures)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    n = args.batched_materials
    ex = cf.ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1), mp_context=mp.get_context("spawn"))
    futs = [ex.submit(_batched_material, i, args.batched_budget_gb * 1e9) for i in range(n)]
    return ex, futs


def batched_bench(args, gen):
    """Batched form: materials of the multi-material generator (atoms 1-8,
    bands 4-26, WS 4^3-9^3, fine meshes 12^3-24^3, window keeping ~4 bands),
    each with 0.5-2e6 k-q pairs of calculate_eph_interaction_ibzk('g'), all in
    one captured hipGraph per precision with tile shapes from the autotuner."""
    import gc
    from elphbolt_amd import bz, interactions
    from elphbolt_amd._lib import ElphError
    from elphbolt_amd.wannier import Wannier
    t0 = time.perf_counter()
    ex, futs = gen
    mats = [f.result() for f in futs]
    ex.shutdown()
    log(f"[bench] batched: {len(mats)} materials generated (waited {time.perf_counter() - t0:.1f} s)")
    rng = np.random.default_rng(BATCH_SEED)
    entries, skipped = [], []
    for m in mats:
        target = float(rng.uniform(0.5e6, 2e6))
        try:
            w = Wannier(m)
            sym = bz.Symmetry(m.lattvecs, m.basis, m.masses, m.wvmesh)
            el = bz.Electrons(w, sym, m.wvmesh, m.enref, m.fsthick)
            ph = bz.Phonons(w, sym, m.phmesh)
        except ElphError as e:
            skipped.append(f"{m.name}: {e}")
            continue
        plans = {pr: interactions.IbzkPlan(w, el, ph, precision=pr) for pr in ("mixed", "fp64")}
        nprocs = plans["fp64"].nprocs
        # IBZ k (in istate order) until the pair target or the output budget
        act_k = plans["fp64"].active.sum(axis=1)
        rows = np.cumsum(act_k)
        # the smallest whole number of IBZ k at or above the target, within 2e6 pairs
        nk_pairs = min(int(np.ceil(target / el.nwv)), int(2e6 // el.nwv))
        nk_out = int(np.searchsorted(rows, args.batched_out_gb * 1e9 / (nprocs * 8), side="right"))
        nk = int(max(1, min(el.nwv_irred, nk_pairs, nk_out)))
        entries.append(dict(mat=m, w=w, el=el, plans=plans, last=nk * m.nb, pairs=nk * el.nwv, target=target))
    log(f"[bench] batched: {len(entries)} plans ready, {len(skipped)} skipped ({time.perf_counter() - t0:.1f} s)")
    gb = sum(e["mat"].gwann_bytes for e in entries) / 1e9
    pairs = np.array([e["pairs"] for e in entries])
    out = {"workload": f"{len(entries)} materials of the multi-material generator (seeds "
                       f"{0xE1B03000:#x}+i, vertex cap {args.batched_budget_gb:g} GB, window ~4 bands): "
                       f"calculate_eph_interaction_ibzk('g') on the first IBZ k of each, up to a target of "
                       f"U[0.5e6, 2e6] k-q pairs or {args.batched_out_gb:g} GB of |g|^2 output per material",
           "materials": len(entries), "skipped": skipped, "resident_vertex_gb": round(gb, 2),
           "pairs_total": int(pairs.sum()), "pairs_per_material": [int(pairs.min()), int(np.median(pairs)),
                                                                   int(pairs.max())],
           "materials_at_pair_target": int(sum(1 for e in entries if e["pairs"] >= 0.98 * e["target"]
                                               or e["last"] == e["el"].nwv_irred * e["mat"].nb)),
           "materials_in_pair_range": int(sum(1 for e in entries if 0.5e6 <= e["pairs"] <= 2e6)),
           "output_gb": round(sum(float((e["plans"]["fp64"].active.sum(axis=1).cumsum()[e["last"] // e["mat"].nb - 1])
                                        * e["plans"]["fp64"].nprocs * 8) for e in entries) / 1e9, 2),
           "nb_range": [min(e["mat"].nb for e in entries), max(e["mat"].nb for e in entries)],
           "nbr_range": [min(e["mat"].nbr for e in entries), max(e["mat"].nbr for e in entries)],
           "nws_range": [min(e["mat"].nwsk for e in entries), max(e["mat"].nwsk for e in entries)],
           "streams": args.batched_streams}
    bufs = [None] * len(entries)
    probe = {}
    for prec in ("mixed", "fp64"):
        b = interactions.Batch(nstreams=args.batched_streams)
        res = []
        for i, e in enumerate(entries):
            c, buf = b.add(e["plans"][prec], 1, e["last"], out=bufs[i])
            bufs[i] = buf
            res.append(c)
        t1 = time.perf_counter()
        b.prepare(autotune=True)
        t_prep = time.perf_counter() - t1
        vert = sum(int(c[c >= 0].sum()) for c in res)
        b.run()
        ms = [b.run() for _ in range(args.batched_steps)]
        tw = time.perf_counter()
        b.run()
        wall_ms = (time.perf_counter() - tw) * 1e3
        cls = b.classes()
        b.close()
        probe[prec] = [[bufs[i].row(r) for r in range(min(2, bufs[i].shape[0]))] for i in range(min(16, len(entries)))]
        out[prec] = {"vertices_per_step": vert, "ms_per_step": float(np.mean(ms)), "ms_steps": [round(x, 2) for x in ms],
                     "vertices_per_s": vert / (np.mean(ms) * 1e-3), "wall_ms_one_replay": round(wall_ms, 2),
                     "prepare_s": round(t_prep, 2),
                     "tile_classes": [{"cols": c["cols"], "kp": c["kp"], "members": c["members"],
                                       "path": "fused" if c["path"] == 0 else "column-tiled",
                                       "cb": c["cb"], "rb": c["rb"], "candidates": c["candidates"],
                                       "tuned_ms": round(c["ms"], 3)} for c in cls]}
        log(f"[bench] batched {prec}: {vert:.3e} vertices per step, {np.mean(ms):.1f} ms per graph replay, "
            f"{vert / (np.mean(ms) * 1e-3):.3e} vertices/s, {len(cls)} size classes (prepare {t_prep:.1f} s)")
    rel = []
    for rm, r6 in zip(probe["mixed"], probe["fp64"]):
        for a, bb in zip(r6, rm):
            sel = np.abs(a) > 1e-6 * np.abs(a).max()
            rel.append(np.abs(bb[sel] - a[sel]) / np.abs(a[sel]))
    rel = np.concatenate(rel) if rel else np.zeros(1)
    out["mixed_vs_fp64"] = round(out["mixed"]["vertices_per_s"] / out["fp64"]["vertices_per_s"], 3)
    out["error_mixed_vs_fp64"] = {"max_rel": float(rel.max()), "median_rel": float(np.median(rel)),
                                  "entries": int(rel.size), "rows": "first 2 in-window states of 16 materials",
                                  "bound_met": bool(rel.max() <= 1e-5)}
    del bufs, entries
    gc.collect()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kblock", type=int, default=0, help="IBZ k points per step (0 = all: the whole job)")
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=96, help="requests of the 200/s stream")
    ap.add_argument("--serving-rate", type=float, default=200.0)
    ap.add_argument("--serving-requests-steady", type=int, default=128,
                    help="requests of the stream below capacity")
    ap.add_argument("--serving-load", type=float, default=0.8,
                    help="offered load of the second stream as a fraction of the measured capacity")
    ap.add_argument("--serving-sigma", type=float, default=1.0)
    ap.add_argument("--serving-pool-gb", type=float, default=128.0)
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--profile", action="store_true",
                    help="config-1 kernels only (no serving, batched or CPU sections; ibzq unless "
                         "--no-ibzq), after a device-to-device copy of --calibration-copy-gb GiB, for "
                         "hardware-counter runs")
    ap.add_argument("--calibration-copy-gb", type=float, default=1.0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-transport", action="store_true")
    ap.add_argument("--no-ibzq", action="store_true")
    ap.add_argument("--ibzq-steps", type=int, default=2)
    ap.add_argument("--transport-steps", type=int, default=3)
    ap.add_argument("--batched-materials", type=int, default=64)
    ap.add_argument("--batched-budget-gb", type=float, default=2.0)
    ap.add_argument("--batched-out-gb", type=float, default=6.0,
                    help="cap on one material's |g|^2 output (the 64 draws need at most ~4 GB each)")
    ap.add_argument("--batched-streams", type=int, default=4)
    ap.add_argument("--batched-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md, not run)"}))
        sys.exit(2)
    if args.profile:
        args.no_batched = args.no_serving = args.no_cpu_baseline = True
    gen = None if args.no_batched else start_batched_generation(args)

    from elphbolt_amd import bz, interactions, synth
    from elphbolt_amd.wannier import Wannier

    t0 = time.perf_counter()
    mat = synth.config(args.config)
    log(f"[bench] material {mat.name}: nb={mat.nb} nbr={mat.nbr} nws={mat.nwsk} mesh={list(mat.wvmesh)} "
        f"enref={mat.enref:.3f} fsthick={mat.fsthick:.3f} ({time.perf_counter() - t0:.1f} s)")
    w = Wannier(mat, device=0)
    sym = bz.Symmetry(mat.lattvecs, mat.basis, mat.masses, mat.wvmesh)
    el = bz.Electrons(w, sym, mat.wvmesh, mat.enref, mat.fsthick)
    ph = bz.Phonons(w, sym, mat.phmesh)
    log(f"[bench] setup: nsymm={sym.nsymm} nwv={el.nwv} nwv_irred={el.nwv_irred} "
        f"nstates_inwindow={el.nstates_inwindow} ({time.perf_counter() - t0:.1f} s)")
    nb, nbr = mat.nb, mat.nbr
    kb = el.nwv_irred if args.kblock <= 0 else min(args.kblock, el.nwv_irred)
    nblocks = max(1, el.nwv_irred // kb)
    nm_avg = float((np.abs(el.ens_irred - el.enref) <= el.fsthick).sum()) / max(1, el.nwv_irred)
    if args.profile:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE calibrate the counters
        nbytes = int(args.calibration_copy_gb * (1 << 30))
        src, dst = w.device_buffer((nbytes // 8,)), w.device_buffer((nbytes // 8,))
        dst.copy_from(src)
        log(f"[bench] calibration copy of {nbytes} bytes done")
        src.free()
        dst.free()
    plans = {pr: interactions.IbzkPlan(w, el, ph, mesh_ref=1, precision=pr) for pr in ("mixed", "fp64")}
    nprocs = plans["fp64"].nprocs
    out = w.device_buffer((kb * nb, nprocs))

    def timed(plan, tag):
        def step(i):
            b = i % nblocks
            counts, _ = plan.run(b * kb * nb + 1, (b + 1) * kb * nb, out=out)
            return counts
        for i in range(args.warmup):
            step(i)
        log(f"[bench] {tag}: {args.warmup} warmup steps done")
        vertices, dom_ms, dom_n, nst_launch = 0, 0.0, 0, 0
        w.synchronize()
        t_start = time.perf_counter()
        for i in range(args.steps):
            counts = step(args.warmup + i)
            nact = int((counts >= 0).sum())
            vertices += int(counts[counts >= 0].sum())
            tm = plan.timing()
            dom_ms += tm["dominant_ms"]
            dom_n += tm["launches"]
            nst_launch += nact
            log(f"[bench] {tag} step {i}: {nact} states, {tm['dominant_ms']:.2f} ms in {tm['launches']} pair "
                f"launches, {tm['all_ms']:.2f} ms device total")
        w.synchronize()
        elapsed = time.perf_counter() - t_start
        avg_states = nst_launch / max(dom_n, 1)
        return dict(value=vertices / elapsed, elapsed=elapsed, vertices=vertices,
                    avg_launch_s=dom_ms / max(dom_n, 1) * 1e-3, states_per_launch=avg_states)

    res_mx = timed(plans["mixed"], "mixed")
    res_64 = timed(plans["fp64"], "fp64")

    # accuracy of the mixed path against the FP64 path on the same states
    ncmp = min(el.nwv_irred, 32) * nb
    a = w.device_buffer((ncmp, nprocs))
    b = w.device_buffer((ncmp, nprocs))
    c64, _ = plans["fp64"].run(1, ncmp, out=a)
    cmx, _ = plans["mixed"].run(1, ncmp, out=b)
    w.synchronize()
    act = c64 >= 0
    assert np.array_equal(c64, cmx)
    ref = a.numpy()[act].ravel()
    got = b.numpy()[act].ravel()
    a.free()
    b.free()
    sel = np.abs(ref) > 1e-6 * np.abs(ref).max()
    rel = np.abs(got[sel] - ref[sel]) / np.abs(ref[sel])
    exact_zero_match = bool(np.array_equal(ref == 0.0, got == 0.0))
    err = {"max_rel": float(rel.max()), "median_rel": float(np.median(rel)), "floor": "1e-6*max|g|^2",
           "entries_compared": int(sel.sum()), "reference": "FP64 path (itself within 1e-12 of the oracle)",
           "zeros_match": exact_zero_match, "bound_met": bool(rel.max() <= 1e-5)}
    log(f"[bench] mixed vs fp64: max rel {err['max_rel']:.2e} median {err['median_rel']:.2e}")

    k2 = pair_rows(mat.rcells_q[:mat.nwsg])
    spl = res_mx["states_per_launch"]
    i8_ops = spl * el.nwv * 10 * 2 * 2.0 * k2 * nb * nbr   # 10 digit products x (Re, Im) columns
    ach_i8 = i8_ops / res_mx["avg_launch_s"] / 1e12
    f64eq_mx = algorithmic_flops(spl, el.nwv, el.nstates_inwindow, nb, nbr, k2) / res_mx["avg_launch_s"] / 1e12
    spl64 = res_64["states_per_launch"]
    fl64 = algorithmic_flops(spl64, el.nwv, el.nstates_inwindow, nb, nbr, k2)
    ach64 = fl64 / res_64["avg_launch_s"] / 1e12
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "ibzk_pair_i8_sic_ws8_k24_traffic.json")
    if mat.name == "sic_like_ws8_k24" and os.path.exists(tfile):
        traffic = json.load(open(tfile))["bytes_per_launch"]   # rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated
    roofline = {"bound": "mfma", "achieved": round(ach_i8, 2), "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                "frac": round(ach_i8 / INT8_PEAK_TOPS, 4), "traffic": traffic, "kernel": "k_pair_i8<3>",
                "avg_launch_ms": round(res_mx["avg_launch_s"] * 1e3, 3),
                "ops_per_launch": i8_ops, "fp64_equivalent_tflops": round(f64eq_mx, 2)}
    fp64 = {"value": res_64["value"], "ms_per_step": res_64["elapsed"] / args.steps * 1e3,
            "roofline": {"bound": "mfma", "achieved": round(ach64, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(ach64 / FP64_PEAK_TFLOPS, 4), "kernel": "k_pair_fp64<3>",
                         "avg_launch_ms": round(res_64["avg_launch_s"] * 1e3, 3), "flops_per_launch": fl64,
                         "achieved_unrotated_count": round(
                             survey_flops(spl64, el.nwv, el.nstates_inwindow, nb, nbr, mat.nwsg, nm_avg)
                             / res_64["avg_launch_s"] / 1e12, 2)}}
    res = {"metric": METRIC, "value": res_mx["value"], "unit": "vertices/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res_mx["elapsed"] / args.steps * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int8+f64", "data": "synthetic",
           "config": {"workload": f"{mat.name}: calculate_eph_interaction_ibzk('g'), all {el.nwv_irred} IBZ k "
                                  f"per step; mixed precision (config 2) timed first, plain FP64 (config 1) after",
                      "nb": nb, "nbr": nbr, "nws": mat.nwsg, "mesh": [int(x) for x in mat.wvmesh],
                      "nwv_window": el.nwv, "nwv_irred": el.nwv_irred, "nstates_inwindow": el.nstates_inwindow,
                      "vertices_per_step": res_mx["vertices"] // args.steps},
           "mixed_vs_fp64": round(res_mx["value"] / res_64["value"], 3), "error_vs_fp64": err,
           "fp64": fp64, "roofline": roofline}
    if not args.no_transport and nblocks == 1:
        res["transport"] = transport_bench(args, w, sym, el, plans, out, kb * nb)
        log_mem("after transport")
    if not args.no_ibzq:
        res["ibzq"] = ibzq_bench(args, w, el, ph, mat)
        log_mem("after ibzq")
    if not args.no_batched:
        log_mem("before batched")
        res["batched"] = batched_bench(args, gen)
        log_mem("after batched")
    if not args.no_serving:
        res["serving"] = serving_bench(args)
        log_mem("after serving")
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle) ...")
        first_active = None
        for s in range(1, el.nwv_irred * nb + 1):
            ik, m = (s - 1) // nb, (s - 1) % nb
            if abs(el.ens_irred[ik, m] - el.enref) <= el.fsthick:
                first_active = s
                break
        res["cpu_baseline"] = cpu_baseline(mat, el, ph, (first_active, min(first_active + 15, el.nwv_irred * nb)),
                                           target_s=args.cpu_seconds)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
