"""Benchmark: partitioned edges/s of the MI355X multilevel k-way partitioner.

Default run (the headline): the batched form on a packed batch of 4096 graphs
(seed 4; vertex counts log-uniform in [500, 50000]; even graphs 2-D Delaunay
meshes, odd graphs layered circuit-like DAGs with edge weights 1..100 and
vertex weights 1..4; k drawn from {2,4,8,16,32}; imbalance 3%).  One step =
one gk_partition_packed call over the whole pack, from a device-resident
packed CSR to device-resident partitions.  Beside it, on the same graphs:
  * the plain path, one gk_partition_device call per graph, the median of 3
    timed calls per graph over all 4096 graphs (speedup_vs_plain);
  * per-graph p50/p99 latency of both;
  * the host-to-host window of the batched form (gk_partition_batch);
  * bit-exactness: batched == plain on every graph, == the CPU oracle on a
    bounded sample (which is also the CPU baseline, with the host named);
  * nested, the plain path on the 2-D Delaunay triangulation of 2^22 points
    (k = 64), with its host-to-host window and an oracle check, once with
    the points numbered in random order and once along a Hilbert curve;
  * nested, the serving form under Poisson arrivals at 200 requests/s for
    60 s (500 distinct graphs, n log-uniform in [1e3, 2e6]): p50/p99 latency
    and sustained edges/s, sampled requests checked alone and by the oracle.
`value` is 0 unless every result is bit-exact, deterministic and within the
3% cap.

  python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-c2] [--no-serve]
  python bench.py --form plain [--workload delaunay|dag|road] [--n N] [--k K]
  python bench.py --form serve [--rate R] [--duration S] [--pool P]
"""
import argparse
import ctypes
import importlib.util
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E peak


def load_pkg():
    spec = importlib.util.spec_from_file_location("gkway_amd", os.path.join(ROOT, "dac_g-kway_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gkway_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_graph(gk, workload, n, seed):
    if workload == "delaunay":
        return gk.delaunay(n, seed)
    if workload == "delaunay_hilbert":
        return gk.delaunay_hilbert(n, seed)
    if workload == "dag":
        return gk.dag(n, 200, seed)
    if workload == "road":
        return gk.road(n, seed)
    raise SystemExit(f"unknown workload {workload}")


def stage_table(gk, stats_list, steps, log_it=True):
    """Per-stage (ms, launches, algorithmic bytes) summed over the given stats."""
    stage = {}
    for i, nm in enumerate(gk.STAGE_NAMES):
        stage[nm] = (sum(s.stage_ms[i] for s in stats_list), sum(s.stage_launches[i] for s in stats_list),
                     sum(s.stage_bytes[i] for s in stats_list))
        ms, cnt, by = stage[nm]
        if cnt and log_it:
            log(f"stage {nm:13s} {ms / steps:9.3f} ms/step {cnt // steps:6d} calls/step "
                f"{by / steps / 1e6:10.1f} MB/step {by / max(ms, 1e-9) / 1e6:8.1f} GB/s")
    return stage


def launch_flags(gk, args):
    """Flags that choose launch mechanics only (results are identical): rounds
    replayed from a hipGraph, launch shapes picked by the autotuner,
    persistent work-stealing launches."""
    return ((gk.GK_FLAG_GRAPH_ROUNDS if args.graph_rounds else 0) | (gk.GK_FLAG_AUTOTUNE if args.autotune else 0)
            | (gk.GK_FLAG_PERSISTENT if args.persistent else 0))


def dominant_stage(gk, L, ctx, stage_step, off, steps=2):
    """Runs `steps` steps with every stage timed, returns the stage with the
    most device time over them and restricts the context's stage timing to
    it."""
    if off:
        return None
    tot = [0.0] * len(gk.STAGE_NAMES)
    for _ in range(steps):
        st = stage_step()
        for i in range(len(gk.STAGE_NAMES)):
            tot[i] += st.stage_ms[i]
    dom = max(range(len(gk.STAGE_NAMES)), key=lambda i: tot[i])
    L.gk_ctx_set_stage_mask(ctx.handle, 1 << dom)
    log(f"dominant stage: {gk.STAGE_NAMES[dom]} ({tot[dom] / steps:.3f} ms per timed-stage step)")
    return gk.STAGE_NAMES[dom]


def roofline_of(stage, workload_tag, dom=None):
    """Roofline of the dominant stage: its algorithmic bytes per call over its
    average event-timed duration."""
    if dom is None:
        dom = max(stage, key=lambda k_: stage[k_][0])
    ms, cnt, by = stage[dom]
    if ms <= 0:
        return None
    gbs = by / (ms / 1e3) / 1e9
    r = {"bound": "hbm", "kernel": dom, "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
         "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": None, "alg_bytes_per_call": int(by / max(cnt, 1)),
         "avg_call_us": round(ms * 1e3 / max(cnt, 1), 2)}
    tf = os.path.join(ROOT, "profiles", f"traffic_{dom}_{workload_tag}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            prof = json.load(f)
        r["traffic"] = prof.get("hbm_bytes_per_call")
        # The stage's kernel time per call in the committed kernel trace of
        # the workload: short launches read longer under events, which carry
        # their dispatch.
        if prof.get("kernel_us_per_call"):
            r["avg_call_us_trace"] = prof["kernel_us_per_call"]
            r["achieved_trace"] = round(r["alg_bytes_per_call"] / (prof["kernel_us_per_call"] * 1e-6) / 1e9, 1)
    return r


def batch_set(gk, count, seed, lo=500, hi=50000, ks=(2, 4, 8, 16, 32)):
    """Packed-batch workload: vertex counts log-uniform in [lo, hi]; even
    graphs are 2-D Delaunay meshes, odd ones layered circuit-like DAGs
    (layers = clamp(round(sqrt(n)), 2, 200), weights 1..100 / 1..4); k drawn
    from ks."""
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)
    kk = rng.choice(np.array(ks), count)
    graphs = []
    for i, n in enumerate(ns):
        s = seed * 1_000_003 + i
        if i % 2 == 0:
            graphs.append(gk.delaunay(int(n), s))
        else:
            graphs.append(gk.dag(int(n), int(min(200, max(2, round(float(np.sqrt(n)))))), s))
    return graphs, [int(k) for k in kk]


def host_info():
    """CPU model and usable cores of the host this runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu": model, "nproc": usable, "cpu_count": os.cpu_count()}


def upload_graphs(gk, torch, graphs):
    """Device copies of each graph (a standalone CSR per graph) and a
    partition buffer per graph, for plain-path calls."""
    I32P = gk.I32P
    ptr = lambda t_: ctypes.cast(t_.data_ptr(), I32P)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).cuda()
    out = []
    for g in graphs:
        keep = (dev(g.xadj), dev(g.adjncy if g.nnz else np.zeros(1, np.int32)),
                dev(g.adjwgt if g.nnz else np.zeros(1, np.int32)), dev(g.vwgt),
                torch.zeros(max(g.n, 1), dtype=torch.int32, device="cuda"))
        c = gk.Csr()
        c.n, c.nnz = g.n, g.nnz
        c.xadj, c.adjncy, c.adjwgt, c.vwgt = (ptr(t_) for t_ in keep[:4])
        out.append((c, keep))
    torch.cuda.synchronize()
    return out


def plain_over_graphs(gk, torch, ctx, graphs, ks, reps):
    """The plain path (one gk_partition_device call per graph, device CSR in,
    device partition out) over every graph: one untimed pass, then `reps`
    timed passes, graph by graph.  Returns the per-pass times [reps, G] and
    each graph's partition."""
    L = gk.lib()
    dgs = upload_graphs(gk, torch, graphs)
    I32P = gk.I32P
    times = np.zeros((reps, len(graphs)))
    for r in range(reps + 1):
        for i, (c, keep) in enumerate(dgs):
            p1 = gk.params(ks[i])
            a = time.perf_counter()
            rc = L.gk_partition_device(ctx.handle, ctypes.byref(c), ctypes.byref(p1),
                                       ctypes.cast(keep[4].data_ptr(), I32P), None, None)
            dt = time.perf_counter() - a
            if rc != 0:
                raise RuntimeError(f"plain path failed on graph {i}: {L.gk_last_error().decode()}")
            if r > 0:
                times[r - 1, i] = dt
        log(f"plain pass {r}: {times[r - 1].sum() if r else 0.0:.2f} s")
    parts = [keep[4].cpu().numpy()[:g.n] for (c, keep), g in zip(dgs, graphs)]
    return times, parts


def run_batch(args, gk, torch, nested_c2=True):
    """Batched form (the headline): all graphs in one packed CSR, one
    gk_partition_packed call per step (device CSR -> device partitions),
    against the plain path on the same graphs."""
    t = time.time()
    graphs, ks = batch_set(gk, args.graphs, args.seed)
    m_tot = sum(g.m for g in graphs)
    log(f"generated {len(graphs)} graphs, n={sum(g.n for g in graphs)} m={m_tot} in {time.time() - t:.1f}s")
    pg, v_off, e_off = gk.pack_graphs(graphs)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).cuda()
    dx, da, dw, dv = dev(pg.xadj), dev(pg.adjncy), dev(pg.adjwgt), dev(pg.vwgt)
    dpart = torch.zeros(pg.n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    I32P = gk.I32P
    ptr = lambda t_: ctypes.cast(t_.data_ptr(), I32P)
    pk = gk.PackedCsr()
    pk.ngraphs, pk.n, pk.nnz = len(graphs), pg.n, pg.nnz
    pk.xadj, pk.adjncy, pk.adjwgt, pk.vwgt = ptr(dx), ptr(da), ptr(dw), ptr(dv)
    pk.v_off = v_off.ctypes.data_as(I32P)
    pk.e_off = e_off.ctypes.data_as(I32P)
    G = len(graphs)
    prm = (gk.Params * G)(*[gk.params(k, flags=launch_flags(gk, args) | (0 if args.no_stage_time else gk.GK_FLAG_STAGE_TIME))
                            for k in ks])
    prm_stage = (gk.Params * G)(*[gk.params(k, flags=launch_flags(gk, args) | gk.GK_FLAG_STAGE_TIME) for k in ks])
    L = gk.lib()
    ctx = gk.Context(0)

    def step(pp=prm):
        st = (gk.Stats * G)()
        rc = L.gk_partition_packed(ctx.handle, ctypes.byref(pk), pp, ptr(dpart), st, None)
        if rc != 0:
            raise RuntimeError(f"gk_partition_packed: {L.gk_strerror(rc).decode()}: {L.gk_last_error().decode()}")
        return st

    for i in range(args.warmup):
        a = time.perf_counter()
        step()
        log(f"warmup {i}: {(time.perf_counter() - a) * 1e3:.1f} ms")
    dom = dominant_stage(gk, L, ctx, lambda: step(prm_stage)[0], args.no_stage_time)
    torch.cuda.synchronize()
    times, runs = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        a = time.perf_counter()
        runs.append(step())
        times.append(time.perf_counter() - a)
        log(f"step {i}: {times[-1] * 1e3:.1f} ms")
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    part = dpart.cpu().numpy()
    last = runs[-1]
    cuts = [s.cut for s in last]
    deterministic = all([s.cut for s in r] == cuts for r in runs)
    lat = np.array([s.t_total_ms for r in runs for s in r])
    worst_imb = max(s.imbalance for s in last)
    # eps = 3% in integers: every part within cap = max(floor(1.03 T / k), ceil(T / k)).
    within_cap = all(s.max_part_wgt <= s.cap for s in last)
    stage, roofline = None, None
    if not args.no_stage_time:
        roofline = roofline_of(stage_table(gk, [r[0] for r in runs], args.steps, log_it=False),
                               f"batch_g{G}_seed{args.seed}", dom)
        L.gk_ctx_set_stage_mask(ctx.handle, 0xFFFFFFFF)
        sruns = [step(prm_stage) for _ in range(args.stage_steps)]
        stage = stage_table(gk, [r[0] for r in sruns], args.stage_steps, log_it=True)
    value = m_tot / (total / args.steps)

    # Host-to-host window of the batched form (host graphs packed and copied
    # in, partitions copied out): gk_partition_batch.
    h2h = []
    for _ in range(2):
        a = time.perf_counter()
        hparts, _ = ctx.partition_batch(graphs, ks)
        h2h.append(time.perf_counter() - a)
    h2h_exact = all(np.array_equal(hp, part[v_off[i]:v_off[i + 1]]) for i, hp in enumerate(hparts))
    del hparts

    # The plain path on every graph: median of `plain_reps` timed calls per
    # graph; its partitions must equal the batched ones.  (plain_reps 0 skips
    # it, for profiling runs of the batched form alone.)
    plain = None
    exact_plain = True
    if args.plain_reps > 0:
        ptimes, pparts = plain_over_graphs(gk, torch, ctx, graphs, ks, args.plain_reps)
        med = np.median(ptimes, axis=0)
        plain_value = m_tot / med.sum()
        rep_values = [m_tot / ptimes[r].sum() for r in range(args.plain_reps)]
        exact_plain = all(np.array_equal(pp, part[v_off[i]:v_off[i + 1]]) for i, pp in enumerate(pparts))
        del pparts
        log(f"plain path on all {G} graphs: {plain_value / 1e6:.2f} M edges/s (passes "
            f"{min(rep_values) / 1e6:.2f}..{max(rep_values) / 1e6:.2f}), bit-exact vs batched={exact_plain}")
        plain = {"edges_per_s": round(plain_value, 1),
                 "per_graph": f"median of {args.plain_reps} timed gk_partition_device calls per graph, all {G} "
                              f"graphs, after one untimed pass",
                 "pass_edges_per_s_min": round(min(rep_values), 1),
                 "pass_edges_per_s_max": round(max(rep_values), 1),
                 "p50_ms": round(float(np.percentile(med * 1e3, 50)), 3),
                 "p99_ms": round(float(np.percentile(med * 1e3, 99)), 3)}

    # The oracle on a bounded sample: bit-exactness and the CPU baseline.
    cpu, exact_oracle, n_oracle = None, None, 0
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle
        exact_oracle = True
        tc, mc = 0.0, 0
        sample = sorted(np.random.default_rng(args.seed + 100).permutation(G)[:args.plain_sample].tolist())
        for i in sample:
            a = time.perf_counter()
            opart, ost = _oracle.partition(graphs[i], ks[i], initpart_threads=1)
            tc += time.perf_counter() - a
            mc += graphs[i].m
            n_oracle += 1
            exact_oracle &= bool(np.array_equal(opart, part[v_off[i]:v_off[i + 1]]))
            if tc > args.cpu_budget:
                break
        cpu = {"value": round(mc / tc, 1), "unit": "edges/s", "cores": 1, "kind": "port",
               "sample": f"{n_oracle} of the {G} graphs (seeded random sample) partitioned one by one "
                         f"by the single-threaded oracle ({tc:.1f} s, {mc} edges)",
               "host": host_info()}
        log(f"oracle: {n_oracle} graphs in {tc:.1f}s, bit_exact={exact_oracle}")

    if args.calibrate_copy:
        calibrate_copy()
    exact = exact_plain and h2h_exact and exact_oracle is not False
    ok = deterministic and exact and within_cap
    out = {
        "metric": "partitioned edges/s at bit-exact cut, eps=3% (batched form vs plain GPU path; p50/p99)",
        "value": round(value, 1) if ok else 0.0,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 3),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "latency": "per-graph completion time inside a batch step (all graphs of all timed steps)",
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generators)",
        "config": {"workload": f"packed batch of {G} graphs (seed {args.seed}): n log-uniform in [500, 50000], "
                               f"even = 2-D Delaunay, odd = layered DAG (w 1..100, vw 1..4); k in {{2,4,8,16,32}}",
                   "n_total": int(pg.n), "m_total": int(m_tot), "eps": 0.03,
                   "window": "device packed CSR -> device partitions",
                   "rounds": "hipGraph replay" if args.graph_rounds else "stream launches",
                   "autotune": bool(args.autotune), "persistent": bool(args.persistent)},
        "plain": plain,
        "speedup_vs_plain": round(value / plain["edges_per_s"], 2) if plain else None,
        "host_to_host_edges_per_s": round(m_tot / min(h2h), 1),
        "worst_imbalance": round(worst_imb, 5),
        "all_parts_within_cap": within_cap,
        "bit_exact": {"batched_vs_plain_all_graphs": exact_plain, "host_to_host_vs_device": h2h_exact,
                      "vs_oracle_sample": exact_oracle, "oracle_sample_graphs": n_oracle},
        "deterministic": deterministic,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    ctx.close()
    if nested_c2 and not args.no_c2:
        c2 = argparse.Namespace(**vars(args))
        c2.workload, c2.n, c2.k, c2.seed = "delaunay", 1 << 22, 64, 1
        c2.steps, c2.warmup, c2.stage_steps = args.c2_steps, 2, 1
        c2.graph_rounds, c2.autotune, c2.persistent = args.plain_launch
        out["plain_config2"] = run_plain(c2, gk, torch, emit=False)
        # The same mesh with its vertices numbered along a Hilbert curve:
        # neighbours get nearby ids, as in the DIMACS10 Delaunay inputs
        # (delaunay_n11: median |u - v| over its edges is 20 of 2048, against
        # about n / 3 for random numbering).
        c2.workload = "delaunay_hilbert"
        out["plain_config2_hilbert"] = run_plain(c2, gk, torch, emit=False)
    if not args.no_serve:
        # Config 5: the serving form under Poisson arrivals at 200 requests/s
        # for 60 s, requests drawn from 500 distinct graphs.
        c5 = argparse.Namespace(**vars(args))
        c5.seed, c5.rate, c5.duration, c5.pool, c5.check = 5, 200.0, 60.0, 500, 24
        out["serve_config5"] = run_serve(c5, gk, torch, emit=False)
    print(json.dumps(out), flush=True)


def calibrate_copy():
    """Elementwise kernel of known traffic: reads 1 GiB and writes 1 GiB per
    dispatch; FETCH_SIZE / WRITE_SIZE of it give the counters' scale."""
    import torch as _t
    a = _t.ones(1 << 28, dtype=_t.int32, device="cuda")
    b = _t.empty_like(a)
    _t.cuda.synchronize()
    for _ in range(3):
        _t.mul(a, 3, out=b)
    _t.cuda.synchronize()
    log("calibration: 3 x elementwise multiply, 1 GiB in + 1 GiB out each")


def serve_pool(gk, count, seed, lo=1000, hi=2_000_000):
    """Distinct request graphs: vertex counts log-uniform in [lo, hi]; 40%
    Delaunay meshes, 40% layered DAGs (edge weights 1..100, vertex weights
    1..4), 20% road-like (average degree 2.1)."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)
    kinds = rng.permutation(np.array(["delaunay"] * (count * 2 // 5) + ["dag"] * (count * 2 // 5) +
                                     ["road"] * (count - 2 * (count * 2 // 5))))

    def make(i):
        n, kind, s = int(ns[i]), kinds[i], seed * 1_000_003 + i
        if kind == "delaunay":
            return gk.delaunay(n, s)
        if kind == "dag":
            return gk.dag(n, int(min(200, max(2, round(float(np.sqrt(n)))))), s)
        return gk.road(n, s)

    # The generators are C calls that release the interpreter lock.
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(make, range(count)))


def run_serve(args, gk, torch, emit=True):
    """Serving form: open-loop Poisson arrivals into gk_serve_*; each request
    is a graph drawn from a pool of distinct graphs with its own k."""
    import threading
    t = time.time()
    pool = serve_pool(gk, args.pool, args.seed)
    log(f"generated a pool of {len(pool)} graphs (n {min(g.n for g in pool)}..{max(g.n for g in pool)}) "
        f"in {time.time() - t:.1f}s")
    rng = np.random.default_rng(args.seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, int(args.rate * args.duration * 1.2) + 16))
    arrivals = arrivals[arrivals < args.duration]
    picks = rng.integers(0, len(pool), len(arrivals))
    ks = rng.choice(np.array([8, 16, 64, 256]), len(arrivals))
    srv = gk.Server(0, hbm_budget=0)
    ids = [None] * len(arrivals)
    t0 = time.perf_counter()

    def submitter():
        for i, at in enumerate(arrivals):
            d = t0 + at - time.perf_counter()
            if d > 0:
                time.sleep(d)
            ids[i] = srv.submit(pool[picks[i]], int(ks[i]))

    backlog = {}

    def submit_all():
        submitter()
        q = srv.stats()
        backlog.update(at_s=time.perf_counter() - t0, queued=int(q.queued), inflight=int(q.inflight),
                       completed=int(q.completed))

    th = threading.Thread(target=submit_all)
    th.start()
    results = {}
    check_every = max(1, len(arrivals) // args.check)
    nxt = 0
    while nxt < len(arrivals):
        if ids[nxt] is None:
            time.sleep(0.001)
            continue
        part, st = srv.wait(ids[nxt])
        if nxt % check_every == 0:
            results[nxt] = part
        nxt += 1
        if nxt % 500 == 0:
            log(f"{nxt} of {len(arrivals)} requests done at {time.perf_counter() - t0:.1f}s")
    th.join()
    wall = time.perf_counter() - t0
    s = srv.stats()
    srv.close()
    m_done = sum(pool[p].m for p in picks)
    # Alone vs in the mix: the sampled requests, each partitioned alone.
    exact_alone = True
    with gk.Context(0) as ctx:
        for i, part in results.items():
            p1, _ = ctx.partition(pool[picks[i]], int(ks[i]))
            exact_alone &= bool(np.array_equal(p1, part))
    exact = exact_alone
    log(f"serve: {len(arrivals)} requests in {wall:.1f}s; p50 {s.p50_ms:.1f} ms p99 {s.p99_ms:.1f} ms; "
        f"{s.edges_per_s / 1e6:.1f} M edges/s; {len(results)} checked alone-vs-mix: {exact}")
    cpu = None
    exact_oracle, ng = None, 0
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle
        tc, mc = 0.0, 0
        exact_oracle = True
        for i in sorted(results):
            g = pool[picks[i]]
            a = time.perf_counter()
            opart, _ = _oracle.partition(g, int(ks[i]), initpart_threads=1)
            tc += time.perf_counter() - a
            exact_oracle &= bool(np.array_equal(opart, results[i]))
            mc += g.m
            ng += 1
            if tc > args.cpu_budget:
                break
        cpu = {"value": round(mc / tc, 1), "unit": "edges/s", "cores": 1, "kind": "port",
               "sample": f"{ng} of the checked requests partitioned one by one by the single-threaded oracle "
                         f"({tc:.1f} s)", "host": host_info()}
        log(f"oracle on {ng} served requests: bit_exact={exact_oracle}")
        exact &= exact_oracle
    ok = exact and s.failed == 0 and s.completed == len(arrivals)
    out = {
        "metric": "sustained partitioned edges/s at bit-exact cut, eps=3% (serving form, continuous batching)",
        "value": round(s.edges_per_s, 1) if ok else 0.0,
        "unit": "edges/s",
        "n_gpus": 1,
        "requests": len(arrivals),
        "p50_ms": round(s.p50_ms, 3),
        "p99_ms": round(s.p99_ms, 3),
        "max_ms": round(s.max_ms, 3),
        "offered_edges_per_s": round(m_done / args.duration, 1),
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "int32",
        "data": f"synthetic: a pool of {len(pool)} distinct seeded graphs, requests draw from it",
        "config": {"workload": f"Poisson arrivals at {args.rate}/s for {args.duration}s; n log-uniform in [1e3, 2e6]; "
                               f"40% Delaunay, 40% DAG (w 1..100), 20% road (deg 2.1); k in {{8,16,64,256}}",
                   "eps": 0.03, "window": "host CSR submitted -> host partition returned"},
        "backlog_when_arrivals_ended": backlog,
        "drain_s": round(wall - float(arrivals[-1]), 2),
        "engine_steps": int(s.steps),
        "footprint_peak_bytes": int(s.footprint_peak),
        "device_peak_bytes": int(s.device_peak),
        "bit_exact": {"alone_vs_mix": exact_alone, "alone_checked": len(results), "vs_oracle_sample": exact_oracle,
                      "oracle_sample_requests": ng},
        "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def run_plain(args, gk, torch, emit=True):
    """Plain path, one graph per call (device CSR -> device partition)."""
    t = time.time()
    g = make_graph(gk, args.workload, args.n, args.seed)
    log(f"generated {args.workload} n={g.n} m={g.m} in {time.time() - t:.1f}s")

    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).cuda()
    dx, da, dw, dv = dev(g.xadj), dev(g.adjncy), dev(g.adjwgt), dev(g.vwgt)
    dpart = torch.zeros(g.n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    I32P = gk.I32P
    ptr = lambda t_: ctypes.cast(t_.data_ptr(), I32P)
    csr = gk.Csr()
    csr.n, csr.nnz = g.n, g.nnz
    csr.xadj, csr.adjncy, csr.adjwgt, csr.vwgt = ptr(dx), ptr(da), ptr(dw), ptr(dv)
    L = gk.lib()

    ctx = gk.Context(0)
    # Stage timing brackets a stage's launches with HIP events on the work
    # stream, and each event adds a gap before the next kernel.  The timed
    # steps therefore time only the dominant stage (the roofline kernel),
    # picked from one all-stage step before them; the full stage table comes
    # from separate steps after them.
    p = gk.params(args.k, flags=launch_flags(gk, args) | (0 if args.no_stage_time else gk.GK_FLAG_STAGE_TIME))
    p_stage = gk.params(args.k, flags=launch_flags(gk, args) | gk.GK_FLAG_STAGE_TIME)

    def step(pp=p):
        st = gk.Stats()
        rc = L.gk_partition_device(ctx.handle, ctypes.byref(csr), ctypes.byref(pp), ptr(dpart),
                                   ctypes.byref(st), None)
        if rc != 0:
            raise RuntimeError(f"gk_partition_device: {L.gk_strerror(rc).decode()}: {L.gk_last_error().decode()}")
        return st

    for i in range(args.warmup):
        st = step()
        log(f"warmup {i}: {st.t_total_ms:.1f} ms cut={st.cut}")
    dom = dominant_stage(gk, L, ctx, lambda: step(p_stage), args.no_stage_time)

    torch.cuda.synchronize()
    times, stats = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        a = time.perf_counter()
        st = step()
        times.append(time.perf_counter() - a)
        stats.append(st)
        log(f"step {i}: {times[-1] * 1e3:.1f} ms cut={st.cut} imb={st.imbalance:.4f}")
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    part = dpart.cpu().numpy()
    cuts = {s.cut for s in stats}
    deterministic = len(cuts) == 1
    st = stats[-1]

    # Roofline from the dominant stage's events over the timed steps; the
    # stage table from separate all-stage steps.
    stage, roofline = None, None
    if not args.no_stage_time:
        roofline = roofline_of(stage_table(gk, stats, args.steps, log_it=False),
                               f"{args.workload}_n{args.n}_k{args.k}", dom)
        L.gk_ctx_set_stage_mask(ctx.handle, 0xFFFFFFFF)
        sstats = [step(p_stage) for _ in range(args.stage_steps)]
        stage = stage_table(gk, sstats, args.stage_steps, log_it=True)

    # Host-to-host window (host CSR in, host partition out, PCIe included):
    # gk_partition, the window to set against a host partitioner.
    h2h = []
    for _ in range(3):
        a = time.perf_counter()
        hpart, _ = ctx.partition(g, args.k)
        h2h.append(time.perf_counter() - a)
    h2h_exact = bool(np.array_equal(hpart, part))

    m = g.m
    ms_step = total / args.steps * 1e3
    value = m / (total / args.steps)
    p50 = statistics.median(times) * 1e3
    p99 = float(np.percentile(np.array(times) * 1e3, 99))

    cpu = None
    bit_exact = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle  # the checker, and the CPU baseline
        t = time.perf_counter()
        opart, ost = _oracle.partition(g, args.k, initpart_threads=1)
        tcpu = time.perf_counter() - t
        bit_exact = bool(np.array_equal(opart, part)) and ost.cut == st.cut
        cpu = {"value": round(m / tcpu, 1), "unit": "edges/s", "cores": 1, "kind": "port",
               "sample": f"one full partition of the same graph by the single-threaded oracle "
                         f"({tcpu:.1f} s, initial partitioner on 1 thread)",
               "cut": int(ost.cut), "host": host_info()}
        log(f"oracle: {tcpu:.1f}s cut={ost.cut} bit_exact={bit_exact}")

    if args.calibrate_copy:
        calibrate_copy()

    ok = deterministic and (bit_exact is not False) and h2h_exact and st.max_part_wgt <= st.cap
    out = {
        "metric": "partitioned edges/s at bit-exact cut, eps=3% (plain path, one graph per call)",
        "value": round(value, 1) if ok else 0.0,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "p50_ms": round(p50, 3),
        "p99_ms": round(p99, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"{args.workload} n={g.n} m={m} seed={args.seed}", "k": args.k,
                   "eps": 0.03, "window": "device CSR -> device partition",
                   "rounds": "hipGraph replay" if args.graph_rounds else "stream launches",
                   "autotune": bool(args.autotune), "persistent": bool(args.persistent)},
        "host_to_host_ms": round(min(h2h) * 1e3, 3),
        "host_to_host_edges_per_s": round(m / min(h2h), 1),
        "cut": int(st.cut),
        "imbalance": round(st.imbalance, 5),
        "levels": int(st.nlevels),
        "coarsest_n": int(st.coarsest_n),
        "phase_ms": {"coarsen": round(st.t_coarsen_ms, 3), "init": round(st.t_init_ms, 3),
                     "refine": round(st.t_refine_ms, 3)},
        "bit_exact_vs_oracle": bit_exact,
        "deterministic": deterministic,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    ctx.close()
    if emit:
        print(json.dumps(out), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (batch: 8, plain: 20)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (batch: 2, plain: 3)")
    ap.add_argument("--form", default="batch", choices=["plain", "batch", "serve"])
    ap.add_argument("--workload", default="delaunay")
    ap.add_argument("--n", type=int, default=1 << 22)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--graphs", type=int, default=4096, help="batch form: graphs in the pack")
    ap.add_argument("--plain-reps", type=int, default=3, help="batch form: timed plain-path calls per graph")
    ap.add_argument("--plain-sample", type=int, default=256, help="batch form: graphs offered to the oracle")
    ap.add_argument("--no-c2", action="store_true", help="batch form: skip the nested 2^22-point Delaunay runs")
    ap.add_argument("--no-serve", action="store_true", help="batch form: skip the nested 60 s serving run")
    ap.add_argument("--c2-steps", type=int, default=10, help="batch form: timed steps of the nested plain run")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="batch/serve forms: seconds of oracle work")
    ap.add_argument("--rate", type=float, default=200.0, help="serve form: Poisson arrival rate (requests/s)")
    ap.add_argument("--duration", type=float, default=20.0, help="serve form: arrival window (s)")
    ap.add_argument("--pool", type=int, default=48, help="serve form: distinct graphs requests draw from")
    ap.add_argument("--check", type=int, default=24, help="serve form: requests re-run alone for bit-exactness")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle (CPU baseline + bit-exact check)")
    ap.add_argument("--no-stage-time", action="store_true")
    ap.add_argument("--graph-rounds", action="store_true",
                    help="replay each level's refinement rounds from a captured hipGraph")
    ap.add_argument("--autotune", action="store_true",
                    help="time candidate launch shapes of matching and contraction on first use")
    ap.add_argument("--persistent", action="store_true",
                    help="run matching, contraction rows and the gain pass as persistent work-stealing launches")
    ap.add_argument("--direct", action="store_true",
                    help="batch form: direct launches and eager rounds (by default the batched form runs "
                         "persistent work-stealing launches, hipGraph-replayed rounds and autotuned launch shapes)")
    ap.add_argument("--stage-steps", type=int, default=3,
                    help="extra steps with per-stage HIP events for the stage table and roofline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain device copy of 1 GiB (PMC calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
                          "(graph-level sharding across GPUs is not implemented)"}))
        sys.exit(2)

    gk = load_pkg()
    import torch

    if args.form == "batch":
        # The batched form's launch mechanics (results identical): persistent
        # work-stealing launches, rounds replayed from captured hipGraphs and
        # autotuned launch shapes, unless --direct.  The nested plain runs
        # keep the plain path's own (direct unless asked for).
        args.plain_launch = (args.graph_rounds, args.autotune, args.persistent)
        if not args.direct:
            args.graph_rounds = args.autotune = args.persistent = True
        args.seed = 4 if args.seed is None else args.seed
        args.steps = 8 if args.steps is None else args.steps
        args.warmup = 2 if args.warmup is None else args.warmup
        run_batch(args, gk, torch)
        return
    if args.form == "serve":
        args.seed = 5 if args.seed is None else args.seed
        run_serve(args, gk, torch)
        return
    args.seed = 1 if args.seed is None else args.seed
    args.steps = 20 if args.steps is None else args.steps
    args.warmup = 3 if args.warmup is None else args.warmup
    run_plain(args, gk, torch)


if __name__ == "__main__":
    main()
