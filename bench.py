#!/usr/bin/env python3
"""Batched search benchmark of the MI355X sieve path on one GPU.

One step = one batched search (route -> LUT scan + RaBitQ estimate + candidate top-k ->
bound cut -> int8 rerank) of the workload's query batch against an index resident in HBM.
The index is built on the GPU first (seeded synthetic corpus in the reference's index format,
not timed). Prints one JSON line: QPS at the cheapest swept configuration reaching
R@10 >= 0.90 against exact flat search, p50/p99 single-query latency, the dominant scan
kernel's roofline (HIP events around its launches, algorithmic work), the CPU oracle (a
restatement of the reference's x86 search) timed on this host on a bounded sample with its
results compared to the GPU's, and, after the 113.5M workload, the 1B x 100-d L2 workload's
record under "workloads".

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload c3] [--then c4|none]
"""
import argparse
import gc
import glob
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "a-game-of-vectors_amd"))
sys.path.insert(0, REPO)

METRIC = "QPS & p50/p99 ms at R@10>=0.90 (113.5M x 768-d); LUT-scan GB/s vs HBM peak"
HBM_PEAK_GBS = 8000.0
# Dense int8 MFMA peak of one MI355X (no sparsity), in tera-ops/s.
MFMA_I8_PEAK_TOPS = 5000.0
# int8 ops the cluster-major MFMA pass spends per (query, slot) pair it scores: the one-hot
# operand of a slot is 64 groups x 16 entries, one multiply-add each against the query's table.
CM_OPS_PER_PAIR = 2 * 64 * 16

# name, rows, mixture components, clusters, queries per batch, swept configurations
# (nprobe:cand:rerank[:bM[xW]][:pD], src/bin/search.rs:180-218).
WORKLOADS = {
    "c4": dict(desc="1B x 100-d int8, L2, 131072 leaves (512 coarse cells x 256), RaBitQ 128-d codes, exact int8 "
                    "rerank on the 100 GB rows in HBM, 10k-query batches", l2=True,
               rows=1_000_000_000, components=131072, topics=16, leaves=(512, 256), queries=10_000,
               sweep=["4:500", "8:1000", "16:1000", "32:1000", "64:2000"]),
    "c4s": dict(desc="200M x 100-d int8, L2, 32768 leaves (256 x 128) (the 1B config's shape, scaled)", l2=True,
                rows=200_000_000, components=32768, topics=16, leaves=(256, 128), queries=10_000,
                sweep=["4:500", "8:1000", "16:1000", "32:1000"]),
    "c3": dict(desc="113.5M x 768-d, 16384 Zipf(1.1) components, 16384 clusters, SOAR, RaBitQ 1-bit + 2-bit tier, "
                    "87 GB int8 rows in HBM, 10k-query batches",
               rows=113_500_000, components=16384, clusters=16384, queries=10_000, weights="zipf", tier2=True,
               topics=131072, topic_spread=0.08, spread=0.05, config_search="64:2000:tiered200",
               sweep=["64:1000:full:b10x0.8", "64:1000:full", "64:2000:tiered200", "64:2000:tiered400",
                      "64:1000:full:b10x0.8:p0.1", "48:1000:full:b10x0.8"]),
    "c2": dict(desc="18.4M x 768-d, 8192 clusters, SOAR, RaBitQ 1-bit, 10k-query batches, margin probing, "
                    "bound-cut rerank",
               rows=18_400_000, components=8192, clusters=8192, queries=10_000,
               sweep=["32:500:full:b10x0.8", "64:500:full:b10x0.8", "128:500:full:b10x0.8",
                      "128:500:full:b10x0.8:p0.1", "128:1000:full:b10x0.8:p0.1", "128:2000:full:b10x0.8"]),
    "c1": dict(desc="1M x 768-d, 1024 clusters, SOAR, RaBitQ 1-bit, 1000 queries, nprobe 32, 1000 candidates",
               rows=1_000_000, components=1024, clusters=1024, queries=1000, config_search="32:1000:full",
               sweep=["16:1000:full", "32:1000:full", "64:1000:full"]),
    "tiny": dict(desc="200k x 768-d, 256 clusters (quick check)", rows=200_000, components=256, clusters=256,
                 queries=2000, sweep=["8:500:full", "16:500:full", "32:500:full"]),
}
# Mixture recipe: query noise N(0, 0.05^2) per dimension throughout. Flat components of spread
# 0.1 at 18.4M rows: recall rises with nprobe and R@10 0.90 is reachable (64 probes, 1000
# candidates); narrower components (0.02-0.045) tie the 10th/11th neighbours closer than the
# 1-bit estimate resolves, wider ones (>= 0.14) spread the similarity outside the 256-d prefix.
# At 113.5M rows under Zipf(1.1) weights flat components hold millions of rows each and no
# setting reaches 0.90 (the exact 256-d prefix finds 99% of the top 10 among its best 2000, the
# 1-bit estimate 70-78%), so that workload adds 131072 topics (spread 0.08) inside the
# components, rows at spread 0.05 around them: R@10 0.92 at 64 probes, 1000 candidates.
SPREAD = 0.1


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def recall10(rows, counts, gt):
    import numpy as np
    from sieve_amd.evaluate import recall_at
    rows = np.asarray(rows).view(np.uint32)
    return float(np.mean([recall_at(10, rows[i, :counts[i]], gt[i]) for i in range(len(gt))]))


def pmc_traffic(kernel, workload, search):
    """HBM bytes per launch of `kernel` on this workload and search configuration, from the
    rocprofv3 --pmc summaries committed under profiles/ (a line `pmc_traffic {json}` holding
    FETCH_SIZE and WRITE_SIZE per dispatch, the copy-probe scale factors and their product);
    None when no profile covers it."""
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                if not line.startswith("pmc_traffic "):
                    continue
                try:
                    d = json.loads(line[len("pmc_traffic "):])
                except ValueError:
                    continue
                if d.get("kernel") == kernel and d.get("workload") == workload and d.get("search") == search:
                    return dict(bytes_per_launch=d["bytes_per_launch"], profile=os.path.relpath(path, REPO),
                                fetch_kib_per_dispatch=d.get("fetch_kib_per_dispatch"),
                                write_kib_per_dispatch=d.get("write_kib_per_dispatch"))
    return None


def host_cpus():
    """The host's CPU facts and the worker count the CPU baseline uses: the CPUs this process may
    run on (affinity, cgroup quota, OMP_NUM_THREADS, whichever is least: on a shared GPU host that
    is one GPU's share of its cores), with nproc and the lscpu model recorded beside it."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    omp = int(omp) if omp and omp.isdigit() and int(omp) > 0 else None
    threads = min(x for x in (affinity, quota, omp) if x)
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), None)
        except OSError:
            pass
    return dict(nproc=nproc, affinity=affinity, cgroup_cpus=quota, omp_num_threads=omp, lscpu_model=model,
                threads=threads)


def ties_equal(g_ids, g_sc, w_ids, w_sc):
    """Two best-first top-k lists agree up to ties: every entry above the lowest kept score of
    either list is in both with a bit-identical score."""
    import numpy as np
    g_sc, w_sc = np.asarray(g_sc, np.float32), np.asarray(w_sc, np.float32)
    if len(g_ids) == 0 or len(w_ids) == 0:
        return len(g_ids) == len(w_ids)
    b = max(g_sc.min(), w_sc.min())
    g = {int(i): s for i, s in zip(g_ids, g_sc) if s > b}
    w = {int(i): s for i, s in zip(w_ids, w_sc) if s > b}
    return g.keys() == w.keys() and all(np.float32(g[i]).view(np.uint32) == np.float32(w[i]).view(np.uint32)
                                        for i in g)


def parity_counts(gs, gr, gn, ws, wr, wn):
    """GPU (gs, gr, gn) against oracle (ws, wr, wn) results of the same queries: lists identical,
    and lists that differ beyond ties."""
    import numpy as np
    m = len(wn)
    same = sum(1 for i in range(m) if gn[i] == wn[i] and np.array_equal(gr[i, :gn[i]], wr[i, :wn[i]])
               and np.array_equal(gs[i, :gn[i]].view(np.uint32), ws[i, :wn[i]].view(np.uint32)))
    bad = sum(1 for i in range(m) if not ties_equal(gr[i, :gn[i]], gs[i, :gn[i]], wr[i, :wn[i]], ws[i, :wn[i]]))
    return {"queries": m, "identical": same, "mismatches_beyond_ties": bad}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--then", default=None,
                    help="a second workload measured after the first in the same run, its record nested under "
                         "'workloads' (default: c4 after c3; 'none' to skip)")
    ap.add_argument("--gt-queries", type=int, default=1000, help="queries with exact ground truth")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample length")
    ap.add_argument("--cpu-latency-seconds", type=float, default=4.0,
                    help="target length of the CPU baseline's single-thread latency sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--latency-queries", type=int, default=300)
    ap.add_argument("--config", default=None, help="skip the sweep and time this configuration")
    ap.add_argument("--sweep", default=None, help="comma-separated configurations to sweep instead of the "
                                                 "workload's list")
    ap.add_argument("--spread", type=float, default=None, help="mixture component spread (data recipe)")
    ap.add_argument("--scan-variant", type=int, default=None,
                    help="use this scan kernel variant instead of autotuning (profiling runs)")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="open-loop serving run per rate (0: skip)")
    ap.add_argument("--serve-rates", default="2000,20000", help="offered Poisson arrival rates, queries/s")
    ap.add_argument("--serve-max-batch", type=int, default=512)
    ap.add_argument("--serve-config", default=None,
                    help="search configuration of the serving runs (default: the timed one)")
    ap.add_argument("--serve-mixed", type=float, default=3.0,
                    help="seconds of the mixed serving load per rate (filters on 30%% of queries, live inserts "
                         "at 100k rows/s with 10%% deletes), run last on the same index (0: skip); rates above "
                         "3000/s run for 10000/rate seconds")
    ap.add_argument("--serve-mixed-rates", default="500,2000,5000,20000",
                    help="offered rates of the mixed serving load")
    ap.add_argument("--serve-mixed-filtered", type=float, default=0.3, help="fraction of filtered queries")
    ap.add_argument("--serve-mixed-writes", type=int, default=1, help="1: live inserts and deletes, 0: none")
    ap.add_argument("--cluster-scan", default=None,
                    help="cluster-major scan settings seed_slots:ext_cap:chunk_blocks:phases (0 = default)")
    ap.add_argument("--cm-sweep", default=None,
                    help="comma-separated cluster-major scan settings (as --cluster-scan) to time on the picked "
                         "configuration with scan variants 10 and 11 before the timed run")
    ap.add_argument("--streams", type=int, default=1,
                    help="host threads, each driving its own stream, over which the timed steps are split")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run an elementwise kernel that reads 1 GiB and writes 1 GiB (the known size that "
                         "scales a profile's FETCH_SIZE/WRITE_SIZE counters), then exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (index sharding is a design "
              "note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if a.copy_probe:
        copy_probe()
        return
    W = WORKLOADS[a.workload]
    out = main_l2(a, W, a.workload) if W.get("l2") else main_ip(a, W, a.workload)
    then = a.then if a.then is not None else ("c4" if a.workload == "c3" else "none")
    if then and then != "none":
        # Everything of the first workload is freed first (the 1B index needs ~130 GB of HBM).
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        log(f"second workload {then}; device free {torch.cuda.mem_get_info()[0] / 1e9:.0f} GB")
        try:
            W2 = WORKLOADS[then]
            a2 = argparse.Namespace(**vars(a))  # the first workload's search overrides do not carry over
            a2.config = a2.sweep = a2.scan_variant = a2.cluster_scan = a2.cm_sweep = None
            sub = main_l2(a2, W2, then) if W2.get("l2") else main_ip(a2, W2, then)
        except Exception as e:  # reported, not fatal: the first record stands
            sub = {"error": f"{type(e).__name__}: {e}"}
            log(f"workload {then} failed: {sub['error']}")
        out["workloads"] = {then: sub}
    print(json.dumps(out), flush=True)


def main_ip(a, W, name):
    """The inner-product path (ivf::Index::search) on a 768-d workload; returns its record."""
    import numpy as np
    import torch
    from sieve_amd import build as B, parse_config, synth
    from sieve_amd.index import Index

    dev = torch.device("cuda:0")
    t_build0 = time.time()
    spread = a.spread if a.spread is not None else W.get("spread", SPREAD)
    spec = synth.MixtureSpec(components=W["components"], seed=20241203, spread=spread,
                             weights=W.get("weights", "lognormal"), topics=W.get("topics", 0),
                             topic_spread=W.get("topic_spread", 0.0))
    mix = synth.Mixture(spec, device=dev)
    log(f"workload {name}: ingest {W['rows']} rows")
    corpus = B.ingest(mix, W["rows"], batch=1 << 20)
    log("build: k-means, SOAR, RaBitQ codes, blocks")
    arrays = B.build(corpus, W["clusters"], sample_n=max(400_000, 64 * W["clusters"]), iters=12, chunk=1 << 19,
                     tier2=W.get("tier2", False))
    q768, q256 = mix.queries(W["queries"])
    nq = W["queries"]
    ngt = min(a.gt_queries, nq)
    log(f"exact ground truth for {ngt} queries")
    gt_ids, _ = B.exact_topk(mix, W["rows"], q768[:ngt], k=10)
    gt = gt_ids.cpu().numpy()
    # The library allocates with hipMalloc, outside torch's caching allocator: hand back the
    # build's freed blocks first.
    torch.cuda.empty_cache()
    log(f"torch holds {torch.cuda.memory_allocated() / 1e9:.1f} GB after the build")
    ix = Index.create(arrays.device(), borrow_rerank=True)
    t_build = time.time() - t_build0
    sizes = np.array(arrays.sizes)
    log(f"index on device: {ix.device_bytes / 1e9:.1f} GB + {corpus.rerank.numel() / 1e9:.1f} GB rows; "
        f"cluster slots mean {sizes.mean():.0f} max {sizes.max()}; build {t_build:.0f} s")
    st = ix.stream()

    def run(p):
        s, r, n = ix.search(q768, q256, p, stream=st)
        return s, r, n

    # Sweep: recall and a 2-step throughput estimate per configuration.
    sweep = []
    cfgs = [a.config] if a.config else (a.sweep.split(",") if a.sweep else W["sweep"])
    for cfg in cfgs:
        _, p = parse_config(cfg, ix.k, k=10)
        s, r, n = run(p)
        torch.cuda.synchronize()
        rec = recall10(r[:ngt].cpu().numpy(), n[:ngt].cpu().numpy().view(np.uint32), gt)
        t0 = time.time()
        for _ in range(2):
            run(p)
        torch.cuda.synchronize()
        qps = 2 * nq / (time.time() - t0)
        stt = st.last_stats()
        sweep.append(dict(config=cfg, recall_at_10=round(rec, 4), qps=round(qps, 1),
                          probed_per_query=round(stt["probed"] / nq, 1),
                          scanned_per_query=round(stt["scanned"] / nq, 0)))
        log(f"sweep {cfg}: R@10 {rec:.4f}, {qps:.0f} QPS, {stt['probed'] / nq:.1f} probes/query")
    ok = [x for x in sweep if x["recall_at_10"] >= 0.90]
    pick = max(ok, key=lambda x: x["qps"]) if ok else max(sweep, key=lambda x: x["recall_at_10"])
    _, p = parse_config(pick["config"], ix.k, k=10)

    def cm_settings(spec):
        v = [int(x) for x in spec.split(":")]
        return dict(seed_slots=v[0], ext_cap=v[1], chunk_blocks=v[2], phases=v[3])

    cm_sweep = []
    if a.cm_sweep:
        for setting in a.cm_sweep.split(","):
            ix.set_cluster_scan(**cm_settings(setting))
            for v in (10, 11):
                ix.set_scan_variant(v)
                run(p)
                ms, resc = [], 0
                for _ in range(3):
                    run(p)
                    stt = st.last_stats()
                    ms.append(stt["scan_ns"] / 1e6)
                    resc += stt["rescanned"]
                cm_sweep.append(dict(settings=setting, variant=v, scan_ms=round(min(ms), 3), rescanned=resc / 3,
                                     id_reads_per_query=round(stt["id_reads"] / nq, 1)))
                log(f"cluster scan {setting} variant {v}: scan {min(ms):.2f} ms, {resc / 3:.0f} rescans, "
                    f"{stt['id_reads'] / nq:.0f} past-threshold slots per query")
        ix.set_cluster_scan()
        ix.set_scan_variant(0)
    if a.cluster_scan:
        ix.set_cluster_scan(**cm_settings(a.cluster_scan))
    variant_ms = None
    if a.scan_variant is None:
        variant, variant_ms = ix.autotune(q768, q256, p, stream=st)
        variant_ms = [round(float(m), 3) for m in variant_ms]
        log(f"autotune: scan variant {variant} ({', '.join(f'{m:.2f}' for m in variant_ms)} ms)")
    else:
        variant = a.scan_variant
        ix.set_scan_variant(variant)
        log(f"scan variant {variant} (fixed)")
    same_as_per_query = None
    if variant >= 10:
        # The cluster-major scan must return the per-query scan's results bit for bit.
        got = [x.cpu().numpy() for x in run(p)]
        ix.set_scan_variant(2)
        want = [x.cpu().numpy() for x in run(p)]
        ix.set_scan_variant(variant)
        same_as_per_query = all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(got, want))
        log(f"results of scan variant {variant} equal the per-query scan's: {same_as_per_query}")
    agree = probe_agreement(ix, arrays.centroids, q256[:ngt], p, st)
    log(f"probe sets equal to the FP64 top-{p.nprobe}: {agree['equal']:.4f} of {ngt} queries "
        f"({agree['equal_up_to_ties']:.4f} up to f32-resolution ties)")

    # Timed region: W warmup + K steps of the batched search, inputs already in HBM. With
    # --streams N, N host threads each drive their own stream (the index is shared read-only,
    # as the C ABI allows), so up to N batches are on the GPU at once; the K steps are split
    # among them.
    import threading
    streams = [st] + [ix.stream() for _ in range(a.streams - 1)]
    for sx in streams:
        for _ in range(a.warmup):
            ix.search(q768, q256, p, stream=sx)
    torch.cuda.synchronize()
    step_stats, lock = [], threading.Lock()

    def worker(sx, count):
        for _ in range(count):
            ix.search(q768, q256, p, stream=sx)
            stt = sx.last_stats()
            with lock:
                step_stats.append(stt)

    counts = [a.steps // a.streams + (1 if i < a.steps % a.streams else 0) for i in range(a.streams)]
    workers = [threading.Thread(target=worker, args=(streams[i], counts[i])) for i in range(a.streams)]
    t0 = time.perf_counter()
    for w in workers:
        w.start()
    for w in workers:
        w.join()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / a.steps
    qps = nq * a.steps / wall
    if a.streams > 1:
        # Stage times of the kernels alone (the roofline): a few steps on one stream.
        step_stats = []
        for _ in range(3):
            worker(st, 1)
    k_steps = len(step_stats)

    def per_step(key):
        return sum(x[key] for x in step_stats) / k_steps

    scan_ns, route_ns, rerank_ns = per_step("scan_ns"), per_step("route_ns"), per_step("rerank_ns")
    scanned, id_reads, probed = per_step("scanned"), per_step("id_reads"), per_step("probed")
    rescanned, reranked = per_step("rescanned"), per_step("reranked")
    cm_pairs, cm_ns, cm_launches = per_step("cm_pairs"), per_step("cm_ns"), per_step("cm_launches")
    cm_staged = per_step("cm_staged")
    s, r, n = run(p)
    torch.cuda.synchronize()
    gpu_res = (s.cpu().numpy(), r.cpu().numpy().view(np.uint32), n.cpu().numpy().view(np.uint32))
    rec = recall10(gpu_res[1][:ngt], gpu_res[2][:ngt], gt)
    scan_path = "cluster-major" if variant >= 10 and nq >= 64 else "per-query"
    # Algorithmic bytes of the scan stage: 32 code + 4 mul bytes per (query, scanned slot), 8
    # (id, err) per slot past the threshold. The cluster-major path reads a code once for a tile
    # of up to 32 queries, so this per-(query, slot) stream can exceed what HBM moves.
    scan_bytes = scanned * 36 + id_reads * 8
    scan_s = scan_ns / 1e9
    effective = scan_bytes / scan_s / 1e9
    log(f"timed {a.steps} steps: {ms_per_step:.2f} ms/step, {qps:.0f} QPS, R@10 {rec:.4f}; "
        f"scan {scan_s * 1e3:.2f} ms ({effective:.0f} GB/s of per-(query, slot) bytes); route "
        f"{route_ns / 1e6:.2f} ms; rerank {rerank_ns / 1e6:.2f} ms")
    if scan_path == "cluster-major" and cm_launches > 0:
        # Dominant kernel: the MFMA pass k_scan_cm, bound by int8 MFMA issue (each code is read
        # once per query tile). achieved = useful int8 ops per launch / mean launch time.
        launch_s = cm_ns / 1e9 / cm_launches
        ops = CM_OPS_PER_PAIR * cm_pairs / cm_launches
        kname = "k_scan_cm2" if variant == 10 else "k_scan_cm<1>"
        tr = pmc_traffic(kname, name, pick["config"])
        achieved = ops / launch_s / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 1), "peak": MFMA_I8_PEAK_TOPS, "unit": "TOP/s",
                    "frac": round(achieved / MFMA_I8_PEAK_TOPS, 4),
                    "traffic": tr["bytes_per_launch"] if tr else None,
                    "kernel": kname, "mfma": "v_mfma_i32_16x16x64_i8 (dense int8)",
                    "ops_per_launch": int(ops), "ops_per_pair": CM_OPS_PER_PAIR,
                    "pairs_per_launch": int(cm_pairs / cm_launches), "launches_per_step": int(cm_launches),
                    "staged_fraction": round(cm_staged / cm_pairs, 5) if cm_pairs else None,
                    "mean_launch_ms": round(launch_s * 1e3, 4),
                    "dram_GBps": round(tr["bytes_per_launch"] / launch_s / 1e9, 1) if tr else None,
                    "traffic_profile": tr["profile"] if tr else None}
        log(f"k_scan_cm: staged {cm_staged / max(cm_pairs, 1):.4f} of the scored pairs for the exact estimate")
        log(f"k_scan_cm: {cm_launches:.0f} launches of {launch_s * 1e3:.2f} ms, {cm_pairs / cm_launches:.3g} pairs "
            f"each = {achieved:.0f} TOP/s int8 ({achieved / MFMA_I8_PEAK_TOPS:.3f} of peak)")
    else:
        tr = pmc_traffic("k_scan_t", name, pick["config"])
        roofline = {"bound": "hbm", "achieved": round(effective, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(effective / HBM_PEAK_GBS, 4), "traffic": tr["bytes_per_launch"] if tr else None,
                    "kernel": "k_scan_t", "algorithmic_bytes_per_launch": int(scan_bytes),
                    "mean_launch_ms": round(scan_s * 1e3, 4), "traffic_profile": tr["profile"] if tr else None}
    # The rerank gathers 768 B per reranked row (plus the 768-byte query).
    rr_bytes = reranked * 768 + nq * 768
    rerank = {"rows_per_step": int(reranked), "bytes_per_step": int(rr_bytes), "ms": round(rerank_ns / 1e6, 4),
              "GBps": round(rr_bytes / (rerank_ns / 1e9) / 1e9, 1) if rerank_ns else None,
              "frac_hbm": round(rr_bytes / (rerank_ns / 1e9) / 1e9 / HBM_PEAK_GBS, 4) if rerank_ns else None}

    # Single-query latency: the plain per-query GPU path (batch of one, synchronous).
    lat = []
    for i in range(min(a.latency_queries, nq)):
        t1 = time.perf_counter()
        ix.search(q768[i:i + 1], q256[i:i + 1], p, stream=st)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t1) * 1e3)
    lat.sort()
    from sieve_amd.evaluate import pct
    p50 = p99 = per_query_qps = None
    if lat:
        p50, p99 = round(pct(lat, 0.5), 3), round(pct(lat, 0.99), 3)
        per_query_qps = round(1e3 / float(np.mean(lat)), 1)
        log(f"single-query latency p50 {p50:.2f} ms p99 {p99:.2f} ms ({per_query_qps:.0f} QPS per-query path)")

    # The same through captured HIP graphs: fixed staging buffers, so every call replays one graph.
    glat = graphed_latency(ix, q768, q256, p, min(a.latency_queries, nq)) if a.latency_queries else []
    gp50 = gp99 = None
    if glat:
        gp50, gp99 = round(pct(glat, 0.5), 3), round(pct(glat, 0.99), 3)
        log(f"single-query latency through graphs p50 {gp50:.2f} ms p99 {gp99:.2f} ms ({len(glat)} replays)")

    serving = []
    ps = p
    if a.serve_config:
        _, ps = parse_config(a.serve_config, ix.k, k=10)
    if a.serve_seconds > 0:
        from sieve_amd.server import Server, open_loop
        q7h, q2h = q768[:2000].cpu().numpy(), q256[:2000].cpu().numpy()
        for rate in (float(x) for x in a.serve_rates.split(",")):
            sv = Server(ix, ps, max_batch=a.serve_max_batch)
            res = open_loop(sv, q7h, q2h, rate=rate, seconds=a.serve_seconds, truth=gt)
            res["slot_bytes"], res["slots"] = sv.stats()["slot_bytes"], sv.stats()["slots"]
            sv.close()
            serving.append(res)
            log(f"serving at {rate:.0f} QPS offered: {res['achieved_qps']:.0f} QPS, p50 {res['p50_ms']:.2f} ms, "
                f"p99 {res['p99_ms']:.2f} ms, mean batch {res['mean_batch']}, R@10 {res.get('recall_at_10')}")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(arrays, q768, q256, p, nq, a, gt, gpu_res)
    mixed = None
    if a.serve_mixed > 0:
        try:
            mixed = serve_mixed(ix, mix, W["rows"], q768, q256, ps, a)
        except Exception as e:  # reported, not fatal: the batched measurement above stands
            mixed = {"error": f"{type(e).__name__}: {e}"}
            log(f"mixed serving failed: {mixed['error']}")

    base_cfg = W.get("config_search")
    base_row = next((x for x in sweep if x["config"] == base_cfg), None) if base_cfg else None
    return {
        "metric": METRIC if name == "c3" else f"QPS & p50/p99 ms at R@10>=0.90 ({W['desc'].split(',')[0]})",
        "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8",
        "data": f"synthetic: seeded anisotropic Gaussian mixture on the unit sphere ({spec.weights} weights"
                + (f", {spec.topics} topics inside the components" if spec.topics else "")
                + "), queries = mixture draws + N(0,0.05^2) noise; index built on the GPU in the reference's format",
        "config": {"workload": W["desc"], "rows": W["rows"], "clusters": ix.k, "queries_per_step": nq,
                   "search": pick["config"], "search_rule": "the fastest swept configuration with R@10 >= 0.90",
                   "config_named_search": base_cfg,
                   "config_named_search_recall_at_10": base_row["recall_at_10"] if base_row else None,
                   "probes_per_query": round(probed / nq, 2), "k": 10, "mixture_spread": spread,
                   "topics": spec.topics, "topic_spread": spec.topic_spread},
        "recall_at_10": round(rec, 4),
        "p50_ms": p50, "p99_ms": p99,
        "p50_ms_graph": gp50, "p99_ms_graph": gp99,
        "per_query_gpu_qps": per_query_qps,
        "serving_config": a.serve_config or pick["config"],
        "serving_open_loop": serving,
        "serving_mixed": mixed,
        "streams": a.streams,
        "scan_variant": variant,
        "scan_variant_ms": variant_ms,
        "cluster_scan": a.cluster_scan,
        "results_equal_per_query_scan": same_as_per_query,
        "cluster_scan_sweep": cm_sweep or None,
        "rescanned_per_step": rescanned,
        "probe_agreement_fp64": agree,
        "stage_ms": {"route": round(route_ns / 1e6, 3), "scan": round(scan_s * 1e3, 3),
                     "rerank": round(rerank_ns / 1e6, 3)},
        "probed_per_query": round(probed / nq, 2),
        "scanned_per_query": round(scanned / nq, 1),
        "roofline": roofline,
        "scan_stage": {"path": scan_path, "ms": round(scan_s * 1e3, 3), "algorithmic_bytes": int(scan_bytes),
                       "effective_GBps": round(effective, 1),
                       "note": "per-(query, slot) bytes / scan stage time (seed, MFMA and merge kernels on the "
                               "cluster-major path); above the HBM peak when codes are shared across queries"},
        "rerank": rerank,
        "cpu_baseline": cpu,
        "sweep": sweep,
        "build_s": round(t_build, 1),
    }


def main_l2(a, W, name):
    """The L2 path (bigann::Index::search, src/bigann.rs:652-747) on the 100-d int8 workload:
    build on the GPU (not timed), R@10 against exact L2 search with ties counted as the
    reference's spacev harness does (src/bin/spacev.rs:108-114), the fastest nprobe:cand at
    R@10 >= 0.90 (and the first at >= 0.96), K timed batches with queries and outputs in HBM.
    Returns its record."""
    import numpy as np
    import torch
    from sieve_amd.evaluate import pct, recall10_ties
    from sieve_amd.index import L2Index, Stream
    from sieve_amd.l2build import L2Mixture, build_l2_large, exact_l2_topk
    t0 = time.time()
    n, nq = W["rows"], W["queries"]
    mix = L2Mixture(W["components"], seed=20241204, topics=W.get("topics", 0))
    log(f"workload {name}: generating {n} rows")
    base = torch.empty(n, 100, dtype=torch.int8, device="cuda")
    mix.fill(base)
    q = mix.queries(nq)
    arr = build_l2_large(base, coarse=W["leaves"][0], fine=W["leaves"][1], log=log)
    torch.cuda.synchronize()
    ngt = min(a.gt_queries, nq)
    gt_i, gt_s = exact_l2_topk(base, q[:ngt], k=32)
    gt_i, gt_s = gt_i.cpu().numpy(), gt_s.cpu().numpy()
    torch.cuda.empty_cache()
    ix = L2Index(dict(leaves=arr.leaves, block_off=arr.block_off, blocks=arr.blocks, ids=arr.ids,
                      factors=arr.factors, rotation=arr.rotation, base=base), borrow_base=True)
    host = None
    if not a.no_cpu:
        log("cpu baseline: copying the L2 index to host memory")
        host = arr.numpy()
    arr_block_off = arr.block_off
    del arr
    torch.cuda.empty_cache()
    t_build = time.time() - t0
    lsz = torch.diff(arr_block_off).float() * 32
    log(f"L2 index on device: {ix.k} leaves (blocks x 32 per leaf: mean {lsz.mean():.0f}, max {lsz.max():.0f}), "
        f"build {t_build:.0f} s")
    st = Stream()

    def recall(r, cnt):
        r = r[:ngt].cpu().numpy().view(np.uint32).astype(np.int64)[:, :10]
        return float(np.mean([recall10_ties(r[i, :cnt[i]], gt_i[i], gt_s[i]) for i in range(ngt)]))

    def run(cfg):
        npb, cand = (int(x) for x in cfg.split(":"))
        return ix.search(q, npb, cand, 10, stream=st)

    sweep = []
    cfgs = [a.config] if a.config else (a.sweep.split(",") if a.sweep else W["sweep"])
    for cfg in cfgs:
        s, r, cnt = run(cfg)
        torch.cuda.synchronize()
        rec = recall(r, cnt[:ngt].cpu().numpy().view(np.uint32))
        t1 = time.time()
        for _ in range(2):
            run(cfg)
        torch.cuda.synchronize()
        qps = 2 * nq / (time.time() - t1)
        stt = st.last_stats()
        sweep.append(dict(config=cfg, recall_at_10=round(rec, 4), qps=round(qps, 1),
                          scanned_per_query=round(stt["scanned"] / nq, 0)))
        log(f"sweep {cfg}: R@10 {rec:.4f} (ties counted), {qps:.0f} QPS, {stt['scanned'] / nq:.0f} slots/query")
    ok = [x for x in sweep if x["recall_at_10"] >= 0.90]
    pick = max(ok, key=lambda x: x["qps"]) if ok else max(sweep, key=lambda x: x["recall_at_10"])
    hi = next((x for x in sweep if x["recall_at_10"] >= 0.96), None)
    # The scan variant: the per-query variants timed on the picked configuration.
    variant_ms = []
    for v in range(10):
        ix.set_scan_variant(v)
        run(pick["config"])
        best = 1e30
        for _ in range(2):
            run(pick["config"])
            best = min(best, st.last_stats()["scan_ns"] / 1e6)
        variant_ms.append(round(best, 3))
    variant = int(np.argmin(variant_ms))
    ix.set_scan_variant(variant)
    log(f"L2 scan variant {variant} ({', '.join(f'{m:.2f}' for m in variant_ms)} ms)")
    for _ in range(a.warmup):
        run(pick["config"])
    torch.cuda.synchronize()
    scan_ns = route_ns = rerank_ns = 0.0
    scanned = id_reads = 0
    t1 = time.perf_counter()
    for _ in range(a.steps):
        run(pick["config"])
        stt = st.last_stats()
        scan_ns += stt["scan_ns"]
        route_ns += stt["route_ns"]
        rerank_ns += stt["rerank_ns"]
        scanned += stt["scanned"]
        id_reads += stt["id_reads"]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    qps = nq * a.steps / wall
    s, r, cnt = run(pick["config"])
    torch.cuda.synchronize()
    gpu_res = (s.cpu().numpy(), r.cpu().numpy().view(np.uint32), cnt.cpu().numpy().view(np.uint32))
    # 16 code bytes + [mul, |x|^2] per scanned slot, the id of each slot past the threshold.
    scan_bytes = (scanned * 24 + id_reads * 4) / a.steps
    scan_s = scan_ns / 1e9 / a.steps
    achieved = scan_bytes / scan_s / 1e9
    log(f"timed {a.steps} steps at {pick['config']}: {wall * 1e3 / a.steps:.2f} ms/step, {qps:.0f} QPS; scan "
        f"{scan_s * 1e3:.2f} ms = {achieved:.0f} GB/s; route {route_ns / 1e6 / a.steps:.2f} ms; rerank "
        f"{rerank_ns / 1e6 / a.steps:.2f} ms")
    lat = []
    npb, cand = (int(x) for x in pick["config"].split(":"))
    for i in range(min(a.latency_queries, nq)):
        t2 = time.perf_counter()
        ix.search(q[i:i + 1], npb, cand, 10, stream=st)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t2) * 1e3)
    lat.sort()
    p50 = round(pct(lat, 0.5), 3) if lat else None
    p99 = round(pct(lat, 0.99), 3) if lat else None
    if lat:
        log(f"single-query latency p50 {p50:.2f} ms p99 {p99:.2f} ms")
    cpu = None
    if host is not None:
        cpu = cpu_baseline_l2(host, q.cpu().numpy(), npb, cand, nq, a, (gt_i, gt_s), gpu_res)
        del host
    tr = pmc_traffic("k_scan_t_l2", name, pick["config"])
    return {
        "metric": "QPS & p50/p99 ms at R@10>=0.90 (1B x 100-d int8, L2); LUT-scan GB/s vs HBM peak",
        "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic: seeded 100-d int8 mixture (uniform component "
        "weights; 16 topics per component at spread 12, rows at spread 4 around a topic, clipped to "
        "[-127, 127]); queries = mixture draws + uniform integer noise in [-6, 6]",
        "config": {"workload": W["desc"], "rows": n, "leaves": ix.k, "queries_per_step": nq,
                   "search": pick["config"], "search_rule": "the fastest swept nprobe:cand with R@10 >= 0.90",
                   "k": 10},
        "recall_at_10": pick["recall_at_10"], "recall_0.96_config": hi, "p50_ms": p50, "p99_ms": p99,
        "stage_ms": {"route": round(route_ns / 1e6 / a.steps, 3), "scan": round(scan_s * 1e3, 3),
                     "rerank": round(rerank_ns / 1e6 / a.steps, 3)},
        "scanned_per_query": round(scanned / a.steps / nq, 1), "scan_variant": variant, "scan_variant_ms": variant_ms,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": tr["bytes_per_launch"] if tr else None,
                     "kernel": f"k_scan_t<32, true, ...> (L2 per-query scan, variant {variant})",
                     "algorithmic_bytes_per_launch": int(scan_bytes), "mean_launch_ms": round(scan_s * 1e3, 4),
                     "traffic_profile": tr["profile"] if tr else None},
        "cpu_baseline": cpu,
        "sweep": sweep, "build_s": round(t_build, 1),
    }


def cpu_baseline(arrays, q768, q256, p, nq, a, gt, gpu_res):
    """The oracle (C++ restatement of the reference's x86 path) on this host: closed-loop QPS with
    whole queries per worker thread over a bounded query sample (src/bin/search.rs:310-332,
    742-748), single-thread p50/p99 (src/bin/search.rs:730-738, percentile :400-402), and its
    results against the GPU's for the same queries."""
    import numpy as np
    from oracle import oracle as O
    from sieve_amd.evaluate import pct
    cpus = host_cpus()
    threads = cpus["threads"]
    log(f"cpu baseline: copying the index to host memory (host: {cpus})")
    host = arrays.numpy()
    oix = O.Index(host)
    q7 = q768.cpu().numpy()
    q2 = q256.cpu().numpy()
    probe_n = min(threads * 4, nq)
    t0 = time.perf_counter()
    oix.search_batch(q7[:probe_n], q2[:probe_n], p, threads=threads)
    per_q = (time.perf_counter() - t0) / probe_n
    m = int(min(nq, max(probe_n, a.cpu_seconds / max(per_q, 1e-6))))
    t0 = time.perf_counter()
    s, r, n, _ = oix.search_batch(q7[:m], q2[:m], p, threads=threads)
    wall = time.perf_counter() - t0
    k = min(m, len(gt))
    rec = recall10(r[:k], n[:k], gt[:k])
    parity = parity_counts(gpu_res[0][:m], gpu_res[1][:m], gpu_res[2][:m], s, r, n)
    log(f"cpu baseline: {m} queries on {threads} threads in {wall:.1f} s = {m / wall:.1f} QPS, R@10 {rec:.4f}; "
        f"vs the GPU: {parity}")
    nl = int(min(nq, max(20, a.cpu_latency_seconds / max(per_q * threads, 1e-6))))
    lat = np.sort(oix.search_latency(q7[:nl], q2[:nl], p).astype(np.float64) / 1e6)
    lp50, lp99 = round(pct(lat, 0.5), 3), round(pct(lat, 0.99), 3)
    log(f"cpu baseline single-thread latency over {nl} queries: p50 {lp50} ms, p99 {lp99} ms")
    return {"value": round(m / wall, 2), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"{m} of the {nq} queries, same index bytes and configuration, closed loop over "
                      f"{threads} threads ({wall:.1f} s)", "recall_at_10": round(rec, 4),
            "host": cpus, "single_thread_p50_ms": lp50, "single_thread_p99_ms": lp99, "latency_queries": nl,
            "parity_vs_oracle": parity}


def cpu_baseline_l2(host, qh, nprobe, cand, nq, a, gt, gpu_res):
    """The oracle's bigann::Index::search on this host for the L2 workload: closed-loop QPS over
    a bounded query sample, single-thread p50/p99, and its results against the GPU's."""
    import numpy as np
    from oracle import oracle as O
    from sieve_amd.evaluate import pct, recall10_ties
    cpus = host_cpus()
    threads = cpus["threads"]
    oix = O.L2Index(host["leaves"], host["block_off"], host["blocks"], host["factors"], host["ids"],
                    host["rotation"], host["base"])
    probe_n = min(threads * 2, nq)
    t0 = time.perf_counter()
    oix.search_batch(qh[:probe_n], nprobe, cand, 10, threads=threads)
    per_q = (time.perf_counter() - t0) / probe_n
    m = int(min(nq, max(probe_n, a.cpu_seconds / max(per_q, 1e-6))))
    t0 = time.perf_counter()
    s, r, n = oix.search_batch(qh[:m], nprobe, cand, 10, threads=threads)
    wall = time.perf_counter() - t0
    gt_i, gt_s = gt
    k = min(m, len(gt_i))
    rec = float(np.mean([recall10_ties(r[i, :n[i]].astype(np.int64), gt_i[i], gt_s[i]) for i in range(k)]))
    parity = parity_counts(gpu_res[0][:m], gpu_res[1][:m], gpu_res[2][:m], s, r, n)
    log(f"cpu baseline (L2): {m} queries on {threads} threads in {wall:.1f} s = {m / wall:.1f} QPS, "
        f"R@10 {rec:.4f}; vs the GPU: {parity}")
    nl = int(min(nq, max(10, a.cpu_latency_seconds / max(per_q * threads, 1e-6))))
    lat = np.sort(oix.search_latency(qh[:nl], nprobe, cand, 10).astype(np.float64) / 1e6)
    lp50, lp99 = round(pct(lat, 0.5), 3), round(pct(lat, 0.99), 3)
    return {"value": round(m / wall, 2), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"{m} of the {nq} queries, same index bytes, {nprobe}:{cand}, closed loop over {threads} "
                      f"threads ({wall:.1f} s)", "recall_at_10": round(rec, 4), "host": cpus,
            "single_thread_p50_ms": lp50, "single_thread_p99_ms": lp99, "latency_queries": nl,
            "parity_vs_oracle": parity}


def serve_mixed(ix, mix, n, q768, q256, p, a):
    """Open-loop Poisson arrivals through the serving loop with a live state attached, at each
    offered rate: 30% of queries carry one of six allowlist filters (selectivity 0.1% .. 100%,
    log-uniform steps; filter::search, plan resolved by size), and a writer inserts 1000 fresh
    rows every 10 ms (100k rows/s) and deletes a tenth of them. Latency percentiles per request
    kind; before the load, the time of one 1000-row insert and one 100-row delete alone."""
    import threading
    import numpy as np
    import torch
    from sieve_amd import abi
    from sieve_amd.build import normalize_rows
    from sieve_amd.evaluate import pct
    from sieve_amd.index import Filter, Live
    from sieve_amd.server import Server
    sel = [0.001, 0.004, 0.016, 0.063, 0.25, 1.0]
    rng = np.random.default_rng(9)
    filters, plans = [], []
    for x in sel:
        mask = (torch.rand(n, generator=torch.Generator().manual_seed(int(x * 1e6))) < x).numpy()
        words = np.packbits(np.concatenate([mask, np.zeros((-n) % 64, bool)]), bitorder="little").view(np.uint64)
        f = Filter(ix, words)
        filters.append(f)
        plans.append(int(abi.lib().sieve_plan_resolve(abi.PLANS["auto"], f.rows(), n)))
    log(f"mixed serving: filters of {[f.rows() for f in filters]} rows, plans {plans}")
    secs = a.serve_mixed
    rates = [float(x) for x in a.serve_mixed_rates.split(",")]
    batch_rows, period = 1000, 0.01
    nw = int(secs / period) + 2
    live = Live(ix, max_rows=(len(rates) * nw + 2) * batch_rows)
    fresh, _ = normalize_rows(mix.rows(n, (len(rates) * nw + 1) * batch_rows).to("cuda"))
    fresh = fresh.cpu().numpy().astype(np.float32)
    q7h, q2h = q768[:2000].cpu().numpy(), q256[:2000].cpu().numpy()
    out = {"filter_selectivity": sel, "filter_plans": plans, "seconds_per_rate": secs,
           "filtered_fraction": a.serve_mixed_filtered, "writes": bool(a.serve_mixed_writes), "rates": []}
    # One write of each kind alone (the device encoder, arena rebuild and tombstone upload).
    ext0 = (2_900_000_000 + np.arange(batch_rows)).astype(np.uint32)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    live.insert(ext0, fresh[:batch_rows])
    out["insert_1000_rows_ms"] = round((time.perf_counter() - t1) * 1e3, 3)
    t1 = time.perf_counter()
    live.delete(ext0[:batch_rows // 10])
    out["delete_100_rows_ms"] = round((time.perf_counter() - t1) * 1e3, 3)
    used = 1
    for rate in rates:
        # high offered rates run shorter: the backlog a saturated mix builds is drained after
        secs = a.serve_mixed if rate <= 3000 else min(a.serve_mixed, 10000.0 / rate)
        sv = Server(ix, p, max_batch=a.serve_max_batch, live=live)
        lat = {"plain": [], "filtered": [], "write": []}
        tickets, lock, done = [], threading.Lock(), threading.Event()
        base = used

        def writer():
            j = 0
            t0w = time.perf_counter()
            while a.serve_mixed_writes and not done.is_set() and j < nw:
                b = base + j
                ext = (3_000_000_000 + b * batch_rows + np.arange(batch_rows)).astype(np.uint32)
                tk = sv.insert(ext, fresh[b * batch_rows:(b + 1) * batch_rows])
                tk2 = sv.delete(ext[rng.choice(batch_rows, batch_rows // 10, replace=False)])
                with lock:
                    tickets.append(("write", tk))
                    tickets.append(("write", tk2))
                j += 1
                time.sleep(max(0.0, t0w + j * period - time.perf_counter()))

        wt = threading.Thread(target=writer)
        wt.start()
        t0 = time.perf_counter()
        nxt, i = t0, 0
        while time.perf_counter() - t0 < secs:
            now = time.perf_counter()
            if now < nxt:
                time.sleep(min(nxt - now, 0.001))
                continue
            j = i % len(q7h)
            if rng.random() < a.serve_mixed_filtered:
                tk = sv.submit_filtered(q7h[j], q2h[j], filters[rng.integers(len(filters))])
                kind = "filtered"
            else:
                tk = sv.submit(q7h[j], q2h[j])
                kind = "plain"
            with lock:
                tickets.append((kind, tk))
            i += 1
            nxt += rng.exponential(1.0 / rate)
        done.set()
        wt.join()
        used = base + nw
        for kind, tk in tickets:
            _, _, info = sv.wait(tk)
            lat[kind].append(info.total_ms)
        wall = time.perf_counter() - t0
        st = sv.stats()
        sv.close()
        r = {"offered_qps": rate, "achieved_qps": round((len(lat["plain"]) + len(lat["filtered"])) / wall, 1),
             "seconds_offered": round(secs, 2), "seconds_to_drain": round(wall, 2), "writes": st["writes"]}
        for k, v in lat.items():
            v.sort()
            if v:
                r[k] = {"requests": len(v), "p50_ms": round(pct(v, 0.5), 3), "p99_ms": round(pct(v, 0.99), 3)}
        out["rates"].append(r)
        log(f"mixed serving at {rate:.0f} QPS offered: {r}")
    info = live.info()
    out["rows_inserted"], out["tombstones"] = info["rows"], info["deleted"]
    live.close()
    return out


def graphed_latency(ix, q768, q256, p, n):
    """Single-query latency on a stream with sieve_stream_set_graphs: each query is copied into
    fixed device staging buffers on the library's own stream, so every call repeats one graph
    key. Returns the latencies of the replayed calls (ms)."""
    import torch
    from sieve_amd.index import Stream
    gst = Stream()
    gst.set_graphs(True)
    ext = torch.cuda.ExternalStream(gst.handle)
    s7 = torch.empty(1, q768.shape[1], dtype=torch.float32, device="cuda")
    s2 = torch.empty(1, q256.shape[1], dtype=torch.float32, device="cuda")
    out = (torch.empty(1, p.k, dtype=torch.float32, device="cuda"),
           torch.empty(1, p.k, dtype=torch.int32, device="cuda"), torch.empty(1, dtype=torch.int32, device="cuda"))
    torch.cuda.synchronize()
    lat = []
    for i in range(n + 2):  # the first call runs eagerly, the second captures
        j = i % q256.shape[0]
        t1 = time.perf_counter()
        with torch.cuda.stream(ext):
            s7.copy_(q768[j:j + 1])
            s2.copy_(q256[j:j + 1])
        ix.search(s7, s2, p, stream=gst, out=out)
        gst.sync()
        dt = (time.perf_counter() - t1) * 1e3
        if gst.last_stats()["graph"] == 2:
            lat.append(dt)
    lat.sort()
    return lat


def copy_probe():
    """One elementwise pass over a 1 GiB float buffer: reads 2^30 bytes, writes 2^30 bytes."""
    import torch
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe": {"kernel": "elementwise neg", "launches": 3, "bytes_read": 1 << 30,
                                     "bytes_written": 1 << 30}}), flush=True)


def probe_agreement(ix, centroids, q256, p, st):
    """Probe sets of the BF16-MFMA route (with its FP32 rescoring band) against the FP64
    top-nprobe of <q, c>, at the configuration's nprobe without margin or budget. `equal`:
    identical sets; `equal_up_to_ties`: every differing centroid scores within the f32 error
    bound of a 256-term sequential dot (gamma_256 * sum |q_i c_i|) of the FP64 cut score."""
    import dataclasses
    import torch
    pf = dataclasses.replace(p, margin=None, budget=None)
    _, c, n = ix.route(q256, pf, stream=st)
    torch.cuda.synchronize()
    q64, c64 = q256.double(), centroids.double()
    s64 = q64 @ c64.T
    bound = (q64.abs() @ c64.abs().T) * (256 * 2.0 ** -24 / (1 - 256 * 2.0 ** -24))
    top = torch.topk(s64, pf.nprobe, dim=1)
    cut = top.values[:, -1:]
    got = torch.zeros_like(s64, dtype=torch.bool)
    got.scatter_(1, c.long(), True)
    want = torch.zeros_like(got)
    want.scatter_(1, top.indices, True)
    diff = got ^ want
    near = (s64 - cut).abs() <= bound + bound.gather(1, top.indices[:, -1:])
    equal = (~diff).all(1) & (n.long() == pf.nprobe)
    ties = (~diff | near).all(1) & (n.long() == pf.nprobe)
    return {"queries": int(q256.shape[0]), "nprobe": pf.nprobe, "equal": round(float(equal.double().mean()), 5),
            "equal_up_to_ties": round(float(ties.double().mean()), 5)}




if __name__ == "__main__":
    main()
