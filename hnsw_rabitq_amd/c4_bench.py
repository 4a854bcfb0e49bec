"""Uneven batched search: inner-product GMM base (4,096 centres, sigma 0.5, unit norm), queries
with per-query ef drawn log-uniform in [16, 512] and k uniform over {1, 10, 100}, all in one
persistent launch captured in a hipGraph and replayed; launch configuration from the autotuner.

    python -m hnsw_rabitq_amd.c4_bench [--n 10000000] [--nq 100000] [--dim 96] [--reps 5]

Prints one JSON line: QPS of the graph replay, per-k recall against exact FP32 brute force on
a sample of queries, the plan, build time.  The default size is the full configuration
(10M x 96, 100k queries); --n scales it down.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--nq", type=int, default=100_000)
    ap.add_argument("--dim", type=int, default=96)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--recall-sample", type=int, default=2000)
    ap.add_argument("--index-cache", default="")
    ap.add_argument("--gpu-build", action="store_true", help="build the graph on the GPU (add_items_gpu)")
    args = ap.parse_args()
    import hnsw_rabitq_amd as H
    from hnsw_rabitq_amd import DeviceArray, _lib, datasets

    t0 = time.time()

    def log(m):
        print(f"[c4 {time.time() - t0:7.1f}s] {m}", file=sys.stderr, flush=True)

    log(f"generating {args.n} x {args.dim} GMM + {args.nq} queries")
    data, queries = datasets.c4_gmm(n=args.n, nq=args.nq, dim=args.dim)
    ef, k = datasets.c4_ef_k(args.nq)
    # larger ef first: the persistent workgroups take queries in order and then finish together
    order = np.argsort(-ef, kind="stable")
    queries, ef, k = np.ascontiguousarray(queries[order]), np.ascontiguousarray(ef[order]), np.ascontiguousarray(k[order])
    idx = H.Index("ip", args.dim)
    tb = time.time()
    if args.index_cache and os.path.exists(args.index_cache):
        idx.load_index(args.index_cache)
    else:
        idx.init_index(args.n, M=16, ef_construction=200, random_seed=100)
        if args.gpu_build:
            idx.add_items_gpu(data, np.arange(args.n, dtype=np.uint64))
            log(f"  inserted {args.n} (graph built on the GPU)")
        else:
            step = 500_000
            for s in range(0, args.n, step):
                idx.add_items(data[s:s + step], np.arange(s, min(args.n, s + step), dtype=np.uint64))
                log(f"  inserted {min(args.n, s + step)}")
        if args.index_cache:
            idx.save_index(args.index_cache)
    build_s = time.time() - tb
    gpu = idx.gpu()
    log(f"GPU index resident: {gpu.nbytes() / 1e9:.2f} GB")
    nq, kmax = args.nq, int(k.max())
    stream = ctypes.c_void_p()
    _lib.call("hg_stream_create", ctypes.byref(stream))
    dq = DeviceArray.from_numpy(queries)
    def_, dk = DeviceArray.from_numpy(ef), DeviceArray.from_numpy(k)
    dl, dd = DeviceArray(nq * kmax * 8), DeviceArray(nq * kmax * 4)
    gpu.autotune(dq, min(nq, 20_000), int(ef.max()), kmax, stream=stream)
    plan = gpu.plan(int(ef.max()), kmax, nq)
    log(f"plan {plan}")
    graph = ctypes.c_void_p()
    _lib.call("hg_graph_capture_search", gpu.handle, dq.ptr, nq, def_.ptr, dk.ptr, int(ef.max()), kmax, kmax,
              dl.ptr, dd.ptr, None, None, stream, ctypes.byref(graph))
    timer = ctypes.c_void_p()
    _lib.call("hg_timer_create", ctypes.byref(timer))
    _lib.call("hg_graph_launch", graph, stream)
    ms = []
    for _ in range(args.reps):
        _lib.call("hg_timer_start", timer, stream)
        _lib.call("hg_graph_launch", graph, stream)
        t = ctypes.c_float()
        _lib.call("hg_timer_stop", timer, stream, ctypes.byref(t))
        ms.append(t.value)
    lab = dl.to_numpy(np.uint64, (nq, kmax))
    # per-query counters of one more launch of the same batch (same rows as the replays), for the
    # algorithmic bytes: query, one upper link row (4*M) per upper hop, one level-0 link row
    # (4*2M) per expansion, one vector row per distance evaluation, the k results
    dc = DeviceArray(nq * 16)
    gpu.search_device(dq, nq, kmax, int(ef.max()), dl, dd, d_counters=dc, d_ef=def_, d_k=dk, stream=stream)
    _lib.call("hg_stream_synchronize", stream)
    ctr = dc.to_numpy(np.int32, (nq, 4)).astype(np.int64)
    vb = 1 if gpu.vector_code == "u8" else 4
    m = 16
    abytes = float((4 * args.dim + 4 * m * ctr[:, 2] + vb * args.dim * ctr[:, 3] + 8 * m * ctr[:, 0]
                    + vb * args.dim * ctr[:, 1] + 12 * k.astype(np.int64)).sum())
    # recall against exact brute force on a sample of queries
    # (every (nq/ns)-th query of the ef-sorted batch, so every ef range is represented)
    ns = min(args.recall_sample, nq)
    sample = np.linspace(0, nq - 1, ns).astype(np.int64)
    bf = H.BFIndex("ip", args.dim)
    bf.init_index(args.n)
    bf.add_items(data, np.arange(args.n, dtype=np.uint64))
    gt, _ = bf.knn_query(queries[sample], k=kmax)
    rec = {}
    for kv in sorted(set(k.tolist())):
        sel = np.nonzero(k[sample] == kv)[0]
        hits = [len(np.intersect1d(lab[sample[i], :kv], gt[i, :kv])) for i in sel]
        rec[str(kv)] = float(np.sum(hits) / (len(sel) * kv)) if len(sel) else None
    t = float(np.median(ms))
    out = dict(workload=f"hnsw_ip_gmm_{args.n}x{args.dim}_q{nq}_ef16-512_k1-10-100", qps=nq / (t * 1e-3),
               ms_per_batch=t, graph_replays=args.reps, plan=plan, recall_by_k=rec, build_s=build_s,
               build="gpu" if args.gpu_build else "cpu",
               ef_mean=float(ef.mean()), k_counts={str(kv): int((k == kv).sum()) for kv in sorted(set(k.tolist()))},
               expansions_per_query=float(ctr[:, 0].mean()), evaluations_per_query=float(ctr[:, 1].mean()),
               vector_rows=gpu.vector_code, algorithmic_bytes_per_batch=abytes,
               algorithmic_gbs=abytes / (t * 1e-3) / 1e9, frac_hbm=abytes / (t * 1e-3) / 8e12)
    _lib.call("hg_graph_free", graph)
    print(json.dumps(out), flush=True)


def run_child(log, gpu_build=True, reps=5, timeout=900):
    """The full uneven batch in a child process (its 10M-row index and brute-force copy are freed
    with it); its progress lines go to this process's stderr as they come.  Returns the child's
    JSON result, or {} when it fails."""
    import subprocess
    cmd = [sys.executable, "-m", "hnsw_rabitq_amd.c4_bench", "--reps", str(reps)]
    if gpu_build:
        cmd.append("--gpu-build")
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=timeout, cwd=ROOT)
    except subprocess.TimeoutExpired:
        log("  uneven batch: child process timed out")
        return {}
    if r.returncode != 0 or not r.stdout.strip():
        log(f"  uneven batch: child process exited with {r.returncode}")
        return {}
    return json.loads(r.stdout.strip().splitlines()[-1])


if __name__ == "__main__":
    main()
