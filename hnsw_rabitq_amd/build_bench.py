"""GPU construction measurement on the 1M x 128 L2 workload: wall time of add_items_gpu (levels
above 0 on the host cores, level 0 on the GPU) beside the CPU build, and recall@10 of the two
graphs at the same ef through the same GPU search."""
import time

import numpy as np


def _recall(lab, gt):
    hits = 0
    for a, b in zip(lab, gt):
        hits += len(np.intersect1d(a, b))
    return hits / (gt.shape[0] * gt.shape[1])


def run(H, cpu_index, data, queries, gt, cpu_build_s, log, efs=(160, 400, 1600), num_threads=0):
    import os
    nt = num_threads or os.cpu_count() or 1
    n, dim = data.shape
    t = time.perf_counter()
    g = H.Index("l2", dim)
    g.init_index(n, M=16, ef_construction=200, random_seed=100)
    g.add_items_gpu(data, np.arange(n, dtype=np.uint64), num_threads=nt)
    build_s = time.perf_counter() - t
    log(f"GPU build: {build_s:.1f} s for {n} x {dim} (CPU build {cpu_build_s if cpu_build_s else float('nan'):.1f} s "
        f"on {nt} threads)")
    out = {"gpu_build_s": build_s, "cpu_build_s": cpu_build_s, "build_threads": nt, "gpu_build_recall": []}
    for ef in efs:
        cpu_index.set_ef(ef)
        g.set_ef(ef)
        rc = _recall(cpu_index.knn_query(queries, k=10)[0], gt)
        rg = _recall(g.knn_query(queries, k=10)[0], gt)
        log(f"  ef={ef}: recall@10 cpu-built {rc:.4f} gpu-built {rg:.4f}")
        out["gpu_build_recall"].append(dict(ef=ef, cpu_built=rc, gpu_built=rg))
    return out
