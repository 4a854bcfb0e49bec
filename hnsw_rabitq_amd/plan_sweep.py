"""Launch-plan sweep of the HNSW search: QPS against resident waves per CU.

Builds the 1M x 128 L2 Gamma index on the GPU (M=16, ef_construction=200), then times the
search of 10,000 queries (k=10) at fixed hash slots for a range of workgroups (one wave, one
query each) per CU.  When the time stops falling before the LDS or VGPR limit on resident waves,
extra waves -- the resource a second wave per query would add -- buy no throughput.

    python -m hnsw_rabitq_amd.plan_sweep [--ef 1344,160] [--per-cu 2,4,6,8,10,12,13,16,24,32]
"""
import argparse
import ctypes
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--ef", default="1344,160")
    ap.add_argument("--per-cu", default="2,4,6,8,10,12,13,16,24,32")
    ap.add_argument("--hash-slots", type=int, default=128)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()

    import hnsw_rabitq_amd as H
    from hnsw_rabitq_amd import _lib, datasets, DeviceArray

    t0 = time.time()
    k = 10
    data = datasets.c2_gamma(n=args.n)
    queries = datasets.c2_gamma(n=args.nq, seed=1)
    idx = H.Index("l2", 128)
    idx.init_index(args.n, M=16, ef_construction=200, random_seed=100)
    idx.add_items_gpu(data)
    print(f"[{time.time() - t0:6.1f}s] index built on the GPU ({args.n} x 128)", flush=True)
    gpu = idx.gpu()
    stream = ctypes.c_void_p()
    _lib.call("hg_stream_create", ctypes.byref(stream))
    timer = ctypes.c_void_p()
    _lib.call("hg_timer_create", ctypes.byref(timer))
    nq = args.nq
    dq = DeviceArray.from_numpy(queries)
    dl, dd, dc = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4), DeviceArray(nq * 16)
    for ef in (int(x) for x in args.ef.split(",")):
        for pc in (int(x) for x in args.per_cu.split(",")):
            gpu.set_plan(ef, k, args.hash_slots, pc)
            gpu.search_device(dq, nq, k, ef, dl, dd, d_counters=dc, stream=stream)
            ms = []
            for _ in range(args.reps):
                _lib.call("hg_timer_start", timer, stream)
                gpu.search_device(dq, nq, k, ef, dl, dd, stream=stream)
                m = ctypes.c_float()
                _lib.call("hg_timer_stop", timer, stream, ctypes.byref(m))
                ms.append(m.value)
            c = dc.to_numpy(np.int32, (nq, 4))
            pl = gpu.plan(ef, k, nq)
            med = float(np.median(ms))
            print(f"ef={ef} hash_slots={args.hash_slots} per_cu={pc}: workgroups={pl['workgroups']} "
                  f"lds/wave={pl['lds_bytes']} B  {med:.2f} ms = {nq / med * 1e3:,.0f} QPS  "
                  f"(expansions {c[:, 0].mean():.0f}, evaluations {c[:, 1].mean():.0f} per query)", flush=True)
    _lib.call("hg_timer_destroy", timer)
    return 0


if __name__ == "__main__":
    sys.exit(main())
