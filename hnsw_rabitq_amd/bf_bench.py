"""Brute-force measurement on the 1M x 768 inner-product workload (unit-normalised Gaussian
mixture, 1,024 centres, sigma 0.35; 10,000 queries; k=100): exact FP32, and BF16 / FP8 MFMA
distance GEMMs with the top-4,096 candidates re-ranked in exact FP32.  Reports GEMM TFLOP/s
(2*Q*N*D over the all-rows GEMM time), end-to-end time, recall@100 against exact FP32 and the
FP64 check of the final lists."""
import ctypes

import numpy as np

from . import datasets


def run(H, lib, n, nq, stream, timer, log, dim=768, k=100, c=4096, library=True):
    from . import DeviceArray
    log(f"brute force: generating {n} x {dim} GMM + {nq} queries")
    data, q = datasets.c3_gmm(n=n, nq=nq, dim=dim)
    bf = H.BFIndex("ip", dim)
    bf.init_index(n)
    bf.add_items(data, np.arange(n, dtype=np.uint64))
    del data
    dq = DeviceArray.from_numpy(q)
    el, ed = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4)
    timer.start()
    lib.call("hg_bf_search_device", bf._h, dq.ptr, nq, k, lib.PREC_FP32, 0, el.ptr, ed.ptr, None, stream)
    exact_ms = timer.stop()
    exact = el.to_numpy(np.uint64, (nq, k))
    flops = 2.0 * nq * n * dim
    out = dict(bf_config=f"ip_gmm_{n}x{dim}_q{nq}_k{k}_c{c}", bf_exact_ms=exact_ms,
               bf_exact_tflops=flops / (exact_ms * 1e-3) / 1e12)
    log(f"  exact FP32: {exact_ms:.1f} ms ({out['bf_exact_tflops']:.1f} TFLOP/s)")
    # the autotuner times every GEMM tile configuration on these queries and keeps the fastest
    # for this index and precision
    for name, prec in (("bf16", lib.PREC_BF16), ("fp8", lib.PREC_FP8)):
        ll, ld = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4)
        times = np.zeros(lib.GEMM_VARIANTS, np.float64)
        chosen = ctypes.c_int()
        lib.call("hg_bf_autotune", bf._h, dq.ptr, nq, k, prec, c, ll.ptr, ld.ptr,
                 ctypes.c_void_p(times.ctypes.data), ctypes.byref(chosen), stream)
        got = ll.to_numpy(np.uint64, (nq, k))
        for v in range(lib.GEMM_VARIANTS):
            log(f"  variant {v} {name}: GEMM {times[v]:.2f} ms = {flops / (times[v] * 1e-3) / 1e12:.0f} TFLOP/s")
            out[f"bf_{name}_gemm_ms_variant{v}"] = float(times[v])
        out[f"bf_{name}_gemm_variant"] = chosen.value
        log(f"  autotuned {name} variant {chosen.value}; last run's rows identical to exact "
            f"{float((got == exact).all(1).mean()):.4f}")
    for name, prec, peak in (("bf16", lib.PREC_BF16, 2500.0), ("fp8", lib.PREC_FP8, 5000.0)):
        ll, ld = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4)
        stats = np.zeros(4, np.float64)
        # first call builds the low-precision copy of the rows
        lib.call("hg_bf_search_device", bf._h, dq.ptr, nq, k, prec, c, ll.ptr, ld.ptr,
                 ctypes.c_void_p(stats.ctypes.data), stream)
        timer.start()
        lib.call("hg_bf_search_device", bf._h, dq.ptr, nq, k, prec, c, ll.ptr, ld.ptr,
                 ctypes.c_void_p(stats.ctypes.data), stream)
        e2e = timer.stop()
        got = ll.to_numpy(np.uint64, (nq, k))
        rec = float(np.mean([len(np.intersect1d(a, b)) for a, b in zip(got, exact)]) / k)
        same = float((got == exact).all(1).mean())
        rep = bf.lowp_report()
        chk = np.zeros(4, np.float64)
        lib.call("hg_bf_fp64_check", bf._h, dq.ptr, nq, k, ll.ptr, ld.ptr, ctypes.c_double(1e-6),
                 ctypes.c_void_p(chk.ctypes.data), stream)
        gemm_tf = flops / (stats[0] * 1e-3) / 1e12
        out.update({f"bf_{name}_gemm_ms": stats[0], f"bf_{name}_gemm_tflops": gemm_tf,
                    f"bf_{name}_gemm_frac_peak": gemm_tf / peak, f"bf_{name}_prep_ms": stats[1],
                    f"bf_{name}_rerank_ms": stats[2], f"bf_{name}_e2e_ms": e2e,
                    f"bf_{name}_e2e_tflops": flops / (e2e * 1e-3) / 1e12, f"bf_{name}_recall100": rec,
                    f"bf_{name}_rows_identical": same, f"bf_{name}_retried_queries": stats[3],
                    f"bf_{name}_fp64_neartie_pairs": chk[0], f"bf_{name}_fp64_order_disagree": chk[1],
                    f"bf_{name}_fp64_max_abs_err": chk[2], f"bf_{name}_fp64_max_rel_err": chk[3],
                    f"bf_{name}_exact_path_overflow": rep["exact_overflow"],
                    f"bf_{name}_exact_path_uncertified": rep["exact_uncertified"],
                    f"bf_{name}_certified_queries": nq - rep["exact_overflow"] - rep["exact_uncertified"],
                    f"bf_{name}_error_bound_max": rep["max_bound"],
                    f"bf_{name}_rerank_neartie_pairs": rep["near_ties"],
                    f"bf_{name}_rerank_neartie_fp64_disagree": rep["near_tie_disagree"],
                    f"bf_{name}_boundary_neartie_pairs": rep["boundary_near_ties"],
                    f"bf_{name}_boundary_fp64_disagree": rep["boundary_disagree"],
                    f"bf_{name}_reranked_rows_per_query": rep["reranked_rows"] / nq})
        log(f"  {name}: GEMM {stats[0]:.2f} ms = {gemm_tf:.0f} TFLOP/s ({100 * gemm_tf / peak:.1f}% of peak); "
            f"prep {stats[1]:.2f} ms, re-rank {stats[2]:.2f} ms, end-to-end {e2e:.2f} ms; recall@100 {rec:.4f}; "
            f"rows identical to exact {same:.4f}; fp64 near-tie pairs {chk[0]:.0f}, disagreements {chk[1]:.0f}; "
            f"certified {nq - rep['exact_overflow'] - rep['exact_uncertified']} of {nq} (bound <= "
            f"{rep['max_bound']:.4f}), exact path {rep['exact_overflow']} overflow + {rep['exact_uncertified']} "
            f"uncertified; k-th boundary near ties {rep['boundary_near_ties']}, FP64 disagreements "
            f"{rep['boundary_disagree']}; exact FP32 rows per query {rep['reranked_rows'] / nq:.1f}")
    if library:
        out.update(library_gemm(n, nq, dim, log))
    return out


def library_gemm(n, nq, dim, log):
    """The vendor GEMM on the same Q x N x D shape (`library_gemm_run`), in a child process:
    torch carries its own HIP runtime, which cannot share a process with this library's."""
    import json
    import subprocess
    import sys
    try:
        r = subprocess.run([sys.executable, "-m", "hnsw_rabitq_amd.bf_bench", "--library-gemm", "--n", str(n),
                            "--nq", str(nq), "--dim", str(dim)], capture_output=True, text=True, timeout=600)
    except subprocess.TimeoutExpired:
        log("  library GEMM: child process timed out")
        return {}
    for line in r.stderr.splitlines():
        if "library" in line:
            log(line.split("] ", 1)[-1])
    if r.returncode != 0 or not r.stdout.strip():
        log(f"  library GEMM: child process exited with {r.returncode}")
        return {}
    return json.loads(r.stdout.strip().splitlines()[-1])


def library_gemm_run(n, nq, dim, log, chunk=1 << 17):
    """The vendor GEMM (torch.matmul -> hipBLASLt; torch._scaled_mm for FP8 e4m3) on the same
    Q x N x D shape, random operands, low-precision output written in row chunks of `chunk`: a
    plain GEMM with no selection, reported beside the distance GEMMs as the library's level on
    this shape.  Skipped when torch is not importable."""
    try:
        import torch
    except ImportError:
        return {}
    dev = "cuda:0"
    g = torch.Generator(device=dev).manual_seed(0)
    qb = torch.randn(nq, dim, device=dev, generator=g).to(torch.bfloat16)
    xb = torch.randn(n, dim, device=dev, generator=g).to(torch.bfloat16)
    res = {w: torch.empty(nq, w, device=dev, dtype=torch.bfloat16) for w in {min(chunk, n), n % chunk or chunk}}
    one = torch.ones((), device=dev)
    q8, x8 = qb.to(torch.float8_e4m3fn), xb.to(torch.float8_e4m3fn)

    def bf16():
        for c in range(0, n, chunk):
            w = min(chunk, n - c)
            torch.matmul(qb, xb[c:c + w].t(), out=res[w])

    def fp8():
        for c in range(0, n, chunk):
            w = min(chunk, n - c)
            torch._scaled_mm(q8, x8[c:c + w].t(), scale_a=one, scale_b=one, out_dtype=torch.bfloat16, out=res[w])

    out = {}
    flops = 2.0 * nq * n * dim
    for name, fn, peak in (("bf16", bf16, 2500.0), ("fp8", fp8, 5000.0)):
        try:
            fn()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record()
            for _ in range(3):
                fn()
            ev1.record()
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / 3
        except RuntimeError as e:  # e.g. no FP8 GEMM for this device in the installed library
            log(f"  library {name} GEMM unavailable: {str(e)[:120]}")
            continue
        tf = flops / (ms * 1e-3) / 1e12
        out[f"bf_library_{name}_gemm_ms"] = ms
        out[f"bf_library_{name}_gemm_frac_peak"] = tf / peak
        log(f"  library {name} GEMM (plain, bf16 output, no selection): {ms:.2f} ms = {tf:.0f} TFLOP/s "
            f"({100 * tf / peak:.1f}% of peak)")
    del qb, xb, res, q8, x8
    torch.cuda.empty_cache()
    return out


def main():
    """Standalone run of the brute-force measurement (one JSON line on stdout):

        python -m hnsw_rabitq_amd.bf_bench [--n 1000000] [--nq 10000] [--dim 768] [--library-gemm | --no-library]
    """
    import argparse
    import json
    import sys
    import time

    import hnsw_rabitq_amd as H
    from hnsw_rabitq_amd import _lib

    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--dim", type=int, default=768)
    ap.add_argument("--library-gemm", action="store_true", help="only time the vendor GEMM on this shape")
    ap.add_argument("--no-library", action="store_true",
                    help="skip the vendor GEMM line (its child process would share a profiler's output)")
    args = ap.parse_args()
    t0 = time.time()

    def log(m):
        print(f"[bf {time.time() - t0:7.1f}s] {m}", file=sys.stderr, flush=True)

    if args.library_gemm:
        print(json.dumps(library_gemm_run(args.n, args.nq, args.dim, log)), flush=True)
        return

    class Timer:
        def __init__(self, stream):
            self.stream = stream
            self.t = ctypes.c_void_p()
            _lib.call("hg_timer_create", ctypes.byref(self.t))

        def start(self):
            _lib.call("hg_timer_start", self.t, self.stream)

        def stop(self):
            ms = ctypes.c_float()
            _lib.call("hg_timer_stop", self.t, self.stream, ctypes.byref(ms))
            return ms.value

    stream = ctypes.c_void_p()
    _lib.call("hg_stream_create", ctypes.byref(stream))
    out = run(H, _lib, n=args.n, nq=args.nq, stream=stream, timer=Timer(stream), log=log, dim=args.dim,
              library=not args.no_library)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
