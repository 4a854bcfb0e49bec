"""Learned early termination of the GPU search (the reference's BaseSearchStopCondition slot,
hnswlib.h:134-150, filled by a small network).

Pipeline (all compute on the GPU through the C ABI):
  1. training queries from the benchmark generator with their own seed;
  2. traces: the plain search at the traced ef, recording every `period` level-0 expansions the
     12 traversal features and the current top-k, plus the search's final top-k;
  3. labels: a record is positive when its top-k already equals the traced search's final top-k
     as a set (the search has converged to what it will return);
  4. training: the project's HIP forward/backward/Adam kernels (hg_stop_model_train_step);
  5. inference: the search kernel evaluates the network every `period` expansions and stops a
     query when sigmoid(logit) >= threshold; otherwise the query runs to the plain fixed-ef end.
"""
import ctypes

import numpy as np

from . import _lib, datasets
from ._lib import call

N_FEAT = 12
HIDDEN = 64


class StopModel:
    def __init__(self, device=0, hidden=HIDDEN):
        self.h = ctypes.c_void_p()
        call("hg_stop_model_create", N_FEAT, hidden, device, ctypes.byref(self.h))
        c = ctypes.c_size_t()
        call("hg_stop_model_param_count", self.h, ctypes.byref(c))
        self.nparam = c.value
        self.hidden = hidden

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        call("hg_stop_model_set_params", self.h, ctypes.c_void_p(p.ctypes.data), p.shape[0])

    def get_params(self):
        p = np.empty(self.nparam, np.float32)
        call("hg_stop_model_get_params", self.h, ctypes.c_void_p(p.ctypes.data), self.nparam)
        return p

    def set_norm(self, mean, inv_std):
        mean = np.ascontiguousarray(mean, np.float32)
        inv_std = np.ascontiguousarray(inv_std, np.float32)
        call("hg_stop_model_set_norm", self.h, ctypes.c_void_p(mean.ctypes.data), ctypes.c_void_p(inv_std.ctypes.data),
             N_FEAT)

    def config(self, period=16, threshold=0.9, min_hops=16):
        call("hg_stop_model_config", self.h, int(period), ctypes.c_float(threshold), int(min_hops))

    def init_params(self, seed=0):
        rng = np.random.default_rng(seed)
        H, I = self.hidden, N_FEAT
        parts = [rng.standard_normal(H * I) * np.sqrt(2.0 / I), np.zeros(H),
                 rng.standard_normal(H * H) * np.sqrt(2.0 / H), np.zeros(H),
                 rng.standard_normal(H) * np.sqrt(1.0 / H), np.zeros(1)]
        self.set_params(np.concatenate(parts).astype(np.float32))

    def train_step(self, dx, dy, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, dgrad=None, stream=None):
        loss = ctypes.c_float()
        call("hg_stop_model_train_step", self.h, dx.ptr, dy.ptr, n, ctypes.c_float(lr), ctypes.c_float(b1),
             ctypes.c_float(b2), ctypes.c_float(eps), ctypes.byref(loss), dgrad.ptr if dgrad else None, stream)
        return loss.value

    def forward(self, dx, n, dlogits, stream=None):
        call("hg_stop_model_forward", self.h, dx.ptr, n, dlogits.ptr, stream)

    def free(self):
        if self.h is not None and self.h.value:
            _lib.load().hg_stop_model_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def collect_traces(gpu, d_queries, nq, ef, k, period, max_records, stream=None):
    """Features [nq][R][12], top-k ids per record [nq][R][k], record counts [nq], final ids [nq][k]."""
    from . import DeviceArray
    feat = DeviceArray(nq * max_records * N_FEAT * 4)
    topk = DeviceArray(nq * max_records * k * 4)
    nrec = DeviceArray(nq * 4)
    final = DeviceArray(nq * k * 4)
    call("hg_gpu_collect_traces", gpu.handle, d_queries.ptr, nq, ef, k, period, max_records, feat.ptr, topk.ptr,
         nrec.ptr, final.ptr, stream)
    f = feat.to_numpy(np.float32, (nq, max_records, N_FEAT))
    t = topk.to_numpy(np.uint32, (nq, max_records, k))
    r = nrec.to_numpy(np.int32, (nq,))
    fin = final.to_numpy(np.uint32, (nq, k))
    return f, t, r, fin


def label_records(topk, nrec, final_ids):
    """Positive when the record's top-k equals the final top-k as a set (the search has already
    converged to what it will return)."""
    nq, R, k = topk.shape
    a = np.sort(topk.astype(np.int64), axis=2)
    b = np.sort(final_ids.astype(np.int64), axis=1)[:, None, :]
    eq = (a == b).all(axis=2)
    valid = np.arange(R)[None, :] < nrec[:, None]
    return eq, valid


def train(model, X, Y, steps, batch, lr, log=None, seed=0, stream=None):
    from . import DeviceArray
    rng = np.random.default_rng(seed)
    dX = DeviceArray.from_numpy(X.astype(np.float32))
    n = X.shape[0]
    losses = []
    dx = DeviceArray(batch * N_FEAT * 4)
    dy = DeviceArray(batch * 4)
    for s in range(steps):
        idx = rng.integers(0, n, batch)
        xb = np.ascontiguousarray(X[idx], np.float32)
        yb = np.ascontiguousarray(Y[idx], np.float32)
        call("hg_memcpy_h2d", dx.ptr, ctypes.c_void_p(xb.ctypes.data), xb.nbytes, stream)
        call("hg_memcpy_h2d", dy.ptr, ctypes.c_void_p(yb.ctypes.data), yb.nbytes, stream)
        loss = model.train_step(dx, dy, batch, lr=lr, stream=stream)
        losses.append(loss)
        if log and (s + 1) % max(1, steps // 5) == 0:
            log(f"    step {s + 1}/{steps} loss {np.mean(losses[-50:]):.4f}")
    del dX
    return losses


def run(gpu, H, lib, data_dim, queries, dq, gt, stream, timer, sweep, log, steps=5, k=10, n_train=200_000,
        period=16, max_records=256, threshold=0.9, chunk=25_000, train_records=4_000_000, plain_qps=None,
        ef_max=4096, ef_trace=None):
    """Train the stop model on traces of the plain GPU search and measure QPS at recall >= 0.95
    against the plain fixed-ef search (same batch, same index)."""
    from . import DeviceArray
    nq = queries.shape[0]
    plain_op = next((s for s in sweep if s["recall"] >= 0.95), sweep[-1])
    if plain_qps is None:
        plain_qps = nq / (plain_op["ms"] * 1e-3)
    # the learned search runs at a larger ef than the plain operating point and stops queries
    # once the network is confident their top-k has converged
    ef_l = min(ef_trace or 2 * plain_op["ef"], ef_max)
    log(f"learned stop: {n_train} training queries, traces at ef={ef_l}, every {period} expansions")
    tq_all = datasets.c2_gamma(n=n_train, seed=2)
    Xs, Ys = [], []
    rng = np.random.default_rng(0)
    for s0 in range(0, n_train, chunk):
        tq = tq_all[s0:s0 + chunk]
        m = tq.shape[0]
        dtq = DeviceArray.from_numpy(tq)
        feats, topk, nrec, final_ids = collect_traces(gpu, dtq, m, ef_l, k, period, max_records, stream)
        eq, valid = label_records(topk, nrec, final_ids)
        X, Y = feats[valid], eq[valid].astype(np.float32)
        keep = min(X.shape[0], train_records * m // n_train)
        sel = rng.choice(X.shape[0], keep, replace=False)
        Xs.append(X[sel])
        Ys.append(Y[sel])
        del dtq, feats, topk
        log(f"  traces {s0 + m}/{n_train}: {int(valid.sum())} records, kept {keep}, positive {Y.mean():.3f}")
    X = np.concatenate(Xs)
    Y = np.concatenate(Ys)
    log(f"  {X.shape[0]} training records, positive fraction {Y.mean():.3f}")
    mean = X.mean(0)
    std = X.std(0) + 1e-6
    model = StopModel()
    model.init_params(0)
    model.set_norm(mean, 1.0 / std)
    Xn = (X - mean) / std
    t0 = __import__("time").time()
    losses = train(model, Xn, Y, steps=2000, batch=4096, lr=2e-3, log=log, stream=stream)
    train_s = __import__("time").time() - t0

    # evaluation on the benchmark queries: the learned search stops a query at the first check
    # where confidence >= threshold, otherwise it runs to the plain fixed-ef end
    dl = DeviceArray(nq * k * 8)
    dd = DeviceArray(nq * k * 4)
    dc = DeviceArray(nq * 16)
    best = None
    table = []
    # candidate ef of the learned search: from just above the plain operating point up to the
    # traced ef (a larger beam costs LDS, hence concurrency)
    base = plain_op["ef"]
    efs = sorted(set(min(ef_max, int(round(base * f / 64.0)) * 64) for f in (1.2, 1.6)) | {ef_l, base})
    for ef in efs:
        gpu.autotune(dq, nq, ef, k, stream=stream)
        for thr in (threshold, 0.5, 0.7, 0.8, 0.95):
            model.config(period=period, threshold=thr, min_hops=period)
            gpu.search_device(dq, nq, k, ef, dl, dd, d_counters=dc, stop_model=model.h, stream=stream)
            hops = float(dc.to_numpy(np.int32, (nq, 4))[:, 0].mean())
            ms = []
            for _ in range(max(1, steps)):
                timer.start()
                gpu.search_device(dq, nq, k, ef, dl, dd, stop_model=model.h, stream=stream)
                ms.append(timer.stop())
            lab = dl.to_numpy(np.uint64, (nq, k))
            rec = float(np.mean([len(np.intersect1d(a, b)) for a, b in zip(lab, gt)]) / k)
            t = float(np.mean(ms))
            row = dict(ef=ef, threshold=thr, recall=rec, ms=t, qps=nq / (t * 1e-3), hops=hops)
            table.append(row)
            log(f"  learned ef={ef} thr={thr}: recall@10={rec:.4f} {t:.2f} ms QPS={nq / (t * 1e-3):.0f} "
                f"hops={hops:.0f}")
            # the configured confidence (0.9) is the reported operating rule
            if thr == threshold and rec >= 0.95 and (best is None or row["qps"] > best["qps"]):
                best = row
    # bound for any stop rule on these traces: each benchmark query stopped at the first record
    # from which its top-k no longer changes (recall = that of the traced search)
    feats, topk, nrec, final_ids = collect_traces(gpu, dq, nq, ef_l, k, period, max_records, stream)
    eq, valid = label_records(topk, nrec, final_ids)
    R = eq.shape[1]
    # first record index after which every valid record is positive
    tail_ok = np.flip(np.logical_and.accumulate(np.flip(eq | ~valid, axis=1), axis=1), axis=1)
    first = np.where(tail_ok.any(axis=1), np.argmax(tail_ok, axis=1), R)
    oracle_hops = np.minimum(first * period, nrec * period).astype(np.float64)
    del feats, topk
    gpu.search_device(dq, nq, k, ef_l, dl, dd, d_counters=dc, stream=stream)
    lab = dl.to_numpy(np.uint64, (nq, k))
    traced_recall = float(np.mean([len(np.intersect1d(a, b)) for a, b in zip(lab, gt)]) / k)
    traced_hops = float(dc.to_numpy(np.int32, (nq, 4))[:, 0].mean())
    log(f"  stop-rule bound: traced ef={ef_l} recall@10={traced_recall:.4f} hops={traced_hops:.0f}; stopping each "
        f"query where its top-{k} stops changing needs {oracle_hops.mean():.0f} hops on average")
    out = dict(learned_table=table, learned_train_loss_first=float(np.mean(losses[:20])),
               learned_train_loss_last=float(np.mean(losses[-50:])), learned_records=int(X.shape[0]),
               learned_positive_fraction=float(Y.mean()), learned_train_s=train_s, learned_trace_ef=ef_l,
               learned_bound_traced_recall=traced_recall, learned_bound_traced_hops=traced_hops,
               learned_bound_converged_hops=float(oracle_hops.mean()),
               plain_hops_at_op=float(plain_op["hops"]))
    if best is not None:
        out.update(learned_qps_at_095=best["qps"], learned_ef=best["ef"], learned_recall=best["recall"],
                   learned_over_plain=best["qps"] / plain_qps)
    model.free()
    return out
