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

    def config_extend(self, ext_ef=0, ext_threshold=0.5):
        """Queries the network is not confident about at the end of their search (confidence <
        ext_threshold) are answered again at ext_ef (0 = no extension)."""
        call("hg_stop_model_config_extend", self.h, int(ext_ef), ctypes.c_float(ext_threshold))

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


def collect_traces(gpu, d_queries, nq, ef, k, period, max_records, stream=None, end_record=False):
    """Features [nq][R][12], top-k ids per record [nq][R][k], record counts [nq], final ids [nq][k].
    With end_record the last record of each query is taken where its search ends."""
    from . import DeviceArray
    feat = DeviceArray(nq * max_records * N_FEAT * 4)
    topk = DeviceArray(nq * max_records * k * 4)
    nrec = DeviceArray(nq * 4)
    final = DeviceArray(nq * k * 4)
    call("hg_gpu_collect_traces", gpu.handle, d_queries.ptr, nq, ef, k, period, max_records, feat.ptr, topk.ptr,
         nrec.ptr, final.ptr, 1 if end_record else 0, stream)
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


def extension_records(gpu, d_queries, nq, ef, ext_ef, k, period, max_records=256, stream=None):
    """Training records of the extended learned search (base ef, extension to ext_ef): the traces
    of the search at ef with its end record and of the search at ext_ef, each record labelled
    positive when its top-k already equals the top-k of the search at ext_ef as a set -- the
    answer the extension would give, so a positive record means the query may stop (or skip the
    extension) without changing its result."""
    f_b, t_b, r_b, _ = collect_traces(gpu, d_queries, nq, ef, k, period, max_records, stream, end_record=True)
    f_e, t_e, r_e, fin_e = collect_traces(gpu, d_queries, nq, ext_ef, k, period, max_records, stream)
    eq_b, valid_b = label_records(t_b, r_b, fin_e)
    eq_e, valid_e = label_records(t_e, r_e, fin_e)
    X = np.concatenate([f_b[valid_b], f_e[valid_e]])
    Y = np.concatenate([eq_b[valid_b], eq_e[valid_e]]).astype(np.float32)
    return X, Y


def train_model(X, Y, steps=2000, batch=4096, lr=2e-3, log=None, stream=None, seed=0):
    """A stop model with feature normalisation from X, trained on (X, Y) by the GPU kernels."""
    mean = X.mean(0)
    std = X.std(0) + 1e-6
    model = StopModel()
    model.init_params(seed)
    model.set_norm(mean, 1.0 / std)
    losses = train(model, (X - mean) / std, Y, steps=steps, batch=min(batch, X.shape[0]), lr=lr, log=log,
                   stream=stream)
    return model, mean.astype(np.float32), (1.0 / std).astype(np.float32), losses


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
    against the plain fixed-ef search (same batch, same index).

    The learned search runs at a base ef with the network evaluated every `period` expansions
    (stop when confident the top-k has converged) and once more where the search ends by its own
    rule (extend to ext_ef when not confident).  Training records: traces of the search at the
    base ef (with its end record) and at ext_ef, labelled by whether the record's top-k already
    equals the ext_ef search's.  The measurement also splits the learned search's cost: the stop
    kernel with a network that never fires against the plain kernel (evaluation and feature
    overhead), and the learned run's per-query stopping points replayed by the plain kernel with
    a hop cap (what the decisions alone are worth)."""
    from . import DeviceArray
    nq = queries.shape[0]
    plain_op = next((s for s in sweep if s["recall"] >= 0.95), sweep[-1])
    if plain_qps is None:
        plain_qps = nq / (plain_op["ms"] * 1e-3)
    base = plain_op["ef"]
    ext = min(ef_trace or int(round(base * 1.5 / 64.0)) * 64, ef_max)
    log(f"learned stop: {n_train} training queries, traces at ef={base} (with end records) and ef={ext}, "
        f"every {period} expansions")
    tq_all = datasets.c2_gamma(n=n_train, seed=2)
    Xs, Ys = [], []
    rng = np.random.default_rng(0)
    for s0 in range(0, n_train, chunk):
        tq = tq_all[s0:s0 + chunk]
        m = tq.shape[0]
        dtq = DeviceArray.from_numpy(tq)
        X, Y = extension_records(gpu, dtq, m, base, ext, k, period, max_records, stream)
        keep = min(X.shape[0], train_records * m // n_train)
        sel = rng.choice(X.shape[0], keep, replace=False)
        Xs.append(X[sel])
        Ys.append(Y[sel])
        del dtq
        log(f"  traces {s0 + m}/{n_train}: {X.shape[0]} records, kept {keep}, positive {Y.mean():.3f}")
    X = np.concatenate(Xs)
    Y = np.concatenate(Ys)
    log(f"  {X.shape[0]} training records, positive fraction {Y.mean():.3f}")
    t0 = __import__("time").time()
    model, _, _, losses = train_model(X, Y, steps=2000, batch=4096, lr=2e-3, log=log, stream=stream)
    train_s = __import__("time").time() - t0

    dl = DeviceArray(nq * k * 8)
    dd = DeviceArray(nq * k * 4)
    dc = DeviceArray(nq * 16)

    def recall():
        lab = dl.to_numpy(np.uint64, (nq, k))
        return float(np.mean([len(np.intersect1d(a, b)) for a, b in zip(lab, gt)]) / k)

    def timed(launch):
        launch(dc)
        hops = dc.to_numpy(np.int32, (nq, 4))[:, 0].copy()
        ms = []
        for _ in range(max(1, steps)):
            timer.start()
            launch(None)
            ms.append(timer.stop())
        return float(np.mean(ms)), hops, recall()

    def learned(ef, model_h):
        return lambda c: gpu.search_device(dq, nq, k, ef, dl, dd, d_counters=c, stop_model=model_h, stream=stream)

    gpu.autotune(dq, nq, base, k, stream=stream)
    # cost split at the base ef: (a) the stop kernel with a network that never fires
    never = StopModel()
    p0 = np.zeros(never.nparam, np.float32)
    p0[-1] = -30.0
    never.set_params(p0)
    never.config(period=period, threshold=threshold, min_hops=period)
    ms_plain, _, rec_plain = timed(lambda c: gpu.search_device(dq, nq, k, base, dl, dd, d_counters=c, stream=stream))
    ms_never, _, _ = timed(learned(base, never.h))
    never.free()
    log(f"  cost split at ef={base}: plain {ms_plain:.2f} ms; stop kernel, network never firing {ms_never:.2f} ms "
        f"(evaluation + features every {period} expansions: {ms_never / ms_plain - 1:+.1%})")
    table = []
    best = None
    st, n_ext = ctypes.c_uint32(), ctypes.c_uint32()
    # base ef of the learned search: the plain operating point, the swept beam sizes between it
    # and ext_ef (a stop rule that ends early needs a larger beam to keep recall 0.95), and ext_ef
    efs = [base] + sorted({s["ef"] for s in sweep if base < s["ef"] < ext}) + [ext]
    for ef in efs:
        if ef != base:
            gpu.autotune(dq, nq, ef, k, stream=stream)
        for thr in (0.8, threshold, 0.95):
            for min_hops in (period, ef // 2):
                for te in ((0.0, 0.3) if ef == base else (0.0,)):
                    model.config(period=period, threshold=thr, min_hops=min_hops)
                    model.config_extend(ext if te > 0 else 0, te if te > 0 else 0.5)
                    t, hops, rec = timed(learned(ef, model.h))
                    lib.call("hg_gpu_search_status", gpu.handle, ctypes.byref(st), ctypes.byref(n_ext))
                    ne = int(n_ext.value) if te > 0 else 0
                    row = dict(ef=ef, threshold=thr, min_hops=min_hops, ext_ef=ext if te > 0 else 0,
                               ext_threshold=te, recall=rec, ms=t, qps=nq / (t * 1e-3), hops=float(hops.mean()),
                               extended=ne)
                    if te == 0:
                        # the same stopping points without the network: the plain kernel with a hop cap
                        dh = DeviceArray.from_numpy(hops.astype(np.int32))
                        t2, _, rec2 = timed(lambda c: lib.call("hg_gpu_search_device_hopcap", gpu.handle, dq.ptr, nq,
                                                               dh.ptr, ef, k, dl.ptr, dd.ptr,
                                                               c.ptr if c else None, stream))
                        row.update(replay_ms=t2, replay_recall=rec2)
                        del dh
                    table.append(row)
                    log(f"  learned ef={ef} thr={thr} min_hops={min_hops} ext={row['ext_ef']}/{te}: recall@10={rec:.4f} "
                        f"{t:.2f} ms QPS={nq / (t * 1e-3):.0f} hops={hops.mean():.0f} extended={ne}"
                        + (f"; hop-capped replay {row['replay_ms']:.2f} ms" if te == 0 else ""))
                    # the configured confidence is the reported operating rule
                    if thr == threshold and rec >= 0.95 and (best is None or row["qps"] > best["qps"]):
                        best = row
    # what any stop rule could reach on these queries: each query stopped at the first record from
    # which its top-k no longer changes, on traces of the search at ext_ef
    feats, topk, nrec, final_ids = collect_traces(gpu, dq, nq, ext, k, period, max_records, stream)
    eq, valid = label_records(topk, nrec, final_ids)
    R = eq.shape[1]
    tail_ok = np.flip(np.logical_and.accumulate(np.flip(eq | ~valid, axis=1), axis=1), axis=1)
    first = np.where(tail_ok.any(axis=1), np.argmax(tail_ok, axis=1), R)
    oracle_hops = np.minimum(first * period, nrec * period).astype(np.float64)
    del feats, topk
    gpu.search_device(dq, nq, k, ext, dl, dd, d_counters=dc, stream=stream)
    traced_recall = recall()
    traced_hops = float(dc.to_numpy(np.int32, (nq, 4))[:, 0].mean())
    log(f"  stop-rule bound: traced ef={ext} recall@10={traced_recall:.4f} hops={traced_hops:.0f}; stopping each "
        f"query where its top-{k} stops changing needs {oracle_hops.mean():.0f} hops on average")
    out = dict(learned_table=table, learned_train_loss_first=float(np.mean(losses[:20])),
               learned_train_loss_last=float(np.mean(losses[-50:])), learned_records=int(X.shape[0]),
               learned_positive_fraction=float(Y.mean()), learned_train_s=train_s, learned_base_ef=base,
               learned_ext_ef=ext, learned_plain_ms_at_base=ms_plain, learned_never_firing_ms_at_base=ms_never,
               learned_bound_traced_recall=traced_recall, learned_bound_traced_hops=traced_hops,
               learned_bound_converged_hops=float(oracle_hops.mean()), plain_hops_at_op=float(plain_op["hops"]))
    if best is not None:
        out.update(learned_qps_at_095=best["qps"], learned_ef=best["ef"], learned_recall=best["recall"],
                   learned_over_plain=best["qps"] / plain_qps, learned_best=best)
    model.free()
    return out
