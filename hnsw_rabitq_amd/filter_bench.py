"""Non-bare-bone search measurement on the 1M x 128 L2 index: a label filter that keeps 10% of
the elements (label % 10 == 0), and 10% of the elements marked deleted (label % 10 == 1).
Ground truth is the exact FP32 top-10 over the returnable rows (GPU brute force over that subset).
Reports recall@10, kernel time and queries/s per ef; the operating point is the smallest ef with
recall@10 >= 0.95.  Then the range search (EpsilonSearchStopCondition) at three min_candidates and
the multi-vector search (MultiVectorSearchStopCondition) at three ef_collection values."""
import numpy as np

FILTER_EFS = (10, 20, 40, 80, 160)
DELETE_EFS = (400, 800, 1600)


def _recall(labels, gt, k):
    hits = 0
    for a, b in zip(labels[:, :k], gt[:, :k]):
        hits += len(np.intersect1d(a, b))
    return hits / (labels.shape[0] * k)


def _exact(H, lib, data, rows, dq, nq, k, stream):
    from . import DeviceArray
    bf = H.BFIndex("l2", data.shape[1])
    bf.init_index(rows.shape[0])
    bf.add_items(data[rows], rows.astype(np.uint64))
    gl, gd = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4)
    lib.call("hg_bf_search_device", bf._h, dq.ptr, nq, k, lib.PREC_FP32, 0, gl.ptr, gd.ptr, None, stream)
    return gl.to_numpy(np.uint64, (nq, k))


def run(H, lib, idx, gpu, data, dq, nq, stream, timer, log, k=10, steps=3):
    from . import DeviceArray
    n = data.shape[0]
    dl, dd, dst = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4), DeviceArray(4)
    dc = DeviceArray(nq * 16)
    labels = np.arange(n, dtype=np.uint64)  # the bench index uses label = row

    def sweep(name, efs, gt, flt):
        pts = []
        for ef in efs:
            gpu.search_device_filtered(dq, nq, k, ef, dl, dd, filter=flt, d_status=dst, stream=stream)
            ms = []
            for _ in range(steps):
                timer.start()
                gpu.search_device_filtered(dq, nq, k, ef, dl, dd, filter=flt, d_status=dst, stream=stream)
                ms.append(timer.stop())
            lab = dl.to_numpy(np.uint64, (nq, k))
            status = int(dst.to_numpy(np.uint32, (1,))[0])
            second = gpu.status()[1]
            gpu.search_device_filtered(dq, nq, k, ef, dl, dd, filter=flt, d_status=dst, d_counters=dc,
                                       stream=stream)
            ctr = dc.to_numpy(np.int32, (nq, 4))
            rec = _recall(lab, gt, k)
            t = float(np.mean(ms))
            log(f"  {name} ef={ef:4d} recall@10={rec:.4f} {t:8.2f} ms QPS={nq / (t * 1e-3):10.0f} "
                f"hops={ctr[:, 0].mean():.0f} evals={ctr[:, 1].mean():.0f} second-pass queries={second} "
                f"status={status}")
            pts.append(dict(ef=ef, recall=rec, ms=t, qps=nq / (t * 1e-3), second_pass=second,
                            hops=float(ctr[:, 0].mean()), evals=float(ctr[:, 1].mean()),
                            overflow=bool(status & 1)))
        op = next((p for p in pts if p["recall"] >= 0.95 and not p["overflow"]), None)
        return pts, op

    out = {}
    # label filter: 10% of the labels
    keep = labels[labels % 10 == 0]
    log(f"filtered search: filter keeps {keep.shape[0]} of {n} labels")
    flt = gpu.make_filter(keep)
    gt = _exact(H, lib, data, keep.astype(np.int64), dq, nq, k, stream)
    pts, op = sweep("filter10%", FILTER_EFS, gt, flt)
    flt.free()
    out["filter10_sweep"] = pts
    out["filter10_qps_at_095"] = op["qps"] if op else None
    out["filter10_ef"] = op["ef"] if op else None

    # deletions: 10% of the elements marked deleted, then unmarked again
    dele = labels[labels % 10 == 1]
    log(f"deleted elements: marking {dele.shape[0]}")
    for lab in dele:
        lib.call("hg_hnsw_mark_deleted", idx.cpu_handle(), int(lab))
    gpu.refresh_deleted(idx.cpu_handle())
    alive = np.nonzero(labels % 10 != 1)[0]
    gt = _exact(H, lib, data, alive, dq, nq, k, stream)
    pts, op = sweep("deleted10%", DELETE_EFS, gt, None)
    for lab in dele:
        lib.call("hg_hnsw_unmark_deleted", idx.cpu_handle(), int(lab))
    gpu.refresh_deleted(idx.cpu_handle())
    out["deleted10_sweep"] = pts
    out["deleted10_qps_at_095"] = op["qps"] if op else None
    out["deleted10_ef"] = op["ef"] if op else None
    out.update(range_search(H, lib, gpu, data, dq, nq, stream, timer, log))
    out.update(multivector(lib, gpu, dq, nq, stream, timer, log))
    return out


def multivector(lib, gpu, dq, nq, stream, timer, log, steps=3):
    """Multi-vector search with documents of 10 consecutive labels (num_docs 10)."""
    from . import DeviceArray
    labs = gpu.labels()
    gd = gpu.make_docs(labs, (labs // 10).astype(np.uint32))
    stride = 128
    dl, dd = DeviceArray(nq * stride * 8), DeviceArray(nq * stride * 4)
    dcnt, dst = DeviceArray(nq * 4), DeviceArray(4)
    out = {"multivector_sweep": []}
    for efc in (10, 40, 160):
        def launch():
            lib.call("hg_gpu_search_multivector_device", gpu.handle, dq.ptr, nq, gd.handle, 10, efc, None, stride,
                     dl.ptr, dd.ptr, dcnt.ptr, None, dst.ptr, 0, stream)
        launch()
        ms = []
        for _ in range(steps):
            timer.start()
            launch()
            ms.append(timer.stop())
        t = float(np.mean(ms))
        cnt = dcnt.to_numpy(np.uint32, (nq,))
        status = int(dst.to_numpy(np.uint32, (1,))[0])
        log(f"  multivector num_docs=10 ef_collection={efc}: {t:8.2f} ms QPS={nq / (t * 1e-3):10.0f} "
            f"results/query {cnt.mean():.1f} status={status}")
        out["multivector_sweep"].append(dict(ef_collection=efc, ms=t, qps=nq / (t * 1e-3),
                                             results_per_query=float(cnt.mean())))
    gd.free()
    return out


def range_search(H, lib, gpu, data, dq, nq, stream, timer, log, kmax=256, steps=3):
    """Range search (EpsilonSearchStopCondition) with the radius at the median 10th-neighbour
    distance: recall of the elements within the radius (exact: the FP32 brute-force top-256,
    which holds every in-range element while fewer than 256 are in range)."""
    from . import DeviceArray
    n = data.shape[0]
    bf = H.BFIndex("l2", data.shape[1])
    bf.init_index(n)
    bf.add_items(data, np.arange(n, dtype=np.uint64))
    gl, gd = DeviceArray(nq * kmax * 8), DeviceArray(nq * kmax * 4)
    lib.call("hg_bf_search_device", bf._h, dq.ptr, nq, kmax, lib.PREC_FP32, 0, gl.ptr, gd.ptr, None, stream)
    del bf
    gt_l = gl.to_numpy(np.uint64, (nq, kmax))
    gt_d = gd.to_numpy(np.float32, (nq, kmax))
    eps = float(np.median(gt_d[:, 9]))
    in_range = gt_d <= eps
    exact_ok = in_range[:, -1] == 0  # fewer than kmax in range: the exact set is known
    stride = kmax
    dl, dd = DeviceArray(nq * stride * 8), DeviceArray(nq * stride * 4)
    dcnt, dst = DeviceArray(nq * 4), DeviceArray(4)
    out = {"range_eps": eps, "range_sweep": []}
    for min_c in (100, 200, 400):
        def launch():
            lib.call("hg_gpu_search_epsilon_device", gpu.handle, dq.ptr, nq, eps, min_c, 4 * min_c, None, stride,
                     dl.ptr, dd.ptr, dcnt.ptr, None, dst.ptr, stream)
        launch()
        ms = []
        for _ in range(steps):
            timer.start()
            launch()
            ms.append(timer.stop())
        cnt = dcnt.to_numpy(np.uint32, (nq,))
        lab = dl.to_numpy(np.uint64, (nq, stride))
        hit = tot = 0
        for i in np.nonzero(exact_ok)[0]:
            truth = gt_l[i][in_range[i]]
            hit += len(np.intersect1d(lab[i, :min(int(cnt[i]), stride)], truth))
            tot += len(truth)
        rec = hit / max(tot, 1)
        t = float(np.mean(ms))
        status = int(dst.to_numpy(np.uint32, (1,))[0])
        log(f"  range eps={eps:.0f} min={min_c:5d} max={4 * min_c}: recall {rec:.4f} "
            f"({tot} in range over {int(exact_ok.sum())} queries) {t:8.2f} ms QPS={nq / (t * 1e-3):10.0f} "
            f"results/query {cnt.mean():.1f} status={status}")
        out["range_sweep"].append(dict(min_candidates=min_c, max_candidates=4 * min_c, recall=rec, ms=t,
                                       qps=nq / (t * 1e-3), results_per_query=float(cnt.mean())))
    return out
