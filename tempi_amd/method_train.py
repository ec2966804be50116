"""Train the learned staging-method model on this node.

    python -m tempi_amd.method_train --out tempi_amd/method_mlp.txt \
        --report tempi_amd/method_mlp.report.json [--train 3000] [--heldout 1000]

Under `mpiexec -n 2` the two ranks also measure the CPU-CPU ping-pong curve
first, and the report then compares the model with the reference's own
choice priced on fully measured tables.

The reference picks how a non-contiguous message travels -- packed on the
device and handed over as a device pointer (DEVICE), packed straight into
pinned host memory (ONE_SHOT), or packed on the device and copied to the host
(STAGED) -- by summing measured table entries for pack, transfer and unpack
(src/internal/sender.cpp:128-273).  This trains the selector's network on
what the choice is about instead: the end-to-end time of each method.

1. Generate seeded messages of device-resident derived types: byte vectors
   with rows of 1 B - 4 KiB at gaps of 1 B - 4 KiB, 256 B - 16 MiB packed, and
   faces of float grids.
2. Time each method the library can run here (ONE_SHOT and STAGED; DEVICE
   only with TEMPI_MPI_GPU_AWARE) on each message: MPI_Irecv + MPI_Isend +
   MPI_Waitall of this process to itself through the interposed calls, the
   median of several exchanges after a warm-up -- the pack, the staging
   copies, the MPI library's transfer and the unpack.
3. Train the network (tempi_mlp_train: the library's HIP forward / backward /
   Adam kernels) with output k predicting method k + 1's log2 time relative
   to the faster method, outputs of methods not timed masked out.
4. On held-out messages, judged on an independent second timing, report the
   regret of the model's choice, of always ONE_SHOT (the choice without
   tables), of always STAGED, and -- with the ping-pong curve -- of the
   table sum; write the weights file the library loads.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import random
import sys
import time

import numpy as np

import tempi_amd as ta
from tempi_amd import mpi as tmpi
from tempi_amd import system

SPAN = 256 << 20  # largest strided span of a message
DEVICE, ONESHOT, STAGED = 1, 2, 3
BYTE = ("named", "MPI_BYTE")
FLOAT = ("named", "MPI_FLOAT")


def _loguniform(rng, lo, hi):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def generate(n, seed):
    """Seeded messages: (type spec, packed bytes, row bytes, stride)."""
    rng = random.Random(seed)
    out = []
    while len(out) < n:
        if rng.random() < 0.1:
            # a face of a float grid: rows of 12 B (x) or 4 KiB-ish (y, z)
            m = rng.choice([32, 64, 128, 256])
            g = 3
            face = rng.choice("xyz")
            full = m + 2 * g
            sub = {"x": [m, m, g], "y": [m, g, m], "z": [g, m, m]}[face]
            spec = ("subarray", [full] * 3, sub, [g] * 3, "C", FLOAT)
            block = 4 * (g if face == "x" else m)
            stride = 4 * full
            packed = 4 * m * m * g
        else:
            packed = _loguniform(rng, 256, 16 << 20)
            block = min(_loguniform(rng, 1, 4096), packed)
            rows = max(1, packed // block)
            packed = rows * block
            gap_max = max(1, min(4096, SPAN // rows - block))
            gap = _loguniform(rng, 1, gap_max) if gap_max > 1 else 1
            stride = block + gap
            if rows < 2:
                continue
            spec = ("vector", rows, block, stride, BYTE)
        out.append((spec, packed, block, stride))
    return out


def time_exchange(m, h, src, dst, method, reps):
    """Median wall time (s) of a self-exchange of one object of type h."""
    ta.lib().tempi_set_datatype_method(method)
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        err, rr = m.irecv(dst.addr(), 1, h, 0, 17)
        assert err == 0, err
        err, sr = m.isend(src.addr(), 1, h, 0, 17)
        assert err == 0, err
        err, _, _ = m.waitall([rr, sr], [h, h])
        assert err == 0, err
        if r:
            times.append(time.perf_counter() - t0)
    return float(np.median(times))


def time_cases(m, cases, methods, src, dst, reps, tag):
    L = ta.lib()
    n = len(cases)
    X = np.zeros((n, L.tempi_mlp_num_features()), np.float32)
    T = np.full((n, len(methods)), np.nan)
    B = np.zeros(n)
    t0 = time.time()
    for i, (spec, packed, block, stride) in enumerate(cases):
        h = m.type_from_spec(spec)
        f = (ctypes.c_float * X.shape[1])()
        ta.check(L.tempi_method_features(packed, block, stride, 1, f), "method_features")
        X[i] = np.frombuffer(f, np.float32)
        for j, meth in enumerate(methods):
            T[i, j] = time_exchange(m, h, src, dst, meth, reps)
        B[i] = packed
        m.free(h)
        if (i + 1) % 500 == 0:
            print(f"[method_train] {tag} timed {i + 1}/{n} in {time.time() - t0:.0f} s",
                  flush=True)
    ta.lib().tempi_set_datatype_method(-1)
    return X, T, B


def regret(T, pick):
    idx = np.arange(len(T))
    best = T.min(1)
    r = T[idx, pick] / best - 1.0
    return {"mean": float(r.mean()), "median": float(np.median(r)),
            "p90": float(np.quantile(r, 0.9)), "within_5pct": float((r <= 0.05).mean()),
            "time_weighted": float(T[idx, pick].sum() / best.sum() - 1.0)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=3000)
    ap.add_argument("--heldout", type=int, default=1000)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=7100)
    ap.add_argument("--heldout-seed", type=int, default=910000)
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", required=True)
    args = ap.parse_args(argv)

    L = ta.lib()
    assert ta.device_count() > 0, "no GPU"
    m = tmpi.get()
    # Under mpiexec with two or more ranks, the ranks first measure the
    # CPU-CPU ping-pong curve together, so that rank 0 can price the
    # reference's table sum on measured tables too; the other ranks then
    # leave.
    rank, size = ctypes.c_int(), ctypes.c_int()
    m.tempi.MPI_Comm_rank.argtypes = [m.H, ctypes.POINTER(ctypes.c_int)]
    m.tempi.MPI_Comm_size.argtypes = [m.H, ctypes.POINTER(ctypes.c_int)]
    m.tempi.MPI_Comm_rank(m.COMM_WORLD, ctypes.byref(rank))
    m.tempi.MPI_Comm_size(m.COMM_WORLD, ctypes.byref(size))
    have_pingpong = False
    if size.value >= 2:
        system.clear()
        system.measure_performance(system.MEASURE_PINGPONG, 0.01)
        have_pingpong = True
        if rank.value != 0:
            m.finalize()
            return 0
    # the library hands device pointers to the MPI library only when told it
    # reads device memory
    gpu_aware = os.environ.get("TEMPI_MPI_GPU_AWARE") is not None
    methods = [ONESHOT, STAGED] + ([DEVICE] if gpu_aware else [])
    src = ta.DeviceBuffer(SPAN + 4096)
    src.fill(1)
    dst = ta.DeviceBuffer(SPAN + 4096)
    dst.fill(2)
    t0 = time.time()
    Xtr, Ttr, Btr = time_cases(m, generate(args.train, args.seed), methods, src, dst,
                               args.reps, "train")
    hcases = generate(args.heldout, args.heldout_seed)
    Xho, Tho, Bho = time_cases(m, hcases, methods, src, dst, args.reps, "heldout")
    _, Tho2, _ = time_cases(m, hcases, methods, src, dst, args.reps, "heldout-retime")
    timing_s = time.time() - t0
    src.free()
    dst.free()

    # the reference's choice: the cheapest sum of measured table entries
    # (only with the ping-pong curve, which needs a second rank)
    table_pick = None
    if have_pingpong:
        system.measure_performance(system.MEASURE_ALL, 0.01)
        L.tempi_method_model_unload()
        ta.lib().tempi_set_datatype_method(-1)
        col = {meth: j for j, meth in enumerate(methods)}
        table_pick = np.array([col.get(L.tempi_p2p_method(1, packed, block, stride), 0)
                               for _, packed, block, stride in hcases])
        L.tempi_method_model_load(None)

    nf, nout = L.tempi_mlp_num_features(), L.tempi_mlp_num_outputs()
    mean = Xtr.mean(0)
    inv = 1.0 / np.maximum(Xtr.std(0), 0.5)
    norm = np.concatenate([mean, inv]).astype(np.float32)
    Xn = ((Xtr - mean) * inv).astype(np.float32)
    # output k = method k + 1; log2 time relative to the faster method
    lt = np.log2(Ttr)
    rel = lt - lt.min(1, keepdims=True)
    target = np.zeros((len(Xn), nout), np.float32)
    mask = np.zeros((len(Xn), nout), np.float32)
    for j, meth in enumerate(methods):
        target[:, meth - 1] = rel[:, j]
        mask[:, meth - 1] = 1.0

    rng = np.random.default_rng(args.seed)
    parts = []
    for s in [(nf, 64), (64,), (64, 64), (64,), (64, nout), (nout,)]:
        parts.append(rng.standard_normal(s) * math.sqrt(2.0 / s[0]) if len(s) == 2
                     else np.zeros(s))
    params = np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)
    assert params.size == L.tempi_mlp_num_params()
    losses = []
    t0 = time.time()
    for epochs, lr in ((args.epochs * 2 // 3, args.lr),
                       (args.epochs - args.epochs * 2 // 3, args.lr / 4)):
        el = np.zeros(epochs, np.float32)
        ta.check(L.tempi_mlp_train(Xn.ctypes.data, target.ctypes.data, mask.ctypes.data,
                                   len(Xn), epochs, ctypes.c_float(lr), args.seed,
                                   params.ctypes.data, el.ctypes.data), "tempi_mlp_train")
        losses += el.tolist()
        print(f"[method_train] {epochs} epochs at lr {lr}: loss {el[0]:.4f} -> {el[-1]:.4f}",
              flush=True)
    train_s = time.time() - t0

    def pick(X):
        pred = np.zeros((len(X), nout), np.float32)
        best = np.zeros(len(X), np.int32)
        ta.check(L.tempi_mlp_predict(params.ctypes.data, norm.ctypes.data,
                                     np.ascontiguousarray(X, np.float32).ctypes.data, len(X),
                                     pred.ctypes.data, best.ctypes.data), "tempi_mlp_predict")
        cols = [meth - 1 for meth in methods]
        return pred[:, cols].argmin(1)  # index into `methods`

    ta.check(L.tempi_mlp_save(args.out.encode(), params.ctypes.data, norm.ctypes.data), "save")
    report = {"methods": [{ONESHOT: "ONESHOT", STAGED: "STAGED", DEVICE: "DEVICE"}[x]
                          for x in methods],
              "train_messages": len(Xtr), "heldout_messages": len(Xho), "reps": args.reps,
              "timing_seconds": round(timing_s, 1), "train_seconds": round(train_s, 2),
              "loss_first_last": [losses[0], losses[-1]]}
    pm_tr, pm_ho = pick(Xtr), pick(Xho)
    fixed = {"always_" + report["methods"][j].lower(): np.full(len(Xho), j)
             for j in range(len(methods))}
    report["train"] = {"model": regret(Ttr, pm_tr)}
    for name, T in (("heldout", Tho), ("heldout_retimed", Tho2)):
        r = {"model": regret(T, pm_ho)}
        for k, v in fixed.items():
            r[k] = regret(T, v)
        if table_pick is not None:
            r["measured_tables"] = regret(T, table_pick)
        if name == "heldout_retimed":
            # the faster method of the first timing, judged on the second
            r["measured_best_first_timing"] = regret(T, Tho.argmin(1))
            idx = np.arange(len(T))
            r["seconds_total"] = {"model": float(T[idx, pm_ho].sum()),
                                  "best": float(T.min(1).sum()),
                                  **{k: float(T[idx, v].sum()) for k, v in fixed.items()}}
            if table_pick is not None:
                r["seconds_total"]["measured_tables"] = float(T[idx, table_pick].sum())
        report[name] = r
    # where the methods trade places: per size band of the held-out set,
    # median time of each method and how often each is the faster
    bands = [(0, 64 << 10), (64 << 10, 1 << 20), (1 << 20, 32 << 20)]
    report["heldout_by_size"] = []
    for lo, hi in bands:
        sel = (Bho >= lo) & (Bho < hi)
        if not sel.any():
            continue
        report["heldout_by_size"].append({
            "packed_bytes": [lo, hi], "messages": int(sel.sum()),
            "median_us": {report["methods"][j]: float(np.median(Tho2[sel, j]) * 1e6)
                          for j in range(len(methods))},
            "faster_share": {report["methods"][j]: float((Tho2[sel].argmin(1) == j).mean())
                             for j in range(len(methods))},
            "model_share": {report["methods"][j]: float((pm_ho[sel] == j).mean())
                            for j in range(len(methods))}})
    report["heldout_share_by_method"] = {
        report["methods"][j]: float((Tho.argmin(1) == j).mean()) for j in range(len(methods))}
    report["model_share_by_method"] = {
        report["methods"][j]: float((pm_ho == j).mean()) for j in range(len(methods))}
    with open(args.report, "w") as fh:
        json.dump(report, fh, indent=1)
    print(json.dumps(report["heldout_retimed"]), flush=True)
    m.finalize()
    return 0


if __name__ == "__main__":
    sys.exit(main())
