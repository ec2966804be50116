"""Benchmark of the derived-datatype pack/unpack path on one MI355X.

Workload (one "step"): the three 3-cell-thick faces of a 518^3 float32 grid
(a 512^3 interior with a 3-cell ghost ring; normal(0,1) values, seed 2),
each described by MPI_Type_create_subarray and committed through the
interposed MPI_Type_commit; a step packs each face and unpacks it back into
a second grid (6 kernels).  The grid stays resident in HBM.  The packs and
unpacks are enqueued with tempi_pack_sb / tempi_unpack_sb on the library
stream -- the kernels MPI_Pack / MPI_Unpack run, without the per-call stream
synchronisation -- and the timed region is K steps between two device
synchronisations.

value   = packed bytes moved per second (pack + unpack, 6 x 3 MiB per step)
roofline: the dominant kernel's algorithmic HBM bytes per launch
          (2 x packed bytes: read once, written once) over its average
          duration, timed with HIP events on the library stream, against
          the 8 TB/s HBM3E peak.
cpu_baseline: the oracle's single-threaded C gather (oracle/oracle.c) on the
          host copy of the same grid, over a bounded sample; the underlying
          MPI library's own MPI_Pack (where the reference sends host buffers,
          src/pack.cpp:51-54) is reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="run warmup + timed steps only (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU's share of the pack path; --gpus {args.gpus} is not "
              "supported (the pack path does not communicate; see DESIGN.md on sharding).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import tempi_amd as ta
    import typecatalog as tc
    from tempi_amd import mpi as tmpi

    assert ta.device_count() > 0, "no GPU"
    m = tmpi.get()
    n, g = 512, 3
    full = n + 2 * g
    log("generating the 518^3 float32 grid (normal(0,1), seed 2)")
    host = np.random.default_rng(2).standard_normal(full ** 3, dtype=np.float32).view(np.uint8)
    grid = ta.DeviceBuffer(host.size)
    grid.upload(host)
    grid2 = ta.DeviceBuffer(host.size)
    grid2.fill(0)

    faces = []
    for face in "xyz":
        h = m.type_from_spec(tc.grid_face_subarray(n, g, face))
        sb, size = m.strided_block(h)
        faces.append({"face": face, "handle": h, "sb": sb, "size": size,
                      "packed": ta.DeviceBuffer(size)})
    step_bytes = sum(2 * f["size"] for f in faces)
    stream = ta.stream()

    def step():
        for f in faces:
            ta.pack_sb(f["sb"], grid.addr(), 1, f["packed"].addr(), f["size"], 0, -1, stream)
        for f in faces:
            ta.unpack_sb(f["sb"], f["packed"].addr(), f["size"], grid2.addr(), 1, 0, -1, stream)

    for _ in range(args.warmup):
        step()
    ta.synchronize(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ta.synchronize(stream)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = step_bytes * args.steps / (t1 - t0) / 1e9
    log(f"{ms_per_step * 1e3:.2f} us/step, {value:.1f} GB/s packed")
    if args.profile_only:
        print(json.dumps({"profile_only": True, "ms_per_step": ms_per_step}))
        return

    # per-kernel durations: HIP events around every launch on the library
    # stream, over as many steps again
    ops = [("pack", f) for f in faces] + [("unpack", f) for f in faces]
    evs = [(ta.Event(), ta.Event()) for _ in ops]
    durs = [[] for _ in ops]
    for _ in range(args.steps):
        for i, (kind, f) in enumerate(ops):
            evs[i][0].record(stream)
            if kind == "pack":
                ta.pack_sb(f["sb"], grid.addr(), 1, f["packed"].addr(), f["size"], 0, -1, stream)
            else:
                ta.unpack_sb(f["sb"], f["packed"].addr(), f["size"], grid2.addr(), 1, 0, -1,
                             stream)
            evs[i][1].record(stream)
        ta.synchronize(stream)
        for i in range(len(ops)):
            durs[i].append(evs[i][0].elapsed_ms(evs[i][1]))
    kernels = []
    for (kind, f), d in zip(ops, durs):
        d = np.array(d) * 1e3  # us
        avg = float(d.mean())
        variant = ta.select_variant(f["sb"], 1, grid.addr(), f["packed"].addr(),
                                    ta.DIR_PACK if kind == "pack" else ta.DIR_UNPACK)
        kernels.append({
            "op": f"{kind}_{f['face']}", "variant": variant,
            "avg_us": round(avg, 3), "median_us": round(float(np.median(d)), 3),
            "algorithmic_bytes": 2 * f["size"],
            "hbm_gbs": round(2 * f["size"] / (avg * 1e-6) / 1e9, 1),
        })
    dom = max(kernels, key=lambda k: k["avg_us"])
    log("per kernel: " + ", ".join(f"{k['op']} {k['avg_us']:.2f}us {k['hbm_gbs']:.0f}GB/s"
                                   for k in kernels))

    # bit-exactness of this run against the oracle
    from oracle import oracle as orc

    bit_exact = True
    for f in faces:
        exp = orc.sb_pack(orc.sb_desc(*f["sb"].as_tuple()), host, 0, 1)
        bit_exact &= bool((f["packed"].download() == exp).all())

    # synchronous MPI_Pack / MPI_Unpack calls (interposed), wall clock
    call_us = {}
    for f in faces:
        ts = []
        for _ in range(50):
            a = time.perf_counter()
            err, pos = m.pack(grid.addr(), 1, f["handle"], f["packed"].addr(), f["size"])
            ts.append(time.perf_counter() - a)
            assert err == 0
        call_us[f["face"]] = round(float(np.median(ts)) * 1e6, 2)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_grid_faces_518.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            t = json.load(fh)
        traffic = t.get(dom["op"])

    cpu = None
    cpu_lib = None
    if not args.no_cpu:
        cores = 1
        log(f"cpu baseline: oracle gather, {args.cpu_seconds:.0f} s sample, 1 thread")
        descs = [orc.sb_desc(*f["sb"].as_tuple()) for f in faces]
        scratch = host.copy()
        moved, t0 = 0, time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < args.cpu_seconds:
            for d, f in zip(descs, faces):
                p = orc.sb_pack(d, host, 0, 1)
                orc.sb_unpack(d, p, scratch, 0, 1)
                moved += 2 * f["size"]
            reps += 1
        el = time.perf_counter() - t0
        cpu = {"value": round(moved / el / 1e9, 3), "unit": "GB/s", "cores": cores,
               "kind": "port",
               "sample": f"oracle sb_pack+sb_unpack of the 3 faces of the 518^3 float32 grid, "
                         f"{reps} steps in {el:.1f} s on 1 host thread"}
        log(f"cpu baseline: underlying MPI library MPI_Pack+MPI_Unpack, "
            f"{args.cpu_seconds / 2:.0f} s sample")
        out = np.zeros(faces[0]["size"], np.uint8)
        moved, t0, reps = 0, time.perf_counter(), 0
        while time.perf_counter() - t0 < args.cpu_seconds / 2:
            for f in faces:
                err, pos = m.pack(host.ctypes.data, 1, f["handle"], out.ctypes.data, f["size"],
                                  library=True)
                err, pos = m.unpack(out.ctypes.data, f["size"], scratch.ctypes.data, 1,
                                    f["handle"], library=True)
                moved += 2 * f["size"]
            reps += 1
        el = time.perf_counter() - t0
        cpu_lib = {"value": round(moved / el / 1e9, 3), "unit": "GB/s", "cores": 1,
                   "library": "MPICH 3.3.2 MPI_Pack/MPI_Unpack (host buffers)",
                   "sample": f"{reps} steps of the 3 faces in {el:.1f} s"}

    result = {
        "metric": "MPI_Pack GB/s and % HBM peak, bit-exact; batched packs/sec vs per-call plain",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic: normal(0,1) float32 grid, seed 2, resident in HBM",
        "config": {
            "workload": "3D subarray faces of a 518^3 float32 grid (512^3 + 3-cell ghost ring): "
                        "pack+unpack of the x, y, z faces (512x512x3, 512x3x512, 3x512x512)",
            "packed_bytes_per_step": step_bytes,
            "path": "tempi_pack_sb/tempi_unpack_sb on the library stream (MPI_Pack's kernels)",
        },
        "bit_exact": bit_exact,
        "roofline": {
            "bound": "hbm",
            "kernel": dom["op"],
            "achieved": dom["hbm_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom["hbm_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "kernels": kernels,
        "mpi_pack_call_us": call_us,
        "cpu_baseline": cpu,
        "cpu_mpi_library": cpu_lib,
    }
    print(json.dumps(result))
    for f in faces:
        m.free(f["handle"])
    m.finalize()


if __name__ == "__main__":
    main()
