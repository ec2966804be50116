"""Benchmark of the derived-datatype pack/unpack path on one MI355X.

Headline workload (one "step"): the three 3-cell-thick faces of a 518^3
float32 grid (a 512^3 interior with a 3-cell ghost ring; normal(0,1) values,
seed 2), each described by MPI_Type_create_subarray and committed through the
interposed MPI_Type_commit.  A step packs the three faces and unpacks them
into a second grid: one batched launch for the packs and one for the unpacks
(tempi_batch_launch on the library stream).  The grids stay resident in HBM;
the timed region is K steps between two device synchronizations.

value    packed bytes moved per second, pack + unpack (6 x 3 MiB per step)
roofline the batch kernel's algorithmic HBM bytes per launch (2 x packed:
         each byte read once and written once) over its average device time
         per launch, from HIP events on the library stream around the timed
         region, against the 8 TB/s HBM3E peak
extras   the same faces as six separate kernels (the MPI_Pack kernels) and
         as synchronous MPI_Pack calls; cold-cache kernel times (caches
         swept by a read-only 1 GiB pass before each step); the halo batch (26
         halos of 64 subdomains of 128^3 float64, radius 3) and 4096 random
         requests, batched vs one call per request; the 2D vector case.
cpu_baseline  the oracle's single-threaded C gather/scatter (oracle/oracle.c)
         on the host copy of the grid, over a bounded sample; the underlying
         MPI library's own MPI_Pack / MPI_Unpack (where the reference sends
         host buffers, src/pack.cpp:51-54) is reported beside it, and the
         oracle's gather/scatter in one process per host core (at most 16).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import random
import subprocess
import tempfile
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_loop(ta, fn, steps, stream):
    """Wall and device time of `steps` calls of fn() between two syncs."""
    a, b = ta.Event(), ta.Event()
    ta.synchronize(stream)
    t0 = time.perf_counter()
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    ta.synchronize(stream)
    t1 = time.perf_counter()
    return t1 - t0, a.elapsed_ms(b) * 1e-3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="warmup + timed headline steps only (for rocprofv3)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="with --profile-only: also run one plain 256 MiB contiguous copy "
                         "(the reference kernel for scaling FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--profile-large-blocks", action="store_true",
                    help="only the 256 MiB 2D packs/unpacks (64 B - 4 KiB blocks, aligned "
                         "and 4 B off), three cold launches each, then the 256 MiB "
                         "contiguous calibration copy (for rocprofv3)")
    ap.add_argument("--profile-sweep", action="store_true",
                    help="the sweep's copies of >= 64-B blocks and >= 16 MiB alone, each pack "
                         "and unpack launched 3 times after a cache sweep (for rocprofv3)")
    ap.add_argument("--profile-random", action="store_true",
                    help="only the 4096 random requests as one batch, three launches back "
                         "to back and three after a cache sweep, then the 256 MiB "
                         "contiguous calibration copy (for rocprofv3)")
    ap.add_argument("--profile-halo", action="store_true",
                    help="only the halo batch (1664 requests), warm and cold launches, then "
                         "the 256 MiB contiguous calibration copy (for rocprofv3)")
    ap.add_argument("--cpu-worker", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args.cpu_worker)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU's share of the pack path; --gpus {args.gpus} is not "
              "supported: packing does not communicate, so N GPUs run N independent copies "
              "of this measurement, each on the subdomains it owns.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import tempi_amd as ta
    import typecatalog as tc
    from tempi_amd import mpi as tmpi

    assert ta.device_count() > 0, "no GPU"
    m = tmpi.get()
    stream = ta.stream()
    if args.profile_large_blocks:
        profile_large_blocks(ta, stream)
        return
    if args.profile_halo:
        profile_halo(ta, m, tc, stream)
        return
    if args.profile_random:
        profile_random(ta, tc, stream)
        return
    if args.profile_sweep:
        profile_sweep(ta, m, tc, stream)
        return
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
    packed_per_step = sum(2 * f["size"] for f in faces)
    pack_batch = ta.Batch([(f["sb"], grid.addr(), f["packed"].addr(), 1, ta.DIR_PACK)
                           for f in faces])
    unpack_batch = ta.Batch([(f["sb"], grid2.addr(), f["packed"].addr(), 1, ta.DIR_UNPACK)
                             for f in faces])

    def step_batched():
        pack_batch.launch(stream)
        unpack_batch.launch(stream)

    for _ in range(args.warmup):
        step_batched()
    wall, dev = timed_loop(ta, step_batched, args.steps, stream)
    ms_per_step = wall * 1e3 / args.steps
    value = packed_per_step * args.steps / wall / 1e9
    batch_avg_us = dev * 1e6 / (2 * args.steps)
    log(f"batched: {ms_per_step * 1e3:.2f} us/step, {value:.1f} GB/s packed, "
        f"batch kernel {batch_avg_us:.2f} us/launch (device)")
    if args.profile_only:
        if args.calibrate_copy:
            # 256 MiB read + 256 MiB written by the contiguous-copy kernel
            cal = 256 << 20
            a, b = ta.DeviceBuffer(cal), ta.DeviceBuffer(cal)
            a.fill(1)
            sb1 = ta.StridedBlock.make(0, cal, [cal], [1])
            ta.pack_sb(sb1, a.addr(), 1, b.addr(), cal, 0, 1, stream)
            ta.synchronize(stream)
        print(json.dumps({"profile_only": True, "ms_per_step": ms_per_step,
                          "batch_avg_us": batch_avg_us}))
        return

    from oracle import oracle as orc

    # bit-exactness of this run against the oracle
    bit_exact = True
    expected = {}
    for f in faces:
        expected[f["face"]] = orc.sb_pack(orc.sb_desc(*f["sb"].as_tuple()), host, 0, 1)
        bit_exact &= bool((f["packed"].download() == expected[f["face"]]).all())

    extras = {}
    # the variant of each face in the two batches (settled when the batches
    # were created: the library times its candidates on the batch itself)
    extras["batch_variants"] = {
        "pack": [pack_batch.variant(i) for i in range(len(faces))],
        "unpack": [unpack_batch.variant(i) for i in range(len(faces))],
    }
    log(f"batch variants (x, y, z faces): {extras['batch_variants']}")
    # the same two batches with the closed-form rule's variants
    rb = [ta.Batch([(f["sb"], grid.addr(), f["packed"].addr(), 1, ta.DIR_PACK)
                    for f in faces], method=ta.SELECT_HEURISTIC),
          ta.Batch([(f["sb"], grid2.addr(), f["packed"].addr(), 1, ta.DIR_UNPACK)
                    for f in faces], method=ta.SELECT_HEURISTIC)]
    for _ in range(args.warmup):
        rb[0].launch(stream)
        rb[1].launch(stream)
    wall_r, dev_r = timed_loop(ta, lambda: (rb[0].launch(stream), rb[1].launch(stream)),
                               args.steps, stream)
    extras["batch_rule_variants"] = {
        "pack": [rb[0].variant(i) for i in range(len(faces))],
        "unpack": [rb[1].variant(i) for i in range(len(faces))],
        "us_per_step": round(wall_r * 1e6 / args.steps, 3),
        "kernel_us_per_launch": round(dev_r * 1e6 / (2 * args.steps), 3)}
    for b in rb:
        b.destroy()
    log(f"rule variants: {extras['batch_rule_variants']}")
    # -- the same step as six single-copy kernels (what MPI_Pack launches)
    def step_per_call():
        for f in faces:
            ta.pack_sb(f["sb"], grid.addr(), 1, f["packed"].addr(), f["size"], 0, -1, stream)
        for f in faces:
            ta.unpack_sb(f["sb"], f["packed"].addr(), f["size"], grid2.addr(), 1, 0, -1, stream)

    for _ in range(args.warmup):
        step_per_call()
    wall_pc, _ = timed_loop(ta, step_per_call, args.steps, stream)
    extras["six_kernels_per_step"] = {
        "us_per_step": round(wall_pc * 1e6 / args.steps, 3),
        "gbs": round(packed_per_step * args.steps / wall_pc / 1e9, 1)}
    # graph replay of the two batches
    for _ in range(args.warmup):
        pack_batch.graph_launch(stream)
        unpack_batch.graph_launch(stream)
    wall_g, _ = timed_loop(ta, lambda: (pack_batch.graph_launch(stream),
                                        unpack_batch.graph_launch(stream)), args.steps, stream)
    extras["batched_graph"] = {"us_per_step": round(wall_g * 1e6 / args.steps, 3),
                               "gbs": round(packed_per_step * args.steps / wall_g / 1e9, 1),
                               "form": "two graph launches per step, one per batch"}
    # the whole step as one graph: the pack batch's launch, then the unpack
    # batch's (tempi_graph_create)
    step_graph = ta.Graph([pack_batch, unpack_batch])
    for _ in range(args.warmup):
        step_graph.launch(stream)
    wall_sg, _ = timed_loop(ta, lambda: step_graph.launch(stream), args.steps, stream)
    extras["step_graph"] = {"us_per_step": round(wall_sg * 1e6 / args.steps, 3),
                            "gbs": round(packed_per_step * args.steps / wall_sg / 1e9, 1),
                            "form": "one graph of both batch launches per step"}
    step_graph.destroy()
    log(f"graph replay: {extras['batched_graph']}, whole step: {extras['step_graph']}")

    # per-kernel device time, each single-copy kernel launched back to back
    kernels = []
    for kind in ("pack", "unpack"):
        for f in faces:
            def one(f=f, kind=kind):
                if kind == "pack":
                    ta.pack_sb(f["sb"], grid.addr(), 1, f["packed"].addr(), f["size"], 0, -1,
                               stream)
                else:
                    ta.unpack_sb(f["sb"], f["packed"].addr(), f["size"], grid2.addr(), 1, 0,
                                 -1, stream)
            _, d = timed_loop(ta, one, 100, stream)
            us = d * 1e6 / 100
            kernels.append({"op": f"{kind}_{f['face']}", "us": round(us, 3),
                            "variant": ta.select_variant(
                                f["sb"], 1, grid.addr(), f["packed"].addr(),
                                ta.DIR_PACK if kind == "pack" else ta.DIR_UNPACK,
                                ta.SELECT_DEFAULT),
                            "hbm_gbs": round(2 * f["size"] / us / 1e3, 1)})
    extras["single_kernels_warm"] = kernels

    # each headline launch between its own pair of events, warm: median and
    # trimean (Q1 + 2 median + Q3) / 4, the reference's summary statistic
    # (src/internal/statistics.cpp:30-39)
    per = {"pack_batch": [], "unpack_batch": []}
    evs = [ta.Event() for _ in range(3)]
    for _ in range(max(args.steps, 100)):
        evs[0].record(stream)
        pack_batch.launch(stream)
        evs[1].record(stream)
        unpack_batch.launch(stream)
        evs[2].record(stream)
        ta.synchronize(stream)
        per["pack_batch"].append(evs[0].elapsed_ms(evs[1]) * 1e3)
        per["unpack_batch"].append(evs[1].elapsed_ms(evs[2]) * 1e3)
    stats = {}
    for k, v in per.items():
        q1, med, q3 = (float(x) for x in np.percentile(v, [25, 50, 75]))
        stats[k] = {"n": len(v), "median_us": round(med, 3),
                    "trimean_us": round((q1 + 2 * med + q3) / 4, 3),
                    "min_us": round(float(np.min(v)), 3)}
    extras["headline_launch_events"] = stats

    # cold caches: a read-only 1 GiB sweep before each step leaves L2 and the
    # Infinity Cache full of clean, unrelated lines
    ev = [ta.Event() for _ in range(4)]
    cold = {"pack_batch_us": [], "unpack_batch_us": []}
    for _ in range(20):
        ta.flush_caches()
        ev[0].record(stream)
        pack_batch.launch(stream)
        ev[1].record(stream)
        unpack_batch.launch(stream)
        ev[2].record(stream)
        ta.synchronize(stream)
        cold["pack_batch_us"].append(ev[0].elapsed_ms(ev[1]) * 1e3)
        cold["unpack_batch_us"].append(ev[1].elapsed_ms(ev[2]) * 1e3)
    half = packed_per_step // 2
    extras["cold_cache_batches"] = {
        k: {"median_us": round(float(np.median(v)), 3),
            "hbm_gbs": round(2 * half / float(np.median(v)) / 1e3, 1)}
        for k, v in cold.items()}

    # the launch floor: the same batch kernel moving one 64-B row, and the
    # x face alone and the y + z faces alone, timed like the headline
    floor = {}
    tiny = ta.DeviceBuffer(64)
    parts = {
        "one_64B_row": [(ta.StridedBlock.make(0, 64, [64], [1]), grid.addr(), tiny.addr(), 1,
                         ta.DIR_PACK)],
        "x_face_only": [(faces[0]["sb"], grid.addr(), faces[0]["packed"].addr(), 1,
                         ta.DIR_PACK)],
        "y_z_faces_only": [(f["sb"], grid.addr(), f["packed"].addr(), 1, ta.DIR_PACK)
                           for f in faces[1:]],
    }
    for name, reqs in parts.items():
        b = ta.Batch(reqs)
        for _ in range(args.warmup):
            b.launch(stream)
        _, dv = timed_loop(ta, lambda b=b: b.launch(stream), args.steps, stream)
        floor[name] = round(dv * 1e6 / args.steps, 3)
        b.destroy()
    # the single-copy kernel moving the same 64-B row: the launch floor
    # without any batch bookkeeping
    one = parts["one_64B_row"][0][0]
    for _ in range(args.warmup):
        ta.pack_sb(one, grid.addr(), 1, tiny.addr(), 64, 0, -1, stream)
    _, dv = timed_loop(ta, lambda: ta.pack_sb(one, grid.addr(), 1, tiny.addr(), 64, 0, -1,
                                              stream), args.steps, stream)
    floor["one_64B_row_single_copy_kernel"] = round(dv * 1e6 / args.steps, 3)
    floor["headline_pack_unpack_average"] = round(batch_avg_us, 3)
    extras["pack_batch_us_per_launch"] = floor
    tiny.free()

    # synchronous MPI_Pack / MPI_Unpack calls (interposed), wall clock
    call_us = {}
    for f in faces:
        ts = []
        for _ in range(50):
            t0 = time.perf_counter()
            err, pos = m.pack(grid.addr(), 1, f["handle"], f["packed"].addr(), f["size"])
            ts.append(time.perf_counter() - t0)
            assert err == 0
        call_us[f"MPI_Pack_{f['face']}"] = round(float(np.median(ts)) * 1e6, 2)
        bit_exact &= bool((f["packed"].download() == expected[f["face"]]).all())
    extras["mpi_call_wall_us"] = call_us

    if not args.no_extras:
        extras["large_blocks_cold"] = bench_large_blocks(ta, stream)
        extras["halo_batch"] = bench_halos(ta, m, tc, orc, stream)
        extras["scattered_rows_cold"] = bench_scattered_rows(ta, stream)
        extras["random_requests"] = bench_random_requests(ta, tc, orc, stream)
        extras["vector_2d"] = bench_vector(ta, m, tc, stream)
        extras["system_tables"] = bench_system_tables(ta)
        extras["vector_sweep"] = bench_sweep(ta, m, tc, stream)
        extras["selector_regret"] = bench_selector_regret(ta)
        extras["isend_irecv_faces"] = bench_p2p(ta, m, faces, grid, grid2, host)
        extras["p2p_small_latency"] = bench_p2p_small(ta, m, tc)
        extras["neighbor_alltoallw_halo"] = bench_neighbor(ta, m, tc)

        extras["reference_mpi_pack_sweep"] = bench_reference_sweep()
        extras["type_commit"] = bench_type_commit()
        extras["first_call_large_copy"] = bench_first_call(ta, m, tc)
        for k in ("halo_batch", "random_requests"):
            bit_exact &= bool(extras[k]["bit_exact"])

    def traffic_of(name):
        tpath = os.path.join(ROOT, "profiles", name)
        if os.path.exists(tpath):
            with open(tpath) as fh:
                return json.load(fh).get("per_launch")
        return None

    traffic = traffic_of("traffic_grid_faces_518.json")

    cpu = cpu_oracle = cpu_all = None
    if not args.no_cpu:
        cpu, cpu_oracle = cpu_baselines(args, m, orc, faces, host, np)
        cpu_all = cpu_all_cores(m, faces, host.size, args.cpu_seconds / 2)

    # each batch launch moves half the step's packed bytes, read once and
    # written once: 2 x (packed_per_step / 2) algorithmic HBM bytes
    achieved_gbs = packed_per_step / (batch_avg_us * 1e-6) / 1e9
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
                        "pack + unpack of the x, y, z faces (512x512x3, 512x3x512, 3x512x512)",
            "packed_bytes_per_step": packed_per_step,
            "path": "tempi_batch_launch: one persistent-queue launch packs the 3 faces, one "
                    "unpacks them (library stream)",
        },
        "bit_exact": bit_exact,
        "selector": "model" if ta.lib().tempi_mlp_loaded() else "closed-form rule",
        "roofline": {
            "bound": "hbm",
            # batches of up to 8 requests launch k_batch_small
            "kernel": "k_batch_small" if len(faces) <= 8 else "k_batch",
            "achieved": round(achieved_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": packed_per_step,
            "avg_launch_us": round(batch_avg_us, 3),
        },
        "extras": extras,
        "cpu_baseline": cpu,
        "cpu_all_cores": cpu_all,
        "cpu_oracle": cpu_oracle,
    }
    halo = extras.get("halo_batch")
    if halo:
        # the largest batched configuration: 26 halos of 64 subdomains of
        # 128^3 float64 in one k_batch launch, 2 x 158 183 424 algorithmic B
        ach = 2 * halo["packed_bytes"] / (halo["batch_kernel_us"] * 1e-6) / 1e9
        result["roofline_halo"] = {
            "bound": "hbm", "kernel": "k_batch",
            "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": traffic_of("traffic_halo_64x128cube_radius3.json"),
            "algorithmic_bytes_per_launch": 2 * halo["packed_bytes"],
            "avg_launch_us": halo["batch_kernel_us"],
            "bit_exact": halo["bit_exact"],
        }
    print(json.dumps(result))
    for f in faces:
        m.free(f["handle"])
    pack_batch.destroy()
    unpack_batch.destroy()
    m.finalize()


def run_default(ta, sb, d, src, dst, n, stream, off=0):
    """One pack or unpack with the library's default choice, synchronized;
    the strided side starts `off` bytes into `src`."""
    if d == ta.DIR_PACK:
        ta.pack_sb(sb, src.addr(off), 1, dst.addr(), n, 0, -1, stream)
    else:
        ta.unpack_sb(sb, dst.addr(), n, src.addr(off), 1, 0, -1, stream)
    ta.synchronize(stream)


def profile_large_blocks(ta, stream, reps=3):
    """Cold launches of each large-block copy with the library's default
    choice (settled by a first call of each shape, before the measured
    launches): `reps` rounds over the copies in a fixed order -- rows of
    64 B - 4 KiB, the strided side at a 64-B boundary and 4 bytes past one
    (_off4) -- then one 256 MiB contiguous copy (variant 1).  The strided-copy
    dispatches after the settling calls are the ones a trace or counter run
    attributes to the copies printed here, in the same order."""
    n = 256 << 20
    src = ta.DeviceBuffer(3 << 30)
    src.fill(3)
    dst = ta.DeviceBuffer(n)
    dst.fill(4)
    shapes = []
    for bl in (64, 256, 1024, 4096):
        rows = n // bl
        stride = 2 * bl + 64
        sb = ta.StridedBlock.make(0, rows * stride, [bl, rows], [1, stride])
        for off in (0, 4):
            for d, name in ((ta.DIR_PACK, "pack"), (ta.DIR_UNPACK, "unpack")):
                run_default(ta, sb, d, src, dst, n, stream, off)  # settles the choice
                shapes.append((f"{name}_bl{bl}" + ("_off4" if off else ""), sb, d, off))
    for rep in range(reps):
        for name, sb, d, off in shapes:
            v = ta.select_variant(sb, 1, src.addr(off), dst.addr(), d, ta.SELECT_DEFAULT)
            ta.flush_caches()
            ev = [ta.Event(), ta.Event()]
            ev[0].record(stream)
            if d == ta.DIR_PACK:
                ta.pack_sb(sb, src.addr(off), 1, dst.addr(), n, 0, v, stream)
            else:
                ta.unpack_sb(sb, dst.addr(), n, src.addr(off), 1, 0, v, stream)
            ev[1].record(stream)
            ta.synchronize(stream)
            print(json.dumps({"copy": name, "rep": rep, "variant": v,
                              "us": round(ev[0].elapsed_ms(ev[1]) * 1e3, 2)}), flush=True)
    src.free()
    cal = ta.DeviceBuffer(n)
    cal.fill(1)
    sb1 = ta.StridedBlock.make(0, n, [n], [1])
    ta.flush_caches()
    ta.pack_sb(sb1, cal.addr(), 1, dst.addr(), n, 0, 1, stream)
    ta.synchronize(stream)
    print(json.dumps({"copy": "contiguous_256MiB", "rep": 0, "variant": 1}), flush=True)


def bench_large_blocks(ta, stream):
    """2D packs and unpacks of 256 MiB with block lengths 64 B - 4 KiB (stride
    = 2 x block + 64 B), each launch timed after a read-only cache sweep
    (the library times copies this large that way), with the library's
    default choice (what MPI_Pack runs) and with the autotuner's (variant,
    grid cap).  Keys ending in _off4 repeat the copies with the strided
    side 4 bytes past a 64-B boundary (float data at an odd element
    offset): every row then starts and ends inside a cache line."""
    n = 256 << 20
    src = ta.DeviceBuffer(3 << 30)
    src.fill(3)
    dst = ta.DeviceBuffer(n)
    dst.fill(4)
    res = {}
    for bl in (64, 256, 1024, 4096):
        rows = n // bl
        stride = 2 * bl + 64
        sb = ta.StridedBlock.make(0, rows * stride, [bl, rows], [1, stride])
        for (d, name), off in [(dn, o) for o in (0, 4) for dn in
                               ((ta.DIR_PACK, "pack"), (ta.DIR_UNPACK, "unpack"))]:
            # the first call settles the default choice for the shape
            run_default(ta, sb, d, src, dst, n, stream, off)
            vd = ta.select_variant(sb, 1, src.addr(off), dst.addr(), d, ta.SELECT_DEFAULT)
            us_d = ta.time_variant(sb, 1, src.addr(off), dst.addr(), d, vd, 0, 5)
            va, ca, _ = ta.autotune(sb, 1, src.addr(off), dst.addr(), d)
            us_a = ta.time_variant(sb, 1, src.addr(off), dst.addr(), d, va, ca, 5)
            gd, ga = 2 * n / us_d / 1e3, 2 * n / us_a / 1e3
            res[f"{name}_bl{bl}" + ("_off4" if off else "")] = {
                "default": {"variant": vd, "us": round(us_d, 2), "hbm_gbs": round(gd, 1),
                            "frac": round(gd / HBM_PEAK_GBS, 4)},
                "autotuned": {"variant": va, "grid_cap": ca, "us": round(us_a, 2),
                              "hbm_gbs": round(ga, 1), "frac": round(ga / HBM_PEAK_GBS, 4)},
            }
    src.free()
    dst.free()
    log(f"large blocks (cold): {res}")
    return res


def profile_sweep(ta, m, tc, stream, ncases=48, reps=3):
    """The sweep cases of >= 64-B blocks and >= 16 MiB packed (seeds 1000 + k,
    as bench_sweep), each packed and then unpacked with the library's
    default choice, `reps` launches each after a read-only 1 GiB sweep (the
    trace: k_empty marks each case and direction, then the settling call's
    kernels, then k_touch and the copy kernel per timed launch), then one
    256 MiB contiguous copy (variant 1, the FETCH_SIZE / WRITE_SIZE
    calibration).  One JSON line per timed launch, in trace order, with the
    case's line bound."""
    for k in range(ncases):
        spec, base, packed_bytes = tc.sweep_type(random.Random(1000 + k))
        if packed_bytes < (16 << 20):
            continue
        h = m.type_from_spec(spec)
        sb, size = m.strided_block(h)
        m.free(h)
        if sb.counts[0] < 64 or base + sb.extent > (8 << 30):
            continue
        src = ta.DeviceBuffer(base + sb.extent + 64)
        dst = ta.DeviceBuffer(size + 64)
        ls, lp = line_bound(sb, src.addr(base) % 128)
        ev = [ta.Event(), ta.Event()]
        for d, name in ((ta.DIR_PACK, "pack"), (ta.DIR_UNPACK, "unpack")):
            def copy(variant, d=d):
                if d == ta.DIR_PACK:
                    ta.pack_sb(sb, src.addr(base), 1, dst.addr(), size, 0, variant, stream)
                else:
                    ta.unpack_sb(sb, dst.addr(), size, src.addr(base), 1, 0, variant, stream)

            # k_empty marks the start of each case and direction in the
            # trace; the first call settles the variant (its trials precede
            # the timed launches, which are the group's last `reps` copies)
            ta.launch_latency_us(1)
            copy(-1)
            ta.synchronize(stream)
            v = ta.select_variant(sb, 1, src.addr(base), dst.addr(), d, ta.SELECT_DEFAULT)
            for _ in range(reps):
                ta.flush_caches()
                ev[0].record(stream)
                copy(v)
                ev[1].record(stream)
                ta.synchronize(stream)
                print(json.dumps({"case": k, "op": name, "variant": v, "block": int(sb.counts[0]),
                                  "packed_bytes": size, "strided_lines": ls, "packed_lines": lp,
                                  "line_bound_frac": round(2 * size / (128 * (ls + lp)), 4),
                                  "us": round(ev[0].elapsed_ms(ev[1]) * 1e3, 2)}), flush=True)
        src.free()
        dst.free()
    cal = 256 << 20
    a, c = ta.DeviceBuffer(cal), ta.DeviceBuffer(cal)
    a.fill(1)
    ta.pack_sb(ta.StridedBlock.make(0, cal, [cal], [1]), a.addr(), 1, c.addr(), cal, 0, 1,
               stream)
    ta.synchronize(stream)


def halo_requests(ta, m, tc, hosts=None):
    """The pack requests of 26 halo subarrays (radius 3, float64) of 64
    subdomains of 128^3, into one packed buffer: (requests, total packed
    bytes, buffers to keep alive).  With `hosts` (64 host arrays) the
    subdomains hold those bytes, else a constant."""
    nsub, lcr, radius, quant = 64, (128, 128, 128), 3, 8
    sub_bytes = (lcr[0] + 6) * (lcr[1] + 6) * (lcr[2] + 6) * quant
    subs = [ta.DeviceBuffer(sub_bytes) for _ in range(nsub)]
    for i, s in enumerate(subs):
        if hosts is not None:
            s.upload(hosts[i])
        else:
            s.fill(7)
    halos = []
    for d in tc.halo_directions():
        h = m.type_from_spec(tc.halo_type(radius, lcr, d, quant))
        sb, size = m.strided_block(h)
        m.free(h)
        halos.append((sb, size))
    total = nsub * sum(sz for _, sz in halos)
    packed = ta.DeviceBuffer(total)
    reqs, off = [], 0
    for s in subs:
        for sb, size in halos:
            reqs.append((sb, s.addr(), packed.addr(off), 1, ta.DIR_PACK))
            off += size
    return reqs, total, subs + [packed]


def profile_halo(ta, m, tc, stream, warm=5, cold=3):
    """The halo batch alone (for rocprofv3): `warm` launches back to back,
    then `cold` launches each after a read-only 1 GiB sweep, then one 256 MiB
    contiguous copy (variant 1, the FETCH_SIZE / WRITE_SIZE calibration).
    The batch's creation launches no copy kernel (its requests are below
    the size at which the library times candidates; the variant selector's
    inference runs as k_mlp_infer), so the k_batch dispatches of a trace are
    these launches, in this order; one JSON line each."""
    reqs, total, keep = halo_requests(ta, m, tc)
    b = ta.Batch(reqs)
    ev = [ta.Event(), ta.Event()]
    for i in range(warm + cold):
        kind = "warm" if i < warm else "cold"
        if kind == "cold":
            ta.flush_caches()
        ta.synchronize(stream)
        ev[0].record(stream)
        b.launch(stream)
        ev[1].record(stream)
        ta.synchronize(stream)
        print(json.dumps({"launch": kind, "us": round(ev[0].elapsed_ms(ev[1]) * 1e3, 2),
                          "requests": len(reqs), "packed_bytes": total}), flush=True)
    b.destroy()
    cal = 256 << 20
    a, c = ta.DeviceBuffer(cal), ta.DeviceBuffer(cal)
    a.fill(1)
    ta.pack_sb(ta.StridedBlock.make(0, cal, [cal], [1]), a.addr(), 1, c.addr(), cal, 0, 1,
               stream)
    ta.synchronize(stream)
    for x in keep + [a, c]:
        x.free()


def profile_random(ta, tc, stream, warm=3, cold=3):
    """The 4096 random requests of bench_random_requests as one batch (for
    rocprofv3): `warm` launches back to back, then `cold` launches each after
    a read-only 1 GiB sweep, then one 256 MiB contiguous copy (variant 1, the
    FETCH_SIZE / WRITE_SIZE calibration).  No request reaches the size at
    which the library times candidates, so the batch's creation launches no
    copy kernel and the k_batch dispatches of a trace are these launches, in
    order; one JSON line each, with the batch's algorithmic bytes and the
    128-B lines its requests touch (summed over requests: they all read one
    source buffer, so lines shared between requests count once per
    request)."""
    geo, span_max = tc.random_requests(4, 4096)
    sbs = [ta.StridedBlock.make(*g) for g in geo]
    total = sum(sb.size() for sb in sbs)
    src = ta.DeviceBuffer(span_max)
    src.fill(7)
    packed = ta.DeviceBuffer(total)
    reqs, off, lines = [], 0, 0
    for sb in sbs:
        reqs.append((sb, src.addr(), packed.addr(off), 1, ta.DIR_PACK))
        ls, lp = line_bound(sb, src.addr() % 128)
        lines += ls + lp
        off += sb.size()
    b = ta.Batch(reqs)
    ev = [ta.Event(), ta.Event()]
    for i in range(warm + cold):
        kind = "warm" if i < warm else "cold"
        if kind == "cold":
            ta.flush_caches()
        ta.synchronize(stream)
        ev[0].record(stream)
        b.launch(stream)
        ev[1].record(stream)
        ta.synchronize(stream)
        print(json.dumps({"launch": kind, "us": round(ev[0].elapsed_ms(ev[1]) * 1e3, 2),
                          "requests": len(reqs), "packed_bytes": total,
                          "lines_touched": lines}), flush=True)
    b.destroy()
    cal = 256 << 20
    a, c = ta.DeviceBuffer(cal), ta.DeviceBuffer(cal)
    a.fill(1)
    ta.pack_sb(ta.StridedBlock.make(0, cal, [cal], [1]), a.addr(), 1, c.addr(), cal, 0, 1,
               stream)
    ta.synchronize(stream)
    for x in (src, packed, a, c):
        x.free()


def bench_halos(ta, m, tc, orc, stream):
    """26 halo subarrays (radius 3, float64) of 64 subdomains of 128^3 of
    seeded bytes: one batched launch / graph replay vs one tempi_pack_sb +
    sync per halo; the packed bytes of the last batched launch are checked
    against the oracle."""
    import numpy as np

    sub_bytes = 134 ** 3 * 8
    hosts = [orc.splitmix_bytes(sub_bytes, 4000 + s) for s in range(64)]
    reqs, total, keep = halo_requests(ta, m, tc, hosts)
    packed = keep[-1]
    t_create = time.perf_counter()
    b = ta.Batch(reqs)
    create_ms = (time.perf_counter() - t_create) * 1e3
    nreq = len(reqs)
    for _ in range(5):
        b.launch(stream)
    reps = 50
    wall_b, dev_b = timed_loop(ta, lambda: b.launch(stream), reps, stream)
    got = packed.download()
    exact, off = True, 0
    for i, r in enumerate(reqs):
        exp = orc.sb_pack(orc.sb_desc(*r[0].as_tuple()), hosts[i // 26], 0, 1)
        exact &= bool((got[off:off + exp.size] == exp).all())
        off += exp.size
    del got
    wall_g, _ = timed_loop(ta, lambda: b.graph_launch(stream), reps, stream)
    # the same requests with the closed-form rule's variants and with the
    # selector's per-copy picks
    br = ta.Batch(reqs, method=ta.SELECT_HEURISTIC)
    bm = ta.Batch(reqs, method=ta.SELECT_MODEL)
    for _ in range(5):
        br.launch(stream)
        bm.launch(stream)
    _, dev_r = timed_loop(ta, lambda: br.launch(stream), reps, stream)
    _, dev_m = timed_loop(ta, lambda: bm.launch(stream), reps, stream)
    picks = {}
    for i in range(26):
        picks[str(i)] = [b.variant(i), br.variant(i), bm.variant(i)]
    br.destroy()
    bm.destroy()

    def per_call():
        o = 0
        for r in reqs:
            ta.pack_sb(r[0], r[1], 1, packed.addr(), total, o, -1, stream)
            ta.synchronize(stream)
            o += r[0].size()

    per_call()
    t0 = time.perf_counter()
    for _ in range(3):
        per_call()
    wall_pc = (time.perf_counter() - t0) / 3

    def per_call_async():
        o = 0
        for r in reqs:
            ta.pack_sb(r[0], r[1], 1, packed.addr(), total, o, -1, stream)
            o += r[0].size()

    wall_pa, _ = timed_loop(ta, per_call_async, 3, stream)
    wall_pa /= 3
    res = {
        "requests": nreq, "packed_bytes": total,
        "create_ms": round(create_ms, 2),
        "batched_packs_per_s": round(nreq * reps / wall_b, 1),
        "graph_packs_per_s": round(nreq * reps / wall_g, 1),
        "per_call_sync_packs_per_s": round(nreq / wall_pc, 1),
        "per_call_async_packs_per_s": round(nreq / wall_pa, 1),
        "batched_us_per_launch": round(wall_b * 1e6 / reps, 2),
        "batch_kernel_us": round(dev_b * 1e6 / reps, 2),
        "batch_kernel_us_rule_variants": round(dev_r * 1e6 / reps, 2),
        "batch_kernel_us_selector_variants": round(dev_m * 1e6 / reps, 2),
        "variants_default_rule_selector_first_subdomain": picks,
        "batch_kernel_hbm_gbs": round(2 * total / (dev_b / reps) / 1e9, 1),
        "bit_exact": exact,
    }
    res["speedup_batched_vs_per_call_sync"] = round(res["batched_packs_per_s"]
                                                    / res["per_call_sync_packs_per_s"], 1)
    res["speedup_batched_vs_per_call_async"] = round(res["batched_packs_per_s"]
                                                     / res["per_call_async_packs_per_s"], 1)
    log(f"halo batch: {res}")
    b.destroy()
    return res


def bench_scattered_rows(ta, stream, reps=5):
    """2^20 rows of 24 B at a 1072-B stride (the pattern of the halo's x
    faces, 1.1 GB span), packed cold (caches swept before each launch) by
    the single-copy kernel, as a one-request batch and as 64 requests of one
    batch: the line-read rate of the pattern in each form (HIP events around
    each launch, median of `reps`)."""
    n, row, stride = 1 << 20, 24, 1072
    span = (n - 1) * stride + row
    src = ta.DeviceBuffer(span)
    src.fill(3)
    dst = ta.DeviceBuffer(n * row)
    sb = ta.StridedBlock.make(0, span, [row, n], [1, stride])
    ev = [ta.Event(), ta.Event()]

    def cold(fn):
        ts = []
        for _ in range(reps):
            ta.flush_caches()
            ta.synchronize(stream)
            ev[0].record(stream)
            fn()
            ev[1].record(stream)
            ta.synchronize(stream)
            ts.append(ev[0].elapsed_ms(ev[1]) * 1e3)
        return sorted(ts)[len(ts) // 2]

    v = ta.select_variant(sb, 1, src.addr(), dst.addr(), ta.DIR_PACK, ta.SELECT_DEFAULT)
    one = ta.Batch([(sb, src.addr(), dst.addr(), 1, ta.DIR_PACK)])
    part = n // 64
    sbk = ta.StridedBlock.make(0, (part - 1) * stride + row, [row, part], [1, stride])
    many = ta.Batch([(sbk, src.addr(k * part * stride), dst.addr(k * part * row), 1,
                      ta.DIR_PACK) for k in range(64)])
    res = {}
    for name, fn in (
            ("single_copy_kernel",
             lambda: ta.pack_sb(sb, src.addr(), 1, dst.addr(), n * row, 0, v, stream)),
            ("batch_of_1", lambda: one.launch(stream)),
            ("batch_of_64", lambda: many.launch(stream))):
        us = cold(fn)
        res[name] = {"us": round(us, 2), "line_reads_per_s": round(n / (us * 1e-6), 1),
                     "line_tbs": round(n * 128 / us / 1e6, 2)}
    ok = bool((dst.download() == 3).all())
    one.destroy()
    many.destroy()
    src.free()
    dst.free()
    res["packed_ok"] = ok
    log(f"scattered 24-B rows (cold): {res}")
    return res


def bench_random_requests(ta, tc, orc, stream, nreq=4096, seed=4):
    """4096 random 1D/2D/3D strided blocks with packed sizes log-uniform in
    [64 B, 16 MiB], all reading one source buffer, packed to disjoint
    outputs: batched vs per call."""
    geo, span_max = tc.random_requests(seed, nreq)
    reqs_geo = [ta.StridedBlock.make(*g) for g in geo]
    total = sum(sb.size() for sb in reqs_geo)
    host = orc.splitmix_bytes(span_max, 46)
    src = ta.DeviceBuffer(span_max)
    src.upload(host)
    packed = ta.DeviceBuffer(total)
    reqs, off = [], 0
    for sb in reqs_geo:
        reqs.append((sb, src.addr(), packed.addr(off), 1, ta.DIR_PACK))
        off += sb.size()
    t_create = time.perf_counter()
    b = ta.Batch(reqs)
    create_ms = (time.perf_counter() - t_create) * 1e3
    b.launch(stream)
    reps = 5
    wall_b, dev_b = timed_loop(ta, lambda: b.launch(stream), reps, stream)
    br = ta.Batch(reqs, method=ta.SELECT_HEURISTIC)
    br.launch(stream)
    _, dev_r = timed_loop(ta, lambda: br.launch(stream), reps, stream)
    br.destroy()
    bm = ta.Batch(reqs, method=ta.SELECT_MODEL)
    bm.launch(stream)
    _, dev_m = timed_loop(ta, lambda: bm.launch(stream), reps, stream)
    bm.destroy()
    b.launch(stream)
    ta.synchronize(stream)
    got = packed.download()
    exact, off = True, 0
    for sb in reqs_geo:
        exp = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), host, 0, 1)
        exact &= bool((got[off:off + exp.size] == exp).all())
        off += exp.size
    del got

    def per_call():
        o = 0
        for sb in reqs_geo:
            ta.pack_sb(sb, src.addr(), 1, packed.addr(), total, o, -1, stream)
            ta.synchronize(stream)
            o += sb.size()

    per_call()
    t0 = time.perf_counter()
    per_call()
    wall_pc = time.perf_counter() - t0
    res = {"requests": nreq, "packed_bytes": total,
           "create_ms": round(create_ms, 2),
           "batched_packs_per_s": round(nreq * reps / wall_b, 1),
           "per_call_sync_packs_per_s": round(nreq / wall_pc, 1),
           "batch_kernel_hbm_gbs": round(2 * total / (dev_b / reps) / 1e9, 1),
           "batch_kernel_us": round(dev_b * 1e6 / reps, 2),
           "batch_kernel_us_rule_variants": round(dev_r * 1e6 / reps, 2),
           "batch_kernel_us_selector_variants": round(dev_m * 1e6 / reps, 2),
           "bit_exact": exact}
    res["speedup"] = round(res["batched_packs_per_s"] / res["per_call_sync_packs_per_s"], 1)
    log(f"random requests: {res}")
    b.destroy()
    return res


def bench_vector(ta, m, tc, stream):
    """MPI_Type_vector(1024, 256, 512, MPI_BYTE): MPI_Pack wall time and
    kernel device time."""
    import numpy as np

    h = m.type_from_spec(("vector", 1024, 256, 512, tc.BYTE))
    sb, size = m.strided_block(h)
    src = ta.DeviceBuffer(524288)
    src.fill(9)
    out = ta.DeviceBuffer(size)
    ts = []
    for _ in range(200):
        t0 = time.perf_counter()
        m.pack(src.addr(), 1, h, out.addr(), size)
        ts.append(time.perf_counter() - t0)
    _, dev = timed_loop(ta, lambda: ta.pack_sb(sb, src.addr(), 1, out.addr(), size, 0, -1,
                                               stream), 200, stream)
    # the same pack into ordinary (pageable) host memory: staged on the
    # device, then copied across
    pageable = np.zeros(size, np.uint8)
    tp = []
    for _ in range(200):
        t0 = time.perf_counter()
        m.pack(src.addr(), 1, h, pageable.ctypes.data, size)
        tp.append(time.perf_counter() - t0)
    ok = bool((pageable == 9).all())
    m.free(h)
    return {"mpi_pack_wall_us": round(float(np.median(ts)) * 1e6, 2),
            "kernel_us": round(dev * 1e6 / 200, 3),
            "mpi_pack_into_pageable_host_wall_us": round(float(np.median(tp)) * 1e6, 2),
            "pageable_ok": ok,
            "packed_bytes": size}


def line_bound(sb, strided_mod, count=1):
    """Distinct 128-B lines a copy touches: (strided-side lines, packed-side
    lines).  `strided_mod` is the strided origin's address modulo 128 (the
    packed buffer starts a line).  Rows are visited in address order (the
    sweep's strides are positive and each dimension's stride exceeds the
    span inside it), so a line shared by two rows is shared by neighbours."""
    import numpy as np

    counts = [int(sb.counts[i]) for i in range(sb.ndims)]
    strides = [int(sb.strides[i]) for i in range(sb.ndims)]
    L = counts[0]
    starts = np.zeros(1, np.int64)
    for c, st in list(zip(counts[1:], strides[1:])) + [(count, int(sb.extent))]:
        starts = ((np.arange(c, dtype=np.int64) * st)[:, None] + starts[None, :]).reshape(-1)
    a = starts + int(sb.start) + strided_mod
    first, last = a // 128, (a + L - 1) // 128
    lines = int((last - first + 1).sum()) - int((first[1:] == last[:-1]).sum())
    packed = L * len(starts)
    return lines, (packed + 127) // 128


def bench_sweep(ta, m, tc, stream, ncases=48, max_span=8 << 30):
    """The 2D vector / 3D hvector-of-vector byte sweep (block lengths 1 B -
    1 KiB, stride = block + U{0..4096}, base offset U{0..15}, 1 KiB - 256 MiB
    packed; case k uses seed 1000 + k), committed through the interposer and
    packed and unpacked with the library's default choice.  Cases whose
    strided span exceeds `max_span` are skipped and counted.  Each copy is
    timed as the library times copies: the mean of 3 launches, back to back
    below 64 MiB, each after a cache sweep from 64 MiB up."""
    import random

    import numpy as np

    cases, skipped = [], 0
    for k in range(ncases):
        spec, base, packed_bytes = tc.sweep_type(random.Random(1000 + k))
        h = m.type_from_spec(spec)
        sb, size = m.strided_block(h)
        m.free(h)
        span = base + sb.extent
        if span > max_span:
            skipped += 1
            continue
        src = ta.DeviceBuffer(span + 64)
        dst = ta.DeviceBuffer(size + 64)
        row = {"block": int(sb.counts[0]) if sb.ndims else size, "packed_bytes": size}
        # the line-granular bound: 128-B lines touched on both sides, each
        # moved once at the HBM peak
        ls, lp = line_bound(sb, src.addr(base) % 128)
        row["lines"] = ls + lp
        row["line_bound_frac"] = round(2 * size / (128 * (ls + lp)), 4)
        for d, name in ((ta.DIR_PACK, "pack"), (ta.DIR_UNPACK, "unpack")):
            # one call settles the default choice, which is then timed the way
            # the library times copies (cold from 64 MiB up)
            if d == ta.DIR_PACK:
                ta.pack_sb(sb, src.addr(base), 1, dst.addr(), size, 0, -1, stream)
            else:
                ta.unpack_sb(sb, dst.addr(), size, src.addr(base), 1, 0, -1, stream)
            ta.synchronize(stream)
            v = ta.select_variant(sb, 1, src.addr(base), dst.addr(), d, ta.SELECT_DEFAULT)
            us = ta.time_variant(sb, 1, src.addr(base), dst.addr(), d, v, 0, 3)
            row[f"{name}_us"] = round(us, 2)
            row[f"{name}_frac"] = round(2 * size / us / 1e3 / HBM_PEAK_GBS, 4)
            row[f"{name}_of_line_bound"] = round(row[f"{name}_frac"] / row["line_bound_frac"], 4)
            row[f"{name}_variant"] = v
        src.free()
        dst.free()
        cases.append(row)
    res = {"cases": len(cases), "skipped_span_over_8GiB": skipped}
    for name in ("pack", "unpack"):
        tot_b = sum(c["packed_bytes"] for c in cases)
        tot_t = sum(c[f"{name}_us"] for c in cases)
        res[f"{name}_packed_gbs_all"] = round(tot_b / tot_t / 1e3, 1)
        big = [c for c in cases if c["block"] >= 64 and c["packed_bytes"] >= (16 << 20)]
        if big:
            fr = [c[f"{name}_frac"] for c in big]
            lb = [c[f"{name}_of_line_bound"] for c in big]
            res[f"{name}_hbm_frac_block_ge64_ge16MiB"] = {
                "n": len(big), "median": round(float(np.median(fr)), 4),
                "min": round(float(np.min(fr)), 4),
                "line_bound_median": round(float(np.median(
                    [c["line_bound_frac"] for c in big])), 4),
                "of_line_bound_median": round(float(np.median(lb)), 4),
                "of_line_bound_min": round(float(np.min(lb)), 4)}
    res["per_case"] = cases
    log(f"vector sweep: { {k: v for k, v in res.items() if k != 'per_case'} }")
    return res


def bench_p2p(ta, m, faces, grid, grid2, host, reps=10):
    """The three faces sent by this process to itself with MPI_Isend /
    MPI_Irecv (all six requests outstanding, then MPI_Waitall), per staging
    method; beside it the same exchange of host buffers through the MPI
    library alone (its own CPU pack and unpack)."""
    import numpy as np

    res = {}
    # 0: the library's own choice (the method model when it is installed)
    names = {2: "oneshot", 3: "staged", 0: "library_choice"}
    for method, name in names.items():
        ta.lib().tempi_set_datatype_method(method)
        ts = []
        for r in range(reps + 2):
            t0 = time.perf_counter()
            reqs = []
            for i, f in enumerate(faces):
                reqs.append(m.irecv(grid2.addr(), 1, f["handle"], 0, 200 + i)[1])
            for i, f in enumerate(faces):
                reqs.append(m.isend(grid.addr(), 1, f["handle"], 0, 200 + i)[1])
            err, _, _ = m.waitall(reqs)
            assert err == 0
            if r >= 2:
                ts.append(time.perf_counter() - t0)
        us = float(np.median(ts)) * 1e6
        res[name] = {"us_per_exchange": round(us, 1),
                     "packed_gbs": round(sum(f["size"] for f in faces) / us / 1e3, 2)}
        if method == 0:
            res[name]["method_per_face"] = [
                {1: "device", 2: "oneshot", 3: "staged"}.get(
                    ta.lib().tempi_p2p_method(1, f["size"], f["sb"].as_tuple()[2][0],
                                              f["sb"].as_tuple()[3][1]), "unknown")
                for f in faces]
    ta.lib().tempi_set_datatype_method(-1)
    dst = np.zeros_like(host)
    ts = []
    for r in range(4):
        t0 = time.perf_counter()
        reqs = []
        for i, f in enumerate(faces):
            reqs.append(m.irecv(dst.ctypes.data, 1, f["handle"], 0, 300 + i)[1])
        for i, f in enumerate(faces):
            reqs.append(m.isend(host.ctypes.data, 1, f["handle"], 0, 300 + i)[1])
        err, _, _ = m.waitall(reqs)
        assert err == 0
        if r >= 1:
            ts.append(time.perf_counter() - t0)
    us = float(np.median(ts)) * 1e6
    res["mpi_library_host_buffers"] = {"us_per_exchange": round(us, 1),
                                       "packed_gbs": round(sum(f["size"] for f in faces) / us / 1e3, 2)}
    log(f"isend/irecv of the faces: {res}")
    return res


def bench_neighbor(ta, m, tc, reps=20):
    """A periodic halo exchange of one 128^3 float64 subdomain (radius 3)
    whose 26 neighbours are all this process: MPI_Neighbor_alltoallw (one
    batched pack, the library's alltoallv, one batched unpack) per staging
    method, against the reference's scheme of one MPI_Isend / MPI_Irecv per
    neighbour through the same library."""
    import numpy as np

    dirs = tc.halo_directions()
    lcr = (128, 128, 128)
    st = [m.type_from_spec(tc.halo_type(3, lcr, d, 8)) for d in dirs]
    rt = [m.type_from_spec(tc.halo_type(3, lcr, tuple(-x for x in d), 8, exterior=True))
          for d in dirs]
    k = len(dirs)
    comm = m.dist_graph([0] * k, [0] * k)
    n = 134 ** 3 * 8
    src, dst = ta.DeviceBuffer(n), ta.DeviceBuffer(n)
    src.fill(1)
    dst.fill(2)
    packed = sum(m.strided_block(h)[1] for h in st)

    def timed(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e6

    def alltoallw():
        assert m.neighbor_alltoallw(src.addr(), [1] * k, [0] * k, st, dst.addr(), [1] * k,
                                    [0] * k, rt, comm) == 0

    def isend_irecv():
        reqs = [m.irecv(dst.addr(), 1, h, 0, 400 + i)[1] for i, h in enumerate(rt)]
        reqs += [m.isend(src.addr(), 1, h, 0, 400 + i)[1] for i, h in enumerate(st)]
        assert m.waitall(reqs)[0] == 0

    res = {"neighbours": k, "packed_bytes": packed}
    for method, name in ((2, "oneshot"), (3, "staged")):
        ta.lib().tempi_set_datatype_method(method)
        res[f"alltoallw_{name}_us"] = round(timed(alltoallw), 1)
        res[f"isend_irecv_{name}_us"] = round(timed(isend_irecv), 1)
    ta.lib().tempi_set_datatype_method(-1)
    m.comm_free(comm)
    for h in st + rt:
        m.free(h)
    log(f"neighbor alltoallw: {res}")
    return res


SWEEP_APP = os.path.join(ROOT, "tests", "c", "mpi_pack_sweep")


def run_sweep_app(args, preload, extra_env=None, timeout=600):
    """The native benchmark program (tests/c/mpi_pack_sweep.cpp) as a child
    process, with libtempi_amd.so preloaded in front of the MPI library or
    not; its JSON lines."""
    import tempi_amd as ta

    env = dict(os.environ)
    if preload:
        env["LD_PRELOAD"] = ta.LIB_PATH + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD")
                                           else "")
    env.update(extra_env or {})
    out = subprocess.run([SWEEP_APP, *args], env=env, capture_output=True, text=True,
                         timeout=timeout)
    if out.returncode != 0:
        raise RuntimeError(f"{SWEEP_APP} {' '.join(args)}: exit {out.returncode}: "
                           f"{out.stderr[-400:]}")
    return [json.loads(line) for line in out.stdout.splitlines() if line.startswith("{")]


def bench_reference_sweep():
    """The reference's MPI_Pack benchmark (bin/bench_mpi_pack.cpp:127-143):
    per-call MPI_Pack and MPI_Unpack wall time (trimean) of 2D byte types --
    packed 1 KiB, 1 MiB, 4 MiB at incount 1 and 2, blocks of 1-512 B at a
    512-B stride, as vector, hvector and subarray -- with device buffers and
    with the packed buffer in mapped host memory through this library, and
    with host buffers through the MPI library alone.  Throughput as the
    reference reports it: packed bytes / pack time."""
    import numpy as np

    if not os.path.exists(SWEEP_APP):
        return {"error": f"{SWEEP_APP} not built"}
    log("reference MPI_Pack sweep: device, mapped host, MPI library on host buffers")
    runs = {"device": run_sweep_app(["pack", "device", "100"], True),
            "mapped": run_sweep_app(["pack", "mapped", "100"], True),
            "mpi_library_host": run_sweep_app(["pack", "host", "5"], False)}
    rows = {}
    for mode, lines in runs.items():
        for r in lines:
            key = (r["count"], r["target"], r["block"])
            row = rows.setdefault(key, {"count": r["count"], "target": r["target"],
                                        "block": r["block"], "bytes": r["bytes"]})
            short = {"2d_byte_vector": "vector", "2d_byte_hvector": "hvector",
                     "2d_byte_subarray": "subarray"}[r["layout"]]
            row[f"{mode}_{short}_pack_us"] = round(r["pack_s"] * 1e6, 2)
            row[f"{mode}_{short}_unpack_us"] = round(r["unpack_s"] * 1e6, 2)
            row[f"{mode}_{short}_pack_mibs"] = (round(r["bytes"] / r["pack_s"] / 2 ** 20, 1)
                                                if r["pack_s"] > 0 else None)
    table = sorted(rows.values(), key=lambda r: (r["target"], r["count"], r["block"]))
    summary = {}
    for target in (1024, 1 << 20, 4 << 20):
        sel = [r for r in table if r["target"] == target]
        sp = [r[f"device_{k}_pack_mibs"] / r[f"mpi_library_host_{k}_pack_mibs"]
              for r in sel for k in ("vector", "hvector", "subarray")
              if r.get(f"device_{k}_pack_mibs") and r.get(f"mpi_library_host_{k}_pack_mibs")]
        dev = [r["bytes"] / (r[f"device_{k}_pack_us"] * 1e-6) / 1e9
               for r in sel for k in ("vector", "hvector", "subarray")
               if r.get(f"device_{k}_pack_us")]
        summary[f"packed_{target}B"] = {
            "device_pack_gbs_median": round(float(np.median(dev)), 2),
            "speedup_vs_mpi_library_geomean": round(float(np.exp(np.mean(np.log(sp)))), 3),
            "speedup_vs_mpi_library_min": round(float(np.min(sp)), 3),
            "speedup_vs_mpi_library_max": round(float(np.max(sp)), 1)}
    res = {"summary": summary, "rows": table,
           "note": "wall time of one synchronous call from a native MPI program; device and "
                   "mapped: this library preloaded; mpi_library_host: the MPI library alone "
                   "on host buffers (5 iterations, as the reference runs its library-only "
                   "case)"}
    log(f"reference MPI_Pack sweep: {summary}")
    return res


def bench_type_commit():
    """MPI_Type_commit wall time (trimean of 200) of the reference's commit
    benchmark types (bin/bench_type_commit.cpp:28-53): through this library
    with the learned selector evaluated at commit (the default), through it
    with the closed-form rule (canonicalization only), and the MPI library
    alone."""
    import numpy as np

    if not os.path.exists(SWEEP_APP):
        return {"error": f"{SWEEP_APP} not built"}
    log("type commit: model, rule, MPI library alone")
    runs = {"with_model": run_sweep_app(["commit", "200"], True),
            "canonicalize_only": run_sweep_app(["commit", "200"], True,
                                               {"TEMPI_SELECT": "heuristic"}),
            "mpi_library": run_sweep_app(["commit", "200"], False)}
    res = {}
    for k, lines in runs.items():
        us = [r["commit_s"] * 1e6 for r in lines]
        res[k] = {"cases": len(us), "median_us": round(float(np.median(us)), 2),
                  "max_us": round(float(np.max(us)), 2)}
        for layout in sorted({r["layout"] for r in lines}):
            res[k][f"{layout}_median_us"] = round(float(np.median(
                [r["commit_s"] * 1e6 for r in lines if r["layout"] == layout])), 2)
    log(f"type commit: { {k: v['median_us'] for k, v in res.items()} }")
    return res


def bench_first_call(ta, m, tc):
    """Wall time of the first synchronous MPI_Pack / MPI_Unpack of a
    256 MiB shape (64-B rows, 2 x 64 + 80 B apart) against later calls:
    under the default policy the first call of a copy of 64 MiB or more
    times the model's candidates once, cold."""
    import numpy as np

    # a stride no other copy of the run uses: the shape is new to the library
    rows, bl = 1 << 22, 64
    spec = ("vector", rows, bl, 2 * bl + 80, tc.BYTE)
    h = m.type_from_spec(spec)
    sb, size = m.strided_block(h)
    src = ta.DeviceBuffer(sb.extent + 64)
    dst = ta.DeviceBuffer(size)
    res = {"packed_bytes": size}
    for name, fn in (("pack", lambda: m.pack(src.addr(), 1, h, dst.addr(), size)),
                     ("unpack", lambda: m.unpack(dst.addr(), size, src.addr(), 1, h))):
        c0 = ta.counters()
        t0 = time.perf_counter()
        err, _ = fn()
        first = time.perf_counter() - t0
        assert err == 0
        later = []
        for _ in range(5):
            t0 = time.perf_counter()
            fn()
            later.append(time.perf_counter() - t0)
        c1 = ta.counters()
        res[name] = {"first_call_ms": round(first * 1e3, 3),
                     "later_call_ms_median": round(float(np.median(later)) * 1e3, 3),
                     "trial_launches": c1["trial_launches"] - c0["trial_launches"]}
    m.free(h)
    src.free()
    dst.free()
    log(f"first call of a 256 MiB shape: {res}")
    return res


def bench_selector_regret(ta, n=512, rounds=3):
    """The learned selector against the exhaustive autotuner on `n` held-out
    copies (the selector's held-out generator, seed 900000: 2D/3D vectors,
    halo pieces, grid faces and random blocks, a tenth of 64-256 MiB timed
    cold).  Every variant of every copy is timed twice, independently (each
    time the median of `rounds` timings): the autotuner picks from the first
    timing, and every choice -- model, autotuner, closed-form rule -- is
    judged on the second.  Regret = t(choice) / t(fastest) - 1."""
    import numpy as np

    from tempi_amd import selector_train as st

    if not ta.lib().tempi_mlp_loaded():
        return {"error": "no selector model loaded"}
    cases = st.generate(n, 900000, 32 << 20)
    src = ta.DeviceBuffer(st.SRC_BYTES)
    src.fill(1)
    dst = ta.DeviceBuffer(st.DST_BYTES + 4096)
    dst.fill(2)
    log(f"selector regret: timing {n} held-out copies x {ta.num_variants()} variants, twice")
    _, T1, H, B = st.time_cases(cases, src, dst, log_every=0, rounds=rounds)
    _, T2, _, _ = st.time_cases(cases, src, dst, log_every=0, rounds=rounds)
    model = np.array([ta.select_variant(sb, count, src.addr(base), dst.addr(ppos), d,
                                        ta.SELECT_MODEL)
                      for sb, base, d, ppos, count in cases], np.int32)
    auto = T1.argmin(1)
    idx = np.arange(len(cases))

    def stats(mask):
        best = T2[mask].min(1)
        out = {"copies": int(mask.sum())}
        for name, pick in (("model", model), ("autotuner", auto), ("rule", H)):
            t = T2[idx[mask], pick[mask]]
            r = t / best - 1
            out[name] = {"mean": round(float(r.mean()), 4),
                         "median": round(float(np.median(r)), 4),
                         "p90": round(float(np.quantile(r, 0.9)), 4),
                         "within_5pct": round(float((r <= 0.05).mean()), 4),
                         "time_weighted": round(float(t.sum() / best.sum() - 1), 4),
                         "packed_gbs": round(float(B[mask].sum() / (t.sum() * 1e-6) / 1e9), 1)}
        return out

    res = {"timings_per_label": rounds, "all": stats(idx >= 0),
           "ge_256KiB": stats(B >= (256 << 10)), "ge_64MiB_cold": stats(B >= st.BIG_MIN)}
    src.free()
    dst.free()
    log(f"selector regret: { {k: v for k, v in res['ge_256KiB'].items() if k != 'copies'} }")
    return res


def bench_p2p_small(ta, m, tc, rounds=300):
    """Round trips to self of small non-contiguous device messages with the
    ONESHOT method (medians of `rounds`): MPI_Isend + MPI_Irecv +
    MPI_Waitall, and a blocking MPI_Send into a posted MPI_Irecv.  Packs and
    unpacks this small return on their own completion signal
    (`signalled_waits`); the received bytes are checked once."""
    import numpy as np

    ta.lib().tempi_set_datatype_method(2)  # ONESHOT
    res = {}
    try:
        for name, spec in (("vector_300x12B", ("vector", 300, 3, 7, tc.FLOAT)),
                           ("vector_1024x256B", ("vector", 1024, 256, 512, tc.BYTE))):
            h = m.type_from_spec(spec)
            sb, size = m.strided_block(h)
            span = sb.extent + 64
            src, dst = ta.DeviceBuffer(span), ta.DeviceBuffer(span)
            host = np.arange(span, dtype=np.uint32).astype(np.uint8)
            src.upload(host)
            dst.fill(0xAB)

            def nonblocking():
                _, rr = m.irecv(dst.addr(), 1, h, 0, 41)
                _, sr = m.isend(src.addr(), 1, h, 0, 41)
                m.waitall([rr, sr], [h, h])

            def blocking():
                _, rr = m.irecv(dst.addr(), 1, h, 0, 42)
                m.send(src.addr(), 1, h, 0, 42)
                m.wait(rr, h)

            nonblocking()
            got = dst.download()
            sig0 = ta.counters()["signalled_waits"]
            row = {}
            for key, f in (("isend_irecv_waitall_us", nonblocking),
                           ("send_into_posted_irecv_us", blocking)):
                for _ in range(30):
                    f()
                ts = []
                for _ in range(rounds):
                    t0 = time.perf_counter()
                    f()
                    ts.append(time.perf_counter() - t0)
                row[key] = round(float(np.median(ts)) * 1e6, 2)
            row["signalled_completions_per_round"] = round(
                (ta.counters()["signalled_waits"] - sig0) / (2 * (rounds + 30)), 2)
            # the received bytes: the sent type map, 0xAB between its blocks
            ref = np.full(span, 0xAB, np.uint8)
            packed = np.zeros(size, np.uint8)
            err, _ = m.pack(host.ctypes.data, 1, h, packed.ctypes.data, size, library=True)
            err2, _ = m.unpack(packed.ctypes.data, size, ref.ctypes.data, 1, h, library=True)
            ok = err == 0 and err2 == 0 and bool((got == ref).all())
            row["bit_exact"] = ok
            res[name] = row
            m.free(h)
    finally:
        ta.lib().tempi_set_datatype_method(0)
    log(f"p2p small latency: {res}")
    return res


def bench_system_tables(ta):
    """Measure the method model's tables (kernel launch, d2h/h2d curves, the
    four 9x9 pack/unpack tables) and report a few of their points: 4 MiB in
    256-B blocks packed to HBM (DEVICE) and to mapped host memory (ONE_SHOT,
    bounded by the host link, not HBM)."""
    from tempi_amd import system as tsys

    tsys.clear()
    # each point: trials of 10 ms until the samples pass the SP 800-90B
    # permutation test, at most 3 (the library's default is 10)
    os.environ["TEMPI_MEASURE_TRIALS"] = "3"
    t0 = time.perf_counter()
    try:
        tsys.measure_performance(tsys.MEASURE_ALL, 0.01)
    finally:
        del os.environ["TEMPI_MEASURE_TRIALS"]
    el = time.perf_counter() - t0
    with tempfile.TemporaryDirectory() as d:
        tsys.export_performance(os.path.join(d, "perf.json"))
        perf = json.load(open(os.path.join(d, "perf.json")))
    flags = [e["iid"] for k in ("d2h", "h2d") for e in perf[k]]
    flags += [e["iid"] for k in tsys.TABLE_KEYS for row in perf[k] for e in row]
    dev = tsys.get_table(tsys.PACK_DEVICE)
    host = tsys.get_table(tsys.PACK_HOST)
    uhost = tsys.get_table(tsys.UNPACK_HOST)
    d2h = tsys.get_curve(tsys.D2H)
    mib4 = 1 << 22
    res = {
        "measure_seconds": round(el, 2),
        "entries_passing_iid_test": round(sum(flags) / len(flags), 3),
        "kernel_launch_us": round(tsys.kernel_launch() * 1e6, 3),
        "pack_device_4MiB_bl256_us": round(dev[8][8] * 1e6, 2),
        "pack_device_64B_bl1_us": round(dev[0][0] * 1e6, 2),
        "oneshot_pack_4MiB_bl256_gbs": round(mib4 / host[8][8] / 1e9, 2),
        "oneshot_unpack_4MiB_bl256_gbs": round(mib4 / uhost[8][8] / 1e9, 2),
        "d2h_8MiB_gbs": round((1 << 23) / d2h[23] / 1e9, 2),
    }
    tsys.clear()
    log(f"system tables: {res}")
    return res


def cpu_baselines(args, m, orc, faces, host, np):
    """The single-core legs, pinned to one core of this process's CPU set
    (restored afterwards): the MPI library's own MPI_Pack / MPI_Unpack on
    host buffers (the headline CPU figure), and the oracle's gather /
    scatter."""
    allowed = os.sched_getaffinity(0)
    core = min(allowed)
    os.sched_setaffinity(0, {core})
    try:
        return _cpu_baselines(args, m, orc, faces, host, np, core)
    finally:
        os.sched_setaffinity(0, allowed)


def _cpu_baselines(args, m, orc, faces, host, np, core):
    log(f"cpu baseline: MPI library MPI_Pack + MPI_Unpack, {args.cpu_seconds:.0f} s sample, "
        f"pinned to core {core}")
    scratch = host.copy()
    out = np.zeros(faces[0]["size"], np.uint8)
    moved, reps = 0, 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < args.cpu_seconds:
        for f in faces:
            m.pack(host.ctypes.data, 1, f["handle"], out.ctypes.data, f["size"], library=True)
            m.unpack(out.ctypes.data, f["size"], scratch.ctypes.data, 1, f["handle"],
                     library=True)
            moved += 2 * f["size"]
        reps += 1
    el = time.perf_counter() - t0
    lib = {"value": round(moved / el / 1e9, 3), "unit": "GB/s", "cores": 1,
           "kind": "reference", "cpu": host_cpu_model(),
           "implementation": f"{mpi_library_name(m)} MPI_Pack / MPI_Unpack on host buffers, "
                             "the reference's path for host data (src/pack.cpp:51-54)",
           "sample": f"pack + unpack of the 3 faces of the 518^3 float32 grid, {reps} steps "
                     f"in {el:.1f} s on 1 host thread pinned to core {core}"}
    log(f"cpu baseline: oracle gather/scatter, {args.cpu_seconds / 2:.0f} s sample")
    descs = [orc.sb_desc(*f["sb"].as_tuple()) for f in faces]
    moved, reps = 0, 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < args.cpu_seconds / 2:
        for d, f in zip(descs, faces):
            p = orc.sb_pack(d, host, 0, 1)
            orc.sb_unpack(d, p, scratch, 0, 1)
            moved += 2 * f["size"]
        reps += 1
    el = time.perf_counter() - t0
    port = {"value": round(moved / el / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "implementation": "the oracle's C gather / scatter (oracle/oracle.c sb_pack, "
                              "sb_unpack)",
            "sample": f"the 3 faces, {reps} steps in {el:.1f} s on 1 host thread pinned to "
                      f"core {core}"}
    return lib, port


def mpi_library_name(m):
    import ctypes

    buf = ctypes.create_string_buffer(8192)
    n = ctypes.c_int()
    try:
        m.libmpi.MPI_Get_library_version(buf, ctypes.byref(n))
        return buf.value.decode(errors="replace").splitlines()[0].strip()
    except Exception:
        return "the MPI library"


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_worker(spec):
    """One process of the all-cores CPU baseline: the MPI library alone (no
    libtempi_amd.so, no GPU), a singleton MPI_Init, the three face subarrays
    committed, its own copy of the grid; reports ready, waits for the go line
    on stdin, then runs MPI_Pack + MPI_Unpack of the faces for the given
    time and prints the bytes it moved."""
    import ctypes

    import numpy as np

    job = json.loads(spec)
    L = ctypes.CDLL(job["libmpi"], mode=ctypes.RTLD_GLOBAL)
    H = ctypes.c_int32 if job["handle_size"] == 4 else ctypes.c_int64
    I3 = ctypes.c_int * 3
    if L.MPI_Init(None, None) != 0:
        raise RuntimeError("MPI_Init failed")
    L.MPI_Type_create_subarray.argtypes = [ctypes.c_int, I3, I3, I3, ctypes.c_int, H,
                                           ctypes.POINTER(H)]
    L.MPI_Type_commit.argtypes = [ctypes.POINTER(H)]
    L.MPI_Pack.argtypes = [ctypes.c_void_p, ctypes.c_int, H, ctypes.c_void_p, ctypes.c_int,
                           ctypes.POINTER(ctypes.c_int), H]
    L.MPI_Unpack.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                             ctypes.c_void_p, ctypes.c_int, H, H]
    types = []
    for sizes, subsizes, starts in job["faces"]:
        t = H()
        L.MPI_Type_create_subarray(3, I3(*sizes), I3(*subsizes), I3(*starts),
                                   job["order_c"], H(job["mpi_float"]), ctypes.byref(t))
        L.MPI_Type_commit(ctypes.byref(t))
        types.append(t)
    # the bytes' values do not change a gather's cost
    host = np.ones(job["grid_bytes"], np.uint8)
    scratch = np.ones(job["grid_bytes"], np.uint8)
    out = np.zeros(max(job["sizes"]), np.uint8)
    comm = H(job["comm_world"])
    print("ready", flush=True)
    sys.stdin.readline()
    moved = 0
    pos = ctypes.c_int()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < job["seconds"]:
        for t, n in zip(types, job["sizes"]):
            pos.value = 0
            L.MPI_Pack(host.ctypes.data, 1, t, out.ctypes.data, n, ctypes.byref(pos), comm)
            pos.value = 0
            L.MPI_Unpack(out.ctypes.data, n, ctypes.byref(pos), scratch.ctypes.data, 1, t,
                         comm)
            moved += 2 * n
    el = time.perf_counter() - t0
    L.MPI_Finalize()
    print(json.dumps({"moved": moved, "seconds": el}), flush=True)


def cpu_all_cores(m, faces, grid_bytes, seconds):
    """MPI_Pack + MPI_Unpack of the faces by the MPI library in one
    single-rank process per core of this host's CPU share (at most 16), all
    timed over the same window; the sum of their rates."""
    nproc = min(16, len(os.sched_getaffinity(0)))
    log(f"cpu baseline: the MPI library in {nproc} processes, {seconds:.0f} s sample")
    n, g = 512, 3
    full = n + 2 * g
    specs = {"x": [n, n, g], "y": [n, g, n], "z": [g, n, n]}
    job = json.dumps({
        "libmpi": m.libmpi._name, "handle_size": ta_handle_size(m),
        "mpi_float": int(m.const("MPI_FLOAT")), "order_c": int(m.ORDER_C),
        "comm_world": int(m.COMM_WORLD), "grid_bytes": int(grid_bytes), "seconds": seconds,
        "faces": [[[full] * 3, specs[f["face"]], [g] * 3] for f in faces],
        "sizes": [f["size"] for f in faces]})
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", job],
                              stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)
             for _ in range(nproc)]
    try:
        for p in procs:
            if p.stdout.readline().strip() != "ready":
                raise RuntimeError("a CPU baseline worker did not start")
        for p in procs:
            p.stdin.write("go\n")
            p.stdin.flush()
        rates = []
        for p in procs:
            r = json.loads(p.stdout.readline())
            rates.append(r["moved"] / r["seconds"])
        for p in procs:
            p.wait(timeout=60)
    except Exception as e:  # reported, not fatal: the GPU figures stand alone
        log(f"cpu all-cores baseline failed: {e}")
        return {"error": str(e)}
    finally:
        for p in procs:
            if p.poll() is None:
                p.kill()
                p.wait()
    return {"value": round(sum(rates) / 1e9, 3), "unit": "GB/s", "cores": nproc,
            "kind": "reference", "cpu": host_cpu_model(),
            "implementation": f"{mpi_library_name(m)} MPI_Pack / MPI_Unpack on host buffers",
            "sample": f"pack + unpack of the 3 faces in {nproc} concurrent single-rank "
                      f"processes (MPI library only), each on its own copy of the grid, "
                      f"{seconds:.0f} s window"}


def ta_handle_size(m):
    import ctypes

    return ctypes.sizeof(m.H)


if __name__ == "__main__":
    main()
