#!/usr/bin/env python3
"""
bench.py -- conservative regridding on one MI355X.

A step is one pass of the hot path over one grid pair: the conservative 2-D
weight build (sweep + coalesce, both CSR layouts) followed by the FP64 apply
of the pair's value stack.  The workload is the plain pair at real size:
a 2049x2049 warped input grid onto a 1537x1537 rotated output grid, 16
images of 2048x2048 Poisson counts, FP64.  Inputs (perturbed coordinates and
values) are resident in HBM before the timed region; the timed region is K
steps bracketed by device synchronisation and timed with HIP events on the
stream the kernels run on.

Printed: one JSON line with the whole-job throughput (grid pairs per second)
plus the apply roofline, the build phases, the FP32 apply and the CPU
baseline (the C oracle run over the same full pair on the host cores).

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "grid-pairs/s build+apply vs plain path, apply HBM GB/s; p50/p99 serve latency"
HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_args():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--apply-reps", type=int, default=20)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--small", action="store_true",
                   help="129x129 -> 101x101, 4 images (the CPU-runnable pair)")
    p.add_argument("--edges-per-lane", type=int, default=0)
    p.add_argument("--block", type=int, default=0)
    p.add_argument("--apply-sweep", action="store_true",
                   help="time the apply over tile widths x images per pass")
    p.add_argument("--no-batched", action="store_true",
                   help="skip the batched-vs-plain comparison over 256 mixed pairs")
    p.add_argument("--batched-pairs", type=int, default=256)
    return p.parse_args()


def apply_bytes(n_in: int, n_out: int, nnz: int, batch: int, s: int) -> int:
    """Algorithmic HBM bytes of one apply launch: values read once, outputs
    written once, every weight (value + int32 column) and the row pointers
    read once."""
    return batch * (n_in + n_out) * s + nnz * (s + 4) + (n_out + 1) * 4


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: the build and apply of a grid pair run on a "
            "single MI355X (multi-GPU sharding is described in DESIGN.md, not run)",
            file=sys.stderr,
        )
        return 2

    from regridding_amd import _util, workloads
    from regridding_amd._conservative import build_2d_device
    from regridding_amd._lib import DeviceArray, EventTimer, check, lib

    if args.small:
        gi, go, values = workloads.plain_small()
        workload = "plain_small: 129x129 warped input (0.3 rad, warp 0.05) -> 101x101 rotated output (0.1 rad), 4 FP64 images"
    else:
        gi, go, values = workloads.plain_large()
        workload = ("plain_large: 2049x2049 warped input (0.3 rad, seeded warp) -> 1537x1537 rotated "
                    "output over [-0.9,0.9] (0.1 rad), 16 FP64 images of 2048x2048 Poisson counts")
    ci, co, *_ = _util.normalize_coordinates(gi, go, perturb=True)
    B = values.shape[0]
    n_in = (ci[0].shape[0] - 1) * (ci[0].shape[1] - 1)
    n_out = (co[0].shape[0] - 1) * (co[0].shape[1] - 1)
    log(f"workload {workload}; n_in={n_in} n_out={n_out} batch={B}")

    L = lib()
    xi, yi = DeviceArray.from_host(ci[0]), DeviceArray.from_host(ci[1])
    xo, yo = DeviceArray.from_host(co[0]), DeviceArray.from_host(co[1])
    vals64 = np.ascontiguousarray(values.reshape(B, n_in))
    dx64 = DeviceArray.from_host(vals64)
    dx32 = DeviceArray.from_host(vals64.astype(np.float32))
    dy64 = DeviceArray.empty((B, n_out), np.float64)
    dy32 = DeviceArray.empty((B, n_out), np.float32)
    dt64 = DeviceArray.empty((B, n_in), np.float64)
    check(L.rg_device_sync())
    tuned = None
    if args.edges_per_lane or args.block:
        tune = dict(edges_per_lane=args.edges_per_lane, block=args.block)
    else:
        # the library's autotuner picks the launch shape for this size class
        from regridding_amd import _autotune
        from regridding_amd._conservative import pack_pairs

        tuned = _autotune.tune(pack_pairs([(ci, co)]), reps=2)
        tune = dict(edges_per_lane=tuned["edges_per_lane"], block=tuned["block"])
        log(f"autotuned build: edges_per_lane={tune['edges_per_lane']} block={tune['block']} "
            f"({tuned['ms']:.3f} ms)")

    def build(timings=False):
        return build_2d_device(xi, yi, ci[0].shape, xo, yo, co[0].shape, None,
                               timings=timings, **tune)

    def step():
        h = build()
        check(L.rg_apply_f64(h.ptr, 0, dx64.ptr, dy64.ptr, B, 0))
        return h

    for w in range(args.warmup):
        step().free()
        check(L.rg_device_sync())
        log(f"warmup {w + 1}/{args.warmup} done")

    timer = EventTimer(0)
    check(L.rg_device_sync())
    wall0 = time.perf_counter()
    timer.start()
    for k in range(args.steps):
        step().free()
    total_ms = timer.stop()
    check(L.rg_device_sync())
    wall_ms = (time.perf_counter() - wall0) * 1e3
    ms_per_step = total_ms / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step (events), wall {wall_ms / args.steps:.3f} ms/step")

    # ---- build phases (events inside the library)
    h = build(timings=True)
    st = h.stats()
    nnz = h.nnz
    log(f"build: total {st['ms_total']:.3f} ms, triples {st['n_triples']}, nnz {nnz}, "
        f"fixups {st['n_fixups']}, restaged {st['n_restaged']}; phases "
        + " ".join(f"{k[3:]}={v:.3f}" for k, v in st.items() if k.startswith('ms_')))
    build_timer = EventTimer(0)
    build_timer.start()
    for _ in range(3):
        build().free()
    build_ms = build_timer.stop() / 3

    # ---- apply kernels (events around repeated launches on stream 0)
    def time_apply(fn, x, y, transpose):
        fn(h.ptr, transpose, x.ptr, y.ptr, B, 0)
        t = EventTimer(0)
        t.start()
        for _ in range(args.apply_reps):
            fn(h.ptr, transpose, x.ptr, y.ptr, B, 0)
        return t.stop() / args.apply_reps

    # plain streaming copy of the input stack: the achievable-bandwidth
    # reference and the known-size kernel for the traffic counters
    n_copy = B * n_in
    dcopy = DeviceArray.empty((n_copy,), np.float64)
    check(L.rg_copy_f64(dx64.ptr, dcopy.ptr, n_copy, 0))
    tc = EventTimer(0)
    tc.start()
    for _ in range(args.apply_reps):
        check(L.rg_copy_f64(dx64.ptr, dcopy.ptr, n_copy, 0))
    ms_copy = tc.stop() / args.apply_reps
    copy_gbps = 2 * n_copy * 8 / (ms_copy * 1e-3) / 1e9
    dcopy.free()
    log(f"copy {2 * n_copy * 8 / 1e9:.3f} GB in {ms_copy:.4f} ms = {copy_gbps:.0f} GB/s")

    sweep = []
    if args.apply_sweep:
        shapes = [(tw, ipp, kib) for tw in (-16, 64, 32, 16, 8) for ipp in (16, 8, 4)
                  for kib in (32, 64)]
        for tw, ipp, kib in shapes:
            if tw < 0 and kib == 64:
                continue
            check(L.rg_set_apply_tuning(tw, ipp, kib))
            ms = time_apply(L.rg_apply_f64, dx64, dy64, 0)
            ms_t = time_apply(L.rg_apply_f64, dy64, dt64, 1)
            sweep.append({"tile_w": tw, "images_per_pass": ipp, "lds_kib": kib,
                          "ms": round(ms, 5), "adjoint_ms": round(ms_t, 5)})
            log(f"apply sweep tile_w={tw} images/pass={ipp} lds={kib}KiB: "
                f"{ms:.4f} ms, adjoint {ms_t:.4f} ms")
        best = min(sweep, key=lambda r: r["ms"])
        check(L.rg_set_apply_tuning(best["tile_w"], best["images_per_pass"], best["lds_kib"]))
        log(f"apply sweep best: {best}")

    ms_apply64 = time_apply(L.rg_apply_f64, dx64, dy64, 0)
    ms_apply32 = time_apply(L.rg_apply_f32, dx32, dy32, 0)
    ms_adjoint64 = time_apply(L.rg_apply_f64, dy64, dt64, 1)
    by64 = apply_bytes(n_in, n_out, nnz, B, 8)
    by32 = apply_bytes(n_in, n_out, nnz, B, 4)
    byt64 = apply_bytes(n_out, n_in, nnz, B, 8)
    gbps64 = by64 / (ms_apply64 * 1e-3) / 1e9
    gbps32 = by32 / (ms_apply32 * 1e-3) / 1e9
    gbpst64 = byt64 / (ms_adjoint64 * 1e-3) / 1e9
    log(f"apply f64 {ms_apply64:.4f} ms = {gbps64:.0f} GB/s; f32 {ms_apply32:.4f} ms = {gbps32:.0f} GB/s; "
        f"adjoint f64 {ms_adjoint64:.4f} ms = {gbpst64:.0f} GB/s")

    result = {
        "metric": METRIC,
        "value": round(1e3 / ms_per_step, 4),
        "unit": "grid-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "n_cells_input": n_in,
            "n_cells_output": n_out,
            "images": B,
            "nnz": nnz,
            "triples": st["n_triples"],
            "edges_per_lane": tune["edges_per_lane"] or 16,
            "block": tune["block"] or 128,
        },
        "roofline": {
            "kernel": "k_apply<double,16> (CSR gather apply, FP64, 16 images)",
            "bound": "hbm",
            "achieved": round(gbps64, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(gbps64 / HBM_PEAK_GBPS, 4),
            "traffic": None,
            "bytes_per_launch": by64,
            "ms_per_launch": round(ms_apply64, 5),
        },
        "build_ms": round(build_ms, 4),
        **({"build_autotune": tuned} if tuned else {}),
        "build_phases_ms": {k[3:]: round(v, 4) for k, v in st.items() if k.startswith("ms_")},
        "build_fixups": st["n_fixups"],
        "build_restaged": st["n_restaged"],
        "apply_f64": {"ms": round(ms_apply64, 5), "GBps": round(gbps64, 1), "frac_hbm": round(gbps64 / HBM_PEAK_GBPS, 4)},
        "apply_f32": {"ms": round(ms_apply32, 5), "GBps": round(gbps32, 1), "frac_hbm": round(gbps32 / HBM_PEAK_GBPS, 4)},
        "adjoint_f64": {"ms": round(ms_adjoint64, 5), "GBps": round(gbpst64, 1), "frac_hbm": round(gbpst64 / HBM_PEAK_GBPS, 4)},
        "wall_ms_per_step": round(wall_ms / args.steps, 4),
        "copy_GBps": round(copy_gbps, 1),
        **({"apply_sweep": sweep} if sweep else {}),
    }

    h.free()
    del dx32, dy32, dt64
    if not args.no_batched:
        result["batched"] = bench_batched(args)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(ci, co, values, build())
    print(json.dumps(result), flush=True)
    return 0


def bench_batched(args) -> dict:
    """Many uneven grid pairs (per-axis vertex counts log-uniform in
    [33, 1025], 1-8 FP64 value arrays each), build + apply, two ways on the
    same device-resident inputs: the plain path (one rg_build_conservative_2d
    and one rg_apply_f64 call per pair, in sequence) and the batched path
    (one rg_build_conservative_2d_batched and one rg_apply_batched_f64 call
    for all pairs).  Both produce the same bits; the ratio of their
    grid-pairs/s is reported."""
    import ctypes

    from regridding_amd import _util, workloads
    from regridding_amd._conservative import build_2d_batched, pack_pairs
    from regridding_amd._lib import DeviceArray, EventTimer, Weights, check, lib, rg_apply_job

    L = lib()
    n_pairs = args.batched_pairs
    t0 = time.perf_counter()
    raw = workloads.batched_pairs(n_pairs=n_pairs)
    items, vals = [], []
    for p in raw:
        ci, co, *_ = _util.normalize_coordinates(p["grid_input"], p["grid_output"], perturb=True)
        items.append((ci, co))
        vals.append(np.ascontiguousarray(p["values"].reshape(p["values"].shape[0], -1)))
    del raw
    packed = pack_pairs(items)
    n_in = [(ci[0].shape[0] - 1) * (ci[0].shape[1] - 1) for ci, _ in items]
    n_out = [(co[0].shape[0] - 1) * (co[0].shape[1] - 1) for _, co in items]
    nb = [v.shape[0] for v in vals]
    x_off = np.concatenate([[0], np.cumsum([v.size for v in vals])])
    y_off = np.concatenate([[0], np.cumsum([b * n for b, n in zip(nb, n_out)])])
    dX = DeviceArray.from_host(np.concatenate([v.ravel() for v in vals]))
    dYp = DeviceArray.empty((int(y_off[-1]),), np.float64)
    dYb = DeviceArray.empty((int(y_off[-1]),), np.float64)
    del vals
    check(L.rg_device_sync())
    log(f"batched workload: {n_pairs} pairs, {sum(n_in)} input cells, {sum(n_out)} output "
        f"cells, {sum(nb)} value arrays (prepared in {time.perf_counter() - t0:.1f} s)")

    def plain():
        for k in range(n_pairs):
            d = packed.desc[k]
            h = ctypes.c_void_p()
            check(L.rg_build_conservative_2d(d.xi, d.yi, d.nxi, d.nyi, d.xo, d.yo, d.nxo, d.nyo,
                                             None, None, ctypes.byref(h), 0))
            check(L.rg_apply_f64(h, 0, dX.ptr + 8 * int(x_off[k]), dYp.ptr + 8 * int(y_off[k]),
                                 nb[k], 0))
            check(L.rg_weights_free(h))

    jobs = (rg_apply_job * n_pairs)()
    from regridding_amd import _autotune

    t1 = time.perf_counter()
    tuned = _autotune.tune(packed, reps=1)
    log(f"autotuned batch build: edges_per_lane={tuned['edges_per_lane']} block={tuned['block']} "
        f"({tuned['ms']:.2f} ms; tuning took {time.perf_counter() - t1:.1f} s)")

    def batched(timings=False):
        hs = build_2d_batched(packed, timings=timings, edges_per_lane=tuned["edges_per_lane"],
                              block=tuned["block"])
        for k, h in enumerate(hs):
            j = jobs[k]
            j.weights, j.transpose, j.batch = h.ptr, 0, nb[k]
            j.values_input = dX.ptr + 8 * int(x_off[k])
            j.values_output = dYb.ptr + 8 * int(y_off[k])
        check(L.rg_apply_batched_f64(jobs, n_pairs, 0))
        return hs

    def timed(fn, reps):
        fn()
        check(L.rg_device_sync())
        t = EventTimer(0)
        t.start()
        for _ in range(reps):
            r = fn()
            if r is not None:
                for h in r:
                    h.free()
        ms = t.stop() / reps
        check(L.rg_device_sync())
        return ms

    ms_plain = timed(plain, 2)
    ms_batched = timed(batched, 3)
    hs = batched(timings=True)
    st = hs[0].stats()
    triples = sum(h.stats()["n_triples"] for h in hs)
    nnz = sum(h.nnz for h in hs)
    for h in hs:
        h.free()
    same = bool(np.array_equal(dYp.to_host(), dYb.to_host()))
    ratio = ms_plain / ms_batched
    log(f"batched: plain {ms_plain:.2f} ms ({n_pairs / ms_plain * 1e3:.0f} pairs/s), batched "
        f"{ms_batched:.2f} ms ({n_pairs / ms_batched * 1e3:.0f} pairs/s), x{ratio:.2f}; "
        f"outputs identical: {same}; batch phases "
        + " ".join(f"{k[3:]}={v:.2f}" for k, v in st.items() if k.startswith("ms_")))
    return {
        "workload": f"{n_pairs} independent pairs, per-axis vertices log-uniform in [33,1025], "
                    "spacing ratio U[0.5,2], rotations U(0,pi/4), 1-8 FP64 Poisson arrays each",
        "pairs": n_pairs,
        "input_cells": int(sum(n_in)),
        "output_cells": int(sum(n_out)),
        "value_arrays": int(sum(nb)),
        "triples": int(triples),
        "nnz": int(nnz),
        "plain_ms": round(ms_plain, 3),
        "plain_pairs_per_s": round(n_pairs / ms_plain * 1e3, 1),
        "batched_ms": round(ms_batched, 3),
        "batched_pairs_per_s": round(n_pairs / ms_batched * 1e3, 1),
        "batched_vs_plain": round(ratio, 3),
        "outputs_bit_identical": same,
        "batch_build_phases_ms": {k[3:]: round(v, 3) for k, v in st.items() if k.startswith("ms_")},
        "batch_autotune": tuned,
    }


def cpu_baseline(ci, co, values, handle) -> dict:
    """The C oracle (the reference's algorithm restated, same row-parallel
    sweep, single-threaded coalesce, element-parallel apply) over the same
    full pair on the host cores; its weights are also compared with the GPU's
    bit for bit."""
    from oracle import oracle

    cores = min(16, len(os.sched_getaffinity(0)))
    oracle.set_threads(cores)
    B = values.shape[0]
    t0 = time.perf_counter()
    flat = oracle.sweep_2d(ci, co)
    t1 = time.perf_counter()
    ii, io, w = oracle.coalesce(*flat)
    t2 = time.perf_counter()
    del flat
    y = oracle.apply(ii, io, w, values.reshape(B, -1), handle.n_out)
    t3 = time.perf_counter()
    gi, go, gw = handle.export()
    parity = bool(np.array_equal(gi, ii) and np.array_equal(go, io) and np.array_equal(gw, w))
    log(f"cpu oracle ({cores} threads): sweep {t1 - t0:.2f} s, coalesce {t2 - t1:.2f} s, "
        f"apply {t3 - t2:.2f} s; GPU weights bit-identical: {parity}")
    del y
    return {
        "value": round(1.0 / (t3 - t0), 6),
        "unit": "grid-pairs/s",
        "cores": cores,
        "kind": "port",
        "sample": "the full pair: sweep (row-parallel), coalesce (1 thread), FP64 apply of all images "
                  "(image-parallel), one run",
        "seconds": {"sweep": round(t1 - t0, 3), "coalesce": round(t2 - t1, 3), "apply": round(t3 - t2, 3)},
        "gpu_weights_bit_identical": parity,
    }


if __name__ == "__main__":
    sys.exit(main())
