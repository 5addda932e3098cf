"""
The serving loop: every request's result is the same alone and in a mixed,
concurrent load (bit for bit: the batched build and apply reproduce the
single-pair kernels exactly), cache hits skip builds, admission under a tight
memory budget still completes everything, and a failing request fails alone.
"""

import threading

import numpy as np
import pytest

from regridding_amd import _util, workloads
from regridding_amd._conservative import build_2d
from regridding_amd._lib import DeviceArray, check, lib
from regridding_amd.serve import Server, grid_key, pinned_empty

pytestmark = pytest.mark.gpu


def _pairs(n, seed):
    out = []
    for p in workloads.batched_pairs(n_pairs=n, seed=seed, n_min=9, n_max=120):
        ci, co, *_ = _util.normalize_coordinates(p["grid_input"], p["grid_output"], perturb=True)
        out.append((ci, co))
    return out


def _reference(ci, co, x, transpose):
    h = build_2d(ci, co)
    D = x.shape[0]
    n_dst = h.n_in if transpose else h.n_out
    dx = DeviceArray.from_host(x)
    dy = DeviceArray.empty((D, n_dst), x.dtype)
    fn = lib().rg_apply_f32 if x.dtype == np.float32 else lib().rg_apply_f64
    check(fn(h.ptr, int(transpose), dx.ptr, dy.ptr, D, 0))
    return dy.to_host()


def _requests(pairs, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        j = int(rng.integers(len(pairs)))
        ci, co = pairs[j]
        transpose = bool(rng.integers(0, 4) == 0)
        n_in = (ci[0].shape[0] - 1) * (ci[0].shape[1] - 1)
        n_out = (co[0].shape[0] - 1) * (co[0].shape[1] - 1)
        D = int(rng.integers(1, 6))
        dt = np.float32 if rng.integers(0, 2) else np.float64
        x = rng.poisson(30, (D, n_out if transpose else n_in)).astype(dt)
        cached = bool(rng.integers(0, 3))  # about two thirds by key
        reqs.append((j, x, transpose, cached))
    return reqs


def test_same_result_alone_and_in_the_mix():
    pairs = _pairs(10, 41)
    keys = [grid_key(ci, co) for ci, co in pairs]
    reqs = _requests(pairs, 120, 5)
    alone = {}
    with Server(workers=1, max_group=1) as srv:  # one request at a time
        for k, (j, x, t, cached) in enumerate(reqs):
            ci, co = pairs[j]
            tk = srv.submit(x, ci, co, key=keys[j] if cached else 0, transpose=t)
            alone[k] = srv.wait(tk)[0].copy()
    for k in (0, 7, 33):  # the alone results are the kernels' own
        j, x, t, _ = reqs[k]
        assert np.array_equal(alone[k], _reference(*pairs[j], x, t))
    with Server(workers=3, max_group=16) as srv:
        tickets = {}

        def producer(ks):
            for k in ks:
                j, x, t, cached = reqs[k]
                ci, co = pairs[j]
                tickets[k] = srv.submit(x, ci, co, key=keys[j] if cached else 0, transpose=t)

        th = [threading.Thread(target=producer, args=(range(i, len(reqs), 2),)) for i in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for k in range(len(reqs)):
            res, _ = srv.wait(tickets[k])
            assert np.array_equal(res, alone[k]), k
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    n_cached_keys = len({reqs[k][0] for k in range(len(reqs)) if reqs[k][3]})
    n_uncached = sum(1 for r in reqs if not r[3])
    assert st["builds"] == n_cached_keys + n_uncached
    assert st["groups"] < len(reqs)  # requests were batched


def test_cached_key_without_grids_and_tight_admission():
    import ctypes

    pairs = _pairs(6, 43)
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().rg_mem_info(ctypes.byref(free), ctypes.byref(total)))
    cache = 256 << 20
    # an admission budget of about 8 MiB: most groups hold one request
    reserve = max(int(free.value) - cache - (8 << 20), 1)
    refs = []
    with Server(workers=2, max_group=8, reserve_bytes=reserve, cache_bytes=cache) as srv:
        tks = []
        for j, (ci, co) in enumerate(pairs):
            x = pinned_empty((3, (ci[0].shape[0] - 1) * (ci[0].shape[1] - 1)), np.float64)
            x[:] = np.random.default_rng(j).normal(size=x.shape)
            tks.append((x, srv.submit(x, ci, co, key=grid_key(ci, co))))
        for j, (x, tk) in enumerate(tks):
            res = srv.wait(tk)[0].copy()
            assert np.array_equal(res, _reference(*pairs[j], np.asarray(x), False))
            refs.append((x, res))
        # by key only: the weights are cached
        for j, (ci, co) in enumerate(pairs):
            x, ref = refs[j]
            tk = srv.submit(x, key=grid_key(ci, co), shapes=(ci[0].shape, co[0].shape))
            assert np.array_equal(srv.wait(tk)[0], ref)
        st = srv.stats()
    assert st["builds"] == len(pairs) and st["cache_hits"] == len(pairs)
    assert st["peak_admitted_bytes"] > 0


def test_failing_request_fails_alone():
    pairs = _pairs(2, 47)
    with Server(workers=1) as srv:
        ci, co = pairs[0]
        x = np.ones((1, (ci[0].shape[0] - 1) * (ci[0].shape[1] - 1)))
        bad = srv.submit(x, key=12345, shapes=(ci[0].shape, co[0].shape))  # nothing cached
        with pytest.raises(RuntimeError, match="needs its grids"):
            srv.wait(bad)
        good = srv.submit(x, ci, co)
        assert np.array_equal(srv.wait(good)[0], _reference(ci, co, x, False))


def test_cached_key_with_wrong_shapes_fails_alone():
    """A request whose shapes disagree with the cached weights of its key
    fails with an argument error (its buffers are sized from its shapes);
    the requests grouped with it complete."""
    from regridding_amd._lib import RGError

    pairs = _pairs(2, 53)
    (ci, co), (cj, cq) = pairs
    key = grid_key(ci, co)
    with Server(workers=1, max_group=8) as srv:
        x = np.ones((2, (ci[0].shape[0] - 1) * (ci[0].shape[1] - 1)))
        first = srv.submit(x, ci, co, key=key)
        ref = srv.wait(first)[0].copy()
        # wrong shapes for the key: the other pair's shapes
        bad_shapes = (cj[0].shape, cq[0].shape)
        xb_ = np.ones((1, (cj[0].shape[0] - 1) * (cj[0].shape[1] - 1)))
        bad = srv.submit(xb_, key=key, shapes=bad_shapes)
        good = srv.submit(x, key=key, shapes=(ci[0].shape, co[0].shape))
        with pytest.raises(RGError, match="cached weights of this key"):
            srv.wait(bad)
        assert np.array_equal(srv.wait(good)[0], ref)
        st = srv.stats()
    assert st["failed"] == 1


def test_dropped_ticket_keeps_buffers_alive():
    """A ticket dropped without wait(): the server holds the request's host
    buffers until it has finished with them."""
    import gc

    pairs = _pairs(1, 59)
    ci, co = pairs[0]
    with Server(workers=1) as srv:
        x = np.random.default_rng(0).random((3, (ci[0].shape[0] - 1) * (ci[0].shape[1] - 1)))
        srv.submit(x.copy(), ci, co)  # ticket and values dropped at once
        gc.collect()
        tk = srv.submit(x, ci, co)
        res = srv.wait(tk)[0]
    assert np.array_equal(res, _reference(ci, co, x, False))
