"""
Batched conservative build (rg_build_conservative_2d_batched): every pair of
a mixed batch must come out bit-identical to its own single-pair build and
to the CPU oracle, whatever else is in the batch.
"""

import numpy as np
import pytest

from oracle import oracle
from regridding_amd import _util, workloads
from regridding_amd._conservative import build_2d, build_2d_batched, pack_pairs
from regridding_amd._lib import DeviceArray, check, lib

pytestmark = pytest.mark.gpu


def _perturbed(grid_in, grid_out, perturb=True, seed=42):
    ci, co, *_ = _util.normalize_coordinates(grid_in, grid_out, perturb=perturb, seed=seed)
    return ci, co


def _mixed_pairs():
    pairs = []
    # shapes from the single-pair parity cases, a minimal 2x2 pair, a pair
    # whose grids do not overlap, and small pairs of the batched workload
    pairs.append(_perturbed(workloads.warped_grid(17, 0.2, 0.03, 0.04),
                            workloads.warped_grid(61, 0.45, 0.0, 0.0, -0.9, 0.9)))
    pairs.append(_perturbed(workloads.warped_grid(90, 0.7, 0.06, 0.02),
                            workloads.warped_grid(12, 0.05, 0.0, 0.0, -0.8, 0.8)))
    pairs.append(_perturbed(workloads.warped_grid((2, 2), 0.1, 0.0, 0.0),
                            workloads.warped_grid((3, 2), 0.2, 0.0, 0.0, -0.7, 0.7)))
    pairs.append(_perturbed(workloads.warped_grid((9, 7), 0.0, 0.0, 0.0, -1.0, -0.5),
                            workloads.warped_grid((8, 8), 0.0, 0.0, 0.0, 0.5, 1.0)))
    pairs.append(_perturbed(workloads.warped_grid((40, 31), 0.3, 0.05, 0.05),
                            workloads.warped_grid((27, 35), -0.2, 0.01, 0.0, 0.2, 1.6)))
    for p in workloads.batched_pairs(n_pairs=6, seed=31, n_min=5, n_max=70):
        pairs.append(_perturbed(p["grid_input"], p["grid_output"]))
    return pairs


def test_batch_matches_single_and_oracle():
    pairs = _mixed_pairs()
    rng = np.random.default_rng(8)
    items = []
    for k, (ci, co) in enumerate(pairs):
        wi = None
        if k % 3 == 1:  # a mix of pairs with and without an input weighting
            wi = rng.uniform(0.5, 2.0, (ci[0].shape[0] - 1, ci[0].shape[1] - 1))
        items.append((ci, co, wi))
    hs = build_2d_batched(items)
    assert len(hs) == len(items)
    for k, ((ci, co, wi), h) in enumerate(zip(items, hs)):
        print(f"pair {k}: {ci[0].shape} -> {co[0].shape}", flush=True)
        got = h.export()
        ref = oracle.weights_2d(ci, co, wi)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        single = build_2d(ci, co, wi)
        assert h.stats()["n_triples"] == single.stats()["n_triples"]
        assert h.stats()["n_chunks"] == single.stats()["n_chunks"]


def test_batch_row_starts_many_rows_per_wave():
    """The row starts with 64 rows per wavefront (a test hook: a batch this
    small takes one row per wavefront), so a wavefront's rows span passes
    and pairs: every pair still matches its own single-pair build."""
    pairs = _mixed_pairs()
    hs = build_2d_batched(pairs, flags=4)
    for (ci, co), h in zip(pairs, hs):
        for a, b in zip(h.export(), build_2d(ci, co).export()):
            assert np.array_equal(a, b)


def test_batch_apply_and_adjoint():
    pairs = _mixed_pairs()
    hs = build_2d_batched(pairs)
    for k, ((ci, co), h) in enumerate(zip(pairs, hs)):
        ii, io, w = h.export()
        x = np.random.default_rng(k).normal(size=(3, h.n_in))
        y = np.random.default_rng(k + 50).normal(size=(2, h.n_out))
        dx, dy = DeviceArray.from_host(x), DeviceArray.empty((3, h.n_out), np.float64)
        check(lib().rg_apply_f64(h.ptr, 0, dx.ptr, dy.ptr, 3, 0))
        assert np.array_equal(dy.to_host(), oracle.apply(ii, io, w, x, h.n_out))
        dyt, dxt = DeviceArray.from_host(y), DeviceArray.empty((2, h.n_in), np.float64)
        check(lib().rg_apply_f64(h.ptr, 1, dyt.ptr, dxt.ptr, 2, 0))
        assert np.array_equal(dxt.to_host(), oracle.apply(io, ii, w, y, h.n_in))


def test_batch_handles_outlive_each_other():
    pairs = _mixed_pairs()[:5]
    ref = [oracle.weights_2d(ci, co) for ci, co in pairs]
    hs = build_2d_batched(pairs)
    order = [3, 0, 4]
    for k in order:
        hs[k].free()
    for k in (1, 2):
        for a, b in zip(hs[k].export(), ref[k]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("K,block", [(1, 64), (4, 256), (32, 512)])
def test_batch_tuning_invariant(K, block):
    pairs = _mixed_pairs()
    packed = pack_pairs(pairs)
    ref = [h.export() for h in build_2d_batched(packed)]
    got = [h.export() for h in build_2d_batched(packed, edges_per_lane=K, block=block)]
    for r, g in zip(ref, got):
        for a, b in zip(r, g):
            assert np.array_equal(a, b)


def test_batch_errors():
    pairs = _mixed_pairs()[:3]
    bad = list(pairs)
    x = np.zeros((1, 5))
    bad.insert(1, ((x, x), pairs[0][1]))
    with pytest.raises(RuntimeError, match=r"pair 1"):
        build_2d_batched(bad)
    lib().rg_inject_failure(b"build")
    try:
        with pytest.raises(RuntimeError, match="injected"):
            build_2d_batched(pairs)
    finally:
        lib().rg_inject_failure(None)
    assert all(h.nnz > 0 for h in build_2d_batched(pairs))


def test_autotune_picks_a_candidate(tmp_path, monkeypatch):
    from regridding_amd import _autotune

    monkeypatch.setenv("RG_TUNE_CACHE", str(tmp_path / "tune.json"))
    packed = pack_pairs(_mixed_pairs())
    choice = _autotune.tune(packed, reps=1, edges_per_lane=(2, 16), blocks=(64, 256), force=True)
    assert choice["edges_per_lane"] in (2, 16) and choice["block"] in (64, 256)
    assert len(choice["table"]) == 4 and all(v > 0 for v in choice["table"].values())
    assert min(choice["table"].values()) == choice["ms"]
    assert _autotune.cached(packed)["edges_per_lane"] == choice["edges_per_lane"]
    ref = [h.export() for h in build_2d_batched(packed)]
    got = [h.export() for h in build_2d_batched(packed, edges_per_lane=choice["edges_per_lane"],
                                                 block=choice["block"])]
    for r, g in zip(ref, got):
        for a, b in zip(r, g):
            assert np.array_equal(a, b)
