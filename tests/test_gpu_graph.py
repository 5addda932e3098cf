"""
The batched build and apply captured in one hipGraph (rg_graph_*): every
replay must give the weights and applied values of the eager batched build,
bit for bit, also after the coordinates and values were changed in place,
and a replay that outgrows the graph's capacities must say so in its status
word instead of writing past them.
"""

import numpy as np
import pytest

from oracle import oracle
from regridding_amd import _util, workloads
from regridding_amd._conservative import build_2d_batched, pack_pairs
from regridding_amd._lib import RGError
from regridding_amd.graph import BuildApplyGraph

pytestmark = pytest.mark.gpu


def _pairs(seed=31, n=6, scale=1.0):
    out = []
    for p in workloads.batched_pairs(n_pairs=n, seed=seed, n_min=5, n_max=60):
        gi = tuple(scale * np.asarray(a) for a in p["grid_input"])
        ci, co, *_ = _util.normalize_coordinates(gi, p["grid_output"], perturb=True, seed=42)
        out.append((ci, co))
    # a pair whose grids do not overlap: no weights, nothing applied
    ci, co, *_ = _util.normalize_coordinates(
        workloads.warped_grid((9, 7), 0.0, 0.0, 0.0, -1.0, -0.5),
        workloads.warped_grid((8, 8), 0.0, 0.0, 0.0, 0.5, 1.0), perturb=True, seed=42)
    out.append((ci, co))
    return out


def _values(pairs, seed, batches=(2, 0, 1, 3)):
    rng = np.random.default_rng(seed)
    vals = []
    for k, (ci, _) in enumerate(pairs):
        b = batches[k % len(batches)]
        vals.append(rng.normal(size=(b, (ci[0].shape[0] - 1) * (ci[0].shape[1] - 1))))
    return vals


def _check_replay(g, pairs, vals):
    outs = g.outputs()
    for k, ((ci, co), v) in enumerate(zip(pairs, vals)):
        print(f"pair {k}: {ci[0].shape} -> {co[0].shape} batch {v.shape[0]}", flush=True)
        ii, io, w = oracle.weights_2d(ci, co)
        got = g.weights(k).export()
        for a, b in zip(got, (ii, io, w)):
            assert np.array_equal(a, b)
        n_out = (co[0].shape[0] - 1) * (co[0].shape[1] - 1)
        assert outs[k].shape == (v.shape[0], n_out)
        assert np.array_equal(outs[k], oracle.apply(ii, io, w, v, n_out))


def test_graph_replays_match_eager_build_and_apply():
    pairs = _pairs()
    packed = pack_pairs(pairs)
    vals = _values(pairs, 3)
    g = BuildApplyGraph(packed, vals, slack=0.25)
    try:
        caps = g.capacity()
        assert caps["triples"] > 0 and caps["weights"] > 0
        eager = build_2d_batched(packed)
        for rep in range(3):
            g.launch()
            assert g.status() == 0
            _check_replay(g, pairs, vals)
            # eager work between the replays
            for h, (ci, co) in zip(eager, pairs):
                assert h.stats()["n_triples"] >= 0
            build_2d_batched(packed)
    finally:
        g.close()


def test_graph_replay_after_in_place_update():
    pairs = _pairs()
    packed = pack_pairs(pairs)
    vals = _values(pairs, 4)
    g = BuildApplyGraph(packed, vals, slack=0.5)
    try:
        g.launch()
        assert g.status() == 0
        _check_replay(g, pairs, vals)
        # same shapes, other coordinates and values
        new_pairs = _pairs(seed=31, scale=1.02)
        new_vals = _values(new_pairs, 9)
        packed.update(new_pairs)
        g.set_values(new_vals)
        g.launch()
        assert g.status() == 0
        _check_replay(g, new_pairs, new_vals)
    finally:
        g.close()


def test_graph_with_overflowing_staging():
    # one staged triple per edge: most chunks keep their triples in the
    # overflow buffer, sized by capacity like the others
    pairs = _pairs()
    packed = pack_pairs(pairs)
    vals = _values(pairs, 6)
    g = BuildApplyGraph(packed, vals, slack=0.25, stage_per_edge=1)
    try:
        assert g.capacity()["overflow"] > 4096
        for rep in range(2):
            g.launch()
            assert g.status() == 0
            _check_replay(g, pairs, vals)
    finally:
        g.close()


def _shifted(pairs, dx):
    return [(ci, (co[0] + dx, co[1])) for ci, co in pairs]


def test_graph_capacity_overrun_is_flagged():
    # sized on grids that do not meet (no triples, no weights) and replayed
    # on grids that do: the replay must raise capacity bits, skip the apply
    # and leave the graph usable for inputs that fit again
    pairs = _pairs()
    apart = _shifted(pairs, 10.0)
    packed = pack_pairs(apart)
    vals = _values(pairs, 5)
    # (one staged triple per edge: the replay also outgrows the overflow buffer)
    g = BuildApplyGraph(packed, vals, slack=0.0, stage_per_edge=1)
    try:
        g.launch()
        assert g.status() == 0
        for o in g.outputs():
            assert np.all(o == 0)
        caps = g.capacity()
        n_triples = sum(oracle.weights_2d(ci, co)[0].size for ci, co in pairs)
        print(f"capacity {caps}, overlapping weights {n_triples}", flush=True)
        assert n_triples > caps["weights"]
        packed.update(pairs)
        g.launch()
        f = g.status()
        print(f"status {f:#x}", flush=True)
        assert f != 0
        with pytest.raises(RGError):
            g.check()
        packed.update(apart)
        g.launch()
        assert g.status() == 0
        for o in g.outputs():
            assert np.all(o == 0)
    finally:
        g.close()
