"""
Pins the oracle's 2-D geometry primitives with the reference's own truth
tables (regridding/_tests/test_geometry.py:112-311 and 432-495), ported as
known answers, and with textbook checks.
"""

import numpy as np
import pytest

from oracle import oracle


@pytest.mark.parametrize(
    "x_p1,y_p1,x_p2,y_p2,x_q1,y_q1,x_q2,y_q2,expected",
    [
        (0, 0, 1, 1, 0, 0, 1, 1, True),
        (0, 0, 1, 1, 2, 0, 3, 1, False),
        (0, 0, 2, 2, 1, 1, 3, 3, True),
        (0, 0, 1, 1, 1, 1, 2, 2, True),
        (-1, -1, 0, 0, 0, 0, 1, 1, True),
        (-1, 1, 0, 0, 0, 0, 1, 1, True),
        (-1, -1, 0, 0, 1, 0, 0, 1, True),
        (1, 1, 0, 0, 2, 2, 1, 1, True),
    ],
)
def test_bounding_boxes_intersect(x_p1, y_p1, x_p2, y_p2, x_q1, y_q1, x_q2, y_q2, expected):
    assert oracle.bbox_intersect(x_p1, y_p1, x_p2, y_p2, x_q1, y_q1, x_q2, y_q2) == expected


@pytest.mark.parametrize(
    "line_1,line_2,expected",
    [
        (((0, 0), (0, 1)), ((1, 0), (1, 1)), False),
        (((0, 0), (1, 0)), ((0, 1), (1, 1)), False),
        (((0, 0), (1, 0)), ((2, -1), (2, +1)), False),
        (((0, -1), (0, +1)), ((-1, 0), (+1, 0)), True),
        (((0, +1), (0, -1)), ((-1, 0), (+1, 0)), True),
        (((0, 0), (1, 1)), ((0, 1), (1, 0)), True),
        (((1, 1), (0, 0)), ((0, 1), (1, 0)), True),
    ],
)
def test_two_line_segments_intersect(line_1, line_2, expected):
    t, u = oracle.seg_params(*line_1, *line_2)
    assert oracle.seg_hit(t, u) == expected


@pytest.mark.parametrize(
    "line_1,line_2,expected",
    [
        (((0, -1), (0, +1)), ((-1, 0), (+1, 0)), (0, 0)),
        (((0, +1), (0, -1)), ((-1, 0), (+1, 0)), (0, 0)),
        (((0, 0), (1, 1)), ((0, 1), (1, 0)), (0.5, 0.5)),
        (((1, 1), (0, 0)), ((0, 1), (1, 0)), (0.5, 0.5)),
    ],
)
def test_two_line_segment_intersection(line_1, line_2, expected):
    (x1, y1), (x2, y2) = line_1
    t, _ = oracle.seg_params(*line_1, *line_2)
    assert (x1 + t * (x2 - x1), y1 + t * (y2 - y1)) == expected


def test_segment_parameters_half_open_and_parallel():
    # t = 0 counts, t = 1 does not (geometry.py:471-472)
    assert oracle.seg_hit(*oracle.seg_params((0, 0), (1, 0), (0, -1), (0, 1)))
    assert not oracle.seg_hit(*oracle.seg_params((0, 0), (1, 0), (1, -1), (1, 1)))
    # parallel, overlapping boxes: det = 0 -> inf/NaN -> no hit
    t, u = oracle.seg_params((0, 0), (2, 0), (1, 0), (3, 0))
    assert not oracle.seg_hit(t, u)
    # disjoint boxes return (inf, inf) before dividing
    assert oracle.seg_params((0, 0), (1, 1), (5, 5), (6, 7)) == (np.inf, np.inf)


def test_segment_parameters_closed_form():
    # t and u solve p1 + t (p2 - p1) = q1 + u (q2 - q1) for random segments
    rng = np.random.default_rng(0)
    for _ in range(200):
        p1, p2, q1, q2 = rng.normal(size=(4, 2))
        t, u = oracle.seg_params(p1, p2, q1, q2)
        if not np.isfinite(t):
            continue
        a = np.array([p2 - p1, q1 - q2]).T
        tt, uu = np.linalg.solve(a, q1 - p1)
        assert t == pytest.approx(tt, rel=1e-9, abs=1e-9)
        assert u == pytest.approx(uu, rel=1e-9, abs=1e-9)


RED_CROSS = [
    (8, 4, False), (7, 4, False), (6, 4, False), (5, 4, True), (4, 4, True), (3.5, 4, True),
    (3, 4, True), (2, 4, True), (1, 4, False), (0, 4, False), (-1, 4, False), (-2, 4, False),
    (-3, 4, False), (-4, 4, False), (-5, 4, False), (-6, 4, False), (-7, 4, False),
    (-8, 4, False), (3, 8, False), (3, 7, False), (3, 6, False), (3, 5, True), (3, 4, True),
    (3, 3.5, True), (3, 3, True), (3, 2, True), (3, 1, False), (3, 0, False), (3, -1, False),
    (3, -2, False), (3, -3, False), (3, -4, False), (3, -5, False), (3, -6, False),
    (3, -7, False), (3, -8, False),
]


@pytest.mark.parametrize("x,y,expected", RED_CROSS)
def test_point_is_inside_red_cross(x, y, expected):
    vx = np.array([2, 2, 3, 3, 4, 4, 3, 3, 2, 2, 1, 1]) + 1
    vy = np.array([2, 1, 1, 2, 2, 3, 3, 4, 4, 3, 3, 2]) + 1
    assert oracle.point_in_polygon(x, y, vx, vy) == expected
    assert oracle.point_in_polygon(x, y, vx[::-1], vy[::-1]) == expected


def test_point_in_polygon_random_convex():
    # random convex polygons: inside iff on the inner side of every edge
    rng = np.random.default_rng(1)
    for _ in range(50):
        ang = np.sort(rng.uniform(0, 2 * np.pi, 7))
        vx, vy = np.cos(ang), np.sin(ang)
        for px, py in rng.uniform(-1.2, 1.2, (40, 2)):
            cross = (np.roll(vx, -1) - vx) * (py - vy) - (np.roll(vy, -1) - vy) * (px - vx)
            if np.min(np.abs(cross)) < 1e-9:
                continue
            assert oracle.point_in_polygon(px, py, vx, vy) == bool(np.all(cross > 0))
