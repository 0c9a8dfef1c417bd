"""The binning rectangle's column bound against the rasteriser it must
contain.  k_front bins an unclipped triangle into the tiles of the columns
[floor(xmin - slack), ceil(xmax + slack)] with

    slack = SW_RECT_MARGIN + (rows + 8) * max(|xmin|, |xmax|, 1) * 2^-20

(kernels.hip: quick_extent, sub_extent), rows = ceil(ymax) - ceil(ymin).
A pixel that fill_triangle_2 covers outside those columns would be lost.
The oracle's rasteriser (renderer.cpp:361-558, the same float sequence the
GPU replays bit for bit) is run on triangles of every shape the workloads
hold and on adversarial ones -- vertices on integers and half-integers,
slivers a few hundred rows tall, vertices far outside the viewport, where
the accumulated error of x += ratio is largest -- and every covered pixel
must lie inside the columns.  The bound is then checked to have teeth: a
rectangle of the bare vertex range, one column short on each side, loses
pixels on these same triangles."""
import ctypes as C

import numpy as np

import oracle
from swegl_amd import _abi as A

f32 = np.float32
SW_RECT_MARGIN = 0.5   # kernels.hip


def columns(v, margin, with_error=True):
    """The kernel's column range for the screen triangle v (3 x 3), float32."""
    xs, ys = [f32(v[k][0]) for k in range(3)], [f32(v[k][1]) for k in range(3)]
    xmin, xmax = min(xs), max(xs)
    iy0, iy2 = int(np.ceil(min(ys))), int(np.ceil(max(ys)))
    X = max(abs(xmin), abs(xmax), f32(1.0))
    err = f32(f32(f32(iy2 - iy0 + 8) * X) * f32(1.0 / 1048576.0)) if with_error else f32(0.0)
    slack = f32(f32(margin) + err)
    return int(np.floor(f32(xmin - slack))), int(np.ceil(f32(xmax + slack)))


def coverage(v, w, h):
    cover = np.zeros((h, w), np.uint32)
    zlast = np.zeros((h, w), np.float32)
    tri = (C.c_float * 9)(*[float(x) for t in v for x in t])
    oracle.lib().oracle_raster_triangle(tri, w, h, A.u32ptr(cover), A.fptr(zlast))
    return cover


def triangles(seed, w, h):
    rng = np.random.default_rng(seed)
    out = []
    z = lambda: float(rng.uniform(0.1, 0.9))
    for _ in range(300):   # the height field's 1-3 pixel triangles, anywhere
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        out.append([(cx + rng.uniform(-2, 2), cy + rng.uniform(-2, 2), z()) for _ in range(3)])
    for _ in range(150):   # vertices on integers and half-integers
        out.append([(float(rng.integers(0, w)) + 0.5 * rng.integers(0, 2),
                     float(rng.integers(0, h)) + 0.5 * rng.integers(0, 2), z()) for _ in range(3)])
    for _ in range(150):   # tall slivers: hundreds of adds of a non-dyadic slope
        x0, x1 = rng.uniform(0, w, 2)
        out.append([(x0, rng.uniform(0, 3), z()), (x1, rng.uniform(h - 3, h), z()),
                    (x0 + rng.uniform(-1.5, 1.5), rng.uniform(0, h), z())])
    for _ in range(150):   # one or two vertices far outside the viewport
        far = 10.0 ** rng.uniform(3, 6)
        out.append([(far * rng.choice([-1, 1]), rng.uniform(-far, far), z()),
                    (rng.uniform(0, w), rng.uniform(0, h), z()),
                    (rng.uniform(0, w), rng.uniform(0, h), z())])
    for _ in range(100):   # medium and large triangles
        out.append([(rng.uniform(-w, 2 * w), rng.uniform(-h, 2 * h), z()) for _ in range(3)])
    return out


def escaped(v, cover, margin, with_error=True):
    ys, xs = np.nonzero(cover)
    if xs.size == 0:
        return 0
    x0, x1 = columns(v, margin, with_error)
    return int(((xs < x0) | (xs > x1)).sum())


def test_binning_columns_hold_every_covered_pixel():
    w, h = 256, 320
    lost = tight_lost = 0
    drawn = 0
    for v in triangles(3, w, h):
        cover = coverage(v, w, h)
        drawn += cover.any()
        lost += escaped(v, cover, SW_RECT_MARGIN)
        tight_lost += escaped(v, cover, -1.0, with_error=False)
    assert drawn > 700
    assert lost == 0, f"{lost} covered pixels outside the binning columns"
    # teeth: one column short on each side of the vertex range loses pixels
    assert tight_lost > 0
