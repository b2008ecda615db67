"""Oracle of the guide-star raytrace (test infrastructure only).

Restates SutraSource::raytrace(SutraAtmos*) (libsutra/src.cpp/
sutra_source.cpp:222-267) and its kernel generic_raytrace / raytrace_krnl
(libsutra/src.cu/sutra_target.cu:25-77): the source screen (nx x ny, x
fastest) is cleared, then every layer adds, at each point (x, y),

    x1 = (x - nx/2) G,  y1 = (y - ny/2) G
    x2 = (cos(t) x1 - sin(t) y1) + nx/2 + xoff/delta
    y2 = (sin(t) x1 + cos(t) y1) + ny/2 + yoff/delta
    xref = x2 delta + dx,  yref = y2 delta + dy
    i = trunc(xref), j = trunc(yref),  u = xref - i, v = yref - j
    value = (1-u)(1-v) L[i, j] + u (1-v) L[i+1, j] + (1-u) v L[i, j+1] + u v L[i+1, j+1]

when i + 1 < dim and j + 1 < dim (the layer is dim x dim), else nothing.
The reference reads before the layer for i < 0 or j < 0; here such points
add nothing (the same as the kernels this oracle checks).  xoff / yoff
include the layer's accumulated wind shift (accumx / accumy).

The coordinates are computed in float32, in the reference's order of
operations (t = thetaML; cos / sin of t taken in float32), so that the
integer parts and the weights are those of a float32 kernel; the
interpolation and the sum over the layers are float64.
"""
import numpy as np


def layer_values(screen, nx, ny, xoff, yoff, G=1.0, thetaML=0.0, dx=0.0, dy=0.0, delta=1.0):
    """The value one layer adds at every point of an nx x ny screen, [nx, ny]
    (x first: the F-order view of the x-fastest screen)."""
    L = np.asarray(screen, dtype=np.float64)
    dim = L.shape[0]
    f = np.float32
    x = np.arange(nx, dtype=f)[:, None]
    y = np.arange(ny, dtype=f)[None, :]
    G, t, dx, dy, delta = f(G), f(thetaML), f(dx), f(dy), f(delta)
    c, s = f(np.cos(t)), f(np.sin(t))
    hx, hy = f(nx) / f(2.0), f(ny) / f(2.0)
    x1 = (x - hx) * G
    y1 = (y - hy) * G
    x2 = ((c * x1 - s * y1) + hx) + f(xoff) / delta
    y2 = ((s * x1 + c * y1) + hy) + f(yoff) / delta
    xref = (x2 * delta + dx).astype(f)
    yref = (y2 * delta + dy).astype(f)
    i = np.trunc(xref).astype(np.int64)
    j = np.trunc(yref).astype(np.int64)
    u = (xref - i.astype(f)).astype(np.float64)
    v = (yref - j.astype(f)).astype(np.float64)
    ok = (i >= 0) & (j >= 0) & (i + 1 < dim) & (j + 1 < dim)
    ic = np.clip(i, 0, dim - 2)
    jc = np.clip(j, 0, dim - 2)
    # L is indexed [x, y] (F-order view of an x-fastest screen)
    val = ((1 - u) * (1 - v) * L[ic, jc] + u * (1 - v) * L[ic + 1, jc]
           + (1 - u) * v * L[ic, jc + 1] + u * v * L[ic + 1, jc + 1])
    return np.where(ok, val, 0.0)


def raytrace(layers, nx, ny):
    """Source screen [nx, ny] (x first) from a list of layers, each a dict
    with 'screen' ([dim, dim], x first) and optional xoff, yoff, G, thetaML,
    dx, dy, delta; summed in list order."""
    out = np.zeros((nx, ny))
    for l in layers:
        out += layer_values(l["screen"], nx, ny, l.get("xoff", 0.0), l.get("yoff", 0.0),
                            l.get("G", 1.0), l.get("thetaML", 0.0), l.get("dx", 0.0),
                            l.get("dy", 0.0), l.get("delta", 1.0))
    return out
