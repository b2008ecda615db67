"""Pins of the raytrace oracle (oracle/raytrace.py) by closed forms of
bilinear interpolation: integer shifts copy the layer exactly, a constant
layer stays constant inside the layer and adds nothing past its edge, a
linear layer a + b x + c y is reproduced exactly at the (rotated, magnified,
compressed, shifted) coordinates, and layers add."""
import numpy as np
import pytest

from oracle import raytrace as rt


def _linear(dim, a, b, c):
    i = np.arange(dim, dtype=np.float64)[:, None]
    j = np.arange(dim, dtype=np.float64)[None, :]
    return a + b * i + c * j


@pytest.mark.parametrize("xoff,yoff", [(0, 0), (3, 5), (17, 2)])
def test_integer_shift_copies_the_layer(xoff, yoff):
    rng = np.random.default_rng(1)
    L = rng.normal(size=(48, 48))
    out = rt.layer_values(L, 24, 20, float(xoff), float(yoff))
    # source point (x, y) reads layer point (x + xoff, y + yoff)
    assert np.array_equal(out, L[xoff:xoff + 24, yoff:yoff + 20])


def test_constant_layer_and_its_edge():
    L = np.full((32, 32), 2.5)
    out = rt.layer_values(L, 32, 32, 0.25, 0.5)
    # xref = x + 0.25: the last point whose right neighbour exists is x = 30
    assert np.allclose(out[:31, :31], 2.5, rtol=0, atol=1e-12)
    assert np.all(out[31, :] == 0.0) and np.all(out[:, 31] == 0.0)


@pytest.mark.parametrize("G,theta,delta,dx,dy", [(1.0, 0.0, 1.0, 0.0, 0.0),
                                                 (1.3, 0.0, 1.0, 0.4, -0.2),
                                                 (1.0, 0.3, 1.0, 0.0, 0.0),
                                                 (0.9, -0.2, 0.8, 1.5, 0.7)])
def test_linear_layer_is_reproduced_at_the_traced_coordinates(G, theta, delta, dx, dy):
    a, b, c = 0.7, 0.031, -0.052
    dim, nx, ny, xoff, yoff = 128, 40, 36, 30.37, 41.81
    out = rt.layer_values(_linear(dim, a, b, c), nx, ny, xoff, yoff, G, theta, dx, dy, delta)
    # coordinates from the formula in float64 (the oracle uses float32)
    x = np.arange(nx, dtype=np.float64)[:, None]
    y = np.arange(ny, dtype=np.float64)[None, :]
    x1, y1 = (x - nx / 2) * G, (y - ny / 2) * G
    xr = (np.cos(theta) * x1 - np.sin(theta) * y1 + nx / 2 + xoff / delta) * delta + dx
    yr = (np.sin(theta) * x1 + np.cos(theta) * y1 + ny / 2 + yoff / delta) * delta + dy
    ref = a + b * xr + c * yr
    assert np.all(out != 0.0)
    # float32 coordinates: ~1e-5 pixel
    assert np.abs(out - ref).max() < 1e-4 * (abs(b) + abs(c)) * 10


def test_layers_add_in_order():
    rng = np.random.default_rng(2)
    l1 = dict(screen=rng.normal(size=(64, 64)), xoff=3.3, yoff=7.9)
    l2 = dict(screen=rng.normal(size=(80, 80)), xoff=11.1, yoff=2.25, G=1.1)
    both = rt.raytrace([l1, l2], 30, 30)
    one = rt.raytrace([l1], 30, 30) + rt.raytrace([l2], 30, 30)
    assert np.allclose(both, one, rtol=0, atol=1e-14)
