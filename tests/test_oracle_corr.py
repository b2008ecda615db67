"""Pins of the correlation-centroider oracle (oracle/corr.py): the map is the
linear cross-correlation summed from its definition, a spot that is the
reference shifted by whole pixels gives exactly that shift, the parabola fit
recovers the vertex of an exact quadratic, and the acquisim gather
(bcube_krnl) picks the listed windows."""
import numpy as np
import pytest

from oracle import corr as oc


def _gauss(n, cx, cy, w=1.3):
    x = np.arange(n)[:, None]
    y = np.arange(n)[None, :]
    return np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * w * w))


def test_map_is_the_linear_cross_correlation():
    rng = np.random.default_rng(1)
    n = 6
    b, r = rng.random((n, n)), rng.random((n, n))
    got = oc.corr_map(b, r)
    ref = np.zeros((2 * n - 1, 2 * n - 1))
    for kx in range(-(n - 1), n):
        for ky in range(-(n - 1), n):
            acc = 0.0
            for mx in range(n):
                for my in range(n):
                    if 0 <= mx + kx < n and 0 <= my + ky < n:
                        acc += b[mx + kx, my + ky] * r[mx, my]
            ref[kx + n - 1, ky + n - 1] = acc
    # the reference's unnormalised inverse transform on the 2n x 2n grid
    assert np.allclose(got, (2 * n) ** 2 * ref, rtol=0, atol=1e-9)


@pytest.mark.parametrize("dx,dy", [(0, 0), (1, -2), (-2, 1), (3, 3)])
def test_whole_pixel_shift_is_measured_exactly(dx, dy):
    # a compact pattern moved inside the window: the spot is exactly the
    # reference shifted, its correlation the (symmetric) autocorrelation
    # moved to lag (dx, dy), whose parabola vertex is that lag
    n = 8
    pat = np.random.default_rng(5).random((3, 3)) + 0.5
    r = np.zeros((n, n))
    r[2:5, 2:5] = pat
    img = np.zeros((16, 8))
    img[2 + dx:5 + dx, 2 + dy:5 + dy] = pat
    img[8:, :] = r
    sl = oc.centroids(img, [0, 8], [0, 0], n, r, n / 2 - 0.5, 0.25)
    assert np.allclose(sl, [0.25 * dx, 0.0, 0.25 * dy, 0.0], rtol=0, atol=1e-9)


def test_parabola_fit_recovers_the_vertex():
    M = oc.interp_mat(3, 3)
    x0, y0 = 0.3, -0.2
    x = np.array([-1, 0, 1] * 3, float)
    y = np.repeat([-1.0, 0.0, 1.0], 3)
    f = 5.0 - 2.0 * (x - x0) ** 2 - 1.5 * (y - y0) ** 2 + 0.4 * (x - x0) * (y - y0)
    c = M @ f
    D = c[2] ** 2 - 4 * c[1] * c[0]
    assert np.isclose((2 * c[1] * c[3] - c[4] * c[2]) / D, x0)
    assert np.isclose((2 * c[0] * c[4] - c[3] * c[2]) / D, y0)


def test_acquisim_gather_known_answer():
    img = np.arange(6 * 6, dtype=np.float64).reshape(6, 6, order="F")  # img[x, y] = x + 6 y
    got = oc.acquisim_bincube(img, [0, 3], [3, 0], 3, 6)
    assert got[1, 2, 0] == 1 + 6 * (2 + 3)
    assert got[2, 0, 1] == (2 + 3) + 6 * 0


def test_product_interp_mat_equals_the_oracle():
    """shwfs.rtc_init.create_interp_mat (the matrix the RTC set-up uploads)
    is the oracle's fit matrix, transposed to [dimx*dimy, 6]."""
    from shwfs import rtc_init
    for dx, dy in ((3, 3), (5, 5), (3, 5)):
        assert np.allclose(rtc_init.create_interp_mat(dx, dy), oc.interp_mat(dx, dy).T)


def test_reference_acquisim_stride_leaves_the_image():
    """Why the product gathers with the image side as row stride: the
    reference's SutraAcquisim::comp_image passes npix * nxsub as `nsub` to
    bcube_krnl, whose row stride npix * nsub is then npix times the image
    side; the last window row of the lowest spots already indexes past the
    image."""
    npix, nxsub = 6, 4
    size = npix * nxsub
    vy_last = size - npix
    stride_ref = npix * (npix * nxsub)
    last = (npix - 1) + (npix - 1 + vy_last) * stride_ref
    assert last >= size * size
    img = np.arange(size * size, dtype=np.float64).reshape(size, size, order="F")
    cube = oc.acquisim_bincube(img, [0], [0], npix, size)
    assert np.array_equal(cube[:, :, 0], img[:npix, :npix])
