"""The brightest-pixel CoG oracle pinned by hand-computed cases (CPU).

Each case fixes one part of the definition (libsutra/src.cu/
sutra_centroider_bpcog.cu:28-78): the threshold is element nbpix of the window
sorted in decreasing order, ties included; pixels at or below it count 0;
pixels outside the image count 0 in the sort; the thresholded values are
written back; the moments use x = p % npix, y = p // npix."""
import numpy as np

from oracle import centroid as oc


def _one(window, nbpix, offset=0.0, scale=1.0):
    """A single 2x2 sub-aperture at the corner of a 2x2 image (x fastest)."""
    img = np.asarray(window, np.float64).reshape((2, 2), order="F")
    return oc.bpcog(img, [0], [0], 2, nbpix, offset, scale)


def test_threshold_is_element_nbpix_of_the_decreasing_order():
    # window (x, y): (0,0)=4, (1,0)=1, (0,1)=3, (1,1)=2; decreasing 4 3 2 1
    sl, inten, out = _one([4, 1, 3, 2], 1)       # t = 3: only the 4 stays, as 1
    assert inten[0] == 1.0
    np.testing.assert_allclose(sl, [0.0, 0.0], atol=1e-12)
    np.testing.assert_array_equal(out.ravel(order="F"), [1, 0, 0, 0])
    sl, inten, out = _one([4, 1, 3, 2], 2)       # t = 2: 4 -> 2 at (0,0), 3 -> 1 at (0,1)
    assert inten[0] == 3.0
    np.testing.assert_allclose(sl, [0.0, 1.0 / (3.0 + 1e-6)], rtol=1e-12)
    np.testing.assert_array_equal(out.ravel(order="F"), [2, 0, 1, 0])


def test_ties_and_the_last_rank():
    sl, inten, out = _one([5, 5, 5, 1], 0)       # t = 5: nothing strictly above
    assert inten[0] == 0.0
    np.testing.assert_array_equal(out.ravel(order="F"), [0, 0, 0, 0])
    sl, inten, out = _one([7, 1, 3, 2], 3)       # t = min = 1: v - 1 everywhere
    np.testing.assert_array_equal(out.ravel(order="F"), [6, 0, 2, 1])
    assert inten[0] == 9.0
    # offset and scale as the other CoG centroiders: (S/(I + 1e-6) - offset) * scale
    sl, _, _ = _one([7, 1, 3, 2], 3, offset=0.5, scale=2.0)
    np.testing.assert_allclose(sl, [(1.0 / (9 + 1e-6) - 0.5) * 2, (3.0 / (9 + 1e-6) - 0.5) * 2],
                               rtol=1e-12)


def test_pixels_outside_the_image_count_zero():
    # a 2x2 window at y = 1 of a 2x2 image: its second row lies past the end
    img = np.array([[9.0, 3.0], [8.0, 4.0]])   # img[x, y]; flat (x fastest): 9 8 3 4
    sl, inten, out = oc.bpcog(img, [0], [1], 2, 1, 0.0, 1.0)
    # window values 3, 4 (row y = 1) and two zeros: decreasing 4 3 0 0, t = 3
    assert inten[0] == 1.0
    np.testing.assert_allclose(sl, [1.0 / (1 + 1e-6), 0.0], rtol=1e-12)
    np.testing.assert_array_equal(out, [[9.0, 0.0], [8.0, 1.0]])
