"""CPU oracle for the Shack-Hartmann WFS hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the C++
``sutra``/``carma`` layer, the ``shwfs`` host package) may import, call or link
anything from this package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker / the CPU
baseline, never as the thing measured or shipped.

The oracle restates, in float64 NumPy (pocketfft for the FFT) and plain Python
loops for the small cases, the algorithm of the reference path:

* ``geometry``  - pupil, valid sub-apertures, index maps (binmap, halfxy, ...)
                  (shesha/shesha/init/geom_init.py, shesha/shesha/util/make_pupil.py)
* ``wfs``       - SutraWfsSH::comp_image (libsutra/src.cpp/sutra_wfs_sh.cpp:337-643)
* ``centroid``  - CoG / weighted CoG / thresholded CoG
                  (libsutra/src.cu/sutra_centroider_{cog,wcog,tcog}.cu)
* ``philox``    - Philox4x32-10 counter RNG (Salmon et al., SC'11) and the
                  Poisson / Gaussian samplers the noise stage draws with
* ``turbulence``- von Karman phase structure function (closed form used by the
                  reference's ``rodconan``, shesha/shesha/util/iterkolmo.py:313-351)
* ``mlp``       - float64 forward/backward of the per-subaperture MLP centroider
* ``raytrace``  - guide-star raytrace through layers (libsutra/src.cpp/
                  sutra_source.cpp:222-267, libsutra/src.cu/sutra_target.cu:25-77)
* ``lgs``       - LGS elongation (libsutra/src.cpp/sutra_lgs.cpp:160-212,
                  libsutra/src.cu/sutra_lgs.cu)
* ``corr``      - correlation centroider (libsutra/src.cpp/
                  sutra_centroider_corr.cpp:190-262) and the acquisim gather
                  (libsutra/src.cu/sutra_acquisim.cu:65-87)

Pinning status of every function is listed in ``PINNING`` below; each entry
names the tests that pin it.  Any function marked "parity unpinned" has no
independent check.
"""

PINNING = {
    "geometry.wfs_size": "pinned: tests/test_oracle_geometry.py (two hand-worked cases, Ntot = Nfft and Ntot > Nfft; stated Nfft/nrebin/Ntot of the configs)",
    "geometry.sh_geometry": "pinned: c1 known answer (52 valid), closed-form halfxy, "
                            "binmap injection/fftshift invariants, brute-force recount",
    "wfs.comp_image": "pinned: brute-force DFT, Parseval, flux normalisation, Fourier-shift "
                      "theorem slopes, flat-wavefront centring (tests/test_oracle_wfs.py)",
    "wfs.extend_grid / geometry.hrmap_table": "pinned: frequency-preserving zero padding "
                                              "(fftshift, centred pad, ifftshift), flux kept "
                                              "(tests/test_oracle_wfs_ext.py)",
    "wfs.convolve_kernel": "pinned: Dirac kernel is the identity, Gaussian kernel equals the "
                           "circular convolution summed from its definition, flux kept "
                           "(tests/test_oracle_wfs_ext.py)",
    "centroid.cog/wcog/tcog": "pinned: scipy.ndimage.center_of_mass (as the reference tests do), "
                              "threshold/epsilon special cases (tests/test_oracle_centroid.py)",
    "philox.philox4x32": "pinned: Random123 known-answer vectors (tests/test_oracle_philox.py)",
    "philox.poisson/normal": "pinned: mean/variance/KS tests against scipy.stats; the integer forms of the PTRS uniform tests checked at every threshold (test_oracle_philox.py)",
    "turbulence.structure_function": "pinned: Kolmogorov limit 6.88 (r/r0)^(5/3) and "
                                     "saturation 2*variance for r >> L0",
    "mlp": "pinned: central finite differences of the float64 loss",
    "raytrace.raytrace": "pinned: integer shifts copy the layer exactly, constant layer and its "
                         "edge, linear layers reproduced at the traced coordinates (rotation, "
                         "magnification, compression, shifts), layers add "
                         "(tests/test_oracle_raytrace.py)",
    "corr.corr_map": "pinned: equals (2n)^2 times the linear cross-correlation summed from its "
                     "definition on random data (tests/test_oracle_corr.py)",
    "corr.interp_mat / corr.centroids": "pinned: the fit recovers the vertex of an exact "
                                        "quadratic, whole-pixel shifts of a pattern are measured "
                                        "exactly through the full centroider",
    "corr.acquisim_bincube": "pinned: known-answer gather of listed windows",
    "wfs.field_stop": "pinned: an all-pass stop returns the pupil field, a zero-frequency stop "
                      "returns the field's mean, brute-force DFT pair on a small grid "
                      "(tests/test_oracle_field_stop.py)",
    "lgs.interp_transf / lgs.profiles": "pinned: exact on constant and linear profiles, clamps "
                                        "at both ends (tests/test_oracle_lgs.py)",
    "lgs.prof2d": "pinned: equals the circular convolution with the beam summed from its "
                  "definition; a Dirac beam returns the profile",
    "lgs.kernels": "pinned: a zero azimuth is the identity up to the clamp of the last two "
                   "samples; the transform is numpy's fft2",
    "lgs.convolve": "pinned: a Dirac kernel leaves the images unchanged (times n^2)",
    "geometry.field_stop_mask": "pinned: centred and symmetric under both flips and the "
                                "transpose, centre pixel kept; equals the product's mask",
}
