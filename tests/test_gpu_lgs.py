"""LGS elongation (SutraLGS, shwfs_lgs_prof2d / shwfs_lgs_kernels, the LGS
convolution of the spot kernel) against the float64 oracle (oracle/lgs.py):
the beam-convolved profiles, the rotated kernels and their transforms of an
8x8 LGS WFS launched from the centre and from the edge, loaded kernels, the
spots, image and CoG slopes of LGS frames, and a Dirac kernel that leaves
the frame of the un-rolled binning window unchanged."""
import numpy as np
import pytest

from conftest import relative_array_error
from oracle import centroid as oc
from oracle import lgs as ol
from oracle import wfs as ow
from shwfs import configs, system

pytestmark = pytest.mark.gpu

f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)  # noqa: E731


def _rel(a, b):
    """Largest difference over the largest magnitude (complex arrays too)."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _lgs_system(lltx):
    cfg = configs.lgs(lltx=lltx)
    cfg["p_wfs"].noise = -1.0
    return system.ShSystem(cfg)


def _oracle_prof2d(w, L):
    return ol.prof2d(f32(w._prof1d), f32(w._profcum), int(L.nprof), f32(w._doffaxis),
                     float(np.float32(L.hg)), float(np.float32(L.pixsize)), float(np.float32(L.h0)),
                     float(np.float32(L.deltah)), np.asarray(w._ftbeam, np.complex64), w._Ntot)


@pytest.mark.parametrize("lltx", [0.0, 4.5])
def test_profiles_and_kernels_match_oracle(lltx):
    s = _lgs_system(lltx)
    w = s.p_wfs
    L = s.wfs.d_gs.d_lgs
    assert s.wfs.lgs and s.wfs.d_gs.lgs and L.npix == w._Ntot and L.nvalid == w._nvalid
    print("lgs lltx %.1f: Ntot %d nvalid %d nprof %d hg %.1f elongation up to %.2f arcsec" %
          (lltx, w._Ntot, w._nvalid, L.nprof, L.hg,
           206265.0 * np.max(w._doffaxis) * 4 * 2.355 * 1000 / L.hg ** 2), flush=True)
    p_dev = np.array(L.d_prof2d)  # [Ntot, nvalid] complex
    p_or = _oracle_prof2d(w, L)
    ep = _rel(p_dev, p_or)
    # kernels from the device's own profiles: the rotation and the transform
    lk_or, ft_or = ol.kernels(p_dev.astype(np.complex128), f32(w._beam), f32(w._azimuth), w._Ntot)
    lk = np.array(L.d_lgskern)
    ft = np.array(L.d_ftlgskern)
    # The rotation clamps a coordinate at n - 2 to the last sample with weight
    # 1 (rotate_krnl): a coordinate within rounding of n - 2 may land on
    # either side in float32, on the device as in the reference, so those
    # samples are compared against both sides.
    n = w._Ntot
    c = n // 2
    u = np.arange(n)[:, None] - c
    v = np.arange(n)[None, :] - c
    near = np.zeros_like(lk, dtype=bool)
    for k, th in enumerate(f32(w._azimuth)):
        tu = u * np.cos(th) + v * np.sin(th) + c
        tv = -u * np.sin(th) + v * np.cos(th) + c
        near[:, :, k] = (np.abs(tu - (n - 2)) < 1e-4) | (np.abs(tv - (n - 2)) < 1e-4)
    ek = _rel(np.where(near, 0, lk), np.where(near, 0, lk_or)) if (~near).any() else 0.0
    ef = _rel(ft, np.fft.fft2(f32(lk), axes=(0, 1)))  # the device's transform of its kernels
    print("  prof2d %.2e  lgskern %.2e (%d samples at the clamp edge)  ftlgskern %.2e" %
          (ep, ek, int(near.sum()), ef), flush=True)
    assert near.mean() < 0.01
    assert ep < 1e-5 and ek < 1e-5 and ef < 1e-5


def test_loaded_kernels_are_transformed():
    s = _lgs_system(0.0)
    w = s.p_wfs
    L = s.wfs.d_gs.d_lgs
    k = np.random.default_rng(7).random((w._Ntot, w._Ntot, w._nvalid)).astype(np.float32)
    L.load_kernels(k)
    assert np.array_equal(np.array(L.d_lgskern), k)
    e = _rel(np.array(L.d_ftlgskern), np.fft.fft2(f32(k), axes=(0, 1)))
    print("load_kernels: ftlgskern %.2e" % e, flush=True)
    assert e < 1e-5


def _frame_vs_oracle(s, phase, ftlgskern):
    t, w = s.tables(), s.p_wfs
    b, _, img = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                              t["jstart"], t["validsubsx"], t["validsubsy"], t["binmap"],
                              t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"], t["npix"],
                              t["nxsub"], ftlgskern=ftlgskern)
    eb = relative_array_error(s.bincube(), b)
    ei = relative_array_error(s.binimg(), img)
    s.centro.get_cog()
    got = np.array(s.wfs.d_slopes).ravel()
    ref, _ = oc.cog(img, w._validsubsx, w._validsubsy, w.npix, s.centro.offset, s.centro.scale)
    return eb, ei, relative_array_error(got, ref), b


@pytest.mark.parametrize("lltx", [0.0, 4.5])
def test_lgs_frame_matches_oracle(lltx):
    s = _lgs_system(lltx)
    phase = system.make_screens(s.cfg, 1)[0]
    s.set_phase(phase)
    s.comp_image(noise=False)
    w = s.p_wfs
    L = s.wfs.d_gs.d_lgs
    if lltx == 0.0:
        # the whole chain in float64: profiles, kernels, frame (with the launch
        # at the centre no rotated sample sits at the clamp edge)
        _, ft_or = ol.kernels(_oracle_prof2d(w, L), f32(w._beam), f32(w._azimuth), w._Ntot)
    else:
        # the frame stage from the device's kernels (their edge samples may
        # round to either side of the clamp, see above)
        ft_or = np.array(L.d_ftlgskern).astype(np.complex128)
    eb, ei, es, b = _frame_vs_oracle(s, phase, ft_or)
    # the spots are elongated: the LGS frame differs from the plain one
    t = s.tables()
    b_plain, _, _ = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                                  t["jstart"], t["validsubsx"], t["validsubsy"], t["binmap"],
                                  t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"], t["npix"],
                                  t["nxsub"], ftlgskern=np.ones_like(ft_or))
    moved = relative_array_error(b_plain, b)
    print("lgs frame lltx %.1f: spots %.2e image %.2e slopes %.2e (elongation changes the spots "
          "by %.2e)" % (lltx, eb, ei, es, moved), flush=True)
    assert moved > 1e-2
    assert eb < 1e-5 and ei < 1e-5 and es < 1e-5


def test_dirac_kernel_frame_equals_the_unconvolved_frame():
    s = _lgs_system(0.0)
    w = s.p_wfs
    n = w._Ntot
    k = np.zeros((n, n, w._nvalid), np.float32)
    k[0, 0, :] = 1.0
    s.wfs.d_gs.d_lgs.load_kernels(k)
    phase = system.make_screens(s.cfg, 1)[0]
    s.set_phase(phase)
    s.comp_image(noise=False)
    eb, ei, es, _ = _frame_vs_oracle(s, phase, np.ones((n, n, w._nvalid)))
    print("dirac lgs kernel: spots %.2e image %.2e slopes %.2e" % (eb, ei, es), flush=True)
    assert eb < 1e-5 and ei < 1e-5 and es < 1e-5
