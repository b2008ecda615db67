"""The vectorised CPU frame (oracle/fast.py, the CPU baseline of bench.py)
against the loop oracle (oracle/wfs.py, oracle/centroid.py): equal to float64
rounding in float64, within 1e-5 of the maximum in float32."""
import numpy as np
import pytest

from oracle import centroid as oc
from oracle import fast
from oracle import wfs as ow
from shwfs import configs, geom, system


def _tables(name):
    c = configs.get(name)
    system.geometry(c)
    w, g = c["p_wfs"], c["p_geom"]
    t = dict(pupil=g._mpupil, scale=np.float32(2 * np.pi / w.Lambda), offsets=w._halfxy,
             istart=w._validpuppixx, jstart=w._validpuppixy, validsubsx=w._validsubsx,
             validsubsy=w._validsubsy, binmap=w._binmap, fluxPerSub=geom.flux_per_valid_subap(w),
             nphot=np.float32(w._nphotons), pdiam=w._pdiam, Nfft=w._Nfft, npix=w.npix,
             nxsub=w.nxsub)
    return c, w, t


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_fast_frame_matches_the_loop_oracle(dtype, tol):
    c, w, t = _tables("small")
    phase = system.make_screens(c, 1)[0]
    b, inten, img = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                                  t["jstart"], t["validsubsx"], t["validsubsy"],
                                  t["binmap"], t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"],
                                  t["npix"], t["nxsub"])
    f = fast.Frame(t, dtype)
    offset, scale = w.npix / 2 - 0.5, w.pixsize
    thr = 0.1 * img.max()
    fb, fs, fimg = f.frame(phase, offset, scale, thr, workers=2)
    ref_s, _, ref_img = oc.tcog(img, w._validsubsx, w._validsubsy, w.npix, offset, scale, thr)
    cube = b.transpose(2, 1, 0).reshape(len(t["istart"]), -1)  # [s, px + py npix]
    assert np.abs(fb - cube).max() / np.abs(cube).max() < tol
    assert np.abs(fs - ref_s).max() / np.abs(ref_s).max() < tol
    assert np.abs(fimg - ref_img).max() / np.abs(ref_img).max() < tol
    cog, _ = f.centroids(f.spots(phase)[2], offset, scale)
    ref_c, _ = oc.cog(img, w._validsubsx, w._validsubsy, w.npix, offset, scale)
    assert np.abs(cog - ref_c).max() / np.abs(ref_c).max() < tol
