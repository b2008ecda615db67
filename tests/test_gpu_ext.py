"""The enlarged HR grid (Ntot != Nfft: |E|^2 placed on the Ntot grid through
hrmap, binned there) and the kernel convolution (kernconv: FFT, product with
ftkernel, inverse FFT before binning) against the float64 oracle: spots,
detector image and CoG slopes within 1e-5 of the largest value."""
import numpy as np
import pytest

from conftest import relative_array_error
from oracle import centroid as oc
from oracle import wfs as ow
from shwfs import configs, system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,kernconv", [("wide", False), ("wide", True), ("small", True),
                                           ("noisy", True)])
def test_enlarged_grid_and_kernel_convolution_match_oracle(name, kernconv):
    cfg = configs.get(name)
    cfg["p_wfs"].noise = -1.0          # noise-free frames
    s = system.ShSystem(cfg)
    phase = system.make_screens(s.cfg, 1)[0]
    s.set_phase(phase)
    s.wfs.set_kernconv(kernconv)
    s.comp_image(noise=False)
    t, w = s.tables(), s.p_wfs
    print("%s: Nfft %d Ntot %d kernconv %d nvalid %d" % (name, w._Nfft, w._Ntot, kernconv,
                                                          w._nvalid), flush=True)
    b, _, img = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                              t["jstart"], t["validsubsx"], t["validsubsy"], t["binmap"],
                              t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"], t["npix"],
                              t["nxsub"], Ntot=w._Ntot, hrmap=w._hrmap,
                              ftkernel=w._ftkernel if kernconv else None)
    eb = relative_array_error(s.bincube(), b)
    ei = relative_array_error(s.binimg(), img)
    es = 0.0
    if cfg["p_centroider"].type == "cog":
        s.centro.get_cog()
        got = np.array(s.wfs.d_slopes).ravel()
        ref, _ = oc.cog(img, w._validsubsx, w._validsubsy, w.npix, s.centro.offset,
                        s.centro.scale)
        es = relative_array_error(got, ref)
    print("  spots %.2e image %.2e CoG slopes %.2e" % (eb, ei, es), flush=True)
    assert eb < 1e-5 and ei < 1e-5 and es < 1e-5
