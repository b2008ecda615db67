"""Pins the geometry oracle and checks the product's geometry against it."""
import numpy as np
import pytest

from oracle import geometry as og
from shwfs import configs, geom, system


def test_small_has_52_valid_subaps():
    # known answer stated for the D=4 m, 8x8 SH workload (fracsub 0.5, cobs 0.12)
    o = og.sh_geometry(64, 8, 8, 8, 2, 16, 16, 0.5, 0.12)
    assert o["nvalid"] == 52
    c = system.geometry(configs.get("small"))
    assert c["p_wfs"]._nvalid == 52


def _config(name):
    # named configs, or "batch<i>": WFS i of the uneven eight-WFS batch
    if name.startswith("batch"):
        return configs.batch()[int(name[5:])]
    return configs.get(name)


# the 80x80 WFS and the three geometries of the uneven batch (74x74, 40x40, 8x8)
@pytest.mark.parametrize("name", ["small", "noisy", "large", "batch0", "batch6", "batch7"])
def test_product_geometry_matches_oracle(name):
    c = system.geometry(_config(name))
    w, g = c["p_wfs"], c["p_geom"]
    o = og.sh_geometry(g.pupdiam, w.nxsub, w._pdiam, w.npix, w._nrebin, w._Nfft, w._Ntot,
                       w.fracsub, c["p_tel"].cobs)
    assert o["nvalid"] == w._nvalid
    assert np.array_equal(o["mpupil"], g._mpupil)
    assert np.array_equal(o["validsubsx"] * w.npix, w._validsubsx)
    assert np.array_equal(o["validsubsy"] * w.npix, w._validsubsy)
    assert np.array_equal(o["istart"], w._validpuppixx)
    assert np.array_equal(o["jstart"], w._validpuppixy)
    assert np.array_equal(o["binmap"], w._binmap)
    assert np.array_equal(o["halfxy"], w._halfxy)
    assert np.array_equal(o["fluxPerSub"], geom.flux_per_valid_subap(w))


def test_large_and_batch_counts():
    assert system.geometry(configs.get("large"))["p_wfs"]._nvalid == 4964
    counts = [system.geometry(c)["p_wfs"]._nvalid for c in configs.batch()]
    assert counts == [4268] * 6 + [1256, 52]


def test_halfxy_closed_form():
    h = og.halfxy_table(16, 32, 8, 4)
    a = np.arange(16)
    assert np.allclose(h, (np.pi * (a[:, None] + a[None, :]) / 32).astype(np.float32), atol=0)
    assert not og.halfxy_table(6, 16, 7, 3).any()   # odd npix and odd nrebin -> zero


@pytest.mark.parametrize("Ntot,npix,nrebin", [(16, 8, 2), (32, 8, 4), (16, 6, 2), (64, 16, 4), (15, 5, 3)])
def test_binmap_invariants(Ntot, npix, nrebin):
    bm = og.binmap_table(Ntot, npix, nrebin)
    assert bm.shape == (nrebin * nrebin, npix * npix)
    flat = bm.ravel()
    assert len(np.unique(flat)) == flat.size           # injection
    assert flat.min() >= 0 and flat.max() < Ntot * Ntot
    # un-shifting gives a square window of side nrebin*npix, centred
    u, v = flat % Ntot, flat // Ntot
    us, vs = (u - Ntot // 2) % Ntot, (v - Ntot // 2) % Ntot
    w = nrebin * npix
    assert us.max() - us.min() + 1 == w and vs.max() - vs.min() + 1 == w
    # each binned pixel is an nrebin x nrebin block, x index fast in p
    for p in range(npix * npix):
        ub, vb = us[p::npix * npix], vs[p::npix * npix]
        ub, vb = (bm[:, p] % Ntot - Ntot // 2) % Ntot, (bm[:, p] // Ntot - Ntot // 2) % Ntot
        assert ub.max() - ub.min() == nrebin - 1 and vb.max() - vb.min() == nrebin - 1
        assert (ub.min() - us.min()) // nrebin == p % npix
        assert (vb.min() - vs.min()) // nrebin == p // npix
    # product table identical
    class W:
        pass
    wfs = W()
    wfs._Ntot, wfs._nrebin, wfs.npix, wfs.gsalt = Ntot, nrebin, npix, 0.0
    assert np.array_equal(geom._binmap(wfs), bm)


def test_fft_goodsize_is_strictly_next_power_of_two():
    assert [og.fft_goodsize(s) for s in (15, 16, 17, 31, 32)] == [16, 32, 32, 32, 64]
    assert [geom.fft_goodsize(s) for s in (15, 16, 17, 31, 32)] == [16, 32, 32, 32, 64]


def test_sizing_rule_known_answers():
    """The sizing rule worked by hand (lambda 0.5 um, 0.5 m sub-apertures,
    1 rad = 206264.8 arcsec):
    * pdiam 16, 0.3 arcsec pixels: first rebin int(0.3 / 0.10313) + 1 = 3,
      Nfft = goodsize(int(32 * 3 / 0.3 * 0.10313)) = goodsize(33) = 64,
      quantum pixel (16 * 0.5e-6 / 0.5 rad) / 64 = 3.30024 / 64 = 0.0515662,
      ratio 5.82 -> 6 (fraction above 1/2), pixel 0.309397; 8 pixels
      (2.48 arcsec) fit in 64 quantum pixels: Ntot = Nfft;
    * pdiam 8, 1 arcsec pixels: first rebin 10, Nfft = goodsize(16) = 32,
      quantum pixel 1.65012 / 32 = 0.0515662 again, ratio 19.39 -> 19,
      pixel 0.979758; 8 pixels span 152 quantum pixels: Ntot =
      goodsize(153) = 256."""
    a = og.wfs_size(8.0, 16, 16, 8, 0.3, 0.5)
    assert (a["Nfft"], a["Ntot"], a["nrebin"]) == (64, 64, 6)
    assert a["pixsize"] == pytest.approx(0.309397, abs=2e-6)
    b = og.wfs_size(8.0, 16, 8, 8, 1.0, 0.5)
    assert (b["Nfft"], b["Ntot"], b["nrebin"]) == (32, 256, 19)
    assert b["pixsize"] == pytest.approx(0.979758, abs=2e-6)
    assert a["qpixsize"] == pytest.approx(0.0515662, abs=2e-7) == b["qpixsize"]


def test_sizing_rule_matches_oracle():
    class W:
        pass
    from shwfs.params import ParamTel, ParamWfs
    for (diam, nxsub, pdiam, npix, pix) in [(8.0, 16, 16, 8, 0.3), (4.0, 8, 8, 8, 0.5), (16.0, 80, 16, 8, 1.0)]:
        w = ParamWfs(nxsub=nxsub, npix=npix, pixsize=pix)
        w._pdiam = pdiam
        geom.init_wfs_size(w, 0.16, ParamTel(diam=diam))
        o = og.wfs_size(diam, nxsub, pdiam, npix, pix, 0.5)
        assert (w._Nfft, w._Ntot, w._nrebin) == (o["Nfft"], o["Ntot"], o["nrebin"])
        assert w.pixsize == pytest.approx(o["pixsize"], rel=1e-12)


def test_product_hrmap_equals_oracle():
    """The enlarged-grid map of the host geometry (shwfs/geom.py) equals the
    oracle's restatement (itself pinned as frequency-preserving zero padding,
    tests/test_oracle_wfs_ext.py)."""
    import numpy as np
    from oracle import geometry as og
    from shwfs import configs, system
    c = system.geometry(configs.get("wide"))
    w = c["p_wfs"]
    assert (w._Nfft, w._Ntot) == (16, 32)
    assert np.array_equal(w._hrmap, og.hrmap_table(16, 32))
    assert w._binmap.max() < 32 * 32
