"""The named workloads of the SH WFS path.

Each entry fixes the telescope, the SH WFS (phase points per sub-aperture,
FFT support, binning and spot size imposed; the pixel scale follows as nrebin
quantum pixels), the phase screens and the centroider.

  small   D=4 m, 64 px pupil, 8x8 SH, pdiam 8, Nfft 16, 2x2 binning to 8x8 px,
          52 valid, noise off, CoG
  noisy   D=8 m, 640 px pupil, 40x40 SH, pdiam 16, Nfft 32, 4x4 to 8x8 px,
          1000 ph/subap Poisson + 0.5 e- readout, 64 screens (r0 0.12 m), WCoG
  large   D=16 m, 1280 px pupil, 80x80 SH, pdiam 16, Nfft 32, 4x4 to 8x8 px,
          noise off, TCoG (threshold 0.1 * max, float64 moments) and CoG
  batch   six 74x74 (pdiam 8, Nfft 16, 6x6 px), one 40x40 (pdiam 16, Nfft 32,
          8x8 px) and one 8x8 (pdiam 32, Nfft 64, 16x16 px) SH, each with its
          own pupil, screen and r0 in 0.10-0.20 m
  lowflux D=8 m, 40x40 SH, pdiam 16, Nfft 32, 8x8 px, 50 ph/subap Poisson +
          3 e- readout
  wide    D=4 m, 8x8 SH, pdiam 8, Nfft 16, 3x3 binning to 8x8 px: the 24-point
          detector window needs the enlarged 32-point HR grid (Ntot 32)
All: annular pupil (central obstruction 0.12), fraction threshold 0.5,
lambda 0.5 um, L0 25 m.
"""
import copy

from .params import ParamCentroider, ParamGeom, ParamTel, ParamWfs

L0_DEFAULT = 25.0


def sh_config(diam, nxsub, pdiam, npix, nfft, nrebin, cobs=0.12, fracsub=0.5, noise=-1.0,
              nphot=None, r0=0.16, L0=L0_DEFAULT, centroider="cog", thresh=0.0, width=None,
              seeds=(1,), name="sh"):
    return dict(
        name=name,
        p_tel=ParamTel(diam=diam, cobs=cobs),
        p_geom=ParamGeom(pupdiam=nxsub * pdiam),
        p_wfs=ParamWfs(nxsub=nxsub, npix=npix, pixsize=0.0, Lambda=0.5, fracsub=fracsub,
                       noise=noise, zerop=1e11, gsmag=0.0, optthroughput=1.0),
        p_centroider=ParamCentroider(type=centroider, thresh=thresh, width=width),
        nfft=nfft, nrebin=nrebin, nphot=nphot, r0=r0, L0=L0, seeds=list(seeds),
    )


def small():
    return sh_config(4.0, 8, 8, 8, 16, 2, r0=0.16, seeds=(1,), name="small")


def noisy():
    return sh_config(8.0, 40, 16, 8, 32, 4, noise=0.5, nphot=1000.0, r0=0.12,
                     centroider="wcog", width=4.0, seeds=range(1, 65), name="noisy")


def large():
    return sh_config(16.0, 80, 16, 8, 32, 4, r0=0.16, centroider="tcog", seeds=(1,),
                     name="large")


def lowflux():
    return sh_config(8.0, 40, 16, 8, 32, 4, noise=3.0, nphot=50.0, r0=0.16, seeds=(1,),
                     name="lowflux")


def batch():
    """Eight independent SH WFS of uneven size (one config per WFS)."""
    r0s = [0.10, 0.12, 0.14, 0.16, 0.18, 0.20, 0.15, 0.13]
    out = [sh_config(8.0, 74, 8, 6, 16, 2, r0=r0s[i], seeds=(i + 1,), name="b74_%d" % i)
           for i in range(6)]
    out.append(sh_config(8.0, 40, 16, 8, 32, 4, r0=r0s[6], seeds=(7,), name="b40"))
    out.append(sh_config(8.0, 8, 32, 16, 64, 4, r0=r0s[7], seeds=(8,), name="b8"))
    return out


def wide():
    return sh_config(4.0, 8, 8, 8, 16, 3, r0=0.16, seeds=(1,), name="wide")


def lgs(lltx=0.0, llty=0.0):
    """8x8 SH on an 8 m telescope with a sodium laser guide star at 90 km
    (one-Gaussian layer), beam FWHM 0.8 arcsec launched from (lltx, llty) m:
    pdiam 16, Nfft = Ntot 32, 4x4 binning to 8x8 px."""
    c = sh_config(8.0, 8, 16, 8, 32, 4, r0=0.16, seeds=(1,), name="lgs")
    w = c["p_wfs"]
    w.gsalt, w.beamsize, w.lltx, w.llty = 90000.0, 0.8, lltx, llty
    return c


WORKLOADS = {"small": small, "noisy": noisy, "large": large, "lowflux": lowflux, "wide": wide,
             "lgs": lgs}


def get(name):
    return copy.deepcopy(WORKLOADS[name]())
