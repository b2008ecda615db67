"""Parameter holders for the SH WFS path.

Field names follow shesha's config classes (shesha/shesha/config/pTel.py,
pGeom.py, pWfs.py, pCentroider.py) so code written against shesha reads the
same; only the fields the SH path reads exist.  Private fields (leading
underscore) are filled by ``geom`` exactly where shesha's init fills them.
"""


class ParamTel:
    def __init__(self, diam=8.0, cobs=0.0):
        self.diam = float(diam)
        self.cobs = float(cobs)
        self.t_spiders = -1.0
        self.spiders_type = None


class ParamGeom:
    def __init__(self, pupdiam=0, zenithangle=0.0):
        self.pupdiam = int(pupdiam)
        self.zenithangle = float(zenithangle)
        self.is_init = False
        self.ssize = 0
        self.cent = 0.0
        self._p1 = self._p2 = self._n = self._n1 = self._n2 = 0
        self._spupil = None
        self._mpupil = None
        self._ipupil = None
        self._pixsize = 0.0


class ParamWfs:
    def __init__(self, nxsub, npix, pixsize, Lambda=0.5, fracsub=0.5, noise=-1.0,
                 gsmag=0.0, zerop=1e11, optthroughput=1.0, xpos=0.0, ypos=0.0, kernel=0.0):
        self.type = "sh"
        self.nxsub = int(nxsub)
        self.npix = int(npix)
        self.pixsize = float(pixsize)
        self.Lambda = float(Lambda)
        self.fracsub = float(fracsub)
        self.noise = float(noise)
        self.gsmag = float(gsmag)
        self.zerop = float(zerop)
        self.optthroughput = float(optthroughput)
        self.xpos = float(xpos)
        self.ypos = float(ypos)
        self.gsalt = 0.0
        self.kernel = float(kernel)
        self.nphotons4imat = 1.0e5
        self.G = 1.0
        self.thetaML = 0.0
        self.dx = 0.0
        self.dy = 0.0
        self.fakecam = False
        self.max_flux_per_pix = 0
        self.max_pix_value = 0
        self.is_low_order = False
        self.nPupils = 1
        # laser guide star (gsalt > 0): altitude (m), launch telescope position
        # (m), beam FWHM (arcsec), return per watt and power, sodium profile
        self.lltx = 0.0
        self.llty = 0.0
        self.beamsize = 0.0
        self.lgsreturnperwatt = 0.0
        self.laserpower = 0.0
        self._altna = None
        self._profna = None
        self.fssize = 0.0       # field stop size (arcsec), 0: none
        self.fstop = "round"    # "round" or "square"
        self._submask = None
        # derived (geom)
        self._pdiam = 0
        self._Nfft = 0
        self._Ntot = 0
        self._nrebin = 0
        self._qpixsize = 0.0
        self._subapd = 0.0
        self._nvalid = 0
        self._nphotons = 0.0
        self._fluxPerSub = None
        self._isvalid = None
        self._validsubsx = None
        self._validsubsy = None
        self._validpuppixx = None
        self._validpuppixy = None
        self._phasemap = None
        self._ttprojmat = None
        self._halfxy = None
        self._hrmap = None
        self._binmap = None
        self._ftkernel = None


class ParamCentroider:
    def __init__(self, type="cog", thresh=0.0, width=None, type_fct="gauss", nwfs=0, nmax=10):
        self.type = type
        self.thresh = float(thresh)
        self.nmax = int(nmax)
        self.sizex = 3
        self.sizey = 3
        self.interpmat = None
        self.width = width
        self.type_fct = type_fct
        self.nwfs = nwfs
        self.filter_TT = False
        self.weights = None
        self._nslope = 0
