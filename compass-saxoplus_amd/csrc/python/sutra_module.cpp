// pybind11 module "sutra": the SH WFS path of the reference's sutra module.
// Class and method names follow python_module/sutra/src/{telescope,source,
// wfs,wfs_sh,sensors,centroider,centroider_tcog,centroider_wcog,rtc}.cpp.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <type_traits>

#include "../sutra/sutra.hpp"

namespace py = pybind11;

template <typename T>
using ArrayFStyle = py::array_t<T, py::array::f_style | py::array::forcecast>;

using Centro = SutraCentroider<float, float>;
using Rtc = SutraRtc<float, float, float>;
using Ctrl = SutraController<float, float>;

static void declare_telescope(py::module& m) {
  py::class_<SutraTelescope>(m, "Telescope")
      .def(py::init([](CarmaContext& c, int64_t n_pup, int64_t npos, ArrayFStyle<float>& pupil,
                       int64_t n_pup_m, ArrayFStyle<float>& pupil_m) {
             if (pupil.size() != n_pup * n_pup || pupil_m.size() != n_pup_m * n_pup_m)
               throw std::runtime_error("Telescope: pupil sizes do not match n_pup / n_pup_m");
             return std::unique_ptr<SutraTelescope>(new SutraTelescope(
                 &c, n_pup, npos, pupil.mutable_data(), n_pup_m, pupil_m.mutable_data()));
           }),
           py::arg("context"), py::arg("n_pup"), py::arg("npos"), py::arg("pupil"),
           py::arg("n_pup_m"), py::arg("pupil_m"))
      .def_property_readonly("device", [](SutraTelescope& t) { return t.device; })
      .def_property_readonly("pup_size", [](SutraTelescope& t) { return t.pup_size; })
      .def_property_readonly("pup_size_m", [](SutraTelescope& t) { return t.pup_size_m; })
      .def_property_readonly("num_eleme_pup", [](SutraTelescope& t) { return t.num_eleme_pup; })
      .def_property_readonly("d_pupil", [](SutraTelescope& t) { return t.d_pupil; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_pupil_m", [](SutraTelescope& t) { return t.d_pupil_m; },
                             py::return_value_policy::reference_internal);
}

static void declare_lgs(py::module& m) {
  using C = hipFloatComplex;
  py::class_<SutraLGS>(m, "LGS")
      .def_property_readonly("device", [](SutraLGS& l) { return l.device; })
      .def_property_readonly("nvalid", [](SutraLGS& l) { return l.nvalid; })
      .def_property_readonly("npix", [](SutraLGS& l) { return l.npix; })
      .def_property_readonly("nprof", [](SutraLGS& l) { return l.nprof; })
      .def_property_readonly("nmaxhr", [](SutraLGS& l) { return l.nmaxhr; })
      .def_property_readonly("hg", [](SutraLGS& l) { return l.hg; })
      .def_property_readonly("h0", [](SutraLGS& l) { return l.h0; })
      .def_property_readonly("deltah", [](SutraLGS& l) { return l.deltah; })
      .def_property_readonly("pixsize", [](SutraLGS& l) { return l.pixsize; })
      .def_property_readonly("d_doffaxis", [](SutraLGS& l) { return l.d_doffaxis; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_azimuth", [](SutraLGS& l) { return l.d_azimuth; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_prof1d", [](SutraLGS& l) { return l.d_prof1d; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_profcum", [](SutraLGS& l) { return l.d_profcum; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_prof2d", [](SutraLGS& l) { return l.d_prof2d; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_beam", [](SutraLGS& l) { return l.d_beam; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_ftbeam", [](SutraLGS& l) { return l.d_ftbeam; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_lgskern", [](SutraLGS& l) { return l.d_lgskern; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_ftlgskern", [](SutraLGS& l) { return l.d_ftlgskern; },
                             py::return_value_policy::reference_internal)
      .def("lgs_init",
           [](SutraLGS& l, int32_t nprof, float hg, float h0, float deltah, float pixsize,
              ArrayFStyle<float>& doffaxis, ArrayFStyle<float>& prof1d,
              ArrayFStyle<float>& profcum, ArrayFStyle<float>& beam,
              py::array_t<std::complex<float>, py::array::f_style | py::array::forcecast>& ftbeam,
              ArrayFStyle<float>& azimuth) {
             if (doffaxis.size() != l.nvalid || azimuth.size() != l.nvalid)
               throw std::runtime_error("lgs_init: doffaxis and azimuth need nvalid values");
             if (prof1d.size() < nprof || profcum.size() < nprof + 1)
               throw std::runtime_error("lgs_init: prof1d needs nprof values, profcum nprof + 1");
             if (beam.size() != l.npix || ftbeam.size() != l.npix)
               throw std::runtime_error("lgs_init: beam and ftbeam need npix values");
             return l.lgs_init(nprof, hg, h0, deltah, pixsize, doffaxis.mutable_data(),
                               prof1d.mutable_data(), profcum.mutable_data(), beam.mutable_data(),
                               reinterpret_cast<C*>(ftbeam.mutable_data()),
                               azimuth.mutable_data());
           },
           py::arg("nprof"), py::arg("hg"), py::arg("h0"), py::arg("deltah"), py::arg("pixsize"),
           py::arg("doffaxis"), py::arg("prof1d"), py::arg("profcum"), py::arg("beam"),
           py::arg("ftbeam"), py::arg("azimuth"))
      .def("load_prof",
           [](SutraLGS& l, ArrayFStyle<float>& prof1d, ArrayFStyle<float>& profcum, float hg,
              float h0, float deltah) {
             if (prof1d.size() < l.nprof || profcum.size() < l.nprof + 1)
               throw std::runtime_error("load_prof: prof1d needs nprof values, profcum nprof + 1");
             return l.load_prof(prof1d.mutable_data(), profcum.mutable_data(), hg, h0, deltah);
           },
           py::arg("prof1d"), py::arg("profcum"), py::arg("hg"), py::arg("h0"), py::arg("deltah"))
      .def("lgs_update", [](SutraLGS& l) { return l.lgs_update(); })
      .def("lgs_makespot", [](SutraLGS& l, int32_t nin) { return l.lgs_makespot(nullptr, nin); },
           py::arg("nin") = 0)
      .def("load_kernels",
           [](SutraLGS& l, ArrayFStyle<float>& k) {
             if (k.size() != l.npix * l.npix * l.nvalid)
               throw std::runtime_error("load_kernels: npix * npix * nvalid values needed");
             return l.load_kernels(k.mutable_data());
           },
           py::arg("lgskern"));
}

static void declare_source(py::module& m) {
  py::class_<SutraSource>(m, "Source")
      .def_property_readonly("device", [](SutraSource& s) { return s.device; })
      .def_property_readonly("type", [](SutraSource& s) { return s.type; })
      .def_property_readonly("posx", [](SutraSource& s) { return s.posx; })
      .def_property_readonly("posy", [](SutraSource& s) { return s.posy; })
      .def_property_readonly("npts", [](SutraSource& s) { return s.npts; })
      .def_property_readonly("mag", [](SutraSource& s) { return s.mag; })
      .def_property_readonly("Lambda", [](SutraSource& s) { return s.lambda; })
      .def_property_readonly("zp", [](SutraSource& s) { return s.zp; })
      .def_property_readonly("scale", [](SutraSource& s) { return s.scale; })
      .def_property_readonly("lgs", [](SutraSource& s) { return s.lgs; })
      .def_property_readonly("d_lgs", [](SutraSource& s) { return s.d_lgs; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_phase",
                             [](SutraSource& s) {
                               s.render_layers();  // pending raytrace layers, if any
                               return s.d_phase->d_screen;
                             },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("xoff", [](SutraSource& s) { return s.xoff; })
      .def_property_readonly("yoff", [](SutraSource& s) { return s.yoff; })
      .def_property_readonly("G", [](SutraSource& s) { return s.G; })
      .def_property_readonly("thetaML", [](SutraSource& s) { return s.thetaML; })
      .def_property_readonly("dx", [](SutraSource& s) { return s.dx; })
      .def_property_readonly("dy", [](SutraSource& s) { return s.dy; })
      .def_property("lazy_raytrace", [](SutraSource& s) { return s.lazy_raytrace; },
                    [](SutraSource& s, bool v) { s.lazy_raytrace = v; },
                    "True: raytrace keeps the layers for the spot kernel's field stage")
      .def_property_readonly("npending", [](SutraSource& s) { return (int)s.pending.size(); })
      .def("add_layer", &SutraSource::add_layer, py::arg("type"), py::arg("idx"), py::arg("xoff"),
           py::arg("yoff"))
      .def("remove_layer", &SutraSource::remove_layer, py::arg("type"), py::arg("idx"))
      .def("raytrace",
           [](SutraSource& s, SutraAtmos& atmos, bool async) { return s.raytrace(&atmos, async); },
           py::arg("atmos"), py::arg("async") = false)
      .def("render_layers", [](SutraSource& s) { return s.render_layers(); })
      .def("reset_phase", &SutraSource::reset_phase)
      .def("set_phase",
           [](SutraSource& s, ArrayFStyle<float>& data) {
             if ((int64_t)data.size() != s.d_phase->d_screen->get_nb_elements())
               throw std::runtime_error("set_phase: wrong phase size");
             return s.set_phase(data.mutable_data());
           },
           py::arg("data"));
}

static void declare_atmos(py::module& m) {
  py::class_<SutraTurbuScreen>(m, "TurbuScreen")
      .def_property_readonly("device", [](SutraTurbuScreen& t) { return t.device; })
      .def_property_readonly("screen_size", [](SutraTurbuScreen& t) { return t.screen_size; })
      .def_property_readonly("r0", [](SutraTurbuScreen& t) { return t.r0; })
      .def_property_readonly("amplitude", [](SutraTurbuScreen& t) { return t.amplitude; })
      .def_property_readonly("altitude", [](SutraTurbuScreen& t) { return t.altitude; })
      .def_property_readonly("windspeed", [](SutraTurbuScreen& t) { return t.windspeed; })
      .def_property_readonly("winddir", [](SutraTurbuScreen& t) { return t.winddir; })
      .def_property_readonly("deltax", [](SutraTurbuScreen& t) { return t.deltax; })
      .def_property_readonly("deltay", [](SutraTurbuScreen& t) { return t.deltay; })
      .def_property_readonly("accumx", [](SutraTurbuScreen& t) { return t.accumx; })
      .def_property_readonly("accumy", [](SutraTurbuScreen& t) { return t.accumy; })
      .def_property_readonly("d_screen", [](SutraTurbuScreen& t) { return t.d_tscreen->d_screen; },
                             py::return_value_policy::reference_internal);

  py::class_<SutraAtmos>(m, "Atmos")
      .def(py::init([](CarmaContext& c, int32_t nscreens, float global_r0,
                       ArrayFStyle<float>& r0_per_layer, ArrayFStyle<int64_t>& size1,
                       ArrayFStyle<int64_t>& size2, ArrayFStyle<float>& altitude,
                       ArrayFStyle<float>& windspeed, ArrayFStyle<float>& winddir,
                       ArrayFStyle<float>& deltax, ArrayFStyle<float>& deltay, int32_t device) {
             for (auto n : {r0_per_layer.size(), size1.size(), size2.size(), altitude.size(),
                            windspeed.size(), winddir.size(), deltax.size(), deltay.size()})
               if (n < nscreens) throw std::runtime_error("Atmos: per-layer arrays shorter than nscreens");
             return std::unique_ptr<SutraAtmos>(new SutraAtmos(
                 &c, nscreens, global_r0, r0_per_layer.mutable_data(), size1.mutable_data(),
                 size2.mutable_data(), altitude.mutable_data(), windspeed.mutable_data(),
                 winddir.mutable_data(), deltax.mutable_data(), deltay.mutable_data(), device));
           }),
           py::arg("context"), py::arg("nscreens"), py::arg("global_r0"),
           py::arg("r0_per_layer"), py::arg("size1"), py::arg("size2"), py::arg("altitude"),
           py::arg("windspeed"), py::arg("winddir"), py::arg("deltax"), py::arg("deltay"),
           py::arg("device"))
      .def_property_readonly("nscreens", [](SutraAtmos& a) { return a.nscreens; })
      .def_property_readonly("r0", [](SutraAtmos& a) { return a.r0; })
      .def_property_readonly("d_screens", [](SutraAtmos& a) { return a.d_screens; },
                             py::return_value_policy::reference_internal);
}

static void declare_wfs(py::module& m) {
  py::class_<SutraWfs>(m, "Wfs")
      .def_property_readonly("device", [](SutraWfs& w) { return w.device; })
      .def_property_readonly("type", [](SutraWfs& w) { return w.type; })
      .def_property_readonly("is_low_order", [](SutraWfs& w) { return w.is_low_order; })
      .def_property_readonly("kernconv", [](SutraWfs& w) { return w.kernconv; })
      .def("set_kernconv", [](SutraWfs& w, bool k) { w.kernconv = k; }, py::arg("kernconv"),
           "convolve the HR image with d_ftkernel before binning (the rtc sets it for imats)")
      .def_property_readonly("d_hrmap", [](SutraWfs& w) { return w.d_hrmap; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_ftkernel", [](SutraWfs& w) { return w.d_ftkernel; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("nxsub", [](SutraWfs& w) { return w.nxsub; })
      .def_property_readonly("nvalid", [](SutraWfs& w) { return w.nvalid; })
      .def_property_readonly("npix", [](SutraWfs& w) { return w.npix; })
      .def_property_readonly("nrebin", [](SutraWfs& w) { return w.nrebin; })
      .def_property_readonly("nfft", [](SutraWfs& w) { return w.nfft; })
      .def_property_readonly("ntot", [](SutraWfs& w) { return w.ntot; })
      .def_property_readonly("npup", [](SutraWfs& w) { return w.npup; })
      .def_property_readonly("nphase", [](SutraWfs& w) { return w.nphase; })
      .def_property_readonly("nmaxhr", [](SutraWfs& w) { return w.nmaxhr; })
      .def_property_readonly("nffthr", [](SutraWfs& w) { return w.nffthr; })
      .def_property_readonly("subapd", [](SutraWfs& w) { return w.subapd; })
      .def_property_readonly("nphot", [](SutraWfs& w) { return w.nphot; })
      .def_property_readonly("nphot4imat", [](SutraWfs& w) { return w.nphot4imat; })
      .def_property_readonly("noise", [](SutraWfs& w) { return w.noise; })
      .def_property_readonly("lgs", [](SutraWfs& w) { return w.lgs; })
      .def_property_readonly("kernconv", [](SutraWfs& w) { return w.kernconv; })
      .def_property_readonly("fakecam", [](SutraWfs& w) { return w.fakecam; })
      .def_property_readonly("max_flux_per_pix", [](SutraWfs& w) { return w.max_flux_per_pix; })
      .def_property_readonly("max_pix_value", [](SutraWfs& w) { return w.max_pix_value; })
      .def_property_readonly("roket", [](SutraWfs& w) { return w.roket; })
#define RO_OBJ(name) \
  .def_property_readonly(#name, [](SutraWfs& w) { return w.name; }, py::return_value_policy::reference_internal)
          RO_OBJ(d_camplipup) RO_OBJ(d_camplifoc) RO_OBJ(d_fttotim) RO_OBJ(d_pupil)
              RO_OBJ(d_bincube) RO_OBJ(d_binimg) RO_OBJ(d_binimg_notnoisy) RO_OBJ(d_intensities)
                  RO_OBJ(d_offsets) RO_OBJ(d_fluxPerSub) RO_OBJ(d_sincar) RO_OBJ(d_hrmap)
                      RO_OBJ(d_camimg) RO_OBJ(d_dark) RO_OBJ(d_flat) RO_OBJ(d_slopes) RO_OBJ(d_gs)
                          RO_OBJ(d_phasemap) RO_OBJ(d_ttprojmat) RO_OBJ(d_ttprojvec)
                              RO_OBJ(d_validsubsx) RO_OBJ(d_validsubsy) RO_OBJ(d_submask)
#undef RO_OBJ
      .def("comp_image", &SutraWfs::comp_image, py::arg("noise") = true)
      .def("slopes_geom", [](SutraWfs& w, int32_t type) { return w.slopes_geom(type); },
           py::arg("type") = 0)
      .def("fill_binimage", &SutraWfs::fill_binimage, py::arg("async") = 0)
      .def("set_pupil",
           [](SutraWfs& w, ArrayFStyle<float>& data) {
             if ((int64_t)data.size() != w.d_pupil->get_nb_elements())
               throw std::runtime_error("set_pupil: wrong dimensions");
             return w.set_pupil(data.mutable_data());
           },
           py::arg("pupil"))
      .def("set_noise", &SutraWfs::set_noise, py::arg("noise"), py::arg("seed"))
      .def("set_fakecam", &SutraWfs::set_fakecam, py::arg("fakecam"))
      .def("set_max_flux_per_pix", &SutraWfs::set_max_flux_per_pix, py::arg("max_flux_per_pix"))
      .def("set_max_pix_value", &SutraWfs::set_max_pix_value, py::arg("max_pix_value"))
      .def("set_binimg",
           [](SutraWfs& w, ArrayFStyle<float>& d, int32_t n) { return w.set_binimg(d.mutable_data(), n); },
           py::arg("binimg"), py::arg("nElem"))
      .def("set_dark",
           [](SutraWfs& w, ArrayFStyle<float>& d, int32_t n) { return w.set_dark(d.mutable_data(), n); },
           py::arg("dark"), py::arg("nElem"))
      .def("set_flat",
           [](SutraWfs& w, ArrayFStyle<float>& d, int32_t n) { return w.set_flat(d.mutable_data(), n); },
           py::arg("flat"), py::arg("nElem"));
}

static void declare_wfs_sh(py::module& m) {
  py::class_<SutraWfsSH, SutraWfs>(m, "SHWFS")
      .def_property_readonly("d_binmap", [](SutraWfsSH& w) { return w.d_binmap; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_validpuppixx", [](SutraWfsSH& w) { return w.d_validpuppixx; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_validpuppixy", [](SutraWfsSH& w) { return w.d_validpuppixy; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_fsamplipup", [](SutraWfsSH& w) { return w.d_fsamplipup; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_fsamplifoc", [](SutraWfsSH& w) { return w.d_fsamplifoc; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("method", &SutraWfsSH::get_method)
      .def_property_readonly("wg_size", &SutraWfsSH::get_wg_size)
      .def_property_readonly("bin_block", [](SutraWfsSH& w) { return w.bin_block; })
      .def_property_readonly("bin_origin", [](SutraWfsSH& w) { return w.bin_origin; })
      .def_property_readonly("d_pupil_unit", [](SutraWfsSH& w) { return w.d_pupil_unit; },
                             py::return_value_policy::reference_internal)
      .def("set_wg_size", &SutraWfsSH::set_wg_size, py::arg("wg_size"),
           "threads per workgroup of the spot kernel: 0 (default), 64, 128, 256")
      .def_property_readonly("noise_call", [](SutraWfsSH& w) { return w.noise_call; })
      .def("set_method", &SutraWfsSH::set_method, py::arg("method"),
           "0: FP32 FFT, 1: MFMA DFT split-BF16x3, 2: MFMA DFT FP16 with per-subap scale")
      .def("load_arrays",
           [](SutraWfsSH& w, ArrayFStyle<int32_t>& phasemap, ArrayFStyle<int32_t>& hrmap,
              ArrayFStyle<int32_t>& binmap, ArrayFStyle<float>& offsets,
              ArrayFStyle<float>& fluxPerSub, ArrayFStyle<int32_t>& validsubsx,
              ArrayFStyle<int32_t>& validsubsy, ArrayFStyle<int32_t>& istart,
              ArrayFStyle<int32_t>& jstart, ArrayFStyle<float>& ttprojmat,
              ArrayFStyle<std::complex<float>>& kernel) {
             const int64_t nv = w.nvalid_tot, np2 = w.nphase * w.nphase;
             if (phasemap.size() < np2 * w.nvalid || binmap.size() != w.nrebin * w.nrebin * w.npix * w.npix ||
                 offsets.size() != np2 || fluxPerSub.size() < nv || validsubsx.size() < nv ||
                 validsubsy.size() < nv || istart.size() < w.nvalid || jstart.size() < w.nvalid ||
                 ttprojmat.size() != np2 * 2 * w.nvalid || kernel.size() != w.ntot * w.ntot ||
                 (w.ntot != w.nfft && hrmap.size() != w.nfft * w.nfft))
               throw std::runtime_error("load_arrays: array sizes do not match the WFS");
             return w.load_arrays(phasemap.mutable_data(), hrmap.mutable_data(),
                                  binmap.mutable_data(), offsets.mutable_data(),
                                  fluxPerSub.mutable_data(), validsubsx.mutable_data(),
                                  validsubsy.mutable_data(), istart.mutable_data(),
                                  jstart.mutable_data(), ttprojmat.mutable_data(),
                                  reinterpret_cast<hipFloatComplex*>(kernel.mutable_data()));
           },
           py::arg("phasemap"), py::arg("hrmap"), py::arg("binmap"), py::arg("offsets"),
           py::arg("fluxPerSub"), py::arg("validsubsx"), py::arg("validsubsy"), py::arg("istart"),
           py::arg("jstart"), py::arg("ttprojmat"), py::arg("kernel"))
      .def("comp_nphot", &SutraWfsSH::comp_nphot, py::arg("ittime"), py::arg("optthroughput"),
           py::arg("diam"), py::arg("nxsub"), py::arg("zerop") = 0, py::arg("gsmag") = 0,
           py::arg("lgsreturnperwatt") = 0, py::arg("laserpower") = 0)
      .def("set_bincube",
           [](SutraWfsSH& w, ArrayFStyle<float>& b, int32_t n) { return w.set_bincube(b.mutable_data(), n); },
           py::arg("bincube"), py::arg("nElem"));
}

static void declare_sensors(py::module& m) {
  py::class_<SutraSensors>(m, "Sensors")
      .def(py::init([](CarmaContext& context, SutraTelescope* d_tel, vector<string> type,
                       int32_t nwfs, ArrayFStyle<int64_t>& nxsub, ArrayFStyle<int64_t>& nvalid,
                       ArrayFStyle<int64_t>& npupils, ArrayFStyle<int64_t>& npix,
                       ArrayFStyle<int64_t>& nphase, ArrayFStyle<int64_t>& nrebin,
                       ArrayFStyle<int64_t>& nfft, ArrayFStyle<int64_t>& ntot,
                       ArrayFStyle<int64_t>& npup, ArrayFStyle<float>& pdiam,
                       ArrayFStyle<float>& nphot, ArrayFStyle<float>& nphot4imat,
                       ArrayFStyle<int32_t>& lgs, ArrayFStyle<bool>& fakecam,
                       ArrayFStyle<int32_t>& max_flux_per_pix, ArrayFStyle<int32_t>& max_pix_value,
                       int32_t device, bool roket) {
             for (auto* a : {&nxsub, &nvalid, &npupils, &npix, &nphase, &nrebin, &nfft, &ntot, &npup})
               if (a->size() < nwfs) throw std::runtime_error("Sensors: per-WFS arrays shorter than nwfs");
             if (pdiam.size() < nwfs || nphot.size() < nwfs || nphot4imat.size() < nwfs ||
                 lgs.size() < nwfs || fakecam.size() < nwfs || max_flux_per_pix.size() < nwfs ||
                 max_pix_value.size() < nwfs)
               throw std::runtime_error("Sensors: per-WFS arrays shorter than nwfs");
             return std::unique_ptr<SutraSensors>(new SutraSensors(
                 &context, d_tel, type, nwfs, nxsub.mutable_data(), nvalid.mutable_data(),
                 npupils.mutable_data(), npix.mutable_data(), nphase.mutable_data(),
                 nrebin.mutable_data(), nfft.mutable_data(), ntot.mutable_data(),
                 npup.mutable_data(), pdiam.mutable_data(), nphot.mutable_data(),
                 nphot4imat.mutable_data(), lgs.mutable_data(), fakecam.mutable_data(),
                 max_flux_per_pix.mutable_data(), max_pix_value.mutable_data(), device, roket));
           }),
           py::arg("context"), py::arg("d_tel"), py::arg("type"), py::arg("nwfs"),
           py::arg("nxsub"), py::arg("nvalid"), py::arg("npupils"), py::arg("npix"),
           py::arg("nphase"), py::arg("nrebin"), py::arg("nfft"), py::arg("ntot"), py::arg("npup"),
           py::arg("pdiam"), py::arg("nphot"), py::arg("nphot4imat"), py::arg("lgs"),
           py::arg("fakecam"), py::arg("max_flux_per_pix"), py::arg("max_pix_value"),
           py::arg("device"), py::arg("roket"), py::keep_alive<1, 3>())
      .def_property_readonly("device", [](SutraSensors& s) { return s.device; })
      .def_property_readonly("roket", [](SutraSensors& s) { return s.roket; })
      .def_property_readonly("nsensors", [](SutraSensors& s) { return s.nsensors(); })
      .def_property_readonly(
          "d_wfs",
          [](SutraSensors& s) {
            py::list l;
            for (auto* w : s.d_wfs) {
              SutraWfsSH* sh = dynamic_cast<SutraWfsSH*>(w);
              l.append(py::cast(sh, py::return_value_policy::reference));
            }
            return l;
          })
      .def_property_readonly("d_camplipup", [](SutraSensors& s) { return s.d_camplipup; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_camplifoc", [](SutraSensors& s) { return s.d_camplifoc; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_fttotim", [](SutraSensors& s) { return s.d_fttotim; },
                             py::return_value_policy::reference_internal)
      .def("initgs",
           [](SutraSensors& s, ArrayFStyle<float>& xpos, ArrayFStyle<float>& ypos,
              ArrayFStyle<float>& lambda, ArrayFStyle<float>& mag, float zerop,
              ArrayFStyle<int64_t>& size, ArrayFStyle<float>& noise, ArrayFStyle<int64_t>& seed,
              ArrayFStyle<float>& G, ArrayFStyle<float>& thetaML, ArrayFStyle<float>& dx,
              ArrayFStyle<float>& dy) {
             const py::ssize_t n = (py::ssize_t)s.nsensors();
             for (py::ssize_t v : {xpos.size(), ypos.size(), lambda.size(), mag.size(), size.size(),
                                   noise.size(), seed.size(), G.size(), thetaML.size(), dx.size(),
                                   dy.size()})
               if (v < n) throw std::runtime_error("initgs: per-WFS arrays shorter than nsensors");
             return s.initgs(xpos.mutable_data(), ypos.mutable_data(), lambda.mutable_data(),
                             mag.mutable_data(), zerop, size.mutable_data(), noise.mutable_data(),
                             seed.mutable_data(), G.mutable_data(), thetaML.mutable_data(),
                             dx.mutable_data(), dy.mutable_data());
           },
           py::arg("xpos"), py::arg("ypos"), py::arg("lambda_um"), py::arg("mag"), py::arg("zerop"),
           py::arg("sizes"), py::arg("noise"), py::arg("seeds"), py::arg("G"), py::arg("thetaML"),
           py::arg("dx"), py::arg("dy"))
      .def("set_noise", &SutraSensors::set_noise, py::arg("nwfs"), py::arg("noise"), py::arg("seed"))
      .def("set_field_stop",
           [](SutraSensors& s, int32_t nwfs, ArrayFStyle<float>& fs, int32_t N) {
             if (fs.size() != (py::ssize_t)N * N)
               throw std::runtime_error("set_field_stop: the mask must hold N * N values");
             return s.set_field_stop(nwfs, fs.mutable_data(), N);
           },
           py::arg("nwfs"), py::arg("field_stop"), py::arg("N"));
}

// Centroider classes for images of Tin (float: _FF, uint16 camera frames: _UF;
// the uint16 image is calibrated into the float image by calibrate_img).
template <typename Tin>
static void declare_centroiders(py::module& m, const std::string& sfx) {
  using Centro = SutraCentroider<Tin, float>;
  py::class_<Centro>(m, ("Centroider" + sfx).c_str())
      .def_property_readonly("device", [](Centro& c) { return c.device; })
      .def_property_readonly("type", [](Centro& c) { return c.get_type(); })
      .def_property_readonly("nslopes", [](Centro& c) { return c.nslopes; })
      .def_property_readonly("wfs", [](Centro& c) { return dynamic_cast<SutraWfsSH*>(c.wfs); },
                             py::return_value_policy::reference)
      .def_property_readonly("nvalid", [](Centro& c) { return c.nvalid; })
      .def_property_readonly("nxsub", [](Centro& c) { return c.nxsub; })
      .def_property_readonly("npix", [](Centro& c) { return c.npix; })
      .def_property_readonly("offset", [](Centro& c) { return c.offset; })
      .def_property_readonly("scale", [](Centro& c) { return c.scale; })
      .def_property_readonly("filter_TT", [](Centro& c) { return c.filter_TT; })
      .def_property_readonly("fp64_moments", [](Centro& c) { return c.fp64_moments; })
#define RO_C(name) \
  .def_property_readonly(#name, [](Centro& c) { return c.name; }, py::return_value_policy::reference_internal)
          RO_C(d_bincube) RO_C(d_intensities) RO_C(d_centroids_ref) RO_C(d_img) RO_C(d_img_raw)
              RO_C(d_validx) RO_C(d_validy) RO_C(d_dark) RO_C(d_flat) RO_C(d_lutPix)
                  RO_C(d_validMask) RO_C(d_ref_Tip) RO_C(d_ref_Tilt) RO_C(d_TT_slopes)
#undef RO_C
      .def("get_cog", [](Centro& c) { return c.get_cog(); })
      .def("load_validpos",
           [](Centro& c, ArrayFStyle<int32_t>& vx, ArrayFStyle<int32_t>& vy, int32_t N) {
             if (vx.size() < N || vy.size() < N) throw std::runtime_error("load_validpos: arrays shorter than N");
             return c.load_validpos(vx.mutable_data(), vy.mutable_data(), N);
           },
           py::arg("validx"), py::arg("validy"), py::arg("N"))
      .def("set_npix", &Centro::set_npix, py::arg("npix"))
      .def("set_nxsub", &Centro::set_nxsub, py::arg("nxsub"))
      .def("set_fp64_moments", &Centro::set_fp64_moments, py::arg("value"),
           "sum flux and first moments in float64")
      .def("load_img",
           [](Centro& c, ArrayFStyle<Tin>& img, int32_t m_, int32_t n, int32_t loc) {
             if (img.size() != (py::ssize_t)m_ * n) throw std::runtime_error("load_img: wrong size");
             return c.load_img(img.mutable_data(), m_, n, loc);
           },
           py::arg("img"), py::arg("m"), py::arg("n"), py::arg("location"))
      .def("load_img",
           [](Centro& c, ArrayFStyle<Tin>& img, int32_t n, int32_t loc) {
             if (img.size() != (py::ssize_t)n * n) throw std::runtime_error("load_img: wrong size");
             return c.load_img(img.mutable_data(), n, loc);
           },
           py::arg("img"), py::arg("n"), py::arg("location") = -1)
      .def("load_img",
           [](Centro& c, CarmaObj<Tin>* img) { return c.load_img(img); }, py::arg("img"))
      .def("calibrate_img", [](Centro& c) { return c.calibrate_img(); })
      .def("set_centroids_ref",
           [](Centro& c, ArrayFStyle<float>& r) {
             if (r.size() != c.nslopes) throw std::runtime_error("set_centroids_ref: wrong size");
             return c.set_centroids_ref(r.mutable_data());
           },
           py::arg("refslopes"))
      .def_property(
          "slope_order", [](Centro& c) { return (int)c.slope_order; },
          [](Centro& c, int o) {
            if (o != 0 && o != 1) throw std::invalid_argument("slope_order: 0 (untied) or 1 (interlaced)");
            c.slope_order = static_cast<SlopeOrder>(o);
          },
          "slope layout (the C++ member of libsutra/include/sutra_centroider.hpp:31-53): 0 untied "
          "(x slopes then y slopes), 1 interlaced (x, y per sub-aperture)")
      .def("set_scale", &Centro::set_scale, py::arg("scale"))
      .def("set_offset", &Centro::set_offset, py::arg("offset"))
      .def("set_dark",
           [](Centro& c, ArrayFStyle<float>& d, int32_t n) { return c.set_dark(d.mutable_data(), n); },
           py::arg("dark"), py::arg("n"))
      .def("set_flat",
           [](Centro& c, ArrayFStyle<float>& d, int32_t n) { return c.set_flat(d.mutable_data(), n); },
           py::arg("flat"), py::arg("n"))
      .def("init_calib", &Centro::init_calib, py::arg("n"), py::arg("m"))
      .def("init_img_raw", &Centro::init_img_raw, py::arg("n"), py::arg("m"))
      .def("set_lutPix",
           [](Centro& c, ArrayFStyle<int32_t>& l, int32_t n) { return c.set_lutPix(l.mutable_data(), n); },
           py::arg("lutPix"), py::arg("n"));

  using Cog = SutraCentroiderCog<Tin, float>;
  py::class_<Cog, Centro>(m, ("CentroiderCOG" + sfx).c_str());

  using Tcog = SutraCentroiderTcog<Tin, float>;
  py::class_<Tcog, Centro>(m, ("CentroiderTCOG" + sfx).c_str())
      .def_property_readonly("threshold", [](Tcog& c) { return c.threshold; })
      .def("set_threshold", &Tcog::set_threshold, py::arg("thresh"));

  using Wcog = SutraCentroiderWcog<Tin, float>;
  py::class_<Wcog, Centro>(m, ("CentroiderWCOG" + sfx).c_str())
      .def_property_readonly("threshold", [](Wcog& c) { return c.threshold; })
      .def_property_readonly("d_weights", [](Wcog& c) { return c.d_weights; },
                             py::return_value_policy::reference_internal)
      .def("init_weights", &Wcog::init_weights)
      .def("load_weights",
           [](Wcog& c, ArrayFStyle<float>& w, int32_t ndim) {
             const py::ssize_t need = (py::ssize_t)c.npix * c.npix * (ndim == 3 ? c.nvalid : 1);
             if (w.size() != need) throw std::runtime_error("load_weights: wrong size");
             return c.load_weights(w.mutable_data(), ndim);
           },
           py::arg("weights"), py::arg("ndim"))
      .def("set_threshold", &Wcog::set_threshold, py::arg("thresh"));

  using Bpcog = SutraCentroiderBpcog<Tin, float>;
  py::class_<Bpcog, Centro>(m, ("CentroiderBPCOG" + sfx).c_str())
      .def_property_readonly("nmax", [](Bpcog& c) { return c.nmax; })
      .def_property_readonly("d_bpix", [](Bpcog& c) { return c.d_bpix; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_bpind", [](Bpcog& c) { return c.d_bpind; },
                             py::return_value_policy::reference_internal)
      .def("set_nmax", &Bpcog::set_nmax, py::arg("nmax"),
           "number of brightest pixels kept per sub-aperture");

  using Corr = SutraCentroiderCorr<Tin, float>;
  py::class_<Corr, Centro>(m, ("CentroiderCORR" + sfx).c_str())
      .def_property_readonly("npix", [](Corr& c) { return c.npix; })
      .def_property_readonly("interp_sizex", [](Corr& c) { return c.interp_sizex; })
      .def_property_readonly("interp_sizey", [](Corr& c) { return c.interp_sizey; })
      .def_property_readonly("d_corrfnct", [](Corr& c) { return c.d_corrfnct; },
                             py::return_value_policy::reference_internal,
                             "reference function(s) in pixels [npix, npix, nvalid]")
      .def_property_readonly("d_corrspot", [](Corr& c) { return c.d_corrspot; },
                             py::return_value_policy::reference_internal,
                             "None: the correlation is summed without spot transforms")
      .def_property_readonly("d_corrnorm", [](Corr& c) { return c.d_corrnorm; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_corrmax", [](Corr& c) { return c.d_corrmax; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_corr", [](Corr& c) { return c.d_corr; },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("d_interpmat", [](Corr& c) { return c.d_interpmat; },
                             py::return_value_policy::reference_internal)
      .def("init_corr",
           [](Corr& c, int32_t isizex, int32_t isizey, ArrayFStyle<float>& interpmat) {
             if (interpmat.size() != (py::ssize_t)isizex * isizey * 6)
               throw std::runtime_error("init_corr: interpmat must hold isizex * isizey * 6 values");
             return c.init_corr(isizex, isizey, interpmat.mutable_data());
           },
           py::arg("isizex"), py::arg("isizey"), py::arg("interpmat"))
      .def("load_corr",
           [](Corr& c, ArrayFStyle<float>& corr, ArrayFStyle<float>& corr_norm, int32_t ndim) {
             const py::ssize_t need = (py::ssize_t)c.npix * c.npix * (ndim == 3 ? c.nvalid : 1);
             const py::ssize_t nn = (py::ssize_t)(2 * c.npix - 1) * (2 * c.npix - 1);
             if (corr.size() != need) throw std::runtime_error("load_corr: wrong corr size");
             if (corr_norm.size() < nn)
               throw std::runtime_error("load_corr: corr_norm needs (2 npix - 1)^2 values");
             return c.load_corr(corr.mutable_data(), corr_norm.mutable_data(), ndim);
           },
           py::arg("corr"), py::arg("corr_norm"), py::arg("ndim"));

  if constexpr (std::is_same<Tin, float>::value) {  // the learned centroider takes float images
    using Mlp = SutraCentroiderMlp<float, float>;
    py::class_<Mlp, Centro>(m, "CentroiderMLP_FF")
        .def_property_readonly("d_params", [](Mlp& c) { return c.d_params; },
                               py::return_value_policy::reference_internal)
        .def_property_readonly("d_grad", [](Mlp& c) { return c.d_grad; },
                               py::return_value_policy::reference_internal)
        .def_property_readonly("d_fallback", [](Mlp& c) { return c.d_fallback; },
                               py::return_value_policy::reference_internal)
        .def_property_readonly("x_scale", [](Mlp& c) { return c.x_scale; })
        .def_property_readonly("flux_range", [](Mlp& c) { return py::make_tuple(c.flux_lo, c.flux_hi); })
        .def_property_readonly("step", [](Mlp& c) { return c.step; })
        .def("init_params", &Mlp::init_params, py::arg("seed"))
        .def("set_input_scale", &Mlp::set_input_scale, py::arg("scale"))
        .def("set_flux_range", &Mlp::set_flux_range, py::arg("lo"), py::arg("hi"))
        .def("loss_grad", &Mlp::loss_grad, py::arg("x"), py::arg("t"), py::arg("n"),
             py::arg("t_scale") = 1.0f, "mean loss; the gradient is left in d_grad")
        .def("forward", &Mlp::forward, py::arg("x"), py::arg("n"), py::arg("y"))
        .def("train",
             [](Mlp& c, CarmaObj<float>* x, CarmaObj<float>* t, int64_t n, float t_scale,
                int32_t batch, int32_t epochs, float lr, uint64_t seed) {
               py::gil_scoped_release nogil;
               return c.train(x, t, n, t_scale, batch, epochs, lr, seed);
             },
             py::arg("x"), py::arg("t"), py::arg("n"), py::arg("t_scale") = 1.0f,
             py::arg("batch") = 4096, py::arg("epochs") = 1, py::arg("lr") = 1e-3f,
             py::arg("seed") = 1u, "Adam epochs over shuffled batches; last epoch's mean loss");
  }
}

static void declare_controller(py::module& m) {
  py::class_<Ctrl>(m, "Controller_FF")
      .def_property_readonly("type", [](Ctrl& c) { return c.type; })
      .def_property_readonly("nslope", [](Ctrl& c) { return c.nslope; })
      .def_property_readonly("nactu", [](Ctrl& c) { return c.nactu; })
      .def_property_readonly("delay", [](Ctrl& c) { return c.delay; })
      .def_property_readonly("gain", [](Ctrl& c) { return c.gain; })
      .def_property_readonly("centro_idx", [](Ctrl& c) { return c.centro_idx; })
      .def_property_readonly("d_centroids", [](Ctrl& c) { return c.d_centroids; },
                             py::return_value_policy::reference_internal);
}

// RTC for images of Tin (Rtc_FFF, Rtc_UFF).
template <typename Tin>
static void declare_rtc(py::module& m, const char* name) {
  using Rtc = SutraRtc<Tin, float, float>;
  py::class_<Rtc>(m, name)
      .def(py::init([]() { return std::unique_ptr<Rtc>(new Rtc()); }))
      .def_property_readonly(
          "d_centro",
          [](Rtc& r) {
            py::list l;
            for (auto* c : r.d_centro) {
              if (auto* b = dynamic_cast<SutraCentroiderBpcog<Tin, float>*>(c))
                l.append(py::cast(b, py::return_value_policy::reference));
              else if (auto* k = dynamic_cast<SutraCentroiderCorr<Tin, float>*>(c))
                l.append(py::cast(k, py::return_value_policy::reference));
              else if (auto* t = dynamic_cast<SutraCentroiderTcog<Tin, float>*>(c))
                l.append(py::cast(t, py::return_value_policy::reference));
              else if (auto* w = dynamic_cast<SutraCentroiderWcog<Tin, float>*>(c))
                l.append(py::cast(w, py::return_value_policy::reference));
              else if (auto* cg = dynamic_cast<SutraCentroiderCog<Tin, float>*>(c))
                l.append(py::cast(cg, py::return_value_policy::reference));
              else
                l.append(py::cast(c, py::return_value_policy::reference));
            }
            return l;
          })
      .def_property_readonly(
          "d_control",
          [](Rtc& r) {
            py::list l;
            for (auto* c : r.d_control) l.append(py::cast(c, py::return_value_policy::reference));
            return l;
          })
      .def("add_centroider",
           [](Rtc& r, CarmaContext& ctx, int64_t nvalid, float offset, float scale, bool filter_TT,
              int64_t device, std::string typec, SutraWfs* wfs) {
             return r.add_centroider(&ctx, nvalid, offset, scale, filter_TT, device, typec, wfs);
           },
           py::arg("context"), py::arg("nvalid"), py::arg("offset"), py::arg("scale"),
           py::arg("filter_TT"), py::arg("device"), py::arg("typec"), py::arg("wfs"),
           py::keep_alive<1, 9>())
      .def("add_centroider",
           [](Rtc& r, CarmaContext& ctx, int64_t nvalid, float offset, float scale, bool filter_TT,
              int64_t device, std::string typec) {
             return r.add_centroider(&ctx, nvalid, offset, scale, filter_TT, device, typec);
           },
           py::arg("context"), py::arg("nvalid"), py::arg("offset"), py::arg("scale"),
           py::arg("filter_TT"), py::arg("device"), py::arg("typec"))
      .def("add_controller",
           [](Rtc& r, CarmaContext& ctx, std::string typec, int64_t device, float delay,
              int32_t nslope, int32_t nactu, py::kwargs kw) {
             std::vector<int32_t> idx;
             int32_t ncentro = 0;
             if (kw.contains("idx_centro")) {
               auto a = py::array_t<int32_t, py::array::forcecast>::ensure(kw["idx_centro"]);
               if (!a) throw std::runtime_error("add_controller: idx_centro must be an integer array");
               idx.assign(a.data(), a.data() + a.size());
             }
             if (kw.contains("ncentro")) ncentro = kw["ncentro"].cast<int32_t>();
             if ((int32_t)idx.size() < ncentro) {
               for (int32_t i = (int32_t)idx.size(); i < ncentro; ++i) idx.push_back(i);
             }
             for (int32_t i : idx)
               if (i < 0) throw std::runtime_error("add_controller: negative idx_centro");
             return r.add_controller(&ctx, typec, device, delay, nslope, nactu, idx.data(), ncentro);
           },
           py::arg("context"), py::arg("typec"), py::arg("device"), py::arg("delay"),
           py::arg("nslope"), py::arg("nactu"))
      .def("remove_centroider", &Rtc::remove_centroider, py::arg("ncentro"))
      .def("do_calibrate_img", &Rtc::do_calibrate_img, py::arg("ncontrol"))
      .def("do_centroids", [](Rtc& r) { return r.do_centroids(); })
      .def("do_centroids", [](Rtc& r, int32_t n) { return r.do_centroids(n); }, py::arg("ncontrol"))
      .def("do_centroids", [](Rtc& r, int32_t n, bool noise) { return r.do_centroids(n, noise); },
           py::arg("ncontrol"), py::arg("noise"));
}

PYBIND11_MODULE(sutra, m) {
  m.doc() = "sutra: Shack-Hartmann WFS hot path on MI355X (gfx950)";
  py::module::import("carma");  // registers the CarmaObj / CarmaContext types
  declare_telescope(m);
  declare_atmos(m);
  declare_lgs(m);
  declare_source(m);
  declare_wfs(m);
  declare_wfs_sh(m);
  declare_sensors(m);
  declare_centroiders<float>(m, "_FF");
  declare_centroiders<uint16_t>(m, "_UF");
  declare_controller(m);
  declare_rtc<float>(m, "Rtc_FFF");
  declare_rtc<uint16_t>(m, "Rtc_UFF");
  m.def("comp_frame",
        [](SutraWfsSH& w, Centro& c, CarmaObj<float>& slopes, bool noise) {
          if (slopes.get_nb_elements() < 2 * w.nvalid) throw std::runtime_error("comp_frame: slope vector too short");
          return comp_frame(&w, &c, c.d_intensities->get_data(), slopes.get_data(), noise);
        },
        py::arg("wfs"), py::arg("centroider"), py::arg("slopes"), py::arg("noise") = true,
        "spots + slopes of one frame in one launch");
  m.def("run_frames",
        [](SutraWfsSH& w, Centro& c, CarmaObj<float>& slopes, int32_t n, bool noise, bool fused) {
          if (slopes.get_nb_elements() < 2 * w.nvalid) throw std::runtime_error("run_frames: slope vector too short");
          if (n < 0) throw std::runtime_error("run_frames: negative frame count");
          py::gil_scoped_release nogil;
          return run_frames(&w, &c, slopes.get_data(), n, noise, fused);
        },
        py::arg("wfs"), py::arg("centroider"), py::arg("slopes"), py::arg("nframes"),
        py::arg("noise") = true, py::arg("fused") = true,
        "issues nframes frames back to back; returns hipEvent-timed milliseconds");
  m.def("run_frames_graph",
        [](SutraWfsSH& w, Centro& c, CarmaObj<float>& slopes, int32_t n, int32_t reps, bool noise) {
          if (slopes.get_nb_elements() < 2 * w.nvalid) throw std::runtime_error("run_frames_graph: slope vector too short");
          py::gil_scoped_release nogil;
          return run_frames_graph(&w, &c, slopes.get_data(), n, reps, noise);
        },
        py::arg("wfs"), py::arg("centroider"), py::arg("slopes"), py::arg("nframes"),
        py::arg("reps"), py::arg("noise") = false,
        "fused frames (noise-free by default) captured in a hipGraph of nframes launches, "
        "replayed reps times; returns hipEvent-timed milliseconds of the reps replays");
  py::class_<SutraAcquisim>(m, "Acquisim")
      .def(py::init([](SutraSensors& sensors, int32_t wfs_num) {
             return new SutraAcquisim(&sensors, wfs_num);
           }),
           py::arg("sensors"), py::arg("wfs_num"), py::keep_alive<1, 2>(),
           "spots of an SH WFS from an external detector image")
      .def_property_readonly("npix", [](SutraAcquisim& a) { return a.npix; })
      .def_property_readonly("nxsub", [](SutraAcquisim& a) { return a.nxsub; })
      .def_property_readonly("nvalid", [](SutraAcquisim& a) { return a.nvalid; })
      .def("set_validsubs",
           [](SutraAcquisim& a, ArrayFStyle<int32_t>& vx, ArrayFStyle<int32_t>& vy) {
             if (vx.size() != vy.size()) throw std::runtime_error("set_validsubs: sizes differ");
             return a.set_validsubs(vx.size(), vx.mutable_data(), vy.mutable_data());
           },
           py::arg("validsubsx"), py::arg("validsubsy"))
      .def("comp_image",
           [](SutraAcquisim& a, ArrayFStyle<float>& img) {
             if (img.ndim() != 2) throw std::runtime_error("comp_image: a 2-D image is expected");
             int64_t dims[3] = {2, img.shape(0), img.shape(1)};
             return a.comp_image(dims, img.mutable_data());
           },
           py::arg("image"), "gathers the spots into the WFS's bincube")
      .def("comp_image_2D",
           [](SutraAcquisim& a, ArrayFStyle<float>& img, ArrayFStyle<int32_t>& num) {
             if (img.ndim() != 2) throw std::runtime_error("comp_image_2D: a 2-D image is expected");
             if (num.size() != a.nxsub * a.nxsub)
               throw std::runtime_error("comp_image_2D: num_ssp must have nxsub^2 entries");
             int64_t dims[3] = {2, img.shape(0), img.shape(1)};
             return a.comp_image_2D(dims, img.mutable_data(), num.mutable_data());
           },
           py::arg("image"), py::arg("num_ssp"));
  py::class_<SutraWfsBatch>(m, "WfsBatch")
      .def(py::init([](CarmaContext& ctx, std::vector<SutraWfsSH*> wfs,
                       std::vector<Centro*> centro, bool noise) {
             return new SutraWfsBatch(&ctx, wfs, centro, noise);
           }),
           py::arg("context"), py::arg("wfs"), py::arg("centroiders") = std::vector<Centro*>(),
           py::arg("noise") = false, py::keep_alive<1, 3>(), py::keep_alive<1, 4>(),
           "several SH WFS (and their fused centroiders, None allowed) computed by one "
           "launch per frame")
      .def_property_readonly("d_slopes",
                             [](SutraWfsBatch& b) {
                               py::list l;
                               for (auto* s : b.d_slopes)
                                 l.append(py::cast(s, py::return_value_policy::reference_internal,
                                                   py::cast(&b)));
                               return l;
                             })
      .def_property_readonly("methods", &SutraWfsBatch::methods)
      .def_property_readonly("splits", &SutraWfsBatch::splits)
      .def_property_readonly("grid", &SutraWfsBatch::grid)
      .def_property_readonly("remap", &SutraWfsBatch::remap)
      .def_property_readonly("frame", &SutraWfsBatch::frame)
      .def("set_tiles", &SutraWfsBatch::set_tiles, py::arg("methods"), py::arg("splits"),
           py::arg("grid") = 0, py::arg("remap") = 0)
      .def("set_frame", &SutraWfsBatch::set_frame, py::arg("frame"))
      .def("autotune",
           [](SutraWfsBatch& b, uint32_t mask, int32_t reps) {
             py::gil_scoped_release nogil;
             return b.autotune(mask, reps);
           },
           py::arg("method_mask") = 3u, py::arg("reps") = 20,
           "times the candidate tiles, keeps the fastest; returns us per frame")
      .def("launch", &SutraWfsBatch::launch)
      .def("capture", &SutraWfsBatch::capture, py::arg("frames"))
      .def("replay",
           [](SutraWfsBatch& b, int32_t n) {
             py::gil_scoped_release nogil;
             return b.replay(n);
           },
           py::arg("ngraphs"), "launches the captured graph ngraphs times; returns ms")
      .def("run",
           [](SutraWfsBatch& b, int32_t n) {
             py::gil_scoped_release nogil;
             return b.run(n);
           },
           py::arg("nframes"), "nframes single-frame launches; returns ms");
  m.def("make_mlp_training_set",
        [](SutraWfsSH& w, CarmaObj<float>& base, float r0_base, float r0_min, float r0_max,
           int32_t nframes, uint64_t seed, int32_t geom_type, CarmaObj<float>& X,
           CarmaObj<float>& T) {
          py::gil_scoped_release nogil;
          return make_mlp_training_set(&w, &base, r0_base, r0_min, r0_max, nframes, seed,
                                       geom_type, &X, &T);
        },
        py::arg("wfs"), py::arg("base"), py::arg("r0_base"), py::arg("r0_min"), py::arg("r0_max"),
        py::arg("nframes"), py::arg("seed"), py::arg("geom_type"), py::arg("X"), py::arg("T"),
        "noisy spots and geometric slopes of nframes shifted / rescaled screen windows");
  m.attr("SHWFS_METHOD_FFT_FP32") = (int)SHWFS_METHOD_FFT_FP32;
  m.attr("SHWFS_METHOD_DFT_BF16X3") = (int)SHWFS_METHOD_DFT_BF16X3;
  m.attr("SHWFS_METHOD_DFT_FP16") = (int)SHWFS_METHOD_DFT_FP16;
  m.attr("__version__") = shwfs_version();
}
