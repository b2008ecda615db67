// pybind11 module "carma": device context, device arrays and timer.
// Mirrors the names of the reference's carma module (python_module/carma/src/
// carma.cpp, context.cpp, obj.hpp:31-190, timer.cpp): context.get_instance*,
// obj_<type>(context, ndarray) with an F-ordered buffer protocol, host2device /
// device2host / copy_from / copy_into / reset, init_prng / prng, timer.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>

#include "../carma/carma.hpp"

namespace py = pybind11;

template <typename T>
struct PyElem {
  using type = T;
};
template <>
struct PyElem<hipFloatComplex> {
  using type = std::complex<float>;
};

template <typename T>
using ArrayF = py::array_t<typename PyElem<T>::type, py::array::f_style | py::array::forcecast>;

template <typename T>
static std::unique_ptr<CarmaObj<T>> obj_from_array(CarmaContext& c, const ArrayF<T>& data) {
  std::vector<int64_t> dims(data.ndim() + 1);
  dims[0] = data.ndim();
  for (py::ssize_t i = 0; i < data.ndim(); ++i) dims[i + 1] = data.shape(i);
  return std::unique_ptr<CarmaObj<T>>(
      new CarmaObj<T>(&c, dims.data(), reinterpret_cast<const T*>(data.data())));
}

template <typename T>
static void declare_obj(py::module& m, const char* name) {
  using Class = CarmaObj<T>;
  using E = typename PyElem<T>::type;
  py::class_<Class>(m, name, py::buffer_protocol())
      .def(py::init(&obj_from_array<T>), py::arg("context").none(false), py::arg("h_data").none(false))
      .def(py::init([](CarmaContext& c, const Class& d) {
             return std::unique_ptr<Class>(new Class(&c, &d));
           }),
           py::arg("context").none(false), py::arg("d_data").none(false))
      .def_buffer([](Class& o) -> py::buffer_info {
        o.sync_h_data();
        const int64_t* dims = o.get_dims();
        std::vector<py::ssize_t> shape(dims[0]), strides(dims[0]);
        py::ssize_t stride = sizeof(T);
        for (int64_t d = 0; d < dims[0]; ++d) {  // F-style
          shape[d] = dims[d + 1];
          strides[d] = stride;
          stride *= shape[d];
        }
        return py::buffer_info(reinterpret_cast<E*>(o.get_h_data()), sizeof(T),
                               py::format_descriptor<E>::format(), dims[0], shape, strides);
      })
      .def("__repr__", &Class::to_string)
      .def_property_readonly("shape",
                             [](Class& o) {
                               std::vector<int64_t> s(o.get_dims() + 1, o.get_dims() + 1 + o.get_dims(0));
                               return py::array_t<int64_t>(s.size(), s.data());
                             })
      .def_property_readonly("nbElem", &Class::get_nb_elements)
      .def_property_readonly("device", &Class::get_device)
      .def_property_readonly("data_ptr", [](Class& o) { return reinterpret_cast<uintptr_t>(o.get_data()); },
                             "device address of the data (for zero-copy interop)")
      .def("host2device",
           [](Class& o, ArrayF<T>& data) {
             if ((int64_t)data.size() != o.get_nb_elements())
               throw std::runtime_error("host2device: size mismatch");
             o.host2device(reinterpret_cast<const T*>(data.data()));
           },
           py::arg("data").none(false))
      .def("device2host",
           [](Class& o, ArrayF<T>& data) {
             if ((int64_t)data.size() != o.get_nb_elements())
               throw std::runtime_error("device2host: size mismatch");
             o.device2host(reinterpret_cast<T*>(data.mutable_data()));
           },
           py::arg("data").none(false))
      .def("copy_from",
           [](Class& o, Class& src, int64_t n) { o.copy_from(src.get_data(), n); },
           py::arg("data"), py::arg("nb_elem"))
      .def("copy_into",
           [](Class& o, Class& dst, int64_t n) { o.copy_into(dst.get_data(), n); },
           py::arg("data"), py::arg("nb_elem"))
      .def("reset", [](Class& o) { return o.reset(); })
      .def("is_rng_init", &Class::is_rng_init)
      .def("init_prng", [](Class& o) { return o.init_prng(); })
      .def("init_prng", [](Class& o, int64_t seed) { return o.init_prng(seed); }, py::arg("seed"))
      .def("prng", [](Class& o, char g) { return o.prng(g); }, py::arg("gtype"))
      .def("prng", [](Class& o, char g, float a) { return o.prng(g, a); }, py::arg("gtype"),
           py::arg("alpha"))
      .def("prng", [](Class& o, char g, float a, float b) { return o.prng(g, a, b); },
           py::arg("gtype"), py::arg("alpha"), py::arg("beta"));
}

PYBIND11_MODULE(carma, m) {
  m.doc() = "carma: HIP device context and arrays (MI355X build)";

  py::class_<CarmaDevice>(m, "device")
      .def_property_readonly("id", &CarmaDevice::get_id)
      .def_property_readonly("name", &CarmaDevice::get_name)
      .def_property_readonly("total_mem", &CarmaDevice::get_total_mem)
      .def_property_readonly("free_mem", &CarmaDevice::get_free_mem)
      .def_property_readonly("cu_count", &CarmaDevice::get_cu_count);

  py::class_<CarmaContext, std::unique_ptr<CarmaContext, py::nodelete>>(m, "context")
      .def_static("get_instance", &CarmaContext::instance, py::return_value_policy::reference)
      .def_static("get_instance_1gpu", &CarmaContext::instance_1gpu,
                  py::return_value_policy::reference, py::arg("dev"))
      .def_static("get_instance_ngpu",
                  [](int32_t n, py::array_t<int32_t> ids) -> CarmaContext& {
                    return CarmaContext::instance_ngpu(n, ids.data());
                  },
                  py::return_value_policy::reference, py::arg("nb_devices"), py::arg("devices_id"))
      .def_property_readonly("ndevice", &CarmaContext::get_ndevice)
      .def_property_readonly("ndevices", &CarmaContext::get_ndevice)
      .def_property_readonly("active_device", &CarmaContext::get_active_device)
      .def_property_readonly("activeRealDevice", &CarmaContext::get_active_real_device)
      .def_property_readonly("hipRuntimeGetVersion", &CarmaContext::get_runtime_version)
      .def_property_readonly("driverVersion", &CarmaContext::get_driver_version)
      .def("get_device", &CarmaContext::get_device, py::return_value_policy::reference, py::arg("dev"))
      .def("set_active_device",
           [](CarmaContext& c, int32_t d) { return c.set_active_device(d, 1); }, py::arg("new_device"))
      .def("set_active_device_force",
           [](CarmaContext& c, int32_t d) { return c.set_active_device_force(d, 1); },
           py::arg("new_device"))
      .def("get_device_info", &CarmaContext::get_device_info, py::arg("dev"));

  declare_obj<float>(m, "obj_float");
  declare_obj<double>(m, "obj_double");
  declare_obj<int32_t>(m, "obj_int32_t");
  declare_obj<uint32_t>(m, "obj_uint");
  declare_obj<uint16_t>(m, "obj_uint16");
  declare_obj<hipFloatComplex>(m, "obj_float_complex");

  py::class_<CarmaTimer>(m, "timer")
      .def(py::init<>())
      .def("start", [](CarmaTimer& t) { t.start(nullptr); })
      .def("stop", [](CarmaTimer& t) { t.stop(nullptr); })
      .def("reset", &CarmaTimer::reset)
      .def_property_readonly("total_time", [](CarmaTimer& t) { return t.elapsed_ms() * 1e-3; });

  m.def("deviceSync", []() { carma_safe_call(hipDeviceSynchronize()); });
  m.attr("__version__") = shwfs_version();
}
