// carma: minimal HIP-native device-array runtime used by the SH WFS path.
//
// Mirrors the parts of the reference's libcarma the path touches
// (libcarma/include/carma_context.hpp:85-161, carma_obj.hpp:97-361,
// carma_timer.hpp:33-66): a per-process device context, an owning device
// array with the reference's dims convention (dims[0] = rank, dims[1..] =
// extents, column-major), host<->device copies, reset/memset and the
// counter-based random fills.  HIP is called directly; no CUDA names.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_complex.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "shwfs.h"

// Throws std::runtime_error on a failed HIP call (the reference only checks
// in DEBUG builds, libcarma/include/carma_utils.hpp:99-114; here always).
#define carma_safe_call(call)                                                            \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

// Turns a negative shwfs status into an exception carrying the library message.
void shwfs_throw_if(int32_t rc, const char* where);

class CarmaDevice {
 public:
  explicit CarmaDevice(int32_t id);
  int32_t get_id() const { return id_; }
  const hipDeviceProp_t& get_properties() const { return prop_; }
  std::string get_name() const { return prop_.name; }
  size_t get_total_mem() const { return prop_.totalGlobalMem; }
  size_t get_free_mem() const;
  int32_t get_cu_count() const { return prop_.multiProcessorCount; }

 private:
  int32_t id_;
  hipDeviceProp_t prop_;
};

class CarmaContext {
 public:
  static CarmaContext& instance();
  static CarmaContext& instance_1gpu(int32_t num_device);
  static CarmaContext& instance_ngpu(int32_t nb_devices, const int32_t* devices_id);
  CarmaContext(const CarmaContext&) = delete;
  CarmaContext& operator=(const CarmaContext&) = delete;

  int32_t get_ndevice() const { return (int32_t)devices_.size(); }
  int32_t get_active_device() const { return active_; }
  int32_t get_active_real_device() const { return devices_.at(active_)->get_id(); }
  CarmaDevice* get_device(int32_t dev) { return devices_.at(dev).get(); }
  int32_t set_active_device(int32_t newDevice, int32_t silent = 1);
  int32_t set_active_device_force(int32_t newDevice, int32_t silent = 1);
  int32_t get_runtime_version() const;
  int32_t get_driver_version() const;
  std::string get_device_info(int32_t dev) const;

 private:
  CarmaContext(const std::vector<int32_t>& ids);
  std::vector<std::unique_ptr<CarmaDevice>> devices_;
  int32_t active_ = 0;
};

// Owning (or borrowing) device array.  dims_data[0] = rank, dims_data[1..rank]
// the extents, column-major as in the reference.
template <class T>
class CarmaObj {
 public:
  CarmaObj(CarmaContext* ctx, const int64_t* dims_data);
  CarmaObj(CarmaContext* ctx, const int64_t* dims_data, const T* h_data);
  CarmaObj(CarmaContext* ctx, const CarmaObj<T>* src);
  ~CarmaObj();
  CarmaObj(const CarmaObj&) = delete;
  CarmaObj& operator=(const CarmaObj&) = delete;

  T* get_data() { return d_data_; }
  const T* get_data() const { return d_data_; }
  T* get_data_at(int64_t index) { return d_data_ + index; }
  operator T*() { return d_data_; }
  const int64_t* get_dims() const { return dims_.data(); }
  int64_t get_dims(int32_t i) const { return dims_.at(i); }
  int64_t get_nb_elements() const { return nb_elem_; }
  int32_t get_device() const { return device_; }
  CarmaContext* get_context() { return ctx_; }

  int32_t host2device(const T* data);
  int32_t device2host(T* data) const;
  int32_t copy_from(const T* d_src, int64_t nb_elem);
  int32_t copy_into(T* d_dst, int64_t nb_elem) const;
  int32_t reset();
  int32_t memset(T value);

  // Counter-based random fills (Philox4x32-10 via libshwfs).  init_prng
  // stores the key and restarts the call counter; every prng call advances it.
  int32_t init_prng() { return init_prng(1234); }
  int32_t init_prng(int64_t seed);
  bool is_rng_init() const { return rng_init_; }
  int32_t prng(char gtype) { return prng(gtype, 1.0f, 0.0f); }
  int32_t prng(char gtype, float alpha) { return prng(gtype, alpha, 0.0f); }
  int32_t prng(char gtype, float alpha, float beta);
  uint64_t rng_seed() const { return seed_; }
  uint32_t rng_call() const { return call_; }
  uint32_t next_rng_call() { return call_++; }

  // host mirror for the Python buffer protocol
  int32_t sync_h_data();
  T* get_h_data() { return h_data_.data(); }
  std::string to_string() const;

 private:
  void alloc(CarmaContext* ctx, const int64_t* dims_data);
  CarmaContext* ctx_ = nullptr;
  int32_t device_ = 0;
  std::vector<int64_t> dims_;
  int64_t nb_elem_ = 0;
  T* d_data_ = nullptr;
  std::vector<T> h_data_;
  bool rng_init_ = false;
  uint64_t seed_ = 0;
  uint32_t call_ = 0;
};

// hipEvent-based timer (libcarma/include/carma_timer.hpp:33-66).
class CarmaTimer {
 public:
  CarmaTimer();
  ~CarmaTimer();
  void start(hipStream_t s = nullptr);
  void stop(hipStream_t s = nullptr);
  void reset() { total_ms_ = 0.0; }
  double elapsed_ms() const { return total_ms_; }

 private:
  hipEvent_t start_, stop_;
  hipStream_t stream_ = nullptr;
  double total_ms_ = 0.0;
};

