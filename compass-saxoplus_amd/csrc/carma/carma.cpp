// carma runtime implementation (host code; HIP runtime API only).
#include "carma.hpp"

#include <mutex>

void shwfs_throw_if(int32_t rc, const char* where) {
  if (rc == SHWFS_OK) return;
  char msg[512];
  shwfs_last_error(msg, sizeof(msg));
  throw std::runtime_error(std::string(where) + ": " + msg);
}

// ---------------------------------------------------------------- CarmaDevice
CarmaDevice::CarmaDevice(int32_t id) : id_(id) {
  carma_safe_call(hipGetDeviceProperties(&prop_, id));
}

size_t CarmaDevice::get_free_mem() const {
  int32_t cur = 0;
  carma_safe_call(hipGetDevice(&cur));
  carma_safe_call(hipSetDevice(id_));
  size_t fr = 0, tot = 0;
  carma_safe_call(hipMemGetInfo(&fr, &tot));
  carma_safe_call(hipSetDevice(cur));
  return fr;
}

// ---------------------------------------------------------------- CarmaContext
static std::mutex g_ctx_mutex;
static std::unique_ptr<CarmaContext> g_ctx;

CarmaContext::CarmaContext(const std::vector<int32_t>& ids) {
  for (int32_t id : ids) devices_.emplace_back(new CarmaDevice(id));
  if (devices_.empty()) throw std::runtime_error("CarmaContext: no device");
  carma_safe_call(hipSetDevice(devices_[0]->get_id()));
  active_ = 0;
}

CarmaContext& CarmaContext::instance() {
  std::lock_guard<std::mutex> lk(g_ctx_mutex);
  if (!g_ctx) {
    int32_t n = 0;
    carma_safe_call(hipGetDeviceCount(&n));
    if (n <= 0) throw std::runtime_error("CarmaContext: no HIP device visible");
    std::vector<int32_t> ids(n);
    for (int32_t i = 0; i < n; ++i) ids[i] = i;
    g_ctx.reset(new CarmaContext(ids));
  }
  return *g_ctx;
}

CarmaContext& CarmaContext::instance_1gpu(int32_t num_device) {
  std::lock_guard<std::mutex> lk(g_ctx_mutex);
  if (!g_ctx) g_ctx.reset(new CarmaContext({num_device}));
  return *g_ctx;
}

CarmaContext& CarmaContext::instance_ngpu(int32_t nb_devices, const int32_t* devices_id) {
  std::lock_guard<std::mutex> lk(g_ctx_mutex);
  if (!g_ctx) g_ctx.reset(new CarmaContext(std::vector<int32_t>(devices_id, devices_id + nb_devices)));
  return *g_ctx;
}

int32_t CarmaContext::set_active_device(int32_t newDevice, int32_t silent) {
  (void)silent;
  if (newDevice == active_) return active_;
  return set_active_device_force(newDevice, silent);
}

int32_t CarmaContext::set_active_device_force(int32_t newDevice, int32_t silent) {
  (void)silent;
  if (newDevice < 0 || newDevice >= get_ndevice())
    throw std::runtime_error("set_active_device: invalid device index");
  carma_safe_call(hipSetDevice(devices_[newDevice]->get_id()));
  active_ = newDevice;
  return active_;
}

int32_t CarmaContext::get_runtime_version() const {
  int v = 0;
  carma_safe_call(hipRuntimeGetVersion(&v));
  return v;
}

int32_t CarmaContext::get_driver_version() const {
  int v = 0;
  carma_safe_call(hipDriverGetVersion(&v));
  return v;
}

std::string CarmaContext::get_device_info(int32_t dev) const {
  const auto& p = devices_.at(dev)->get_properties();
  std::ostringstream os;
  os << "device " << dev << ": " << p.name << " (" << p.gcnArchName << "), " << p.multiProcessorCount
     << " CUs, " << (p.totalGlobalMem >> 20) << " MiB";
  return os.str();
}

// ---------------------------------------------------------------- CarmaObj
template <class T>
void CarmaObj<T>::alloc(CarmaContext* ctx, const int64_t* dims_data) {
  ctx_ = ctx;
  device_ = ctx->get_active_device();
  const int64_t rank = dims_data[0];
  if (rank < 1) throw std::runtime_error("CarmaObj: rank must be >= 1");
  dims_.assign(dims_data, dims_data + rank + 1);
  nb_elem_ = 1;
  for (int64_t i = 1; i <= rank; ++i) {
    if (dims_data[i] < 0) throw std::runtime_error("CarmaObj: negative extent");
    nb_elem_ *= dims_data[i];
  }
  if (nb_elem_ > 0) carma_safe_call(hipMalloc((void**)&d_data_, sizeof(T) * nb_elem_));
}

template <class T>
CarmaObj<T>::CarmaObj(CarmaContext* ctx, const int64_t* dims_data) {
  alloc(ctx, dims_data);
  reset();
}

template <class T>
CarmaObj<T>::CarmaObj(CarmaContext* ctx, const int64_t* dims_data, const T* h_data) {
  alloc(ctx, dims_data);
  host2device(h_data);
}

template <class T>
CarmaObj<T>::CarmaObj(CarmaContext* ctx, const CarmaObj<T>* src) {
  alloc(ctx, src->get_dims());
  copy_from(src->get_data(), nb_elem_);
}

template <class T>
CarmaObj<T>::~CarmaObj() {
  if (d_data_) (void)hipFree(d_data_);
}

template <class T>
int32_t CarmaObj<T>::host2device(const T* data) {
  if (nb_elem_ == 0) return EXIT_SUCCESS;
  carma_safe_call(hipMemcpy(d_data_, data, sizeof(T) * nb_elem_, hipMemcpyHostToDevice));
  return EXIT_SUCCESS;
}

template <class T>
int32_t CarmaObj<T>::device2host(T* data) const {
  if (nb_elem_ == 0) return EXIT_SUCCESS;
  carma_safe_call(hipMemcpy(data, d_data_, sizeof(T) * nb_elem_, hipMemcpyDeviceToHost));
  return EXIT_SUCCESS;
}

template <class T>
int32_t CarmaObj<T>::copy_from(const T* d_src, int64_t nb) {
  if (nb > nb_elem_) throw std::runtime_error("CarmaObj::copy_from: too many elements");
  if (nb > 0) carma_safe_call(hipMemcpy(d_data_, d_src, sizeof(T) * nb, hipMemcpyDeviceToDevice));
  return EXIT_SUCCESS;
}

template <class T>
int32_t CarmaObj<T>::copy_into(T* d_dst, int64_t nb) const {
  if (nb > nb_elem_) throw std::runtime_error("CarmaObj::copy_into: too many elements");
  if (nb > 0) carma_safe_call(hipMemcpy(d_dst, d_data_, sizeof(T) * nb, hipMemcpyDeviceToDevice));
  return EXIT_SUCCESS;
}

template <class T>
int32_t CarmaObj<T>::reset() {
  if (nb_elem_ > 0) carma_safe_call(hipMemset(d_data_, 0, sizeof(T) * nb_elem_));
  return EXIT_SUCCESS;
}

template <class T>
int32_t CarmaObj<T>::memset(T value) {
  std::vector<T> h((size_t)nb_elem_, value);
  return host2device(h.data());
}

template <class T>
int32_t CarmaObj<T>::init_prng(int64_t seed) {
  seed_ = (uint64_t)seed;
  call_ = 0;
  rng_init_ = true;
  return EXIT_SUCCESS;
}

template <class T>
int32_t CarmaObj<T>::prng(char gtype, float alpha, float beta) {
  if constexpr (std::is_same<T, float>::value) {
    if (!rng_init_) init_prng(1234);
    shwfs_throw_if(shwfs_prng(d_data_, nb_elem_, gtype, alpha, beta, seed_, call_++, 3u, nullptr),
                   "CarmaObj::prng");
    return EXIT_SUCCESS;
  } else {
    (void)gtype;
    (void)alpha;
    (void)beta;
    throw std::runtime_error("CarmaObj::prng is implemented for float arrays only");
  }
}

template <class T>
int32_t CarmaObj<T>::sync_h_data() {
  h_data_.resize((size_t)nb_elem_);
  carma_safe_call(hipDeviceSynchronize());
  return device2host(h_data_.data());
}

template <class T>
std::string CarmaObj<T>::to_string() const {
  std::ostringstream os;
  os << "CarmaObj<" << sizeof(T) << "B>(";
  for (size_t i = 1; i < dims_.size(); ++i) os << (i > 1 ? ", " : "") << dims_[i];
  os << ") on device " << device_;
  return os.str();
}

template class CarmaObj<float>;
template class CarmaObj<double>;
template class CarmaObj<int32_t>;
template class CarmaObj<uint32_t>;
template class CarmaObj<uint16_t>;
template class CarmaObj<hipFloatComplex>;

// ---------------------------------------------------------------- CarmaTimer
CarmaTimer::CarmaTimer() {
  carma_safe_call(hipEventCreate(&start_));
  carma_safe_call(hipEventCreate(&stop_));
}

CarmaTimer::~CarmaTimer() {
  (void)hipEventDestroy(start_);
  (void)hipEventDestroy(stop_);
}

void CarmaTimer::start(hipStream_t s) {
  stream_ = s;
  carma_safe_call(hipEventRecord(start_, s));
}

void CarmaTimer::stop(hipStream_t s) {
  carma_safe_call(hipEventRecord(stop_, s));
  carma_safe_call(hipEventSynchronize(stop_));
  float ms = 0.f;
  carma_safe_call(hipEventElapsedTime(&ms, start_, stop_));
  total_ms_ += ms;
}
