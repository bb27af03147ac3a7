// device.hpp -- single-GPU device handle and RAII streams.
//
// Interface follows kites/device/stream.cuh:12-61 and nv_gpu.cuh:21-88.  The
// device type is `gpu`; `nv_gpu` is an alias so reference user code compiles
// unchanged.  Streams are created at the default priority.  The multi-device
// group (nv_gpus, nv_gpu.cuh:95-121) is not provided: this engine runs one
// graph on one GPU.
#ifndef KITES_DEVICE_HPP_
#define KITES_DEVICE_HPP_

#include <utility>

#include "common.hpp"

namespace kites {

class stream {
  int residing_device_id;
  hipStream_t strm;

 public:
  stream() = delete;
  stream(stream const&) = delete;
  stream& operator=(stream const&) = delete;
  stream& operator=(stream&&) = delete;
  stream(stream&& in) : residing_device_id{in.residing_device_id}, strm{in.strm} {
    in.strm = nullptr;
  }
  explicit stream(int dev) : residing_device_id(dev) {
    KITES_HIP_CHECK(hipSetDevice(dev));
    KITES_HIP_CHECK(hipStreamCreateWithFlags(&strm, hipStreamNonBlocking));
  }
  ~stream() {
    if (strm != nullptr) {
      (void)hipSetDevice(residing_device_id);
      (void)hipStreamDestroy(strm);
    }
  }
  void sync() {
    KITES_HIP_CHECK(hipSetDevice(residing_device_id));
    KITES_HIP_CHECK(hipStreamSynchronize(strm));
  }
  hipStream_t& get() { return strm; }
  hipStream_t* get_ptr() { return &strm; }
  int getResidingDeviceId() const { return residing_device_id; }
};

class gpu {
  int devIdx;
  hipDeviceProp_t devProp;
  stream ssEven, ssOdd;

 public:
  explicit gpu(int idx) : devIdx{idx}, ssEven{idx}, ssOdd{idx} {
    KITES_HIP_CHECK(hipGetDeviceProperties(&devProp, idx));
    KITES_HIP_CHECK(hipSetDevice(idx));
  }
  gpu() : gpu{0} {}
  gpu(gpu const&) = delete;
  gpu& operator=(gpu const&) = delete;
  gpu& operator=(gpu&&) = delete;
  gpu(gpu&& in)
      : devIdx{in.devIdx}, devProp(in.devProp), ssEven{std::move(in.ssEven)},
        ssOdd{std::move(in.ssOdd)} {}
  ~gpu() = default;

  int getDevIdx() const { return devIdx; }
  hipDeviceProp_t& getDevProp() { return devProp; }
  int numComputeUnits() const { return devProp.multiProcessorCount; }
  void setAsActive() const { KITES_HIP_CHECK(hipSetDevice(devIdx)); }
  void sync() {
    setAsActive();
    KITES_HIP_CHECK(hipDeviceSynchronize());
  }
  stream& getSsEven() { return ssEven; }
  stream& getSsOdd() { return ssOdd; }
  // (used MB, total MB), as nv_gpu::queryMemUsage (nv_gpu.cuh:78-86).
  std::pair<double, double> queryMemUsage() const {
    setAsActive();
    std::size_t fr = 0, tot = 0;
    KITES_HIP_CHECK(hipMemGetInfo(&fr, &tot));
    return {(tot - fr) / 1e6, tot / 1e6};
  }
};

using nv_gpu = gpu;

inline void sync_with(gpu& d) { d.sync(); }

}  // namespace kites

#endif  // KITES_DEVICE_HPP_
