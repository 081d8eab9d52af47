// Minimal stand-ins for the DspCore (drv_gpu_lib) types that appear in the
// signatures of the spectrum facades.  DspCore itself is not part of this
// library; only what a caller of FFTProcessorROCm / ComplexToMagPhaseROCm /
// SpectrumProcessorROCm needs to name is defined here.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

namespace drv_gpu_lib {

// Backends the factory understands (the ROCm ones map to the gfx950 library).
enum class BackendType { OPENCL, ROCm, OPENCLandROCm, AUTO };

// Where AllMaxima results go: host vectors, caller-owned device buffers, or both.
enum class OutputDestination { CPU, GPU, ALL };

// One timed stage, filled from hipEvent pairs.
struct ROCmProfilingData {
  double elapsed_ms = 0.0;
};

// The facades are built on a non-owning backend: a device index and a HIP stream.
class IBackend {
 public:
  virtual ~IBackend() = default;
  virtual int GetDeviceIndex() const = 0;
  virtual void* GetNativeQueue() const = 0;  // hipStream_t
  virtual std::string GetDeviceName() const = 0;
};

// Concrete ROCm backend: owns one non-blocking stream on one device.
class ROCmBackend : public IBackend {
 public:
  ROCmBackend();
  ~ROCmBackend() override;
  ROCmBackend(const ROCmBackend&) = delete;
  ROCmBackend& operator=(const ROCmBackend&) = delete;
  void Initialize(int device_index);
  bool IsInitialized() const { return stream_ != nullptr; }
  int GetDeviceIndex() const override { return device_; }
  void* GetNativeQueue() const override { return stream_; }
  std::string GetDeviceName() const override { return name_; }
  size_t GetGlobalMemorySize() const { return mem_; }

 private:
  int device_ = 0;
  void* stream_ = nullptr;
  std::string name_;
  size_t mem_ = 0;
};

struct ROCmCore {
  static int GetAvailableDeviceCount();
};

}  // namespace drv_gpu_lib
