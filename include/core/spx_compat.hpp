// Stand-ins for the DspCore (drv_gpu_lib) types that callers of the spectrum
// facades use.  DspCore itself is not part of this library; what the
// reference's callers of FFTProcessorROCm / ComplexToMagPhaseROCm /
// SpectrumProcessorROCm name on a backend is defined here with the same
// names and meaning: device memory helpers (tests/test_fft_processor_rocm.hpp:216-226,
// tests/test_complex_to_mag_phase_rocm.hpp:138-221), GetDeviceInfo().name
// (tests/test_fft_matrix_rocm.hpp:448, tests/test_gate3_fft_profiler_v2.hpp:133),
// GetCore().GetWarpSize() (src/fft_func/src/all_maxima_pipeline_rocm.cpp:82) and
// the ROCmProfilingData timestamps (include/spectrum/utils/rocm_profiling_helpers.hpp:60-78).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

namespace drv_gpu_lib {

// Backends the factory understands (the ROCm ones map to the gfx950 library).
enum class BackendType { OPENCL, ROCm, OPENCLandROCm, AUTO };

// Where AllMaxima results go: host vectors, caller-owned device buffers, or both.
enum class OutputDestination { CPU, GPU, ALL };

// One timed stage.  The facades fill it from hipEvent pairs the way the
// reference's MakeROCmDataFromEvents does: start_ns = 0, end_ns =
// complete_ns = the stage's duration in ns, queued/submit 0; kind is the
// stage's code (0 upload, 1 compute, 2 download) and op_string its name.
struct ROCmProfilingData {
  uint64_t queued_ns = 0;
  uint64_t submit_ns = 0;
  uint64_t start_ns = 0;
  uint64_t end_ns = 0;
  uint64_t complete_ns = 0;
  uint32_t kind = 0;
  std::string op_string;
};

// Device properties a caller reads from a backend.
struct DeviceInfo {
  std::string name;
  size_t global_mem_bytes = 0;
  int compute_units = 0;
  int warp_size = 64;
  int device_index = 0;
};

// The device-level queries behind IBackend::GetCore().
class ROCmCore {
 public:
  static int GetAvailableDeviceCount();
  int GetWarpSize() const { return warp_size_; }
  int GetComputeUnits() const { return compute_units_; }
  int warp_size_ = 64;
  int compute_units_ = 0;
};

// The facades are built on a non-owning backend: a device index and a HIP
// stream, plus device-memory helpers on that device.
class IBackend {
 public:
  virtual ~IBackend() = default;
  virtual int GetDeviceIndex() const = 0;
  virtual void* GetNativeQueue() const = 0;  // hipStream_t
  virtual std::string GetDeviceName() const = 0;
  virtual DeviceInfo GetDeviceInfo() const = 0;
  virtual const ROCmCore& GetCore() const = 0;
  virtual size_t GetGlobalMemorySize() const = 0;
  // hipMalloc / hipFree on the backend's device (std::runtime_error on failure)
  virtual void* Allocate(size_t bytes) = 0;
  virtual void Free(void* ptr) = 0;
  // copies ordered on the backend's stream, complete on return
  virtual void MemcpyHostToDevice(void* dst, const void* src, size_t bytes) = 0;
  virtual void MemcpyDeviceToHost(void* dst, const void* src, size_t bytes) = 0;
  virtual void MemcpyDeviceToDevice(void* dst, const void* src, size_t bytes) = 0;
  virtual void Synchronize() = 0;
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
  std::string GetDeviceName() const override { return info_.name; }
  DeviceInfo GetDeviceInfo() const override { return info_; }
  const ROCmCore& GetCore() const override { return core_; }
  size_t GetGlobalMemorySize() const override { return info_.global_mem_bytes; }
  void* Allocate(size_t bytes) override;
  void Free(void* ptr) override;
  void MemcpyHostToDevice(void* dst, const void* src, size_t bytes) override;
  void MemcpyDeviceToHost(void* dst, const void* src, size_t bytes) override;
  void MemcpyDeviceToDevice(void* dst, const void* src, size_t bytes) override;
  void Synchronize() override;

 private:
  int device_ = 0;
  void* stream_ = nullptr;
  DeviceInfo info_;
  ROCmCore core_;
};

}  // namespace drv_gpu_lib
