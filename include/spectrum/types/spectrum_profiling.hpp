#pragma once
namespace antenna_fft {
struct ProfilingData {
  double upload_time_ms = 0.0;
  double fft_time_ms = 0.0;
  double post_kernel_time_ms = 0.0;
  double download_time_ms = 0.0;
  double total_time_ms = 0.0;
};
}  // namespace antenna_fft
