// pybind11 module `dsp_spectrum`: the reference's Python surface for the FFT
// and spectrum-maxima path (python/dsp_spectrum_module.cpp:35-68 and the
// py_*_rocm.hpp bindings there) over the gfx950 facades.  Same class names,
// method names, keyword names, defaults, return shapes and dict keys; the GIL
// is released during compute.  Extra (no reference counterpart):
// FusedSpectrumPeaksROCm, the one-launch window->FFT->|X|->threshold->peaks path.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <cstring>
#include <hip/hip_runtime.h>
#include <spectrum/complex_to_mag_phase_rocm.hpp>
#include <spectrum/fft_processor_rocm.hpp>
#include <spectrum/processors/spectrum_processor_rocm.hpp>
#include <spectrum/utils/cpu_fft.hpp>
#include <spectrum_mi355x.h>
#include <stdexcept>
#include <vector>

namespace py = pybind11;
using cvec = std::vector<std::complex<float>>;
using carr = py::array_t<std::complex<float>, py::array::c_style | py::array::forcecast>;

template <typename T>
py::array_t<T> to_numpy(std::vector<T>&& v) {
  auto* h = new std::vector<T>(std::move(v));
  py::capsule cap(h, [](void* p) { delete static_cast<std::vector<T>*>(p); });
  return py::array_t<T>({(py::ssize_t)h->size()}, {(py::ssize_t)sizeof(T)}, h->data(), cap);
}
template <typename T>
py::array_t<T> to_numpy_2d(std::vector<T>&& v, size_t rows, size_t cols) {
  auto* h = new std::vector<T>(std::move(v));
  py::capsule cap(h, [](void* p) { delete static_cast<std::vector<T>*>(p); });
  return py::array_t<T>({(py::ssize_t)rows, (py::ssize_t)cols}, {(py::ssize_t)(cols * sizeof(T)), (py::ssize_t)sizeof(T)},
                        h->data(), cap);
}

class ROCmGPUContext {
 public:
  explicit ROCmGPUContext(int device_index) { backend_.Initialize(device_index); }
  std::string device_name() const { return backend_.GetDeviceName(); }
  int device_index() const { return backend_.GetDeviceIndex(); }
  drv_gpu_lib::IBackend* backend() { return &backend_; }

 private:
  drv_gpu_lib::ROCmBackend backend_;
};

// ---------------------------------------------------------------- FFTProcessorROCm
class PyFFTProcessorROCm {
 public:
  explicit PyFFTProcessorROCm(ROCmGPUContext& ctx) : fft_(ctx.backend()) {}

  static void shape(const py::buffer_info& b, uint32_t& beams, uint32_t& n) {
    const size_t total = (size_t)b.size;
    if (b.ndim == 2) {
      if (beams == 0) beams = (uint32_t)b.shape[0];
      if (n == 0) n = (uint32_t)b.shape[1];
    } else {
      if (beams == 0) beams = 1;
      if (n == 0) n = (uint32_t)(total / beams);
    }
  }

  py::array_t<std::complex<float>> process_complex(carr data, float sample_rate, uint32_t beam_count,
                                                   uint32_t n_point) {
    auto b = data.request();
    shape(b, beam_count, n_point);
    auto* p = static_cast<std::complex<float>*>(b.ptr);
    cvec v(p, p + b.size);
    fft_processor::FFTProcessorParams prm;
    prm.beam_count = beam_count;
    prm.n_point = n_point;
    prm.sample_rate = sample_rate;
    prm.output_mode = fft_processor::FFTOutputMode::COMPLEX;
    std::vector<fft_processor::FFTComplexResult> r;
    {
      py::gil_scoped_release rel;
      r = fft_.ProcessComplex(v, prm);
    }
    if (r.size() == 1) return to_numpy(std::move(r[0].spectrum));
    const uint32_t N = r[0].nFFT;
    std::vector<std::complex<float>> flat;
    flat.reserve(r.size() * N);
    for (auto& x : r) flat.insert(flat.end(), x.spectrum.begin(), x.spectrum.end());
    return to_numpy_2d(std::move(flat), r.size(), N);
  }

  py::dict process_mag_phase(carr data, float sample_rate, uint32_t beam_count, uint32_t n_point, bool include_freq) {
    auto b = data.request();
    shape(b, beam_count, n_point);
    auto* p = static_cast<std::complex<float>*>(b.ptr);
    cvec v(p, p + b.size);
    fft_processor::FFTProcessorParams prm;
    prm.beam_count = beam_count;
    prm.n_point = n_point;
    prm.sample_rate = sample_rate;
    prm.output_mode = include_freq ? fft_processor::FFTOutputMode::MAGNITUDE_PHASE_FREQ
                                   : fft_processor::FFTOutputMode::MAGNITUDE_PHASE;
    std::vector<fft_processor::FFTMagPhaseResult> r;
    {
      py::gil_scoped_release rel;
      r = fft_.ProcessMagPhase(v, prm);
    }
    py::dict out;
    out["nFFT"] = r[0].nFFT;
    out["sample_rate"] = r[0].sample_rate;
    if (r.size() == 1) {
      out["magnitude"] = to_numpy(std::move(r[0].magnitude));
      out["phase"] = to_numpy(std::move(r[0].phase));
      if (include_freq && !r[0].frequency.empty()) out["frequency"] = to_numpy(std::move(r[0].frequency));
      return out;
    }
    const uint32_t N = r[0].nFFT;
    std::vector<float> m, ph, f;
    for (auto& x : r) {
      m.insert(m.end(), x.magnitude.begin(), x.magnitude.end());
      ph.insert(ph.end(), x.phase.begin(), x.phase.end());
      if (include_freq) f.insert(f.end(), x.frequency.begin(), x.frequency.end());
    }
    out["magnitude"] = to_numpy_2d(std::move(m), r.size(), N);
    out["phase"] = to_numpy_2d(std::move(ph), r.size(), N);
    if (!f.empty()) out["frequency"] = to_numpy_2d(std::move(f), r.size(), N);
    return out;
  }

  py::dict get_profiling() const {
    auto p = fft_.GetProfilingData();
    py::dict d;
    d["upload_ms"] = p.upload_time_ms;
    d["fft_ms"] = p.fft_time_ms;
    d["post_processing_ms"] = p.post_processing_time_ms;
    d["download_ms"] = p.download_time_ms;
    d["total_ms"] = p.total_time_ms;
    return d;
  }
  uint32_t nfft() const { return fft_.GetNFFT(); }

 private:
  fft_processor::FFTProcessorROCm fft_;
};

// ---------------------------------------------------------------- SpectrumMaximaFinderROCm
class PySpectrumMaximaFinderROCm {
 public:
  explicit PySpectrumMaximaFinderROCm(ROCmGPUContext& ctx) : proc_(ctx.backend()) {}

  py::object process(carr signal, uint32_t n_point, float sample_rate, uint32_t antenna_count, uint32_t repeat_count) {
    auto b = signal.request();
    if (antenna_count == 0) antenna_count = b.ndim == 2 ? (uint32_t)b.shape[0] : 1;
    auto* p = static_cast<std::complex<float>*>(b.ptr);
    cvec v(p, p + b.size);
    ensure_init(antenna_count, n_point, sample_rate, repeat_count, antenna_fft::PeakSearchMode::ONE_PEAK);
    std::vector<antenna_fft::SpectrumResult> r;
    {
      py::gil_scoped_release rel;
      r = proc_.ProcessFromCPU(v);
    }
    auto beam = [](const antenna_fft::SpectrumResult& s) {
      py::dict d;
      d["antenna_id"] = s.antenna_id;
      d["freq_hz"] = s.interpolated.refined_frequency;
      d["magnitude"] = s.interpolated.magnitude;
      d["phase"] = s.interpolated.phase;
      d["index"] = s.interpolated.index;
      d["freq_offset"] = s.interpolated.freq_offset;
      return d;
    };
    if (r.size() == 1) return beam(r[0]);
    py::list out;
    for (auto& s : r) out.append(beam(s));
    return out;
  }

  py::object find_all_maxima(carr fft_data, uint32_t nFFT, float sample_rate, uint32_t beam_count,
                             uint32_t search_start, uint32_t search_end) {
    auto b = fft_data.request();
    if (beam_count == 0) beam_count = b.ndim == 2 ? (uint32_t)b.shape[0] : 1;
    auto* p = static_cast<std::complex<float>*>(b.ptr);
    cvec v(p, p + b.size);
    ensure_init(beam_count, nFFT, sample_rate, 1, antenna_fft::PeakSearchMode::ALL_MAXIMA);
    antenna_fft::AllMaximaResult r;
    {
      py::gil_scoped_release rel;
      r = proc_.AllMaximaFromCPU(v, beam_count, nFFT, sample_rate, antenna_fft::OutputDestination::CPU, search_start,
                                 search_end);
    }
    return make_all(r, beam_count);
  }

  py::object find_all_maxima_from_signal(carr signal, uint32_t n_point, float sample_rate, uint32_t antenna_count,
                                         uint32_t repeat_count, uint32_t search_start, uint32_t search_end) {
    auto b = signal.request();
    if (antenna_count == 0) antenna_count = b.ndim == 2 ? (uint32_t)b.shape[0] : 1;
    auto* p = static_cast<std::complex<float>*>(b.ptr);
    cvec v(p, p + b.size);
    ensure_init(antenna_count, n_point, sample_rate, repeat_count, antenna_fft::PeakSearchMode::ALL_MAXIMA);
    antenna_fft::AllMaximaResult r;
    {
      py::gil_scoped_release rel;
      r = proc_.FindAllMaximaFromCPU(v, antenna_fft::OutputDestination::CPU, search_start, search_end);
    }
    return make_all(r, antenna_count);
  }

  bool initialized() const { return proc_.IsInitialized(); }
  py::dict get_params() const {
    const auto& p = proc_.GetParams();
    py::dict d;
    d["antenna_count"] = p.antenna_count;
    d["n_point"] = p.n_point;
    d["repeat_count"] = p.repeat_count;
    d["sample_rate"] = p.sample_rate;
    d["nFFT"] = p.nFFT;
    return d;
  }

 private:
  void ensure_init(uint32_t antennas, uint32_t n_point, float fs, uint32_t repeat, antenna_fft::PeakSearchMode mode) {
    const auto& p = proc_.GetParams();
    if (proc_.IsInitialized() && p.antenna_count == antennas && p.n_point == n_point && p.sample_rate == fs &&
        p.repeat_count == repeat && p.peak_mode == mode)
      return;
    antenna_fft::SpectrumParams prm;
    prm.antenna_count = antennas;
    prm.n_point = n_point;
    prm.sample_rate = fs;
    prm.repeat_count = repeat;
    prm.peak_mode = mode;
    proc_.Initialize(prm);
  }

  static py::object make_all(const antenna_fft::AllMaximaResult& r, uint32_t beams) {
    auto beam = [](const antenna_fft::AllMaximaBeamResult& b) {
      std::vector<uint32_t> pos;
      std::vector<float> mag, freq;
      for (const auto& m : b.maxima) {
        pos.push_back(m.index);
        mag.push_back(m.magnitude);
        freq.push_back(m.refined_frequency);
      }
      py::dict d;
      d["antenna_id"] = b.antenna_id;
      d["positions"] = to_numpy(std::move(pos));
      d["magnitudes"] = to_numpy(std::move(mag));
      d["frequencies"] = to_numpy(std::move(freq));
      d["num_maxima"] = b.num_maxima;
      return d;
    };
    if (beams == 1 && r.beams.size() == 1) return beam(r.beams[0]);
    py::list out;
    for (const auto& b : r.beams) out.append(beam(b));
    return out;
  }

  antenna_fft::SpectrumProcessorROCm proc_;
};

// ---------------------------------------------------------------- ComplexToMagROCm
class PyComplexToMagROCm {
 public:
  explicit PyComplexToMagROCm(ROCmGPUContext& ctx) : proc_(ctx.backend()) {}
  py::object process_magnitude(carr data, uint32_t beam_count, float norm_coeff) {
    auto b = data.request();
    const size_t total = (size_t)b.size;
    if (beam_count == 0) beam_count = b.ndim == 2 ? (uint32_t)b.shape[0] : 1;
    if (beam_count == 0 || total % beam_count != 0)
      throw std::invalid_argument("process_magnitude: total=" + std::to_string(total) +
                                  " not divisible by beam_count=" + std::to_string(beam_count));
    const uint32_t n_point = (uint32_t)(total / beam_count);
    fft_processor::MagPhaseParams prm;
    prm.beam_count = beam_count;
    prm.n_point = n_point;
    prm.norm_coeff = norm_coeff;
    void* dev = nullptr;
    if (hipMalloc(&dev, total * 8) != hipSuccess) throw std::runtime_error("process_magnitude: hipMalloc failed");
    if (hipMemcpy(dev, b.ptr, total * 8, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(dev);
      throw std::runtime_error("process_magnitude: hipMemcpy H2D failed");
    }
    std::vector<fft_processor::MagnitudeResult> r;
    try {
      py::gil_scoped_release rel;
      r = proc_.ProcessMagnitude(dev, prm, total * 8);
    } catch (...) {
      (void)hipFree(dev);
      throw;
    }
    (void)hipFree(dev);
    if (r.size() == 1) return to_numpy(std::move(r[0].magnitude));
    std::vector<float> flat;
    for (auto& x : r) flat.insert(flat.end(), x.magnitude.begin(), x.magnitude.end());
    return to_numpy_2d(std::move(flat), r.size(), n_point);
  }

 private:
  fft_processor::ComplexToMagPhaseROCm proc_;
};

// numpy dtype of the 32-byte MaxValue record (include/spectrum/types/spectrum_result_types.hpp:34-43 there)
static py::dtype maxvalue_dtype() {
  py::list f;
  for (const char* n : {"index"}) f.append(py::make_tuple(n, "<u4"));
  for (const char* n : {"real", "imag", "magnitude", "phase", "freq_offset", "refined_frequency"})
    f.append(py::make_tuple(n, "<f4"));
  f.append(py::make_tuple("pad", "<u4"));
  return py::dtype::from_args(f);
}

// ---------------------------------------------------------------- fused north-star path
class PyFusedSpectrumPeaksROCm {
 public:
  explicit PyFusedSpectrumPeaksROCm(ROCmGPUContext& ctx) : dev_(ctx.device_index()) {
    if (spx_create(dev_, &c_) != SPX_OK) throw std::runtime_error("FusedSpectrumPeaksROCm: spx_create failed");
    spx_set_stream(c_, ctx.backend()->GetNativeQueue());
  }
  ~PyFusedSpectrumPeaksROCm() {
    for (void* p : {in_, out_})
      if (p) (void)hipFree(p);
    if (c_) spx_destroy(c_);
  }

  py::dict process(carr signal, float sample_rate, uint32_t nfft, const std::string& window, float nsigma,
                   uint32_t max_peaks, bool parabolic) {
    auto b = signal.request();
    const uint32_t beams = b.ndim == 2 ? (uint32_t)b.shape[0] : 1;
    const uint32_t n_point = (uint32_t)(b.size / std::max<uint32_t>(beams, 1));
    if (nfft == 0) nfft = n_point;
    uint32_t w = SPX_WINDOW_NONE;
    if (window == "hann") w = SPX_WINDOW_HANN;
    else if (window == "hamming") w = SPX_WINDOW_HAMMING;
    else if (window == "blackman") w = SPX_WINDOW_BLACKMAN;
    else if (window != "none") throw std::invalid_argument("window must be none/hann/hamming/blackman");
    const size_t in_bytes = (size_t)b.size * 8;
    const size_t rec_bytes = (size_t)beams * std::max<uint32_t>(max_peaks, 1) * 32;
    const size_t out_bytes = rec_bytes + (size_t)beams * (8 + 32 + 16);
    grow(&in_, &in_cap_, in_bytes);
    grow(&out_, &out_cap_, out_bytes);
    char* o = (char*)out_;
    spx_fused_desc d;
    std::memset(&d, 0, sizeof(d));
    d.input = in_;
    d.batch = beams;
    d.n_point = n_point;
    d.nfft = nfft;
    d.window = w;
    d.dtype = SPX_DTYPE_C64;
    d.sample_rate = sample_rate;
    d.peak_mode = SPX_PEAKS_THRESHOLD;
    d.nsigma = nsigma;
    d.max_peaks = max_peaks;
    d.refine = parabolic ? SPX_REFINE_PARABOLIC : SPX_REFINE_NONE;
    d.peaks_out = (spx_max_value*)o;
    d.counts_out = (uint32_t*)(o + rec_bytes);
    d.gmax_out = (spx_max_value*)(o + rec_bytes + (size_t)beams * 8);
    d.stats_out = (float*)(o + rec_bytes + (size_t)beams * 40);
    std::vector<char> host(out_bytes);
    {
      py::gil_scoped_release rel;
      check(spx_memcpy_h2d(c_, in_, b.ptr, in_bytes), "upload");
      check(spx_fused_peaks(c_, &d), "spx_fused_peaks");
      check(spx_memcpy_d2h(c_, host.data(), out_, out_bytes), "download");
    }
    py::dtype mv = maxvalue_dtype();
    py::dict out;
    py::array peaks(mv, std::vector<py::ssize_t>{(py::ssize_t)beams, (py::ssize_t)std::max<uint32_t>(max_peaks, 1)});
    std::memcpy(peaks.mutable_data(), host.data(), rec_bytes);
    py::array_t<uint32_t> counts(std::vector<py::ssize_t>{(py::ssize_t)beams, (py::ssize_t)2});
    std::memcpy(counts.mutable_data(), host.data() + rec_bytes, (size_t)beams * 8);
    py::array gmax(mv, std::vector<py::ssize_t>{(py::ssize_t)beams});
    std::memcpy(gmax.mutable_data(), host.data() + rec_bytes + (size_t)beams * 8, (size_t)beams * 32);
    py::array_t<float> stats(std::vector<py::ssize_t>{(py::ssize_t)beams, (py::ssize_t)4});
    std::memcpy(stats.mutable_data(), host.data() + rec_bytes + (size_t)beams * 40, (size_t)beams * 16);
    out["peaks"] = peaks;
    out["counts"] = counts;
    out["global_max"] = gmax;
    out["stats"] = stats;
    return out;
  }

 private:
  void check(int rc, const char* what) {
    if (rc == SPX_OK) return;
    std::string m = std::string(what) + ": " + spx_last_error(c_);
    if (rc == SPX_EINVAL) throw std::invalid_argument(m);
    throw std::runtime_error(m);
  }
  void grow(void** p, size_t* cap, size_t bytes) {
    if (*cap >= bytes) return;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    if (hipMalloc(p, bytes) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    *cap = bytes;
  }
  int dev_;
  spx_ctx* c_ = nullptr;
  void* in_ = nullptr;
  size_t in_cap_ = 0;
  void* out_ = nullptr;
  size_t out_cap_ = 0;
};

PYBIND11_MODULE(dsp_spectrum, m) {
  m.doc() = "dsp_spectrum on MI355X (gfx950): batched FFT, magnitude/phase and spectrum maxima";
  py::class_<ROCmGPUContext>(m, "ROCmGPUContext", "ROCm GPU context (one device, one stream).")
      .def(py::init<int>(), py::arg("device_index") = 0)
      .def_property_readonly("device_name", &ROCmGPUContext::device_name)
      .def_property_readonly("device_index", &ROCmGPUContext::device_index);

  py::class_<PyFFTProcessorROCm>(m, "FFTProcessorROCm", "Batched zero-padded FFT on gfx950 kernels.")
      .def(py::init<ROCmGPUContext&>(), py::arg("ctx"), py::keep_alive<1, 2>())
      .def("process_complex", &PyFFTProcessorROCm::process_complex, py::arg("data"), py::arg("sample_rate"),
           py::arg("beam_count") = 0, py::arg("n_point") = 0)
      .def("process_mag_phase", &PyFFTProcessorROCm::process_mag_phase, py::arg("data"), py::arg("sample_rate"),
           py::arg("beam_count") = 0, py::arg("n_point") = 0, py::arg("include_freq") = true)
      .def("get_profiling", &PyFFTProcessorROCm::get_profiling)
      .def_property_readonly("nfft", &PyFFTProcessorROCm::nfft)
      .def("__repr__", [](const PyFFTProcessorROCm& s) {
        return "<FFTProcessorROCm nfft=" + std::to_string(s.nfft()) + ">";
      });

  py::class_<PySpectrumMaximaFinderROCm>(m, "SpectrumMaximaFinderROCm", "Spectrum maxima search on gfx950 kernels.")
      .def(py::init<ROCmGPUContext&>(), py::arg("ctx"), py::keep_alive<1, 2>())
      .def("process", &PySpectrumMaximaFinderROCm::process, py::arg("signal"), py::arg("n_point"),
           py::arg("sample_rate"), py::arg("antenna_count") = 0, py::arg("repeat_count") = 1)
      .def("find_all_maxima", &PySpectrumMaximaFinderROCm::find_all_maxima, py::arg("fft_data"), py::arg("nFFT"),
           py::arg("sample_rate"), py::arg("beam_count") = 0, py::arg("search_start") = 0, py::arg("search_end") = 0)
      .def("find_all_maxima_from_signal", &PySpectrumMaximaFinderROCm::find_all_maxima_from_signal,
           py::arg("signal"), py::arg("n_point"), py::arg("sample_rate"), py::arg("antenna_count") = 0,
           py::arg("repeat_count") = 1, py::arg("search_start") = 0, py::arg("search_end") = 0)
      .def_property_readonly("initialized", &PySpectrumMaximaFinderROCm::initialized)
      .def("get_params", &PySpectrumMaximaFinderROCm::get_params)
      .def("__repr__", [](const PySpectrumMaximaFinderROCm& s) {
        return std::string("<SpectrumMaximaFinderROCm initialized=") + (s.initialized() ? "True" : "False") + ">";
      });

  py::class_<PyComplexToMagROCm>(m, "ComplexToMagROCm", "Magnitude-only conversion of complex IQ data.")
      .def(py::init<ROCmGPUContext&>(), py::arg("ctx"), py::keep_alive<1, 2>())
      .def("process_magnitude", &PyComplexToMagROCm::process_magnitude, py::arg("data"), py::arg("beam_count") = 0,
           py::arg("norm_coeff") = 1.0f)
      .def("__repr__", [](const PyComplexToMagROCm&) { return std::string("<ComplexToMagROCm>"); });

  py::class_<PyFusedSpectrumPeaksROCm>(m, "FusedSpectrumPeaksROCm",
                                       "One-launch window->FFT->|X|->mean+k*std->local maxima->first K.")
      .def(py::init<ROCmGPUContext&>(), py::arg("ctx"), py::keep_alive<1, 2>())
      .def("process", &PyFusedSpectrumPeaksROCm::process, py::arg("signal"), py::arg("sample_rate"),
           py::arg("nfft") = 0, py::arg("window") = "hann", py::arg("nsigma") = 6.0f, py::arg("max_peaks") = 32,
           py::arg("parabolic") = true);

  m.attr("MAXVALUE_DTYPE") = maxvalue_dtype();

  // Host FFT utilities of the reference module (python/py_cpu_fft.hpp:28-106):
  // spectrum_utils::CpuFFT, computed on the CPU by design, never by the GPU path.
  using farr = py::array_t<float, py::array::c_style | py::array::forcecast>;
  m.def("cpu_fft_c2c", [](carr data) {
    auto b = data.request();
    auto* p = static_cast<std::complex<float>*>(b.ptr);
    return to_numpy(spectrum_utils::CpuFFT::ForwardC2C(cvec(p, p + b.size)));
  }, py::arg("data"), "Forward FFT complex->complex on the host (unnormalised).");
  m.def("cpu_ifft_c2c", [](carr data) {
    auto b = data.request();
    auto* p = static_cast<std::complex<float>*>(b.ptr);
    return to_numpy(spectrum_utils::CpuFFT::InverseC2C(cvec(p, p + b.size)));
  }, py::arg("data"), "Inverse FFT complex->complex on the host, scaled by 1/N.");
  m.def("cpu_fft_r2c", [](farr data) {
    auto b = data.request();
    auto* p = static_cast<float*>(b.ptr);
    return to_numpy(spectrum_utils::CpuFFT::ForwardR2C(std::vector<float>(p, p + b.size)));
  }, py::arg("data"), "Forward FFT real->complex on the host, N/2+1 bins.");
  m.def("cpu_fft_r2c_full", [](farr data) {
    auto b = data.request();
    auto* p = static_cast<float*>(b.ptr);
    return to_numpy(spectrum_utils::CpuFFT::ForwardR2C_Full(std::vector<float>(p, p + b.size)));
  }, py::arg("data"), "Forward FFT real->complex on the host, N bins with the Hermitian mirror.");
  m.def("magnitude", [](carr spectrum, bool squared) {
    auto b = spectrum.request();
    auto* p = static_cast<std::complex<float>*>(b.ptr);
    return to_numpy(spectrum_utils::CpuFFT::Magnitude(cvec(p, p + b.size), squared));
  }, py::arg("spectrum"), py::arg("squared") = false, "|X| (squared=False) or |X|^2 of a host spectrum.");
}
