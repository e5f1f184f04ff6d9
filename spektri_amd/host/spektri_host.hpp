// C++ host side of the FCFB + Welch path over the C ABI (spektri_gpu.h).
//
// Mirrors the reference's Rust DSP interface so that its main loop reads
// the same way: DspParams / InputBufferSize / DspState::init
// (spektri/src/dsp.rs:24-112), FilterParams / BinNumbers / freq_to_bins*
// (spektri/src/dsp/fcfb.rs:206-285), Metadata and the record and topic
// byte formats (spektri/src/dsp/data.rs:8-171), Output (file part of
// spektri/src/dsp/output.rs:14-62) and the input formats
// (spektri/src/inputformats.rs:12-62).  DspState::submit takes the raw
// bytes of one read_exact (main.rs:176) instead of the converted f32
// buffer with overlap (conversion and the overlap carry run on the GPU) and
// returns before the GPU is done; DspState::write_results then writes that
// buffer's records, which is what process_real / process_complex do in one
// call (dsp.rs:114-160).  ZeroMQ publishing is not part of this build;
// outputs go to files.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <cctype>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "spektri_gpu.h"

namespace spektri {

// ---- input formats (DataFormat codes, data.rs:81-95) ----------------------
enum class InputFormat : uint32_t {
  S8 = SPK_FMT_S8, S16le = SPK_FMT_S16LE, S16be = SPK_FMT_S16BE, F32le = SPK_FMT_F32LE,
  F32be = SPK_FMT_F32BE, U8 = SPK_FMT_U8, Cs8 = SPK_FMT_CS8, Cs16le = SPK_FMT_CS16LE,
  Cs16be = SPK_FMT_CS16BE, Cf32le = SPK_FMT_CF32LE, Cf32be = SPK_FMT_CF32BE, Cu8 = SPK_FMT_CU8,
};

inline bool iequals(const std::string& a, const char* b) {
  if (a.size() != strlen(b)) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
  return true;
}

// Command-line names are the enum variant names, matched case-insensitively
// (clap's arg_enum!, main.rs:31: "--inputformat=s16le").
inline std::optional<InputFormat> parse_input_format(const std::string& s) {
  static const std::pair<const char*, InputFormat> names[] = {
      {"S8", InputFormat::S8},         {"S16le", InputFormat::S16le}, {"S16be", InputFormat::S16be},
      {"F32le", InputFormat::F32le},   {"F32be", InputFormat::F32be}, {"U8", InputFormat::U8},
      {"Cs8", InputFormat::Cs8},       {"Cs16le", InputFormat::Cs16le}, {"Cs16be", InputFormat::Cs16be},
      {"Cf32le", InputFormat::Cf32le}, {"Cf32be", InputFormat::Cf32be}, {"Cu8", InputFormat::Cu8}};
  for (const auto& n : names)
    if (iequals(s, n.first)) return n.second;
  return std::nullopt;
}

// inputformats.rs:12-27
inline size_t bytes_per_input_sample(InputFormat f) {
  switch (f) {
    case InputFormat::U8: case InputFormat::S8: return 1;
    case InputFormat::S16le: case InputFormat::S16be: case InputFormat::Cu8: case InputFormat::Cs8: return 2;
    case InputFormat::F32le: case InputFormat::F32be: case InputFormat::Cs16le: case InputFormat::Cs16be: return 4;
    default: return 8;  // Cf32le, Cf32be
  }
}
// inputformats.rs:29-44 (bit 6 of the code marks complex formats)
inline bool is_input_format_complex(InputFormat f) { return ((uint32_t)f & 0x40u) != 0; }
// inputformats.rs:47-62
inline float input_format_scaling(InputFormat f) {
  switch (f) {
    case InputFormat::U8: case InputFormat::S8: case InputFormat::Cu8: case InputFormat::Cs8: return 1.0f / 127.0f;
    case InputFormat::S16le: case InputFormat::S16be: case InputFormat::Cs16le: case InputFormat::Cs16be:
      return 1.0f / 32767.0f;
    default: return 1.0f;
  }
}

// ---- DSP parameters (dsp.rs:24-42, spectrum.rs:29-32, fcfb.rs:206-216) ----
enum class SpectrumFormat { U8, U16 };

struct OutputParams {
  std::optional<std::string> filename;
};

struct FilterParams {
  double fs_out;  // output sample rate
  double fc_out;  // output center frequency
  OutputParams output;
};

struct DspParams {
  bool complex = false;
  double fs_in = 0.0;
  double fc_in = 0.0;
  size_t fft_size = 16384;
  float scaling = 1.0f;
  size_t ffts_per_buf = 8;
  SpectrumFormat spectrum_format = SpectrumFormat::U8;
  uint32_t spectrum_averages = 2000;
  std::vector<FilterParams> filters;
};

struct InputBufferSize {
  size_t new_samples;  // `new` in the reference
  size_t overlap;
  size_t total;
};

struct FftInfo {
  double fs, fc;
  size_t size;
  bool complex;
};

struct BinNumbers {
  size_t bins;
  ptrdiff_t first;
};

// fcfb.rs:225-285, through the library (one implementation of the rounding)
inline std::optional<BinNumbers> freq_to_bins(const FftInfo& fi, double fs_out, double fc_out) {
  uint32_t bins = 0;
  int64_t first = 0;
  if (spk_freq_to_bins(fi.fs, fi.fc, (uint32_t)fi.size, fs_out, fc_out, &bins, &first) != SPK_OK) return std::nullopt;
  return BinNumbers{bins, (ptrdiff_t)first};
}
inline std::pair<double, double> bins_to_freq(const FftInfo& fi, const BinNumbers& bn) {
  double fs = 0, fc = 0;
  spk_bins_to_freq(fi.fs, fi.fc, (uint32_t)fi.size, (uint32_t)bn.bins, (int64_t)bn.first, &fs, &fc);
  return {fs, fc};
}
inline std::optional<BinNumbers> freq_to_bins_exact(const FftInfo& fi, double fs_out, double fc_out) {
  uint32_t bins = 0;
  int64_t first = 0;
  if (spk_freq_to_bins_exact(fi.fs, fi.fc, (uint32_t)fi.size, fs_out, fc_out, &bins, &first) != SPK_OK)
    return std::nullopt;
  return BinNumbers{bins, (ptrdiff_t)first};
}
// Fcfb::nearest_freq (fcfb.rs:108-115)
inline std::optional<std::pair<double, double>> nearest_freq(const FftInfo& fi, double fs_out, double fc_out) {
  auto bn = freq_to_bins(fi, fs_out, fc_out);
  if (!bn) return std::nullopt;
  return bins_to_freq(fi, *bn);
}

// ---- records and topics (data.rs:8-171) -----------------------------------
struct Metadata {
  uint64_t seq;        // sequence number of the processing block
  timespec systemtime;  // when the block was received
};

inline void put_le(std::vector<uint8_t>& b, uint64_t v, int n) {
  for (int i = 0; i < n; ++i) b.push_back((uint8_t)(v >> (8 * i)));
}
inline void put_f64(std::vector<uint8_t>& b, double v) {
  uint64_t u;
  memcpy(&u, &v, 8);
  put_le(b, u, 8);
}

// data.rs:110-132: u64 seq, u64 seconds, u32 nanoseconds, u32 reserved (LE)
inline void serialize_metadata(std::vector<uint8_t>& b, const Metadata& m, uint64_t seq) {
  put_le(b, seq, 8);
  put_le(b, (uint64_t)m.systemtime.tv_sec, 8);
  put_le(b, (uint64_t)m.systemtime.tv_nsec, 4);
  put_le(b, 0, 4);
}

constexpr uint8_t kProtocolVersion = 2;  // data.rs:99
// data.rs:137-153
inline std::vector<uint8_t> serialize_signal_topic(double fs, double fc) {
  std::vector<uint8_t> b = {kProtocolVersion, 0x40, (uint8_t)SPK_FMT_CF32LE, 0, 0, 0, 0, 0};
  put_f64(b, fs);
  put_f64(b, fc);
  return b;
}
// data.rs:156-171 (the format byte is U8 for both spectrum formats)
inline std::vector<uint8_t> serialize_spectrum_topic(double fd, double f0) {
  std::vector<uint8_t> b = {kProtocolVersion, 0x60, (uint8_t)SPK_FMT_U8, 0, 0, 0, 0, 0};
  put_f64(b, fd);
  put_f64(b, f0);
  return b;
}

// The file part of Output (output.rs:14-62): records are appended to the
// file; a file that cannot be created is reported and skipped.
class Output {
 public:
  Output(const OutputParams& p, std::vector<uint8_t> topic) : topic_(std::move(topic)) {
    if (p.filename) {
      file_ = fopen(p.filename->c_str(), "wb");
      if (!file_) fprintf(stderr, "Could not create file: %s\n", p.filename->c_str());
    }
  }
  Output(const Output&) = delete;
  Output(Output&& o) noexcept : file_(o.file_), topic_(std::move(o.topic_)) { o.file_ = nullptr; }
  ~Output() {
    if (file_) fclose(file_);
  }
  void write(const uint8_t* data, size_t n) {
    if (file_) {
      fwrite(data, 1, n, file_);
      fflush(file_);
    }
  }
  void write(const std::vector<uint8_t>& buf) { write(buf.data(), buf.size()); }
  const std::vector<uint8_t>& topic() const { return topic_; }

 private:
  FILE* file_ = nullptr;
  std::vector<uint8_t> topic_;
};

// ---- DspState (dsp.rs:44-160 over the C ABI) ------------------------------
class DspState {
 public:
  static std::pair<std::unique_ptr<DspState>, InputBufferSize> init(const DspParams& p, InputFormat fmt) {
    spk_config cfg{};
    cfg.fft_size = (uint32_t)p.fft_size;
    cfg.input_format = (uint32_t)fmt;
    cfg.scaling = p.scaling;
    cfg.ffts_per_buf = (uint32_t)p.ffts_per_buf;
    cfg.spectrum_averages = p.spectrum_averages;
    cfg.max_blocks = 0;
    cfg.flags = p.spectrum_format == SpectrumFormat::U8 ? SPK_FLAG_SPECTRUM_U8 : SPK_FLAG_SPECTRUM_U16;
    spk_ctx* ctx = nullptr;
    const spk_status st = spk_create(&cfg, &ctx);
    if (st != SPK_OK)
      throw std::runtime_error("spk_create failed (status " + std::to_string((int)st) + ")");
    std::unique_ptr<DspState> d(new DspState(ctx, p));
    spk_buffer_size bs{};
    d->check(spk_buffer_size_of(ctx, &bs));
    for (const auto& f : p.filters) d->add_filter(f);
    // SpectrumAccumulator::init (spectrum.rs:52-79): records go to standard output
    d->spectrum_out_.reset(new Output(OutputParams{std::string("/dev/stdout")},
                                      serialize_spectrum_topic(p.fs_in / (double)p.fft_size, p.fc_in)));
    return {std::move(d), InputBufferSize{bs.new_samples, bs.overlap, bs.total}};
  }

  ~DspState() {
    if (ctx_) spk_destroy(ctx_);
  }

  // Fcfb::add_filter (fcfb.rs:41-64): invalid parameters are reported and skipped.
  void add_filter(const FilterParams& p) {
    const auto bn = freq_to_bins_exact(info_, p.fs_out, p.fc_out);
    if (!bn) {
      fprintf(stderr, "Invalid filter parameters\n");
      return;
    }
    uint32_t id = 0;
    const spk_status st = spk_add_channel(ctx_, (int64_t)bn->first, (uint32_t)bn->bins, nullptr, &id);
    if (st != SPK_OK) {
      fprintf(stderr, "Error creating filter: %s\n", spk_last_error(ctx_));
      return;
    }
    filters_.push_back(Filter{id, bn->bins, Output(p.output, serialize_signal_topic(p.fs_out, p.fc_out))});
  }

  // The pinned buffer the next process_raw call uploads from; read_exact
  // into it to skip a host copy.
  uint8_t* input_buffer(size_t* nbytes) {
    void* ptr = nullptr;
    check(spk_input_buffer(ctx_, &ptr, nbytes));
    return static_cast<uint8_t*>(ptr);
  }

  // Enqueue one buffer of raw input (process_real / process_complex,
  // dsp.rs:114-160) without waiting; write_results() emits its records.
  void submit(const uint8_t* raw, size_t nbytes, const Metadata& m) {
    check(spk_process_async(ctx_, raw, nbytes));
    pending_ = m;
  }

  // Records of the last submitted buffer, as the reference writes them:
  // per filter one record of metadata (seq = buffer seq) and F blocks of
  // samples (fcfb.rs:66-100), and every finished Welch record with
  // metadata (seq = record counter) and the quantised bins
  // (spectrum.rs:113-160).
  void write_results() {
    if (!pending_) return;
    const Metadata m = *pending_;
    pending_.reset();
    // Channel records arrive packed with their metadata from the GPU, one
    // per filter in id order (filters_ is in id order: ids are handed out
    // in sequence and only accepted filters are kept).
    const uint64_t secs = (uint64_t)m.systemtime.tv_sec;
    const uint32_t nanos = (uint32_t)m.systemtime.tv_nsec;
    if (!filters_.empty()) {
      uint32_t nch = 0;
      size_t cbytes = 0;
      check_or_capacity(spk_channel_records_packed(ctx_, m.seq, secs, nanos, nullptr, 0, &nch, &cbytes));
      chan_.resize(cbytes);
      check(spk_channel_records_packed(ctx_, m.seq, secs, nanos, chan_.data(), chan_.size(), &nch, &cbytes));
      if (nch != filters_.size()) throw std::runtime_error("channel record count mismatch");
      uint32_t nb = 0;
      check(spk_last_blocks(ctx_, &nb));
      size_t o = 0;
      for (auto& f : filters_) {
        const size_t size = 24 + (size_t)8 * nb * (6 * (f.bins / 8));
        f.output.write(chan_.data() + o, size);
        o += size;
      }
      if (o != cbytes) throw std::runtime_error("channel record size mismatch");
    }
    // Spectrum records arrive packed with their metadata from the GPU.
    uint32_t nrec = 0;
    size_t nbytes = 0;
    check_or_capacity(spk_spectrum_records_packed(ctx_, spectrum_seq_, secs, nanos, nullptr, 0, &nrec, &nbytes));
    if (!nrec) return;
    spec_.resize(nbytes);
    check(spk_spectrum_records_packed(ctx_, spectrum_seq_, secs, nanos, spec_.data(), spec_.size(), &nrec, &nbytes));
    spectrum_seq_ += nrec;
    const size_t each = nbytes / nrec;
    for (uint32_t r = 0; r < nrec; ++r) spectrum_out_->write(spec_.data() + (size_t)r * each, each);
  }

  size_t filter_count() const { return filters_.size(); }

 private:
  struct Filter {
    uint32_t id;
    size_t bins;
    Output output;
  };

  DspState(spk_ctx* ctx, const DspParams& p)
      : ctx_(ctx), params_(p), info_{p.fs_in, p.fc_in, p.fft_size, p.complex} {}

  void check(spk_status st) {
    if (st != SPK_OK) throw std::runtime_error(std::string("spektri GPU call failed: ") + spk_last_error(ctx_));
  }
  // A size query (cap 0) reports SPK_ERR_CAPACITY when there is anything to copy.
  void check_or_capacity(spk_status st) {
    if (st != SPK_ERR_CAPACITY) check(st);
  }

  spk_ctx* ctx_ = nullptr;
  DspParams params_;
  FftInfo info_;
  std::vector<Filter> filters_;
  std::unique_ptr<Output> spectrum_out_;
  std::optional<Metadata> pending_;
  uint64_t spectrum_seq_ = 0;
  std::vector<uint8_t> chan_;
  std::vector<uint8_t> spec_;
};

}  // namespace spektri
