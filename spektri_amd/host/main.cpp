// spektri-gpu: the reference's command-line program (spektri/src/main.rs)
// with the DSP on the GPU.  Same options, same input framing (read_exact of
// F * 3N/4 raw samples from standard input, main.rs:113-193), same record
// bytes: spectrum records on standard output (spectrum.rs:75-78) and
// channel records in the files named by `file=` (output.rs:24-62).
// ZeroMQ publishing is not built: --zmqbind is accepted and reported.
//
// The loop overlaps the read of buffer i+1 (straight into the library's
// pinned slot) with the GPU work of buffer i, then writes buffer i's
// records and submits buffer i+1.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <map>
#include <string>
#include <vector>

#include "spektri_host.hpp"

namespace {

void usage() {
  fprintf(stderr,
          "usage: spektri-gpu [options]\n"
          "  -s, --samplerate=HZ          Input sample rate\n"
          "  -f, --centerfreq=HZ          Input center frequency (default 0)\n"
          "  -n, --fftsize=SIZE           FFT size (default 16384)\n"
          "      --fftbuf=NUMBER          Number of FFTs in each input buffer (default 8)\n"
          "  -a, --averages=NUMBER        Number of FFTs averaged for spectrum (default 2000)\n"
          "  -I, --inputformat=FORMAT     Input signal format (S8 S16le S16be F32le F32be U8\n"
          "                               Cs8 Cs16le Cs16be Cf32le Cf32be Cu8)\n"
          "      --spectrumformat=FORMAT  Spectrum output format (U8 or U16, default U8)\n"
          "      --filters=PARAMETERS...  Filter parameters fs=HZ:fc=HZ[:file=PATH]\n"
          "      --zmqbind=ADDRESS...     ZeroMQ binding addresses (not built here)\n");
}

[[noreturn]] void die(const std::string& msg) {
  fprintf(stderr, "error: %s\n", msg.c_str());
  usage();
  exit(2);
}

double to_f64(const std::string& opt, const std::string& v) {
  char* end = nullptr;
  const double d = strtod(v.c_str(), &end);
  if (v.empty() || *end) die("invalid value '" + v + "' for " + opt);
  return d;
}
unsigned long to_uint(const std::string& opt, const std::string& v) {
  char* end = nullptr;
  const unsigned long u = strtoul(v.c_str(), &end, 10);
  if (v.empty() || *end || v[0] == '-') die("invalid value '" + v + "' for " + opt);
  return u;
}

// parse_filter_params (main.rs:69-90): key=value pairs separated by ':'.
spektri::FilterParams parse_filter_params(const std::string& s) {
  std::map<std::string, std::string> m;
  size_t pos = 0;
  while (pos <= s.size()) {
    const size_t end = s.find(':', pos);
    const std::string kv = s.substr(pos, end == std::string::npos ? std::string::npos : end - pos);
    const size_t eq = kv.find('=');
    if (eq == std::string::npos) die("filter parameter '" + kv + "' is not key=value");
    m[kv.substr(0, eq)] = kv.substr(eq + 1);
    if (end == std::string::npos) break;
    pos = end + 1;
  }
  if (!m.count("fs") || !m.count("fc")) die("filter '" + s + "' needs fs and fc");
  spektri::FilterParams f{to_f64("fs", m["fs"]), to_f64("fc", m["fc"]), {}};
  if (m.count("file")) f.output.filename = m["file"];
  return f;
}

bool read_exact(FILE* in, uint8_t* buf, size_t n) {
  size_t got = 0;
  while (got < n) {
    const size_t r = fread(buf + got, 1, n - got, in);
    if (r == 0) return false;
    got += r;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  spektri::DspParams p;
  std::optional<spektri::InputFormat> fmt;
  bool have_fs = false;
  std::vector<std::string> zmqbind;

  // option name -> value, in the forms --name=value, --name value, -x value, -xvalue
  std::vector<std::string> args(argv + 1, argv + argc);
  for (size_t i = 0; i < args.size(); ++i) {
    std::string a = args[i], name, val;
    bool has_val = false;
    if (a == "-h" || a == "--help") {
      usage();
      return 0;
    }
    if (a.rfind("--", 0) == 0) {
      const size_t eq = a.find('=');
      name = a.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
      if (eq != std::string::npos) {
        val = a.substr(eq + 1);
        has_val = true;
      }
    } else if (a.size() >= 2 && a[0] == '-') {
      static const std::map<char, std::string> shorts = {
          {'s', "samplerate"}, {'f', "centerfreq"}, {'n', "fftsize"}, {'a', "averages"}, {'I', "inputformat"}};
      const auto it = shorts.find(a[1]);
      if (it == shorts.end()) die("unknown option " + a);
      name = it->second;
      if (a.size() > 2) {
        val = a.substr(a[2] == '=' ? 3 : 2);
        has_val = true;
      }
    } else {
      die("unexpected argument " + a);
    }
    const bool multi = name == "filters" || name == "zmqbind";
    std::vector<std::string> vals;
    if (has_val) vals.push_back(val);
    if (!has_val || multi) {
      while (i + 1 < args.size() && !(args[i + 1].size() > 1 && args[i + 1][0] == '-')) {
        vals.push_back(args[++i]);
        if (!multi) break;
      }
    }
    if (vals.empty()) die("option --" + name + " needs a value");
    if (name == "samplerate") {
      p.fs_in = to_f64(name, vals[0]);
      have_fs = true;
    } else if (name == "centerfreq") {
      p.fc_in = to_f64(name, vals[0]);
    } else if (name == "fftsize") {
      p.fft_size = to_uint(name, vals[0]);
    } else if (name == "fftbuf") {
      p.ffts_per_buf = to_uint(name, vals[0]);
    } else if (name == "averages") {
      p.spectrum_averages = (uint32_t)to_uint(name, vals[0]);
    } else if (name == "inputformat") {
      fmt = spektri::parse_input_format(vals[0]);
      if (!fmt) die("invalid input format '" + vals[0] + "'");
    } else if (name == "spectrumformat") {
      if (spektri::iequals(vals[0], "U8")) p.spectrum_format = spektri::SpectrumFormat::U8;
      else if (spektri::iequals(vals[0], "U16")) p.spectrum_format = spektri::SpectrumFormat::U16;
      else die("invalid spectrum format '" + vals[0] + "'");
    } else if (name == "filters") {
      for (const auto& v : vals) p.filters.push_back(parse_filter_params(v));
    } else if (name == "zmqbind") {
      zmqbind = vals;
    } else {
      die("unknown option --" + name);
    }
  }
  // main.rs:31,39: the input format and the sample rate are required
  if (!fmt) die("--inputformat is required");
  if (!have_fs) die("--samplerate is required");
  if (!zmqbind.empty()) fprintf(stderr, "note: ZeroMQ output is not built; use file= outputs\n");
  p.complex = spektri::is_input_format_complex(*fmt);
  p.scaling = spektri::input_format_scaling(*fmt);

  try {
    auto [dsp, bufsize] = spektri::DspState::init(p, *fmt);
    const size_t nbytes = spektri::bytes_per_input_sample(*fmt) * bufsize.new_samples;
    uint64_t seq = 0;
    size_t slot_bytes = 0;
    uint8_t* slot = dsp->input_buffer(&slot_bytes);
    if (slot_bytes != nbytes) throw std::runtime_error("input slot size mismatch");
    if (!read_exact(stdin, slot, nbytes)) return 0;
    timespec now{};
    clock_gettime(CLOCK_REALTIME, &now);
    dsp->submit(slot, nbytes, spektri::Metadata{seq, now});
    for (;;) {
      slot = dsp->input_buffer(&slot_bytes);  // the other pinned slot
      const bool more = read_exact(stdin, slot, nbytes);
      dsp->write_results();  // waits for the GPU work of the last buffer
      if (!more) break;
      clock_gettime(CLOCK_REALTIME, &now);
      dsp->submit(slot, nbytes, spektri::Metadata{++seq, now});
    }
  } catch (const std::exception& e) {
    fprintf(stderr, "spektri-gpu: %s\n", e.what());
    return 1;
  }
  return 0;
}
