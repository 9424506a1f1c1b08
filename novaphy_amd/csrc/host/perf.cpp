// perf.cpp -- PerformanceMonitor: per-frame phase timings under the
// reference's names, aggregate stats and Chrome-trace JSON export
// (src/sim/performance_monitor.cpp:83-289 behaviour).
#include <algorithm>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "npx_host.hpp"

namespace novaphy {

void PerformanceMonitor::reset() {
  epoch_ = Clock::now();
  frame_active_ = false;
  frame_count_ = 0;
  last_total_ms_ = 0;
  cur_phase_.clear();
  cur_metrics_.clear();
  last_metrics_.clear();
  cur_dur_.clear();
  agg_.clear();
  frames_.clear();
}

void PerformanceMonitor::begin_frame() {
  cur_phase_.clear();
  cur_metrics_.clear();
  cur_dur_.clear();
  cur_gpu_.clear();
  frame_active_ = enabled || trace_enabled;
  if (frame_active_) {
    frame_start_ = Clock::now();
    cur_gpu_ts_ = us(frame_start_);
  }
}

void PerformanceMonitor::end_frame() {
  if (!frame_active_) return;
  const auto end = Clock::now();
  last_total_ms_ = std::chrono::duration<double, std::milli>(end - frame_start_).count();
  for (const auto& [name, ms] : cur_phase_) {
    Agg& a = agg_[name];
    a.total += ms;
    a.last = ms;
    a.max = std::max(a.max, ms);
    a.samples += 1;
  }
  last_metrics_ = cur_metrics_;
  frame_count_ += 1;
  if (trace_enabled && trace_frame_capacity > 0) {
    Frame f;
    f.dur = cur_dur_;
    f.gpu = cur_gpu_;
    f.gpu_ts = cur_gpu_ts_;
    f.ts = us(end);
    f.metrics = last_metrics_;
    f.metrics.push_back({"frame_count", (double)frame_count_});
    f.metrics.push_back({"frame_total_ms", last_total_ms_});
    frames_.push_back(std::move(f));
    while ((int)frames_.size() > trace_frame_capacity) frames_.pop_front();
  }
  frame_active_ = false;
}

void PerformanceMonitor::record_phase(const std::string& name, Clock::time_point a, Clock::time_point b) {
  if (!frame_active_) return;
  cur_phase_[name] += std::chrono::duration<double, std::milli>(b - a).count();
  if (trace_enabled && trace_frame_capacity > 0)
    cur_dur_.push_back({name, us(a), std::chrono::duration_cast<std::chrono::microseconds>(b - a).count()});
}

void PerformanceMonitor::record_phase_ms(const std::string& name, double ms) {
  if (!frame_active_) return;
  cur_phase_[name] += ms;
  if (trace_enabled && trace_frame_capacity > 0) cur_dur_.push_back({name, us(Clock::now()), (long long)(ms * 1e3)});
}

void PerformanceMonitor::record_gpu_kernel(const std::string& name, double start_ms, double dur_ms, double gbps) {
  if (!frame_active_ || !trace_enabled || trace_frame_capacity <= 0) return;
  cur_gpu_.push_back({name, start_ms, dur_ms, gbps});
}

void PerformanceMonitor::record_metric(std::string_view name, double v) {
  if (!frame_active_) return;
  for (auto& m : cur_metrics_)
    if (m.name == name) {
      m.value = v;
      return;
    }
  cur_metrics_.push_back({std::string(name), v});
}

std::vector<PerformancePhaseStat> PerformanceMonitor::phase_stats() const {
  std::vector<PerformancePhaseStat> out;
  for (const auto& [name, a] : agg_)
    out.push_back({name, a.last, a.samples ? a.total / a.samples : 0.0, a.max, a.samples});
  return out;  // std::map keeps names sorted
}

static std::string esc(const std::string& s) {
  std::ostringstream o;
  for (unsigned char c : s) {
    if (c == '"' || c == '\\')
      o << '\\' << c;
    else if (c < 0x20)
      o << "\\u" << std::hex << std::setw(4) << std::setfill('0') << (int)c << std::dec;
    else
      o << c;
  }
  return o.str();
}

void PerformanceMonitor::write_trace_json(const std::string& path) const {
  std::filesystem::path p(path);
  if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path());
  std::ofstream out(p, std::ios::out | std::ios::trunc);
  if (!out.is_open()) throw std::runtime_error("Failed to open output file: " + path);
  out << "{\n  \"traceEvents\": [\n";
  bool first = true;
  auto sep = [&]() {
    if (!first) out << ",\n";
    first = false;
  };
  bool named = false;
  for (const Frame& f : frames_) {
    for (const Dur& d : f.dur) {
      sep();
      out << "    {\"name\":\"" << esc(d.name) << "\",\"ph\":\"X\",\"ts\":" << d.ts << ",\"dur\":" << d.dur
          << ",\"pid\":1,\"tid\":0}";
    }
    if (!f.gpu.empty() && !named) {  // name the GPU track once
      sep();
      out << "    {\"name\":\"thread_name\",\"ph\":\"M\",\"pid\":1,\"tid\":1,\"args\":{\"name\":\"gpu\"}}";
      named = true;
    }
    for (const GpuK& k : f.gpu) {
      // GPU kernels on their own track, placed from the frame start at their
      // device offsets; algorithmic GB/s as a counter
      const long long ts = f.gpu_ts + (long long)(k.start_ms * 1e3);
      sep();
      out << "    {\"name\":\"" << esc(k.name) << "\",\"ph\":\"X\",\"ts\":" << ts << ",\"dur\":"
          << std::max(1LL, (long long)(k.dur_ms * 1e3)) << ",\"pid\":1,\"tid\":1,\"args\":{\"GB/s\":"
          << std::setprecision(6) << k.gbps << "}}";
      if (k.gbps > 0.0) {
        sep();
        out << "    {\"name\":\"gpu.GB/s\",\"ph\":\"C\",\"ts\":" << ts << ",\"pid\":1,\"tid\":1,\"args\":{\""
            << esc(k.name) << "\":" << std::setprecision(6) << k.gbps << "}}";
      }
    }
    sep();
    out << "    {\"name\":\"frame.metrics\",\"ph\":\"C\",\"ts\":" << f.ts << ",\"pid\":1,\"tid\":0,\"args\":{";
    for (size_t i = 0; i < f.metrics.size(); ++i) {
      if (i) out << ",";
      out << "\"" << esc(f.metrics[i].name) << "\":" << std::setprecision(17) << f.metrics[i].value;
    }
    out << "}}";
  }
  out << "\n  ]\n}\n";
}

}  // namespace novaphy
