// Cost models, hardware profile (with JSON persistence) and fusion config.
#include "cost_model.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>

namespace cast {

double SizeOnlyCostModel::cost_of(const QuantumGate &g) const {
  const size_t size = trajectory_mode ? g.n_qubits() : g.effective_n_qubits();
  if (size > max_size)
    return 1.0;
  if (g.arithmetic_intensity(zero_tol) > double(max_ai))
    return 1.0;
  return 1e-10;
}

double compute_bw_slope(double peak_bw_gib_s, size_t scalar_bytes) {
  return peak_bw_gib_s * double(uint64_t(1) << 30) / (2.0 * double(scalar_bytes)) / 1e9;
}

double compute_crossover(double peak_gflops_s, double bw_slope) {
  if (bw_slope > 0.0 && peak_gflops_s > 0.0)
    return peak_gflops_s / bw_slope;
  return std::numeric_limits<double>::infinity();
}

HardwareProfile HardwareProfile::from_roofline(ProfileConfig c, double bw, double gf,
                                               size_t scalar_bytes) {
  HardwareProfile p;
  p.config = std::move(c);
  p.peak_bw_gib_s = bw;
  p.peak_gflops_s = gf;
  p.bw_slope = compute_bw_slope(bw, scalar_bytes);
  p.crossover_ai = compute_crossover(gf, p.bw_slope);
  return p;
}

HardwareProfile HardwareProfile::from_measurements(ProfileConfig c, double bw, double gf,
                                                   double slope) {
  HardwareProfile p;
  p.config = std::move(c);
  p.peak_bw_gib_s = bw;
  p.peak_gflops_s = gf;
  p.bw_slope = slope;
  p.crossover_ai = compute_crossover(gf, slope);
  return p;
}

double HardwareAdaptiveCostModel::cost_of(const QuantumGate &g) const {
  const size_t size = trajectory_mode ? g.n_qubits() : g.effective_n_qubits();
  if (size > max_size)
    return std::numeric_limits<double>::infinity();
  return std::max(1.0, g.arithmetic_intensity(zero_tol) / crossover_ai);
}

FusionConfig FusionConfig::size_only(size_t max_size) {
  FusionConfig c;
  c.size_max = max_size;
  c.cost_model = std::make_shared<SizeOnlyCostModel>(max_size, ~size_t(0), 0.0, false);
  return c;
}

FusionConfig FusionConfig::size_only_trajectory(size_t max_size) {
  FusionConfig c;
  c.size_max = max_size;
  c.trajectory_mode = true;
  c.cost_model = std::make_shared<SizeOnlyCostModel>(max_size, ~size_t(0), 0.0, true);
  return c;
}

FusionConfig FusionConfig::hardware_adaptive(const HardwareProfile &p, size_t max_size,
                                             double zero_tol) {
  FusionConfig c;
  c.size_max = max_size;
  c.zero_tol = zero_tol;
  c.cost_model = std::make_shared<HardwareAdaptiveCostModel>(p, max_size, false, zero_tol);
  return c;
}

FusionConfig FusionConfig::hardware_adaptive_trajectory(const HardwareProfile &p,
                                                        size_t max_size, double zero_tol) {
  FusionConfig c;
  c.size_max = max_size;
  c.zero_tol = zero_tol;
  c.trajectory_mode = true;
  c.cost_model = std::make_shared<HardwareAdaptiveCostModel>(p, max_size, true, zero_tol);
  return c;
}

size_t FusionLog::n_accepted() const {
  size_t n = 0;
  for (const auto &d : decisions)
    n += d.accepted;
  return n;
}

size_t FusionLog::n_rejected() const { return decisions.size() - n_accepted(); }

std::string FusionLog::summary() const {
  std::ostringstream os;
  os << "  Fusion log: " << decisions.size() << " decisions (" << n_accepted() << " accepted, "
     << n_rejected() << " rejected)\n";
  char line[160];
  std::snprintf(line, sizeof(line), "  %8s %6s %8s %10s %10s %10s %8s\n", "Size", "NCand",
                "ProdQ", "OldCost", "NewCost", "Benefit", "Accept");
  os << line;
  for (const auto &d : decisions) {
    std::snprintf(line, sizeof(line), "  %8zu %6zu %8zu %10.3f %10.3f %10.3f %8s\n", d.cdd_size,
                  d.n_candidates, d.product_n_qubits, d.old_cost, d.new_cost, d.benefit,
                  d.accepted ? "yes" : "no");
    os << line;
  }
  return os.str();
}

// ── JSON (the HardwareProfile schema of cost_model.rs:128-225, plus a "hip"
//    device variant) ──────────────────────────────────────────────────────

namespace {

std::string jstr(const std::string &s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\')
      o += '\\';
    o += c;
  }
  return o + "\"";
}

std::string jnum(double v) {
  if (std::isinf(v))
    return "1e308";
  char b[64];
  std::snprintf(b, sizeof(b), "%.17g", v);
  return b;
}

struct JVal {
  enum Kind { Null, Num, Str, Obj, Arr, Bool } kind = Null;
  double num = 0;
  std::string str;
  std::map<std::string, JVal> obj;
  std::vector<JVal> arr;
  const JVal &at(const std::string &k) const {
    auto it = obj.find(k);
    if (it == obj.end())
      throw Error("profile JSON: missing key '" + k + "'");
    return it->second;
  }
  bool has(const std::string &k) const { return obj.count(k) != 0; }
};

struct JParser {
  const std::string &s;
  size_t i = 0;
  void ws() {
    while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i])))
      ++i;
  }
  char peek() {
    ws();
    if (i >= s.size())
      throw Error("profile JSON: unexpected end");
    return s[i];
  }
  void expect(char c) {
    if (peek() != c)
      throw Error(std::string("profile JSON: expected '") + c + "'");
    ++i;
  }
  JVal parse() {
    const char c = peek();
    JVal v;
    if (c == '{') {
      v.kind = JVal::Obj;
      ++i;
      if (peek() == '}') {
        ++i;
        return v;
      }
      for (;;) {
        JVal k = parse();
        if (k.kind != JVal::Str)
          throw Error("profile JSON: object key must be a string");
        expect(':');
        v.obj[k.str] = parse();
        if (peek() == ',') {
          ++i;
          continue;
        }
        expect('}');
        return v;
      }
    }
    if (c == '[') {
      v.kind = JVal::Arr;
      ++i;
      if (peek() == ']') {
        ++i;
        return v;
      }
      for (;;) {
        v.arr.push_back(parse());
        if (peek() == ',') {
          ++i;
          continue;
        }
        expect(']');
        return v;
      }
    }
    if (c == '"') {
      v.kind = JVal::Str;
      ++i;
      while (i < s.size() && s[i] != '"') {
        if (s[i] == '\\' && i + 1 < s.size())
          ++i;
        v.str += s[i++];
      }
      if (i >= s.size())
        throw Error("profile JSON: unterminated string");
      ++i;
      return v;
    }
    if (s.compare(i, 4, "true") == 0 || s.compare(i, 5, "false") == 0) {
      v.kind = JVal::Bool;
      v.num = s[i] == 't';
      i += s[i] == 't' ? 4 : 5;
      return v;
    }
    if (s.compare(i, 4, "null") == 0) {
      i += 4;
      return v;
    }
    size_t end = i;
    while (end < s.size() && (std::isdigit(static_cast<unsigned char>(s[end])) || s[end] == '-' ||
                              s[end] == '+' || s[end] == '.' || s[end] == 'e' || s[end] == 'E'))
      ++end;
    if (end == i)
      throw Error("profile JSON: unexpected character");
    v.kind = JVal::Num;
    v.num = std::stod(s.substr(i, end - i));
    i = end;
    return v;
  }
};

} // namespace

std::string HardwareProfile::to_json() const {
  std::ostringstream os;
  os << "{\n  \"config\": {\n    \"device\": {\"backend\": " << jstr(config.backend)
     << ", \"name\": " << jstr(config.name);
  if (config.backend == "cpu")
    os << ", \"n_threads\": " << config.n_threads;
  else if (config.backend == "cuda")
    os << ", \"sm_major\": " << config.sm_major << ", \"sm_minor\": " << config.sm_minor;
  else
    os << ", \"arch\": " << jstr(config.arch);
  os << "},\n    \"precision\": " << jstr(config.precision) << ",\n    \"n_qubits\": "
     << config.n_qubits << "\n  },\n";
  os << "  \"peak_bw_gib_s\": " << jnum(peak_bw_gib_s) << ",\n";
  os << "  \"peak_gflops_s\": " << jnum(peak_gflops_s) << ",\n";
  os << "  \"bw_slope\": " << jnum(bw_slope) << ",\n";
  os << "  \"crossover_ai\": " << jnum(crossover_ai) << ",\n";
  os << "  \"raw\": [";
  for (size_t i = 0; i < raw.size(); ++i)
    os << (i ? ",\n    " : "\n    ") << "{\"ai\": " << jnum(raw[i].ai)
       << ", \"gflops_s\": " << jnum(raw[i].gflops_s) << ", \"gib_s\": " << jnum(raw[i].gib_s)
       << "}";
  os << (raw.empty() ? "]\n}\n" : "\n  ]\n}\n");
  return os.str();
}

HardwareProfile HardwareProfile::from_json(const std::string &text) {
  JParser p{text};
  const JVal root = p.parse();
  HardwareProfile h;
  const JVal &cfg = root.at("config");
  const JVal &dev = cfg.at("device");
  h.config.backend = dev.at("backend").str;
  if (dev.has("name"))
    h.config.name = dev.at("name").str;
  if (dev.has("arch"))
    h.config.arch = dev.at("arch").str;
  if (dev.has("n_threads"))
    h.config.n_threads = uint32_t(dev.at("n_threads").num);
  if (dev.has("sm_major"))
    h.config.sm_major = uint32_t(dev.at("sm_major").num);
  if (dev.has("sm_minor"))
    h.config.sm_minor = uint32_t(dev.at("sm_minor").num);
  h.config.precision = cfg.at("precision").str;
  h.config.n_qubits = uint32_t(cfg.at("n_qubits").num);
  h.peak_bw_gib_s = root.at("peak_bw_gib_s").num;
  h.peak_gflops_s = root.at("peak_gflops_s").num;
  h.bw_slope = root.at("bw_slope").num;
  h.crossover_ai = root.at("crossover_ai").num;
  if (root.has("raw"))
    for (const auto &e : root.at("raw").arr)
      h.raw.push_back({e.at("ai").num, e.at("gflops_s").num, e.at("gib_s").num});
  return h;
}

void HardwareProfile::save(const std::string &path) const {
  std::ofstream f(path);
  if (!f)
    throw Error("cannot open " + path + " for writing");
  f << to_json();
}

HardwareProfile HardwareProfile::load(const std::string &path) {
  std::ifstream f(path);
  if (!f)
    throw Error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return from_json(ss.str());
}

} // namespace cast
