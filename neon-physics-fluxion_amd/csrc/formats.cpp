// On-disk formats of the HAPR pipeline: the BLIF reader
// (HAPR_fluxion_v4/src/parser/blif.rs:13-122) and the DEF writer
// (HAPR_fluxion_v4/src/parser/def.rs:12-41). Host code only.
//
// BLIF semantics follow blif.rs line by line: a gate per .names / .latch /
// .gate line named g<k> in file order (.latch sequential, the rest
// combinational; 2.0 x 10.0); .names drives its last token and reads the
// others, .latch reads token 1 and drives token 2, .gate drives the nets of
// its O= / Y= pins and reads every other name=net pin. A later driver of a
// net name replaces an earlier one (HashMap::insert). A net is its driver
// followed by its readers in gate order; nets of one gate are dropped.
// Readers-only nets (primary inputs) follow with two or more readers.
//
// Net order. The reference creates nets by iterating two HashMaps
// (blif.rs:96-116), whose order changes from run to run. Here driven nets
// come first, in the order their names were first driven, then readers-only
// nets in the order their names were first read. Any fixed order gives the
// same nets; the order only decides the adjacency lists' neighbour order.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <unordered_map>
#include <vector>

#include "fluxion_hip.h"

struct fx_circuit {
  std::string name;
  float chip_width = 1000.0f, chip_height = 1000.0f, row_height = 10.0f;
  std::vector<std::string> gate_names;
  std::vector<uint8_t> gate_types;  // 0 combinational, 1 sequential
  std::vector<std::string> net_names;
  std::vector<int32_t> net_ptr{0};
  std::vector<int32_t> net_pins;
};

namespace fx {
// Thread-local error record of the library (fx_last_error).
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
}  // namespace fx

namespace {

int set_err(int code, const std::string& msg) { return fx::fail(code, "%s", msg.c_str()); }

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && is_space(s[i])) ++i;
    size_t j = i;
    while (j < s.size() && !is_space(s[j])) ++j;
    if (j > i) out.emplace_back(s, i, j - i);
    i = j;
  }
  return out;
}

// Path::file_stem: the file name without its last extension; a name that
// starts with its only dot is kept whole.
std::string file_stem(const std::string& path) {
  size_t b = path.find_last_of('/');
  std::string f = b == std::string::npos ? path : path.substr(b + 1);
  if (f == "..") return f;
  const size_t d = f.find_last_of('.');
  if (d == std::string::npos || d == 0) return f;
  return f.substr(0, d);
}

}  // namespace

extern "C" int fx_blif_read(const char* path, fx_circuit** out) {
  fx::clear_error();
  if (!path || !out) return set_err(FX_E_INVALID, "fx_blif_read: bad arguments");
  *out = nullptr;
  std::ifstream in(path, std::ios::binary);
  if (!in) return set_err(FX_E_INVALID, std::string("Failed to open BLIF: ") + path);
  auto* c = new fx_circuit();
  c->name = file_stem(path);
  // net name -> index into `nets` (first-seen order); driver per net
  std::unordered_map<std::string, int32_t> net_id;
  std::vector<std::string> nets;
  std::vector<int32_t> driver;     // -1: not driven
  std::vector<int32_t> first_drv;  // sequence number of the first drive, -1 if never
  std::vector<int32_t> first_read;
  std::vector<std::vector<int32_t>> readers;
  int32_t drive_seq = 0, read_seq = 0;
  auto id_of = [&](const std::string& s) {
    auto it = net_id.find(s);
    if (it != net_id.end()) return it->second;
    const int32_t id = (int32_t)nets.size();
    net_id.emplace(s, id);
    nets.push_back(s);
    driver.push_back(-1);
    first_drv.push_back(-1);
    first_read.push_back(-1);
    readers.emplace_back();
    return id;
  };
  auto drive = [&](const std::string& s, int32_t g) {
    const int32_t id = id_of(s);
    driver[id] = g;
    if (first_drv[id] < 0) first_drv[id] = drive_seq++;
  };
  auto read = [&](const std::string& s, int32_t g) {
    const int32_t id = id_of(s);
    readers[id].push_back(g);
    if (first_read[id] < 0) first_read[id] = read_seq++;
  };
  std::string line;
  int64_t lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const std::vector<std::string> tok = split_ws(line);
    if (tok.empty() || tok[0][0] == '#') continue;
    const std::string& d = tok[0];
    if (d != ".names" && d != ".latch" && d != ".gate") continue;
    const int32_t g = (int32_t)c->gate_names.size();
    if (g == std::numeric_limits<int32_t>::max()) {
      delete c;
      return set_err(FX_E_INVALID, "BLIF: too many gates");
    }
    c->gate_names.push_back("g" + std::to_string(g));
    c->gate_types.push_back(d == ".latch" ? 1 : 0);
    if (d == ".names") {
      if (tok.size() >= 2) {
        // Inputs are recorded before the output is driven; both orders give
        // the same nets.
        for (size_t k = 1; k + 1 < tok.size(); ++k) read(tok[k], g);
        drive(tok.back(), g);
      }
    } else if (d == ".latch") {
      if (tok.size() >= 3) {
        read(tok[1], g);
        drive(tok[2], g);
      }
    } else {
      if (tok.size() < 2) {
        // blif.rs:73 slices tokens[2..], which panics for a bare ".gate".
        delete c;
        return set_err(FX_E_INVALID, "BLIF line " + std::to_string(lineno) +
                                         ": .gate without a gate type");
      }
      for (size_t k = 2; k < tok.size(); ++k) {
        const std::string& t = tok[k];
        const size_t eq = t.find('=');
        if (eq == std::string::npos) continue;
        const std::string net = t.substr(eq + 1);
        if (t.compare(0, 2, "O=") == 0 || t.compare(0, 2, "Y=") == 0)
          drive(net, g);
        else
          read(net, g);
      }
    }
  }
  if (in.bad()) {
    delete c;
    return set_err(FX_E_INVALID, std::string("error reading BLIF: ") + path);
  }
  // Driven nets by first drive, then readers-only nets by first read.
  std::vector<int32_t> seq_drv(drive_seq, -1), seq_read(read_seq, -1);
  for (int32_t id = 0; id < (int32_t)nets.size(); ++id) {
    if (first_drv[id] >= 0) seq_drv[first_drv[id]] = id;
    else if (first_read[id] >= 0) seq_read[first_read[id]] = id;
  }
  auto add_net = [&](int32_t id, bool with_driver) {
    const size_t k = readers[id].size() + (with_driver ? 1 : 0);
    if (k < 2) return;
    c->net_names.push_back(nets[id]);
    if (with_driver) c->net_pins.push_back(driver[id]);
    c->net_pins.insert(c->net_pins.end(), readers[id].begin(), readers[id].end());
    c->net_ptr.push_back((int32_t)c->net_pins.size());
  };
  for (int32_t id : seq_drv)
    if (id >= 0) add_net(id, true);
  for (int32_t id : seq_read)
    if (id >= 0) add_net(id, false);
  if (c->net_pins.size() >= (size_t)std::numeric_limits<int32_t>::max()) {
    delete c;
    return set_err(FX_E_INVALID, "BLIF: too many pins");
  }
  *out = c;
  return FX_OK;
}

extern "C" void fx_circuit_free(fx_circuit* c) { delete c; }

extern "C" int fx_circuit_info(const fx_circuit* c, int32_t* num_gates, int32_t* num_nets,
                               int64_t* num_pins, float* chip_width, float* chip_height) {
  if (!c) return set_err(FX_E_INVALID, "fx_circuit_info: null circuit");
  if (num_gates) *num_gates = (int32_t)c->gate_names.size();
  if (num_nets) *num_nets = (int32_t)c->net_names.size();
  if (num_pins) *num_pins = (int64_t)c->net_pins.size();
  if (chip_width) *chip_width = c->chip_width;
  if (chip_height) *chip_height = c->chip_height;
  return FX_OK;
}

extern "C" int fx_circuit_netlist(const fx_circuit* c, int32_t* net_ptr, int32_t* net_pins,
                                  uint8_t* gate_types) {
  if (!c) return set_err(FX_E_INVALID, "fx_circuit_netlist: null circuit");
  if (net_ptr) std::memcpy(net_ptr, c->net_ptr.data(), 4 * c->net_ptr.size());
  if (net_pins && !c->net_pins.empty())
    std::memcpy(net_pins, c->net_pins.data(), 4 * c->net_pins.size());
  if (gate_types && !c->gate_types.empty())
    std::memcpy(gate_types, c->gate_types.data(), c->gate_types.size());
  return FX_OK;
}

extern "C" const char* fx_circuit_name(const fx_circuit* c) { return c ? c->name.c_str() : nullptr; }

extern "C" const char* fx_circuit_net_name(const fx_circuit* c, int32_t net) {
  if (!c || net < 0 || net >= (int32_t)c->net_names.size()) return nullptr;
  return c->net_names[net].c_str();
}

extern "C" const char* fx_circuit_gate_name(const fx_circuit* c, int32_t gate) {
  if (!c || gate < 0 || gate >= (int32_t)c->gate_names.size()) return nullptr;
  return c->gate_names[gate].c_str();
}

// Rust `f32 as i64`: truncation toward zero, NaN -> 0, saturating at the
// i64 range.
static long long as_i64(float v) {
  if (v != v) return 0;
  if (v >= 9223372036854775807.0f) return std::numeric_limits<long long>::max();
  if (v <= -9223372036854775808.0f) return std::numeric_limits<long long>::min();
  return (long long)v;
}

extern "C" int fx_def_write(const char* path, const char* design, int32_t num_gates,
                            const char* const* gate_names, const float* x, const float* y,
                            float chip_width, float chip_height) {
  fx::clear_error();
  if (!path || !design || num_gates < 0 || (num_gates && (!x || !y)))
    return set_err(FX_E_INVALID, "fx_def_write: bad arguments");
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_err(FX_E_INVALID, std::string("Failed to create DEF: ") + path);
  std::vector<char> buf(1 << 20);
  std::setvbuf(f, buf.data(), _IOFBF, buf.size());
  std::fprintf(f, "VERSION 5.8 ;\n");
  std::fprintf(f, "DIVIDERCHAR \"/\" ;\n");
  std::fprintf(f, "BUSBITCHARS \"[]\" ;\n");
  std::fprintf(f, "DESIGN %s ;\n", design);
  std::fprintf(f, "UNITS DISTANCE MICRONS 1000 ;\n");
  // Products in f32, then the saturating cast (def.rs:22-23, 29-30).
  std::fprintf(f, "\nDIEAREA ( 0 0 ) ( %lld %lld ) ;\n", as_i64(chip_width * 1000.0f),
               as_i64(chip_height * 1000.0f));
  std::fprintf(f, "\nCOMPONENTS %d ;\n", num_gates);
  for (int32_t i = 0; i < num_gates; ++i) {
    const float px = x[i] * 1000.0f, py = y[i] * 1000.0f;
    if (gate_names)
      std::fprintf(f, "    - %s GATE_TYPE + PLACED ( %lld %lld ) N ;\n", gate_names[i], as_i64(px),
                   as_i64(py));
    else
      std::fprintf(f, "    - g%d GATE_TYPE + PLACED ( %lld %lld ) N ;\n", i, as_i64(px), as_i64(py));
  }
  std::fprintf(f, "END COMPONENTS ;\n");
  std::fprintf(f, "\nEND DESIGN\n");
  const bool ok = !std::ferror(f);
  if (std::fclose(f) != 0 || !ok) return set_err(FX_E_INVALID, std::string("error writing DEF: ") + path);
  return FX_OK;
}
