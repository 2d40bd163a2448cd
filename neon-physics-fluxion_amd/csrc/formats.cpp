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
#include <limits>
#include <string>
#include <string_view>
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

// Whitespace-separated tokens of one line, as views into the file buffer.
void split_ws(std::string_view s, std::vector<std::string_view>& out) {
  out.clear();
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && is_space(s[i])) ++i;
    size_t j = i;
    while (j < s.size() && !is_space(s[j])) ++j;
    if (j > i) out.push_back(s.substr(i, j - i));
    i = j;
  }
}

// Net names -> ids in first-seen order: open addressing with linear probing
// over a power-of-two table of ids; names are views into the file buffer.
class NameTable {
 public:
  // The id of `s`, adding it as names.size() when new.
  int32_t find_or_add(std::string_view s, std::vector<std::string_view>& names, bool* added) {
    if (2 * (names.size() + 1) > slot_.size()) grow(names);
    const uint64_t h = hash(s);
    for (size_t i = h & mask_;; i = (i + 1) & mask_) {
      const int32_t id = slot_[i];
      if (id < 0) {
        slot_[i] = (int32_t)names.size();
        hashes_.push_back(h);
        names.push_back(s);
        *added = true;
        return slot_[i];
      }
      if (hashes_[id] == h && names[id] == s) {
        *added = false;
        return id;
      }
    }
  }

 private:
  static uint64_t hash(std::string_view s) {  // FNV-1a, then a final mix
    uint64_t h = 1469598103934665603ull;
    for (unsigned char ch : s) h = (h ^ ch) * 1099511628211ull;
    h ^= h >> 29;
    h *= 0xbf58476d1ce4e5b9ull;
    return h ^ (h >> 32);
  }
  void grow(const std::vector<std::string_view>& names) {
    const size_t cap = slot_.empty() ? 1024 : 2 * slot_.size();
    slot_.assign(cap, -1);
    mask_ = cap - 1;
    for (int32_t id = 0; id < (int32_t)names.size(); ++id) {
      size_t i = hashes_[id] & mask_;
      while (slot_[i] >= 0) i = (i + 1) & mask_;
      slot_[i] = id;
    }
  }
  std::vector<int32_t> slot_;
  std::vector<uint64_t> hashes_;
  size_t mask_ = 0;
};

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
  // The whole file in memory; tokens and net names are views into it.
  std::string text;
  {
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_err(FX_E_INVALID, std::string("Failed to open BLIF: ") + path);
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return set_err(FX_E_INVALID, std::string("error reading BLIF: ") + path);
  }
  auto* c = new fx_circuit();
  c->name = file_stem(path);
  // net name -> index into `nets` (first-seen order); driver per net
  NameTable net_id;
  std::vector<std::string_view> nets;
  std::vector<int32_t> driver;     // -1: not driven
  std::vector<int32_t> first_drv;  // sequence number of the first drive, -1 if never
  std::vector<int32_t> first_read;
  std::vector<int32_t> read_net, read_gate;  // reads in file order
  int32_t drive_seq = 0, read_seq = 0;
  auto id_of = [&](std::string_view s) {
    bool added = false;
    const int32_t id = net_id.find_or_add(s, nets, &added);
    if (added) {
      driver.push_back(-1);
      first_drv.push_back(-1);
      first_read.push_back(-1);
    }
    return id;
  };
  auto drive = [&](std::string_view s, int32_t g) {
    const int32_t id = id_of(s);
    driver[id] = g;
    if (first_drv[id] < 0) first_drv[id] = drive_seq++;
  };
  auto read = [&](std::string_view s, int32_t g) {
    const int32_t id = id_of(s);
    read_net.push_back(id);
    read_gate.push_back(g);
    if (first_read[id] < 0) first_read[id] = read_seq++;
  };
  std::vector<std::string_view> tok;
  int64_t lineno = 0;
  const std::string_view all(text);
  for (size_t pos = 0; pos < all.size();) {
    size_t eol = all.find('\n', pos);
    if (eol == std::string_view::npos) eol = all.size();
    const std::string_view line = all.substr(pos, eol - pos);
    pos = eol + 1;
    ++lineno;
    split_ws(line, tok);
    if (tok.empty() || tok[0][0] == '#') continue;
    const std::string_view d = tok[0];
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
        const std::string_view t = tok[k];
        const size_t eq = t.find('=');
        if (eq == std::string_view::npos) continue;
        const std::string_view net = t.substr(eq + 1);
        if (t.substr(0, 2) == "O=" || t.substr(0, 2) == "Y=")
          drive(net, g);
        else
          read(net, g);
      }
    }
  }
  // Readers of each net in file order: a counting sort of the reads by net.
  const int32_t nn = (int32_t)nets.size();
  std::vector<int64_t> rptr(nn + 1, 0);
  for (int32_t id : read_net) ++rptr[id + 1];
  for (int32_t id = 0; id < nn; ++id) rptr[id + 1] += rptr[id];
  std::vector<int32_t> readers(read_net.size());
  {
    std::vector<int64_t> cur(rptr.begin(), rptr.end() - 1);
    for (size_t k = 0; k < read_net.size(); ++k) readers[cur[read_net[k]]++] = read_gate[k];
  }
  // Driven nets by first drive, then readers-only nets by first read.
  std::vector<int32_t> seq_drv(drive_seq, -1), seq_read(read_seq, -1);
  for (int32_t id = 0; id < nn; ++id) {
    if (first_drv[id] >= 0) seq_drv[first_drv[id]] = id;
    else if (first_read[id] >= 0) seq_read[first_read[id]] = id;
  }
  auto add_net = [&](int32_t id, bool with_driver) {
    const int64_t nr = rptr[id + 1] - rptr[id];
    const int64_t k = nr + (with_driver ? 1 : 0);
    if (k < 2) return;
    c->net_names.emplace_back(nets[id]);
    if (with_driver) c->net_pins.push_back(driver[id]);
    c->net_pins.insert(c->net_pins.end(), readers.begin() + rptr[id], readers.begin() + rptr[id + 1]);
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

extern "C" int fx_circuit_net_names(const fx_circuit* c, char* buf, int64_t buf_len,
                                    int64_t* needed) {
  fx::clear_error();
  if (!c || !needed) return set_err(FX_E_INVALID, "fx_circuit_net_names: bad arguments");
  int64_t total = 0;
  for (const std::string& n : c->net_names) total += (int64_t)n.size() + 1;
  *needed = total;
  if (!buf) return FX_OK;
  if (buf_len < total) return set_err(FX_E_INVALID, "fx_circuit_net_names: buffer too small");
  for (const std::string& n : c->net_names) {
    std::memcpy(buf, n.c_str(), n.size() + 1);
    buf += n.size() + 1;
  }
  return FX_OK;
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
