// gp_mps.cpp -- MPS text to a host model for device ingest (gp_mps_parse,
// gp_mps_read, gp_mps_free).  Host code: reading a text file is CPU work in
// the reference too; the device takes over at gp_mps_upload.
//
// Format rules follow mps_reader::read_mps (cpp/mps_reader/src/io/mpsreader.cpp)
// so that every file reads to the same LP and fails with the same message:
//  * the text is read as free format first and, if that fails, as fixed
//    format; when both fail the free-format message is reported (:692-708);
//  * cards: a line whose first character is not blank is a section header;
//    blank lines and lines whose first non-blank character is '*' or '&' are
//    skipped; fixed-format data fields sit in columns 2-3, 5-12, 15-22, 25-36,
//    40-47 and 50-61 and a fixed-format NAME takes columns 15 onwards (:136-187);
//  * numbers: integer digits accumulate as v = 10 v + d, the k-th fraction
//    digit adds d * 0.1^k with 0.1^k a running product, an exponent scales by
//    pow(10, e); more than 18 mantissa digits, a token without digits or any
//    other character hands the token to the C library (strtod, 'd'/'D' read
//    as 'E'), which must consume all of it (:191-298);
//  * sections (:546-621): OBJSENSE (MAX rejected), ROWS (first letter of the
//    type; the first N row is the objective, other N rows are dropped),
//    COLUMNS ('MARKER' lines toggle integer marking; integrality itself is
//    dropped), RHS (an RHS on the objective sets the objective constant to its
//    negative), RANGES, BOUNDS (FR MI PL BV LO UP FX LI UI); only the first
//    RHS, RANGES and BOUNDS vector is read; QUADOBJ / QMATRIX are rejected;
//  * bounds left unset (:523-542): [0, inf), [0, 1] for a column first seen
//    between integer markers, lower -inf when only a negative upper is given.
//
// Design: the text is held once in memory and every field and name is a view
// into it -- no per-token strings; names are interned in hash tables of
// views; the matrix leaves as entries in file order (column, row, value)
// with duplicates unsummed.  Summing duplicates in file order and building
// A (CSR) and A^T happen on the device (gp_lp_upload_entries).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <string_view>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/gpupre.h"

namespace gp {
void set_last_error(const std::string& msg);
}

namespace {

using View = std::string_view;

constexpr double kInf = std::numeric_limits<double>::infinity();
const double kUnset = std::numeric_limits<double>::quiet_NaN();

// Every malformed file is reported as std::invalid_argument, as in the reference.
struct MpsError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

bool is_blank(char ch) { return std::isspace(static_cast<unsigned char>(ch)) != 0; }

View strip(View s) {
  while (!s.empty() && is_blank(s.front())) s.remove_prefix(1);
  while (!s.empty() && is_blank(s.back())) s.remove_suffix(1);
  return s;
}

// characters [from, to) of a line (0-based), stripped
View columns(View line, size_t from, size_t to) {
  if (line.size() <= from) return {};
  return strip(line.substr(from, std::min(line.size(), to) - from));
}

void split_words(View line, std::vector<View>& out) {
  size_t p = 0;
  while (p < line.size()) {
    while (p < line.size() && is_blank(line[p])) ++p;
    const size_t b = p;
    while (p < line.size() && !is_blank(line[p])) ++p;
    if (p > b) out.push_back(line.substr(b, p - b));
  }
}

// ---- cards --------------------------------------------------------------------
struct Card {
  int line = 0;  // 1-based line number in the text
  bool header = false;
  std::vector<View> field;
};

// Walks the text line by line (std::getline's split: '\n' ends a line, a
// final line without '\n' counts when it is not empty) and yields the cards.
class CardReader {
 public:
  CardReader(View text, bool fixed) : text_(text), fixed_(fixed) {}

  bool next(Card& card) {
    View line;
    while (next_line(line)) {
      if (skipped(line)) continue;
      card.line = line_no_;
      card.header = !is_blank(line.front());
      card.field.clear();
      if (card.header) {
        split_words(line, card.field);
        header_fields(line, card.field);
      } else if (fixed_) {
        static const size_t kFixed[6][2] = {{1, 3}, {4, 12}, {14, 22}, {24, 36}, {39, 47}, {49, 61}};
        for (const auto& w : kFixed) {
          const View f = columns(line, w[0], w[1]);
          if (!f.empty()) card.field.push_back(f);
        }
      } else {
        split_words(line, card.field);
      }
      if (!card.field.empty()) return true;
    }
    return false;
  }

 private:
  bool next_line(View& line) {
    if (pos_ >= text_.size()) return false;
    const size_t nl = text_.find('\n', pos_);
    const size_t end = nl == View::npos ? text_.size() : nl;
    line = text_.substr(pos_, end - pos_);
    pos_ = end + 1;
    ++line_no_;
    return true;
  }

  static bool skipped(View line) {
    for (const char ch : line)
      if (!is_blank(ch)) return ch == '*' || ch == '&';
    return true;
  }

  void header_fields(View line, std::vector<View>& f) const {
    if (f.size() >= 2 && f[0] == "OBJECT" && f[1] == "BOUND") {
      f.assign(1, View("OBJECT BOUND"));
      return;
    }
    if (fixed_ && f[0] == "NAME") {
      const View name = columns(line, 14, line.size());
      if (!name.empty()) f.assign({View("NAME"), name});
    }
  }

  View text_;
  bool fixed_;
  size_t pos_ = 0;
  int line_no_ = 0;
};

// ---- numbers ------------------------------------------------------------------
// The C library's reading, with the std::stod failure modes: no conversion
// ("stod", invalid_argument), out of range ("stod", out_of_range), trailing
// characters (the MPS message).
double library_number(View token) {
  std::string s(token);
  for (char& ch : s)
    if (ch == 'd' || ch == 'D') ch = 'E';
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str()) throw std::invalid_argument("stod");
  if (errno == ERANGE) throw std::out_of_range("stod");
  if (static_cast<size_t>(end - s.c_str()) != s.size())
    throw MpsError("invalid numeric token in MPS: " + std::string(token));
  return v;
}

bool is_digit(char ch) { return ch >= '0' && ch <= '9'; }

double mps_number(View t) {
  const size_t n = t.size();
  size_t i = 0;
  const bool minus = n > 0 && t[0] == '-';
  if (n > 0 && (t[0] == '-' || t[0] == '+')) i = 1;
  double v = 0.0;
  double place = 0.1;  // 0.1^k for the k-th fraction digit
  int digits = 0;
  for (; i < n && is_digit(t[i]); ++i) {
    v = v * 10.0 + static_cast<double>(t[i] - '0');
    if (++digits > 18) return library_number(t);
  }
  if (i < n && t[i] == '.') {
    for (++i; i < n && is_digit(t[i]); ++i) {
      v += static_cast<double>(t[i] - '0') * place;
      place *= 0.1;
      if (++digits > 18) return library_number(t);
    }
  }
  if (digits == 0) return library_number(t);
  if (i < n) {
    if (t[i] != 'e' && t[i] != 'E' && t[i] != 'd' && t[i] != 'D') return library_number(t);
    ++i;
    const bool down = i < n && t[i] == '-';
    if (i < n && (t[i] == '-' || t[i] == '+')) ++i;
    if (i == n) return library_number(t);
    uint32_t e = 0;  // wraps like the reference's int accumulator on x86-64
    for (; i < n; ++i) {
      if (!is_digit(t[i])) return library_number(t);
      e = 10u * e + static_cast<uint32_t>(t[i] - '0');
    }
    const double scale = std::pow(10.0, static_cast<double>(static_cast<int32_t>(e)));
    v = down ? v / scale : v * scale;
  }
  return minus ? -v : v;
}

// ---- the LP as the cards describe it -------------------------------------------
class MpsLp {
 public:
  // false at ENDATA
  bool header(const Card& c) {
    const View h = c.field[0];
    if (h == "NAME") {
      if (c.field.size() >= 2) name_ = c.field[1];
      return true;
    }
    if (h == "ENDATA") return false;
    static const std::pair<View, Part> kParts[] = {{"OBJSENSE", Part::kSense}, {"ROWS", Part::kRows},
                                                   {"COLUMNS", Part::kColumns}, {"RHS", Part::kRhs},
                                                   {"BOUNDS", Part::kBounds},   {"RANGES", Part::kRanges}};
    for (const auto& [label, part] : kParts) {
      if (h == label) {
        part_ = part;
        return true;
      }
    }
    if (h == "QUADOBJ" || h == "QMATRIX")
      throw MpsError("quadratic MPS sections are not supported by the MPS reader");
    throw MpsError("unknown MPS section header: " + std::string(h));
  }

  void data(const Card& c) {
    switch (part_) {
      case Part::kSense:
        if (c.field[0] == "MAX") throw MpsError("MPS OBJSENSE MAX is not supported by the MPS reader");
        return;
      case Part::kRows: return row(c);
      case Part::kColumns: return column_entries(c);
      case Part::kRhs: return rhs(c);
      case Part::kRanges: return ranges(c);
      case Part::kBounds: return bounds(c);
      case Part::kNone: break;
    }
    throw MpsError("MPS data entry outside a supported section");
  }

  void finish() {
    if (row_kind_.empty() && col_of_.empty()) throw MpsError("empty MPS model");
    for (size_t j = 0; j < lb_.size(); ++j) {
      const bool no_lb = std::isnan(lb_[j]), no_ub = std::isnan(ub_[j]);
      if (no_lb && no_ub) {
        lb_[j] = 0.0;
        ub_[j] = marked_col_[j] ? 1.0 : kInf;
      } else if (no_lb) {
        lb_[j] = ub_[j] < 0.0 ? -kInf : 0.0;
      } else if (no_ub) {
        ub_[j] = kInf;
      }
    }
  }

  void emit(gp_mps_model* out) const;

 private:
  enum class Part { kNone, kSense, kRows, kColumns, kRhs, kBounds, kRanges };
  struct RowRef {
    int32_t index;  // -1: an N row
    char kind;
  };

  static std::string at_line(const char* what, const Card& c) {
    return std::string("invalid ") + what + " entry at line " + std::to_string(c.line);
  }

  const RowRef& row_ref(View name, const char* section) const {
    const auto it = row_of_.find(name);
    if (it == row_of_.end()) throw MpsError(std::string("unknown MPS row in ") + section + ": " + std::string(name));
    return it->second;
  }

  // first vector name met in a section is the one read
  static bool chosen(View& vector, View name) {
    if (vector.empty()) vector = name;
    return name == vector;
  }

  void row(const Card& c) {
    if (c.field.size() < 2) throw MpsError(at_line("ROWS", c));
    const char kind = c.field[0][0];
    const View name = c.field[1];
    if (kind == 'N') {
      if (objective_.empty()) objective_ = name;
      row_of_.emplace(name, RowRef{-1, 'N'});
      return;
    }
    if (kind != 'E' && kind != 'G' && kind != 'L') throw MpsError(std::string("unsupported MPS row type: ") + kind);
    row_of_.emplace(name, RowRef{static_cast<int32_t>(row_kind_.size()), kind});
    row_kind_.push_back(kind);
    row_lo_.push_back(kind == 'L' ? -kInf : 0.0);
    row_hi_.push_back(kind == 'G' ? kInf : 0.0);
  }

  int32_t column_index(View name) {
    const auto [it, added] = col_of_.emplace(name, static_cast<int32_t>(cost_.size()));
    if (added) {
      cost_.push_back(0.0);
      lb_.push_back(kUnset);
      ub_.push_back(kUnset);
      marked_col_.push_back(marked_ ? 1 : 0);
    }
    return it->second;
  }

  void column_entries(const Card& c) {
    if (c.field.size() >= 3 && c.field[1] == "'MARKER'") {
      if (c.field[2] == "'INTORG'") marked_ = true;
      else if (c.field[2] == "'INTEND'") marked_ = false;
      return;
    }
    if (c.field.size() < 3) throw MpsError(at_line("COLUMNS", c));
    const int32_t j = column_index(c.field[0]);
    for (size_t k = 1; k + 1 < c.field.size(); k += 2) {
      const double v = mps_number(c.field[k + 1]);
      const RowRef& r = row_ref(c.field[k], "COLUMNS");
      if (r.index >= 0) {
        entry_col_.push_back(j);
        entry_row_.push_back(r.index);
        entry_val_.push_back(v);
      } else if (c.field[k] == objective_) {
        cost_[static_cast<size_t>(j)] += v;
      }
    }
  }

  void rhs(const Card& c) {
    if (c.field.size() < 3) throw MpsError(at_line("RHS", c));
    if (!chosen(rhs_vector_, c.field[0])) return;
    for (size_t k = 1; k + 1 < c.field.size(); k += 2) {
      const RowRef& r = row_ref(c.field[k], "RHS");
      if (r.index < 0) {
        if (c.field[k] == objective_) obj_constant_ = -mps_number(c.field[k + 1]);
        continue;
      }
      const double v = mps_number(c.field[k + 1]);
      const size_t i = static_cast<size_t>(r.index);
      if (r.kind != 'G') row_hi_[i] = v;
      if (r.kind != 'L') row_lo_[i] = v;
    }
  }

  void ranges(const Card& c) {
    if (c.field.size() < 3) throw MpsError(at_line("RANGES", c));
    if (!chosen(range_vector_, c.field[0])) return;
    for (size_t k = 1; k + 1 < c.field.size(); k += 2) {
      const double v = mps_number(c.field[k + 1]);
      const RowRef& r = row_ref(c.field[k], "RANGES");
      if (r.index < 0) continue;
      const size_t i = static_cast<size_t>(r.index);
      switch (r.kind) {
        case 'E':
          (v >= 0.0 ? row_hi_[i] : row_lo_[i]) += v;
          break;
        case 'L': row_lo_[i] = row_hi_[i] - std::fabs(v); break;
        default: row_hi_[i] = row_lo_[i] + std::fabs(v); break;
      }
    }
  }

  void bounds(const Card& c) {
    if (c.field.size() < 3) throw MpsError(at_line("BOUNDS", c));
    if (!chosen(bound_vector_, c.field[1])) return;
    const auto it = col_of_.find(c.field[2]);
    if (it == col_of_.end()) throw MpsError("unknown MPS column in BOUNDS: " + std::string(c.field[2]));
    const size_t j = static_cast<size_t>(it->second);
    const View type = c.field[0];
    // bound types without a value
    if (type == "FR" || type == "MI" || type == "PL" || type == "BV") {
      if (type != "PL") lb_[j] = type == "BV" ? 0.0 : -kInf;
      if (type != "MI") ub_[j] = type == "BV" ? 1.0 : kInf;
      return;
    }
    if (c.field.size() < 4) throw MpsError("missing value for MPS bound type: " + std::string(type));
    const double v = mps_number(c.field[3]);
    const bool lower = type == "LO" || type == "LI" || type == "FX";
    const bool upper = type == "UP" || type == "UI" || type == "FX";
    if (!lower && !upper) throw MpsError("unsupported MPS bound type: " + std::string(type));
    if (lower) lb_[j] = v;
    if (upper) ub_[j] = v;
  }

  Part part_ = Part::kNone;
  View name_, objective_, rhs_vector_, range_vector_, bound_vector_;
  bool marked_ = false;
  std::unordered_map<View, RowRef> row_of_;
  std::unordered_map<View, int32_t> col_of_;
  std::vector<char> row_kind_;
  std::vector<double> row_lo_, row_hi_;
  std::vector<double> cost_, lb_, ub_;
  std::vector<uint8_t> marked_col_;
  std::vector<int32_t> entry_col_, entry_row_;
  std::vector<double> entry_val_;
  double obj_constant_ = 0.0;
};

template <class T>
T* host_copy(const std::vector<T>& v) {
  T* p = static_cast<T*>(std::malloc(sizeof(T) * (v.empty() ? 1 : v.size())));
  if (!p) throw std::bad_alloc();
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

void MpsLp::emit(gp_mps_model* out) const {
  if (entry_val_.size() > static_cast<size_t>(INT32_MAX))
    throw std::length_error("MPS matrix entries exceed the int32 offset range");
  gp_mps_model m{};
  try {
    m.rows = static_cast<int32_t>(row_kind_.size());
    m.cols = static_cast<int32_t>(cost_.size());
    m.entries = static_cast<int64_t>(entry_val_.size());
    m.entry_col = host_copy(entry_col_);
    m.entry_row = host_copy(entry_row_);
    m.entry_val = host_copy(entry_val_);
    m.c = host_copy(cost_);
    m.AL = host_copy(row_lo_);
    m.AU = host_copy(row_hi_);
    m.l = host_copy(lb_);
    m.u = host_copy(ub_);
    m.obj_constant = obj_constant_;
    m.name = static_cast<char*>(std::malloc(name_.size() + 1));
    if (!m.name) throw std::bad_alloc();
    std::memcpy(m.name, name_.data(), name_.size());
    m.name[name_.size()] = '\0';
  } catch (...) {
    gp_mps_free(&m);
    throw;
  }
  *out = m;
}

std::unique_ptr<MpsLp> read_as(View text, bool fixed) {
  auto lp = std::make_unique<MpsLp>();
  CardReader reader(text, fixed);
  Card card;
  while (reader.next(card)) {
    if (!card.header) lp->data(card);
    else if (!lp->header(card)) break;
  }
  lp->finish();
  return lp;
}

template <class F>
int host_guarded(F&& body) {
  try {
    body();
    return GP_OK;
  } catch (const std::invalid_argument& e) {
    gp::set_last_error(e.what());
    return GP_ERR_INVALID_ARGUMENT;
  } catch (const std::length_error& e) {
    gp::set_last_error(e.what());
    return GP_ERR_LENGTH;
  } catch (const std::bad_alloc&) {
    gp::set_last_error("host allocation failed");
    return GP_ERR_ALLOCATION;
  } catch (const std::exception& e) {
    gp::set_last_error(e.what());
    return GP_ERR_RUNTIME;
  }
}

}  // namespace

extern "C" int gp_mps_parse(const char* text, size_t length, gp_mps_model* out) {
  return host_guarded([&] {
    if (!out || (!text && length)) throw std::invalid_argument("gp_mps_parse: null argument");
    std::memset(out, 0, sizeof(*out));
    const View all(text ? text : "", length);
    std::unique_ptr<MpsLp> lp;
    try {
      lp = read_as(all, false);
    } catch (const std::exception& free_format) {
      const std::string message = free_format.what();
      try {
        lp = read_as(all, true);
      } catch (const std::exception&) {
        throw MpsError(message);
      }
    }
    lp->emit(out);
  });
}

extern "C" int gp_mps_read(const char* path, gp_mps_model* out) {
  std::string text;
  const int rc = host_guarded([&] {
    if (!path || !out) throw std::invalid_argument("gp_mps_read: null argument");
    std::memset(out, 0, sizeof(*out));
    std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "rb"), &std::fclose);
    if (!f) throw std::runtime_error(std::string("failed to open ") + path);
    if (std::fseek(f.get(), 0, SEEK_END) == 0) {
      const long size = std::ftell(f.get());
      if (size > 0) text.reserve(static_cast<size_t>(size));
      std::rewind(f.get());
    }
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f.get())) > 0) text.append(buf, got);
    if (std::ferror(f.get())) throw std::runtime_error(std::string("failed to read ") + path);
  });
  if (rc != GP_OK) return rc;
  return gp_mps_parse(text.data(), text.size(), out);
}

extern "C" void gp_mps_free(gp_mps_model* model) {
  if (!model) return;
  std::free(model->entry_col);
  std::free(model->entry_row);
  std::free(model->entry_val);
  std::free(model->c);
  std::free(model->AL);
  std::free(model->AU);
  std::free(model->l);
  std::free(model->u);
  std::free(model->name);
  std::memset(model, 0, sizeof(*model));
}
