#include "mps.hpp"

#include <algorithm>
#include <atomic>
#include <climits>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <numeric>
#include <sstream>
#include <thread>
#include <type_traits>
#include <unordered_set>

#include <sys/mman.h>

namespace smy::host {
namespace {

constexpr double kInfiniteBound = 1e30;  // |value| >= this reads as infinite (mps_reader.cpp)

enum class Sec : std::uint8_t { kNone, kObjSense, kObjName, kRows, kColumns, kRhs, kRanges, kBounds, kEnd };
enum class Kind : std::uint8_t { kE, kL, kG };

bool blank(char c) { return c == ' ' || c == '\t'; }

bool iequals(std::string_view a, std::string_view b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i) {
    const auto low = [](char c) { return (c >= 'A' && c <= 'Z') ? static_cast<char>(c + 32) : c; };
    if (low(a[i]) != low(b[i])) return false;
  }
  return true;
}

struct Tokens {
  static constexpr int kMax = 8;
  std::string_view t[kMax];
  int n = 0;
  std::string_view operator[](int k) const { return t[k]; }
};

// Free MPS: whitespace-separated fields (more than kMax fields are counted, not kept).
void tokenize_free(std::string_view line, Tokens& tk) {
  tk.n = 0;
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && blank(line[i])) ++i;
    const size_t start = i;
    while (i < line.size() && !blank(line[i])) ++i;
    if (i > start) {
      if (tk.n < Tokens::kMax) tk.t[tk.n] = line.substr(start, i - start);
      ++tk.n;
    }
  }
}

// Fixed MPS: the type field (columns 2-3) and five data fields, trimmed; empty fields are skipped.
void tokenize_fixed(std::string_view line, Tokens& tk) {
  static constexpr size_t kStart[] = {1, 4, 14, 24, 39, 49};
  static constexpr size_t kEnd[] = {3, 12, 22, 36, 47, 61};
  tk.n = 0;
  for (size_t f = 0; f < 6; ++f) {
    if (kStart[f] >= line.size()) break;
    std::string_view field = line.substr(kStart[f], std::min(kEnd[f], line.size()) - kStart[f]);
    while (!field.empty() && blank(field.front())) field.remove_prefix(1);
    while (!field.empty() && blank(field.back())) field.remove_suffix(1);
    if (!field.empty()) tk.t[tk.n++] = field;
  }
}

double read_number(std::string_view s, long long line) {
  std::string_view body = s;
  if (!body.empty() && body.front() == '+') body.remove_prefix(1);
  double v = 0.0;
  const auto [ptr, ec] = std::from_chars(body.data(), body.data() + body.size(), v);
  if (ec != std::errc() || ptr != body.data() + body.size())
    throw MpsError("invalid number '" + std::string(s) + "'", line);
  if (v >= kInfiniteBound) return kInf;
  if (v <= -kInfiniteBound) return -kInf;
  return v;
}

// An uninitialized array of a trivial type, 2 MiB aligned and marked for transparent huge pages:
// the name tables and the line index are read at random, and huge pages spare those reads most of
// their TLB misses.
template <class T>
class HugeArray {
 public:
  HugeArray() = default;
  explicit HugeArray(size_t count) { allocate(count); }
  HugeArray(const HugeArray&) = delete;
  HugeArray& operator=(const HugeArray&) = delete;
  HugeArray(HugeArray&& o) noexcept : data_(o.data_), size_(o.size_) { o.data_ = nullptr, o.size_ = 0; }
  HugeArray& operator=(HugeArray&& o) noexcept {
    std::swap(data_, o.data_);
    std::swap(size_, o.size_);
    return *this;
  }
  ~HugeArray() { std::free(data_); }

  void allocate(size_t count) {
    static_assert(std::is_trivially_copyable_v<T>);
    std::free(data_);
    constexpr size_t kAlign = size_t{2} << 20;
    const size_t bytes = std::max<size_t>(kAlign, (count * sizeof(T) + kAlign - 1) / kAlign * kAlign);
    data_ = static_cast<T*>(std::aligned_alloc(kAlign, bytes));
    if (data_ == nullptr) throw std::bad_alloc();
    madvise(data_, bytes, MADV_HUGEPAGE);
    size_ = count;
  }
  T* data() { return data_; }
  const T* data() const { return data_; }
  size_t size() const { return size_; }
  T& operator[](size_t i) { return data_[i]; }
  const T& operator[](size_t i) const { return data_[i]; }

 private:
  T* data_ = nullptr;
  size_t size_ = 0;
};

// Names to integers: open addressing with linear probing over a power-of-two array of slots
// that keep the name's 64-bit hash and its first 8 bytes, so that a lookup of a short name reads
// one slot (one cache line) and no text. Names are views into the parsed text.
class NameTable {
 public:
  static constexpr int kAbsent = INT32_MIN;

  void reserve(size_t count) {
    size_t cap = 16;
    while (cap < 2 * count) cap <<= 1;
    if (cap > slots_.size()) rehash(cap);
  }
  // Loads the slot a key hashes to into the cache ahead of its lookup.
  void prefetch(std::string_view key) const {
    if (!slots_.size()) return;
    __builtin_prefetch(&slots_[hash(key) & mask_]);
  }
  size_t size() const { return size_; }
  bool empty() const { return size_ == 0; }

  int find(std::string_view key) const {
    if (slots_.size() == 0) return kAbsent;
    const uint64_t h = hash(key), pre = prefix(key);
    for (size_t i = h & mask_;; i = (i + 1) & mask_) {
      const Slot& s = slots_[i];
      if (s.len == kEmpty) return kAbsent;
      if (s.hash == h && s.prefix == pre && same(s, key)) return s.value;
    }
  }

  // Sets the value of a key that is in the table.
  void assign(std::string_view key, int value) {
    const uint64_t h = hash(key), pre = prefix(key);
    for (size_t i = h & mask_;; i = (i + 1) & mask_) {
      Slot& s = slots_[i];
      if (s.hash == h && s.prefix == pre && same(s, key)) {
        s.value = value;
        return;
      }
    }
  }

  // The value stored for key, or `value` after storing it; inserted says which.
  int insert(std::string_view key, int value, bool& inserted) {
    if (2 * (size_ + 1) > slots_.size()) rehash(std::max<size_t>(16, 2 * slots_.size()));
    const uint64_t h = hash(key), pre = prefix(key);
    for (size_t i = h & mask_;; i = (i + 1) & mask_) {
      Slot& s = slots_[i];
      if (s.len == kEmpty) {
        s = Slot{h, pre, key.data(), static_cast<uint32_t>(key.size()), value};
        ++size_;
        inserted = true;
        return value;
      }
      if (s.hash == h && s.prefix == pre && same(s, key)) {
        inserted = false;
        return s.value;
      }
    }
  }

 private:
  static constexpr uint32_t kEmpty = UINT32_MAX;
  struct Slot {
    uint64_t hash;
    uint64_t prefix;
    const char* key;
    uint32_t len;
    int32_t value;
  };

  static uint64_t prefix(std::string_view k) {
    uint64_t v = 0;
    std::memcpy(&v, k.data(), std::min<size_t>(8, k.size()));
    return v;
  }
  static bool same(const Slot& s, std::string_view k) {
    return s.len == k.size() && (k.size() <= 8 || std::memcmp(s.key + 8, k.data() + 8, k.size() - 8) == 0);
  }
  // 8-byte words combined by multiply-xorshift steps, the length folded in.
  static uint64_t hash(std::string_view k) {
    uint64_t h = 0x9e3779b97f4a7c15ull ^ k.size();
    size_t i = 0;
    for (; i + 8 <= k.size(); i += 8) {
      uint64_t w;
      std::memcpy(&w, k.data() + i, 8);
      h = (h ^ w) * 0xff51afd7ed558ccdull;
      h ^= h >> 32;
    }
    if (i < k.size()) {
      uint64_t w = 0;
      std::memcpy(&w, k.data() + i, k.size() - i);
      h = (h ^ w) * 0xc4ceb9fe1a85ec53ull;
    }
    h ^= h >> 29;
    h *= 0xbf58476d1ce4e5b9ull;
    return h ^ (h >> 32);
  }

  void rehash(size_t cap) {
    HugeArray<Slot> old(cap);
    for (size_t i = 0; i < cap; ++i) old[i] = Slot{0, 0, nullptr, kEmpty, 0};
    std::swap(old, slots_);
    mask_ = cap - 1;
    for (size_t k = 0; k < old.size(); ++k) {
      const Slot& s = old[k];
      if (s.len == kEmpty) continue;
      size_t i = s.hash & mask_;
      while (slots_[i].len != kEmpty) i = (i + 1) & mask_;
      slots_[i] = s;
    }
  }

  HugeArray<Slot> slots_;
  size_t mask_ = 0;
  size_t size_ = 0;
};

struct Entry {
  int row;  // -1: the objective
  double value;
};

// A run of COLUMNS lines naming one column: its matrix entries are entries[first, first + count)
// of its part and its objective entries objective[obj_first, obj_first + obj_count); integer: the
// marker state at the run (-1: the part's starting state).
struct Run {
  std::string_view name;
  size_t first, count;
  size_t obj_first, obj_count;
  int integer;
};

struct Part {
  std::vector<Run> runs;
  std::vector<Entry> entries;
  std::vector<double> objective;
  int last_marker = -1;  // the state the part ends in (-1: unchanged)
  bool failed = false;
  long long error_line = 0;
  std::string error;
};

// fn(q) for q in [0, count) on count threads (the calling thread runs the last).
template <class Fn>
void run_parallel(int count, const Fn& fn) {
  std::vector<std::thread> pool;
  for (int q = 0; q + 1 < count; ++q) pool.emplace_back([&fn, q] { fn(q); });
  if (count > 0) fn(count - 1);
  for (std::thread& t : pool) t.join();
}

class Parser {
 public:
  Parser(std::string_view text, bool fixed, int threads) : text_(text), fixed_(fixed), threads_(threads) {}

  MpsModel run() {
    split_lines();
    size_t k = 0;
    while (k < line_count_ && sec_ != Sec::kEnd) {
      const std::string_view line = lines_[k];
      line_no_ = static_cast<long long>(k) + 1;
      if (line.empty() || line.front() == '*') {
        ++k;
        continue;
      }
      const bool header = !blank(line.front());
      tokenize(line, header);
      if (tk_.n == 0) {
        ++k;
        continue;
      }
      if (header) {
        parse_header();
        ++k;
        if (sec_ == Sec::kColumns)
          k = parse_columns(k);
        else if (sec_ == Sec::kRhs || sec_ == Sec::kRanges || sec_ == Sec::kBounds)
          k = parse_records(k);
      } else {
        parse_data();
        ++k;
      }
    }
    // The last line read: ENDATA's, or else every line (one more after a final '\n').
    if (sec_ != Sec::kEnd)
      line_no_ = static_cast<long long>(line_count_) + (!text_.empty() && text_.back() == '\n' ? 1 : 0);
    return finish();
  }

 private:
  [[noreturn]] void fail(const std::string& what) const { throw MpsError(what, line_no_); }

  // The text's lines (a final '\r' removed), found on several threads for a large text.
  void split_lines() {
    const size_t size = text_.size();
    const int parts = size < (size_t{8} << 20) ? 1 : threads_;
    // Part q covers the lines that start in [cut[q], cut[q + 1]).
    std::vector<size_t> cut(static_cast<size_t>(parts) + 1, size);
    cut[0] = 0;
    for (int q = 1; q < parts; ++q) {
      size_t c = size * static_cast<size_t>(q) / static_cast<size_t>(parts);
      c = std::max(c, cut[q - 1]);
      const void* hit = c < size ? std::memchr(text_.data() + c, '\n', size - c) : nullptr;
      cut[q] = hit ? static_cast<size_t>(static_cast<const char*>(hit) - text_.data()) + 1 : size;
    }
    const auto lines_in = [this](size_t b, size_t e, std::string_view* out) {
      size_t count = 0;
      size_t pos = b;
      while (pos < e) {
        const void* hit = std::memchr(text_.data() + pos, '\n', e - pos);
        const size_t end = hit ? static_cast<size_t>(static_cast<const char*>(hit) - text_.data()) : e;
        if (out != nullptr) {
          std::string_view line = text_.substr(pos, end - pos);
          if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
          out[count] = line;
        }
        ++count;
        pos = end + 1;
      }
      return count;
    };
    std::vector<size_t> first(static_cast<size_t>(parts) + 1, 0);
    run_parallel(parts, [&](int q) { first[q + 1] = lines_in(cut[q], cut[q + 1], nullptr); });
    for (int q = 0; q < parts; ++q) first[q + 1] += first[q];
    // A text that ends without '\n' after its last line, or is empty, still has that line.
    // An empty text is one empty line.
    line_count_ = std::max<size_t>(1, first[parts]);
    lines_.allocate(line_count_);
    lines_[0] = std::string_view();
    run_parallel(parts, [&](int q) { lines_in(cut[q], cut[q + 1], lines_.data() + first[q]); });
  }

  void tokenize(std::string_view line, bool header) {
    if (fixed_ && !header)
      tokenize_fixed(line, tk_);
    else
      tokenize_free(line, tk_);
  }

  void parse_header() {
    const std::string_view key = tk_[0];
    if (key == "NAME") {
      if (tk_.n > 1) out_.name = std::string(tk_[1]);
      sec_ = Sec::kNone;
    } else if (key == "OBJSENSE") {
      sec_ = Sec::kObjSense;
      if (tk_.n > 1) set_sense(tk_[1]);
    } else if (key == "OBJNAME") {
      sec_ = Sec::kObjName;
      if (tk_.n > 1) set_objective_name(tk_[1]);
    } else if (key == "ROWS") {
      sec_ = Sec::kRows;
    } else if (key == "COLUMNS") {
      sec_ = Sec::kColumns;
    } else if (key == "RHS") {
      sec_ = Sec::kRhs;
    } else if (key == "RANGES") {
      sec_ = Sec::kRanges;
    } else if (key == "BOUNDS") {
      sec_ = Sec::kBounds;
    } else if (key == "ENDATA") {
      sec_ = Sec::kEnd;
    } else if (key == "QUADOBJ" || key == "QMATRIX" || key == "QSECTION" || key == "QCMATRIX") {
      fail("section " + std::string(key) + " (quadratic terms) is not supported by the LP path");
    } else if (key == "SOS" || key == "INDICATORS" || key == "CSECTION") {
      fail("section " + std::string(key) + " is not supported");
    } else {
      fail("unknown section '" + std::string(key) + "'");
    }
  }

  void parse_data() {
    switch (sec_) {
      case Sec::kObjSense: set_sense(tk_[0]); break;
      case Sec::kObjName: set_objective_name(tk_[0]); break;
      case Sec::kRows: parse_row(); break;
      case Sec::kRhs:     // handled by parse_records
      case Sec::kRanges:
      case Sec::kBounds:
      case Sec::kColumns:  // handled by parse_columns
      case Sec::kNone:
      case Sec::kEnd: fail("data line outside of a section");
    }
  }

  void set_sense(std::string_view s) {
    if (iequals(s, "MAX") || iequals(s, "MAXIMIZE"))
      out_.model.maximize = true;
    else if (iequals(s, "MIN") || iequals(s, "MINIMIZE"))
      out_.model.maximize = false;
    else
      fail("unknown objective sense '" + std::string(s) + "'");
  }

  void set_objective_name(std::string_view name) {
    if (!rows_.empty()) fail("OBJNAME must precede ROWS");
    objective_ = name;
  }

  void parse_row() {
    if (tk_.n < 2) fail("ROWS entry needs a type and a name");
    const std::string_view type = tk_[0], name = tk_[1];
    bool inserted = false;
    rows_.insert(name, kPending, inserted);
    if (!inserted) fail("duplicate row '" + std::string(name) + "'");
    if (iequals(type, "N")) {
      if (objective_.empty()) objective_ = name;
      set_row(name, name != objective_ ? kFreeRow : kObjectiveRow);
      return;
    }
    Kind kind;
    if (iequals(type, "E"))
      kind = Kind::kE;
    else if (iequals(type, "L"))
      kind = Kind::kL;
    else if (iequals(type, "G"))
      kind = Kind::kG;
    else
      fail("unknown row type '" + std::string(type) + "'");
    set_row(name, static_cast<int>(kinds_.size()));
    kinds_.push_back(kind);
    out_.row_names.emplace_back(name);
  }

  // Codes of the row table besides constraint row indices.
  static constexpr int kPending = -4, kObjectiveRow = -3, kFreeRow = -2;

  void set_row(std::string_view name, int code) { rows_.assign(name, code); }

  // -1 for the objective, -2 for dropped free rows; throws for unknown rows.
  int find_row(std::string_view name, long long line) const {
    if (name == objective_) return -1;
    const int code = rows_.find(name);
    if (code >= 0 || code == kFreeRow) return code;
    throw MpsError("unknown row '" + std::string(name) + "'", line);
  }

  // Prefetches the row-table slot of a free-MPS COLUMNS line's first row name.
  void prefetch_row(std::string_view line) const {
    size_t i = 0;
    for (int field = 0; field < 2; ++field) {
      while (i < line.size() && blank(line[i])) ++i;
      const size_t start = i;
      while (i < line.size() && !blank(line[i])) ++i;
      if (field == 1 && i > start) rows_.prefetch(line.substr(start, i - start));
    }
  }

  int find_col(std::string_view name) const {
    const int j = cols_index_.find(name);
    if (j == NameTable::kAbsent) fail("unknown column '" + std::string(name) + "'");
    return j;
  }

  // The COLUMNS section from line index k to the next header, in parts on several threads.
  size_t parse_columns(size_t k) {
    size_t end = k;
    while (end < line_count_) {
      const std::string_view line = lines_[end];
      if (!line.empty() && line.front() != '*' && !blank(line.front())) break;
      ++end;
    }
    const size_t count = end - k;
    const int parts = count < 65536 ? 1 : std::min<int>(threads_, static_cast<int>(count / 16384) + 1);
    const size_t base = parts_.size();
    parts_.resize(base + static_cast<size_t>(parts));
    run_parallel(parts, [&](int q) {
      const size_t b = k + count * q / parts, e = k + count * (q + 1) / parts;
      parse_column_lines(parts_[base + q], b, e);
    });
    // The earliest error, as a sequential reading reports it.
    for (size_t q = base; q < parts_.size(); ++q)
      if (parts_[q].failed) throw MpsError(parts_[q].error, parts_[q].error_line);
    number_columns(base);
    return end;
  }

  void parse_column_lines(Part& out, size_t b, size_t e) const {
    Tokens tk;
    int state = -1;  // marker state inside this part (-1: the part's starting state)
    out.entries.reserve(e - b + 16);
    constexpr size_t kAhead = 8;  // lines between a row-slot prefetch and its lookup
    try {
      for (size_t k = b; k < e; ++k) {
        if (k + kAhead < e && !fixed_) prefetch_row(lines_[k + kAhead]);
        const std::string_view line = lines_[k];
        const long long no = static_cast<long long>(k) + 1;
        if (line.empty() || line.front() == '*') continue;
        if (fixed_)
          tokenize_fixed(line, tk);
        else
          tokenize_free(line, tk);
        if (tk.n == 0) continue;
        if (tk.n >= 3 && tk[1] == "'MARKER'") {
          if (tk[2] == "'INTORG'")
            state = 1;
          else if (tk[2] == "'INTEND'")
            state = 0;
          else
            throw MpsError("unknown marker " + std::string(tk[2]), no);
          out.last_marker = state;
          continue;
        }
        if (tk.n != 3 && tk.n != 5) throw MpsError("COLUMNS entry needs 3 or 5 fields", no);
        if (out.runs.empty() || out.runs.back().name != tk[0] || out.runs.back().integer != state)
          out.runs.push_back(Run{tk[0], out.entries.size(), 0, out.objective.size(), 0, state});
        Run& run = out.runs.back();
        for (int f = 1; f + 1 < tk.n; f += 2) {
          const int row = find_row(tk[f], no);
          const double value = read_number(tk[f + 1], no);
          if (row >= 0) {
            out.entries.push_back(Entry{row, value});
            ++run.count;
          } else if (row == -1) {
            out.objective.push_back(value);
            ++run.obj_count;
          }
        }
      }
    } catch (const MpsError& err) {
      out.failed = true;
      out.error_line = err.line();
      const std::string what = err.what();
      const std::string prefix = "line " + std::to_string(err.line()) + ": ";
      out.error = what.compare(0, prefix.size(), prefix) == 0 ? what.substr(prefix.size()) : what;
    }
  }

  // Columns numbered in the order they first appear, their integrality from the markers, their
  // objective entries summed in file order; the runs are kept for the matrix (finish).
  void number_columns(size_t base) {
    Model& md = out_.model;
    int state = integer_state_;
    size_t total = 0;
    for (size_t q = base; q < parts_.size(); ++q) total += parts_[q].runs.size();
    refs_.reserve(refs_.size() + total);
    cols_index_.reserve(cols_index_.size() + total);
    for (size_t q = base; q < parts_.size(); ++q) {
      const Part& p = parts_[q];
      for (size_t r = 0; r < p.runs.size(); ++r) {
        const Run& run = p.runs[r];
        if (r + 16 < p.runs.size()) cols_index_.prefetch(p.runs[r + 16].name);
        bool inserted = false;
        const int col = cols_index_.insert(run.name, static_cast<int>(out_.col_names.size()), inserted);
        if (inserted) {
          const int integer = run.integer >= 0 ? run.integer : state;
          out_.col_names.emplace_back(run.name);
          md.obj.push_back(0.0);
          md.col_lower.push_back(0.0);
          md.col_upper.push_back(kInf);
          out_.col_integer.push_back(static_cast<char>(integer == 1));
          lower_set_.push_back(0);
        }
        for (size_t t = 0; t < run.obj_count; ++t) md.obj[col] += p.objective[run.obj_first + t];
        refs_.push_back(Ref{col, static_cast<uint32_t>(q), static_cast<uint32_t>(r)});
      }
      if (p.last_marker >= 0) state = p.last_marker;
    }
    integer_state_ = state;
  }

  // The CSC matrix from the runs: entries of each column in file order, then each column sorted by
  // row (stably), duplicates summed in that order and zero sums dropped.
  void build_matrix() {
    Model& md = out_.model;
    const size_t n = out_.col_names.size();
    if (build_matrix_in_order()) return;
    std::vector<int64_t> start(n + 1, 0);
    for (const Ref& f : refs_) start[static_cast<size_t>(f.col) + 1] += static_cast<int64_t>(run_of(f).count);
    std::partial_sum(start.begin(), start.end(), start.begin());
    const size_t nnz = static_cast<size_t>(start[n]);
    // Where each run's entries go: sequential cursors, then the copies on several threads.
    std::vector<int64_t> at(refs_.size());
    {
      std::vector<int64_t> cursor(start.begin(), start.end() - 1);
      for (size_t r = 0; r < refs_.size(); ++r) {
        at[r] = cursor[refs_[r].col];
        cursor[refs_[r].col] += static_cast<int64_t>(run_of(refs_[r]).count);
      }
    }
    std::vector<Entry> all(nnz);
    const int parts = nnz < (size_t{1} << 18) ? 1 : threads_;
    run_parallel(parts, [&](int q) {
      const size_t b = refs_.size() * q / parts, e = refs_.size() * (q + 1) / parts;
      for (size_t r = b; r < e; ++r) {
        const Run& run = run_of(refs_[r]);
        const Entry* src = parts_[refs_[r].part].entries.data() + run.first;
        std::copy(src, src + run.count, all.begin() + at[r]);
      }
    });
    // Sort, sum and drop within each column; kept[j] entries stay at the column's front.
    std::vector<int64_t> kept(n + 1, 0);
    run_parallel(n < 4096 ? 1 : threads_, [&](int q) {
      const int pc = n < 4096 ? 1 : threads_;
      const size_t b = n * q / pc, e = n * (q + 1) / pc;
      for (size_t j = b; j < e; ++j) {
        Entry* c = all.data() + start[j];
        const size_t len = static_cast<size_t>(start[j + 1] - start[j]);
        bool sorted = true;
        for (size_t t = 1; t < len && sorted; ++t) sorted = c[t - 1].row < c[t].row;
        if (!sorted)
          std::stable_sort(c, c + len, [](const Entry& x, const Entry& y) { return x.row < y.row; });
        size_t w = 0;
        for (size_t r = 0; r < len;) {
          const int row = c[r].row;
          double sum = 0.0;
          for (; r < len && c[r].row == row; ++r) sum += c[r].value;
          if (sum != 0.0) c[w++] = Entry{row, sum};
        }
        kept[j + 1] = static_cast<int64_t>(w);
      }
    });
    std::partial_sum(kept.begin(), kept.end(), kept.begin());
    md.A.rows = md.m;
    md.A.cols = md.n;
    md.A.index.resize(static_cast<size_t>(kept[n]));
    md.A.value.resize(static_cast<size_t>(kept[n]));
    run_parallel(n < 4096 ? 1 : threads_, [&](int q) {
      const int pc = n < 4096 ? 1 : threads_;
      const size_t b = n * q / pc, e = n * (q + 1) / pc;
      for (size_t j = b; j < e; ++j) {
        const Entry* c = all.data() + start[j];
        for (int64_t t = 0; t < kept[j + 1] - kept[j]; ++t) {
          md.A.index[static_cast<size_t>(kept[j] + t)] = c[t].row;
          md.A.value[static_cast<size_t>(kept[j] + t)] = c[t].value;
        }
      }
    });
    md.A.start = std::move(kept);
  }

  // The usual file: the runs come in column order (a column's lines together; a column cut by
  // the split into parts has two runs), its rows ascend strictly and no value is zero, so the runs
  // are the CSC columns as they stand: copied while that is checked. Returns false when it does
  // not hold (the general build then starts over).
  bool build_matrix_in_order() {
    Model& md = out_.model;
    const size_t n = out_.col_names.size();
    std::vector<int64_t> start(n + 1, 0);
    std::vector<size_t> first_ref(n + 1, refs_.size());
    int prev_col = -1;
    for (size_t r = 0; r < refs_.size(); ++r) {
      const int col = refs_[r].col;
      if (col != prev_col && col != prev_col + 1) return false;
      if (col != prev_col) first_ref[col] = r;
      prev_col = col;
      start[static_cast<size_t>(col) + 1] += static_cast<int64_t>(run_of(refs_[r]).count);
    }
    if (prev_col + 1 != static_cast<int>(n)) return false;
    std::partial_sum(start.begin(), start.end(), start.begin());
    md.A.rows = md.m;
    md.A.cols = md.n;
    md.A.index.resize(static_cast<size_t>(start[n]));
    md.A.value.resize(static_cast<size_t>(start[n]));
    const int pc = n < 4096 ? 1 : threads_;
    std::vector<char> clean(static_cast<size_t>(pc), 1);
    run_parallel(pc, [&](int q) {
      bool ok = true;
      for (size_t j = n * q / pc; j < n * (q + 1) / pc && ok; ++j) {
        int* index = md.A.index.data() + start[j];
        double* value = md.A.value.data() + start[j];
        int prev = -1;
        for (size_t r = first_ref[j]; r < first_ref[j + 1]; ++r) {
          const Run& run = run_of(refs_[r]);
          const Entry* c = parts_[refs_[r].part].entries.data() + run.first;
          for (size_t t = 0; t < run.count; ++t) {
            ok = ok && c[t].value != 0.0 && c[t].row > prev;
            prev = c[t].row;
            *index++ = c[t].row;
            *value++ = c[t].value;
          }
        }
      }
      clean[q] = ok;
    });
    for (char c : clean)
      if (!c) return false;
    md.A.start = std::move(start);
    return true;
  }

  struct Ref {
    int col;
    uint32_t part, run;
  };
  const Run& run_of(const Ref& f) const { return parts_[f.part].runs[f.run]; }

  // RHS / RANGES pairs (row -1: the objective, -2: a dropped free row) and BOUNDS entries, parsed
  // from the lines of one part of a section.
  struct RowValue {
    int row;
    double value;
  };
  struct BoundRec {
    int col;
    int rule;
    double value;
  };
  struct Records {
    std::vector<RowValue> pairs;
    std::vector<BoundRec> bounds;
    bool failed = false;
    long long error_line = 0;
    std::string error;
  };

  // A RHS, RANGES or BOUNDS section from line index k to the next header: its lines parsed into
  // records on several threads, then applied in file order.
  size_t parse_records(size_t k) {
    size_t end = k;
    while (end < line_count_) {
      const std::string_view line = lines_[end];
      if (!line.empty() && line.front() != '*' && !blank(line.front())) break;
      ++end;
    }
    const size_t count = end - k;
    const int parts = count < 65536 ? 1 : std::min<int>(threads_, static_cast<int>(count / 16384) + 1);
    std::vector<Records> rec(static_cast<size_t>(parts));
    const Sec sec = sec_;
    run_parallel(parts, [&](int q) {
      const size_t b = k + count * q / parts, e = k + count * (q + 1) / parts;
      Records& out = rec[q];
      Tokens tk;
      try {
        for (size_t i = b; i < e; ++i) {
          const std::string_view line = lines_[i];
          if (line.empty() || line.front() == '*') continue;
          if (fixed_)
            tokenize_fixed(line, tk);
          else
            tokenize_free(line, tk);
          if (tk.n == 0) continue;
          const long long no = static_cast<long long>(i) + 1;
          if (sec == Sec::kBounds)
            scan_bound(tk, no, out.bounds);
          else
            scan_pairs(tk, no, sec == Sec::kRanges, out.pairs);
        }
      } catch (const MpsError& err) {
        out.failed = true;
        out.error_line = err.line();
        const std::string what = err.what();
        const std::string prefix = "line " + std::to_string(err.line()) + ": ";
        out.error = what.compare(0, prefix.size(), prefix) == 0 ? what.substr(prefix.size()) : what;
      }
    });
    // Records before the earliest error are applied (as a sequential reading would have), then it
    // is reported.
    for (const Records& r : rec) {
      apply(sec, r);
      if (r.failed) throw MpsError(r.error, r.error_line);
    }
    return end;
  }

  void scan_pairs(const Tokens& tk, long long no, bool is_range, std::vector<RowValue>& out) const {
    const int first = tk.n % 2 == 1 ? 1 : 0;  // an optional set name makes the count odd
    if (tk.n - first != 2 && tk.n - first != 4)
      throw MpsError(is_range ? "RANGES entry has the wrong number of fields"
                              : "RHS entry has the wrong number of fields", no);
    for (int f = first; f + 1 < tk.n; f += 2) {
      const int row = find_row(tk[f], no);
      const double value = read_number(tk[f + 1], no);
      if (is_range && row < 0) throw MpsError("RANGES entry on an objective or free row", no);
      out.push_back(RowValue{row, value});
    }
  }

  void apply(Sec sec, const Records& r) {
    if (rhs_.size() < kinds_.size()) {
      rhs_.resize(kinds_.size(), 0.0);
      range_.resize(kinds_.size(), 0.0);
      has_range_.resize(kinds_.size(), 0);
    }
    if (sec == Sec::kRhs) {
      for (const RowValue& x : r.pairs) {
        if (x.row == -1)
          out_.model.offset = -x.value;
        else if (x.row >= 0)
          rhs_[x.row] = x.value;
      }
    } else if (sec == Sec::kRanges) {
      for (const RowValue& x : r.pairs) {
        range_[x.row] = x.value;
        has_range_[x.row] = 1;
      }
    } else {
      for (const BoundRec& b : r.bounds) apply_bound(b);
    }
  }

  // BOUNDS rules: what each type does to the column's lower and upper bounds (kKeep, the entry's
  // value, 0, 1 or an infinity), whether it counts as setting the lower bound, whether it makes
  // the column integer. UP and UI with a negative value and no lower bound set also make the
  // lower bound -inf.
  enum Act : std::uint8_t { kKeep, kValue, kZero, kOne, kMinusInf, kPlusInf };
  struct BoundRule {
    const char* type;
    Act lower, upper;
    bool sets_lower, integer;
  };
  static constexpr BoundRule kBoundRules[] = {
      {"UP", kKeep, kValue, false, false},       {"UI", kKeep, kValue, false, true},
      {"LO", kValue, kKeep, true, false},        {"LI", kValue, kKeep, true, true},
      {"FX", kValue, kValue, true, false},       {"FR", kMinusInf, kPlusInf, true, false},
      {"MI", kMinusInf, kKeep, true, false},     {"PL", kKeep, kPlusInf, false, false},
      {"BV", kZero, kOne, true, true},
  };

  static double act(Act a, double old, double value) {
    switch (a) {
      case kValue: return value;
      case kZero: return 0.0;
      case kOne: return 1.0;
      case kMinusInf: return -kInf;
      case kPlusInf: return kInf;
      case kKeep: break;
    }
    return old;
  }

  void scan_bound(const Tokens& tk, long long no, std::vector<BoundRec>& out) const {
    if (tk.n < 2) throw MpsError("BOUNDS entry is too short", no);
    const std::string_view type = tk[0];
    int rule = -1;
    for (int r = 0; r < static_cast<int>(std::size(kBoundRules)); ++r)
      if (iequals(type, kBoundRules[r].type)) rule = r;
    // An unsupported type is reported after its column and value are read, as the reference does.
    const bool needs_value =
        rule < 0 || kBoundRules[rule].lower == kValue || kBoundRules[rule].upper == kValue;
    // TYPE [set] COLUMN [value]: the set name is there when the field count says so, or, for a
    // type without a value, when the third of three fields names a column.
    int col_pos = 1;
    switch (tk.n) {
      case 2:
        if (needs_value) throw MpsError("BOUNDS entry has the wrong number of fields", no);
        break;
      case 3: col_pos = (!needs_value && cols_index_.find(tk[2]) != NameTable::kAbsent) ? 2 : 1; break;
      case 4: col_pos = 2; break;
      default: throw MpsError("BOUNDS entry has the wrong number of fields", no);
    }
    const int col = cols_index_.find(tk[col_pos]);
    if (col == NameTable::kAbsent) throw MpsError("unknown column '" + std::string(tk[col_pos]) + "'", no);
    const double value = needs_value ? read_number(tk[col_pos + 1], no) : 0.0;
    if (rule < 0) {
      if (iequals(type, "SC")) throw MpsError("semi-continuous bounds (SC) are not supported", no);
      throw MpsError("unknown bound type '" + std::string(type) + "'", no);
    }
    out.push_back(BoundRec{col, rule, value});
  }

  void apply_bound(const BoundRec& b) {
    const BoundRule& rule = kBoundRules[b.rule];
    double& lower = out_.model.col_lower[b.col];
    double& upper = out_.model.col_upper[b.col];
    const bool was_set = lower_set_[b.col] != 0;
    lower = act(rule.lower, lower, b.value);
    upper = act(rule.upper, upper, b.value);
    if (rule.upper == kValue && rule.lower == kKeep && b.value < 0.0 && lower == 0.0 && !was_set)
      lower = -kInf;
    if (rule.sets_lower) lower_set_[b.col] = 1;
    if (rule.integer) out_.col_integer[b.col] = 1;
  }

  // Row bounds from the row's type, its RHS b and its range r (when given): E is [b, b] widened
  // by |r| on the side of r's sign, L is [b - |r|, b], G is [b, b + |r|].
  static void row_bounds(Kind kind, double b, bool ranged, double r, double& lo, double& up) {
    lo = kind == Kind::kL ? -kInf : b;
    up = kind == Kind::kG ? kInf : b;
    if (!ranged) return;
    const double w = std::fabs(r);
    const bool widen_up = kind == Kind::kG || (kind == Kind::kE && r >= 0.0);
    if (widen_up)
      up = b + w;
    else
      lo = b - w;
  }

  MpsModel finish() {
    Model& md = out_.model;
    if (objective_.empty() && kinds_.empty() && out_.col_names.empty()) fail("no ROWS or COLUMNS found");
    line_no_ = 0;  // what follows is about the whole model
    const size_t m = kinds_.size(), n = out_.col_names.size();
    rhs_.resize(m, 0.0);
    range_.resize(m, 0.0);
    has_range_.resize(m, 0);
    md.row_lower.resize(m);
    md.row_upper.resize(m);
    for (size_t i = 0; i < m; ++i)
      row_bounds(kinds_[i], rhs_[i], has_range_[i] != 0, range_[i], md.row_lower[i], md.row_upper[i]);
    md.m = static_cast<int>(m);
    md.n = static_cast<int>(n);
    build_matrix();
    // Model::validate (src/core/model.cpp:59-99), as far as an LP has it.
    if (!std::isfinite(md.offset)) fail("objective offset is not finite");
    for (size_t j = 0; j < n; ++j)
      if (!std::isfinite(md.obj[j])) fail("objective coefficient " + std::to_string(j) + " is not finite");
    for (double v : md.A.value)
      if (!std::isfinite(v)) fail("constraint matrix contains a non-finite value");
    check_bounds("column", md.col_lower, md.col_upper);
    check_bounds("row", md.row_lower, md.row_upper);
    return std::move(out_);
  }

  void check_bounds(const char* what, const std::vector<double>& lo, const std::vector<double>& up) const {
    for (size_t i = 0; i < lo.size(); ++i) {
      if (std::isnan(lo[i]) || std::isnan(up[i])) fail(std::string(what) + " " + std::to_string(i) + " has a NaN bound");
      if (lo[i] == kInf || up[i] == -kInf)
        fail(std::string(what) + " " + std::to_string(i) + " has an infinite bound on the wrong side");
    }
  }

  std::string_view text_;
  bool fixed_;
  int threads_;
  HugeArray<std::string_view> lines_;
  size_t line_count_ = 0;
  Tokens tk_;
  Sec sec_ = Sec::kNone;
  long long line_no_ = 0;
  MpsModel out_;
  std::string_view objective_;
  NameTable rows_, cols_index_;
  std::vector<Kind> kinds_;
  std::vector<double> rhs_, range_;
  std::vector<char> has_range_, lower_set_;
  int integer_state_ = 0;
  std::vector<Part> parts_;
  std::vector<Ref> refs_;
};

int thread_count(int threads) {
  if (threads > 0) return threads;
  const unsigned hw = std::thread::hardware_concurrency();
  return static_cast<int>(std::max(1u, std::min(16u, hw)));
}

void put_number(std::string& out, double v) {
  if (v == kInf) {
    out += "1e30";
    return;
  }
  if (v == -kInf) {
    out += "-1e30";
    return;
  }
  char buf[32];
  const auto r = std::to_chars(buf, buf + sizeof buf, v);
  out.append(buf, r.ptr);
}

}  // namespace

MpsModel read_mps_text(std::string_view text, int threads) {
  const int t = thread_count(threads);
  try {
    return Parser(text, false, t).run();
  } catch (const MpsError& free_error) {
    try {
      return Parser(text, true, t).run();
    } catch (const MpsError&) {
      throw free_error;  // the free-format message is usually the more useful one
    }
  }
}

MpsModel read_mps_file(const std::string& path, int threads) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open '" + path + "'");
  std::ostringstream buffer;
  buffer << in.rdbuf();
  const std::string text = std::move(buffer).str();
  return read_mps_text(text, threads);
}

std::string write_mps(const Model& md, const MpsModel* names) {
  // Free MPS cannot hold a name with a blank (fixed MPS can): such a set of names is replaced by
  // generated ones.
  const auto usable = [](const std::vector<std::string>& v, int count) {
    if (v.size() != static_cast<size_t>(count)) return false;
    for (const std::string& s : v)
      if (s.empty() || s.find_first_of(" \t") != std::string::npos) return false;
    return true;
  };
  const bool named_rows = names != nullptr && usable(names->row_names, md.m);
  const bool named_cols = names != nullptr && usable(names->col_names, md.n);
  const bool marked = names != nullptr && names->col_integer.size() == static_cast<size_t>(md.n);
  const auto row_name = [&](int i) { return named_rows ? names->row_names[i] : "R" + std::to_string(i); };
  const auto col_name = [&](int j) { return named_cols ? names->col_names[j] : "C" + std::to_string(j); };
  // The objective row's name: OBJ, with underscores added until no constraint row has it.
  std::string obj = "OBJ";
  if (named_rows) {
    std::unordered_set<std::string_view> taken(names->row_names.begin(), names->row_names.end());
    while (taken.count(obj) != 0) obj += "_";
  }
  std::string out;
  out.reserve(static_cast<size_t>(md.A.nnz()) * 28 + static_cast<size_t>(md.n + md.m) * 24 + 256);
  const auto line = [&out](std::initializer_list<std::string_view> fields, const double* value) {
    for (std::string_view f : fields) {
      out += ' ';
      out += f;
    }
    if (value != nullptr) {
      out += ' ';
      put_number(out, *value);
    }
    out += '\n';
  };
  out += "NAME ";
  out += names != nullptr && !names->name.empty() ? names->name : std::string("MODEL");
  out += "\n";
  if (md.maximize) out += "OBJSENSE\n    MAX\n";
  out += "ROWS\n";
  line({"N", obj}, nullptr);
  // Row types: E for lower == upper, L for (-inf, upper], G otherwise (a free row is G with an
  // infinite RHS of -1e30). A ranged row is written from whichever end reproduces the other
  // exactly through rhs +- |range|: G (rhs = lower) when lower + (upper - lower) == upper, else L.
  enum RowForm : char { kEq, kLe, kGe };
  std::vector<char> form(static_cast<size_t>(md.m));
  for (int i = 0; i < md.m; ++i) {
    const double lo = md.row_lower[i], up = md.row_upper[i];
    if (lo == up)
      form[i] = kEq;
    else if (lo == -kInf)
      form[i] = up < kInf ? kLe : kGe;
    else if (up == kInf || lo + (up - lo) == up)
      form[i] = kGe;
    else
      form[i] = up - (up - lo) == lo ? kLe : kGe;
    line({form[i] == kEq ? "E" : form[i] == kLe ? "L" : "G", row_name(i)}, nullptr);
  }
  out += "COLUMNS\n";
  bool in_integer = false;
  for (int j = 0; j < md.n; ++j) {
    const bool integer = marked && names->col_integer[j] != 0;
    if (integer != in_integer) {
      line({"MARKER", "'MARKER'", integer ? "'INTORG'" : "'INTEND'"}, nullptr);
      in_integer = integer;
    }
    const std::string c = col_name(j);
    // A column with no entries is still listed (by its objective entry) so that it exists.
    if (md.obj[j] != 0.0 || md.A.start[j] == md.A.start[j + 1]) line({c, obj}, &md.obj[j]);
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p)
      line({c, row_name(md.A.index[p])}, &md.A.value[p]);
  }
  if (in_integer) line({"MARKER", "'MARKER'", "'INTEND'"}, nullptr);
  out += "RHS\n";
  if (md.offset != 0.0) {
    const double v = -md.offset;
    line({"RHS", obj}, &v);
  }
  std::string ranges;
  for (int i = 0; i < md.m; ++i) {
    const double lo = md.row_lower[i], up = md.row_upper[i];
    const double rhs = form[i] == kGe ? lo : up;
    if (rhs != 0.0) line({"RHS", row_name(i)}, &rhs);
    if (form[i] != kEq && lo > -kInf && up < kInf) {
      ranges += " RNG " + row_name(i) + " ";
      put_number(ranges, up - lo);
      ranges += '\n';
    }
  }
  if (!ranges.empty()) out += "RANGES\n" + ranges;
  out += "BOUNDS\n";
  for (int j = 0; j < md.n; ++j) {
    const double lo = md.col_lower[j], up = md.col_upper[j];
    const std::string c = col_name(j);
    if (lo == up) {
      line({"FX", "BND", c}, &lo);
    } else if (lo == -kInf && up == kInf) {
      line({"FR", "BND", c}, nullptr);
    } else {
      if (lo == -kInf)
        line({"MI", "BND", c}, nullptr);
      else if (lo != 0.0 || up < 0.0)  // LO 0 keeps a negative UP from freeing the lower end
        line({"LO", "BND", c}, &lo);
      if (up < kInf) line({"UP", "BND", c}, &up);
    }
  }
  out += "ENDATA\n";
  return out;
}

std::string solution_text(const std::string& status, bool verified, double objective,
                          const std::vector<std::string>& col_names,
                          const std::vector<std::string>& row_names,
                          const std::vector<double>& col_value, const std::vector<double>& col_dual,
                          const std::vector<double>& row_activity,
                          const std::vector<double>& row_dual,
                          const std::vector<double>& certificate,
                          const std::vector<double>& ray) {
  // apps/cli/main.cpp:137-198, numbers with 17 significant digits.
  std::ostringstream out;
  out.precision(17);
  out << "status " << status << "\n";
  out << "verified " << (verified ? "yes" : "no") << "\n";
  if (std::isfinite(objective)) out << "objective " << objective << "\n";
  const size_t n = col_value.size(), m = row_activity.size();
  const auto cname = [&](size_t j) { return j < col_names.size() ? col_names[j] : "c" + std::to_string(j); };
  const auto rname = [&](size_t i) { return i < row_names.size() ? row_names[i] : "r" + std::to_string(i); };
  const bool duals = !row_dual.empty();
  if (n > 0) {
    out << "\ncolumns " << n << "\n# name value" << (duals ? " reduced_cost" : "") << "\n";
    for (size_t j = 0; j < n; ++j) {
      out << cname(j) << " " << col_value[j];
      if (duals) out << " " << col_dual[j];
      out << "\n";
    }
    out << "\nrows " << m << "\n# name activity" << (duals ? " dual" : "") << "\n";
    for (size_t i = 0; i < m; ++i) {
      out << rname(i) << " " << row_activity[i];
      if (duals) out << " " << row_dual[i];
      out << "\n";
    }
  }
  if (!certificate.empty()) {
    out << "\ninfeasibility_certificate " << certificate.size() << "\n";
    for (size_t i = 0; i < certificate.size(); ++i) out << rname(i) << " " << certificate[i] << "\n";
  }
  if (!ray.empty()) {
    out << "\nunbounded_ray " << ray.size() << "\n";
    for (size_t j = 0; j < ray.size(); ++j) out << cname(j) << " " << ray[j] << "\n";
  }
  return out.str();
}

}  // namespace smy::host
