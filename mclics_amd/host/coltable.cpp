// Column-table reader of the custom-geometry and initial-condition files.
#include "coltable.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

namespace {

constexpr std::string_view kBlank = " \t\r";

std::string_view trim(std::string_view s) {
  const size_t a = s.find_first_not_of(kBlank);
  if (a == std::string_view::npos) return {};
  return s.substr(a, s.find_last_not_of(kBlank) - a + 1);
}

bool same_name(std::string_view cell, const char *name) {
  const size_t n = strlen(name);
  return cell.size() == n && strncasecmp(cell.data(), name, n) == 0;
}

}  // namespace

ColumnTable::ColumnTable(const std::string &path, int status) : in_(path), path_(path), status_(status) {
  if (!in_) {
    perror(path_.c_str());
    exit(2);
  }
}

void ColumnTable::fail(const char *fmt, ...) const {
  char what[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(what, sizeof(what), fmt, ap);
  va_end(ap);
  fprintf(stderr, "  [ERROR] %s, line %d: %s\n", path_.c_str(), line_, what);
  exit(status_);
}

bool ColumnTable::read_line() {
  if (!std::getline(in_, text_)) return false;
  line_++;
  return true;
}

// Cells are views into text_, cut at every comma and trimmed; a line of n commas has n + 1 cells.
void ColumnTable::cut_cells() {
  cells_.clear();
  std::string_view rest(text_);
  for (;;) {
    const size_t comma = rest.find(',');
    cells_.push_back(trim(rest.substr(0, comma)));
    if (comma == std::string_view::npos) break;
    rest.remove_prefix(comma + 1);
  }
}

void ColumnTable::bind(const std::vector<const char *> &names, const char *required) {
  if (!read_line()) {
    line_ = 1;
    fail("empty file (a header line is expected).");
  }
  cut_cells();
  width_ = cells_.size();
  names_ = names;
  bound_.assign(names.size(), -1);
  std::string missing;
  for (size_t c = 0; c < names.size(); c++) {
    for (size_t h = 0; h < width_; h++)
      if (same_name(cells_[h], names[c])) bound_[c] = (int)h;  // the last matching header cell wins
    if (bound_[c] < 0) {
      if (!missing.empty()) missing += ", ";
      missing += names[c];
    }
  }
  if (!missing.empty()) fail("missing column(s) in the header: %s (required: %s).", missing.c_str(), required);
}

bool ColumnTable::next(double *out) {
  do {
    if (!read_line()) return false;
  } while (trim(text_).empty());
  cut_cells();
  if (cells_.size() != width_) fail("%zu fields, but the header has %zu.", cells_.size(), width_);
  for (size_t c = 0; c < bound_.size(); c++) {
    const std::string cell(cells_[bound_[c]]);
    char *end = nullptr;
    out[c] = strtod(cell.c_str(), &end);
    if (cell.empty() || *end != '\0') fail("column %s is not a number: '%s'.", names_[c], cell.c_str());
  }
  return true;
}
