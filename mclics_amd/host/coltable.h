// A comma-separated table read by column name: the format of the custom-geometry and initial-condition files.
//
// The first line is a header naming the columns. A reader binds the columns it needs by name (case-insensitive, any
// order, other columns ignored) and then pulls one numeric record per non-blank line. Every format error is reported
// as "  [ERROR] <file>, line <n>: <what>" on stderr and ends the program with the status the table was opened with.
#ifndef MCLICS_COLTABLE_H_
#define MCLICS_COLTABLE_H_

#include <fstream>
#include <string>
#include <string_view>
#include <vector>

class ColumnTable {
 public:
  // Opens `path`; a file that cannot be opened ends the program with status 2 (perror), as any input file does.
  // `status` is the exit status of every later format error.
  ColumnTable(const std::string &path, int status);

  // Reads the header and binds `names` in order. Missing names are listed in one error on line 1, followed by
  // "(required: <required>)".
  void bind(const std::vector<const char *> &names, const char *required);

  // Position of the i-th bound column in the header.
  int position(int i) const { return bound_[i]; }

  // Reads the next non-blank line into `out` (one double per bound column). Returns false at the end of the file.
  bool next(double *out);

  // Reports an error on the current line and exits with the table's status.
  [[noreturn]] void fail(const char *fmt, ...) const __attribute__((format(printf, 2, 3)));

  const std::string &path() const { return path_; }

 private:
  bool read_line();
  void cut_cells();

  std::ifstream in_;
  std::string path_;
  int status_;
  int line_ = 0;
  std::string text_;
  std::vector<std::string_view> cells_;
  std::vector<const char *> names_;
  std::vector<int> bound_;
  size_t width_ = 0;  // number of header cells
};

#endif
