// Text formatting of results files (host code): GEMMA .assoc.txt rows and the
// %.10g matrices of .cXX.txt / .eigenD.txt / .eigenU.txt, formatted by several
// threads into one caller buffer.
//
// The numbers go through snprintf, whose %.6e / %.3f / %.10g conversions are
// correctly rounded (round-half-even on the exact binary value), as Python's
// format() is: a row is byte-identical to the reference's f-strings
// (src/jamma/lmm/io.py:12-147, results.py:222-286, kinship/io.py,
// lmm/eigen_io.py). NaN is printed "nan" whatever its sign bit, as Python does
// (C would print "-nan" for a negative NaN).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/jamma_hip.h"

namespace {

inline void put_num(std::string& s, const char* fmt, double v) {
  if (std::isnan(v)) {
    s += "nan";
    return;
  }
  char buf[64];
  const int k = std::snprintf(buf, sizeof(buf), fmt, v);
  s.append(buf, (size_t)k);
}

template <typename RowFn>
int64_t format_parallel(int64_t n_rows, int n_threads, char* out, int64_t out_cap, RowFn row) {
  if (n_rows <= 0) return 0;
  int nt = n_threads > 0 ? n_threads : 1;
  if (nt > 64) nt = 64;
  if ((int64_t)nt > n_rows) nt = (int)n_rows;
  std::vector<std::string> parts((size_t)nt);
  auto work = [&](int t) {
    const int64_t r0 = n_rows * t / nt, r1 = n_rows * (t + 1) / nt;
    std::string& s = parts[(size_t)t];
    s.reserve((size_t)(r1 - r0) * 160);
    for (int64_t r = r0; r < r1; ++r) row(r, s);
  };
  if (nt == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    th.reserve((size_t)nt);
    for (int t = 0; t < nt; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
  }
  int64_t total = 0;
  for (const auto& s : parts) total += (int64_t)s.size();
  if (total > out_cap) return -total;
  int64_t pos = 0;
  for (const auto& s : parts) {
    std::memcpy(out + pos, s.data(), s.size());
    pos += (int64_t)s.size();
  }
  return total;
}

}  // namespace

extern "C" int64_t jm_format_assoc_rows(int64_t n_rows, const char* pre, const int64_t* pre_off,
                                        const int64_t* n_miss, const char* mid,
                                        const int64_t* mid_off, const double* af,
                                        const double* stats, int n_stats, int64_t ld_stats,
                                        char* out, int64_t out_cap, int n_threads) {
  if (n_rows < 0 || n_stats < 0 || (n_rows > 0 && (!pre || !pre_off || !n_miss || !mid ||
                                                   !mid_off || !af || !out ||
                                                   (n_stats > 0 && !stats))))
    return INT64_MIN;
  return format_parallel(n_rows, n_threads, out, out_cap, [&](int64_t r, std::string& s) {
    s.append(pre + pre_off[r], (size_t)(pre_off[r + 1] - pre_off[r]));
    s += std::to_string(n_miss[r]);
    s += '\t';
    s.append(mid + mid_off[r], (size_t)(mid_off[r + 1] - mid_off[r]));
    put_num(s, "%.3f", af[r]);
    const double* st = stats + r * ld_stats;
    for (int k = 0; k < n_stats; ++k) {
      s += '\t';
      put_num(s, "%.6e", st[k]);
    }
    s += '\n';
  });
}

extern "C" int64_t jm_format_g10_rows(int64_t n_rows, int64_t n_cols, const double* a, int64_t lda,
                                      char* out, int64_t out_cap, int n_threads) {
  if (n_rows < 0 || n_cols < 0 || (n_rows > 0 && n_cols > 0 && (!a || !out))) return INT64_MIN;
  return format_parallel(n_rows, n_threads, out, out_cap, [&](int64_t r, std::string& s) {
    const double* row = a + r * lda;
    for (int64_t c = 0; c < n_cols; ++c) {
      if (c) s += '\t';
      put_num(s, "%.10g", row[c]);
    }
    s += '\n';
  });
}

// ---------------------------------------------------------------------------
// Parsing whitespace-separated numeric text (GEMMA .cXX.txt / .eigenU.txt /
// .eigenD.txt): the lines are found first, then row ranges are converted by
// several threads with strtod (correctly rounded, as NumPy's loadtxt is).
// ---------------------------------------------------------------------------
extern "C" int64_t jm_parse_text_rows(const char* buf, int64_t len, double* out, int64_t out_cap,
                                      int64_t* n_rows, int64_t* n_cols, int n_threads) {
  if (!buf || len < 0 || !n_rows || !n_cols || (out_cap > 0 && !out)) return -1;
  // line starts of the non-blank lines
  std::vector<int64_t> starts, ends;
  int64_t pos = 0;
  while (pos < len) {
    const char* nl = (const char*)std::memchr(buf + pos, '\n', (size_t)(len - pos));
    const int64_t e = nl ? (int64_t)(nl - buf) : len;
    bool blank = true;
    for (int64_t i = pos; i < e; ++i)
      if (buf[i] != ' ' && buf[i] != '\t' && buf[i] != '\r') {
        blank = false;
        break;
      }
    if (!blank) {
      starts.push_back(pos);
      ends.push_back(e);
    }
    pos = e + 1;
  }
  const int64_t rows = (int64_t)starts.size();
  *n_rows = rows;
  *n_cols = 0;
  if (rows == 0) return 0;
  auto count_tokens = [&](int64_t r) {
    int64_t c = 0;
    bool in = false;
    for (int64_t i = starts[(size_t)r]; i < ends[(size_t)r]; ++i) {
      const bool ws = buf[i] == ' ' || buf[i] == '\t' || buf[i] == '\r';
      if (!ws && !in) ++c;
      in = !ws;
    }
    return c;
  };
  const int64_t cols = count_tokens(0);
  *n_cols = cols;
  if (rows * cols > out_cap) return -2;
  int nt = n_threads > 0 ? (n_threads > 64 ? 64 : n_threads) : 1;
  if ((int64_t)nt > rows) nt = (int)rows;
  std::vector<int> bad((size_t)nt, 0);
  auto work = [&](int t) {
    const int64_t r0 = rows * t / nt, r1 = rows * (t + 1) / nt;
    std::string tok;
    for (int64_t r = r0; r < r1 && !bad[(size_t)t]; ++r) {
      int64_t i = starts[(size_t)r];
      const int64_t e = ends[(size_t)r];
      int64_t c = 0;
      while (true) {
        while (i < e && (buf[i] == ' ' || buf[i] == '\t' || buf[i] == '\r')) ++i;
        if (i >= e) break;
        int64_t j = i;
        while (j < e && buf[j] != ' ' && buf[j] != '\t' && buf[j] != '\r') ++j;
        if (c >= cols) {
          bad[(size_t)t] = 1;
          break;
        }
        tok.assign(buf + i, (size_t)(j - i));  // NUL-terminated copy for strtod
        char* endp = nullptr;
        const double v = std::strtod(tok.c_str(), &endp);
        if (endp != tok.c_str() + tok.size()) {
          bad[(size_t)t] = 2;
          break;
        }
        out[r * cols + c] = v;
        ++c;
        i = j;
      }
      if (!bad[(size_t)t] && c != cols) bad[(size_t)t] = 1;
    }
  };
  if (nt == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
  }
  for (int b : bad)
    if (b) return b == 1 ? -3 : -4;
  return rows * cols;
}
