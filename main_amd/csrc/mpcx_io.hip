// Files of the reference's restart protocol (host I/O only, no device work):
//   timestep.dat         main_pattern.f95:130-136 (read) and :243-249 (write)
//   mt_state.dat         mtmod.f90:152-229 (mtsavef / mtgetf with form 'u')
//   motor_positions.dat  initial_mod.f95:79,118 (write), restart_mod.f95:410-414 (read)
// and the fixed-form text records of selkov_code (Output_SelkovDimer.f90,
// Restart.f90, Main_SelkovDimer.f90:108-113,152-158).
//
// List-directed files are written in gfortran's layout (the reference's
// Makefile builds with gfortran): each item is a separator blank followed by
// the item's field, I11 for a default integer, I20 for integer(8) and, for
// real(8), G25.17E3 (F20.d plus five blanks for 0.1 <= |x| < 1e17, else
// 1PE25.16E3).  17 significant digits give every double back exactly.  The
// readers accept any list-directed form (blanks or commas as separators,
// r*c repeat counts, D or E or letterless exponents).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "../../include/mpcx.h"
#include "mpcx_host.h"

using mpcx::set_error;

namespace {

struct File {
  FILE* f = nullptr;
  File(const char* path, const char* mode) : f(fopen(path, mode)) {}
  ~File() { if (f) fclose(f); }
};

// ---------------------------------------------------- list-directed input
// Reads the values of one READ statement at a time: a statement starts on a
// fresh record; what is left of its last record is skipped.
class ListReader {
 public:
  explicit ListReader(FILE* f) : f_(f) {}
  void begin() { have_ = false; rep_ = 0; }
  // next value of the current statement; false at the end of the file or a '/'
  bool next(std::string& tok) {
    if (rep_ > 0) { rep_--; tok = reptok_; return true; }
    for (;;) {
      if (!have_) {
        if (!getline()) return false;
        have_ = true;
        pos_ = 0;
      }
      while (pos_ < line_.size() && (line_[pos_] == ' ' || line_[pos_] == '\t' || line_[pos_] == '\r')) pos_++;
      if (pos_ >= line_.size()) { have_ = false; continue; }
      char c = line_[pos_];
      if (c == '/') return false;
      if (c == ',') { bad_ = true; return false; }   // null value
      size_t b = pos_;
      while (pos_ < line_.size() && line_[pos_] != ' ' && line_[pos_] != '\t' && line_[pos_] != ',' &&
             line_[pos_] != '/' && line_[pos_] != '\r')
        pos_++;
      tok = line_.substr(b, pos_ - b);
      while (pos_ < line_.size() && (line_[pos_] == ' ' || line_[pos_] == '\t' || line_[pos_] == '\r')) pos_++;
      if (pos_ < line_.size() && line_[pos_] == ',') pos_++;
      size_t star = tok.find('*');
      if (star != std::string::npos) {   // r*c
        char* end = nullptr;
        long r = strtol(tok.c_str(), &end, 10);
        if (end != tok.c_str() + star || r < 1 || star + 1 >= tok.size()) { bad_ = true; return false; }
        reptok_ = tok.substr(star + 1);
        rep_ = (int)r - 1;
        tok = reptok_;
      }
      return true;
    }
  }
  bool bad() const { return bad_; }

 private:
  bool getline() {
    line_.clear();
    int ch;
    bool any = false;
    while ((ch = fgetc(f_)) != EOF) {
      any = true;
      if (ch == '\n') return true;
      line_.push_back((char)ch);
    }
    return any;
  }
  FILE* f_;
  std::string line_, reptok_;
  size_t pos_ = 0;
  bool have_ = false, bad_ = false;
  int rep_ = 0;
};

bool parse_int(const std::string& t, long long& v) {
  if (t.empty()) return false;
  char* end = nullptr;
  v = strtoll(t.c_str(), &end, 10);
  return end == t.c_str() + t.size();
}

// Fortran real forms: [sign] digits [. digits] [exp], exp = (E|D|Q)[sign]digits
// or a bare sign and digits
bool parse_real(const std::string& t, double& v) {
  if (t.empty()) return false;
  std::string s;
  for (size_t i = 0; i < t.size(); i++) {
    char c = t[i];
    if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'E';
    if ((c == '+' || c == '-') && i > 0 && t[i - 1] != 'e' && t[i - 1] != 'E' && t[i - 1] != 'd' &&
        t[i - 1] != 'D' && t[i - 1] != 'q' && t[i - 1] != 'Q')
      s.push_back('E');
    s.push_back(c);
  }
  for (char c : s)
    if (!(c == '+' || c == '-' || c == '.' || c == 'E' || c == 'e' || (c >= '0' && c <= '9'))) return false;
  char* end = nullptr;
  v = strtod(s.c_str(), &end);
  return end == s.c_str() + s.size();
}

// --------------------------------------------------- list-directed output
// one real(8) item: separator blank + G25.17E3 field
void put_real(std::string& out, double x) {
  char e[64], buf[96];
  snprintf(e, sizeof e, "%.16E", x);   // d.ddddddddddddddddE+xx, 17 significant digits
  const char* ep = strchr(e, 'E');
  const int e10 = ep ? atoi(ep + 1) : 0;
  if (x == 0.0 || (e10 >= -1 && e10 <= 16)) {   // 0.1 <= |x| < 1e17 after rounding: F20.(16-e10), 5 blanks
    const int dec = x == 0.0 ? 16 : 16 - e10;
    char f[64];
    snprintf(f, sizeof f, "%.*f", dec, x);
    snprintf(buf, sizeof buf, " %20s     ", f);
  } else {   // 1PE25.16E3
    char m[32];
    size_t ml = (size_t)(ep - e);
    memcpy(m, e, ml);
    m[ml] = 0;
    char f[64];
    snprintf(f, sizeof f, "%sE%c%03d", m, e10 < 0 ? '-' : '+', e10 < 0 ? -e10 : e10);
    snprintf(buf, sizeof buf, " %25s", f);
  }
  out += buf;
}

// -------------------------------------------------- fixed-form output
// Ew.d (0.ddddE+xx) and ESw.d (d.dddE+xx) of the Fortran standard with the
// default exponent width: E and two digits, or a sign and three digits when
// the exponent exceeds 99.  Rounding is to nearest (the processor default).
bool fmt_e(char* out, size_t cap, double x, int w, int d, bool es) {
  char e[64];
  const int sig = es ? d + 1 : d;
  snprintf(e, sizeof e, "%.*E", sig - 1, x);   // m.mmmE+xx with sig digits
  const char* ep = strchr(e, 'E');
  if (!ep) return false;
  int e10 = atoi(ep + 1);
  std::string digits;
  for (const char* c = e; c < ep; c++)
    if (*c >= '0' && *c <= '9') digits.push_back(*c);
  const bool neg = e[0] == '-';
  std::string body;
  int ex;
  if (es) {
    body = digits.substr(0, 1) + "." + digits.substr(1);
    ex = x == 0.0 ? 0 : e10;
  } else {
    body = "0." + digits;
    ex = x == 0.0 ? 0 : e10 + 1;
  }
  char xp[16];
  const int ax = ex < 0 ? -ex : ex;
  if (ax <= 99) snprintf(xp, sizeof xp, "E%c%02d", ex < 0 ? '-' : '+', ax);
  else if (ax <= 999) snprintf(xp, sizeof xp, "%c%03d", ex < 0 ? '-' : '+', ax);
  else return false;
  std::string f = (neg ? "-" : "") + body + xp;
  if ((int)f.size() > w) {   // the optional leading zero of E editing goes first
    if (!es && f.size() == (size_t)w + 1) f.erase(neg ? 1 : 0, 1);
    else return false;
  }
  if ((size_t)w + 1 > cap) return false;
  snprintf(out, cap, "%*s", w, f.c_str());
  return true;
}

}  // namespace

// ================================================================ timestep.dat
extern "C" int mpcx_timestep_write(const char* path, int32_t cstep, int64_t n_a, int64_t n_b, int64_t n_s) {
  if (!path) return set_error(MPCX_E_ARG, "null path");
  File f(path, "w");
  if (!f.f) return set_error(MPCX_E_ARG, "timestep.dat: cannot open for writing");
  fprintf(f.f, "%12d\n%21lld\n%21lld\n%21lld\n", (int)cstep, (long long)n_a, (long long)n_b, (long long)n_s);
  return ferror(f.f) ? set_error(MPCX_E_ARG, "timestep.dat: write failed") : 0;
}

extern "C" int mpcx_timestep_read(const char* path, int32_t* cstep, int64_t* n_a, int64_t* n_b, int64_t* n_s) {
  if (!path || !cstep || !n_a || !n_b || !n_s) return set_error(MPCX_E_ARG, "null argument");
  File f(path, "r");
  if (!f.f) return set_error(MPCX_E_ARG, "timestep.dat: cannot open");
  ListReader r(f.f);
  long long v[4];
  for (int i = 0; i < 4; i++) {
    std::string t;
    r.begin();
    if (!r.next(t) || !parse_int(t, v[i])) return set_error(MPCX_E_ARG, "timestep.dat: an integer is missing");
  }
  if (v[0] < INT32_MIN || v[0] > INT32_MAX) return set_error(MPCX_E_ARG, "timestep.dat: cstep out of range");
  *cstep = (int32_t)v[0]; *n_a = v[1]; *n_b = v[2]; *n_s = v[3];
  return 0;
}

// ================================================================ mt_state.dat
// Sequential unformatted records of mtsavef: [n] mti(1:T) [n] [m] mt(1:624,1:T) [m],
// 4-byte little-endian record markers holding the record lengths in bytes.
extern "C" int mpcx_mt_state_write(const char* path, const uint32_t* mt, const int32_t* mti, int32_t nstreams) {
  if (!path || !mt || !mti) return set_error(MPCX_E_ARG, "null argument");
  if (nstreams < 1 || nstreams > 65536) return set_error(MPCX_E_ARG, "mt_state.dat: 1 to 65536 streams");
  File f(path, "wb");
  if (!f.f) return set_error(MPCX_E_ARG, "mt_state.dat: cannot open for writing");
  const uint32_t n1 = 4u * (uint32_t)nstreams, n2 = 624u * 4u * (uint32_t)nstreams;
  bool ok = fwrite(&n1, 4, 1, f.f) == 1 && fwrite(mti, 4, (size_t)nstreams, f.f) == (size_t)nstreams &&
            fwrite(&n1, 4, 1, f.f) == 1 && fwrite(&n2, 4, 1, f.f) == 1 &&
            fwrite(mt, 4, (size_t)624 * nstreams, f.f) == (size_t)624 * nstreams && fwrite(&n2, 4, 1, f.f) == 1;
  return ok ? 0 : set_error(MPCX_E_ARG, "mt_state.dat: write failed");
}

extern "C" int mpcx_mt_state_read(const char* path, uint32_t* mt, int32_t* mti, int32_t max_streams,
                                  int32_t* nstreams) {
  if (!path || !nstreams) return set_error(MPCX_E_ARG, "null argument");
  File f(path, "rb");
  if (!f.f) return set_error(MPCX_E_ARG, "mt_state.dat: cannot open");
  uint32_t n1 = 0, t1 = 0, n2 = 0, t2 = 0;
  if (fread(&n1, 4, 1, f.f) != 1 || n1 == 0 || n1 % 4 != 0 || n1 > 4u * 65536u)
    return set_error(MPCX_E_ARG, "mt_state.dat: bad first record marker");
  const int32_t T = (int32_t)(n1 / 4);
  *nstreams = T;
  if (!mt || !mti) return 0;   // size query
  if (T > max_streams) return set_error(MPCX_E_ARG, "mt_state.dat: more streams than the buffers hold");
  if (fread(mti, 4, (size_t)T, f.f) != (size_t)T || fread(&t1, 4, 1, f.f) != 1 || t1 != n1)
    return set_error(MPCX_E_ARG, "mt_state.dat: truncated first record");
  if (fread(&n2, 4, 1, f.f) != 1 || n2 != 624u * 4u * (uint32_t)T)
    return set_error(MPCX_E_ARG, "mt_state.dat: second record is not mt(624,T)");
  if (fread(mt, 4, (size_t)624 * T, f.f) != (size_t)624 * T || fread(&t2, 4, 1, f.f) != 1 || t2 != n2)
    return set_error(MPCX_E_ARG, "mt_state.dat: truncated second record");
  return 0;
}

// =========================================================== motor_positions.dat
// cm is motorCM(numdims,3), column-major: cm[k + d*nd]
extern "C" int mpcx_motor_positions_write(const char* path, const double* cm, int32_t nd) {
  if (!path || !cm || nd < 1) return set_error(MPCX_E_ARG, "bad argument");
  File f(path, "w");
  if (!f.f) return set_error(MPCX_E_ARG, "motor_positions.dat: cannot open for writing");
  for (int k = 0; k < nd; k++) {
    std::string line;
    for (int d = 0; d < 3; d++) {
      if (!isfinite(cm[k + (size_t)d * nd])) return set_error(MPCX_E_ARG, "motor_positions.dat: value not finite");
      put_real(line, cm[k + (size_t)d * nd]);
    }
    fprintf(f.f, "%s\n", line.c_str());
  }
  return ferror(f.f) ? set_error(MPCX_E_ARG, "motor_positions.dat: write failed") : 0;
}

extern "C" int mpcx_motor_positions_read(const char* path, double* cm, int32_t nd) {
  if (!path || !cm || nd < 1) return set_error(MPCX_E_ARG, "bad argument");
  File f(path, "r");
  if (!f.f) return set_error(MPCX_E_ARG, "motor_positions.dat: cannot open");
  ListReader r(f.f);
  for (int k = 0; k < nd; k++) {
    r.begin();
    for (int d = 0; d < 3; d++) {
      std::string t;
      double v;
      if (!r.next(t) || !parse_real(t, v)) return set_error(MPCX_E_ARG, "motor_positions.dat: a value is missing");
      cm[k + (size_t)d * nd] = v;
    }
  }
  return 0;
}

// ============================================================ fixed-form text
// One record of edit descriptors from a small set: 'I' (Iw), 'E' (Ew.d), 'S'
// (ESw.d) and 'X' (nX), given as a spec string such as "I8 X1 E23.16 X1 ..."
// with the values in order (integers as doubles holding whole numbers).  As
// in Fortran, X only positions: a trailing X adds no blanks to the record.
extern "C" int mpcx_format_record(const char* spec, const double* values, int32_t nvalues, char* out,
                                  int64_t cap) {
  if (!spec || !out || cap < 2) return set_error(MPCX_E_ARG, "bad argument");
  std::string line;
  size_t skip = 0;
  int vi = 0;
  const char* p = spec;
  while (*p) {
    while (*p == ' ') p++;
    if (!*p) break;
    const char kind = *p++;
    char* end = nullptr;
    const long w = strtol(p, &end, 10);
    if (end == p || w < 1 || w > 64) return set_error(MPCX_E_ARG, "format: bad width");
    p = end;
    long d = 0;
    if (*p == '.') {
      d = strtol(p + 1, &end, 10);
      if (end == p + 1 || d < 1 || d > 30) return set_error(MPCX_E_ARG, "format: bad digits");
      p = end;
    }
    char buf[96];
    if (kind == 'X') {   // positioning only: blanks appear when a later item is written
      skip += (size_t)w;
      continue;
    }
    if (vi >= nvalues || !values) return set_error(MPCX_E_ARG, "format: more descriptors than values");
    const double v = values[vi++];
    line.append(skip, ' ');
    skip = 0;
    if (kind == 'I') {
      if (v != floor(v) || fabs(v) > 9.2e18) return set_error(MPCX_E_ARG, "format: I value is not an integer");
      char t[32];
      snprintf(t, sizeof t, "%lld", (long long)v);
      if ((long)strlen(t) > w) line.append((size_t)w, '*');
      else { snprintf(buf, sizeof buf, "%*s", (int)w, t); line += buf; }
    } else if (kind == 'E' || kind == 'S') {
      if (!fmt_e(buf, sizeof buf, v, (int)w, (int)d, kind == 'S')) line.append((size_t)w, '*');
      else line += buf;
    } else {
      return set_error(MPCX_E_ARG, "format: unknown descriptor");
    }
  }
  if (vi != nvalues) return set_error(MPCX_E_ARG, "format: more values than descriptors");
  if ((int64_t)line.size() + 1 > cap) return set_error(MPCX_E_ARG, "format: output buffer too small");
  memcpy(out, line.c_str(), line.size() + 1);
  return 0;
}
