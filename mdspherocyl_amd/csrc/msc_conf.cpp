// conf.dat in the reference's format (print_conf / chrg_conf,
// v25/functions.cpp:1181-1296): host-only readers and writers, so that the
// reference and this library can hand states to each other.
//
// Layout of the file:
//   L[0] L[1] L[2]                      (default stream format, 6 digits)
//   per particle, species-major:
//     r p ori(3x3, row-major) l " | " mass (dipole: pos ori mu " . ")* clust
//   per thermostat (2): s strn
// Every floating-point value after the first line is written fixed with 12
// decimals, as the reference's stream does after setprecision(12).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/msc.h"

namespace {

// whitespace-separated tokens of a whole file
struct Tokens {
  std::vector<std::string> t;
  size_t i = 0;
  bool load(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::string cur;
    int c;
    while ((c = std::fgetc(f)) != EOF) {
      if (c == ' ' || c == '\t' || c == '\n' || c == '\r') {
        if (!cur.empty()) t.push_back(cur), cur.clear();
      } else {
        cur.push_back((char)c);
      }
    }
    if (!cur.empty()) t.push_back(cur);
    std::fclose(f);
    return true;
  }
  bool num(double *v) {
    if (i >= t.size()) return false;
    char *end = nullptr;
    errno = 0;
    *v = std::strtod(t[i].c_str(), &end);
    if (errno || !end || *end) return false;
    ++i;
    return true;
  }
  bool integer(int *v) {
    if (i >= t.size()) return false;
    char *end = nullptr;
    const long x = std::strtol(t[i].c_str(), &end, 10);
    if (!end || *end) return false;
    *v = (int)x;
    ++i;
    return true;
  }
  bool word(const char *w) {
    if (i >= t.size() || t[i] != w) return false;
    ++i;
    return true;
  }
};

}  // namespace

extern "C" {

int msc_write_conf(const char *path, int N, int nd, const double L[3], const double *r, const double *p,
                   const double *ori, const double *l, const double *mass, const double *dpos, const double *dori,
                   const double *dmu, const int *clust, const double thermo[4]) {
  if (!path || N < 0 || nd < 0 || !L || (N > 0 && (!r || !p || !ori || !l || !mass)) ||
      (N > 0 && nd > 0 && (!dpos || !dori || !dmu)) || !thermo)
    return MSC_E_ARG;
  FILE *f = std::fopen(path, "wb");
  if (!f) return MSC_E_ARG;
  std::fprintf(f, "%g %g %g\n", L[0], L[1], L[2]);
  for (int g = 0; g < N; ++g) {
    for (int a = 0; a < 3; ++a) std::fprintf(f, "%.12f ", r[3 * g + a]);
    for (int a = 0; a < 3; ++a) std::fprintf(f, "%.12f ", p[3 * g + a]);
    for (int k = 0; k < 9; ++k) std::fprintf(f, "%.12f ", ori[9 * g + k]);
    std::fprintf(f, "%.12f %.12f %.12f  | %.12f ", l[3 * g], l[3 * g + 1], l[3 * g + 2], mass[g]);
    for (int w = 0; w < nd; ++w) {
      const size_t q = (size_t)g * nd + w;
      for (int a = 0; a < 3; ++a) std::fprintf(f, "%.12f ", dpos[3 * q + a]);
      for (int a = 0; a < 3; ++a) std::fprintf(f, "%.12f ", dori[3 * q + a]);
      std::fprintf(f, "%.12f . ", dmu[q]);
    }
    std::fprintf(f, "%d\n", clust ? clust[g] : 0);
  }
  for (int i = 0; i < 2; ++i) std::fprintf(f, "%.12f %.12f\n", thermo[2 * i], thermo[2 * i + 1]);
  const bool ok = std::ferror(f) == 0;
  return (std::fclose(f) == 0 && ok) ? MSC_OK : MSC_E_ARG;
}

int msc_read_conf(const char *path, int N, int nd, double L[3], double *r, double *p, double *ori, double *l,
                  double *mass, double *dpos, double *dori, double *dmu, int *clust, double thermo[4]) {
  if (!path || N < 0 || nd < 0 || !L || !thermo) return MSC_E_ARG;
  Tokens T;
  if (!T.load(path)) return MSC_E_ARG;
  for (int a = 0; a < 3; ++a)
    if (!T.num(&L[a])) return MSC_E_STATE;
  for (int g = 0; g < N; ++g) {
    double v[18];
    for (int k = 0; k < 18; ++k)
      if (!T.num(&v[k])) return MSC_E_STATE;
    for (int a = 0; a < 3; ++a) {
      if (r) r[3 * g + a] = v[a];
      if (p) p[3 * g + a] = v[3 + a];
      if (l) l[3 * g + a] = v[15 + a];
    }
    if (ori)
      for (int k = 0; k < 9; ++k) ori[9 * g + k] = v[6 + k];
    double m;
    if (!T.word("|") || !T.num(&m)) return MSC_E_STATE;
    if (mass) mass[g] = m;
    for (int w = 0; w < nd; ++w) {
      const size_t q = (size_t)g * nd + w;
      double d[7];
      for (int k = 0; k < 7; ++k)
        if (!T.num(&d[k])) return MSC_E_STATE;
      if (!T.word(".")) return MSC_E_STATE;
      for (int a = 0; a < 3; ++a) {
        if (dpos) dpos[3 * q + a] = d[a];
        if (dori) dori[3 * q + a] = d[3 + a];
      }
      if (dmu) dmu[q] = d[6];
    }
    int cl;
    if (!T.integer(&cl)) return MSC_E_STATE;
    if (clust) clust[g] = cl;
  }
  for (int k = 0; k < 4; ++k)
    if (!T.num(&thermo[k])) return MSC_E_STATE;
  return T.i == T.t.size() ? MSC_OK : MSC_E_STATE;
}

}  // extern "C"
