// Test stand-ins for the particle containers of the reference (v25/Matrix.hpp,
// v25/classes.hpp): the member functions the adapter calls, with the
// reference's names and parameter types, over plain storage.  They let the
// adapter (mdspherocyl_amd/adapter) be compiled and run without the
// reference's headers, which need GSL.  Class and template names equal the
// reference's, so the adapter's symbols mangle as they would against the
// real headers.  TEST INFRASTRUCTURE: no product code includes this file.
#pragma once

#include <cmath>
#include <deque>
#include <vector>

using namespace std;

template <typename T>
class Dynvec {
  std::vector<T> m_v;

 public:
  Dynvec() {}
  explicit Dynvec(int n) : m_v(n) {}
  int get_size() const { return (int)m_v.size(); }
  void set_size(int n) { m_v.assign(n, T()); }
  T &operator[](int i) { return m_v[i]; }
  T const &operator[](int i) const { return m_v[i]; }
};

template <typename T>
class Matrix {
  int m_rows = 1, m_cols = 1;
  std::vector<T> m_d = std::vector<T>(1);

 public:
  Matrix() {}
  Matrix(int rows, int cols) : m_rows(rows), m_cols(cols), m_d((size_t)rows * cols) {}
  Matrix(const Matrix<T> &) = default;
  void get_size(int &rows, int &cols) const {
    rows = m_rows;
    cols = m_cols;
  }
  void set_size(int rows, int cols) {
    m_rows = rows;
    m_cols = cols;
    m_d.assign((size_t)rows * cols, T());
  }
  T *operator[](int i) { return m_d.data() + (size_t)i * m_cols; }
  T const *operator[](int i) const { return m_d.data() + (size_t)i * m_cols; }
  // the reference's assignment keeps the target's size (and refuses others)
  void operator=(const Matrix<T> &o) {
    if (o.m_rows == m_rows && o.m_cols == m_cols) m_d = o.m_d;
  }
};

class part_num {
  int m_Ncomp = 1, m_Tpart = 0;
  std::vector<int> m_Npart;

 public:
  void set(const std::vector<int> &counts) {
    m_Npart = counts;
    m_Ncomp = (int)counts.size();
    m_Tpart = 0;
    for (int c : counts) m_Tpart += c;
  }
  int get_Ncomp() const { return m_Ncomp; }
  int get_Tpart() const { return m_Tpart; }
  int get_Npart(int comp) const { return m_Npart[comp]; }
};

class Dipole {
  double m_u = 0.0, m_r[3] = {0, 0, 0}, m_o[3] = {0, 0, 1};

 public:
  double get_dip() const { return m_u; }
  void set_dip(double u) { m_u = u; }
  void get_pos(double r[3]) const { r[0] = m_r[0], r[1] = m_r[1], r[2] = m_r[2]; }
  void set_pos(const double r[3]) { m_r[0] = r[0], m_r[1] = r[1], m_r[2] = r[2]; }
  void get_ori(double o[3]) const { o[0] = m_o[0], o[1] = m_o[1], o[2] = m_o[2]; }
  void set_ori(const double o[3]) { m_o[0] = o[0], m_o[1] = o[1], m_o[2] = o[2]; }
};

class Spherocyl {
  double m_mass = 1, m_L = 0, m_D = 1;
  double m_r[3] = {0, 0, 0}, m_p[3] = {0, 0, 0}, m_l[3] = {0, 0, 0}, m_F[3] = {0, 0, 0}, m_T[3] = {0, 0, 0};
  Matrix<double> m_ori = Matrix<double>(3, 3);
  std::vector<double> m_pos_constr;  // signed distance of each dipole along the axis
  std::vector<Dipole> m_u;           // lab-frame dipoles

 public:
  double get_mass() const { return m_mass; }
  void set_mass(double m) { m_mass = m; }
  void get_shape(double sh[2]) const { sh[0] = m_L, sh[1] = m_D; }
  void set_shape(const double sh[2]) { m_L = sh[0], m_D = sh[1]; }
  int get_ndips() const { return (int)m_u.size(); }
  Dipole get_dip(int i) const { return m_u[i]; }
  void get_pos(double r[3]) const { r[0] = m_r[0], r[1] = m_r[1], r[2] = m_r[2]; }
  void set_pos(const double r[3]) { m_r[0] = r[0], m_r[1] = r[1], m_r[2] = r[2]; }
  void get_mom(double p[3]) const { p[0] = m_p[0], p[1] = m_p[1], p[2] = m_p[2]; }
  void set_mom(const double p[3]) { m_p[0] = p[0], m_p[1] = p[1], m_p[2] = p[2]; }
  void get_angm(double l[3]) const { l[0] = m_l[0], l[1] = m_l[1], l[2] = m_l[2]; }
  void set_angm(const double l[3]) { m_l[0] = l[0], m_l[1] = l[1], m_l[2] = l[2]; }
  void get_F(double f[3]) const { f[0] = m_F[0], f[1] = m_F[1], f[2] = m_F[2]; }
  void set_F(const double f[3]) { m_F[0] = f[0], m_F[1] = f[1], m_F[2] = f[2]; }
  void add_F(const double f[3]) { m_F[0] += f[0], m_F[1] += f[1], m_F[2] += f[2]; }
  void get_T(double t[3]) const { t[0] = m_T[0], t[1] = m_T[1], t[2] = m_T[2]; }
  void set_T(const double t[3]) { m_T[0] = t[0], m_T[1] = t[1], m_T[2] = t[2]; }
  void add_T(const double t[3]) { m_T[0] += t[0], m_T[1] += t[1], m_T[2] += t[2]; }
  void get_omat(Matrix<double> &o) const { o = m_ori; }
  void set_omat(const Matrix<double> &o) { m_ori = o; }
  void get_lev(int i, double d[3]) const {
    for (int a = 0; a < 3; ++a) d[a] = m_pos_constr[i] * m_ori[2][a];
  }
  // body-frame dipoles: positions dpc along the axis, orientations doc in
  // body coordinates, moments mu; lab frame r + dpc u and O^T doc
  void set_dipoles(int nd, const double *dpc, const double *doc, const double *mu) {
    m_pos_constr.assign(dpc, dpc + nd);
    m_u.assign(nd, Dipole());
    for (int w = 0; w < nd; ++w) {
      double r[3], o[3];
      for (int a = 0; a < 3; ++a) {
        r[a] = m_r[a] + dpc[w] * m_ori[2][a];
        o[a] = m_ori[0][a] * doc[3 * w] + m_ori[1][a] * doc[3 * w + 1] + m_ori[2][a] * doc[3 * w + 2];
      }
      m_u[w].set_pos(r);
      m_u[w].set_ori(o);
      m_u[w].set_dip(mu[w]);
    }
  }
};

struct cross {
  Dynvec<int> C;
  Dynvec<int> P;
  Dynvec<bool> T;
};

class Thermostat;
class accumulator;
