// Minimum segment-segment distance between two spherocylinders, FP64,
// following spherocyl_dist (v25/functions.cpp:89-489) operation for
// operation so that list and contact decisions are bit-identical to the
// reference.  The translation unit is compiled with -ffp-contract=off: no
// multiply-add may be fused here.
#pragma once

#include "msc_common.hpp"

namespace msc {

MSC_HD double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
MSC_HD void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
MSC_HD void seg_vec(const double rAB[3], double s, double t, double LA, const double uA[3],
                    double LB, const double uB[3], double d[3]) {
  for (int i = 0; i < 3; ++i) d[i] = rAB[i] + (s - 0.5) * LA * uA[i] - (t - 0.5) * LB * uB[i];
}
MSC_HD double clamp01(double x) {
  if (x > 1) x = 1;
  else if (x < 0) x = 0;
  return x;
}

// rA/rB: absolute centre positions (only the intersecting-lines branch reads
// them), uA/uB: axes, rAB: minimum-image rA - rB.  Returns d^2, writes the
// joining vector and (s, t) or the reference's sentinel codes.
MSC_HD double segdist(const double rA[3], const double uA[3], double LA, const double rB[3],
                      const double uB[3], double LB, const double Lb[3], const double rAB[3],
                      double dAB[3], double st[2]) {
  double doto = dot3(uA, uB), dotrA = dot3(rAB, uA), dotrB = dot3(rAB, uB);
  double a = LA * LA;
  double b = doto * LA * LB;
  double c = LB * LB;
  double d = dotrA * LA - a / 2 + b / 2;
  double e = dotrB * LB + c / 2 - b / 2;
  double det = a * c - b * b;
  double s = b * e - c * d;
  double t = a * e - b * d;
  double dis[3], dsq = 0.0;
  if (det > 1.e-12) {
    double s0 = s / det, t0 = t / det;
    seg_vec(rAB, s0, t0, LA, uA, LB, uB, dis);
    double dsq0 = dot3(dis, dis);
    if (dsq0 < 1.e-12) {
      // the infinite lines meet: endpoint search on absolute positions
      double pp[4][3], dv[4][3], dd[4];
      for (int k = 0; k < 3; ++k) {
        pp[0][k] = rA[k] + LA * uA[k] / 2;
        pp[1][k] = rA[k] - LA * uA[k] / 2;
        pp[2][k] = rB[k] + LB * uB[k] / 2;
        pp[3][k] = rB[k] - LB * uB[k] / 2;
      }
      for (int k = 0; k < 3; ++k) {
        dv[0][k] = wrapd(pp[0][k] - pp[2][k], Lb[k]);
        dv[1][k] = wrapd(pp[0][k] - pp[3][k], Lb[k]);
        dv[2][k] = wrapd(pp[1][k] - pp[2][k], Lb[k]);
        dv[3][k] = wrapd(pp[1][k] - pp[3][k], Lb[k]);
      }
      int im = 0;
      double mind = 1.e6;
      for (int q = 0; q < 4; ++q) {
        dd[q] = dot3(dv[q], dv[q]);
        if (dd[q] < mind) { im = q; mind = dd[q]; }
      }
      // case im: first t, the pp entry subtracted from pp[1] (projected on
      // A), second s, the pp entry subtracted from pp[3] (projected on B);
      // selected with conditionals so nothing is indexed at run time
      const bool even = (im == 0 || im == 2);
      const int tfirst = even ? 1 : 0;
      const int ssec = (im < 2) ? 1 : 0;
      double vec[3], dsel[3];
      for (int k = 0; k < 3; ++k) {
        double pa = even ? pp[2][k] : pp[3][k];
        vec[k] = wrapd(pp[1][k] - pa, Lb[k]);
        dsel[k] = im == 0 ? dv[0][k] : (im == 1 ? dv[1][k] : (im == 2 ? dv[2][k] : dv[3][k]));
      }
      t = tfirst;
      s = dot3(vec, uA);
      s /= -LA;
      if (s >= 0 && s <= 1) {
        seg_vec(rAB, s, t, LA, uA, LB, uB, dAB);
        dsq = dot3(dAB, dAB);
      } else {
        s = ssec;
        for (int k = 0; k < 3; ++k) vec[k] = wrapd(pp[3][k] - ((im < 2) ? pp[0][k] : pp[1][k]), Lb[k]);
        t = dot3(vec, uB);
        t /= -LB;
        if (t >= 0 && t <= 1) {
          seg_vec(rAB, s, t, LA, uA, LB, uB, dAB);
          dsq = dot3(dAB, dAB);
        } else {
          t = tfirst;
          for (int k = 0; k < 3; ++k) dAB[k] = dsel[k];
          dsq = dot3(dAB, dAB);
        }
      }
    } else {
      // nine regions of the unconstrained optimum relative to [0,1]^2
      int region;
      if (s >= 0) {
        if (s <= det) region = (t >= 0) ? ((t <= det) ? 5 : 6) : 4;
        else region = (t >= 0) ? ((t <= det) ? 2 : 3) : 1;
      } else {
        region = (t >= 0) ? ((t <= det) ? 8 : 9) : 7;
      }
      if (region == 5) {
        s /= det; t /= det;
        seg_vec(rAB, s, t, LA, uA, LB, uB, dAB);
        dsq = dot3(dAB, dAB);
      } else if (region == 2 || region == 8 || region == 6 || region == 4) {
        if (region == 2) { s = 1; t = clamp01((e + b) / c); }
        else if (region == 8) { s = 0; t = clamp01(e / c); }
        else if (region == 6) { t = 1; s = clamp01(-(d - b) / a); }
        else { t = 0; s = clamp01(-d / a); }
        seg_vec(rAB, s, t, LA, uA, LB, uB, dAB);
        dsq = dot3(dAB, dAB);
      } else {
        double s1, t1;
        if (region == 1 || region == 3) { s = 1; t = clamp01((e + b) / c); }
        else if (region == 7) { t = 0; s = clamp01(-d / a); }
        else { t = 1; s = clamp01(-(d - b) / a); }
        seg_vec(rAB, s, t, LA, uA, LB, uB, dAB);
        dsq = dot3(dAB, dAB);
        if (region == 1) { t1 = 0; s1 = clamp01(-d / a); }
        else if (region == 3) { t1 = 1; s1 = clamp01(-(d - b) / a); }
        else { s1 = 0; t1 = clamp01(e / c); }
        seg_vec(rAB, s1, t1, LA, uA, LB, uB, dis);
        double dsq1 = dot3(dis, dis);
        if (dsq1 < dsq) {
          for (int k = 0; k < 3; ++k) dAB[k] = dis[k];
          s = s1; t = t1; dsq = dsq1;
        }
      }
    }
  } else {
    double le = (LA + LB) / 2;
    if (fabs(dotrA) < le) {
      for (int k = 0; k < 3; ++k) dAB[k] = rAB[k] - dotrA * uA[k];
      dsq = dot3(dAB, dAB);
      s = t = -1.e+6;
    } else {
      double d2[3], dsq0;
      double sg = (doto > 0) ? 1.0 : -1.0;  // head-tail (+) or head-head (-)
      for (int k = 0; k < 3; ++k) {
        if (doto > 0) {
          dAB[k] = rAB[k] + LA * uA[k] / 2 + LB * uB[k] / 2;
          d2[k] = rAB[k] - LA * uA[k] / 2 - LB * uB[k] / 2;
        } else {
          dAB[k] = rAB[k] + LA * uA[k] / 2 - LB * uB[k] / 2;
          d2[k] = rAB[k] - LA * uA[k] / 2 + LB * uB[k] / 2;
        }
      }
      dsq = dot3(dAB, dAB);
      dsq0 = dot3(d2, d2);
      s = t = (sg > 0) ? -1.e+10 : -3.e+10;
      if (dsq0 < dsq) {
        for (int k = 0; k < 3; ++k) dAB[k] = d2[k];
        dsq = dsq0;
        s = t = (sg > 0) ? -2.e+10 : -4.e+10;
      }
    }
  }
  st[0] = s;
  st[1] = t;
  return dsq;
}

}  // namespace msc
