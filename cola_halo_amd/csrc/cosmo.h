// Host-side cosmology scalars of the COLA path: input power spectrum table,
// sigma8 check, linear growth by quadrature (power.c) and the COLA growth /
// time-stepping factors (cola.c).  These are per-step scalars computed once on
// the host and passed to the kernels by value.
#pragma once

#include <vector>

namespace cola {

// power.c:26-29 table of (log10 k, log10 Delta^2), Delta^2 = k^3 P / (2 pi^2)
struct PowerTable {
  std::vector<double> logk, logD;
  double norm = 1.0;
  bool empty() const { return logk.empty(); }
  void set(const double* k, const double* p, int n);   // power.c:62-101
  int read(const char* filename);                       // 0 or 3000
  double operator()(double k) const;                    // power.c:120-151
  double sigma8() const;                                // power.c:103-118, 204-224
};

// Gauss hypergeometric function for -1 < z <= 0 (cola.c:169-179 arguments).
double hyp2f1(double a, double b, double c, double z);

// power.c:172-202: D(aend)/D(astart) with D = H(a) * int_0^a (a/(Om + OL a^3))^1.5
double growth_factor(double astart, double aend, double omega_m, double omega_lambda);

// Eisenstein & Hu 1998 no-wiggle transfer function T0(k), k in h/Mpc.
double eh98_nowiggle_transfer(double k, double omega_m, double omega_b, double h);
// Table of P(k) = A k^ns T0^2 at n log-spaced k, A set so PowerTable::sigma8() == sigma8.
void eh98_nowiggle_table(double omega_m, double omega_b, double h, double ns, double sigma8, double kmin,
                         double kmax, int n, double* k, double* p);

// cola.c:58-292.  Om is cola.c's file-static float (cola.c:58).
struct ColaScalars {
  float Om;
  explicit ColaScalars(float om) : Om(om) {}
  float growthDtemp(float a) const;          // cola.c:156-194
  float growthD(float a) const;              // cola.c:196-198
  float Qfactor(float a) const;              // cola.c:201-203
  float growthD2(float a) const;             // cola.c:208-216
  float growthD2v(float a) const;            // cola.c:219-223
  float decayD(float a) const;               // cola.c:225-227
  double DprimeQ(double a, float nGrowth) const;  // cola.c:229-234
  double Sq(double ai, double af, double aRef) const;    // cola.c:262-281
  double Sphi(double ai, double af, double aRef) const;  // cola.c:287-292
};

// Kernel arguments of cola_kick (cola.c:77-91): v += dda*(-1.5 Om f - (q1 dx1 + q2 dx2))
struct KickFactors {
  float dda, q1, q2;
  double m15om;  // -1.5*Om, applied in double as cola.c:105-107 does
};
// cola_drift (cola.c:123-134): x += v dyyy + (dx1 da1 + dx2 da2)
struct DriftFactors {
  float dyyy, da1, da2;
};
// cola_set_snapshot (cola.c:353-383)
struct SnapshotFactors {
  KickFactors kick;
  DriftFactors drift;
  float vfac, Dv, Dv2;
};

KickFactors kick_factors(float Om, float a_x, float a_v, float avel1);
DriftFactors drift_factors(float Om, float a_x, float a_v, float apos1);
SnapshotFactors snapshot_factors(float Om, float a_x, float a_v, double aout);
// set_noncola_initial (cola.c:301-313)
SnapshotFactors noncola_initial_factors(float Om, float a_x);

}  // namespace cola
