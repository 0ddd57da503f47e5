// Host-side cosmology scalars.  The reference computes these with GSL
// (QAG quadrature, 2F1) and libm in a specific mix of float and double; the
// results are stored to float before they reach the particle loops, so the
// float/double promotions below follow the reference expression by
// expression (cola.c, power.c) to reproduce the same float scalars.
#include "cosmo.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <vector>

#include "gk_nodes.h"

namespace cola {

namespace {

constexpr double kPi = 3.1415926535897932384626433832795;

// ---- Gauss-Kronrod rule on one interval (QUADPACK QK41 / QK61) ----------
struct Rule {
  int n;  // number of non-negative Kronrod nodes (21 or 31)
  const double* xgk;
  const double* wgk;
  const double* wg;
};
const Rule kGK41{21, gk41_xgk, gk41_wgk, gk41_wg};
const Rule kGK61{31, gk61_xgk, gk61_wgk, gk61_wg};

struct Estimate {
  double result, abserr, resabs, resasc;
};

double scaled_error(double err, double resabs, double resasc)
{
  err = std::fabs(err);
  if (resasc != 0 && err != 0) {
    const double s = std::pow(200 * err / resasc, 1.5);
    err = s < 1 ? resasc * s : resasc;
  }
  if (resabs > DBL_MIN / (50 * DBL_EPSILON)) err = std::max(err, 50 * DBL_EPSILON * resabs);
  return err;
}

template <class F>
Estimate apply_rule(const Rule& r, const F& f, double a, double b)
{
  double lo[31], hi[31];
  const double c = 0.5 * (a + b);
  const double h = 0.5 * (b - a);
  const double fc = f(c);
  const int n = r.n;
  double gauss = (n % 2 == 0) ? fc * r.wg[n / 2 - 1] : 0.0;
  double kron = fc * r.wgk[n - 1];
  double sabs = std::fabs(kron);
  // Gauss nodes sit at the odd indices of xgk
  for (int j = 0; j < (n - 1) / 2; ++j) {
    const int g = 2 * j + 1;
    const double dx = h * r.xgk[g];
    const double f1 = f(c - dx), f2 = f(c + dx);
    lo[g] = f1;
    hi[g] = f2;
    gauss += r.wg[j] * (f1 + f2);
    kron += r.wgk[g] * (f1 + f2);
    sabs += r.wgk[g] * (std::fabs(f1) + std::fabs(f2));
  }
  for (int j = 0; j < n / 2; ++j) {
    const int k = 2 * j;
    const double dx = h * r.xgk[k];
    const double f1 = f(c - dx), f2 = f(c + dx);
    lo[k] = f1;
    hi[k] = f2;
    kron += r.wgk[k] * (f1 + f2);
    sabs += r.wgk[k] * (std::fabs(f1) + std::fabs(f2));
  }
  const double mean = 0.5 * kron;
  double sasc = r.wgk[n - 1] * std::fabs(fc - mean);
  for (int j = 0; j < n - 1; ++j) sasc += r.wgk[j] * (std::fabs(lo[j] - mean) + std::fabs(hi[j] - mean));
  const double err = (kron - gauss) * h;
  const double ah = std::fabs(h);
  Estimate e;
  e.result = kron * h;
  e.resabs = sabs * ah;
  e.resasc = sasc * ah;
  e.abserr = scaled_error(err, e.resabs, e.resasc);
  return e;
}

// Adaptive bisection of the interval with the largest error estimate until
// the summed error meets max(epsabs, epsrel |I|) (QUADPACK DQAGE).
template <class F>
double qag(const F& f, double a, double b, double epsabs, double epsrel, int limit, int key)
{
  const Rule& rule = key == 6 ? kGK61 : kGK41;
  const Estimate e0 = apply_rule(rule, f, a, b);
  double tol = std::max(epsabs, epsrel * std::fabs(e0.result));
  if (e0.abserr <= 50 * DBL_EPSILON * e0.resabs && e0.abserr > tol) return e0.result;
  if ((e0.abserr <= tol && e0.abserr != e0.resasc) || e0.abserr == 0.0) return e0.result;
  if (limit == 1) return e0.result;

  struct Piece {
    double a, b, r, e;
  };
  std::vector<Piece> pcs;
  pcs.reserve(std::min(limit, 4096));
  pcs.push_back({a, b, e0.result, e0.abserr});
  double area = e0.result, errsum = e0.abserr;
  int it = 1, rt1 = 0, rt2 = 0;
  bool stop = false;
  do {
    size_t w = 0;
    for (size_t i = 1; i < pcs.size(); ++i)
      if (pcs[i].e > pcs[w].e) w = i;
    const Piece p = pcs[w];
    const double mid = 0.5 * (p.a + p.b);
    const Estimate l = apply_rule(rule, f, p.a, mid);
    const Estimate r = apply_rule(rule, f, mid, p.b);
    const double area12 = l.result + r.result;
    const double err12 = l.abserr + r.abserr;
    errsum += err12 - p.e;
    area += area12 - p.r;
    if (l.resasc != l.abserr && r.resasc != r.abserr) {
      if (std::fabs(p.r - area12) <= 1.0e-5 * std::fabs(area12) && err12 >= 0.99 * p.e) ++rt1;
      if (it >= 10 && err12 > p.e) ++rt2;
    }
    tol = std::max(epsabs, epsrel * std::fabs(area));
    if (errsum > tol) {
      if (rt1 >= 6 || rt2 >= 20) stop = true;
      const double tiny = (1 + 100 * DBL_EPSILON) * (std::fabs(mid) + 1000 * DBL_MIN);
      if (std::fabs(p.a) <= tiny && std::fabs(p.b) <= tiny) stop = true;
    }
    // the half with the larger error keeps the slot, the other is appended
    if (r.abserr > l.abserr) {
      pcs[w] = {mid, p.b, r.result, r.abserr};
      pcs.push_back({p.a, mid, l.result, l.abserr});
    } else {
      pcs[w] = {p.a, mid, l.result, l.abserr};
      pcs.push_back({mid, p.b, r.result, r.abserr});
    }
    ++it;
  } while (it < limit && !stop && errsum > tol);
  double sum = 0;
  for (const Piece& q : pcs) sum += q.r;
  return sum;
}

}  // namespace

// 2F1(a,b;c;z), -1 < z <= 0.  Pfaff's transformation
//   2F1(a,b;c;z) = (1-z)^(-a) 2F1(a, c-b; c; z/(z-1))
// maps z to w = z/(z-1) in [0, 1/2), where the Maclaurin series converges at
// least like 2^-n.
double hyp2f1(double a, double b, double c, double z)
{
  const double w = z / (z - 1.0);
  const double bb = c - b;
  long double term = 1.0L, sum = 1.0L;
  for (int n = 0; n < 400; ++n) {
    term *= ((long double)a + n) * ((long double)bb + n) / (((long double)c + n) * (n + 1)) * w;
    sum += term;
    if (fabsl(term) <= 1e-21L * fabsl(sum)) break;
  }
  return (double)(std::pow(1.0L - (long double)z, -(long double)a) * sum);
}

// ---- power.c ----------------------------------------------------------------
void PowerTable::set(const double* k, const double* p, int n)
{
  const double fac = 1.0 / (2.0 * kPi * kPi);
  logk.resize(n);
  logD.resize(n);
  for (int i = 0; i < n; ++i) {
    logk[i] = std::log10(k[i]);
    logD[i] = std::log10(fac * k[i] * k[i] * k[i] * p[i]);
  }
  norm = 1.0;
}

int PowerTable::read(const char* filename)
{
  std::FILE* fp = std::fopen(filename, "r");
  if (!fp) return 3000;
  std::vector<double> k, p;
  double a, b;
  while (std::fscanf(fp, " %lg %lg ", &a, &b) == 2) {
    k.push_back(a);
    p.push_back(b);
  }
  std::fclose(fp);
  set(k.data(), p.data(), (int)k.size());
  return 0;
}

double PowerTable::operator()(double k) const
{
  const double lk = std::log10(k);
  const int n = (int)logk.size();
  if (lk < logk[0] || lk > logk[n - 1]) return 0;
  // bracketing bin by bisection (power.c:127-136)
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (hi + lo) / 2;
    if (lk < logk[mid])
      hi = mid;
    else
      lo = mid;
  }
  const double u = (lk - logk[lo]) / (logk[hi] - logk[lo]);
  const double ld = (1 - u) * logD[lo] + u * logD[hi];
  return norm * std::pow(10.0, ld) / (4.0 * kPi * k * k * k);
}

double PowerTable::sigma8() const
{
  const double R = 8.0;
  auto integrand = [&](double k) {
    const double kr = R * k, kr2 = kr * kr, kr3 = kr2 * kr;
    if (kr < 1e-8) return 0.0;
    const double w = 3 * (std::sin(kr) / kr3 - std::cos(kr) / kr2);
    return 4 * kPi * k * k * w * w * (*this)(k);
  };
  return std::sqrt(qag(integrand, 0, 500.0 * 1 / R, 0, 1.0e-4, 100000, 4));
}

double growth_factor(double astart, double aend, double Om, double OL)
{
  auto growth = [&](double a) {
    const double hubble_a = std::sqrt(Om / (a * a * a) + (1 - Om - OL) / (a * a) + OL);
    auto integrand = [&](double x) { return std::pow(x / (Om + (1 - Om - OL) * x + OL * x * x * x), 1.5); };
    return hubble_a * qag(integrand, 0, a, 0, 1.0e-8, 100000, 4);
  };
  return growth(aend) / growth(astart);
}

// ---- Eisenstein & Hu (1998) no-wiggle spectrum ------------------------------
double eh98_nowiggle_transfer(double k, double omega_m, double omega_b, double h)
{
  const double theta = 2.728 / 2.7;  // T_cmb / 2.7 K
  const double om = omega_m * h * h, ob = omega_b * h * h, fb = omega_b / omega_m;
  const double s = 44.5 * std::log(9.83 / om) / std::sqrt(1.0 + 10.0 * std::pow(ob, 0.75));  // eq. 26, Mpc
  const double alpha = 1.0 - 0.328 * std::log(431.0 * om) * fb + 0.38 * std::log(22.3 * om) * fb * fb;  // eq. 31
  const double kmpc = k * h;
  const double ks = 0.43 * kmpc * s;
  const double gamma = omega_m * h * (alpha + (1.0 - alpha) / (1.0 + ks * ks * ks * ks));  // eq. 30
  const double q = k * theta * theta / gamma;                                                // eq. 28
  const double L0 = std::log(2.0 * std::exp(1.0) + 1.8 * q);
  const double C0 = 14.2 + 731.0 / (1.0 + 62.5 * q);
  return L0 / (L0 + C0 * q * q);  // eq. 29
}

void eh98_nowiggle_table(double omega_m, double omega_b, double h, double ns, double sigma8, double kmin,
                         double kmax, int n, double* k, double* p)
{
  for (int i = 0; i < n; ++i) {
    k[i] = kmin * std::pow(kmax / kmin, n > 1 ? (double)i / (n - 1) : 0.0);
    const double t = eh98_nowiggle_transfer(k[i], omega_m, omega_b, h);
    p[i] = std::pow(k[i], ns) * t * t;
  }
  PowerTable tab;
  tab.set(k, p, n);
  const double s = tab.sigma8();
  const double A = (sigma8 / s) * (sigma8 / s);
  for (int i = 0; i < n; ++i) p[i] *= A;
}

// ---- cola.c -----------------------------------------------------------------
float ColaScalars::growthDtemp(const float a) const
{
  // float x, hyperP, hyperM and the double literals below are the reference's
  float x = -Om / (Om - 1.0) / (a * a * a);
  float P = 0, M = 0;
  if (std::fabs(x - 1.0) < 1.e-3) {
    const double t = -1.0 + x;
    P = 0.859596768064608 - 0.1016599912520404 * t + 0.025791094277821357 * std::pow(t, 2) -
        0.008194025861121475 * std::pow(t, 3) + 0.0029076305993447644 * std::pow(t, 4) -
        0.0011025426387159761 * std::pow(t, 5) + 0.00043707304964624546 * std::pow(t, 6) -
        0.0001788889964687831 * std::pow(t, 7);
    M = 1.1765206505266006 + 0.15846194123099624 * t - 0.014200487494738975 * std::pow(t, 2) +
        0.002801728034399257 * std::pow(t, 3) - 0.0007268267888593511 * std::pow(t, 4) +
        0.00021801569226706922 * std::pow(t, 5) - 0.00007163321597397065 * std::pow(t, 6) +
        0.000025063737576245116 * std::pow(t, 7);
  } else {
    if (x < 1.0) {
      P = hyp2f1(1.0 / 2.0, 2.0 / 3.0, 5.0 / 3.0, -x);
      M = hyp2f1(-1.0 / 2.0, 2.0 / 3.0, 5.0 / 3.0, -x);
    }
    x = 1.0 / x;
    if ((x < 1.0) && (x > 1.0 / 30)) {
      P = hyp2f1(-1.0 / 6.0, 0.5, 5.0 / 6.0, -x);
      P *= 4 * std::sqrt((double)x);
      P += -3.4494794123063873799 * std::pow((double)x, 2.0 / 3.0);
      M = hyp2f1(-7.0 / 6.0, -0.5, -1.0 / 6.0, -x);
      M *= 4.0 / 7.0 / std::sqrt((double)x);
      M += std::pow((double)x, 2.0 / 3.0) * (-1.4783483195598803057);
    }
    if (x <= 1.0 / 30.0) {
      const double xd = x;
      P = 3.9999999999999996 * std::sqrt(xd) - 3.4494794123063865 * std::pow(xd, 0.6666666666666666) +
          0.3999999999999999 * std::pow(xd, 1.5) - 0.13636363636363635 * std::pow(xd, 2.5) +
          0.07352941176470587 * std::pow(xd, 3.5) - 0.04755434782608695 * std::pow(xd, 4.5) +
          0.033943965517241374 * std::pow(xd, 5.5) - 0.02578125 * std::pow(xd, 6.5) +
          0.020436356707317072 * std::pow(xd, 7.5) - 0.01671324384973404 * std::pow(xd, 8.5) +
          0.013997779702240564 * std::pow(xd, 9.5) - 0.011945562847590041 * std::pow(xd, 10.5) +
          0.01035003662109375 * std::pow(xd, 11.5) - 0.009080577904069926 * std::pow(xd, 12.5);
      M = 0.5714285714285715 / std::sqrt(xd) + 2.000000000000001 * std::sqrt(xd) -
          1.4783483195598794 * std::pow(xd, 0.66666666666666666) + 0.10000000000000002 * std::pow(xd, 1.5) -
          0.022727272727272735 * std::pow(xd, 2.5) + 0.009191176470588237 * std::pow(xd, 3.5) -
          0.004755434782608697 * std::pow(xd, 4.5) + 0.002828663793103449 * std::pow(xd, 5.5) -
          0.0018415178571428578 * std::pow(xd, 6.5) + 0.0012772722942073172 * std::pow(xd, 7.5) -
          0.0009285135472074472 * std::pow(xd, 8.5) + 0.0006998889851120285 * std::pow(xd, 9.5) -
          0.0005429801294359111 * std::pow(xd, 10.5) + 0.0004312515258789064 * std::pow(xd, 11.5) -
          0.00034925299631038194 * std::pow(xd, 12.5);
    }
  }
  const double ad = a, om = Om, Pd = P, Md = M;
  if (a > 0.2)
    return std::sqrt(1.0 + (-1.0 + std::pow(ad, -3)) * om) *
           (3.4494794123063873799 * std::pow(-1.0 + 1.0 / om, 0.666666666666666666666666666) +
            (Pd * (4 * std::pow(ad, 3) * (-1.0 + om) - om) - 7.0 * std::pow(ad, 3) * Md * (-1.0 + om)) /
                (std::pow(ad, 5) * (-1.0 + om) - std::pow(ad, 2) * om));
  // small-a series of the reference (cola.c:193); (1 - Om) and (-1 + Om) are float there
  const double omc = (float)(1 - Om), omm = (float)(-1 + Om);
  return (ad * std::pow(omc, 1.5) *
          (1291467969 * std::pow(ad, 12) * std::pow(omm, 4) + 1956769650 * std::pow(ad, 9) * std::pow(omm, 3) * om +
           8000000000LL * std::pow(ad, 3) * omm * std::pow(om, 3) + 37490640625LL * std::pow(om, 4))) /
         (1.5625e10 * std::pow(om, 5));
}

float ColaScalars::growthD(const float a) const { return growthDtemp(a) / growthDtemp(1.0f); }

float ColaScalars::Qfactor(const float a) const
{
  return std::sqrt(Om / (a * a * a) + 1.0 - Om) * a * a * a;
}

float ColaScalars::growthD2(const float a) const
{
  auto d2t = [&](float b) -> float {
    const float d = growthD(b);
    const float omega = Om / (Om + (1.0 - Om) * b * b * b);
    return d * d * std::pow((double)omega, -1.0 / 143.);
  };
  return d2t(a) / d2t(1.0f);
}

float ColaScalars::growthD2v(const float a) const
{
  const float d2 = growthD2(a);
  const float omega = Om / (Om + (1.0 - Om) * a * a * a);
  return Qfactor(a) * (d2 / a) * 2.0 * std::pow((double)omega, 6.0 / 11.);
}

float ColaScalars::decayD(const float a) const { return std::sqrt(Om / (a * a * a) + 1.0 - Om); }

double ColaScalars::DprimeQ(double a, float nGrowth) const
{
  const float nDecay = 0.0f;
  const float Nn = 6.0 * std::pow(1.0 - Om, 1.5) / growthDtemp(1.0f);
  const float af = (float)a;
  return std::pow((double)decayD(af), -1.0 + nDecay) * std::pow((double)growthD(af), -1.0 + nGrowth) *
         (nGrowth * Nn - (3.0 * (nDecay + nGrowth) * Om * growthD(af)) / (2. * a));
}

// cola.c:59-61: nLPT = -2.5f, fullT = 1
static const double kNLPT = (double)-2.5f;

double ColaScalars::Sq(double ai, double af, double aRef) const
{
  auto integrand = [&](double a) { return std::pow(a, kNLPT) / (double)Qfactor((float)a); };
  return qag(integrand, ai, af, 0, 1e-5, 5000, 6) / std::pow(aRef, kNLPT);
}

double ColaScalars::Sphi(double ai, double af, double aRef) const
{
  return (std::pow(af, kNLPT) - std::pow(ai, kNLPT)) * aRef / (double)Qfactor((float)aRef) /
         (kNLPT * std::pow(aRef, kNLPT - 1));
}

KickFactors kick_factors(float Om, float a_x, float a_v, float avel1)
{
  const ColaScalars s(Om);
  const float AI = a_v, A = a_x, AF = avel1;
  const float Om143 = std::pow((double)(Om / (Om + (1 - Om) * A * A * A)), 1.0 / 143.0);
  KickFactors k;
  k.dda = s.Sphi(AI, AF, A);
  const float g = s.growthD(A);
  k.q2 = 1.5 * Om * g * g * (1.0 + 7.0 / 3.0 * Om143);
  k.q1 = 1.5 * Om * g;
  k.m15om = -1.5 * Om;
  return k;
}

DriftFactors drift_factors(float Om, float a_x, float a_v, float apos1)
{
  const ColaScalars s(Om);
  const float A = a_x, AC = a_v, AF = apos1;
  DriftFactors d;
  d.dyyy = s.Sq(A, AF, AC);
  d.da1 = s.growthD(AF) - s.growthD(A);
  d.da2 = s.growthD2(AF) - s.growthD2(A);
  return d;
}

SnapshotFactors snapshot_factors(float Om, float a_x, float a_v, double aout)
{
  const ColaScalars s(Om);
  SnapshotFactors f;
  f.vfac = 100.0f / aout;
  const float AI = a_v, A = a_x, AF = aout;
  const float Om143 = std::pow((double)(Om / (Om + (1 - Om) * A * A * A)), 1.0 / 143.0);
  f.kick.dda = s.Sphi(AI, AF, A);
  const float g = s.growthD(A);
  f.kick.q1 = 1.5 * Om * g;
  f.kick.q2 = 1.5 * Om * g * g * (1.0 + 7.0 / 3.0 * Om143);
  f.kick.m15om = -1.5 * Om;
  f.Dv = s.DprimeQ(aout, 1.0f);
  f.Dv2 = s.growthD2v((float)aout);
  const float AC = a_v;
  f.drift.dyyy = s.Sq(A, AF, AC);
  f.drift.da1 = s.growthD(AF) - s.growthD(A);
  f.drift.da2 = s.growthD2(AF) - s.growthD2(A);
  return f;
}

SnapshotFactors noncola_initial_factors(float Om, float a_x)
{
  const ColaScalars s(Om);
  SnapshotFactors f{};
  const float aout = a_x;
  f.vfac = 100.0f / aout;
  f.Dv = s.DprimeQ(aout, 1.0f);
  f.Dv2 = s.growthD2v(aout);
  return f;
}

}  // namespace cola
