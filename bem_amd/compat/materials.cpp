// Material library of the compat layer: the media the reference's registry
// names (src/util/Dielectrics.cpp:15-63), each model restated from its
// formula; the SiN rows are the reference's measured table (sin_table.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>

#include "bem_compat.h"
#include "sin_table.h"

namespace {
// src/util/SI.cpp
const Real kHbar = 1.05457148e-34, kKb = 1.3806503e-23, kEV = 1.60217646e-19, kMe = 9.109e-31,
           kEps0 = 8.85418782e-12;
Real omega_of(Real eV) { return eV * kEV / kHbar; }

// Interval of the table holding x by bisection (src/util/MathUtils.cpp:351-367):
// k with tab[k] <= x < tab[k+1], clamped to the first and last interval, so
// values outside the table extrapolate linearly.
int search_table(const double (*tab)[3], int n, Real x) {
  int lo = 0;
  while (n > 2) {
    const int k = n / 2;
    if (x >= tab[lo + k][0]) {
      lo += k;
      n -= k;
    } else {
      n = k + 1;
    }
  }
  return (n == 2 && x == tab[lo + 1][0]) ? lo + 1 : lo;
}
Real interpolate(const double (*tab)[3], int n, int col, Real x) {
  // x equal to the last energy selects the last interval (the bisection
  // alone would point past the table)
  const int k = std::min(search_table(tab, n, x), n - 2);
  const Real h = tab[k + 1][0] - tab[k][0];
  const Real slope = (tab[k + 1][col] - tab[k][col]) / h;
  return tab[k][col] + slope * (x - tab[k][0]);
}
}  // namespace

// src/util/SiN.cpp:3352-3363: (n + i k)^2, n and k linear in the table.
Complex SiN::eps(Real eV) {
  const Complex sq(interpolate(kSiN, kSiNRows, 1, eV), interpolate(kSiN, kSiNRows, 2, eV));
  return sq * sq;
}
Complex SiN::mu(Real) { return Complex(1., 0.); }

// src/util/SiC.cpp:4-23
SiC::SiC() { setTemperature(300); }
void SiC::setTemperature(Real T) {
  epsinf = 6.7 * std::exp(5e-5 * (T - 300));
  wL = 182.7e12 - 5.463e9 * (T - 300);
  wT = 149.5e12 - 4.106e9 * (T - 300);
  gamma = 6.6e11 * (1 + 2.0 / (std::exp(kHbar * wT / (2 * kKb * T)) - 1));
}
Complex SiC::eps(Real eV) {
  const Real w = omega_of(eV);
  return epsinf * (wL * wL - w * w - Complex(0, gamma * w)) / (wT * wT - w * w - Complex(0, gamma * w));
}
Complex SiC::mu(Real) { return Complex(1., 0.); }

// src/util/Al.cpp:4-16
Complex Al::eps(Real eV) {
  const Real epsInf = 2.2281, epsLorentz = 1.30601, omega0 = 1.50690679e14, gamma0 = 3.76734192e13;
  const Real w = omega_of(eV);
  return epsInf + epsLorentz * omega0 * omega0 / (omega0 * omega0 - w * w - Complex(0, 2.0 * w * gamma0));
}
Complex Al::mu(Real) { return Complex(1., 0.); }

// src/util/Au.cpp:4-15
Complex Au::eps(Real eV) {
  const Real eps0 = 1.0, omegap = 1.366592804e16, omegac = 4.08407045e13;
  const Real w = omega_of(eV);
  return eps0 - omegap * omegap / (w * (w + Complex(0, omegac)));
}
Complex Au::mu(Real) { return Complex(1., 0.); }

// src/util/SiliconDoped.cpp:6-44: Drude term with the carrier mobility of
// the Masetti model, n-type for N < 0, p-type for N > 0, 11.7 undoped.
Complex SiliconDoped::eps(Real eV) {
  const Real w = omega_of(eV);
  if (N == 0.0) return 11.7;
  Real mu, m;
  if (N < 0.0) {
    const Real mu1 = .00685, mumax = .1414, mu2 = .00561, Cr = 9.20e22, Cs = 3.41e26, alpha = .711, beta = 1.98;
    m = .27 * kMe;
    mu = mu1 + (mumax - mu1) / (1 + std::pow(std::fabs(N) / Cr, alpha)) - mu2 / (1 + std::pow(Cs / std::fabs(N), beta));
  } else {
    const Real mu1 = .00449, mumax = .04705, mu2 = .0029, Cr = 2.23e23, Cs = 6.10e26, alpha = .719, beta = 2.00,
               pc = 9.23e24;
    m = .37 * kMe;
    mu = mu1 * std::exp(-pc / N) + mumax / (1 + std::pow(N / Cr, alpha)) - mu2 / (1 + std::pow(Cs / N, beta));
  }
  const Real omegap = kEV * std::sqrt(std::fabs(N) / m / kEps0);
  const Real gamma = kEV / m / mu;
  return (Real)11.7 - omegap * omegap / (w * (w + Complex(0., gamma)));
}
Complex SiliconDoped::mu(Real) { return Complex(1., 0.); }

const std::string *Params::find(const std::string &k) const {
  for (auto &kv : params)
    if (kv.first == k) return &kv.second;
  return nullptr;
}

Complex parseComplex(const std::string &s) {
  const char *p = s.c_str();
  char *e;
  const Real re = std::strtod(p, &e);
  if (e == p) throw bem::Error(BEM_E_INVALID, "not a complex number: '" + s + "'");
  while (*e == ' ') e++;
  if (*e == 0) return Complex(re, 0.0);
  if (*e != '+' && *e != '-') throw bem::Error(BEM_E_INVALID, "not a complex number: '" + s + "'");
  const char *q = e;
  const Real im = std::strtod(q, &e);
  if (e == q) throw bem::Error(BEM_E_INVALID, "not a complex number: '" + s + "'");
  if (*e == 'i' || *e == 'j') e++;
  while (*e == ' ') e++;
  if (*e) throw bem::Error(BEM_E_INVALID, "not a complex number: '" + s + "'");
  return Complex(re, im);
}

// src/util/Dielectrics.cpp:21-63: one instance per (name, parameters).
Dielectric *Dielectrics::getDielectric(const std::string &name, const Params &params) {
  static std::mutex mu;
  static std::map<std::pair<std::string, std::vector<std::pair<std::string, std::string>>>,
                  std::unique_ptr<Dielectric>>
      registry;
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair(name, params.params);
  auto it = registry.find(key);
  if (it != registry.end()) return it->second.get();
  std::unique_ptr<Dielectric> d;
  if (params.params.empty() && name == "SiN") d.reset(new SiN());
  else if (params.params.empty() && name == "SiC") d.reset(new SiC());
  else if (params.params.empty() && name == "Al") d.reset(new Al());
  else if (params.params.empty() && name == "Au") d.reset(new Au());
  else if (name == "SiliconDoped") {
    const std::string *n = params.find("N");
    d.reset(new SiliconDoped(n ? std::strtod(n->c_str(), nullptr) : 0.0));
  } else {
    const std::string *e = params.find("eps"), *m = params.find("mu");
    d.reset(new ConstantDielectric(e ? parseComplex(*e) : Complex(1.0, 0.0), m ? parseComplex(*m) : Complex(1.0, 0.0)));
  }
  Dielectric *out = d.get();
  registry.emplace(key, std::move(d));
  return out;
}
