// trees.cpp -- merger-tree forests on the host (C ABI of include/gal_tree.h):
// cosmology and sigma(M), a Parkinson-Cole-Helly forest builder, and the
// reference's tree walk recorded as a schedule of node-evolution intervals
// and events, split into waves for the GPU.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <atomic>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "../../include/gal_tree.h"
#include "constants.hpp"
#include "schedule.hpp"

// Host arithmetic is evaluated as written (no fused multiply-add), so that
// times and halo parameters are reproducible by any IEEE double evaluation of
// the same expressions.
#pragma clang fp contract(off)

namespace {

using gal::kGravity;
using gal::kMpcPerKmsToGyr;
using gal::kPi;

constexpr double kInf = std::numeric_limits<double>::infinity();

// ---- cosmology ----------------------------------------------------------------
// Flat LCDM: a(t) = (Om/OL)^(1/3) sinh^(2/3)(1.5 sqrt(OL) H0 t); linear growth
// D(a) = 5/2 Om E(a) int_0^a da'/(a' E(a'))^3 normalised to D(1) = 1;
// delta_c(t) = (3/20)(12 pi)^(2/3) (1 + 0.0123 log10 Om(a)) / D(a) (spherical
// collapse, Navarro, Frenk & White 1997 fit); sigma(M) from the Eisenstein & Hu
// (1998) zero-baryon-wiggle transfer function with a top-hat window.
struct Cosmology {
  double om, ol, ob, h0, s8, ns, tcmb;
  double h0_gyr;     // H0 in 1/Gyr
  double rho_mean0;  // Msun / Mpc^3
  double amp = 1.0;  // P(k) normalisation
  std::vector<double> ln_m, ln_s, dln_s;   // S = sigma^2 on a ln M grid
  std::vector<double> ln_a_tab, growth_tab;
  std::vector<double> t_tab, delta_tab;    // delta_c on a time grid (decreasing)

  explicit Cosmology(const gal_cosmology& c)
      : om(c.omega_matter), ol(c.omega_dark_energy), ob(c.omega_baryon), h0(c.hubble_constant), s8(c.sigma_8),
        ns(c.index_spectral), tcmb(c.temperature_cmb > 0.0 ? c.temperature_cmb : 2.7255) {
    h0_gyr = h0 / kMpcPerKmsToGyr;
    rho_mean0 = om * 3.0 * h0 * h0 / (8.0 * kPi * kGravity);
    build_growth();
    build_delta();
    build_sigma();
  }

  double e2(double a) const { return om / (a * a * a) + ol; }
  double a_of_t(double t) const { return std::cbrt(om / ol) * std::pow(std::sinh(1.5 * std::sqrt(ol) * h0_gyr * t), 2.0 / 3.0); }
  double t_of_a(double a) const { return std::asinh(std::pow(a / std::cbrt(om / ol), 1.5)) / (1.5 * std::sqrt(ol) * h0_gyr); }
  double hubble_time(double t) const { return 1.0 / (h0_gyr * std::sqrt(e2(a_of_t(t)))); }

  void build_growth() {
    // cumulative integral of 1/(a E)^3 on a fine ln a grid (trapezoid in ln a)
    const int n = 4000;
    const double la0 = std::log(1.0e-5), la1 = std::log(2.0);
    std::vector<double> la(n + 1), cum(n + 1, 0.0);
    auto f = [&](double lna) {
      const double a = std::exp(lna);
      const double ae = a * std::sqrt(e2(a));
      return a / (ae * ae * ae);  // da = a dln a
    };
    for (int i = 0; i <= n; ++i) la[i] = la0 + (la1 - la0) * i / n;
    // analytic start: for a << 1, int_0^a da/(a E)^3 = a^(5/2) / (2.5 Om^(3/2))
    cum[0] = std::pow(std::exp(la0), 2.5) / (2.5 * std::pow(om, 1.5));
    for (int i = 1; i <= n; ++i) cum[i] = cum[i - 1] + 0.5 * (f(la[i - 1]) + f(la[i])) * (la[i] - la[i - 1]);
    ln_a_tab = la;
    growth_tab.resize(n + 1);
    for (int i = 0; i <= n; ++i) {
      const double a = std::exp(la[i]);
      growth_tab[i] = 2.5 * om * std::sqrt(e2(a)) * cum[i];
    }
    const double d1 = growth(1.0);
    for (double& g : growth_tab) g /= d1;
  }
  double growth(double a) const {
    const double la = std::log(a);
    const int n = (int)ln_a_tab.size() - 1;
    double x = (la - ln_a_tab[0]) / (ln_a_tab[n] - ln_a_tab[0]) * n;
    int i = std::min(std::max((int)x, 0), n - 1);
    const double w = x - i;
    return growth_tab[i] * (1.0 - w) + growth_tab[i + 1] * w;
  }
  double delta_c_exact(double t) const {
    const double a = a_of_t(t);
    const double om_a = om / (a * a * a) / e2(a);
    return 0.15 * std::pow(12.0 * kPi, 2.0 / 3.0) * (1.0 + 0.0123 * std::log10(om_a)) / growth(a);
  }
  void build_delta() {
    const int n = 4000;
    t_tab.resize(n + 1);
    delta_tab.resize(n + 1);
    const double l0 = std::log(0.01), l1 = std::log(30.0);
    for (int i = 0; i <= n; ++i) {
      t_tab[i] = std::exp(l0 + (l1 - l0) * i / n);
      delta_tab[i] = delta_c_exact(t_tab[i]);
    }
  }
  double delta_c(double t) const { return delta_c_exact(t); }
  double time_of_delta(double d) const {
    // delta decreases with t: bisect the table, then refine by secant steps
    int lo = 0, hi = (int)t_tab.size() - 1;
    if (d >= delta_tab[lo]) return t_tab[lo];
    if (d <= delta_tab[hi]) return t_tab[hi];
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (delta_tab[mid] > d) lo = mid; else hi = mid;
    }
    double ta = t_tab[lo], tb = t_tab[hi], da = delta_tab[lo] - d, db = delta_tab[hi] - d;
    for (int it = 0; it < 60; ++it) {
      const double tm = 0.5 * (ta + tb);
      const double dm = delta_c_exact(tm) - d;
      if (dm == 0.0) return tm;
      if ((dm > 0.0) == (da > 0.0)) { ta = tm; da = dm; } else { tb = tm; db = dm; }
      if (tb - ta < 1.0e-13 * tb) break;
    }
    (void)db;
    return 0.5 * (ta + tb);
  }

  double transfer(double k) const {  // k in 1/Mpc
    const double h = h0 / 100.0, theta = tcmb / 2.7;
    const double om_h2 = om * h * h, ob_h2 = ob * h * h, fb = ob / om;
    const double s = 44.5 * std::log(9.83 / om_h2) / std::sqrt(1.0 + 10.0 * std::pow(ob_h2, 0.75));
    const double alpha = 1.0 - 0.328 * std::log(431.0 * om_h2) * fb + 0.38 * std::log(22.3 * om_h2) * fb * fb;
    const double ks = 0.43 * k * s;
    const double gamma = om * h * (alpha + (1.0 - alpha) / (1.0 + ks * ks * ks * ks));
    const double q = k / h * theta * theta / gamma;
    const double l0 = std::log(2.0 * std::exp(1.0) + 1.8 * q);
    const double c0 = 14.2 + 731.0 / (1.0 + 62.5 * q);
    return l0 / (l0 + c0 * q * q);
  }
  double sigma2_radius(double r) const {  // unnormalised: int k^3 P W^2 dln k / 2 pi^2
    const double lk0 = std::log(1.0e-6), lk1 = std::log(300.0 / r);
    const int n = 6000;  // even (Simpson)
    const double dl = (lk1 - lk0) / n;
    double sum = 0.0;
    for (int i = 0; i <= n; ++i) {
      const double k = std::exp(lk0 + dl * i);
      const double x = k * r;
      double w;
      if (x < 1.0e-3) w = 1.0 - x * x / 10.0;
      else w = 3.0 * (std::sin(x) - x * std::cos(x)) / (x * x * x);
      const double t = transfer(k);
      const double f = k * k * k * std::pow(k, ns) * t * t * w * w;
      sum += f * ((i == 0 || i == n) ? 1.0 : ((i & 1) ? 4.0 : 2.0));
    }
    return sum * dl / 3.0 / (2.0 * kPi * kPi);
  }
  void build_sigma() {
    const double r8 = 8.0 / (h0 / 100.0);
    amp = s8 * s8 / sigma2_radius(r8);
    const int n = 500;
    const double l0 = std::log(1.0e-4), l1 = std::log(1.0e17);
    ln_m.resize(n + 1);
    ln_s.resize(n + 1);
    dln_s.resize(n + 1);
    for (int i = 0; i <= n; ++i) {
      ln_m[i] = l0 + (l1 - l0) * i / n;
      const double r = std::cbrt(3.0 * std::exp(ln_m[i]) / (4.0 * kPi * rho_mean0));
      ln_s[i] = std::log(amp * sigma2_radius(r));
    }
    for (int i = 0; i <= n; ++i) {
      const int a = std::max(i - 1, 0), b = std::min(i + 1, n);
      dln_s[i] = (ln_s[b] - ln_s[a]) / (ln_m[b] - ln_m[a]);
    }
  }
  // S(M) = sigma^2 and dS/dln M (S times dln S/dln M), interpolated in ln M
  void variance(double m, double& s, double& ds) const {
    const int n = (int)ln_m.size() - 1;
    const double x = (std::log(m) - ln_m[0]) / (ln_m[n] - ln_m[0]) * n;
    const int i = std::min(std::max((int)x, 0), n - 1);
    const double w = x - i;
    const double ls = ln_s[i] * (1.0 - w) + ln_s[i + 1] * w;
    const double dls = dln_s[i] * (1.0 - w) + dln_s[i + 1] * w;
    s = std::exp(ls);
    ds = s * dls;
  }
  double sigma(double m) const {
    double s, ds;
    variance(m, s, ds);
    return std::sqrt(s);
  }
};

// Halo virial radius [Mpc] and velocity [km/s] (Bryan & Norman 1998 contrast,
// the same definition the node RHS uses).
void virial(const Cosmology& c, double mass, double t, double& r, double& v) {
  const double a = c.a_of_t(t);
  const double e2 = c.e2(a);
  const double om_a = c.om / (a * a * a) / e2;
  const double x = om_a - 1.0;
  const double delta_c = 18.0 * kPi * kPi + 82.0 * x - 39.0 * x * x;
  const double rho_crit = 3.0 * c.h0 * c.h0 * e2 / (8.0 * kPi * kGravity);
  r = std::cbrt(3.0 * mass / (4.0 * kPi * delta_c * rho_crit));
  v = std::sqrt(kGravity * mass / r);
}

// ---- Parkinson, Cole & Helly (2008) branching -----------------------------
// Progenitor rate per unit q = M1/M2 per unit delta step (extended
// Press-Schechter with the PCH factor G):
//   R(q) = (1/sqrt(2 pi)) |dS1/dln M1| / q^2 / (S1 - S2)^(3/2)
//          * G0 (sigma1/sigma2)^gamma1 (delta2/sigma2)^gamma2
// The split probability per step is dw * int_{q_res}^{1/2} R dq and the
// sub-resolution accretion fraction dw * int_0^{q_res} q R dq.  The factor
// (delta2/sigma2)^gamma2 is taken out of both integrals, which are then
// functions of M2 alone and are tabulated once per forest.
struct Branching {
  const Cosmology& c;
  gal_pch_params p;
  std::vector<double> ln_m2, i_split, i_accrete;
  double ln_m_lo = 0.0, ln_m_hi = 0.0;

  Branching(const Cosmology& cos, const gal_pch_params& pp, double m_max) : c(cos), p(pp) {
    const int n = 400;
    ln_m_lo = std::log(p.mass_resolution);
    ln_m_hi = std::log(std::max(m_max, 4.0 * p.mass_resolution) * 1.01);
    ln_m2.resize(n + 1);
    i_split.resize(n + 1);
    i_accrete.resize(n + 1);
    for (int i = 0; i <= n; ++i) {
      ln_m2[i] = ln_m_lo + (ln_m_hi - ln_m_lo) * i / n;
      integrals(std::exp(ln_m2[i]), i_split[i], i_accrete[i]);
    }
  }
  // R(q) without the (delta2/sigma2)^gamma2 factor
  double rate(double q, double m2, double s2) const {
    double s1, ds1;
    c.variance(q * m2, s1, ds1);
    const double d = s1 - s2;
    if (!(d > 0.0)) return 0.0;
    return (1.0 / std::sqrt(2.0 * kPi)) * std::fabs(ds1) / (q * q) / (d * std::sqrt(d)) * p.g0 *
           std::pow(std::sqrt(s1 / s2), p.gamma1);
  }
  void integrals(double m2, double& split, double& accrete) const {
    double s2, ds2;
    c.variance(m2, s2, ds2);
    const double q_res = p.mass_resolution / m2;
    split = 0.0;
    if (q_res < 0.5) {  // Simpson in ln q
      const int n = 400;
      const double l0 = std::log(q_res), l1 = std::log(0.5), dl = (l1 - l0) / n;
      for (int i = 0; i <= n; ++i) {
        const double q = std::exp(l0 + dl * i);
        split += q * rate(q, m2, s2) * ((i == 0 || i == n) ? 1.0 : ((i & 1) ? 4.0 : 2.0));
      }
      split *= dl / 3.0;
    }
    accrete = 0.0;
    {
      const double q_min = std::exp(c.ln_m[0] + 1.0e-9) / m2;
      const double q_top = std::min(q_res, 0.5);
      if (q_min < q_top) {
        const int n = 800;
        const double l0 = std::log(q_min), l1 = std::log(q_top), dl = (l1 - l0) / n;
        for (int i = 0; i <= n; ++i) {
          const double q = std::exp(l0 + dl * i);
          accrete += q * q * rate(q, m2, s2) * ((i == 0 || i == n) ? 1.0 : ((i & 1) ? 4.0 : 2.0));
        }
        accrete *= dl / 3.0;
      }
    }
  }
  void tabulated(double m2, double& split, double& accrete) const {
    const int n = (int)ln_m2.size() - 1;
    const double x = (std::log(m2) - ln_m_lo) / (ln_m_hi - ln_m_lo) * n;
    const int i = std::min(std::max((int)x, 0), n - 1);
    const double w = std::min(std::max(x - i, 0.0), 1.0);
    split = i_split[i] * (1.0 - w) + i_split[i + 1] * w;
    accrete = i_accrete[i] * (1.0 - w) + i_accrete[i + 1] * w;
  }
  // Draw q in [q_res, 1/2] from R(q) by inversion of its cumulative integral.
  double draw_q(double m2, double u) const {
    double s2, ds2;
    c.variance(m2, s2, ds2);
    const double q_res = p.mass_resolution / m2;
    const int n = 256;
    const double l0 = std::log(q_res), l1 = std::log(0.5), dl = (l1 - l0) / n;
    double cum[n + 1];
    double prev = std::exp(l0) * rate(std::exp(l0), m2, s2);
    cum[0] = 0.0;
    for (int i = 1; i <= n; ++i) {
      const double q = std::exp(l0 + dl * i);
      const double f = q * rate(q, m2, s2);
      cum[i] = cum[i - 1] + 0.5 * (prev + f) * dl;
      prev = f;
    }
    const double target = u * cum[n];
    int i = 0;
    while (i < n - 1 && cum[i + 1] < target) ++i;
    const double span = cum[i + 1] - cum[i];
    const double w = span > 0.0 ? (target - cum[i]) / span : 0.5;
    return std::exp(l0 + dl * (i + w));
  }
};

}  // namespace

struct gal_forest {
  int n_trees = 0;
  std::vector<int64_t> tree_offset;
  std::vector<int32_t> parent, first_child, sibling;
  std::vector<double> mass, time;
};

namespace {

// Build one tree (Cole2000 construction with the PCH rates) into f.
void build_tree(const Cosmology& c, const Branching& br, double m0, double t0, std::mt19937_64& rng,
                gal_forest& f) {
  std::uniform_real_distribution<double> uni(0.0, 1.0);
  const gal_pch_params& p = br.p;
  const double delta_earliest = c.delta_c(p.time_earliest);
  struct Node { int32_t parent, first_child, sibling; double mass, delta; };
  std::vector<Node> nodes;
  nodes.push_back({-1, -1, -1, m0, c.delta_c(t0)});
  std::vector<int> branches{0};
  while (!branches.empty()) {
    int cur = branches.back();
    branches.pop_back();
    while (cur >= 0) {
      const double m_node = nodes[cur].mass;
      double acc_cum = 0.0, m_b = m_node, d_b = nodes[cur].delta;
      int next = -1;
      for (;;) {
        if (m_b <= p.mass_resolution || d_b >= delta_earliest * (1.0 - 1.0e-12)) {
          if (acc_cum > 0.0) {
            nodes.push_back({cur, -1, -1, m_node * (1.0 - acc_cum), d_b});
            nodes[cur].first_child = (int)nodes.size() - 1;
          }
          next = -1;
          break;
        }
        double s2, ds2;
        c.variance(m_b, s2, ds2);
        double i_split, i_acc;
        br.tabulated(m_b, i_split, i_acc);
        const double g2 = std::pow(d_b / std::sqrt(s2), p.gamma2);
        const double rate_split = i_split * g2, rate_acc = i_acc * g2;
        double s_half, ds_half;
        c.variance(0.5 * m_b, s_half, ds_half);
        double dw = p.epsilon1 * std::sqrt(2.0 * std::max(s_half - s2, 0.0));
        if (rate_split > 0.0) dw = std::min(dw, p.merge_probability / rate_split);
        bool snap_acc = false, snap_early = false;
        if (rate_acc > 0.0 && (p.accretion_limit - acc_cum) / rate_acc <= dw) {
          dw = (p.accretion_limit - acc_cum) / rate_acc;
          snap_acc = true;
        }
        if (delta_earliest - d_b <= dw) {
          dw = delta_earliest - d_b;
          snap_early = true;
          snap_acc = false;
        }
        const bool branch = rate_split > 0.0 && uni(rng) < dw * rate_split;
        const double d_new = snap_early ? delta_earliest : d_b + dw;
        acc_cum = snap_acc ? p.accretion_limit : acc_cum + dw * rate_acc;
        if (branch) {
          const double q = br.draw_q(m_b, uni(rng));
          double m1 = q * m_b, m2 = m_node - m1;
          m1 *= 1.0 - acc_cum;
          m2 *= 1.0 - acc_cum;
          const int i1 = (int)nodes.size();
          nodes.push_back({cur, -1, -1, m1, d_new});
          if (m2 > p.mass_resolution) {
            const int i2 = (int)nodes.size();
            nodes.push_back({cur, -1, -1, m2, d_new});
            const int big = m2 > m1 ? i2 : i1, small = m2 > m1 ? i1 : i2;
            nodes[cur].first_child = big;
            nodes[big].sibling = small;
            branches.push_back(small);
            next = big;
          } else {
            nodes[cur].first_child = i1;
            next = i1;
          }
          break;
        }
        if (acc_cum >= p.accretion_limit) {
          nodes.push_back({cur, -1, -1, m_node * (1.0 - acc_cum), d_new});
          nodes[cur].first_child = (int)nodes.size() - 1;
          next = (int)nodes.size() - 1;
          break;
        }
        m_b = m_node * (1.0 - acc_cum);
        d_b = d_new;
      }
      cur = next;
    }
  }
  const int64_t base = (int64_t)f.parent.size();
  for (const Node& n : nodes) {
    f.parent.push_back(n.parent < 0 ? -1 : (int32_t)(base + n.parent));
    f.first_child.push_back(n.first_child < 0 ? -1 : (int32_t)(base + n.first_child));
    f.sibling.push_back(n.sibling < 0 ? -1 : (int32_t)(base + n.sibling));
    f.mass.push_back(n.mass);
    f.time.push_back(n.parent < 0 ? t0 : c.time_of_delta(n.delta));
  }
  f.tree_offset.push_back((int64_t)f.parent.size());
  f.n_trees += 1;
}

// ---- the tree walk ------------------------------------------------------------
struct Walk {
  const gal_forest& f;
  const Cosmology& c;
  gal_walk_params w;
  std::vector<gal_schedule_op>& ops;
  int64_t lo, hi;  // this tree's nodes
  std::vector<int32_t> parent, first_child, sibling, first_sat, last_sat;
  std::vector<char> is_sat, alive;
  std::vector<double> time, merge_time;
  std::vector<double> par;  // node parameters (5 per node) of the node's current halo
  int32_t root;

  Walk(const gal_forest& ff, const Cosmology& cc, const gal_walk_params& ww, std::vector<gal_schedule_op>& o,
       int64_t l, int64_t h)
      : f(ff), c(cc), w(ww), ops(o), lo(l), hi(h) {
    const int64_t n = hi - lo;
    parent.resize(n); first_child.resize(n); sibling.resize(n); first_sat.assign(n, -1); last_sat.assign(n, -1);
    is_sat.assign(n, 0); alive.assign(n, 1); time.resize(n); merge_time.assign(n, kInf); par.assign((size_t)GAL_NPARAM_NODE * n, 0.0);
    for (int64_t i = 0; i < n; ++i) {
      auto loc = [&](int32_t g) { return g < 0 ? -1 : (int32_t)(g - lo); };
      parent[i] = loc(f.parent[lo + i]);
      first_child[i] = loc(f.first_child[lo + i]);
      sibling[i] = loc(f.sibling[lo + i]);
      time[i] = f.time[lo + i];
    }
    root = 0;
    for (int64_t i = 0; i < n; ++i) {
      // halo of node i until it reaches its parent: the primary progenitor
      // grows linearly to the parent's mass (and spin), others keep theirs
      const double m = f.mass[lo + i], t = f.time[lo + i];
      double r, v;
      virial(c, m, t, r, v);
      const double j = w.spin_parameter * m * v * r;
      double m_rate = 0.0, j_rate = 0.0;
      const int32_t p = parent[i];
      if (p >= 0 && first_child[p] == (int32_t)i && f.time[lo + p] > t) {
        const double mp = f.mass[lo + p], tp = f.time[lo + p];
        double rp, vp;
        virial(c, mp, tp, rp, vp);
        const double jp = w.spin_parameter * mp * vp * rp;
        m_rate = (mp - m) / (tp - t);
        j_rate = (jp - j) / (tp - t);
      }
      double* q = &par[(size_t)GAL_NPARAM_NODE * i];
      q[GAL_NP_TIME_REFERENCE] = t;
      q[GAL_NP_MASS_REFERENCE] = m;
      q[GAL_NP_MASS_RATE] = m_rate;
      q[GAL_NP_SPIN_RATE] = j_rate;
      q[GAL_NP_IS_SATELLITE] = 0.0;
    }
  }

  int32_t glob(int32_t i) const { return i < 0 ? -1 : (int32_t)(lo + i); }
  bool primary(int32_t n) const { return parent[n] >= 0 && first_child[parent[n]] == n; }
  double halo_mass(int32_t n, double t) const {
    const double* q = &par[(size_t)GAL_NPARAM_NODE * n];
    return q[GAL_NP_MASS_REFERENCE] + q[GAL_NP_MASS_RATE] * (t - q[GAL_NP_TIME_REFERENCE]);
  }
  int32_t descend(int32_t n) const {
    while (first_sat[n] >= 0 || first_child[n] >= 0) n = first_sat[n] >= 0 ? first_sat[n] : first_child[n];
    return n;
  }
  // mergerTreeWalkerAllNodes%next from node n (walkers/all_nodes/_class.F90:101-170)
  int32_t walk_next(int32_t n) const {
    if (n == root) return -1;
    if (sibling[n] >= 0) return descend(sibling[n]);
    if (is_sat[n]) {
      const int32_t p = parent[n];
      if (first_child[p] >= 0) return descend(first_child[p]);
      return p;
    }
    return parent[n];
  }
  // Satellite lists are singly linked through sibling[] in arrival order,
  // with the tail kept so that appending (and handing a whole list to a new
  // host, as a promotion does for thousands of satellites) costs no walk.
  void append_satellite(int32_t host, int32_t s) {
    sibling[s] = -1;
    parent[s] = host;
    if (first_sat[host] < 0) first_sat[host] = s;
    else sibling[last_sat[host]] = s;
    last_sat[host] = s;
  }
  void remove_satellite(int32_t host, int32_t s) {
    if (first_sat[host] == s) {
      first_sat[host] = sibling[s];
      if (first_sat[host] < 0) last_sat[host] = -1;
      sibling[s] = -1;
      return;
    }
    int32_t k = first_sat[host];
    while (k >= 0 && sibling[k] != s) k = sibling[k];
    if (k >= 0) {
      sibling[k] = sibling[s];
      if (last_sat[host] == s) last_sat[host] = k;
    }
    sibling[s] = -1;
  }
  // Every satellite of `from` joins the end of `to`'s list, in order.
  void move_satellites(int32_t from, int32_t to) {
    const int32_t head = first_sat[from];
    if (head < 0) return;
    for (int32_t s = head; s >= 0; s = sibling[s]) parent[s] = to;
    if (first_sat[to] < 0) first_sat[to] = head;
    else sibling[last_sat[to]] = head;
    last_sat[to] = last_sat[from];
    first_sat[from] = last_sat[from] = -1;
  }
  void emit(int kind, int32_t a, int32_t b, double t0, double t1, const double* np) {
    gal_schedule_op op{};
    op.kind = kind;
    op.a = glob(a);
    op.b = glob(b);
    op.wave = 0;
    op.t0 = t0;
    op.t1 = t1;
    if (np) for (int k = 0; k < GAL_NPARAM_NODE; ++k) op.node_par[k] = np[k];
    ops.push_back(op);
  }
  // Jiang, Jing, Faltenbacher, Lin & Li (2008) merging time
  // (satellites/merging/timescale/Jiang2008.F90) on a fixed orbit.
  double merging_time(int32_t host, int32_t sat, double t) const {
    const double mh = halo_mass(host, t), ms = halo_mass(sat, t);
    const double q = mh / ms;
    if (!(q > 0.0)) return 0.0;
    double r, v;
    virial(c, mh, t, r, v);
    const double t_dyn = kMpcPerKmsToGyr * r / v;
    const double cc = 0.43, a = 0.94, b = 0.60, d = 0.60;
    return w.merging_multiplier * t_dyn * std::sqrt(w.orbit_radius_ratio) *
           ((a * std::pow(w.orbit_circularity, b) + d) / 2.0 / cc) * q / std::log(1.0 + q);
  }
  bool evolvable(int32_t n, double final_time) const {
    return parent[n] >= 0 && first_child[n] < 0 && time[n] < w.time_end &&
           (time[n] < final_time || (is_sat[n] && time[n] <= final_time));
  }
  // standardTimeEvolveTo (merger_trees/evolver/standard.F90:762-1035) for the
  // standard timestepper (simple + satellite) of evolve/timesteps/standard.F90.
  double evolve_to(int32_t n, bool& merge_task) const {
    merge_task = false;
    const double t = time[n];
    double to = w.time_end;
    const double step_simple = std::min(w.timestep_relative * c.hubble_time(t), w.timestep_absolute) + t;
    double step_sat = kInf;
    bool task_sat = false;
    if (is_sat[n] && merge_time[n] < kInf) {
      const double tm = merge_time[n];
      const double off = std::min(w.satellite_offset_absolute, w.satellite_offset_relative * tm);
      const int32_t host = parent[n];
      if (time[host] < tm - off && parent[host] >= 0) {
        step_sat = std::max(tm - t - 0.5 * off, 0.0) + t;
      } else {
        step_sat = tm;
        task_sat = first_child[host] < 0;
      }
    }
    const bool sat_wins = step_sat <= step_simple;
    const double step = sat_wins ? step_sat : step_simple;
    if (step <= to) {
      to = step;
      merge_task = sat_wins && task_sat;
    }
    if (to == t) return to;
    if (!is_sat[n]) {
      if (parent[n] >= 0 && time[parent[n]] < to) to = time[parent[n]];
    } else {
      const int32_t host = parent[n];
      const double th = parent[host] >= 0 ? time[host] : std::max(time[host], t);
      double limit;
      if (first_child[host] >= 0) {
        limit = std::max(th, t);
      } else {
        const double step_host = std::min(w.timestep_host_relative * c.hubble_time(th), w.timestep_host_absolute);
        limit = std::max(th + step_host, t);
      }
      if (limit < to) to = limit;
    }
    if (to == t) return to;
    if (primary(n)) {
      for (int32_t s = sibling[n]; s >= 0; s = sibling[s])
        if (std::max(t, time[s]) < to) to = std::max(t, time[s]);
    }
    if (to == t) return to;
    if (t < to) {
      for (int32_t s = first_sat[n]; s >= 0; s = sibling[s]) {
        if (time[s] < to) {
          to = std::max(time[s], t);
          if (to == t) break;
        }
      }
    }
    if (to == t) return to;
    for (int32_t s = first_sat[n]; s >= 0; s = sibling[s])  // mergees: satellites merging into n
      if (merge_time[s] < kInf && std::max(merge_time[s], t) < to) to = std::max(merge_time[s], t);
    return to;
  }

  // Returns false if a walk finds evolvable nodes but none can move forward
  // (the reference reports such a tree as deadlocked).
  bool run() {
    const double final_time = time[root];
    bool did_evolve = true;
    while (did_evolve) {
      did_evolve = false;
      bool changed = false;
      const size_t n_ops = ops.size();
      int32_t node = descend(root);
      while (node >= 0) {
        const int32_t parent_next = parent[node], sibling_next = sibling[node];
        int32_t next = -1;
        bool next_found = false;
        bool exists = true;
        if (evolvable(node, final_time)) {
          did_evolve = true;
          bool progressed = false;
          bool merge_task;
          const double t_end = evolve_to(node, merge_task);
          if (t_end > time[node]) {
            progressed = true;
            emit(GAL_OP_EVOLVE, node, -1, time[node], t_end, &par[(size_t)GAL_NPARAM_NODE * node]);
            time[node] = t_end;
          }
          if (merge_task) {
            // satelliteMergerProcess (evolve/timesteps/satellite.F90:210-278)
            const int32_t host = parent[node];
            emit(GAL_OP_MERGE_GALAXY, node, host, time[node], time[node], nullptr);
            move_satellites(node, host);
            remove_satellite(host, node);
            alive[node] = 0;
            exists = false;
          }
          if (exists && parent[node] >= 0 && time[node] >= time[parent[node]]) {
            const int32_t p = parent[node];
            if (!primary(node)) {
              if (!is_sat[node]) {
                // halo merger: standardMerge -> singleLevelHierarchyProcess
                int32_t k = first_child[p];
                while (sibling[k] != node) k = sibling[k];
                sibling[k] = sibling[node];
                sibling[node] = -1;
                append_satellite(p, node);
                move_satellites(node, p);
                is_sat[node] = 1;
                merge_time[node] = time[node] + merging_time(first_child[p], node, time[node]);
                double* q = &par[(size_t)GAL_NPARAM_NODE * node];
                q[GAL_NP_MASS_REFERENCE] = halo_mass(node, time[node]);
                q[GAL_NP_TIME_REFERENCE] = time[node];
                q[GAL_NP_MASS_RATE] = 0.0;
                q[GAL_NP_SPIN_RATE] = 0.0;
                q[GAL_NP_IS_SATELLITE] = 1.0;
                emit(GAL_OP_BECOME_SATELLITE, node, p, time[node], merge_time[node], nullptr);
              }
            } else if (sibling[node] < 0) {
              // standardPromote (node_evolver/standard.F90:1241-1327)
              emit(GAL_OP_PROMOTE, node, p, time[node], time[p], nullptr);
              int32_t host = p;
              while (is_sat[host]) host = parent[host];
              move_satellites(node, host);
              first_child[p] = -1;
              alive[node] = 0;
              exists = false;
              next = p;
              next_found = true;
            }
          }
          if (!next_found && exists && progressed) next = node;
        }
        if (next < 0) {
          if (exists) {
            next = walk_next(node);
          } else if (!next_found) {
            if (sibling_next >= 0 || parent_next >= 0) {
              next = descend(sibling_next >= 0 ? sibling_next : parent_next);
            } else {
              next = -1;
            }
          }
        }
        node = next;
      }
      changed = ops.size() > n_ops;
      if (did_evolve && !changed) return false;
    }
    return true;
  }
};

// Run fn(i) for i in [0, n) on the host's cores (trees are independent);
// OMP_NUM_THREADS, when set, caps the number of threads.
template <typename F>
void parallel_trees(int n, F fn) {
  int nt = (int)std::thread::hardware_concurrency();
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) nt = std::min(nt, v);
  }
  nt = std::max(1, std::min(nt, n));
  std::atomic<int> next{0};
  std::atomic<bool> failed{false};
  auto work = [&]() {
    try {
      for (int i = next++; i < n; i = next++) fn(i);
    } catch (...) {
      failed = true;
    }
  };
  if (nt == 1) {
    work();
  } else {
    std::vector<std::thread> pool;
    for (int k = 0; k < nt; ++k) pool.emplace_back(work);
    for (auto& th : pool) th.join();
  }
  if (failed) throw std::bad_alloc();
}

}  // namespace


extern "C" {

int gal_cosmology_sigma(const gal_cosmology* cosmo, int n, const double* mass, double* sigma) {
  if (!cosmo || n < 0 || (n > 0 && (!mass || !sigma))) return GAL_E_INVALID;
  Cosmology c(*cosmo);
  for (int i = 0; i < n; ++i) sigma[i] = c.sigma(mass[i]);
  return GAL_OK;
}

int gal_cosmology_delta_critical(const gal_cosmology* cosmo, int n, const double* time, double* delta) {
  if (!cosmo || n < 0 || (n > 0 && (!time || !delta))) return GAL_E_INVALID;
  Cosmology c(*cosmo);
  for (int i = 0; i < n; ++i) delta[i] = c.delta_c(time[i]);
  return GAL_OK;
}

int gal_forest_build_pch(const gal_cosmology* cosmo, const gal_pch_params* pch, int n_trees, const double* root_mass,
                         double time_root, uint64_t seed, gal_forest** out) {
  if (!cosmo || !pch || n_trees < 0 || (n_trees > 0 && !root_mass) || !out) return GAL_E_INVALID;
  if (!(pch->mass_resolution > 0.0) || !(pch->merge_probability > 0.0) || !(pch->accretion_limit > 0.0) ||
      pch->accretion_limit >= 1.0 || !(pch->time_earliest > 0.0) || pch->time_earliest >= time_root)
    return GAL_E_INVALID;
  *out = nullptr;
  try {
    Cosmology c(*cosmo);
    double m_max = 0.0;
    for (int i = 0; i < n_trees; ++i) m_max = std::max(m_max, root_mass[i]);
    Branching br(c, *pch, m_max);
    // each tree from its own generator, built in parallel, then concatenated in order
    std::vector<gal_forest> parts(n_trees);
    parallel_trees(n_trees, [&](int i) {
      std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ull + (uint64_t)i * 0xD1B54A32D192ED03ull + 1ull);
      build_tree(c, br, root_mass[i], time_root, rng, parts[i]);
    });
    auto f = std::make_unique<gal_forest>();
    f->tree_offset.push_back(0);
    for (gal_forest& p : parts) {
      const int32_t base = (int32_t)f->parent.size();
      auto shift = [&](int32_t k) { return k < 0 ? -1 : k + base; };
      for (size_t k = 0; k < p.parent.size(); ++k) {
        f->parent.push_back(shift(p.parent[k]));
        f->first_child.push_back(shift(p.first_child[k]));
        f->sibling.push_back(shift(p.sibling[k]));
      }
      f->mass.insert(f->mass.end(), p.mass.begin(), p.mass.end());
      f->time.insert(f->time.end(), p.time.begin(), p.time.end());
      f->tree_offset.push_back((int64_t)f->parent.size());
      f->n_trees += 1;
      p = gal_forest();
    }
    *out = f.release();
  } catch (const std::bad_alloc&) {
    return GAL_E_NOMEM;
  }
  return GAL_OK;
}

int gal_forest_from_arrays(int64_t n_nodes, int n_trees, const int64_t* tree_offset, const int32_t* parent,
                           const int32_t* first_child, const int32_t* sibling, const double* mass,
                           const double* time, gal_forest** out) {
  if (!out || n_nodes < 0 || n_trees < 0 || !tree_offset) return GAL_E_INVALID;
  if (n_nodes > 0 && (!parent || !first_child || !sibling || !mass || !time)) return GAL_E_INVALID;
  if (tree_offset[0] != 0 || tree_offset[n_trees] != n_nodes) return GAL_E_INVALID;
  for (int t = 0; t < n_trees; ++t) {
    if (tree_offset[t + 1] <= tree_offset[t]) return GAL_E_INVALID;
    for (int64_t i = tree_offset[t]; i < tree_offset[t + 1]; ++i) {
      const int32_t links[3] = {parent[i], first_child[i], sibling[i]};
      for (int32_t l : links)
        if (l != -1 && (l < tree_offset[t] || l >= tree_offset[t + 1])) return GAL_E_INVALID;
    }
    if (parent[tree_offset[t]] != -1) return GAL_E_INVALID;
  }
  auto f = std::make_unique<gal_forest>();
  f->n_trees = n_trees;
  f->tree_offset.assign(tree_offset, tree_offset + n_trees + 1);
  f->parent.assign(parent, parent + n_nodes);
  f->first_child.assign(first_child, first_child + n_nodes);
  f->sibling.assign(sibling, sibling + n_nodes);
  f->mass.assign(mass, mass + n_nodes);
  f->time.assign(time, time + n_nodes);
  *out = f.release();
  return GAL_OK;
}

int gal_forest_size(const gal_forest* f, int64_t* n_nodes, int* n_trees) {
  if (!f) return GAL_E_INVALID;
  if (n_nodes) *n_nodes = (int64_t)f->parent.size();
  if (n_trees) *n_trees = f->n_trees;
  return GAL_OK;
}

int gal_forest_arrays(const gal_forest* f, int64_t* tree_offset, int32_t* parent, int32_t* first_child,
                      int32_t* sibling, double* mass, double* time) {
  if (!f) return GAL_E_INVALID;
  const size_t n = f->parent.size();
  if (tree_offset) std::memcpy(tree_offset, f->tree_offset.data(), sizeof(int64_t) * (f->n_trees + 1));
  if (parent) std::memcpy(parent, f->parent.data(), sizeof(int32_t) * n);
  if (first_child) std::memcpy(first_child, f->first_child.data(), sizeof(int32_t) * n);
  if (sibling) std::memcpy(sibling, f->sibling.data(), sizeof(int32_t) * n);
  if (mass) std::memcpy(mass, f->mass.data(), sizeof(double) * n);
  if (time) std::memcpy(time, f->time.data(), sizeof(double) * n);
  return GAL_OK;
}

int gal_forest_destroy(gal_forest* f) {
  delete f;
  return GAL_OK;
}

int gal_schedule_build(const gal_forest* f, const gal_cosmology* cosmo, const gal_walk_params* walk,
                       gal_schedule** out) {
  if (!f || !cosmo || !walk || !out) return GAL_E_INVALID;
  *out = nullptr;
  try {
    Cosmology c(*cosmo);
    auto s = std::make_unique<gal_schedule>();
    std::vector<std::vector<gal_schedule_op>> parts(f->n_trees);
    std::vector<char> ok(f->n_trees, 1);
    parallel_trees(f->n_trees, [&](int t) {
      Walk wk(*f, c, *walk, parts[t], f->tree_offset[t], f->tree_offset[t + 1]);
      ok[t] = wk.run() ? 1 : 0;
    });
    for (int t = 0; t < f->n_trees; ++t)
      if (!ok[t]) return GAL_E_INVALID;
    size_t total = 0;
    for (const auto& p : parts) total += p.size();
    s->ops.reserve(total);
    for (auto& p : parts) {
      s->ops.insert(s->ops.end(), p.begin(), p.end());
      std::vector<gal_schedule_op>().swap(p);
    }
    // waves: one more than the latest earlier operation on the same node(s)
    std::vector<int32_t> last(f->parent.size(), -1);
    int n_waves = 0;
    for (gal_schedule_op& op : s->ops) {
      // becoming a satellite changes no state: it orders only its own node
      const bool uses_b = op.b >= 0 && op.kind != GAL_OP_BECOME_SATELLITE;
      int wv = last[op.a] + 1;
      if (uses_b) wv = std::max(wv, last[op.b] + 1);
      op.wave = wv;
      last[op.a] = wv;
      if (uses_b) last[op.b] = wv;
      n_waves = std::max(n_waves, wv + 1);
    }
    s->n_waves = n_waves;
    *out = s.release();
  } catch (const std::bad_alloc&) {
    return GAL_E_NOMEM;
  }
  return GAL_OK;
}

int gal_schedule_size(const gal_schedule* s, int64_t* n_ops, int* n_waves) {
  if (!s) return GAL_E_INVALID;
  if (n_ops) *n_ops = (int64_t)s->ops.size();
  if (n_waves) *n_waves = s->n_waves;
  return GAL_OK;
}

int gal_schedule_ops(const gal_schedule* s, gal_schedule_op* ops) {
  if (!s || (!ops && !s->ops.empty())) return GAL_E_INVALID;
  if (!s->ops.empty()) std::memcpy(ops, s->ops.data(), sizeof(gal_schedule_op) * s->ops.size());
  return GAL_OK;
}

int gal_schedule_node_gyr(const gal_schedule* s, double* node_gyr) {
  if (!s || !node_gyr) return GAL_E_INVALID;
  double sum = 0.0;
  for (const gal_schedule_op& o : s->ops)
    if (o.kind == GAL_OP_EVOLVE) sum += o.t1 - o.t0;
  *node_gyr = sum;
  return GAL_OK;
}

int gal_schedule_destroy(gal_schedule* s) {
  delete s;
  return GAL_OK;
}

int gal_forest_initial_state(const gal_forest* f, const gal_cosmology* cosmo, const gal_walk_params* walk,
                             double metal_fraction, double black_hole_seed, int ld, double* y) {
  if (!f || !cosmo || !walk || !y || ld < (int64_t)f->parent.size()) return GAL_E_INVALID;
  Cosmology c(*cosmo);
  const double fb = cosmo->omega_baryon / cosmo->omega_matter;
  const size_t n = f->parent.size();
  for (int v = 0; v < GAL_NVAR; ++v) std::memset(y + (size_t)v * ld, 0, sizeof(double) * n);
  for (size_t i = 0; i < n; ++i) {
    if (f->first_child[i] >= 0) continue;
    const double m = f->mass[i], t = f->time[i];
    double r, v;
    virial(c, m, t, r, v);
    const double j = walk->spin_parameter * m * v * r;
    auto at = [&](int var) -> double& { return y[(size_t)var * ld + i]; };
    at(GAL_V_HOT_MASS) = fb * m;
    at(GAL_V_HOT_ABUNDANCES) = metal_fraction * fb * m;
    at(GAL_V_HOT_ANGULAR_MOMENTUM) = fb * j;
    at(GAL_V_HOT_OUTER_RADIUS) = r;
    at(GAL_V_DISK_FRACTION_MASS_RETAINED) = 1.0;
    at(GAL_V_BH_MASS) = black_hole_seed;
    at(GAL_V_SPIN_ANGULAR_MOMENTUM) = j;
    at(GAL_V_SATELLITE_BOUND_MASS) = m;
  }
  return GAL_OK;
}

}  // extern "C"
