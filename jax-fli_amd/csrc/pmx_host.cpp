// Host-side numerics of the pmx Python layer that are too slow as Python loops: the linear-growth
// ODE table behind growth_factor / growth_rate / Gf / gp / dGfa (the jaxpm.growth and
// jax_cosmo.background functions the solvers' coefficients come from, src/jax_fli/pm/solvers.py:8-9).
// Plain C++ on the CPU (no device code); the Python layer uses it through ctypes and keeps an
// operation-for-operation identical Python version of the same loop.
#include "pmx_cosmo.h"

#include <algorithm>
#include <cmath>

namespace {

struct Bg {
  double om, ok, ode, w0, wa;
};

// y = (D, G = dD/dln a);  dG/dln a = -(2 + dlnE/dlna) G + 1.5 Om D / (a^3 E^2)
inline void rhs(const Bg& c, double x, double D, double G, double& dD, double& dG) {
  const double a = std::exp(x);
  const double p = -3.0 * (1.0 + c.w0 + c.wa);
  const double fde = std::pow(a, p) * std::exp(-3.0 * c.wa * (1.0 - a));
  const double om = c.om * std::pow(a, -3.0);
  const double ok = c.ok * std::pow(a, -2.0);
  const double e2 = om + ok + c.ode * fde;
  const double dlne = 0.5 * (-3.0 * om - 2.0 * ok + c.ode * fde * (p + 3.0 * c.wa * a)) / e2;
  dD = G;
  dG = -(2.0 + dlne) * G + 1.5 * om * D / e2;
}

}  // namespace

extern "C" {

// RK4 in ln a from x0 with step h over n_tot steps, starting from the matter-dominated growing mode
// D = G = d0 (= exp(x0)); D and G are normalised by D at node nstep_norm (a = 1); dG = dG/dln a at the
// normalised nodes. Outputs have n_tot + 1 entries. Returns 0.
int pmx_host_growth_table(double Omega_m, double Omega_k, double Omega_de, double w0, double wa, double x0, double d0,
                          double h, int nstep_norm, int n_tot, double* D_out, double* G_out, double* dG_out) {
  const Bg c{Omega_m, Omega_k, Omega_de, w0, wa};
  double D = d0, G = d0;
  D_out[0] = D;
  G_out[0] = G;
  for (int i = 0; i < n_tot; ++i) {
    const double x = x0 + h * i;
    double k1d, k1g, k2d, k2g, k3d, k3g, k4d, k4g;
    rhs(c, x, D, G, k1d, k1g);
    rhs(c, x + 0.5 * h, D + 0.5 * h * k1d, G + 0.5 * h * k1g, k2d, k2g);
    rhs(c, x + 0.5 * h, D + 0.5 * h * k2d, G + 0.5 * h * k2g, k3d, k3g);
    rhs(c, x + h, D + h * k3d, G + h * k3g, k4d, k4g);
    D += (h / 6.0) * (k1d + 2 * k2d + 2 * k3d + k4d);
    G += (h / 6.0) * (k1g + 2 * k2g + 2 * k3g + k4g);
    D_out[i + 1] = D;
    G_out[i + 1] = G;
  }
  const double norm = D_out[nstep_norm];
  for (int i = 0; i <= n_tot; ++i) {
    D_out[i] = D_out[i] / norm;
    G_out[i] = G_out[i] / norm;
    double dd, dg;
    rhs(c, x0 + h * i, D_out[i], G_out[i], dd, dg);
    dG_out[i] = dg;
  }
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ request host preparation
namespace pmxc {

namespace {
constexpr double kAInit = 1e-5, kAMax = 2.0, kTcmb = 2.726, kPi = 3.141592653589793;
constexpr int kNStep = 2048;  // RK4 steps from kAInit to a = 1

// sum of a strided run in the order numpy's pairwise summation uses (blocks of 8 partial sums up to
// 128 elements, halves above)
double pairwise_sum(const double* a, long n, long stride) {
  if (n < 8) {
    double r = -0.0;
    for (long i = 0; i < n; ++i) r += a[i * stride];
    return r;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j * stride];
    long i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += a[(i + j) * stride];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i * stride];
    return res;
  }
  long n2 = n / 2;
  n2 -= n2 % 8;
  return pairwise_sum(a, n2, stride) + pairwise_sum(a + n2 * stride, n - n2, stride);
}
}  // namespace

Background::Background(const Params& p) : p_(p) {
  x0_ = std::log(kAInit);
  h_ = -x0_ / kNStep;
  const int n_tot = kNStep + (int)std::ceil(std::log(kAMax) / h_);
  D_.resize(n_tot + 1);
  G_.resize(n_tot + 1);
  dG_.resize(n_tot + 1);
  pmx_host_growth_table(p.Omega_m(), p.Omega_k, p.Omega_de(), p.w0, p.wa, x0_, kAInit, h_, kNStep, n_tot, D_.data(),
                        G_.data(), dG_.data());
}

double Background::E(double a) const {
  const double fde = std::pow(a, -3.0 * (1.0 + p_.w0 + p_.wa)) * std::exp(-3.0 * p_.wa * (1.0 - a));
  return std::sqrt(p_.Omega_m() * std::pow(a, -3.0) + p_.Omega_k * std::pow(a, -2.0) + p_.Omega_de() * fde);
}

// cubic Hermite interpolation of (D, dD/dlna) and (G, dG/dlna) in ln a on the table's uniform grid
void Background::interp(double a, double& d, double& g) const {
  const double x = std::log(a);
  const double h = (x0_ + h_ * 1.0) - (x0_ + h_ * 0.0);
  long i = (long)((x - (x0_ + h_ * 0.0)) / h);
  i = std::max(0L, std::min(i, (long)D_.size() - 2));
  const double xi = x0_ + h_ * (double)i;
  const double t = (x - xi) / h;
  const double h00 = (1 + 2 * t) * ((1 - t) * (1 - t));
  const double h10 = t * ((1 - t) * (1 - t));
  const double h01 = t * t * (3 - 2 * t);
  const double h11 = t * t * (t - 1);
  d = h00 * D_[i] + h10 * h * G_[i] + h01 * D_[i + 1] + h11 * h * G_[i + 1];
  g = h00 * G_[i] + h10 * h * dG_[i] + h01 * G_[i + 1] + h11 * h * dG_[i + 1];
}

double Background::D(double a) const {
  double d, g;
  interp(a, d, g);
  return d;
}
double Background::G(double a) const {
  double d, g;
  interp(a, d, g);
  return g;
}
double Background::Gf(double a) const { return G(a) * (a * a) * E(a); }
double Background::gp(double a) const { return G(a) / a; }
double Background::dGfa(double a) const { return 1.5 * p_.Omega_m() * D(a) / ((a * a) * E(a)); }

// EH98 (Eisenstein & Hu 1998) with its k-independent scales computed once per cosmology.
struct EH98 {
  double fb, fc, k_eq, sh_d, k_silk, alpha_c, beta_c, alpha_b, beta_node, beta_b;
  explicit EH98(const Params& c) {
    const double th2 = std::pow(kTcmb / 2.7, 2.0);
    const double h = c.h;
    const double w_m = c.Omega_m() * h * h, w_b = c.Omega_b * h * h;
    fb = c.Omega_b / c.Omega_m();
    fc = c.Omega_c / c.Omega_m();
    k_eq = 7.46e-2 * w_m / th2 / h;
    const double b1 = 0.313 * std::pow(w_m, -0.419) * (1.0 + 0.607 * std::pow(w_m, 0.674));
    const double b2 = 0.238 * std::pow(w_m, 0.223);
    // eq. 4: the last factor takes Omega_b h^2
    const double z_d =
        1291.0 * std::pow(w_m, 0.251) / (1.0 + 0.659 * std::pow(w_m, 0.828)) * (1.0 + b1 * std::pow(w_b, b2));
    const double z_eq = 2.50e4 * w_m / (th2 * th2);
    const double R_eq = 31.5 * w_b / (th2 * th2) * (1.0e3 / z_eq);
    const double R_d = 31.5 * w_b / (th2 * th2) * (1.0e3 / z_d);
    sh_d = 2.0 / (3.0 * k_eq) * std::sqrt(6.0 / R_eq) *
           std::log((std::sqrt(1.0 + R_d) + std::sqrt(R_eq + R_d)) / (1.0 + std::sqrt(R_eq)));
    k_silk = 1.6 * std::pow(w_b, 0.52) * std::pow(w_m, 0.73) * (1.0 + std::pow(10.4 * w_m, -0.95)) / h;
    const double a1 = std::pow(46.9 * w_m, 0.670) * (1.0 + std::pow(32.1 * w_m, -0.532));
    const double a2 = std::pow(12.0 * w_m, 0.424) * (1.0 + std::pow(45.0 * w_m, -0.582));
    alpha_c = std::pow(a1, -fb) * std::pow(a2, -std::pow(fb, 3.0));
    const double bb1 = 0.944 / (1.0 + std::pow(458.0 * w_m, -0.708));
    const double bb2 = std::pow(0.395 * w_m, -0.0266);
    beta_c = 1.0 / (1.0 + bb1 * (std::pow(fc, bb2) - 1.0));
    const double y = (1.0 + z_eq) / (1.0 + z_d);
    const double x = std::sqrt(1.0 + y);
    const double G = y * (-6.0 * x + (2.0 + 3.0 * y) * std::log((x + 1.0) / (x - 1.0)));
    alpha_b = 2.07 * k_eq * sh_d * std::pow(1.0 + R_d, -0.75) * G;
    beta_node = 8.41 * std::pow(w_m, 0.435);
    beta_b = 0.5 + fb + (3.0 - 2.0 * fb) * std::sqrt(std::pow(17.2 * w_m, 2.0) + 1.0);
  }
  double T(double k) const {
    const double q = k / (13.41 * k_eq);
    const double q108 = std::pow(q, 1.08);
    auto t_tilde = [&](double alpha, double beta) {
      const double L = std::log(2.718281828459045 + 1.8 * beta * q);
      const double C = 14.2 / alpha + 386.0 / (1.0 + 69.9 * q108);
      return L / (L + C * q * q);
    };
    const double f = 1.0 / (1.0 + std::pow(k * sh_d / 5.4, 4.0));
    const double Tc = f * t_tilde(1.0, beta_c) + (1.0 - f) * t_tilde(alpha_c, beta_c);
    const double tilde_s = sh_d / std::pow(1.0 + std::pow(beta_node / (k * sh_d), 3.0), 1.0 / 3.0);
    const double sx = kPi * (k * tilde_s / kPi);
    const double sinc = std::sin(sx) / sx;
    const double Tb = (t_tilde(1.0, 1.0) / (1.0 + std::pow(k * sh_d / 5.2, 2.0)) +
                       alpha_b / (1.0 + std::pow(beta_b / (k * sh_d), 3.0)) * std::exp(-std::pow(k / k_silk, 1.4))) *
                      sinc;
    return fb * Tb + fc * Tc;
  }
};

double eisenstein_hu(const Params& c, double k) { return EH98(c).T(k); }

double sigma8_norm(const Params& c) {
  // sigma^2(R = 8) = 1/(2 pi^2) int dlnk k^3 T^2 k^ns W(kR)^2, Simpson over 4097 points of ln k
  const int n = (1 << 12) + 1;
  const double lo = std::log(1e-5), hi = std::log(1e3);
  const double step = (hi - lo) / (double)(n - 1);
  const EH98 eh(c);
  std::vector<double> lnk(n), f(n);
  for (int i = 0; i < n; ++i) lnk[i] = (double)i * step + lo;
  lnk[n - 1] = hi;
  for (int i = 0; i < n; ++i) {
    const double k = std::exp(lnk[i]);
    const double x = k * 8.0;
    const double w = 3.0 * (std::sin(x) - x * std::cos(x)) / std::pow(x, 3.0);
    const double T = eh.T(k);
    f[i] = std::pow(k, 3.0) * (T * T) * std::pow(k, c.n_s) * w * w / (2.0 * std::pow(kPi, 2.0));
  }
  const double h = lnk[1] - lnk[0];
  const double odd = pairwise_sum(f.data() + 1, (n - 1) / 2, 2);   // f[1:-1:2]
  const double even = pairwise_sum(f.data() + 2, (n - 3) / 2, 2);  // f[2:-1:2]
  const double s = (h / 3.0) * (f[0] + f[n - 1] + 4.0 * odd + 2.0 * even);
  return std::pow(c.sigma8, 2.0) / s;
}

std::vector<std::pair<float, float>> time_grid(double t0, const std::vector<float>& ts, double dt0, double t1) {
  const int max_steps = (int)std::ceil(t1 / dt0) + 5;
  const float tol = 1e-6f, dt = (float)dt0;
  float t = (float)t0;
  std::vector<std::pair<float, float>> steps;
  for (float tt : ts) {
    int n = 0;
    while (t < tt && n < max_steps) {
      float tn = t + dt;
      if (tn > tt - tol) tn = tt;
      steps.emplace_back(t, tn);
      t = tn;
      ++n;
    }
  }
  return steps;
}

RequestPlan plan_request(const Params& p, int n, double box, double a0_in, double a1_in, double dt0_in) {
  RequestPlan r;
  const Background bg(p);
  // the colouring table: 128 log-spaced k in [1e-4, 10] h/Mpc (initial.py:171-173), P(k) * ncell / V
  const double norm = sigma8_norm(p);
  const EH98 eh(p);
  const double D1 = bg.D(1.0);
  const double cell = (double)((long long)n * n * n) / std::pow(box, 3.0);
  const double step = (1.0 - (-4.0)) / 127.0;
  for (int i = 0; i < 128; ++i) {
    const double y = i == 127 ? 1.0 : (double)i * step + -4.0;
    const double k = std::pow(10.0, y);
    const double T = eh.T(k);
    r.ktab.push_back(k);
    r.ptab.push_back(norm * std::pow(k, p.n_s) * (T * T) * (D1 * D1) * cell);
  }
  // LPT1 at a0: dx = D(a0) F, p = a0^2 f E dx (arguments rounded to float32 at the ABI)
  const double a = (double)(float)a0_in;
  r.lpt_growth = (float)bg.D(a);
  r.lpt_vel = (float)(((a * a) * bg.growth_rate(a)) * bg.E(a));
  r.force_scale = 1.5 * p.Omega_m();
  // KDK: the initial half kick (solvers.py:136-147), then drift + double kick per step (151-202)
  const float dt = (float)dt0_in;
  const float nan = std::nanf("");
  std::vector<Op> raw;
  {
    const float ai = (float)a0_in;
    const float af = (float)std::pow(a0_in * (a0_in + dt0_in), 0.5);
    const double kick = (bg.Gf(af) - bg.Gf(ai)) / bg.dGfa(ai);
    const double ac2 = (double)(ai * ai);  // a float32 square, as numpy evaluates it
    raw.push_back(Op{nan, (float)(kick / (ac2 * bg.E(ai))), nan});
  }
  for (const auto& st : time_grid(a0_in, std::vector<float>{(float)a1_in}, dt0_in, a1_in)) {
    const float t0 = st.first, t1 = st.second;
    const float t0t1 = std::sqrt(t0 * t1);
    const double drift = (bg.D(t1) - bg.D(t0)) / bg.gp(t0t1);
    const double cd = drift / (std::pow((double)t0t1, 3.0) * bg.E(t0t1));
    const float t2 = t1 + dt;
    const float t1t2 = std::sqrt(t1 * t2);
    const double dgf = bg.dGfa(t1), gf1 = bg.Gf(t1);
    const double k1 = (gf1 - bg.Gf(t0t1)) / dgf;
    const double k2 = (bg.Gf(t1t2) - gf1) / dgf;
    const double ck = (k1 + k2) / (((double)t1 * (double)t1) * bg.E(t1));
    raw.push_back(Op{(float)cd, (float)ck, nan});
  }
  // fold each leading drift into the previous step's readout (integrate's fused schedule)
  for (size_t i = 0; i + 1 < raw.size(); ++i) {
    if (std::isnan(raw[i].post) && !std::isnan(raw[i + 1].pre)) {
      raw[i].post = raw[i + 1].pre;
      raw[i + 1].pre = nan;
    }
  }
  r.ops = raw;
  return r;
}

}  // namespace pmxc

extern "C" {

// Host preparation of one forward-simulation request (the server's, exported for checks): params =
// {Omega_c, Omega_b, h, n_s, sigma8, Omega_k, w0, wa}; fills ktab/ptab (128 each), lpt = {growth,
// vel factor}, ops (max_ops x 3 floats: leading drift, kick, trailing drift; NaN = none), nops and
// force_scale. Returns 0, or 1 when more than max_ops steps are needed.
int pmx_host_request_plan(const double* params, int n, double box, double a0, double a1, double dt0, double* ktab,
                          double* ptab, float* lpt, float* ops, int max_ops, int* nops, double* force_scale) {
  const pmxc::Params p{params[0], params[1], params[2], params[3], params[4], params[5], params[6], params[7]};
  const pmxc::RequestPlan r = pmxc::plan_request(p, n, box, a0, a1, dt0);
  *nops = (int)r.ops.size();
  if ((int)r.ops.size() > max_ops) return 1;
  for (int i = 0; i < 128; ++i) {
    ktab[i] = r.ktab[i];
    ptab[i] = r.ptab[i];
  }
  lpt[0] = r.lpt_growth;
  lpt[1] = r.lpt_vel;
  for (size_t i = 0; i < r.ops.size(); ++i) {
    ops[3 * i] = r.ops[i].pre;
    ops[3 * i + 1] = r.ops[i].kick;
    ops[3 * i + 2] = r.ops[i].post;
  }
  *force_scale = r.force_scale;
  return 0;
}

}  // extern "C"
