// Host-side numerics of the pmx Python layer that are too slow as Python loops: the linear-growth
// ODE table behind growth_factor / growth_rate / Gf / gp / dGfa (the jaxpm.growth and
// jax_cosmo.background functions the solvers' coefficients come from, src/jax_fli/pm/solvers.py:8-9).
// Plain C++ on the CPU (no device code); the Python layer uses it through ctypes and keeps an
// operation-for-operation identical Python version of the same loop.
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
