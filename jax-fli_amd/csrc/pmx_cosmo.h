// Host-side cosmology of one forward simulation (plain C++, no device code): the expansion rate,
// the linear growth table and the FastPM growth functions, the Eisenstein & Hu (1998) linear power
// spectrum with its sigma8 normalisation, the float32 time grid of the integration loop and the
// kick/drift coefficients of EfficientDriftDoubleKick. These are the host numerics behind
// jfli.nbody / lpt / gaussian_initial_conditions (src/jax_fli/pm/solvers.py:118-213,
// pm/integrate.py:17-33, initial.py:124-199), written once for the request server so that a request
// is prepared without the Python layer. The operations follow the Python layer's
// (jax-fli_amd/cosmology.py, pm/solvers.py, pm/integrate.py) term for term.
#pragma once
#include <cstddef>
#include <utility>
#include <vector>

namespace pmxc {

struct Params {
  double Omega_c, Omega_b, h, n_s, sigma8, Omega_k, w0, wa;
  double Omega_m() const { return Omega_c + Omega_b; }
  double Omega_de() const { return 1.0 - Omega_k - Omega_m(); }
};

class Background {
 public:
  explicit Background(const Params& p);
  double E(double a) const;
  double D(double a) const;   // linear growth, D(1) = 1
  double G(double a) const;   // dD / dln a
  double Gf(double a) const;  // a^3 E dD/da = G a^2 E
  double gp(double a) const;  // dD/da = G / a
  double dGfa(double a) const;  // 1.5 Omega_m D / (a^2 E)
  double growth_rate(double a) const { return G(a) / D(a); }
  const Params& params() const { return p_; }

 private:
  void interp(double a, double& d, double& g) const;
  Params p_;
  double x0_, h_;
  std::vector<double> D_, G_, dG_;
};

// EH98 transfer function at k (h/Mpc), and P(k) = sigma8-normalised k^ns T^2 D(a)^2.
double eisenstein_hu(const Params& p, double k);
double sigma8_norm(const Params& p);

// float32 step grid of integrate(): (t_curr, t_next) pairs from t0 to each snapshot time, restarting
// at every snapshot; clip tolerance 1e-6.
std::vector<std::pair<float, float>> time_grid(double t0, const std::vector<float>& ts, double dt0, double t1);

// One fused PM step of the request loop: optional leading drift, kick, optional trailing drift
// (absent coefficients are NaN).
struct Op {
  float pre, kick, post;
};

// Everything the device needs for one request: the colouring's P(k) table (times ncell / volume),
// the LPT factors, the KDK steps with each leading drift folded into the previous step's readout,
// and the force scale 1.5 Omega_m.
struct RequestPlan {
  std::vector<double> ktab, ptab;
  float lpt_growth, lpt_vel;
  std::vector<Op> ops;
  double force_scale;
};
RequestPlan plan_request(const Params& p, int n, double box, double a0, double a1, double dt0);

}  // namespace pmxc
