// node_host.cpp - host (FP64) build of the per-node arithmetic in lbm_node.h.
//
// The GPU kernels instantiate lbm_node.h with T = float; this file
// instantiates the very same templates with T = double and exports them with
// a C ABI so the CPU test-suite can check the algebra the kernels run against
// the FP64 oracle without a GPU.  Not used by the solver itself.
#include "lbm_node.h"

namespace {
fdn::Consts<double> consts_from(const double c[9]) {
    fdn::Consts<double> k;
    k.inv_tau_g = c[0];
    k.eta_n = c[1];
    k.M = c[2];
    k.a = c[3];
    k.kappa = c[4];
    k.gamma = c[5];
    k.psi_wall = c[6];
    k.gx = c[7];
    k.gy = c[8];
    k.psi_visc = 1e300;
    return k;
}
}  // namespace

extern "C" {

// 1 / tau_mix with the viscosity interpolation's |psi| bound (<= 0: none).
double fdn_inv_tau_mix_f64(double rho, double psi, const double c[9], double bound) {
    fdn::Consts<double> k = consts_from(c);
    if (bound > 0) k.psi_visc = bound;
    return fdn::inv_tau_mix(rho, psi, k);
}

// n[9] psi on the node and its neighbours (C E N W S NE NW SW SE) -> out[3] = dx, dy, lap
void fdn_stencil_f64(const double n[9], double out[3]) { fdn::stencil(n, out[0], out[1], out[2]); }

// Moments (rho, psi, mu, ux, uy, p, 1/tau_mix, Fx, Fy) from post-stream f.
void fdn_moments_f64(const double f[9], double psi, double dpx, double dpy, double lap, const double c[9],
                     double mom[9]) {
    const fdn::Mom<double> m = fdn::moments(f, psi, dpx, dpy, lap, consts_from(c));
    const double v[9] = {m.rho, m.psi, m.mu, m.ux, m.uy, m.p, m.itau, m.fx, m.fy};
    for (int i = 0; i < 9; ++i) mom[i] = v[i];
}

// feq, geq for given moments (same layout as fdn_moments_f64's output).
void fdn_equilibria_f64(const double mom[9], const double c[9], double feq[9], double geq[9]) {
    const fdn::Mom<double> m{mom[0], mom[1], mom[2], mom[3], mom[4], mom[5], mom[6], mom[7], mom[8]};
    fdn::equilibria(m, consts_from(c), feq, geq);
}

// In-place collision of f and g with the given moments.
void fdn_collide_f64(double f[9], double g[9], const double mom[9], const double c[9]) {
    const fdn::Mom<double> m{mom[0], mom[1], mom[2], mom[3], mom[4], mom[5], mom[6], mom[7], mom[8]};
    fdn::collide(f, g, m, consts_from(c));
}

void fdn_inlet_f64(double f[9], double g[9], double u, double px, double py) {
    fdn::inlet_f(f, u, px, py);
    fdn::inlet_g(g);
}

void fdn_outlet_f64(double f[9], double g[9], double u, double px, double py, int pressure, double rho_out,
                    double c3) {
    fdn::outlet_f(f, u, px, py, pressure != 0, rho_out, c3);
    fdn::outlet_g(g);
}

void fdn_inlet_corner_f64(double f[9], double g[9], double rho_nb, int top) {
    if (top) {
        fdn::inlet_corner_top_f(f, rho_nb);
        fdn::inlet_corner_top_g(g);
    } else {
        fdn::inlet_corner_bottom_f(f, rho_nb);
        fdn::inlet_corner_bottom_g(g);
    }
}

double fdn_inlet_rho_f64(const double f[9], double u, double px) { return fdn::inlet_rho(f, u, px); }

}  // extern "C"
