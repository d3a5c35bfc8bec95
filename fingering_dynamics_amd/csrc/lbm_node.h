// lbm_node.h - per-node arithmetic of the two-phase free-energy D2Q9 step.
//
// Shared by the gfx950 kernels (T = float) and by a host build used in the
// CPU tests (T = double), so the algebra the GPU runs is checked against the
// FP64 oracle without a GPU.  Everything here is a pure function of its
// arguments.
//
// Reference formulas (lattice_boltzmann/fingering_periodic.py):
//   stencils      d_x psi, d_y psi, lap psi            :214-256
//   mu            a psi (1 - psi^2) - kappa lap psi    :141-149
//   velocity      (sum f e + mu grad psi / 2) / rho    :126-139
//   pressure      rho/3 + psi mu                       :123-124
//   tau_mix       3 * 2 nu1 nu2/(nu1(1-psi)+nu2(1+psi)) + 1/2, nu1 = Eta_n/rho,
//                 nu2 = M Eta_n/rho                    :201-208
//   equilibria    feq_i = w_i (A_i + rho Q_i), geq_i = w_i (B_i + psi Q_i),
//                 Q_i = 3 e.u + 4.5 (e.u)^2 - 1.5 u.u  :155-192
//   forcing       F_i = w_i (1 - 1/(2 tau_mix)) [(3(e-u) + 9 (e.u) e) . Force],
//                 Force = mu grad psi                  :194-199
//   collision     f* = f - (f - feq)/tau_mix + F,  g* = g - (g - geq)/tau   :258-264
//   Zou-He        inlet :268-302, outlet :305-324; walled corners
//                 lattice_boltzmann/fingering.py:334-364, :387-390
#pragma once

#if defined(__HIPCC__)
#define FD_HD __host__ __device__ __forceinline__
#else
#define FD_HD inline
#endif

namespace fdn {

// Reciprocal: the hardware approximation (v_rcp_f32, 1 ulp) in FP32 device
// code, an exact division elsewhere.
FD_HD float rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
FD_HD double rcp(double x) { return 1.0 / x; }

// D2Q9 order of fingering_periodic.py:50-89.
constexpr int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

template <typename T>
struct W9 {
    static constexpr T w0 = T(4) / T(9);
    static constexpr T w1 = T(1) / T(9);
    static constexpr T w5 = T(1) / T(36);
    static FD_HD T w(int i) { return i == 0 ? w0 : (i < 5 ? w1 : w5); }
};

template <typename T>
struct Consts {
    T inv_tau_g;   // 1/tau of the g equation
    T eta_n, M;    // viscosities
    T a, kappa, gamma;
    T psi_wall;
    T gx, gy;      // body force (0 in the reference)
    T psi_visc;    // |psi| bound of the viscosity interpolation (huge: none)
};

// psi on the node and its 8 neighbours in D2Q9 order:
// n[0]=C n[1]=E n[2]=N n[3]=W n[4]=S n[5]=NE n[6]=NW n[7]=SW n[8]=SE
template <typename T>
FD_HD void stencil(const T n[9], T& dx, T& dy, T& lap) {
    dx = (T(4) * (n[1] - n[3]) + n[5] - n[6] - n[7] + n[8]) * (T(1) / T(12));
    dy = (T(4) * (n[2] - n[4]) + n[5] + n[6] - n[7] - n[8]) * (T(1) / T(12));
    lap = (T(-20) * n[0] + T(4) * (n[2] + n[4] + n[1] + n[3]) + n[5] + n[6] + n[7] + n[8]) *
          (T(1) / T(6));
}

template <typename T>
FD_HD T chem_pot(T psi, T lap, const Consts<T>& c) {
    return c.a * psi * (T(1) - psi * psi) - c.kappa * lap;
}

template <typename T>
FD_HD T inv_tau_mix(T rho, T psi, const Consts<T>& c) {
    // 2 nu1 nu2 / (nu1 (1-psi) + nu2 (1+psi)) with nu1 = eta/rho, nu2 = M eta/rho
    // = 2 M eta / (rho ((1-psi) + M (1+psi))).  The denominator vanishes at
    // psi = -(1+M)/(M-1) (-1.14 for M = 15), which overshoots past -1 near
    // wetting solids reach; with c.psi_visc = 1 the interpolation uses psi
    // clamped to the pure phases.  With the default huge bound both
    // comparisons fail and psi passes unchanged.
    const T pv = psi < -c.psi_visc ? -c.psi_visc : (psi > c.psi_visc ? c.psi_visc : psi);
    const T tau = T(6) * c.M * c.eta_n * rcp(rho * ((T(1) - pv) + c.M * (T(1) + pv))) + T(0.5);
    return rcp(tau);
}

// tau_mix - 1/2 = 6 M eta / (rho ((1-psi) + M (1+psi))), with the same psi
// bound as inv_tau_mix: the health term of the watchdog.  BGK needs
// tau_mix > 1/2; a value <= 0 (or infinite) means psi has crossed the pole
// of the interpolation at psi = -(1+M)/(M-1) (fingering_periodic.py:201-208).
template <typename T>
FD_HD T tau_excess(T rho, T psi, const Consts<T>& c) {
    const T pv = psi < -c.psi_visc ? -c.psi_visc : (psi > c.psi_visc ? c.psi_visc : psi);
    return T(6) * c.M * c.eta_n * rcp(rho * ((T(1) - pv) + c.M * (T(1) + pv)));
}

// Moments the collision of one node uses.
template <typename T>
struct Mom {
    T rho, psi, mu, ux, uy, p, itau, fx, fy;
};

// Moments from the node's post-stream populations (fingering_periodic.py:470-479).
template <typename T>
FD_HD Mom<T> moments(const T f[9], T psi, T dpx, T dpy, T lap, const Consts<T>& c) {
    Mom<T> m;
    m.rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    const T jx = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    const T jy = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    m.psi = psi;
    m.mu = chem_pot(psi, lap, c);
    m.fx = m.mu * dpx + c.gx;
    m.fy = m.mu * dpy + c.gy;
    const T ir = rcp(m.rho);
    m.ux = (jx + T(0.5) * m.fx) * ir;
    m.uy = (jy + T(0.5) * m.fy) * ir;
    m.p = m.rho * (T(1) / T(3)) + psi * m.mu;
    m.itau = inv_tau_mix(m.rho, psi, c);
    return m;
}

template <typename T>
FD_HD void equilibria(const Mom<T>& m, const Consts<T>& c, T feq[9], T geq[9]) {
    const T usq = m.ux * m.ux + m.uy * m.uy;
    const T A = T(3) * m.p;
    const T B = T(3) * c.gamma * m.mu;
    // (x - 3 (1 - w0) y) / w0 with w0 = 4/9: (x - 5/3 y) * 9/4
    const T A0 = (m.rho - T(5) / T(3) * m.p) * T(2.25);
    const T B0 = (m.psi - T(5) / T(3) * c.gamma * m.mu) * T(2.25);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T eu = T(EX[i]) * m.ux + T(EY[i]) * m.uy;
        const T q = T(3) * eu + T(4.5) * eu * eu - T(1.5) * usq;
        const T w = W9<T>::w(i);
        feq[i] = w * ((i == 0 ? A0 : A) + m.rho * q);
        geq[i] = w * ((i == 0 ? B0 : B) + m.psi * q);
    }
}

// In-place BGK collision of both distributions with Guo forcing, written
// over the direction pairs (i, opp(i)) that share |e.u| and |e.F|:
//   f*_i = f_i - (f_i - feq_i)/tau_mix + F_i,  g*_i = g_i - (g_i - geq_i)/tau
//   feq_i = w_i (A_i + rho Q_i),  geq_i = w_i (B_i + psi Q_i)
//   Q_{i,opp} = S +- 3 eu,             S = 4.5 eu^2 - 1.5 u.u
//   F_{i,opp} = w_i fac (T +- 3 eF),   T = 9 eu eF - 3 u.F
template <typename T>
FD_HD void collide(T f[9], T g[9], const Mom<T>& m, const Consts<T>& c) {
    const T usq = m.ux * m.ux + m.uy * m.uy;
    const T uF = m.ux * m.fx + m.uy * m.fy;
    const T it = m.itau, ig = c.inv_tau_g;
    const T fac = T(1) - T(0.5) * it;
    const T A = T(3) * m.p;
    const T B = T(3) * c.gamma * m.mu;
    // rest direction: Q0 = -1.5 u.u, F_0 = w0 fac (-3 u.F)
    {
        const T A0 = (m.rho - T(5) / T(3) * m.p) * T(2.25);
        const T B0 = (m.psi - T(5) / T(3) * c.gamma * m.mu) * T(2.25);
        const T q0 = T(-1.5) * usq;
        const T w = W9<T>::w0;
        const T feq = w * (A0 + m.rho * q0);
        const T geq = w * (B0 + m.psi * q0);
        f[0] = f[0] - (f[0] - feq) * it + w * fac * (T(-3) * uF);
        g[0] = g[0] - (g[0] - geq) * ig;
    }
    // pairs (1,3) (2,4) (5,7) (6,8): e = (1,0) (0,1) (1,1) (-1,1)
    constexpr int PI[4] = {1, 2, 5, 6};
    constexpr int PO[4] = {3, 4, 7, 8};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = PI[k], o = PO[k];
        const T ex = T(EX[i]), ey = T(EY[i]);
        const T w = W9<T>::w(i);
        const T eu = ex * m.ux + ey * m.uy;
        const T eF = ex * m.fx + ey * m.fy;
        const T S = T(4.5) * eu * eu - T(1.5) * usq;
        const T a3 = T(3) * eu;
        const T Fs = w * fac * (T(9) * eu * eF - T(3) * uF);
        const T Fa = w * fac * (T(3) * eF);
        const T fp = w * (A + m.rho * (S + a3)), fm = w * (A + m.rho * (S - a3));
        const T gp = w * (B + m.psi * (S + a3)), gm = w * (B + m.psi * (S - a3));
        f[i] = f[i] - (f[i] - fp) * it + (Fs + Fa);
        f[o] = f[o] - (f[o] - fm) * it + (Fs - Fa);
        g[i] = g[i] - (g[i] - gp) * ig;
        g[o] = g[o] - (g[o] - gm) * ig;
    }
}

// ---------------------------------------------------------------- Zou-He
// Inlet density from the known populations (fingering_periodic.py:276-277).
template <typename T>
FD_HD T inlet_rho(const T f[9], T u, T px) {
    return (f[0] + f[2] + f[4] + T(2) * (f[3] + f[6] + f[7]) - T(0.5) * px) * rcp(T(1) - u);
}

// Inlet, rows with the standard formulas (fingering_periodic.py:276-302).
template <typename T>
FD_HD void inlet_f(T f[9], T u, T px, T py) {
    const T r = inlet_rho(f, u, px);
    const T ur = u * r;
    f[1] = f[3] + (T(2) / T(3)) * ur - px * (T(1) / T(6));
    f[5] = f[7] - T(0.5) * (f[2] - f[4]) + ur * (T(1) / T(6)) - px * (T(1) / T(6)) - T(0.25) * py;
    f[8] = f[6] + T(0.5) * (f[2] - f[4]) + ur * (T(1) / T(6)) - px * (T(1) / T(6)) + T(0.25) * py;
}

template <typename T>
FD_HD void inlet_g(T g[9]) {
    // w_k psi_in / (w1 + w5 + w8): w1/(1/6) = 2/3, w5/(1/6) = 1/6
    const T s = T(1) - (g[0] + g[2] + g[3] + g[4] + g[6] + g[7]);
    g[1] = (T(2) / T(3)) * s;
    g[5] = (T(1) / T(6)) * s;
    g[8] = (T(1) / T(6)) * s;
}

// Walled inlet corner nodes (fingering.py:334-364); rho_nb is the inlet
// density of the adjacent row (row 1 for the bottom corner, row H-2 for the
// top corner).
template <typename T>
FD_HD void inlet_corner_bottom_f(T f[9], T rho_nb) {
    f[1] = f[3];
    f[2] = f[4];
    f[5] = f[7];
    f[6] = T(0.5) * (rho_nb - (f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[7]));
    f[8] = f[6];
}

template <typename T>
FD_HD void inlet_corner_bottom_g(T g[9]) {
    // w6 (1 - sum) / (w6 + w8) = (1 - sum) / 2
    g[1] = g[3];
    g[2] = g[4];
    g[5] = g[7];
    g[6] = T(0.5) * (T(1) - (g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[7]));
    g[8] = g[6];
}

template <typename T>
FD_HD void inlet_corner_top_f(T f[9], T rho_nb) {
    f[1] = f[3];
    f[4] = f[2];
    f[8] = f[6];
    f[5] = T(0.5) * (rho_nb - (f[0] + f[1] + f[2] + f[3] + f[4] + f[6] + f[8]));
    f[7] = f[5];
}

template <typename T>
FD_HD void inlet_corner_top_g(T g[9]) {
    g[1] = g[3];
    g[4] = g[2];
    g[8] = g[6];
    g[5] = T(0.5) * (T(1) - (g[0] + g[1] + g[2] + g[3] + g[4] + g[6] + g[8]));
    g[7] = g[5];
}

// Outlet (fingering_periodic.py:305-324).  Velocity outlet: u given, rho
// solved.  Pressure outlet (build-original): rho given, u solved from the
// same mass balance.  c3 is 2/3 (fingering_periodic.py:317) or 1.5
// (fingering.py:378).
template <typename T>
FD_HD void outlet_f(T f[9], T u, T px, T py, bool pressure, T rho_out, T c3) {
    const T known = f[0] + f[2] + f[4] + T(2) * (f[1] + f[5] + f[8]) + T(0.5) * px;
    T r;
    if (pressure) {
        r = rho_out;
        u = known * rcp(rho_out) - T(1);
    } else {
        r = known * rcp(T(1) + u);
    }
    const T ur = u * r;
    f[3] = f[1] - c3 * ur + px * (T(1) / T(6));
    f[6] = f[8] - T(0.5) * (f[2] - f[4]) - ur * (T(1) / T(6)) + T(0.25) * py + px * (T(1) / T(6));
    f[7] = f[5] + T(0.5) * (f[2] - f[4]) - ur * (T(1) / T(6)) - T(0.25) * py + px * (T(1) / T(6));
}

template <typename T>
FD_HD void outlet_g(T g[9]) {
    const T s = T(-1) - (g[0] + g[1] + g[2] + g[4] + g[5] + g[8]);
    g[3] = (T(2) / T(3)) * s;
    g[6] = (T(1) / T(6)) * s;
    g[7] = (T(1) / T(6)) * s;
}

}  // namespace fdn
