"""Oracle restatement of the physical -> lattice unit conversion (TEST INFRASTRUCTURE ONLY).

Follows lattice_boltzmann/fingering_periodic.py:15-44 (periodic variant) and
lattice_boltzmann/fingering.py:16-50 (walled variant, whose gamma uses the
channel height H instead of the fixed length 20, fingering.py:43).

Pinned by: the closed-form values printed in the tests (tau = 1/(3-sqrt 3),
nu0 = (tau-1/2)/3, sigma = 0.0625, kappa = 0.75*sigma*xi, a = -2 kappa/xi^2)
and the identity Eta_n == nu0 that the reference's unit system implies when
C_rho = 1000 and Eta = 1e-3 (Eta_n = Eta*C_t/(C_rho*C_W^2) = R_Nu*C_t/C_W^2 = v0).
"""
import math


def derived_constants(Pe=15.0, C_W=1.5e-7, Ca=2.0 * 7.33e-3, M=20.0, Eta=0.001,
                      R_sigma=0.045, C_rho=1.0e3, rho0=1.0, xi=2.0,
                      gamma_length=20.0):
    """Return the lattice constants of fingering_periodic.py:15-44.

    gamma_length is the length in ``gamma = u0*L/((-a*Pe)*(tau-0.5))``:
    20 in fingering_periodic.py:40, H in fingering.py:43.
    """
    R_Nu = Eta / 1000.0                                   # :25
    tau = 1.0 / (3.0 - math.sqrt(3.0))                    # :26
    v0 = (tau - 0.5) / 3.0                                # :32
    C_t = v0 / R_Nu * (C_W ** 2)                          # :33
    Eta_n = Eta / (C_rho * (C_W ** 2) / C_t)              # :34
    sigma = R_sigma * (C_t ** 2) / (C_rho * (C_W ** 3))   # :35
    u0 = Ca * sigma / (rho0 * v0)                         # :36
    kappa = 0.75 * sigma * xi                             # :38
    a = -2.0 * kappa / (xi ** 2)                          # :39
    gamma = u0 * gamma_length / ((-a * Pe) * (tau - 0.5))  # :40
    return dict(tau=tau, eta_n=Eta_n, M=M, kappa=kappa, a=a, gamma=gamma,
                u0=u0, sigma=sigma, rho0=rho0, xi=xi)
