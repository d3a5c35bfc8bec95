"""FP64 numpy oracle of the two-phase free-energy D2Q9 step (TEST INFRASTRUCTURE ONLY).

Op-for-op restatement of lattice_boltzmann/fingering_periodic.py (the hot
loop :454-479 and everything it calls) and, behind flags, of the walled
sibling lattice_boltzmann/fingering.py.  It keeps the reference's cost
structure on purpose (compressed fluid arrays gathered with a boolean mask,
``np.roll`` streaming, deep copies of the post-collision populations), so the
same code is both the parity checker and the timed CPU baseline ("port").

Per-step order (fingering_periodic.py:455-479):
    collide (F, feq, geq, f*, g*)  ->  f_behind/g_behind copies
    ->  stream (np.roll, periodic in both axes)
    ->  half-way bounce-back on the per-direction link masks
    ->  [walled] top/bottom half-way walls (fingering.py:573)
    ->  Zou-He inlet (x = 0)  ->  Zou-He outlet (x = W-1)
    ->  rho, psi, grad psi, lap psi, mu, ux, uy, p, tau_mix

Quirks kept (each is exercised by a test):
  * init with mu = 0 (fingering_periodic.py:111-121) so the first collision
    is the identity; the walled init computes mu and uy first
    (fingering.py:121-127).
  * Zou-He reads grad psi and mu_plain of the *previous* step's psi
    (fingering_periodic.py:273-275, 309-311) and covers whole columns.
  * solid nodes are never collided but are streamed (store-and-forward).
  * stencils see psi_wall on solids, ghost columns +1 (x=-1) and -1 (x=W),
    periodic y (fingering_periodic.py:94-95, 214-256); the walled variant
    adds psi_wall ghost rows (fingering.py:224, 245, 266).

Build-original extensions, flagged and off by default: periodic x (no ghost
columns, no Zou-He), a constant body force G added to the interfacial force
mu*grad psi, and a pressure (density) outlet.

Pinning (tests/test_oracle_*.py): D2Q9 moment identities; equilibrium
moment invariants; Guo forcing moments; per-node conservation of sum g;
np.roll brute force on a 5x7 grid; stencils exact on quadratics and on
brute-force ghost/wrap cases; Zou-He closed-form invariants (sum f = rho_in,
momentum balance with the P/2 correction, sum g = +1 / -1); tau_mix at
psi = +-1; Poiseuille profile for the body-force channel.
"""
import copy

import numpy as np

E = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
              [1, 1], [-1, 1], [-1, -1], [1, -1]])          # fingering_periodic.py:50-89
WGT = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9,
                1 / 36, 1 / 36, 1 / 36, 1 / 36])
OPP = (0, 3, 4, 1, 2, 7, 8, 5, 6)                            # bounce_back.py:103-167


def stream(f, g):
    """fingering_periodic.py:327-343: push by np.roll, periodic in both axes."""
    for arr in (f, g):
        arr[1] = np.roll(arr[1], 1, axis=1)
        arr[2] = np.roll(arr[2], 1, axis=0)
        arr[3] = np.roll(arr[3], -1, axis=1)
        arr[4] = np.roll(arr[4], -1, axis=0)
        arr[5] = np.roll(np.roll(arr[5], 1, axis=1), 1, axis=0)
        arr[6] = np.roll(np.roll(arr[6], -1, axis=1), 1, axis=0)
        arr[7] = np.roll(np.roll(arr[7], -1, axis=1), -1, axis=0)
        arr[8] = np.roll(np.roll(arr[8], 1, axis=1), -1, axis=0)


def bounce_back(links, f_behind, g_behind, f, g):
    """Half-way bounce-back (bounce_back.py:25-167): for each direction i,
    overwrite f[i], g[i] on its link mask with the same node's post-collision
    opposite population."""
    for i in range(1, 9):
        m = links[i]
        f[i][m] = f_behind[OPP[i]][m]
        g[i][m] = g_behind[OPP[i]][m]


def bottom_top_wall(f_behind, g_behind, f, g):
    """fingering_periodic.py:354-366 / fingering.py:432-451."""
    f[2][0, :] = f_behind[4][0, :]
    g[2][0, :] = g_behind[4][0, :]
    f[4][-1, :] = f_behind[2][-1, :]
    g[4][-1, :] = g_behind[2][-1, :]
    f[5][0, :] = f_behind[7][0, :]
    g[5][0, :] = g_behind[7][0, :]
    f[6][0, :] = f_behind[8][0, :]
    g[6][0, :] = g_behind[8][0, :]
    f[7][-1, :] = f_behind[5][-1, :]
    g[7][-1, :] = g_behind[5][-1, :]
    f[8][-1, :] = f_behind[6][-1, :]
    g[8][-1, :] = g_behind[6][-1, :]


def inlet_profile(H, u0, kind):
    """Inlet/outlet velocity column: Gaussian (fingering_periodic.py:270-271)
    or constant (fingering.py:299)."""
    if kind == "gauss":
        temp = np.array([i * 3 / (H / 2) for i in range(int(-H / 2), int(H / 2))])
        return u0 * np.exp(-(temp ** 2) / 2)
    return np.full(H, float(u0))


class Lbm64:
    """Reference-shaped FP64 solver state (``Compute`` of fingering_periodic.py:47-324).

    mask    bool[H][W], True on fluid nodes (fingering_periodic.py:450-452)
    links   bool[9][H][W] per-direction half-way bounce-back masks
    y_walls   False: periodic y (fingering_periodic.py); True: half-way
              walls (bottom_top_wall), psi_wall ghost rows, inlet on rows
              1:-1 plus corner nodes (fingering.py:298-390, 573)
    wall_rows "edge": bottom_top_wall on the full stacks, i.e. rows 0 and H-1
              (the function as written, fingering.py:432-451); "shipped":
              the call exactly as fingering.py:573 makes it, whose [:, 1:-1]
              slices rows, so the walls act on rows 1 and H-2
    x_periodic  build-original: periodic x, no ghost columns, no Zou-He
    init      "mu0" (fingering_periodic.py:111-121) or "walled" (fingering.py:121-127)
    """

    def __init__(self, mask, links, *, tau, eta_n, M, kappa, a, gamma, u0, psi_wall,
                 rho0=1.0, n_inv_cols=5, inlet="gauss", outlet="velocity",
                 outlet_f3_coef=2.0 / 3.0, y_walls=False, x_periodic=False,
                 init="mu0", body_force=(0.0, 0.0), rho_out=1.0, psi0=None, rho_init=None,
                 wall_rows="edge", psi_visc_bound=0.0):
        H, W = mask.shape
        self.psi_visc_bound = float(psi_visc_bound)
        self.H, self.W = H, W
        self.mask = mask
        self.block_mask = np.logical_not(mask)
        self.links = links
        self.tau, self.eta_n, self.M = tau, eta_n, M
        self.kappa, self.a, self.gamma = kappa, a, gamma
        self.u0, self.psi_wall = u0, psi_wall
        self.inlet, self.outlet = inlet, outlet
        self.outlet_f3_coef = outlet_f3_coef
        self.y_walls, self.x_periodic = y_walls, x_periodic
        self.wall_rows = wall_rows
        self.G = (float(body_force[0]), float(body_force[1]))
        self.rho_out = rho_out
        self.e, self.w = E, WGT
        # fingering_periodic.py:90-95
        if psi0 is None:
            self.psi = np.full((H, W), -1.0)
            self.psi[:, :n_inv_cols] = 1.0
        else:
            self.psi = np.array(psi0, dtype=np.float64, copy=True)
        self.psi[self.block_mask] = psi_wall
        self.left_wall = np.full((H, 1), 1.0)
        self.right_wall = np.full((H, 1), -1.0)
        self.top_bottom_wall = np.full((1, W if x_periodic else W + 2), psi_wall)
        # :101-113
        if rho_init is None:
            self.rho = np.ones((H, W))[mask] * rho0
        else:
            self.rho = np.asarray(rho_init, dtype=np.float64)[mask]
        self.ux = np.zeros((H, W))[mask]
        self.uy = np.zeros((H, W))[mask]
        self.f = np.array([np.zeros((H, W)) for _ in range(9)])
        self.g = np.array([np.zeros((H, W)) for _ in range(9)])
        self.feq = np.array([np.zeros((H, W))[mask] for _ in range(9)])
        self.geq = np.array([np.zeros((H, W))[mask] for _ in range(9)])
        self.mu = np.zeros((H, W))[mask]
        self.F = np.array([np.zeros((H, W))[mask] for _ in range(9)])
        # :114-121 (mu stays 0 -> p = rho/3, g0 = psi, g_i = 0)
        self.nabla_psix = self.getNabla_psix()
        self.nabla_psiy = self.getNabla_psiy()
        self.nabla_psi2 = self.getNabla_psi2()
        self.p = self.getP()
        if init == "walled":                                  # fingering.py:122-123
            self.mu = self.getMu()
            self.uy = self.getUy()
        self.mix_tau = self.getMix_tau()
        for i in range(9):
            self.f[i][self.mask] = self.getfeq(i)
            self.g[i][self.mask] = self.getgeq(i)
        self.t = 0

    # ---------------------------------------------------------- moments
    def getP(self):                                           # :123-124
        return 1 / 3 * self.rho + self.psi[self.mask] * self.mu

    def _force(self):
        """Total force per node: mu * grad psi (+ G, build-original)."""
        fx = self.mu * self.nabla_psix[self.mask]
        fy = self.mu * self.nabla_psiy[self.mask]
        return fx + self.G[0], fy + self.G[1]

    def getUx(self):                                          # :126-131
        temp = np.zeros((self.H, self.W))[self.mask]
        for i in range(9):
            temp += self.f[i][self.mask] * self.e[i][0]
        if self.G[0] == 0.0:
            return (temp + self.mu * self.nabla_psix[self.mask] / 2) / self.rho
        return (temp + self._force()[0] / 2) / self.rho

    def getUy(self):                                          # :134-139
        temp = np.zeros((self.H, self.W))[self.mask]
        for i in range(9):
            temp += self.f[i][self.mask] * self.e[i][1]
        if self.G[1] == 0.0:
            return (temp + self.mu * self.nabla_psiy[self.mask] / 2) / self.rho
        return (temp + self._force()[1] / 2) / self.rho

    def getMu(self):                                          # :141-144
        ones = np.ones((self.H, self.W))[self.mask]
        return (self.a * self.psi[self.mask] * (ones - np.power(self.psi[self.mask], 2))
                - self.kappa * self.nabla_psi2[self.mask])

    def getMu_plain(self):                                    # :146-149
        ones = np.ones((self.H, self.W))
        return self.a * self.psi * (ones - np.power(self.psi, 2)) - self.kappa * self.nabla_psi2

    def getRho(self):                                         # :151-152
        return np.sum(self.f, axis=0)[self.mask]

    def getA0(self):                                          # :155-157
        return (self.rho - 3.0 * (1.0 - self.w[0]) * self.p) / self.w[0]

    def getA1_8(self):                                        # :159-161
        return 3 * self.p

    def getB0(self):                                          # :163-165
        return (self.psi[self.mask] - 3.0 * (1.0 - self.w[0]) * self.gamma * self.mu) / self.w[0]

    def getB1_8(self):                                        # :167-169
        return 3 * self.gamma * self.mu

    def _quad(self, n):
        eu = self.e[n][0] * self.ux + self.e[n][1] * self.uy
        return 3 * eu + 4.5 * eu ** 2 - 1.5 * (self.ux ** 2 + self.uy ** 2)

    def getfeq(self, n):                                      # :171-181
        A = self.getA0() if n == 0 else self.getA1_8()
        return self.w[n] * (A + self.rho * self._quad(n))

    def getgeq(self, n):                                      # :183-192
        B = self.getB0() if n == 0 else self.getB1_8()
        return self.w[n] * (B + self.psi[self.mask] * self._quad(n))

    def getLarge_F(self, n):                                  # :194-199
        ex, ey = self.e[n]
        eu = ex * self.ux + ey * self.uy
        if self.G == (0.0, 0.0):
            return (self.mu * self.w[n] * (1 - 1 / (2 * self.mix_tau))
                    * (((ex - self.ux) * 3 + ex * eu * 9) * self.nabla_psix[self.mask]
                       + ((ey - self.uy) * 3 + ey * eu * 9) * self.nabla_psiy[self.mask]))
        fx, fy = self._force()
        return (self.w[n] * (1 - 1 / (2 * self.mix_tau))
                * (((ex - self.ux) * 3 + ex * eu * 9) * fx + ((ey - self.uy) * 3 + ey * eu * 9) * fy))

    def getMix_tau(self):                                     # :201-208
        ones = np.ones((self.H, self.W))[self.mask]
        v1 = self.eta_n / self.rho
        v2 = self.eta_n * self.M / self.rho
        psi = self.psi[self.mask]
        if self.psi_visc_bound > 0:
            # build-original option: the interpolation sees psi in [-b, b]
            psi = np.clip(psi, -self.psi_visc_bound, self.psi_visc_bound)
        mix_v = np.divide(2 * v1 * v2, (v1 * (ones - psi) + v2 * (ones + psi)))
        return 3 * mix_v + 0.5

    def udpatePsi(self):                                      # :210-212
        self.psi[self.mask] = np.sum(self.g, axis=0)[self.mask]
        self.psi[self.block_mask] = self.psi_wall

    # --------------------------------------------------------- stencils
    def _padded_psi(self):
        """psi with psi_wall on solids, ghost columns (+1 | -1) unless x is
        periodic, and psi_wall ghost rows when the channel is walled."""
        psi_with_block = copy.deepcopy(self.psi)
        psi_with_block[self.block_mask] = self.psi_wall
        if self.x_periodic:
            temp = psi_with_block
        else:
            temp = np.hstack((self.left_wall, np.hstack((psi_with_block, self.right_wall))))
        if self.y_walls:
            temp = np.vstack((self.top_bottom_wall, np.vstack((temp, self.top_bottom_wall))))
        return temp

    def _crop(self, arr):
        rs = slice(1, -1) if self.y_walls else slice(None)
        cs = slice(None) if self.x_periodic else slice(1, -1)
        return arr[rs, cs]

    def getNabla_psix(self):                                  # :214-226
        temp = self._padded_psi()
        r = np.roll
        f = np.zeros((self.H, self.W))
        f += 4 * self._crop(r(temp, -1, axis=1))
        f += -4 * self._crop(r(temp, 1, axis=1))
        f += self._crop(r(r(temp, -1, axis=1), -1, axis=0))
        f += -self._crop(r(r(temp, 1, axis=1), -1, axis=0))
        f += -self._crop(r(r(temp, 1, axis=1), 1, axis=0))
        f += self._crop(r(r(temp, -1, axis=1), 1, axis=0))
        return f / 12

    def getNabla_psiy(self):                                  # :228-240
        temp = self._padded_psi()
        r = np.roll
        f = np.zeros((self.H, self.W))
        f += 4 * self._crop(r(temp, -1, axis=0))
        f += -4 * self._crop(r(temp, 1, axis=0))
        f += self._crop(r(r(temp, -1, axis=1), -1, axis=0))
        f += self._crop(r(r(temp, 1, axis=1), -1, axis=0))
        f += -self._crop(r(r(temp, 1, axis=1), 1, axis=0))
        f += -self._crop(r(r(temp, -1, axis=1), 1, axis=0))
        return f / 12

    def getNabla_psi2(self):                                  # :242-256
        temp = self._padded_psi()
        r = np.roll
        f = np.zeros((self.H, self.W))
        f += -20 * self._crop(temp)
        f += 4 * self._crop(r(temp, -1, axis=0))
        f += 4 * self._crop(r(temp, -1, axis=1))
        f += 4 * self._crop(r(temp, 1, axis=1))
        f += 4 * self._crop(r(temp, 1, axis=0))
        f += self._crop(r(r(temp, -1, axis=1), -1, axis=0))
        f += self._crop(r(r(temp, 1, axis=1), -1, axis=0))
        f += self._crop(r(r(temp, 1, axis=1), 1, axis=0))
        f += self._crop(r(r(temp, -1, axis=1), 1, axis=0))
        return f / 6

    # -------------------------------------------------------- collision
    def getF(self, i):                                        # :258-260
        return self.f[i][self.mask] - 1.0 / self.mix_tau * (self.f[i][self.mask] - self.feq[i]) + self.F[i]

    def getG(self, i):                                        # :262-264
        return self.g[i][self.mask] - 1.0 / self.tau * (self.g[i][self.mask] - self.geq[i])

    def collide(self):                                        # :455-460
        for j in range(9):
            self.F[j] = self.getLarge_F(j)
            self.feq[j] = self.getfeq(j)
            self.geq[j] = self.getgeq(j)
            self.f[j][self.mask] = self.getF(j)
            self.g[j][self.mask] = self.getG(j)

    # --------------------------------------------------------- Zou-He
    def _zou_he_P(self, col):
        """P_x = d_x psi * mu_plain, P_y = d_y psi * mu_plain on one column,
        from the not-yet-updated psi (fingering_periodic.py:273-275)."""
        psi_x = self.getNabla_psix()[:, col]
        psi_y = self.getNabla_psiy()[:, col]
        mu = self.getMu_plain()[:, col]
        return psi_x * mu, psi_y * mu

    def zou_he_boundary_inlet(self):
        """fingering_periodic.py:268-302 (all rows) or fingering.py:298-364
        (rows 1:-1 plus the two corner nodes) when the channel is walled."""
        f, g, w = self.f, self.g, self.w
        ux = inlet_profile(self.H, self.u0, self.inlet)
        psi_x = self.getNabla_psix()
        psi_y = self.getNabla_psiy()
        mu = self.getMu_plain()
        Px = psi_x[:, 0] * mu[:, 0]
        Py = psi_y[:, 0] * mu[:, 0]
        rho_in = 1 / (1 - ux) * (f[0][:, 0] + f[2][:, 0] + f[4][:, 0]
                                 + 2 * (f[3][:, 0] + f[6][:, 0] + f[7][:, 0]) - Px / 2)
        psi_in = 1.0 - (g[0][:, 0] + g[2][:, 0] + g[3][:, 0] + g[4][:, 0] + g[6][:, 0] + g[7][:, 0])
        rows = slice(1, -1) if self.y_walls else slice(None)
        wsum = w[1] + w[5] + w[8]
        u, R, px, py = ux[rows], rho_in[rows], Px[rows], Py[rows]
        f[1][rows, 0] = f[3][rows, 0] + 2 / 3 * u * R - px / 6
        g[1][rows, 0] = w[1] * psi_in[rows] / wsum
        f[5][rows, 0] = f[7][rows, 0] - 0.5 * (f[2][rows, 0] - f[4][rows, 0]) + 1.0 / 6.0 * u * R - px / 6 - py / 4
        g[5][rows, 0] = w[5] * psi_in[rows] / wsum
        f[8][rows, 0] = f[6][rows, 0] + 0.5 * (f[2][rows, 0] - f[4][rows, 0]) + 1.0 / 6.0 * u * R - px / 6 + py / 4
        g[8][rows, 0] = w[8] * psi_in[rows] / wsum
        if not self.y_walls:
            return
        # left bottom corner node (fingering.py:334-348)
        f[1][0, 0] = f[3][0, 0]
        g[1][0, 0] = g[3][0, 0]
        f[2][0, 0] = f[4][0, 0]
        g[2][0, 0] = g[4][0, 0]
        f[5][0, 0] = f[7][0, 0]
        g[5][0, 0] = g[7][0, 0]
        f[6][0, 0] = 0.5 * (rho_in[1] - (f[0][0, 0] + f[1][0, 0] + f[2][0, 0] + f[3][0, 0]
                                         + f[4][0, 0] + f[5][0, 0] + f[7][0, 0]))
        g[6][0, 0] = w[6] * (1.0 - (g[0][0, 0] + g[1][0, 0] + g[2][0, 0] + g[3][0, 0]
                                    + g[4][0, 0] + g[5][0, 0] + g[7][0, 0])) / (w[6] + w[8])
        f[8][0, 0] = f[6][0, 0]
        g[8][0, 0] = g[6][0, 0]
        # left top corner node (fingering.py:350-364)
        f[1][-1, 0] = f[3][-1, 0]
        g[1][-1, 0] = g[3][-1, 0]
        f[4][-1, 0] = f[2][-1, 0]
        g[4][-1, 0] = g[2][-1, 0]
        f[8][-1, 0] = f[6][-1, 0]
        g[8][-1, 0] = g[6][-1, 0]
        f[5][-1, 0] = 0.5 * (rho_in[-2] - (f[0][-1, 0] + f[1][-1, 0] + f[2][-1, 0] + f[3][-1, 0]
                                           + f[4][-1, 0] + f[6][-1, 0] + f[8][-1, 0]))
        g[5][-1, 0] = w[5] * (1.0 - (g[0][-1, 0] + g[1][-1, 0] + g[2][-1, 0] + g[3][-1, 0]
                                     + g[4][-1, 0] + g[6][-1, 0] + g[8][-1, 0])) / (w[5] + w[7])
        f[7][-1, 0] = f[5][-1, 0]
        g[7][-1, 0] = g[5][-1, 0]

    def zou_he_boundary_outlet(self):
        """fingering_periodic.py:305-324 (velocity outlet, f3 coefficient 2/3);
        fingering.py:369-390 uses 1.5 and fixes the two corner nodes.  The
        pressure outlet (outlet="pressure") is build-original: rho is
        prescribed and the same relation is solved for u."""
        f, g, w = self.f, self.g, self.w
        psi_x = self.getNabla_psix()[:, -1]
        psi_y = self.getNabla_psiy()[:, -1]
        mu = self.getMu_plain()[:, -1]
        Px = psi_x * mu
        Py = psi_y * mu
        known = f[0][:, -1] + f[2][:, -1] + f[4][:, -1] + 2 * (f[1][:, -1] + f[5][:, -1] + f[8][:, -1]) + Px / 2
        if self.outlet == "pressure":
            rho_o = np.full(self.H, float(self.rho_out))
            ux = known / rho_o - 1.0
        else:
            ux = inlet_profile(self.H, self.u0, self.inlet)
            rho_o = 1 / (1 + ux) * known
        psi_out = -1.0 - (g[0][:, -1] + g[1][:, -1] + g[2][:, -1] + g[4][:, -1] + g[5][:, -1] + g[8][:, -1])
        wsum = w[3] + w[6] + w[7]
        f[3][:, -1] = f[1][:, -1] - self.outlet_f3_coef * ux * rho_o + Px / 6
        g[3][:, -1] = w[3] * psi_out / wsum
        f[6][:, -1] = f[8][:, -1] - 0.5 * (f[2][:, -1] - f[4][:, -1]) - 1.0 / 6.0 * ux * rho_o + Py / 4 + Px / 6
        g[6][:, -1] = w[6] * psi_out / wsum
        f[7][:, -1] = f[5][:, -1] + 0.5 * (f[2][:, -1] - f[4][:, -1]) - 1.0 / 6.0 * ux * rho_o - Py / 4 + Px / 6
        g[7][:, -1] = w[7] * psi_out / wsum
        if self.y_walls:                                      # fingering.py:387-390
            f[2][0, -1] = f[4][0, -1]
            g[2][0, -1] = g[4][0, -1]
            f[4][-1, -1] = f[2][-1, -1]
            g[4][-1, -1] = g[2][-1, -1]

    # ---------------------------------------------------------- driver
    def moments(self):                                        # :470-479
        self.rho = self.getRho()
        self.udpatePsi()
        self.nabla_psix = self.getNabla_psix()
        self.nabla_psiy = self.getNabla_psiy()
        self.nabla_psi2 = self.getNabla_psi2()
        self.mu = self.getMu()
        self.ux = self.getUx()
        self.uy = self.getUy()
        self.p = self.getP()
        self.mix_tau = self.getMix_tau()

    def step(self, n=1):
        """n iterations of the main loop body (fingering_periodic.py:454-479)."""
        for _ in range(n):
            self.collide()
            f_behind = copy.deepcopy(self.f)
            g_behind = copy.deepcopy(self.g)
            stream(self.f, self.g)
            bounce_back(self.links, f_behind, g_behind, self.f, self.g)
            if self.y_walls:
                if self.wall_rows == "shipped":
                    # fingering.py:573 passes f[:, 1:-1] of the (9, H, W) stacks:
                    # that slices ROWS, so the walls act on rows 1 and H-2.
                    bottom_top_wall(f_behind[:, 1:-1], g_behind[:, 1:-1],
                                    self.f[:, 1:-1], self.g[:, 1:-1])
                else:
                    bottom_top_wall(f_behind, g_behind, self.f, self.g)
            if not self.x_periodic:
                self.zou_he_boundary_inlet()
                self.zou_he_boundary_outlet()
            self.moments()
            self.t += 1

    @classmethod
    def from_params(cls, p, mask, links, psi0=None, rho0=None):
        """Build from any object with the attributes of the product's
        parameter record (H, W, tau, eta_n, M, a, kappa, gamma, psi_wall,
        u0, rho0, rho_out, outlet_f3_coef, gx, gy, y_bc, x_bc, profile,
        outlet, init_mode, wall_lo, wall_hi, n_inv_cols) using the integer
        codes of include/fdlbm.h."""
        wall_hi = p.H - 1 if p.wall_hi < 0 else p.wall_hi
        return cls(
            mask, links, tau=p.tau, eta_n=p.eta_n, M=p.M, kappa=p.kappa, a=p.a, gamma=p.gamma,
            u0=p.u0, psi_wall=p.psi_wall, rho0=p.rho0, n_inv_cols=p.n_inv_cols,
            inlet="const" if p.profile == 1 else "gauss",
            outlet="pressure" if p.outlet == 1 else "velocity",
            outlet_f3_coef=p.outlet_f3_coef, y_walls=p.y_bc == 1, x_periodic=p.x_bc == 1,
            init="walled" if p.init_mode == 1 else "mu0", body_force=(p.gx, p.gy),
            rho_out=p.rho_out, psi0=psi0, rho_init=rho0,
            wall_rows="edge" if (p.wall_lo == 0 and wall_hi == p.H - 1) else "shipped",
            psi_visc_bound=getattr(p, "psi_visc_bound", 0.0))

    def set_populations(self, f, g):
        """Replace f and g (post-stream, [9][H][W]) and recompute the moments
        from them: the reference's in-place writes of cm.f / cm.g
        (fingering_periodic.py:459-460, 464-467) followed by its moment
        getters (:470-479).  The next step collides these populations."""
        self.f = np.array(f, dtype=np.float64, copy=True)
        self.g = np.array(g, dtype=np.float64, copy=True)
        self.moments()

    def health(self):
        """(number of fluid nodes whose rho or psi is not finite,
        min over fluid nodes of mix_tau - 1/2): the failure watchdog.  BGK
        needs mix_tau > 1/2 (fingering_periodic.py:201-208, :258-260); the
        reference's only guard is np.seterr(all='raise') (:497)."""
        bad = ~np.isfinite(self.rho) | ~np.isfinite(self.psi[self.mask])
        with np.errstate(invalid="ignore"):
            tex = np.asarray(self.mix_tau, dtype=np.float64) - 0.5
        tex = tex[~np.isnan(tex)]
        return int(bad.sum()), float(tex.min()) if tex.size else float("inf")

    # ---------------------------------------------------------- outputs
    def field(self, name):
        """Full [H][W] view of a compressed moment (0 on solid nodes)."""
        out = np.zeros((self.H, self.W))
        out[self.mask] = getattr(self, name)
        return out

    def mass(self):
        """FP64 sums over fluid nodes: (sum rho, sum psi)."""
        return float(np.sum(self.rho)), float(np.sum(self.psi[self.mask]))
