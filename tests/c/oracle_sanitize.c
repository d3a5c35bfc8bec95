/*
 * Drives oracle/c/lbm64.c through every boundary configuration it has
 * (periodic, walls, Zou-He inlet with constant or Gaussian profile, velocity
 * or pressure outlet, walled initialisation, the tau_mix psi bound) and
 * every exported call, on small odd-sized domains.  Built by
 * tests/test_oracle_sanitize.py with gcc's address and undefined-behaviour
 * sanitizers, which abort on the first report.
 *
 * Prints one line per case: "case <k> fluid <n> mass0 <sum rho> nonfinite <n>"
 * where mass0 is the density sum after the first stretch of iterations
 * (before the perturbation), which a fully periodic case must keep at n * rho0.
 */
#include "../../oracle/c/lbm64.c"
#include <stdio.h>

static unsigned long long rs = 88172645463325252ULL;
static double rnd(void) {
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return (double)(rs >> 11) / 9007199254740992.0;
}

static int run(int k, int H, int W, int y_walls, int x_periodic, int gauss, int pressure, int init_walled,
               double bound, int steps) {
    /* lattice_boltzmann/fingering_periodic.py:26-40 values (Pe 15, xi 2), M 15 */
    o64_params p;
    memset(&p, 0, sizeof p);
    p.H = H;
    p.W = W;
    p.tau = 0.7886751345948129;
    p.eta_n = 0.0962250448649376;
    p.M = 15;
    p.a = -0.046875;
    p.kappa = 0.09375;
    p.gamma = 0.93824;
    p.psi_wall = -0.5;
    p.u0 = 1e-3;
    p.rho_out = 1.0;
    p.c3 = 2.0 / 3;
    p.rho0 = 1.0;
    p.y_walls = y_walls;
    p.x_periodic = x_periodic;
    p.gauss = gauss;
    p.pressure = pressure;
    p.init_walled = init_walled;
    p.wall_lo = 0;
    p.wall_hi = H - 1;
    p.psi_visc_bound = bound;
    size_t N = (size_t)H * W;
    unsigned char *fl = malloc(N), *ln = calloc(9 * N, 1);
    double* psi = malloc(N * sizeof(double));
    int nfl = 0;
    for (size_t n = 0; n < N; ++n) {
        fl[n] = rnd() > 0.05;
        nfl += fl[n];
        psi[n] = tanh(((double)W / 3 - (double)(n % W)) / 2.0);
    }
    /* a link on every fluid node whose upstream neighbour (periodic) is solid */
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            for (int i = 1; i < 9; ++i) {
                int ys = ((y - EY[i]) % H + H) % H, xs = ((x - EX[i]) % W + W) % W;
                ln[(size_t)i * N + (size_t)y * W + x] = fl[(size_t)y * W + x] && !fl[(size_t)ys * W + xs];
            }
    void* h = o64_create(&p, fl, ln, psi, NULL);
    double m0[2], m1[2];
    o64_step(h, steps);
    o64_mass(h, m0);
    o64_perturb(h, 1e-7, 3);
    o64_round32(h);
    o64_step(h, 2);
    double *a = malloc(N * sizeof(double)), *b = malloc(N * sizeof(double)), *c = malloc(N * sizeof(double)),
           *d = malloc(N * sizeof(double)), *f = malloc(9 * N * sizeof(double)), *g = malloc(9 * N * sizeof(double));
    o64_fields(h, a, b, c, d);
    o64_pops(h, f, g);
    o64_mass(h, m1);
    o64_destroy(h);
    int bad = 0;
    for (size_t n = 0; n < N; ++n) bad += !isfinite(a[n]) || !isfinite(b[n]) || !isfinite(c[n]) || !isfinite(d[n]);
    for (size_t n = 0; n < 9 * N; ++n) bad += !isfinite(f[n]) || !isfinite(g[n]);
    bad += !isfinite(m1[0]) || !isfinite(m1[1]);
    printf("case %d fluid %d mass0 %.17g nonfinite %d\n", k, nfl, m0[0], bad);
    free(fl);
    free(ln);
    free(psi);
    free(a);
    free(b);
    free(c);
    free(d);
    free(f);
    free(g);
    return bad;
}

int main(void) {
    int bad = 0;
    bad += run(0, 5, 7, 0, 1, 0, 0, 0, 0.0, 30);    /* periodic in both axes */
    bad += run(1, 3, 9, 1, 0, 1, 1, 1, 0.0, 30);    /* walls, Gaussian inlet, pressure outlet, H = 3 */
    bad += run(2, 16, 24, 1, 0, 0, 0, 1, 1.0, 40);  /* walls, constant inlet, velocity outlet, psi bound */
    bad += run(3, 16, 24, 0, 0, 1, 1, 0, 0.0, 40);  /* periodic y, Zou-He in x */
    bad += run(4, 17, 31, 1, 1, 0, 0, 1, 0.9, 40);  /* walls, periodic x */
    bad += run(5, 1, 1, 0, 1, 0, 0, 0, 0.0, 5);     /* one node */
    return bad != 0;
}
