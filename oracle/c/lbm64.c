/*
 * lbm64.c - FP64 C restatement of the two-phase D2Q9 step (TEST INFRASTRUCTURE ONLY).
 *
 * The same algorithm and step order as oracle/lbm64.py (the numpy
 * restatement of lattice_boltzmann/fingering_periodic.py:454-479 and, behind
 * flags, lattice_boltzmann/fingering.py), written as plain loops so long
 * trajectories (thousands of iterations) finish in seconds.  It is pinned to
 * the numpy oracle node for node in tests/test_oracle_c.py.  Only the tests
 * and the CPU-baseline leg of bench.py load it; the product never does.
 *
 * Per iteration (fingering_periodic.py:455-479):
 *   collide fluid nodes (:455-460) -> copy of the post-collision state
 *   (:464-465) -> stream, periodic in both axes (:327-343) -> half-way
 *   bounce-back on the per-direction link masks (bounce_back.py:25-167) ->
 *   walls (fingering.py:573 / :432-451) -> Zou-He inlet (:268-302) and
 *   outlet (:305-324) -> moments (:470-479).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Arithmetic and storage type: double for the oracle; built a second time
   with -DO64_REAL=float as the FP32 evaluation of the same algorithm (the
   yardstick for how far FP32 arithmetic itself moves the trajectory). */
#ifndef O64_REAL
#define O64_REAL double
#endif
typedef O64_REAL real_t;

static const int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const real_t WT[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

typedef struct {
    int H, W;
    double tau, eta_n, M, a, kappa, gamma, psi_wall, u0, rho_out, c3, gx, gy, rho0;
    int y_walls;     /* 1: walls + psi_wall ghost rows + walled Zou-He corners */
    int x_periodic;  /* 1: periodic x, no ghost columns, no Zou-He */
    int gauss;       /* inlet profile: 1 Gaussian (fingering_periodic.py:270), 0 constant */
    int pressure;    /* outlet: 1 prescribed density, 0 velocity */
    int init_walled; /* 1: mu and uy at init (fingering.py:121-127) */
    int wall_lo, wall_hi;
    double psi_visc_bound; /* > 0: psi clamped to [-b, b] in mix_tau (build-original option) */
} o64_params;

typedef struct {
    o64_params p;
    int H, W, N;
    unsigned char *fluid, *links; /* links: [9][H][W] */
    real_t *f, *g, *fb, *gb;      /* [9][H][W] */
    real_t *psi, *rho, *ux, *uy, *mu, *p_, *tau_mix, *dx, *dy, *lap; /* [H][W] */
    real_t *uprof;                /* [H] */
    long t;
} o64;

#define IDX(y, x) ((size_t)(y) * o->W + (x))
#define POP(a, i, y, x) (a)[(size_t)(i) * o->N + IDX(y, x)]

static int wrapi(int v, int n) { return ((v % n) + n) % n; }

/* psi with psi_wall on solids and the stencil's ghosts */
static real_t psi_at(const o64* o, int y, int x) {
    if (o->p.y_walls && (y < 0 || y >= o->H)) return o->p.psi_wall;
    y = wrapi(y, o->H);
    if (!o->p.x_periodic) {
        if (x < 0) return 1.0;
        if (x >= o->W) return -1.0;
    }
    x = wrapi(x, o->W);
    return o->fluid[IDX(y, x)] ? o->psi[IDX(y, x)] : o->p.psi_wall;
}

static void stencils(o64* o) { /* fingering_periodic.py:214-256 */
    /* psi padded by one ghost ring: psi_wall on solids, +1 / -1 ghost
       columns (or periodic x), psi_wall ghost rows (or periodic y) */
    const int H = o->H, W = o->W, P = W + 2;
    real_t* t = malloc(sizeof(real_t) * (size_t)(H + 2) * P);
    for (int y = -1; y <= H; ++y)
        for (int x = -1; x <= W; ++x) t[(size_t)(y + 1) * P + (x + 1)] = psi_at(o, y, x);
    for (int y = 0; y < H; ++y) {
        const real_t* rs = t + (size_t)y * P + 1;
        const real_t* rc = rs + P;
        const real_t* rn = rc + P;
        for (int x = 0; x < W; ++x) {
            real_t C = rc[x], E = rc[x + 1], Wn = rc[x - 1], N = rn[x], S = rs[x];
            real_t NE = rn[x + 1], NW = rn[x - 1], SW = rs[x - 1], SE = rs[x + 1];
            o->dx[IDX(y, x)] = (4 * (E - Wn) + NE - NW - SW + SE) / 12;
            o->dy[IDX(y, x)] = (4 * (N - S) + NE + NW - SW - SE) / 12;
            o->lap[IDX(y, x)] = (-20 * C + 4 * (N + E + Wn + S) + NE + NW + SW + SE) / 6;
        }
    }
    free(t);
}

static real_t mu_plain(const o64* o, size_t n, real_t psi) { /* :146-149 */
    return o->p.a * psi * (1 - psi * psi) - o->p.kappa * o->lap[n];
}

static real_t mix_tau(const o64* o, real_t rho, real_t psi) { /* :201-208 */
    real_t v1 = o->p.eta_n / rho, v2 = o->p.eta_n * o->p.M / rho;
    const real_t b = (real_t)o->p.psi_visc_bound;
    if (b > 0) psi = psi < -b ? -b : (psi > b ? b : psi);
    return 3 * (2 * v1 * v2 / (v1 * (1 - psi) + v2 * (1 + psi))) + 0.5;
}

static void equilibria(const o64* o, size_t n, real_t feq[9], real_t geq[9]) { /* :155-192 */
    real_t rho = o->rho[n], psi = o->psi[n], ux = o->ux[n], uy = o->uy[n], p = o->p_[n], mu = o->mu[n];
    for (int i = 0; i < 9; ++i) {
        real_t eu = EX[i] * ux + EY[i] * uy;
        real_t q = 3 * eu + 4.5 * eu * eu - 1.5 * (ux * ux + uy * uy);
        real_t A = i == 0 ? (rho - 3.0 * (1.0 - WT[0]) * p) / WT[0] : 3 * p;
        real_t B = i == 0 ? (psi - 3.0 * (1.0 - WT[0]) * o->p.gamma * mu) / WT[0] : 3 * o->p.gamma * mu;
        feq[i] = WT[i] * (A + rho * q);
        geq[i] = WT[i] * (B + psi * q);
    }
}

static void collide(o64* o) { /* :455-460, getLarge_F :194-199, getF/getG :258-264 */
    for (int y = 0; y < o->H; ++y)
        for (int x = 0; x < o->W; ++x) {
            size_t n = IDX(y, x);
            if (!o->fluid[n]) continue;
            real_t feq[9], geq[9];
            equilibria(o, n, feq, geq);
            real_t ux = o->ux[n], uy = o->uy[n], tm = o->tau_mix[n];
            real_t Fx = o->mu[n] * o->dx[n] + o->p.gx, Fy = o->mu[n] * o->dy[n] + o->p.gy;
            for (int i = 0; i < 9; ++i) {
                real_t eu = EX[i] * ux + EY[i] * uy;
                real_t F = WT[i] * (1 - 1 / (2 * tm)) *
                           (((EX[i] - ux) * 3 + EX[i] * eu * 9) * Fx + ((EY[i] - uy) * 3 + EY[i] * eu * 9) * Fy);
                real_t fi = POP(o->f, i, y, x), gi = POP(o->g, i, y, x);
                POP(o->f, i, y, x) = fi - 1.0 / tm * (fi - feq[i]) + F;
                POP(o->g, i, y, x) = gi - 1.0 / o->p.tau * (gi - geq[i]);
            }
        }
}

static void moments(o64* o) { /* :470-479 */
    for (size_t n = 0; n < (size_t)o->N; ++n)
        if (o->fluid[n]) {
            real_t s = 0, ps = 0;
            for (int i = 0; i < 9; ++i) {
                s += o->f[(size_t)i * o->N + n];
                ps += o->g[(size_t)i * o->N + n];
            }
            o->rho[n] = s;
            o->psi[n] = ps;
        } else {
            o->psi[n] = o->p.psi_wall;
        }
    stencils(o);
    for (size_t n = 0; n < (size_t)o->N; ++n) {
        if (!o->fluid[n]) continue;
        real_t psi = o->psi[n];
        o->mu[n] = o->p.a * psi * (1 - psi * psi) - o->p.kappa * o->lap[n];
        real_t jx = 0, jy = 0;
        for (int i = 0; i < 9; ++i) {
            jx += o->f[(size_t)i * o->N + n] * EX[i];
            jy += o->f[(size_t)i * o->N + n] * EY[i];
        }
        o->ux[n] = (jx + (o->mu[n] * o->dx[n] + o->p.gx) / 2) / o->rho[n];
        o->uy[n] = (jy + (o->mu[n] * o->dy[n] + o->p.gy) / 2) / o->rho[n];
        o->p_[n] = o->rho[n] / 3 + psi * o->mu[n];
        o->tau_mix[n] = mix_tau(o, o->rho[n], psi);
    }
}

static void stream(o64* o) { /* :327-343: f_i(y, x) <- f*_i(y - ey, x - ex) */
    const int H = o->H, W = o->W;
    for (int i = 0; i < 9; ++i)
        for (int y = 0; y < H; ++y) {
            const int ys = wrapi(y - EY[i], H);
            const real_t* fs = o->fb + (size_t)i * o->N + (size_t)ys * W;
            const real_t* gs = o->gb + (size_t)i * o->N + (size_t)ys * W;
            real_t* fd = o->f + (size_t)i * o->N + (size_t)y * W;
            real_t* gd = o->g + (size_t)i * o->N + (size_t)y * W;
            for (int x = 0; x < W; ++x) {
                int xs = x - EX[i];
                xs = xs < 0 ? xs + W : (xs >= W ? xs - W : xs);
                fd[x] = fs[xs];
                gd[x] = gs[xs];
            }
        }
}

static void bounce(o64* o) { /* bounce_back.py: same node's opposite post-collision value */
    for (int i = 1; i < 9; ++i)
        for (size_t n = 0; n < (size_t)o->N; ++n)
            if (o->links[(size_t)i * o->N + n]) {
                o->f[(size_t)i * o->N + n] = o->fb[(size_t)OPP[i] * o->N + n];
                o->g[(size_t)i * o->N + n] = o->gb[(size_t)OPP[i] * o->N + n];
            }
}

static void walls(o64* o) { /* bottom_top_wall on rows wall_lo / wall_hi */
    for (int x = 0; x < o->W; ++x) {
        int lo = o->p.wall_lo, hi = o->p.wall_hi;
        POP(o->f, 2, lo, x) = POP(o->fb, 4, lo, x);
        POP(o->g, 2, lo, x) = POP(o->gb, 4, lo, x);
        POP(o->f, 4, hi, x) = POP(o->fb, 2, hi, x);
        POP(o->g, 4, hi, x) = POP(o->gb, 2, hi, x);
        POP(o->f, 5, lo, x) = POP(o->fb, 7, lo, x);
        POP(o->g, 5, lo, x) = POP(o->gb, 7, lo, x);
        POP(o->f, 6, lo, x) = POP(o->fb, 8, lo, x);
        POP(o->g, 6, lo, x) = POP(o->gb, 8, lo, x);
        POP(o->f, 7, hi, x) = POP(o->fb, 5, hi, x);
        POP(o->g, 7, hi, x) = POP(o->gb, 5, hi, x);
        POP(o->f, 8, hi, x) = POP(o->fb, 6, hi, x);
        POP(o->g, 8, hi, x) = POP(o->gb, 6, hi, x);
    }
}

#define F(i, y) POP(o->f, i, y, col)
#define G(i, y) POP(o->g, i, y, col)

static void zou_he(o64* o) {
    const int H = o->H;
    real_t *Px = malloc(sizeof(real_t) * H), *Py = malloc(sizeof(real_t) * H), *rin = malloc(sizeof(real_t) * H);
    /* inlet (fingering_periodic.py:268-302; fingering.py:298-364) */
    int col = 0;
    for (int y = 0; y < H; ++y) {
        size_t n = IDX(y, col);
        real_t psi = psi_at(o, y, col), mu = mu_plain(o, n, psi);
        Px[y] = o->dx[n] * mu;
        Py[y] = o->dy[n] * mu;
        rin[y] = 1 / (1 - o->uprof[y]) * (F(0, y) + F(2, y) + F(4, y) + 2 * (F(3, y) + F(6, y) + F(7, y)) - Px[y] / 2);
    }
    const real_t wsi = WT[1] + WT[5] + WT[8];
    for (int y = (o->p.y_walls ? 1 : 0); y < (o->p.y_walls ? H - 1 : H); ++y) {
        real_t u = o->uprof[y], r = rin[y];
        real_t psi_in = 1.0 - (G(0, y) + G(2, y) + G(3, y) + G(4, y) + G(6, y) + G(7, y));
        F(1, y) = F(3, y) + 2.0 / 3 * u * r - Px[y] / 6;
        G(1, y) = WT[1] * psi_in / wsi;
        F(5, y) = F(7, y) - 0.5 * (F(2, y) - F(4, y)) + 1.0 / 6.0 * u * r - Px[y] / 6 - Py[y] / 4;
        G(5, y) = WT[5] * psi_in / wsi;
        F(8, y) = F(6, y) + 0.5 * (F(2, y) - F(4, y)) + 1.0 / 6.0 * u * r - Px[y] / 6 + Py[y] / 4;
        G(8, y) = WT[8] * psi_in / wsi;
    }
    if (o->p.y_walls) {
        int y = 0; /* bottom corner, fingering.py:334-348 */
        F(1, y) = F(3, y); G(1, y) = G(3, y);
        F(2, y) = F(4, y); G(2, y) = G(4, y);
        F(5, y) = F(7, y); G(5, y) = G(7, y);
        F(6, y) = 0.5 * (rin[1] - (F(0, y) + F(1, y) + F(2, y) + F(3, y) + F(4, y) + F(5, y) + F(7, y)));
        G(6, y) = WT[6] * (1.0 - (G(0, y) + G(1, y) + G(2, y) + G(3, y) + G(4, y) + G(5, y) + G(7, y))) / (WT[6] + WT[8]);
        F(8, y) = F(6, y); G(8, y) = G(6, y);
        y = H - 1; /* top corner, fingering.py:350-364 */
        F(1, y) = F(3, y); G(1, y) = G(3, y);
        F(4, y) = F(2, y); G(4, y) = G(2, y);
        F(8, y) = F(6, y); G(8, y) = G(6, y);
        F(5, y) = 0.5 * (rin[H - 2] - (F(0, y) + F(1, y) + F(2, y) + F(3, y) + F(4, y) + F(6, y) + F(8, y)));
        G(5, y) = WT[5] * (1.0 - (G(0, y) + G(1, y) + G(2, y) + G(3, y) + G(4, y) + G(6, y) + G(8, y))) / (WT[5] + WT[7]);
        F(7, y) = F(5, y); G(7, y) = G(5, y);
    }
    /* outlet (fingering_periodic.py:305-324; fingering.py:369-390) */
    col = o->W - 1;
    const real_t wso = WT[3] + WT[6] + WT[7];
    for (int y = 0; y < H; ++y) {
        size_t n = IDX(y, col);
        real_t psi = psi_at(o, y, col), mu = mu_plain(o, n, psi);
        real_t px = o->dx[n] * mu, py = o->dy[n] * mu;
        real_t known = F(0, y) + F(2, y) + F(4, y) + 2 * (F(1, y) + F(5, y) + F(8, y)) + px / 2;
        real_t u, r;
        if (o->p.pressure) {
            r = o->p.rho_out;
            u = known / r - 1.0;
        } else {
            u = o->uprof[y];
            r = 1 / (1 + u) * known;
        }
        real_t psi_out = -1.0 - (G(0, y) + G(1, y) + G(2, y) + G(4, y) + G(5, y) + G(8, y));
        F(3, y) = F(1, y) - o->p.c3 * u * r + px / 6;
        G(3, y) = WT[3] * psi_out / wso;
        F(6, y) = F(8, y) - 0.5 * (F(2, y) - F(4, y)) - 1.0 / 6.0 * u * r + py / 4 + px / 6;
        G(6, y) = WT[6] * psi_out / wso;
        F(7, y) = F(5, y) + 0.5 * (F(2, y) - F(4, y)) - 1.0 / 6.0 * u * r - py / 4 + px / 6;
        G(7, y) = WT[7] * psi_out / wso;
    }
    if (o->p.y_walls) {
        F(2, 0) = F(4, 0); G(2, 0) = G(4, 0);
        F(4, H - 1) = F(2, H - 1); G(4, H - 1) = G(2, H - 1);
    }
    free(Px);
    free(Py);
    free(rin);
}

/* ------------------------------------------------------------------ API */
void* o64_create(const o64_params* p, const unsigned char* fluid, const unsigned char* links, const double* psi0,
                 const double* rho0) {
    o64* o = calloc(1, sizeof(o64));
    o->p = *p;
    o->H = p->H;
    o->W = p->W;
    o->N = p->H * p->W;
    size_t N = o->N;
    o->fluid = malloc(N);
    memcpy(o->fluid, fluid, N);
    o->links = malloc(9 * N);
    memcpy(o->links, links, 9 * N);
    real_t** a[] = {&o->f, &o->g, &o->fb, &o->gb};
    for (int k = 0; k < 4; ++k) *a[k] = calloc(9 * N, sizeof(real_t));
    real_t** b[] = {&o->psi, &o->rho, &o->ux, &o->uy, &o->mu, &o->p_, &o->tau_mix, &o->dx, &o->dy, &o->lap};
    for (int k = 0; k < 10; ++k) *b[k] = calloc(N, sizeof(real_t));
    o->uprof = malloc(sizeof(real_t) * o->H);
    for (int y = 0; y < o->H; ++y) {
        if (p->gauss) {
            double t = (double)(y - o->H / 2) * 3 / ((double)o->H / 2);
            o->uprof[y] = (real_t)(p->u0 * exp(-(t * t) / 2));
        } else {
            o->uprof[y] = p->u0;
        }
    }
    /* initial state, fingering_periodic.py:90-121 (fingering.py:95-127) */
    for (size_t n = 0; n < N; ++n) {
        o->psi[n] = (real_t)(fluid[n] ? psi0[n] : p->psi_wall);
        o->rho[n] = (real_t)(fluid[n] ? (rho0 ? rho0[n] : p->rho0) : 0.0);
    }
    stencils(o);
    for (size_t n = 0; n < N; ++n) {
        if (!fluid[n]) continue;
        o->p_[n] = o->rho[n] / 3; /* getP with mu = 0 */
        if (p->init_walled) {
            o->mu[n] = p->a * o->psi[n] * (1 - o->psi[n] * o->psi[n]) - p->kappa * o->lap[n];
            o->uy[n] = (0 + (o->mu[n] * o->dy[n] + p->gy) / 2) / o->rho[n];
        }
        o->tau_mix[n] = mix_tau(o, o->rho[n], o->psi[n]);
        real_t feq[9], geq[9];
        equilibria(o, n, feq, geq);
        for (int i = 0; i < 9; ++i) {
            o->f[(size_t)i * N + n] = feq[i];
            o->g[(size_t)i * N + n] = geq[i];
        }
    }
    return o;
}

void o64_step(void* h, int nsteps) {
    o64* o = h;
    for (int s = 0; s < nsteps; ++s) {
        collide(o);
        memcpy(o->fb, o->f, sizeof(real_t) * 9 * o->N);
        memcpy(o->gb, o->g, sizeof(real_t) * 9 * o->N);
        stream(o);
        bounce(o);
        if (o->p.y_walls) walls(o);
        if (!o->p.x_periodic) zou_he(o);
        moments(o);
        o->t++;
    }
}

/* Multiply every fluid population by (1 + eps * r), r uniform in [-1, 1]
   from a xorshift stream: rounding-sized noise, so the spread it causes in
   the FP64 trajectory measures how this flow amplifies FP32 rounding. */
void o64_perturb(void* h, double eps, unsigned long long seed) {
    o64* o = h;
    unsigned long long s = seed * 0x9E3779B97F4A7C15ULL + 1;
    for (int i = 0; i < 9; ++i)
        for (size_t n = 0; n < (size_t)o->N; ++n) {
            if (!o->fluid[n]) continue;
            for (int k = 0; k < 2; ++k) {
                s ^= s << 13;
                s ^= s >> 7;
                s ^= s << 17;
                double r = (double)(s >> 11) / 9007199254740992.0 * 2.0 - 1.0;
                real_t* a = k ? o->g : o->f;
                a[(size_t)i * o->N + n] *= (real_t)(1.0 + eps * r);
            }
        }
}

/* Round every population to the nearest FP32 value (storage precision of
   an FP32 solver), keeping the FP64 arithmetic of the step itself. */
void o64_round32(void* h) {
    o64* o = h;
    for (size_t n = 0; n < 9 * (size_t)o->N; ++n) {
        o->f[n] = (real_t)(float)o->f[n];
        o->g[n] = (real_t)(float)o->g[n];
    }
}

/* psi full; rho, ux, uy with 0 on solid nodes */
void o64_fields(void* h, double* psi, double* rho, double* ux, double* uy) {
    o64* o = h;
    for (size_t n = 0; n < (size_t)o->N; ++n) {
        psi[n] = o->psi[n];
        rho[n] = o->fluid[n] ? o->rho[n] : 0.0;
        ux[n] = o->fluid[n] ? o->ux[n] : 0.0;
        uy[n] = o->fluid[n] ? o->uy[n] : 0.0;
    }
}

void o64_pops(void* h, double* f, double* g) {
    o64* o = h;
    for (size_t n = 0; n < 9 * (size_t)o->N; ++n) {
        f[n] = o->f[n];
        g[n] = o->g[n];
    }
}

void o64_mass(void* h, double out[2]) {
    o64* o = h;
    double r = 0, s = 0;
    for (size_t n = 0; n < (size_t)o->N; ++n)
        if (o->fluid[n]) {
            r += (double)o->rho[n];
            s += (double)o->psi[n];
        }
    out[0] = r;
    out[1] = s;
}

void o64_destroy(void* h) {
    o64* o = h;
    if (!o) return;
    free(o->fluid);
    free(o->links);
    free(o->f);
    free(o->g);
    free(o->fb);
    free(o->gb);
    real_t* b[] = {o->psi, o->rho, o->ux, o->uy, o->mu, o->p_, o->tau_mix, o->dx, o->dy, o->lap};
    for (int k = 0; k < 10; ++k) free(b[k]);
    free(o->uprof);
    free(o);
}
