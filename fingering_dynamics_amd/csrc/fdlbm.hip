// fdlbm.hip - gfx950 kernels and C ABI of the two-phase D2Q9 lattice-Boltzmann
// step of lattice_boltzmann/fingering_periodic.py:454-479.
//
// One fused pull-scheme kernel per iteration.  Launch k turns the
// post-collision populations of iteration k-1 into those of iteration k:
//   pull (stream, np.roll semantics, periodic in both axes)
//   -> half-way bounce-back from the node's link bits (bounce_back.py:25-167)
//   -> top/bottom walls (fingering.py:432-451, 573)
//   -> Zou-He inlet/outlet (fingering_periodic.py:268-324)
//   -> rho, psi, grad psi, lap psi, mu, u, p, tau_mix (:470-479)
//   -> collision of iteration k (:455-460) -> store.
// The psi stencil needs psi of the 8 neighbours after the same pull, so each
// workgroup recomputes psi (g pull + boundary conditions) one column and one
// row beyond its tile and keeps the last rows in an LDS ring.
//
// HBM layout: one allocation per ping-pong set holding 18 planes
// (f0..f8, g0..g8), each [H][Wp] float with the row pitch Wp a multiple of
// 64 floats (256 B); one uint16 flag per node (bits 0-7 bounce-back links of
// directions 1..8, bit 8 solid); a 4 x H side buffer with P_x = d_x psi * mu
// and P_y = d_y psi * mu of columns 0 and W-1 taken from the previous
// iteration's psi, which the Zou-He update reads (fingering_periodic.py:273-275).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fdlbm.h"
#include "lbm_node.h"

namespace {

constexpr unsigned kSolid = 0x100u;

struct KParams {
    int H, W, Wp;
    int Q;          // plane stride in floats (H * Wp)
    int rows;       // rows per workgroup chunk
    int nstrips;    // workgroup strips across x
    int y_walls, x_periodic, pressure, wall_lo, wall_hi;
    float rho_out, c3;
    fdn::Consts<float> c;
};

struct KBufs {
    const float* __restrict__ src;     // post-collision populations (18 planes)
    float* __restrict__ dst;           // next post-collision populations / post-stream copy
    const uint16_t* __restrict__ flags;
    const float* __restrict__ side_in; // [4][H] P of the previous psi
    float* __restrict__ side_out;
    const float* __restrict__ uprof;   // [H] inlet/outlet velocity
    float* __restrict__ fields;        // [4][H][Wp] psi, rho, ux, uy
    double* __restrict__ partials;     // [gridDim][3]
};

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Pull the 9 populations of one distribution (plane base P0 = 0 for f, 9
// for g) into node (y, x), then apply the node's bounce-back links and the
// walls.  All sources are post-collision values of the previous iteration.
__device__ __forceinline__ void pull(const float* __restrict__ src, const KParams& k, int P0,
                                     int y, int x, unsigned fl, float v[9]) {
    const int yn = (y + 1 == k.H) ? 0 : y + 1;
    const int ys = (y == 0) ? k.H - 1 : y - 1;
    const int xe = (x + 1 == k.W) ? 0 : x + 1;
    const int xw = (x == 0) ? k.W - 1 : x - 1;
    const float* __restrict__ b = src + P0 * k.Q;
    const int rc = y * k.Wp, rn = yn * k.Wp, rs = ys * k.Wp;
    const int Q = k.Q;
    v[0] = b[rc + x];
    v[1] = b[Q + rc + xw];
    v[2] = b[2 * Q + rs + x];
    v[3] = b[3 * Q + rc + xe];
    v[4] = b[4 * Q + rn + x];
    v[5] = b[5 * Q + rs + xw];
    v[6] = b[6 * Q + rs + xe];
    v[7] = b[7 * Q + rn + xe];
    v[8] = b[8 * Q + rn + xw];
    const unsigned links = fl & FD_LINK_BITS;
    if (links) {
#pragma unroll
        for (int i = 1; i < 9; ++i)
            if (links & (1u << (i - 1))) v[i] = b[fdn::OPP[i] * Q + rc + x];
    }
    if (k.y_walls) {
        if (y == k.wall_lo) {
            v[2] = b[4 * Q + rc + x];
            v[5] = b[7 * Q + rc + x];
            v[6] = b[8 * Q + rc + x];
        }
        if (y == k.wall_hi) {
            v[4] = b[2 * Q + rc + x];
            v[7] = b[5 * Q + rc + x];
            v[8] = b[6 * Q + rc + x];
        }
    }
}

// Zou-He (and walled corner) update of g at an inlet/outlet node.
__device__ __forceinline__ void bc_g(const KParams& k, int y, int x, float g[9]) {
    if (k.x_periodic) return;
    if (x == 0) {
        if (k.y_walls && y == 0)
            fdn::inlet_corner_bottom_g(g);
        else if (k.y_walls && y == k.H - 1)
            fdn::inlet_corner_top_g(g);
        else
            fdn::inlet_g(g);
    }
    if (x == k.W - 1) {
        fdn::outlet_g(g);
        if (k.y_walls) {
            if (y == 0) g[2] = g[4];
            if (y == k.H - 1) g[4] = g[2];
        }
    }
}

// Inlet density of row r (walled corner nodes use the adjacent row's).
__device__ __noinline__ float inlet_rho_row(const KParams& k, const KBufs& b, int r) {
    float f[9];
    pull(b.src, k, 0, r, 0, b.flags[r * k.Wp], f);
    return fdn::inlet_rho(f, b.uprof[r], b.side_in[r]);
}

__device__ __forceinline__ void bc_f(const KParams& k, const KBufs& b, int y, int x, float f[9]) {
    if (k.x_periodic) return;
    if (x == 0) {
        if (k.y_walls && y == 0)
            fdn::inlet_corner_bottom_f(f, inlet_rho_row(k, b, 1));
        else if (k.y_walls && y == k.H - 1)
            fdn::inlet_corner_top_f(f, inlet_rho_row(k, b, k.H - 2));
        else
            fdn::inlet_f(f, b.uprof[y], b.side_in[y], b.side_in[k.H + y]);
    }
    if (x == k.W - 1) {
        fdn::outlet_f(f, b.uprof[y], b.side_in[2 * k.H + y], b.side_in[3 * k.H + y],
                      k.pressure != 0, k.rho_out, k.c3);
        if (k.y_walls) {
            if (y == 0) f[2] = f[4];
            if (y == k.H - 1) f[4] = f[2];
        }
    }
}

// psi of node (y, x) after pull + bounce-back + walls + Zou-He, with the
// stencil's ghost values outside the domain: psi_wall ghost rows when the
// channel is walled (fingering.py:224), +1 / -1 ghost columns at x = -1 / W
// unless x is periodic (fingering_periodic.py:94-95, 218).  Also returns the
// node's g and flags for in-domain nodes.
__device__ __forceinline__ float psi_node(const KParams& k, const KBufs& b, int y, int x,
                                          float g[9], unsigned& fl) {
    fl = kSolid;
    if (k.y_walls && (y < 0 || y >= k.H)) return k.c.psi_wall;
    y = wrap(y, k.H);
    if (x < 0 || x >= k.W) {
        if (!k.x_periodic || x < -1 || x > k.W) return x < 0 ? 1.0f : -1.0f;
        x = wrap(x, k.W);
    }
    fl = b.flags[y * k.Wp + x];
    pull(b.src, k, 9, y, x, fl, g);
    bc_g(k, y, x, g);
    if (fl & kSolid) return k.c.psi_wall;
    return g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6] + g[7] + g[8];
}

template <int NT>
__device__ __forceinline__ void block_sum3(double s0, double s1, double s2, double* out) {
    __shared__ double sm[3][NT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, 64);
        s1 += __shfl_xor(s1, off, 64);
        s2 += __shfl_xor(s2, off, 64);
    }
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    if ((t & 63) == 0) {
        sm[0][t >> 6] = s0;
        sm[1][t >> 6] = s1;
        sm[2][t >> 6] = s2;
    }
    __syncthreads();
    if (t == 0) {
        double a0 = 0, a1 = 0, a2 = 0;
        for (int w = 0; w < NT / 64; ++w) {
            a0 += sm[0][w];
            a1 += sm[1][w];
            a2 += sm[2][w];
        }
        out[0] = a0;
        out[1] = a1;
        out[2] = a2;
    }
}

// MODE 0: one iteration (collide and store the next post-collision state).
// MODE 1: moments only - fields psi/rho/ux/uy, FP64 partial sums, and (if
//         b.dst) the post-stream populations; the state is not advanced.
template <int WX, int MODE>
__global__ __launch_bounds__(64 * WX) void step_kernel(KParams k, KBufs b) {
    constexpr int NT = 64 * WX;
    constexpr int SW = NT - 2;
    __shared__ float ring[4][NT];
    const int t = threadIdx.x;
    const int strip = blockIdx.x % k.nstrips;
    const int chunk = blockIdx.x / k.nstrips;
    const int x = strip * SW - 1 + t;
    const bool interior = (t >= 1) && (t <= SW) && (x < k.W);
    const int y0 = chunk * k.rows;
    const int y1 = min(y0 + k.rows, k.H);

    float gc[9], gp[9];
    unsigned flc, flp;
    double s_rho = 0.0, s_psi = 0.0, s_bad = 0.0;

    float psi_m = psi_node(k, b, y0 - 1, x, gp, flp);
    float psi_c = psi_node(k, b, y0, x, gc, flc);
    ring[0][t] = psi_m;
    ring[1][t] = psi_c;

    for (int y = y0; y < y1; ++y) {
        const int r = y - y0;
        const float psi_p = psi_node(k, b, y + 1, x, gp, flp);
        ring[(r + 2) & 3][t] = psi_p;
        __syncthreads();
        if (interior) {
            const float* Rm = ring[r & 3];
            const float* Rc = ring[(r + 1) & 3];
            const float* Rp = ring[(r + 2) & 3];
            const float n[9] = {psi_c,     Rc[t + 1], psi_p,     Rc[t - 1], psi_m,
                                Rp[t + 1], Rp[t - 1], Rm[t - 1], Rm[t + 1]};
            float dpx, dpy, lap;
            fdn::stencil(n, dpx, dpy, lap);
            float f[9];
            pull(b.src, k, 0, y, x, flc, f);
            bc_f(k, b, y, x, f);
            const int o = y * k.Wp + x;
            const bool solid = (flc & kSolid) != 0;
            if (MODE == 0) {
                if (!k.x_periodic && (x == 0 || x == k.W - 1)) {
                    const float mu = fdn::chem_pot(psi_c, lap, k.c);
                    const int s = (x == 0) ? 0 : 2;
                    b.side_out[s * k.H + y] = dpx * mu;
                    b.side_out[(s + 1) * k.H + y] = dpy * mu;
                }
                if (!solid) {
                    const fdn::Mom<float> m = fdn::moments(f, psi_c, dpx, dpy, lap, k.c);
                    fdn::collide(f, gc, m, k.c);
                }
#pragma unroll
                for (int i = 0; i < 9; ++i) {
                    b.dst[i * k.Q + o] = f[i];
                    b.dst[(9 + i) * k.Q + o] = gc[i];
                }
            } else {
                float rho = 0.f, ux = 0.f, uy = 0.f;
                if (!solid) {
                    const fdn::Mom<float> m = fdn::moments(f, psi_c, dpx, dpy, lap, k.c);
                    rho = m.rho;
                    ux = m.ux;
                    uy = m.uy;
                    s_rho += (double)rho;
                    s_psi += (double)psi_c;
                    if (!isfinite(rho) || !isfinite(psi_c)) s_bad += 1.0;
                }
                b.fields[o] = psi_c;
                b.fields[k.Q + o] = rho;
                b.fields[2 * k.Q + o] = ux;
                b.fields[3 * k.Q + o] = uy;
                if (b.dst) {
#pragma unroll
                    for (int i = 0; i < 9; ++i) {
                        b.dst[i * k.Q + o] = f[i];
                        b.dst[(9 + i) * k.Q + o] = gc[i];
                    }
                }
            }
        }
        psi_m = psi_c;
        psi_c = psi_p;
#pragma unroll
        for (int i = 0; i < 9; ++i) gc[i] = gp[i];
        flc = flp;
    }
    if (MODE == 1) block_sum3<NT>(s_rho, s_psi, s_bad, b.partials + 3 * blockIdx.x);
}

// Initial state (fingering_periodic.py:90-121, or fingering.py:95-127 with
// the walled init), one thread per node.
//   PRE = 1: pre-collision populations f0/g0 into b.dst and the N = 0
//            fields + partial sums.
//   PRE = 0: post-collision populations of iteration 0 into b.dst and the
//            side buffer P of the initial psi into b.side_out.
template <int PRE>
__global__ __launch_bounds__(256) void init_kernel(KParams k, KBufs b, const float* __restrict__ psi0,
                                                   const float* __restrict__ rho0, int walled_init) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    double s_rho = 0.0, s_psi = 0.0, s_bad = 0.0;
    if (x < k.W && y < k.H) {
        auto P = [&](int yy, int xx) -> float {
            if (k.y_walls && (yy < 0 || yy >= k.H)) return k.c.psi_wall;
            yy = wrap(yy, k.H);
            if (xx < 0 || xx >= k.W) {
                if (!k.x_periodic) return xx < 0 ? 1.0f : -1.0f;
                xx = wrap(xx, k.W);
            }
            const int o = yy * k.Wp + xx;
            return (b.flags[o] & kSolid) ? k.c.psi_wall : psi0[o];
        };
        const float n[9] = {P(y, x),         P(y, x + 1),     P(y + 1, x),
                            P(y, x - 1),     P(y - 1, x),     P(y + 1, x + 1),
                            P(y + 1, x - 1), P(y - 1, x - 1), P(y - 1, x + 1)};
        float dpx, dpy, lap;
        fdn::stencil(n, dpx, dpy, lap);
        const float psi = n[0];
        const int o = y * k.Wp + x;
        const bool solid = (b.flags[o] & kSolid) != 0;
        float f[9], g[9];
        fdn::Mom<float> m{};
        if (solid) {
#pragma unroll
            for (int i = 0; i < 9; ++i) f[i] = g[i] = 0.f;
        } else {
            m.rho = rho0[o];
            m.psi = psi;
            m.mu = walled_init ? fdn::chem_pot(psi, lap, k.c) : 0.f;
            m.fx = m.mu * dpx + k.c.gx;
            m.fy = m.mu * dpy + k.c.gy;
            m.ux = 0.f;
            m.uy = walled_init ? (0.5f * m.fy) / m.rho : 0.f;
            m.p = m.rho * (1.f / 3.f);  // evaluated before mu (both inits)
            m.itau = fdn::inv_tau_mix(m.rho, psi, k.c);
            fdn::equilibria(m, k.c, f, g);
            if (!PRE) fdn::collide(f, g, m, k.c);
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            b.dst[i * k.Q + o] = f[i];
            b.dst[(9 + i) * k.Q + o] = g[i];
        }
        if (PRE) {
            b.fields[o] = psi;
            b.fields[k.Q + o] = solid ? 0.f : m.rho;
            b.fields[2 * k.Q + o] = solid ? 0.f : m.ux;
            b.fields[3 * k.Q + o] = solid ? 0.f : m.uy;
            if (!solid) {
                s_rho = (double)m.rho;
                s_psi = (double)psi;
                if (!isfinite(m.rho) || !isfinite(psi)) s_bad = 1.0;
            }
        } else if (!k.x_periodic && (x == 0 || x == k.W - 1)) {
            const float mu = fdn::chem_pot(psi, lap, k.c);
            const int s = (x == 0) ? 0 : 2;
            b.side_out[s * k.H + y] = dpx * mu;
            b.side_out[(s + 1) * k.H + y] = dpy * mu;
        }
    }
    if (PRE) block_sum3<256>(s_rho, s_psi, s_bad, b.partials + 3 * (blockIdx.y * gridDim.x + blockIdx.x));
}

__global__ __launch_bounds__(256) void reduce_partials(const double* __restrict__ part, int n,
                                                       double* __restrict__ out) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        s0 += part[3 * i];
        s1 += part[3 * i + 1];
        s2 += part[3 * i + 2];
    }
    block_sum3<256>(s0, s1, s2, out);
}

// ------------------------------------------------------------------ host
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define FD_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(FD_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

}  // namespace

struct fd_handle {
    fd_params p{};
    KParams k{};
    int device = 0;
    int Wp = 0;
    size_t Q = 0;
    float* pop[2] = {nullptr, nullptr};
    uint16_t* flags = nullptr;
    float* side[2] = {nullptr, nullptr};
    float* uprof = nullptr;
    float* psi0 = nullptr;
    float* rho0 = nullptr;
    float* fields = nullptr;
    double* partials = nullptr;
    double* mass = nullptr;
    int npart = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cur = 0, sc = 0;
    int64_t N = 0;
    bool inited = false;
    int wx = 4;
    int rows = 32;
    int timed_launches = 0;
    bool have_timing = false;
};

namespace {

int check_params(const fd_params* p) {
    if (!p) return fail(FD_E_ARG, "params is NULL");
    if (p->H < 3 || p->W < 3) return fail(FD_E_ARG, "H and W must be >= 3");
    if ((int64_t)p->H * (((int64_t)p->W + 63) / 64 * 64) * 18 >= (int64_t)1 << 31)
        return fail(FD_E_ARG, "domain too large for 32-bit plane offsets");
    if (p->y_bc != FD_Y_PERIODIC && p->y_bc != FD_Y_WALLS) return fail(FD_E_ARG, "bad y_bc");
    if (p->x_bc != FD_X_ZOUHE && p->x_bc != FD_X_PERIODIC) return fail(FD_E_ARG, "bad x_bc");
    if (p->profile != FD_PROFILE_GAUSS && p->profile != FD_PROFILE_CONST)
        return fail(FD_E_ARG, "bad profile");
    if (p->outlet != FD_OUTLET_VELOCITY && p->outlet != FD_OUTLET_PRESSURE)
        return fail(FD_E_ARG, "bad outlet");
    if (p->init_mode != FD_INIT_MU0 && p->init_mode != FD_INIT_WALLED)
        return fail(FD_E_ARG, "bad init_mode");
    // The Gaussian column of fingering_periodic.py:270 has H-1 samples for odd
    // H, which the reference cannot broadcast against an H-row column.
    if (p->x_bc == FD_X_ZOUHE && p->profile == FD_PROFILE_GAUSS && (p->H % 2))
        return fail(FD_E_ARG, "the Gaussian inlet profile needs an even H");
    if (p->y_bc == FD_Y_WALLS &&
        (p->wall_lo < 0 || p->wall_lo >= p->H || p->wall_hi < 0 || p->wall_hi >= p->H))
        return fail(FD_E_ARG, "wall rows outside [0, H)");
    if (!(p->tau > 0.5) || !(p->eta_n > 0) || !(p->M > 0)) return fail(FD_E_ARG, "bad tau/eta_n/M");
    return FD_OK;
}

void fill_kparams(fd_handle* h) {
    const fd_params& p = h->p;
    KParams& k = h->k;
    k.H = p.H;
    k.W = p.W;
    k.Wp = h->Wp;
    k.Q = (int)h->Q;
    k.y_walls = p.y_bc == FD_Y_WALLS;
    k.x_periodic = p.x_bc == FD_X_PERIODIC;
    k.pressure = p.outlet == FD_OUTLET_PRESSURE;
    k.wall_lo = p.wall_lo;
    k.wall_hi = p.wall_hi;
    k.rho_out = (float)p.rho_out;
    k.c3 = (float)p.outlet_f3_coef;
    k.c.inv_tau_g = (float)(1.0 / p.tau);
    k.c.eta_n = (float)p.eta_n;
    k.c.M = (float)p.M;
    k.c.a = (float)p.a;
    k.c.kappa = (float)p.kappa;
    k.c.gamma = (float)p.gamma;
    k.c.psi_wall = (float)p.psi_wall;
    k.c.gx = (float)p.gx;
    k.c.gy = (float)p.gy;
}

int strip_width(int wx) { return 64 * wx - 2; }

int launch_step(fd_handle* h, int mode, const KBufs& b) {
    KParams k = h->k;
    k.rows = h->rows;
    k.nstrips = (k.W + strip_width(h->wx) - 1) / strip_width(h->wx);
    const int nchunks = (k.H + k.rows - 1) / k.rows;
    const dim3 grid(k.nstrips * nchunks);
    const dim3 block(64 * h->wx);
#define FD_LAUNCH(WX)                                                                        \
    do {                                                                                     \
        if (mode == 0)                                                                       \
            hipLaunchKernelGGL((step_kernel<WX, 0>), grid, block, 0, h->stream, k, b);       \
        else                                                                                 \
            hipLaunchKernelGGL((step_kernel<WX, 1>), grid, block, 0, h->stream, k, b);       \
    } while (0)
    switch (h->wx) {
        case 1: FD_LAUNCH(1); break;
        case 2: FD_LAUNCH(2); break;
        default: FD_LAUNCH(4); break;
    }
#undef FD_LAUNCH
    FD_HIP(hipGetLastError());
    return FD_OK;
}

int moments_grid_blocks(const fd_handle* h) {
    const int nstrips = (h->p.W + strip_width(h->wx) - 1) / strip_width(h->wx);
    return nstrips * ((h->p.H + h->rows - 1) / h->rows);
}

int ensure_partials(fd_handle* h, int n) {
    if (n <= h->npart) return FD_OK;
    if (h->partials) FD_HIP(hipFree(h->partials));
    h->partials = nullptr;
    FD_HIP(hipMalloc(&h->partials, sizeof(double) * 3 * (size_t)n));
    h->npart = n;
    return FD_OK;
}

// Fields (and optionally post-stream populations into `pops`) of the current
// iteration, plus the FP64 sums in h->mass.
int compute_fields(fd_handle* h, float* pops) {
    if (!h->inited) return fail(FD_E_STATE, "fd_init has not been called");
    KBufs b{};
    b.flags = h->flags;
    b.uprof = h->uprof;
    b.fields = h->fields;
    b.dst = pops;
    int nblk;
    if (h->N == 0) {
        const dim3 grid((h->p.W + 63) / 64, (h->p.H + 3) / 4);
        nblk = grid.x * grid.y;
        int rc = ensure_partials(h, nblk);
        if (rc) return rc;
        b.partials = h->partials;
        if (!pops) b.dst = h->pop[1 - h->cur];
        hipLaunchKernelGGL((init_kernel<1>), grid, dim3(64, 4), 0, h->stream, h->k, b, h->psi0, h->rho0,
                           h->p.init_mode == FD_INIT_WALLED ? 1 : 0);
        FD_HIP(hipGetLastError());
    } else {
        nblk = moments_grid_blocks(h);
        int rc = ensure_partials(h, nblk);
        if (rc) return rc;
        b.partials = h->partials;
        b.src = h->pop[h->cur];
        b.side_in = h->side[h->sc];
        rc = launch_step(h, 1, b);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(256), 0, h->stream, h->partials, nblk, h->mass);
    FD_HIP(hipGetLastError());
    return FD_OK;
}

int copy_plane_out(fd_handle* h, float* dst, const float* plane) {
    FD_HIP(hipMemcpy2DAsync(dst, sizeof(float) * h->p.W, plane, sizeof(float) * h->Wp,
                            sizeof(float) * h->p.W, h->p.H, hipMemcpyDeviceToHost, h->stream));
    return FD_OK;
}

}  // namespace

extern "C" {

const char* fd_last_error(void) { return g_err.c_str(); }

void fd_destroy(fd_handle* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (int i = 0; i < 2; ++i) {
        if (h->pop[i]) (void)hipFree(h->pop[i]);
        if (h->side[i]) (void)hipFree(h->side[i]);
    }
    if (h->flags) (void)hipFree(h->flags);
    if (h->uprof) (void)hipFree(h->uprof);
    if (h->psi0) (void)hipFree(h->psi0);
    if (h->rho0) (void)hipFree(h->rho0);
    if (h->fields) (void)hipFree(h->fields);
    if (h->partials) (void)hipFree(h->partials);
    if (h->mass) (void)hipFree(h->mass);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int fd_create(const fd_params* p, const uint8_t* solid, const uint8_t* links, fd_handle** out) {
    if (!out) return fail(FD_E_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    if (!solid || !links) return fail(FD_E_ARG, "solid/links is NULL");
    int dev = 0;
    FD_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    FD_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(FD_E_NODEV, std::string("libfdlbm is built for gfx950, device is ") + prop.gcnArchName);

    fd_handle* h = new (std::nothrow) fd_handle();
    if (!h) return fail(FD_E_ARG, "out of host memory");
    h->p = *p;
    h->device = dev;
    h->Wp = (p->W + 63) / 64 * 64;
    h->Q = (size_t)p->H * h->Wp;
    fill_kparams(h);
    // A large domain favours 4-wave strips; small ones need more workgroups.
    h->wx = p->W >= 1024 ? 4 : (p->W >= 256 ? 2 : 1);
    h->rows = p->H >= 1024 ? 32 : 16;

    const size_t Q = h->Q, H = p->H, W = p->W;
    std::vector<uint16_t> fl(Q, (uint16_t)kSolid);  // padding columns: solid, never read
    for (size_t y = 0; y < H; ++y)
        for (size_t x = 0; x < W; ++x)
            fl[y * h->Wp + x] = (uint16_t)((links[y * W + x] & FD_LINK_BITS) | (solid[y * W + x] ? kSolid : 0u));
    std::vector<float> up(H);
    for (size_t y = 0; y < H; ++y) {
        if (p->profile == FD_PROFILE_GAUSS) {
            // fingering_periodic.py:270-271: t = i*3/(H/2), i = y - H/2
            const double i = (double)((long)y - (long)(H / 2));
            const double tt = i * 3.0 / ((double)H / 2.0);
            up[y] = (float)(p->u0 * std::exp(-(tt * tt) / 2.0));
        } else {
            up[y] = (float)p->u0;
        }
    }

#define FD_HIP_H(call)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            fd_destroy(h);                                                                \
            return fail(FD_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
        }                                                                                 \
    } while (0)
    FD_HIP_H(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    FD_HIP_H(hipEventCreate(&h->ev0));
    FD_HIP_H(hipEventCreate(&h->ev1));
    for (int i = 0; i < 2; ++i) {
        FD_HIP_H(hipMalloc(&h->pop[i], sizeof(float) * 18 * Q));
        FD_HIP_H(hipMalloc(&h->side[i], sizeof(float) * 4 * H));
        FD_HIP_H(hipMemsetAsync(h->pop[i], 0, sizeof(float) * 18 * Q, h->stream));
        FD_HIP_H(hipMemsetAsync(h->side[i], 0, sizeof(float) * 4 * H, h->stream));
    }
    FD_HIP_H(hipMalloc(&h->flags, sizeof(uint16_t) * Q));
    FD_HIP_H(hipMalloc(&h->uprof, sizeof(float) * H));
    FD_HIP_H(hipMalloc(&h->psi0, sizeof(float) * Q));
    FD_HIP_H(hipMalloc(&h->rho0, sizeof(float) * Q));
    FD_HIP_H(hipMalloc(&h->fields, sizeof(float) * 4 * Q));
    FD_HIP_H(hipMalloc(&h->mass, sizeof(double) * 3));
    FD_HIP_H(hipMemcpyAsync(h->flags, fl.data(), sizeof(uint16_t) * Q, hipMemcpyHostToDevice, h->stream));
    FD_HIP_H(hipMemcpyAsync(h->uprof, up.data(), sizeof(float) * H, hipMemcpyHostToDevice, h->stream));
    FD_HIP_H(hipStreamSynchronize(h->stream));
#undef FD_HIP_H
    *out = h;
    return FD_OK;
}

int fd_init(fd_handle* h, const double* psi0, const double* rho0) {
    if (!h || !psi0) return fail(FD_E_ARG, "handle/psi0 is NULL");
    FD_HIP(hipSetDevice(h->device));
    const size_t H = h->p.H, W = h->p.W, Wp = h->Wp;
    std::vector<float> ps(h->Q, 0.f), rh(h->Q, 0.f);
    for (size_t y = 0; y < H; ++y)
        for (size_t x = 0; x < W; ++x) {
            ps[y * Wp + x] = (float)psi0[y * W + x];
            rh[y * Wp + x] = rho0 ? (float)rho0[y * W + x] : 1.0f;
        }
    FD_HIP(hipMemcpyAsync(h->psi0, ps.data(), sizeof(float) * h->Q, hipMemcpyHostToDevice, h->stream));
    FD_HIP(hipMemcpyAsync(h->rho0, rh.data(), sizeof(float) * h->Q, hipMemcpyHostToDevice, h->stream));
    KBufs b{};
    b.flags = h->flags;
    b.dst = h->pop[0];
    b.side_out = h->side[0];
    b.uprof = h->uprof;
    const dim3 grid((h->p.W + 63) / 64, (h->p.H + 3) / 4);
    hipLaunchKernelGGL((init_kernel<0>), grid, dim3(64, 4), 0, h->stream, h->k, b, h->psi0, h->rho0,
                       h->p.init_mode == FD_INIT_WALLED ? 1 : 0);
    FD_HIP(hipGetLastError());
    FD_HIP(hipStreamSynchronize(h->stream));  // host staging vectors go out of scope
    h->cur = 0;
    h->sc = 0;
    h->N = 0;
    h->inited = true;
    h->have_timing = false;
    return FD_OK;
}

int fd_step(fd_handle* h, int32_t nsteps) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    if (nsteps < 0) return fail(FD_E_ARG, "nsteps < 0");
    if (!h->inited) return fail(FD_E_STATE, "fd_init has not been called");
    FD_HIP(hipSetDevice(h->device));
    // The buffers hold the post-collision state of iteration max(N-1, 0);
    // reaching iteration N' needs launches up to N'-1.
    const int64_t have = std::max<int64_t>(h->N - 1, 0);
    const int64_t want = std::max<int64_t>(h->N + nsteps - 1, 0);
    const int64_t nl = want - have;
    if (nl > 0) FD_HIP(hipEventRecord(h->ev0, h->stream));
    for (int64_t i = 0; i < nl; ++i) {
        KBufs b{};
        b.src = h->pop[h->cur];
        b.dst = h->pop[1 - h->cur];
        b.flags = h->flags;
        b.side_in = h->side[h->sc];
        b.side_out = h->side[1 - h->sc];
        b.uprof = h->uprof;
        int rc = launch_step(h, 0, b);
        if (rc) return rc;
        h->cur ^= 1;
        h->sc ^= 1;
    }
    if (nl > 0) {
        FD_HIP(hipEventRecord(h->ev1, h->stream));
        h->timed_launches = (int)nl;
        h->have_timing = true;
    }
    h->N += nsteps;
    return FD_OK;
}

int64_t fd_steps_done(const fd_handle* h) { return h ? h->N : -1; }

int fd_synchronize(fd_handle* h) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    FD_HIP(hipStreamSynchronize(h->stream));
    return FD_OK;
}

double fd_last_step_ms(const fd_handle* h) {
    if (!h || !h->have_timing || h->timed_launches <= 0) return -1.0;
    if (hipEventSynchronize(h->ev1) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess) return -1.0;
    return (double)ms / h->timed_launches;
}

int fd_set_rows_per_wave(fd_handle* h, int32_t rows) {
    if (!h || rows < 1) return fail(FD_E_ARG, "rows must be >= 1");
    h->rows = rows;
    return FD_OK;
}

int fd_get_fields(fd_handle* h, float* psi, float* rho, float* ux, float* uy) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    FD_HIP(hipSetDevice(h->device));
    int rc = compute_fields(h, nullptr);
    if (rc) return rc;
    float* outs[4] = {psi, rho, ux, uy};
    for (int i = 0; i < 4; ++i)
        if (outs[i] && (rc = copy_plane_out(h, outs[i], h->fields + i * h->Q))) return rc;
    FD_HIP(hipStreamSynchronize(h->stream));
    return FD_OK;
}

int fd_get_populations(fd_handle* h, float* f, float* g) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    FD_HIP(hipSetDevice(h->device));
    // The idle ping-pong set receives the post-stream copy; the next launch
    // overwrites it anyway.
    float* tmp = h->pop[1 - h->cur];
    int rc = compute_fields(h, tmp);
    if (rc) return rc;
    for (int i = 0; i < 9; ++i) {
        if (f && (rc = copy_plane_out(h, f + (size_t)i * h->p.H * h->p.W, tmp + i * h->Q))) return rc;
        if (g && (rc = copy_plane_out(h, g + (size_t)i * h->p.H * h->p.W, tmp + (9 + i) * h->Q))) return rc;
    }
    FD_HIP(hipStreamSynchronize(h->stream));
    return FD_OK;
}

int fd_phase_mass(fd_handle* h, double out[3]) {
    if (!h || !out) return fail(FD_E_ARG, "handle/out is NULL");
    FD_HIP(hipSetDevice(h->device));
    int rc = compute_fields(h, nullptr);
    if (rc) return rc;
    FD_HIP(hipMemcpyAsync(out, h->mass, sizeof(double) * 3, hipMemcpyDeviceToHost, h->stream));
    FD_HIP(hipStreamSynchronize(h->stream));
    return FD_OK;
}

// State blob: header (magic, H, W, N) + side[4H] + 18 planes [H][W].
int64_t fd_state_bytes(const fd_handle* h) {
    if (!h) return -1;
    return 32 + (int64_t)sizeof(float) * (4 * (int64_t)h->p.H + 18 * (int64_t)h->p.H * h->p.W);
}

int fd_save_state(fd_handle* h, void* buf, int64_t nbytes) {
    if (!h || !buf) return fail(FD_E_ARG, "handle/buf is NULL");
    if (!h->inited) return fail(FD_E_STATE, "fd_init has not been called");
    if (nbytes < fd_state_bytes(h)) return fail(FD_E_ARG, "buffer too small");
    FD_HIP(hipSetDevice(h->device));
    int64_t hdr[4] = {0x4644424d53544154LL, h->p.H, h->p.W, h->N};
    std::memcpy(buf, hdr, 32);
    char* p = (char*)buf + 32;
    FD_HIP(hipMemcpyAsync(p, h->side[h->sc], sizeof(float) * 4 * h->p.H, hipMemcpyDeviceToHost, h->stream));
    p += sizeof(float) * 4 * h->p.H;
    for (int i = 0; i < 18; ++i) {
        int rc = copy_plane_out(h, (float*)p, h->pop[h->cur] + i * h->Q);
        if (rc) return rc;
        p += sizeof(float) * (size_t)h->p.H * h->p.W;
    }
    FD_HIP(hipStreamSynchronize(h->stream));
    return FD_OK;
}

int fd_load_state(fd_handle* h, const void* buf, int64_t nbytes) {
    if (!h || !buf) return fail(FD_E_ARG, "handle/buf is NULL");
    if (nbytes < fd_state_bytes(h)) return fail(FD_E_ARG, "buffer too small");
    int64_t hdr[4];
    std::memcpy(hdr, buf, 32);
    if (hdr[0] != 0x4644424d53544154LL || hdr[1] != h->p.H || hdr[2] != h->p.W || hdr[3] < 1)
        return fail(FD_E_ARG, "state blob does not match this handle (or was saved at step 0)");
    FD_HIP(hipSetDevice(h->device));
    const char* p = (const char*)buf + 32;
    FD_HIP(hipMemcpyAsync(h->side[h->sc], p, sizeof(float) * 4 * h->p.H, hipMemcpyHostToDevice, h->stream));
    p += sizeof(float) * 4 * h->p.H;
    for (int i = 0; i < 18; ++i) {
        FD_HIP(hipMemcpy2DAsync(h->pop[h->cur] + i * h->Q, sizeof(float) * h->Wp, p, sizeof(float) * h->p.W,
                                sizeof(float) * h->p.W, h->p.H, hipMemcpyHostToDevice, h->stream));
        p += sizeof(float) * (size_t)h->p.H * h->p.W;
    }
    FD_HIP(hipStreamSynchronize(h->stream));
    h->N = hdr[3];
    h->inited = true;
    h->have_timing = false;
    return FD_OK;
}

}  // extern "C"
