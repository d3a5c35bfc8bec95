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
// 64 floats (256 B); one uint32 flag word per node (slot / explicit
// bounce-back links, solid bit); a 4 x H side buffer with P_x = d_x psi * mu
// and P_y = d_y psi * mu of columns 0 and W-1 taken from the previous
// iteration's psi, which the Zou-He update reads (fingering_periodic.py:273-275).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <initializer_list>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "fdlbm.h"
#include "lbm_node.h"

namespace {

// Per-node flags (uint16):
//   bits 0-7   links: direction i (bit i-1) is bounced back, i.e. pulled
//              from the node's own post-collision opposite population
//              (half-way bounce-back, bounce_back.py:25-167); the top and
//              bottom walls (fingering.py:432-451, 573) are links of their
//              rows, set by the host
//   bit 8      solid
//   bit 9      dead piece (see kDeadPiece)
// In the inert-solid mode (leak-free geometry) solid nodes are never read
// and store zeros (or nothing, in 64-byte groups of inert solids);
// otherwise they stream their populations on unchanged (store-and-forward,
// fingering_periodic.py:455-466).
constexpr unsigned kLinks = 0xFFu;
constexpr unsigned kSolid = 1u << 8;
// bit 9: the node's aligned 16-column piece (64 B of a plane row) holds
// only inert solid nodes (or columns beyond W): nothing ever reads it, so
// no launch writes it.  Set by the host in the inert-solid mode.
constexpr unsigned kDeadPiece = 1u << 9;
// A vector offset beyond any descriptor's range: such buffer loads return 0
// and such stores are dropped, without a memory access.
constexpr unsigned kOOB = 0x7FFFFFFCu;
// Cache policy of the population stores: non-temporal (gfx950 "nt"); the
// next read of a stored value is one full iteration later.
constexpr int kStoreNT = 2;

struct KParams {
    int H, W, Wp;
    int Q;          // plane stride in floats (H * Wp)
    int rows;       // rows per wave
    int nsb;        // workgroups across x in this launch (each covers WPB strips of 64 columns)
    int sb0;        // first workgroup column of this launch (0 unless a column window)
    int y_walls, x_periodic, pressure;
    int inert;      // solid nodes are never read (leak-free geometry)
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
    double* __restrict__ partials;     // per wave [3]
};

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Row geometry shared by every lane of a wave (y is uniform across a
// wave): the three source rows of a pull, with periodic wrap.
struct Rows {
    int s, c, n;  // y-1, y, y+1 (wrapped), already multiplied by the pitch
};

__device__ __forceinline__ Rows rows_of(const KParams& k, int y) {
    Rows r;
    r.c = y * k.Wp;
    r.s = ((y == 0) ? k.H - 1 : y - 1) * k.Wp;
    r.n = ((y + 1 == k.H) ? 0 : y + 1) * k.Wp;
    return r;
}

// Buffer-resource access to a population set: the descriptor covers the
// whole 18-plane allocation; the vector offset carries the lane's column
// (and, for pulls, the plane) in bytes, the scalar offset the part that is
// the same for every lane of the wave.  Out-of-range offsets read 0.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const float* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, kRsrcWord3);
}

// Per-lane byte offsets of the nine pull sources inside the f planes:
// direction i reads plane i at column x - e_ix, i.e. ((x - e_ix) + i*Q)*4,
// wrapped in x when x is periodic (np.roll semantics).  For Zou-He x the
// column -1 / W reads of the first / last column land inside the
// allocation (planes >= 1, padding columns) and are rewritten by Zou-He.
// (links bit b set) ? a : c, as a sign-extended one-bit field extract and
// a bitfield insert: two VALU operations and no compare into VCC (the
// compiler turns the same expression written in C into and + compare +
// select).
__device__ __forceinline__ unsigned sel_bit(unsigned links, int b, unsigned a, unsigned c) {
    unsigned m, r;
    asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(m) : "v"(links), "i"(b));
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(c));
    return r;
}
__device__ __forceinline__ float sel_bit(unsigned links, int b, float a, float c) {
    return __builtin_bit_cast(float, sel_bit(links, b, __builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, c)));
}

struct PullOff {
    unsigned v[9];
    unsigned xb;  // the lane's own column in bytes
};

__device__ __forceinline__ PullOff pull_offsets(const KParams& k, unsigned x) {
    PullOff po;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        int xs = (int)x - fdn::EX[i];
        if (k.x_periodic) xs = wrap(xs, k.W);
        po.v[i] = (unsigned)(xs + i * k.Q) * 4u;
    }
    po.xb = x * 4u;
    return po;
}

// Streaming loads of one distribution (plane base P0 = 0 for f, 9 for g)
// into the node whose pull offsets are po, at the rows of r, with the
// node's bounced-back links already applied: a linked direction i reads
// the node's own post-collision f_opp(i) (plane opp(i), row y) instead of
// its upstream neighbour - the same single load at another address.  Lanes
// with !on read 0 without a memory access: loads are predicated by offset,
// never by a branch, so every path through the march issues the same loads
// and the waits for them stay exact.
__device__ __forceinline__ void pull(rsrc_t src, const KParams& k, int P0, const Rows& r, const PullOff& po,
                                     unsigned links, bool on, float v[9]) {
    const int gq = P0 * k.Q;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int row = fdn::EY[i] > 0 ? r.s : (fdn::EY[i] < 0 ? r.n : r.c);
        unsigned vo = po.v[i];
        if (i > 0) {
            // own node, plane opp(i), row y, relative to this load's scalar offset
            const unsigned d = (unsigned)(fdn::OPP[i] * k.Q + r.c - row) * 4u;
            vo = sel_bit(links, i - 1, po.xb + d, vo);
        }
        v[i] = __builtin_bit_cast(float,
                                  __builtin_amdgcn_raw_buffer_load_b32(src, on ? vo : kOOB, (row + gq) * 4, 0));
    }
}

// Zou-He (and walled corner) update of g at an inlet/outlet node.  The
// candidate updates are computed on full copies and merged with per-element
// selects: branches that store to different elements of one array make the
// optimiser merge the stores through a pointer phi, which forces the array
// out of registers.
__device__ __forceinline__ void bc_g(const KParams& k, int y, int x, float g[9]) {
    if (k.x_periodic) return;
    if (x == 0) {
        const bool bot = k.y_walls && y == 0;
        const bool top = k.y_walls && y == k.H - 1;
        float n[9], cb[9], ct[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) n[i] = cb[i] = ct[i] = g[i];
        fdn::inlet_g(n);
        fdn::inlet_corner_bottom_g(cb);
        fdn::inlet_corner_top_g(ct);
#pragma unroll
        for (int i = 0; i < 9; ++i) g[i] = bot ? cb[i] : (top ? ct[i] : n[i]);
    }
    if (x == k.W - 1) {
        fdn::outlet_g(g);
        const float g2 = (k.y_walls && y == 0) ? g[4] : g[2];
        const float g4 = (k.y_walls && y == k.H - 1) ? g[2] : g[4];
        g[2] = g2;
        g[4] = g4;
    }
}

// Inlet density of row r (walled corner nodes use the adjacent row's).
__device__ __forceinline__ float inlet_rho_row(const KParams& k, const KBufs& b, rsrc_t src, int r) {
    float f[9];
    pull(src, k, 0, rows_of(k, r), pull_offsets(k, 0u), b.flags[r * k.Wp] & kLinks, true, f);
    return fdn::inlet_rho(f, b.uprof[r], b.side_in[r]);
}

__device__ __forceinline__ void bc_f(const KParams& k, const KBufs& b, rsrc_t src, int y, int x, float f[9]) {
    if (k.x_periodic) return;
    if (x == 0) {
        const bool bot = k.y_walls && y == 0;
        const bool top = k.y_walls && y == k.H - 1;
        float rho_nb = 0.f;
        if (bot || top) rho_nb = inlet_rho_row(k, b, src, bot ? 1 : k.H - 2);
        float n[9], cb[9], ct[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) n[i] = cb[i] = ct[i] = f[i];
        fdn::inlet_f(n, b.uprof[y], b.side_in[y], b.side_in[k.H + y]);
        fdn::inlet_corner_bottom_f(cb, rho_nb);
        fdn::inlet_corner_top_f(ct, rho_nb);
#pragma unroll
        for (int i = 0; i < 9; ++i) f[i] = bot ? cb[i] : (top ? ct[i] : n[i]);
    }
    if (x == k.W - 1) {
        fdn::outlet_f(f, b.uprof[y], b.side_in[2 * k.H + y], b.side_in[3 * k.H + y],
                      k.pressure != 0, k.rho_out, k.c3);
        const float f2 = (k.y_walls && y == 0) ? f[4] : f[2];
        const float f4 = (k.y_walls && y == k.H - 1) ? f[2] : f[4];
        f[2] = f2;
        f[4] = f4;
    }
}

// FP64 reduction of the per-node health sums: s0 = sum rho, s1 = sum psi,
// s2 = non-finite count (sums), m3 = min(tau_mix - 1/2) (minimum).
struct Red4 {
    double s0, s1, s2, m3;
};

__device__ __forceinline__ Red4 red4_init() { return Red4{0.0, 0.0, 0.0, HUGE_VAL}; }

__device__ __forceinline__ Red4 wave_red4(Red4 r) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        r.s0 += __shfl_xor(r.s0, off, 64);
        r.s1 += __shfl_xor(r.s1, off, 64);
        r.s2 += __shfl_xor(r.s2, off, 64);
        r.m3 = fmin(r.m3, __shfl_xor(r.m3, off, 64));
    }
    return r;
}

__device__ __forceinline__ void put_red4(double* out, const Red4& r) {
    out[0] = r.s0;
    out[1] = r.s1;
    out[2] = r.s2;
    out[3] = r.m3;
}

template <int NT>
__device__ __forceinline__ void block_red4(Red4 r, double* out) {
    __shared__ double sm[4][NT / 64];
    r = wave_red4(r);
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    if ((t & 63) == 0) {
        sm[0][t >> 6] = r.s0;
        sm[1][t >> 6] = r.s1;
        sm[2][t >> 6] = r.s2;
        sm[3][t >> 6] = r.m3;
    }
    __syncthreads();
    if (t == 0) {
        Red4 a = red4_init();
        for (int w = 0; w < NT / 64; ++w) {
            a.s0 += sm[0][w];
            a.s1 += sm[1][w];
            a.s2 += sm[2][w];
            a.m3 = fmin(a.m3, sm[3][w]);
        }
        put_red4(out, a);
    }
}

// Health terms of one fluid node: rho and psi into the sums, a non-finite
// rho or psi into the count, tau_mix - 1/2 (FP32) into the minimum.  NaN
// drops out of the minimum (fmin); the count catches it.
__device__ __forceinline__ void red4_add(Red4& r, float rho, float psi, float tex) {
    r.s0 += (double)rho;
    r.s1 += (double)psi;
    if (!isfinite(rho) || !isfinite(psi)) r.s2 += 1.0;
    r.m3 = fmin(r.m3, (double)tex);
}

// Where a lane's column sits: inside the domain, on a ghost column of the
// psi stencil, or unused (beyond x = W in the last strip).
struct Lane {
    unsigned x;      // column (wrapped for periodic x), valid if inside
    bool inside;     // pulls from memory
    float ghost;     // psi value when !inside
};

__device__ __forceinline__ Lane lane_of(const KParams& k, int x) {
    Lane l;
    l.inside = true;
    l.ghost = 0.f;
    if (x < 0 || x >= k.W) {
        if (k.x_periodic) {
            x = ((x % k.W) + k.W) % k.W;
        } else {
            // ghost columns +1 at x = -1 and -1 at x = W (fingering_periodic.py:94-95)
            l.inside = false;
            l.ghost = (x < 0) ? 1.0f : -1.0f;
        }
    }
    l.x = (unsigned)x;
    return l;
}

// Flags of row y (wrapped) for this lane; ghost rows and ghost columns read
// as solid.  The load is unconditional (at a clamped, valid address) and
// the ghost value selected afterwards.
__device__ __forceinline__ unsigned flags_of(const KParams& k, const KBufs& b, const Lane& ln, int y) {
    const bool ghost = !ln.inside || (k.y_walls && (y < 0 || y >= k.H));
    const int yw = (y < 0) ? y + k.H : (y >= k.H ? y - k.H : y);
    const unsigned v = b.flags[yw * k.Wp + (ln.inside ? ln.x : 0u)];
    return ghost ? kSolid : v;
}

// Collision (MODE 0) or field output (MODE 1) of node (y, ln.x), given its
// pulled f (links applied, Zou-He not yet), its finished g and flags, and
// psi of the node and its eight neighbours in n[] (C, E, N, W, S, NE, NW,
// SW, SE).
//   on   the node is the lane's own (false beyond the domain: nothing is
//        written for it)
template <int MODE>
__device__ __forceinline__ void finish_node(const KParams& k, const KBufs& b, rsrc_t src, rsrc_t dst,
                                            const Lane& ln, const PullOff& po, int y, bool on, const float n[9],
                                            float f[9], const float gc[9], unsigned flc, Red4& red) {
    const float pc = n[0];
    float dpx, dpy, lap;
    fdn::stencil(n, dpx, dpy, lap);
    const int rc = y * k.Wp;
    bc_f(k, b, src, y, (int)ln.x, f);
    float g[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = gc[i];
    const bool solid = (flc & kSolid) != 0;
    if (MODE == 0) {
        if (on && !k.x_periodic && (ln.x == 0 || ln.x == (unsigned)k.W - 1)) {
            const float mu = fdn::chem_pot(pc, lap, k.c);
            const int s = (ln.x == 0) ? 0 : 2;
            b.side_out[s * k.H + y] = dpx * mu;
            b.side_out[(s + 1) * k.H + y] = dpy * mu;
        }
        // Collided on every lane, so the pulled registers are read on every
        // path and the stores below never wait for an older store.  On a
        // solid lane the collision is made the identity - zero relaxation
        // rates, zero force and velocity, finite moments - which streams
        // its populations on unchanged (store-and-forward) or keeps the
        // zeros of an inert solid.
        {
            fdn::Mom<float> m = fdn::moments(f, pc, dpx, dpy, lap, k.c);
            fdn::Consts<float> cl = k.c;
            if (solid) {
                m.rho = m.psi = m.mu = m.ux = m.uy = m.p = m.itau = m.fx = m.fy = 0.f;
                cl.inv_tau_g = 0.f;
            }
            fdn::collide(f, g, m, cl);
        }
        // Stores are predicated by offset like the loads, and written in
        // whole 64-byte pieces: a group of 16 lanes (64 B of one plane row)
        // stores all of its nodes - inert solid nodes their zeros - unless
        // all 16 are inert solids, which nothing reads.  A piece written only
        // in part costs far more HBM time than one written whole.
        const unsigned long long live = __ballot(!(solid && k.inert));
        const bool st = on && ((live >> (__lane_id() & 48u)) & 0xFFFFull) != 0ull;
        const unsigned vo = st ? po.xb : kOOB;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, f[i]), dst, vo, (rc + i * k.Q) * 4,
                                                  kStoreNT);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, g[i]), dst, vo,
                                                  (rc + (9 + i) * k.Q) * 4, kStoreNT);
        }
    } else if (on) {
        const unsigned o = rc + ln.x;
        float rho = 0.f, ux = 0.f, uy = 0.f;
        if (!solid) {
            const fdn::Mom<float> m = fdn::moments(f, pc, dpx, dpy, lap, k.c);
            rho = m.rho;
            ux = m.ux;
            uy = m.uy;
            red4_add(red, rho, pc, fdn::tau_excess(rho, pc, k.c));
        }
        b.fields[o] = pc;
        b.fields[k.Q + o] = rho;
        b.fields[2 * k.Q + o] = ux;
        b.fields[3 * k.Q + o] = uy;
        if (b.dst) {
            // inert solid nodes carry no populations (they were not pulled: zeros)
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                b.dst[i * k.Q + o] = f[i];
                b.dst[(9 + i) * k.Q + o] = g[i];
            }
        }
    }
}

// ---------------------------------------------------------------- wave march
// One wavefront per aligned 64-column strip and chunk of rows; no LDS and no
// barrier, so every wave streams on its own.  Lane l of strip s owns column
// x = 64 s + l.  psi of the two columns beside the strip (x0 - 1, x0 + 64) is
// pulled by lanes 0 and 63 in addition to their own column; psi of a lane's
// left and right neighbours comes by cross-lane shuffle.  The march keeps
// psi of rows y - 1, y, y + 1 and g of row y in registers.
//
// Each row issues all of its loads first - the flags two rows ahead, g of
// row y + 1, f of row y, the halo g - and every loaded value is consumed
// within the row, so no load is outstanding when the row's stores go out
// and the next row's loads follow them without a wait.  The flags are
// fetched a row ahead because the link bits choose the load addresses.

// psi of row yy at the lane's column and at its left / right neighbours
struct PsiRow {
    float c, l, r;
};

// Whether g of node (yy, ln) is loaded: an in-domain row and column, and
// not an inert solid (flags fl of that node).
__device__ __forceinline__ bool g_on(const KParams& k, const Lane& ln, int yy, unsigned fl) {
    return ln.inside && !(k.y_walls && (yy < 0 || yy >= k.H)) && !((fl & kSolid) && k.inert);
}

// Zou-He on the g pulled for row yy (lanes with on), then psi of the node:
// psi_wall on ghost rows and solid nodes, the ghost value on ghost columns,
// else the sum of g (fingering_periodic.py:470).
__device__ __forceinline__ float finish_g(const KParams& k, const Lane& ln, int yy, unsigned fl, bool on,
                                          float g[9]) {
    if (on) bc_g(k, wrap(yy, k.H), (int)ln.x, g);
    const float s = g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6] + g[7] + g[8];
    const bool ghost_row = k.y_walls && (yy < 0 || yy >= k.H);
    return ghost_row ? k.c.psi_wall : (!ln.inside ? ln.ghost : ((fl & kSolid) ? k.c.psi_wall : s));
}

// psi of row yy of the column beside the strip (lanes 0 and 63), flags fl.
__device__ __forceinline__ float psi_halo(const KParams& k, rsrc_t src, const Lane& lh, int yy, unsigned fl) {
    float g[9];
    const bool on = g_on(k, lh, yy, fl);
    pull(src, k, 9, rows_of(k, wrap(yy, k.H)), pull_offsets(k, lh.x), fl & kLinks, on, g);
    return finish_g(k, lh, yy, fl, on, g);
}

// Completes a PsiRow from the lane's own psi and the halo value (valid in
// lanes 0 and 63); every lane of the wave must execute it.
__device__ __forceinline__ PsiRow psi_row_of(float pcen, float halo, int lane) {
    const float up = __shfl_up(pcen, 1, 64);    // lane i <- lane i - 1
    const float dn = __shfl_down(pcen, 1, 64);  // lane i <- lane i + 1
    PsiRow p;
    p.c = pcen;
    p.l = (lane == 0) ? halo : up;
    p.r = (lane == 63) ? halo : dn;
    return p;
}

// MODE 0: one iteration (collide and store the next post-collision state).
// MODE 1: moments only - fields psi/rho/ux/uy, FP64 partial sums, and (if
//         b.dst) the post-stream populations; the state is not advanced.
// Workgroups of WPB waves cover WPB adjacent strips of one chunk of k.rows
// rows.  Blocks are numbered so that each XCD (workgroups are dealt
// round-robin over the eight XCDs) gets a contiguous band of chunks:
// neighbouring strips and the halo rows of neighbouring chunks then share
// that XCD's L2.  The FP64 partial sums of MODE 1 are per wave, at
// partials[4 * (WPB * L + wave)] (Red4).
// XCD-banded logical index of this workgroup: workgroups are dealt
// round-robin over the eight XCDs, so XCD j runs blocks j, j+8, ...; they
// are renumbered to one contiguous range per XCD.
__device__ __forceinline__ int banded_block() {
    const int nb = gridDim.x, q = nb >> 3, rem = nb & 7, xcd = blockIdx.x & 7;
    return xcd * q + min(xcd, rem) + (blockIdx.x >> 3);
}

// The march of logical workgroup L of one domain (see step_kernel).
template <int WPB, int MODE>
__device__ __forceinline__ void march_block(const KParams& k, const KBufs& b, int L) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = L / k.nsb;
    const int strip = (k.sb0 + L - chunk * k.nsb) * WPB + wave;
    double* part = b.partials + 4 * (WPB * L + wave);
    const int x0 = strip * 64;
    if (x0 >= k.W) {
        if (MODE == 1 && lane == 0) put_red4(part, red4_init());
        return;
    }
    const int xcol = x0 + lane;
    const Lane ln = lane_of(k, xcol);
    const Lane lh = lane_of(k, lane == 0 ? x0 - 1 : x0 + 64);
    const bool hl = (lane == 0) || (lane == 63);
    const bool interior = xcol < k.W;
    const int y0 = chunk * k.rows;
    const int y1 = min(y0 + k.rows, k.H);

    const unsigned bytes = 18u * (unsigned)k.Q * 4u;
    const rsrc_t src = make_rsrc(b.src, bytes);
    const rsrc_t dst = make_rsrc(b.dst, b.dst ? bytes : 0u);
    Red4 red = red4_init();
    const PullOff po = pull_offsets(k, ln.x);

    // Even chunks march up, odd chunks down: the two chunks beside a chunk
    // boundary then pull its rows at the same time (both first or both
    // last), so the second pull of those rows hits the L2.
    const int d = (chunk & 1) ? -1 : 1;
    const int ys = (d > 0) ? y0 : y1 - 1;

    // prologue: psi of rows ys - d (behind) and ys, g of row ys
    float gc[9];
    PsiRow pb, pc;
    {
        const int yb = ys - d;
        const unsigned fb = flags_of(k, b, ln, yb), fbh = flags_of(k, b, lh, yb);
        const bool on = g_on(k, ln, yb, fb);
        pull(src, k, 9, rows_of(k, wrap(yb, k.H)), po, fb & kLinks, on, gc);
        float h = 0.f;
        if (hl) h = psi_halo(k, src, lh, yb, fbh);
        pb = psi_row_of(finish_g(k, ln, yb, fb, on, gc), h, lane);
    }
    unsigned flc = flags_of(k, b, ln, ys);
    {
        const unsigned fch = flags_of(k, b, lh, ys);
        const bool on = g_on(k, ln, ys, flc);
        pull(src, k, 9, rows_of(k, ys), po, flc & kLinks, on, gc);
        float h = 0.f;
        if (hl) h = psi_halo(k, src, lh, ys, fch);
        pc = psi_row_of(finish_g(k, ln, ys, flc, on, gc), h, lane);
    }
    unsigned fla = flags_of(k, b, ln, ys + d);
    unsigned flah = flags_of(k, b, lh, ys + d);

    for (int j = 0, y = ys; j < y1 - y0; ++j, y += d) {
        // row y + d is the one ahead; its psi completes the stencil of row y
        const int ya = y + d;
        const unsigned fl2 = flags_of(k, b, ln, ya + d);
        const unsigned fl2h = flags_of(k, b, lh, ya + d);
        float ga[9], f[9];
        const bool ga_on = g_on(k, ln, ya, fla);
        pull(src, k, 9, rows_of(k, wrap(ya, k.H)), po, fla & kLinks, ga_on, ga);
        const bool f_on = interior && !((flc & kSolid) && k.inert);
        pull(src, k, 0, rows_of(k, y), po, flc & kLinks, f_on, f);
        float h = 0.f;
        if (hl) h = psi_halo(k, src, lh, ya, flah);
        const PsiRow pa = psi_row_of(finish_g(k, ln, ya, fla, ga_on, ga), h, lane);
        const PsiRow& pn = (d > 0) ? pa : pb;  // row y + 1
        const PsiRow& ps = (d > 0) ? pb : pa;  // row y - 1
        const float nb9[9] = {pc.c, pc.r, pn.c, pc.l, ps.c, pn.r, pn.l, ps.l, ps.r};
        finish_node<MODE>(k, b, src, dst, ln, po, y, interior, nb9, f, gc, interior ? flc : kSolid, red);
        pb = pc;
        pc = pa;
#pragma unroll
        for (int i = 0; i < 9; ++i) gc[i] = ga[i];
        flc = fla;
        fla = fl2;
        flah = fl2h;
    }
    if (MODE == 1) {
        red = wave_red4(red);
        if (lane == 0) put_red4(part, red);
    }
}

template <int WPB, int MODE>
__global__ __launch_bounds__(64 * WPB) void step_kernel(KParams k, KBufs b) {
    march_block<WPB, MODE>(k, b, banded_block());
}

// ---------------------------------------------------- two-iteration march
// One launch advances the state by two iterations: S_n -> S_n+2, with
// S_n+1 kept in registers and never written to HBM, so each iteration moves
// half the bytes of a single-iteration launch.
//
// A wave owns a strip of kFuseCols output columns [xb0, xb0 + 58) and its
// 64 lanes hold columns a0 = xb0 - 3 ... xb0 + 60.  Stage A is the
// single-iteration march (S_n from HBM -> S_n+1 of the strip, P_n+1 of the
// Zou-He columns), valid on lanes 1..62 (its psi stencil takes the
// neighbours' psi from the neighbouring lanes); stage B marches two rows behind it
// and pulls S_n+1 from A's last rows: g from registers, where a population
// moving in x comes from the neighbouring lane by a DPP lane shift
// (wave_shr:1 / wave_shl:1), f from a four-row ring in LDS read at the
// shifted lane.  B needs A's output two columns either side of a node (one
// for the f pull, one more for the g pulls of the psi stencil), so it is
// valid on lanes 3..60: strips overlap by six columns of stage-A work.
// For the chunk of output rows [y0, y1) stage A runs rows y0-2 ... y1+1,
// upwards in even chunks and downwards in odd ones (B then lags two rows
// behind in the march direction).
//
// Every node of either stage runs the single-iteration kernel's arithmetic
// on the same values in the same order, so the result equals two
// single-iteration launches bit for bit.
constexpr int kFuseCols = 58;

__device__ __forceinline__ float lane_from_left(float v) {  // lane l <- lane l-1 (DPP wave_shr:1)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_from_right(float v) {  // lane l <- lane l+1 (DPP wave_shl:1)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}
// psi of a row at this lane and its left / right neighbour lanes (lanes 0
// and 63 get 0 for the missing neighbour; their stencils are not used)
__device__ __forceinline__ PsiRow lane_row(float c) {
    PsiRow p;
    p.c = c;
    p.l = lane_from_left(c);
    p.r = lane_from_right(c);
    return p;
}

// value v of the lane at column x - ex (ex in {-1, 0, 1})
__device__ __forceinline__ float lane_pull(float v, int ex) {
    return ex > 0 ? lane_from_left(v) : (ex < 0 ? lane_from_right(v) : v);
}

// Stage B's pull of one distribution for the node of row y at this lane
// (g) from stage A's post-collision rows y-1 (lo, only directions 2, 5, 6
// are read), y (mid) and y+1 (hi); a linked direction i takes the node's own
// mid[opp(i)] (half-way bounce-back).  An inert solid node has every link
// bit set and stage A left zeros in it, so it pulls zeros, as the
// single-iteration kernel's unloaded lanes do; lanes on ghost columns pull
// values that are never used.
__device__ __forceinline__ void reg_pull(const float lo[9], const float mid[9], const float hi[9], unsigned links,
                                         float v[9]) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const float* row = fdn::EY[i] > 0 ? lo : (fdn::EY[i] < 0 ? hi : mid);
        float x = lane_pull(row[i], fdn::EX[i]);
        if (i > 0) x = sel_bit(links, i - 1, mid[fdn::OPP[i]], x);
        v[i] = x;
    }
}

// Stage A's post-collision f of the last four rows, for stage B, in LDS:
// per wave 4 slots (row & 3) x 9 planes x 64 lanes, with one float of
// padding at either end of the wave's region for the lane-shifted reads of
// lanes 0 and 63 (their values are never used).
constexpr int kRingSlot = 9 * 64;
constexpr int kRingRegion = 4 * kRingSlot + 2;

// Stage B's pull of f for the node of row y at this lane from the ring:
// direction i from row y - e_iy at lane - e_ix, or, linked, the node's own
// opposite population of row y (half-way bounce-back) - one LDS read at one
// of two addresses.  Inert solid nodes pull zeros (see reg_pull).
__device__ __forceinline__ void lds_pull(const float* ring, int y, int lane, unsigned links, float v[9]) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        int idx = ((y - fdn::EY[i]) & 3) * kRingSlot + i * 64 + lane - fdn::EX[i];
        if (i > 0) idx = (int)sel_bit(links, i - 1, (unsigned)((y & 3) * kRingSlot + fdn::OPP[i] * 64 + lane), (unsigned)idx);
        v[i] = ring[idx];
    }
}

// Inlet density (fingering_periodic.py:276-277) of the node of row y at
// column 0 from stage A's rows y-1, y, y+1 in the ring: the adjacent-row
// density of the walled inlet corners (fingering.py:334-364).
__device__ __forceinline__ float ring_inlet_rho(const float* ring, int y, int lane, unsigned links, float u,
                                                float px) {
    float f[9];
    lds_pull(ring, y, lane, links, f);
    return fdn::inlet_rho(f, u, px);
}

// Stage B's Zou-He / walled-corner update of f at node (y, x): as bc_f,
// with the side values P (of stage A's psi) and the corner density given.
__device__ __forceinline__ void reg_bc_f(const KParams& k, const KBufs& b, int y, int x, float f[9], float px,
                                         float py, float rho_nb) {
    if (k.x_periodic) return;
    if (x == 0) {
        const bool bot = k.y_walls && y == 0;
        const bool top = k.y_walls && y == k.H - 1;
        float n[9], cb[9], ct[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) n[i] = cb[i] = ct[i] = f[i];
        fdn::inlet_f(n, b.uprof[y], px, py);
        fdn::inlet_corner_bottom_f(cb, rho_nb);
        fdn::inlet_corner_top_f(ct, rho_nb);
#pragma unroll
        for (int i = 0; i < 9; ++i) f[i] = bot ? cb[i] : (top ? ct[i] : n[i]);
    }
    if (x == k.W - 1) {
        fdn::outlet_f(f, b.uprof[y], px, py, k.pressure != 0, k.rho_out, k.c3);
        const float f2 = (k.y_walls && y == 0) ? f[4] : f[2];
        const float f4 = (k.y_walls && y == k.H - 1) ? f[2] : f[4];
        f[2] = f2;
        f[4] = f4;
    }
}

// Moments and collision of one node (identity on solid lanes), as in
// finish_node MODE 0; returns mu for the side values.
__device__ __forceinline__ float collide_node(const KParams& k, float f[9], float g[9], float pc, float dpx,
                                              float dpy, float lap, bool solid) {
    fdn::Mom<float> m = fdn::moments(f, pc, dpx, dpy, lap, k.c);
    const float mu = m.mu;
    fdn::Consts<float> cl = k.c;
    if (solid) {
        m.rho = m.psi = m.mu = m.ux = m.uy = m.p = m.itau = m.fx = m.fy = 0.f;
        cl.inv_tau_g = 0.f;
    }
    fdn::collide(f, g, m, cl);
    return mu;
}

// Two iterations of logical workgroup L of a launch over one domain.
template <int WPB>
__device__ __forceinline__ void fused_march(const KParams& k, const KBufs& b, int L) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = L / k.nsb;
    const int strip = (L - chunk * k.nsb) * WPB + wave;
    const int xb0 = strip * kFuseCols;
    if (xb0 >= k.W) return;
    const int a0 = xb0 - 3;
    const int xa = a0 + lane;
    const Lane ln = lane_of(k, xa);
    const bool bout = lane >= 3 && lane < 3 + kFuseCols && xa < k.W;   // stage B stores this lane
    const bool zx = !k.x_periodic;
    const bool has_in = zx && xb0 == 0;                                  // column 0 among the outputs
    const int y0 = chunk * k.rows;
    const int y1 = min(y0 + k.rows, k.H);

    const unsigned bytes = 18u * (unsigned)k.Q * 4u;
    const rsrc_t src = make_rsrc(b.src, bytes);
    const rsrc_t dst = make_rsrc(b.dst, bytes);
    const PullOff po = pull_offsets(k, ln.x);
    __shared__ float fring_all[WPB * kRingRegion];
    float* const fring = fring_all + wave * kRingRegion + 1;

    // Even chunks march up (d = 1), odd chunks down (d = -1).  Stage A's
    // rows beyond the chunk (two on either side) are read by both chunks
    // beside a boundary; with alternating directions both read them at the
    // same time (both first or both last), so the second read hits the L2.
    const int d = (chunk & 1) ? -1 : 1;
    const int ys = d > 0 ? y0 - 2 : y1 + 1;   // stage A's first row
    const int yf = d > 0 ? y0 : y1 - 1;       // stage B's first row

    // ---- stage A prologue: psi_n+1 of rows ys-d, ys and g of row ys
    float gc[9];
    PsiRow pb, pc;
    {
        const int yb = ys - d;
        const unsigned fb = flags_of(k, b, ln, yb);
        const bool on = g_on(k, ln, yb, fb);
        pull(src, k, 9, rows_of(k, wrap(yb, k.H)), po, fb & kLinks, on, gc);
        pb = lane_row(finish_g(k, ln, yb, fb, on, gc));
    }
    unsigned flc = flags_of(k, b, ln, ys);
    {
        const bool on = g_on(k, ln, ys, flc);
        pull(src, k, 9, rows_of(k, wrap(ys, k.H)), po, flc & kLinks, on, gc);
        pc = lane_row(finish_g(k, ln, ys, flc, on, gc));
    }
    unsigned fla = flags_of(k, b, ln, ys + d);

    // ---- the hand-over: A's post-collision f of rows r, r-d, r-2d, r-3d
    // in the LDS ring (slot row & 3); its g of rows r (Ga), r-d (Gb), r-2d
    // (Gc: only the three directions pulled from two rows back are read) in
    // registers; A's flags of rows r, r-d, r-2d; A's side values P of the
    // same rows
    float Ga[9], Gb[9], Gc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Ga[i] = Gb[i] = Gc[i] = 0.f;
    unsigned fr0 = kSolid, fr1 = kSolid, fr2 = kSolid;
    float pxa = 0.f, pya = 0.f, pxb = 0.f, pyb = 0.f, pxc = 0.f, pyc = 0.f;
    // ---- stage B state: psi_n+2 of rows y-d (qs), y (qc); g of row y (gq);
    // inlet density of the row behind (walled inlet corners)
    PsiRow qs{}, qc{};
    float gq[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) gq[i] = 0.f;
    float rho_prev = 0.f;

    const int nj = (y1 - y0) + 4;
    for (int j = 0; j < nj; ++j) {
        const int r = ys + j * d;    // stage A row
        const bool arow = !(k.y_walls && (r < 0 || r >= k.H));
        const int rw = wrap(r, k.H);
        // ---------------- stage A, row r
        float Gn[9], pxn = 0.f, pyn = 0.f;
        {
            const int ya = r + d;
            const unsigned fl2 = flags_of(k, b, ln, ya + d);
            float ga[9], f[9];
            const bool ga_on = g_on(k, ln, ya, fla);
            pull(src, k, 9, rows_of(k, wrap(ya, k.H)), po, fla & kLinks, ga_on, ga);
            const bool f_on = arow && ln.inside && !((flc & kSolid) && k.inert);
            pull(src, k, 0, rows_of(k, rw), po, flc & kLinks, f_on, f);
            const PsiRow pa = lane_row(finish_g(k, ln, ya, fla, ga_on, ga));
            const PsiRow pn = d > 0 ? pa : pb;  // row r + 1
            const PsiRow ps = d > 0 ? pb : pa;  // row r - 1
            const float nb9[9] = {pc.c, pc.r, pn.c, pc.l, ps.c, pn.r, pn.l, ps.l, ps.r};
            float dpx, dpy, lap;
            fdn::stencil(nb9, dpx, dpy, lap);
            if (arow) bc_f(k, b, src, rw, (int)ln.x, f);
#pragma unroll
            for (int i = 0; i < 9; ++i) Gn[i] = gc[i];
            const bool solid = !arow || !ln.inside || (flc & kSolid) != 0;
            const float mu = collide_node(k, f, Gn, pc.c, dpx, dpy, lap, solid);
            float* const slot = fring + (r & 3) * kRingSlot + lane;
#pragma unroll
            for (int i = 0; i < 9; ++i) slot[i * 64] = f[i];
            // P = d psi * mu of this row's psi at the Zou-He columns (the
            // value the next iteration's Zou-He reads)
            pxn = dpx * mu;
            pyn = dpy * mu;
            pb = pc;
            pc = pa;
#pragma unroll
            for (int i = 0; i < 9; ++i) gc[i] = ga[i];
            fr2 = fr1;
            fr1 = fr0;
            fr0 = arow && ln.inside ? flc : kSolid;
            flc = fla;
            fla = fl2;
        }
        // ---------------- hand-over
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            Gc[i] = Gb[i];
            Gb[i] = Ga[i];
            Ga[i] = Gn[i];
        }
        pxc = pxb;
        pyc = pyb;
        pxb = pxa;
        pyb = pya;
        pxa = pxn;
        pya = pyn;
        if (j < 2) continue;
        // ---------------- stage B: psi_n+2 and g of row r-d
        PsiRow qn;
        float gn[9];
        {
            const int yq = r - d;
            const bool on = ln.inside && !(k.y_walls && (yq < 0 || yq >= k.H)) && !((fr1 & kSolid) && k.inert);
            float lo[9], hi[9];  // A's rows yq - 1 and yq + 1
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                lo[i] = d > 0 ? Gc[i] : Ga[i];
                hi[i] = d > 0 ? Ga[i] : Gc[i];
            }
            reg_pull(lo, Gb, hi, fr1 & kLinks, gn);
            qn = lane_row(finish_g(k, ln, yq, fr1, on, gn));
        }
        // walled inlet corners need the inlet density of the adjacent row:
        // the row behind stage B's first row is one it has not marched
        if (j == 3 && has_in && k.y_walls && yf - d >= 0 && yf - d < k.H)
            rho_prev = ring_inlet_rho(fring, yf - d, lane, fr2 & kLinks, b.uprof[yf - d], pxc);
        if (j >= 4) {
            // ------------ stage B: collision of row y = r-2d
            const int y = r - 2 * d;
            const unsigned fl = fr2;
            const bool solid = (fl & kSolid) != 0;
            float f[9];
            lds_pull(fring, y, lane, fl & kLinks, f);
            float rho_nb = rho_prev;
            if (has_in) {
                // the corner met first (row 0 marching up, H-1 marching
                // down) takes its neighbour row (ahead) from the ring
                if (k.y_walls && y == (d > 0 ? 0 : k.H - 1))
                    rho_nb = ring_inlet_rho(fring, y + d, lane, fr1 & kLinks, b.uprof[y + d], pxb);
                rho_prev = fdn::inlet_rho(f, b.uprof[y], pxc);
            }
            const PsiRow qa = d > 0 ? qn : qs;  // row y + 1
            const PsiRow qb = d > 0 ? qs : qn;  // row y - 1
            const float nb9[9] = {qc.c, qc.r, qa.c, qc.l, qb.c, qa.r, qa.l, qb.l, qb.r};
            float dpx, dpy, lap;
            fdn::stencil(nb9, dpx, dpy, lap);
            reg_bc_f(k, b, y, (int)ln.x, f, pxc, pyc, rho_nb);
            float g[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) g[i] = gq[i];
            const float mu = collide_node(k, f, g, qc.c, dpx, dpy, lap, solid || !ln.inside);
            if (bout && zx && (ln.x == 0 || ln.x == (unsigned)k.W - 1)) {
                const int sd = (ln.x == 0) ? 0 : 2;
                b.side_out[sd * k.H + y] = dpx * mu;
                b.side_out[(sd + 1) * k.H + y] = dpy * mu;
            }
            const unsigned vo = (bout && !(fl & kDeadPiece)) ? ln.x * 4u : kOOB;
            const int rc = y * k.Wp;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, f[i]), dst, vo, (rc + i * k.Q) * 4,
                                                      0);
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, g[i]), dst, vo,
                                                      (rc + (9 + i) * k.Q) * 4, 0);
            }
        }
        qs = qc;
        qc = qn;
#pragma unroll
        for (int i = 0; i < 9; ++i) gq[i] = gn[i];
    }
}

template <int WPB>
__global__ __launch_bounds__(64 * WPB) void fused_kernel(KParams k, KBufs b) {
    fused_march<WPB>(k, b, banded_block());
}

// ---------------------------------------------------- three-iteration march
// One launch advances the state by three iterations, S_n -> S_n+3: stage A
// as above; stage B (two rows behind A) keeps S_n+2 on chip - g in a second
// register ring, f in a second LDS ring - instead of storing it; stage C
// (two rows behind B) pulls S_n+2 from them, collides and stores S_n+3.
// Each stage needs two columns more of the stage before it, so the 64
// lanes hold columns x0 - 5 ... x0 + 58 and stage C's 54 output columns
// are lanes 5..58.  For the chunk [y0, y1) stage A runs rows y0-4 ... y1+3
// (upwards) and stage B rows y0-2 ... y1+1, which may lie beyond a walled domain's
// edge rows (ghost rows: psi_wall, no populations) or wrap around a
// periodic one.  The arithmetic of every node is the single-iteration
// kernel's, in the same order, so the result equals three single-iteration
// launches bit for bit.
constexpr int kFuse3Cols = 54;

// Rows per wave of the three-iteration kernel.  Measured on 4096 x 2048 (76
// strips, 2,048 concurrent waves), ms per iteration: 12 rows 0.153, 14
// 0.151, 15 0.150, 16 0.1455, 17 0.148, 18 0.148, 20 0.148, 22 0.152, 24
// 0.155, 26 0.154, 32 0.156.  Taller chunks save stage A's 8 extra rows per
// chunk but run slower per row; shorter ones pay more extra rows.
constexpr int kFuse3Rows = 16;

// rings_all: the workgroup's LDS rings, WPB * 2 * kRingRegion floats,
// declared by the kernel (one array whatever copies of the march it holds).
template <int WPB, bool EDGE>
__device__ __forceinline__ void fused3_march(const KParams& k, const KBufs& b, int chunk, int strip,
                                             float* rings_all) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int xc0 = strip * kFuse3Cols;
    if (xc0 >= k.W) return;
    const int xa = xc0 - 5 + lane;
    // !EDGE: every lane is a domain column other than 0 and W-1 (no ghost
    // column, no wrap, no Zou-He column), so the lane predicates are
    // constant and the Zou-He code is left out
    Lane ln = lane_of(k, xa);
    if (!EDGE) {
        ln.inside = true;
        ln.x = (unsigned)xa;
        ln.ghost = 0.f;
    }
    const bool cout = lane >= 5 && lane < 5 + kFuse3Cols && (!EDGE || xa < k.W);   // stage C stores this lane
    const bool zx = !k.x_periodic;
    const bool has_in = EDGE && zx && xc0 == 0;                                     // column 0 among the outputs
    const int y0 = chunk * k.rows;
    const int y1 = min(y0 + k.rows, k.H);

    const unsigned bytes = 18u * (unsigned)k.Q * 4u;
    const rsrc_t src = make_rsrc(b.src, bytes);
    const rsrc_t dst = make_rsrc(b.dst, bytes);
    const PullOff po = pull_offsets(k, ln.x);
    float* const fring = rings_all + wave * 2 * kRingRegion + 1;    // A -> B
    float* const fring2 = fring + kRingRegion;                       // B -> C

    // Every chunk marches up.  This kernel is bound by its arithmetic, not
    // by HBM, so the L2 reuse that alternating directions buy the
    // two-iteration kernel is worth less here than the direction selects
    // it costs (measured on 4096 x 2048 at 1 x 26: 0.1617 ms per iteration
    // up-only, 0.1641 alternating).  The row expressions keep d for the
    // reading.
    constexpr int d = 1;
    const int ys = d > 0 ? y0 - 4 : y1 + 3;   // stage A's first row
    const int yfb = ys + 2 * d;               // stage B's first row
    const int yfc = ys + 4 * d;               // stage C's first row

    // ---- stage A prologue: psi_n+1 of rows ys-d, ys and g of row ys
    // GA: g pulled by stage A, two slots: at a step of phase P (= j mod 2),
    // slot P holds row r (for the collision) and slot 1-P row r+d (for psi);
    // the pulls of row r+2d go to slot P once its row r is consumed, so the
    // loads in flight are never moved between registers.
    float GA[2][9];
    PsiRow pb, pc;
    {
        const int yb = ys - d;
        const unsigned fb = flags_of(k, b, ln, yb);
        const bool on = g_on(k, ln, yb, fb);
        pull(src, k, 9, rows_of(k, wrap(yb, k.H)), po, fb & kLinks, on, GA[0]);
        pb = lane_row(finish_g(k, ln, yb, fb, EDGE && on, GA[0]));
    }
    unsigned flc = flags_of(k, b, ln, ys);
    {
        const bool on = g_on(k, ln, ys, flc);
        pull(src, k, 9, rows_of(k, wrap(ys, k.H)), po, flc & kLinks, on, GA[0]);
        pc = lane_row(finish_g(k, ln, ys, flc, EDGE && on, GA[0]));
    }
    unsigned fla = flags_of(k, b, ln, ys + d);
    // Flags are loaded two rows ahead: the pulls issued after stage A need
    // the flags of row r + 2d, and a flag load issued at the top of the same
    // step would be younger than stage C's stores of the step before, which
    // the wait for it would then cover too (0.1543 -> 0.1455 ms per iteration
    // on 4096 x 2048 at 1 x 16).
    unsigned flb = flags_of(k, b, ln, ys + 2 * d);
    // Stage A's pulls run one row ahead: the loads for row r + d are issued
    // right after stage A of row r, so stages B and C of this row cover
    // their latency.  The first row's are issued here.
    float f_n[9];
    pull(src, k, 9, rows_of(k, wrap(ys + d, k.H)), po, fla & kLinks, true, GA[1]);
    pull(src, k, 0, rows_of(k, wrap(ys, k.H)), po, flc & kLinks, true, f_n);
    // Stages B and C read their f rings one row ahead as well (the rows
    // they read are complete one row earlier).
    float fb_n[9], fc_n[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) fb_n[i] = fc_n[i] = 0.f;

    // A -> B: g of A's rows r (Ga), r-d (Gb), r-2d (Gc); side values P of
    // the same rows.  Flags of rows r, r-d, ..., r-4d (ghost rows solid).
    float Ga[9], Gb[9], Gc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Ga[i] = Gb[i] = Gc[i] = 0.f;
    unsigned fr0 = kSolid, fr1 = kSolid, fr2 = kSolid, fr3 = kSolid, fr4 = kSolid;
    float pxa = 0.f, pya = 0.f, pxb = 0.f, pyb = 0.f, pxc = 0.f, pyc = 0.f;
    // stage B: psi_n+2 of rows behind (qs) and current (qc), g of the current row
    PsiRow qs{}, qc{};
    float gq[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) gq[i] = 0.f;
    float rho_prev = 0.f;
    // B -> C: g of B's rows r-2d (Ha), r-3d (Hb), r-4d (Hc) and their P
    float Ha[9], Hb[9], Hc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Ha[i] = Hb[i] = Hc[i] = 0.f;
    float qxa = 0.f, qya = 0.f, qxb = 0.f, qyb = 0.f, qxc = 0.f, qyc = 0.f;
    // stage C: psi_n+3 of rows behind (us) and current (uc), g of the current row
    PsiRow us{}, uc{};
    float gu[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) gu[i] = 0.f;
    float rho_prev_c = 0.f;

    const int nj = (y1 - y0) + 8;
    auto step = [&](auto phase, const int j) __attribute__((always_inline)) {
        constexpr int P = decltype(phase)::value;
        const int r = ys + j * d;    // stage A row
        const bool arow = !(k.y_walls && (r < 0 || r >= k.H));
        const int rw = wrap(r, k.H);
        // ---------------- stage A, row r
        float Gn[9], pxn = 0.f, pyn = 0.f;
        {
            const int ya = r + d;
            const unsigned fl3 = flags_of(k, b, ln, ya + 2 * d);
            // every lane loads: what inert solids, ghost rows and lanes
            // beyond the domain load is never used by a node of the domain
            // (psi and the collision of those lanes do not depend on it, and
            // every pull of a domain node from them is a bounce-back link or
            // rewritten by Zou-He), and inert solid nodes hold zeros in
            // memory, which they keep; so no load needs a mask here
            float (&ga)[9] = GA[1 - P];   // g pulled at row ya (issued one row ahead)
            float (&f)[9] = f_n;          // f pulled at row r
            const bool ga_on = g_on(k, ln, ya, fla);
            const PsiRow pa = lane_row(finish_g(k, ln, ya, fla, EDGE && ga_on, ga));
            const PsiRow pn = d > 0 ? pa : pb;  // row r + 1
            const PsiRow ps = d > 0 ? pb : pa;  // row r - 1
            const float nb9[9] = {pc.c, pc.r, pn.c, pc.l, ps.c, pn.r, pn.l, ps.l, ps.r};
            float dpx, dpy, lap;
            fdn::stencil(nb9, dpx, dpy, lap);
            if (EDGE && arow) bc_f(k, b, src, rw, (int)ln.x, f);
#pragma unroll
            for (int i = 0; i < 9; ++i) Gn[i] = GA[P][i];
            const bool solid = !arow || !ln.inside || (flc & kSolid) != 0;
            const float mu = collide_node(k, f, Gn, pc.c, dpx, dpy, lap, solid);
            float* const slot = fring + (r & 3) * kRingSlot + lane;
#pragma unroll
            for (int i = 0; i < 9; ++i) slot[i * 64] = f[i];
            pxn = dpx * mu;
            pyn = dpy * mu;
            pb = pc;
            pc = pa;
            fr4 = fr3;
            fr3 = fr2;
            fr2 = fr1;
            fr1 = fr0;
            fr0 = arow && ln.inside ? flc : kSolid;
            flc = fla;
            fla = flb;
            flb = fl3;
            // the next row's pulls (rows wrap, so the loads past the
            // chunk's last row stay inside the allocation; they are unused)
            pull(src, k, 9, rows_of(k, wrap(ya + d, k.H)), po, fla & kLinks, true, GA[P]);
            pull(src, k, 0, rows_of(k, wrap(r + d, k.H)), po, flc & kLinks, true, f_n);
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            Gc[i] = Gb[i];
            Gb[i] = Ga[i];
            Ga[i] = Gn[i];
        }
        pxc = pxb;
        pyc = pyb;
        pxb = pxa;
        pyb = pya;
        pxa = pxn;
        pya = pyn;
        if (j < 2) return;
        // ---------------- stage B: psi_n+2 and g of row r-d
        {
            PsiRow qn;
            float gn[9];
            {
                const int yq = r - d;
                const bool on = ln.inside && !(k.y_walls && (yq < 0 || yq >= k.H)) && !((fr1 & kSolid) && k.inert);
                float lo[9], hi[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) {
                    lo[i] = d > 0 ? Gc[i] : Ga[i];
                    hi[i] = d > 0 ? Ga[i] : Gc[i];
                }
                reg_pull(lo, Gb, hi, fr1 & kLinks, gn);
                qn = lane_row(finish_g(k, ln, yq, fr1, EDGE && on, gn));
            }
            if (j == 3 && has_in && k.y_walls && yfb - d >= 0 && yfb - d < k.H)
                rho_prev = ring_inlet_rho(fring, yfb - d, lane, fr2 & kLinks, b.uprof[yfb - d], pxc);
            float Hn[9], qxn = 0.f, qyn = 0.f;
            if (j >= 4) {
                // -------- stage B: collision of row y = r-2d into the rings
                const int y = r - 2 * d;
                const bool brow = !(k.y_walls && (y < 0 || y >= k.H));
                const int yw = wrap(y, k.H);
                const unsigned fl = fr2;
                const bool solid = (fl & kSolid) != 0;
                float f[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) f[i] = fb_n[i];
                float rho_nb = rho_prev;
                if (has_in && brow) {
                    if (k.y_walls && y == (d > 0 ? 0 : k.H - 1))
                        rho_nb = ring_inlet_rho(fring, y + d, lane, fr1 & kLinks, b.uprof[y + d], pxb);
                    rho_prev = fdn::inlet_rho(f, b.uprof[yw], pxc);
                }
                const PsiRow qa = d > 0 ? qn : qs;
                const PsiRow qb = d > 0 ? qs : qn;
                const float nb9[9] = {qc.c, qc.r, qa.c, qc.l, qb.c, qa.r, qa.l, qb.l, qb.r};
                float dpx, dpy, lap;
                fdn::stencil(nb9, dpx, dpy, lap);
                if (EDGE && brow) reg_bc_f(k, b, yw, (int)ln.x, f, pxc, pyc, rho_nb);
#pragma unroll
                for (int i = 0; i < 9; ++i) Hn[i] = gq[i];
                const float mu = collide_node(k, f, Hn, qc.c, dpx, dpy, lap, solid || !ln.inside);
                float* const slot = fring2 + (y & 3) * kRingSlot + lane;
#pragma unroll
                for (int i = 0; i < 9; ++i) slot[i * 64] = f[i];
                qxn = dpx * mu;
                qyn = dpy * mu;
            } else {
#pragma unroll
                for (int i = 0; i < 9; ++i) Hn[i] = 0.f;
            }
            qs = qc;
            qc = qn;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                gq[i] = gn[i];
                Hc[i] = Hb[i];
                Hb[i] = Ha[i];
                Ha[i] = Hn[i];
            }
            qxc = qxb;
            qyc = qyb;
            qxb = qxa;
            qyb = qya;
            qxa = qxn;
            qya = qyn;
            // B's pull of the next row, r - d: rows r - 2d .. r are in the ring
            lds_pull(fring, r - d, lane, fr1 & kLinks, fb_n);
        }
        if (j < 6) return;
        // ---------------- stage C: psi_n+3 and g of row r-3d
        PsiRow un;
        float gv[9];
        {
            const int yq = r - 3 * d;
            const bool on = ln.inside && !(k.y_walls && (yq < 0 || yq >= k.H)) && !((fr3 & kSolid) && k.inert);
            float lo[9], hi[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                lo[i] = d > 0 ? Hc[i] : Ha[i];
                hi[i] = d > 0 ? Ha[i] : Hc[i];
            }
            reg_pull(lo, Hb, hi, fr3 & kLinks, gv);
            un = lane_row(finish_g(k, ln, yq, fr3, EDGE && on, gv));
        }
        if (j == 6 && has_in && k.y_walls && yfc - d >= 0 && yfc - d < k.H)
            rho_prev_c = ring_inlet_rho(fring2, yfc - d, lane, fr3 & kLinks, b.uprof[yfc - d], qxb);
        if (j >= 8) {
            // ------------ stage C: collision of row y = r-4d, stored
            const int y = r - 4 * d;
            const unsigned fl = fr4;
            const bool solid = (fl & kSolid) != 0;
            float f[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) f[i] = fc_n[i];
            float rho_nb = rho_prev_c;
            if (has_in) {
                if (k.y_walls && y == (d > 0 ? 0 : k.H - 1))
                    rho_nb = ring_inlet_rho(fring2, y + d, lane, fr3 & kLinks, b.uprof[y + d], qxb);
                rho_prev_c = fdn::inlet_rho(f, b.uprof[y], qxc);
            }
            const PsiRow ua = d > 0 ? un : us;
            const PsiRow ub = d > 0 ? us : un;
            const float nb9[9] = {uc.c, uc.r, ua.c, uc.l, ub.c, ua.r, ua.l, ub.l, ub.r};
            float dpx, dpy, lap;
            fdn::stencil(nb9, dpx, dpy, lap);
            if (EDGE) reg_bc_f(k, b, y, (int)ln.x, f, qxc, qyc, rho_nb);
            float g[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) g[i] = gu[i];
            const float mu = collide_node(k, f, g, uc.c, dpx, dpy, lap, solid || !ln.inside);
            if (EDGE && cout && zx && (ln.x == 0 || ln.x == (unsigned)k.W - 1)) {
                const int sd = (ln.x == 0) ? 0 : 2;
                b.side_out[sd * k.H + y] = dpx * mu;
                b.side_out[(sd + 1) * k.H + y] = dpy * mu;
            }
            const unsigned vo = (cout && !(fl & kDeadPiece)) ? ln.x * 4u : kOOB;
            const int rc = y * k.Wp;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, f[i]), dst, vo, (rc + i * k.Q) * 4,
                                                      0);
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, g[i]), dst, vo,
                                                      (rc + (9 + i) * k.Q) * 4, 0);
            }
        }
        us = uc;
        uc = un;
#pragma unroll
        for (int i = 0; i < 9; ++i) gu[i] = gv[i];
        // C's pull of the next row, r - 3d: B wrote row r - 2d above
        lds_pull(fring2, r - 3 * d, lane, fr3 & kLinks, fc_n);
    };
    for (int j = 0; j < nj; j += 2) {
        step(std::integral_constant<int, 0>(), j);
        if (j + 1 < nj) step(std::integral_constant<int, 1>(), j + 1);
    }
}

// The launch of the three-iteration march, one wave per (chunk, strip).
// Interior strips (all 64 lanes are columns 1 ... W-2: no Zou-He column,
// ghost column or periodic wrap) run a copy of the march without the edge
// code; strip 0 and the last one or two run the full copy.  Measured on
// 4096 x 2048 at 1 x 26: 0.149 ms per iteration against 0.156 with the full
// copy everywhere.  Both copies share the workgroup's one LDS ring array.
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void fused3_kernel(KParams k, KBufs b) {
    __shared__ float rings[WPB * 2 * kRingRegion];
    const int L = banded_block();
    const int chunk = L / k.nsb;
    const int strip = (L - chunk * k.nsb) * WPB + (int)(threadIdx.x >> 6);
    const int x0 = strip * kFuse3Cols;
    if (x0 >= 6 && x0 + 60 <= k.W)
        fused3_march<WPB, false>(k, b, chunk, strip, rings);
    else
        fused3_march<WPB, true>(k, b, chunk, strip, rings);
}

// One domain of a batched launch (an ensemble): its constants, flags,
// inlet profile, both population and side-buffer sets, the parity of its
// handle when the table was built, and its first logical workgroup.
struct EnsDomain {
    KParams k;
    const uint16_t* flags;
    const float* uprof;
    float* pop[2];
    float* side[2];
    int cur0, sc0;
    int blk0;
    int pad;
};

// Iteration of every domain of an ensemble in one launch.  Logical
// workgroups (XCD-banded as in step_kernel) are laid out domain after
// domain; a workgroup finds its domain by binary search over blk0 (scalar
// loads: the index is uniform) and marches it with that domain's constants
// and the buffer set selected by `parity` (launches alternate 0, 1, ...).
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void ens_kernel(const EnsDomain* __restrict__ doms, int ndom, int parity) {
    const int L = banded_block();
    int lo = 0, hi = ndom - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (doms[mid].blk0 <= L) lo = mid; else hi = mid - 1;
    }
    const EnsDomain& e = doms[lo];
    const int pc = e.cur0 ^ parity, ps = e.sc0 ^ parity;
    KBufs b{};
    b.src = e.pop[pc];
    b.dst = e.pop[pc ^ 1];
    b.flags = e.flags;
    b.side_in = e.side[ps];
    b.side_out = e.side[ps ^ 1];
    b.uprof = e.uprof;
    march_block<WPB, 0>(e.k, b, L - e.blk0);
}

// Two (DEPTH 2) or three (DEPTH 3) iterations of every domain of an
// ensemble in one launch: as ens_kernel, over a table laid out for that
// march.
template <int WPB, int DEPTH>
__global__ __launch_bounds__(64 * WPB) void ens_fused_kernel(const EnsDomain* __restrict__ doms, int ndom,
                                                             int parity) {
    const int L = banded_block();
    int lo = 0, hi = ndom - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (doms[mid].blk0 <= L) lo = mid; else hi = mid - 1;
    }
    const EnsDomain& e = doms[lo];
    const int pc = e.cur0 ^ parity, ps = e.sc0 ^ parity;
    KBufs b{};
    b.src = e.pop[pc];
    b.dst = e.pop[pc ^ 1];
    b.flags = e.flags;
    b.side_in = e.side[ps];
    b.side_out = e.side[ps ^ 1];
    b.uprof = e.uprof;
    if (DEPTH == 3) {
        const int Ld = L - e.blk0, chunk = Ld / e.k.nsb;
        __shared__ float rings[WPB * 2 * kRingRegion];
        const int strip = (Ld - chunk * e.k.nsb) * WPB + (int)(threadIdx.x >> 6);
        const int x0 = strip * kFuse3Cols;
        if (x0 >= 6 && x0 + 60 <= e.k.W)
            fused3_march<WPB, false>(e.k, b, chunk, strip, rings);
        else
            fused3_march<WPB, true>(e.k, b, chunk, strip, rings);
    }
    else
        fused_march<WPB>(e.k, b, L - e.blk0);
}

// Zeros on every solid node of an inert-solid handle's population set:
// the state the step kernels keep on them (restored blobs are made to
// agree).
__global__ __launch_bounds__(256) void zero_inert_solids(float* __restrict__ pop, const uint16_t* __restrict__ flags,
                                                         int H, int W, int Wp, int64_t Q) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H) return;
    const int64_t o = (int64_t)y * Wp + x;
    if (!(flags[o] & kSolid)) return;
#pragma unroll
    for (int i = 0; i < 18; ++i) pop[i * Q + o] = 0.f;
}

// Initial state (fingering_periodic.py:90-121, or fingering.py:95-127 with
// the walled init), one thread per node.
//   PRE = 1: pre-collision populations f0/g0 into b.dst and the N = 0
//            fields + partial sums.
//   PRE = 0: post-collision populations of iteration 0 into b.dst and the
//            side buffer P of the initial psi into b.side_out.
//   ux0/uy0 (may be NULL): restart from given fields - the velocity is
//            taken as given and mu from psi, so a running flow continues
//            (fd_init_fields); with PRE = 1 the N = 0 state of such a start.
template <int PRE>
__global__ __launch_bounds__(256) void init_kernel(KParams k, KBufs b, const float* __restrict__ psi0,
                                                   const float* __restrict__ rho0, int walled_init,
                                                   const float* __restrict__ ux0 = nullptr,
                                                   const float* __restrict__ uy0 = nullptr) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    Red4 red = red4_init();
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
            if (ux0) {
                m.mu = fdn::chem_pot(psi, lap, k.c);
                m.fx = m.mu * dpx + k.c.gx;
                m.fy = m.mu * dpy + k.c.gy;
                m.ux = ux0[o];
                m.uy = uy0[o];
                m.p = m.rho * (1.f / 3.f) + psi * m.mu;
            }
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
            if (!solid) red4_add(red, m.rho, psi, fdn::tau_excess(m.rho, psi, k.c));
        } else if (!k.x_periodic && (x == 0 || x == k.W - 1)) {
            const float mu = fdn::chem_pot(psi, lap, k.c);
            const int s = (x == 0) ? 0 : 2;
            b.side_out[s * k.H + y] = dpx * mu;
            b.side_out[(s + 1) * k.H + y] = dpy * mu;
        }
    }
    if (PRE) block_red4<256>(red, b.partials + 4 * (blockIdx.y * gridDim.x + blockIdx.x));
}

// State from given post-stream populations of iteration N (what
// fd_get_populations returns: after bounce-back, walls and Zou-He), one
// thread per node.  `staged` holds them as 18 planes [H][Wp].
//   COLLIDE = 0: fields (psi, rho, ux, uy) and FP64 partial sums of
//                iteration N into b.fields / b.partials.
//   COLLIDE = 1: the collision of iteration N into b.dst (the post-collision
//                state the next launch pulls from) and P = d psi * mu of
//                columns 0 and W-1 into b.side_out; on inert solid nodes the
//                staged populations are cleared (nothing reads them).
// psi of a neighbour is the sum of its staged g - the value the step kernel
// forms after its own pull - so the collision equals the step kernel's bit
// for bit.
template <int COLLIDE>
__global__ __launch_bounds__(256) void restage_kernel(KParams k, KBufs b, float* __restrict__ staged) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    Red4 red = red4_init();
    if (x < k.W && y < k.H) {
        auto P = [&](int yy, int xx) -> float {
            if (k.y_walls && (yy < 0 || yy >= k.H)) return k.c.psi_wall;
            yy = wrap(yy, k.H);
            if (xx < 0 || xx >= k.W) {
                if (!k.x_periodic) return xx < 0 ? 1.0f : -1.0f;
                xx = wrap(xx, k.W);
            }
            const int o = yy * k.Wp + xx;
            if (b.flags[o] & kSolid) return k.c.psi_wall;
            const float* gp = staged + 9 * (size_t)k.Q + o;
            return gp[0] + gp[k.Q] + gp[2 * k.Q] + gp[3 * k.Q] + gp[4 * k.Q] + gp[5 * k.Q] + gp[6 * k.Q] +
                   gp[7 * k.Q] + gp[8 * k.Q];
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
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            f[i] = staged[i * (size_t)k.Q + o];
            g[i] = staged[(9 + i) * (size_t)k.Q + o];
        }
        if (COLLIDE) {
            if (solid) {
                // never collided: store-and-forward (or, inert, nothing)
                if (k.inert) {
#pragma unroll
                    for (int i = 0; i < 9; ++i) {
                        f[i] = g[i] = 0.f;
                        staged[i * (size_t)k.Q + o] = 0.f;
                        staged[(9 + i) * (size_t)k.Q + o] = 0.f;
                    }
                }
            } else {
                const fdn::Mom<float> m = fdn::moments(f, psi, dpx, dpy, lap, k.c);
                fdn::collide(f, g, m, k.c);
                if (!k.x_periodic && (x == 0 || x == k.W - 1)) {
                    const int s = (x == 0) ? 0 : 2;
                    b.side_out[s * k.H + y] = dpx * m.mu;
                    b.side_out[(s + 1) * k.H + y] = dpy * m.mu;
                }
            }
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                b.dst[i * (size_t)k.Q + o] = f[i];
                b.dst[(9 + i) * (size_t)k.Q + o] = g[i];
            }
        } else {
            float rho = 0.f, ux = 0.f, uy = 0.f;
            if (!solid) {
                const fdn::Mom<float> m = fdn::moments(f, psi, dpx, dpy, lap, k.c);
                rho = m.rho;
                ux = m.ux;
                uy = m.uy;
                red4_add(red, rho, psi, fdn::tau_excess(rho, psi, k.c));
            }
            b.fields[o] = psi;
            b.fields[k.Q + o] = rho;
            b.fields[2 * k.Q + o] = ux;
            b.fields[3 * k.Q + o] = uy;
        }
    }
    if (!COLLIDE) block_red4<256>(red, b.partials + 4 * (blockIdx.y * gridDim.x + blockIdx.x));
}

// Streaming copy, the HBM yardstick: grid-stride 16-byte non-temporal
// loads and stores (the fastest plain copy measured on MI355X).
__global__ __launch_bounds__(256) void fd_copy_kernel(const float4* __restrict__ a, float4* __restrict__ b,
                                                      int64_t n) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(a) + i);
        __builtin_nontemporal_store(v, reinterpret_cast<v4f*>(b) + i);
    }
}

__global__ __launch_bounds__(256) void reduce_partials(const double* __restrict__ part, int n,
                                                       double* __restrict__ out) {
    Red4 r = red4_init();
    for (int i = threadIdx.x; i < n; i += 256) {
        r.s0 += part[4 * i];
        r.s1 += part[4 * i + 1];
        r.s2 += part[4 * i + 2];
        r.m3 = fmin(r.m3, part[4 * i + 3]);
    }
    block_red4<256>(r, out);
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

constexpr int kStageInit = 1;
constexpr int kStagePops = 2;

struct fd_handle {
    fd_params p{};
    KParams k{};
    int device = 0;
    int Wp = 0;
    size_t Q = 0;
    float* pop[2] = {nullptr, nullptr};
    uint16_t* flags = nullptr;
    bool inert = false;
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
    hipEvent_t evs = nullptr;  // orders the handle's stream against an ensemble's
    int cur = 0, sc = 0;
    int64_t N = 0;
    bool inited = false;
    // What the buffers hold:
    //   0          pop[cur] = post-collision state of iteration N-1 (N >= 1)
    //   kStageInit fd_init / fd_init_fields: pop[cur] = post-collision state
    //              of iteration N (= 0); the fields of N come from the staged
    //              psi0/rho0 (and velocity)
    //   kStagePops fd_set_populations: pop[cur] = post-collision state of
    //              iteration N, pop[1-cur] = the given post-stream
    //              populations of N, which the reads of N return
    int stage = 0;
    bool init_u = false;  // started by fd_init_fields*: the velocity is staged in fields planes 2, 3
    int wx = 4;     // waves (64-column strips) per workgroup
    int rows = 16;  // rows per wave
    int fuse = 2;   // iterations per launch of fd_step (1, 2: fused_kernel, 3: fused3_kernel)
    bool use_graph = false;            // fd_step replays captured launch pairs
    hipGraphExec_t gexec = nullptr;    // the captured pair
    int gkey[5] = {0, 0, 0, 0, 0};     // its launch configuration
    int fwx = 1;    // fused_kernel: waves per workgroup
    int frows = 12; // fused_kernel: output rows per wave
    int f3wx = 1;   // fused3_kernel: waves per workgroup
    int f3rows = kFuse3Rows; // fused3_kernel: output rows per wave
    int timed_launches = 0;
    bool have_timing = false;
    int ens_refs = 0;  // ensembles holding this handle: fd_destroy refuses while > 0
};

namespace {

int check_params(const fd_params* p) {
    if (!p) return fail(FD_E_ARG, "params is NULL");
    if (p->H < 3 || p->W < 3) return fail(FD_E_ARG, "H and W must be >= 3");
    // buffer descriptors address the 18-plane set with byte offsets below 2^31
    // (the out-of-range offset used for masked lanes sits above that)
    if ((int64_t)p->H * (((int64_t)p->W + 63) / 64 * 64) * 18 * 4 >= ((int64_t)1 << 31) - 4096)
        return fail(FD_E_ARG, "domain too large: H * (W rounded up to 64) must stay below 29.8M nodes");
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

// Direction i of node (y, x) is rewritten after the pull by a boundary
// condition (Zou-He columns, walls, walled inlet/outlet corners), so the
// pulled value never matters there.
bool overwritten(const fd_params& p, int y, int x, int i) {
    if (p.x_bc == FD_X_ZOUHE) {
        if (x == 0 && (i == 1 || i == 5 || i == 8)) return true;
        if (x == p.W - 1 && (i == 3 || i == 6 || i == 7)) return true;
        if (p.y_bc == FD_Y_WALLS) {
            if (x == 0 && y == 0 && (i == 2 || i == 6)) return true;
            if (x == 0 && y == p.H - 1 && (i == 4 || i == 7)) return true;
            if (x == p.W - 1 && y == 0 && i == 2) return true;
            if (x == p.W - 1 && y == p.H - 1 && i == 4) return true;
        }
    }
    if (p.y_bc == FD_Y_WALLS) {
        if (y == p.wall_lo && (i == 2 || i == 5 || i == 6)) return true;
        if (y == p.wall_hi && (i == 4 || i == 7 || i == 8)) return true;
    }
    return false;
}

// Pack the node flags and decide whether solid nodes are inert: that needs
// every solid -> fluid link either bounced or rewritten by a boundary
// condition (no leak of a solid node's streamed populations) and no solid
// node on a Zou-He column.  The link bits of a node are its bounce-back
// links (bounce_back.py:25-167, over-bounced links included) plus, on the
// wall rows, the three directions the wall reflects (fingering.py:432-451):
// both pull the node's own opposite population.
constexpr unsigned kWallLoLinks = (1u << 1) | (1u << 4) | (1u << 5);  // directions 2, 5, 6
constexpr unsigned kWallHiLinks = (1u << 3) | (1u << 6) | (1u << 7);  // directions 4, 7, 8

bool pack_flags(const fd_params& p, const uint8_t* solid, const uint8_t* links, int Wp, std::vector<uint16_t>& fl) {
    const int H = p.H, W = p.W;
    auto sol = [&](int y, int x) { return solid[(size_t)((y + H) % H) * W + ((x + W) % W)] != 0; };
    bool inert = true;
    for (int y = 0; y < H && inert; ++y) {
        if (p.x_bc == FD_X_ZOUHE && (sol(y, 0) || sol(y, W - 1))) inert = false;
        for (int x = 0; x < W && inert; ++x) {
            if (sol(y, x)) continue;
            const unsigned lk = links[(size_t)y * W + x];
            for (int i = 1; i < 9; ++i) {
                const bool up = sol(y - fdn::EY[i], x - fdn::EX[i]);
                if (up && !((lk >> (i - 1)) & 1u) && !overwritten(p, y, x, i)) {
                    inert = false;
                    break;
                }
            }
        }
    }
    fl.assign((size_t)H * Wp, kSolid);  // padding columns: solid, never read
    const bool walls = p.y_bc == FD_Y_WALLS;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            unsigned v = links[(size_t)y * W + x] & FD_LINK_BITS;
            if (walls && y == p.wall_lo) v |= kWallLoLinks;
            if (walls && y == p.wall_hi) v |= kWallHiLinks;
            // an inert solid node is never loaded; all its link bits set make
            // a register pull of its populations read its own (zero) ones
            if (sol(y, x)) v = inert ? (kSolid | kLinks) : (v | kSolid);
            fl[(size_t)y * Wp + x] = (uint16_t)v;
        }
    if (inert)
        for (int y = 0; y < H; ++y)
            for (int x0 = 0; x0 < Wp; x0 += 16) {
                bool dead = true;
                for (int x = x0; x < x0 + 16 && dead; ++x) dead = (fl[(size_t)y * Wp + x] & kSolid) != 0;
                if (dead)
                    for (int x = x0; x < x0 + 16; ++x) fl[(size_t)y * Wp + x] |= (uint16_t)kDeadPiece;
            }
    return inert;
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
    k.inert = h->inert ? 1 : 0;
    k.rho_out = (float)p.rho_out;
    k.c3 = (float)p.outlet_f3_coef;
    k.c.inv_tau_g = (float)(1.0 / p.tau);
    k.c.eta_n = (float)p.eta_n;
    k.c.M = (float)p.M;
    k.c.a = (float)p.a;
    k.c.kappa = (float)p.kappa;
    k.c.gamma = (float)p.gamma;
    k.c.psi_wall = (float)p.psi_wall;
    k.c.psi_visc = p.psi_visc_bound > 0 ? (float)p.psi_visc_bound : 3.0e38f;
    k.c.gx = (float)p.gx;
    k.c.gy = (float)p.gy;
}

// Workgroups of h->wx waves (strips of 64 columns each), h->rows rows;
// workgroup columns [sb_lo, sb_hi) of the domain (sb_hi < 0: all).
int all_sb(const fd_handle* h) { return ((h->p.W + 63) / 64 + h->wx - 1) / h->wx; }

int step_grid(const fd_handle* h, KParams* kp, int sb_lo = 0, int sb_hi = -1) {
    if (sb_hi < 0) sb_hi = all_sb(h);
    const int nsb = sb_hi - sb_lo;
    const int nchunks = (h->p.H + h->rows - 1) / h->rows;
    if (kp) {
        kp->nsb = nsb;
        kp->sb0 = sb_lo;
    }
    return nsb * nchunks;
}

int launch_step(fd_handle* h, int mode, const KBufs& b, int sb_lo = 0, int sb_hi = -1,
                hipStream_t stream = nullptr) {
    KParams k = h->k;
    k.rows = h->rows;
    const dim3 grid(step_grid(h, &k, sb_lo, sb_hi));
    const dim3 block(64 * h->wx);
    hipStream_t st = stream ? stream : h->stream;
#define FD_LAUNCH(WPB)                                                                       \
    do {                                                                                     \
        if (mode == 0)                                                                       \
            hipLaunchKernelGGL((step_kernel<WPB, 0>), grid, block, 0, st, k, b);             \
        else                                                                                 \
            hipLaunchKernelGGL((step_kernel<WPB, 1>), grid, block, 0, st, k, b);             \
    } while (0)
    switch (h->wx) {
        case 1: FD_LAUNCH(1); break;
        case 2: FD_LAUNCH(2); break;
        case 8: FD_LAUNCH(8); break;
        default: FD_LAUNCH(4); break;
    }
#undef FD_LAUNCH
    FD_HIP(hipGetLastError());
    return FD_OK;
}

// The iteration whose post-collision state pop[cur] holds.
int64_t launched_through(const fd_handle* h) { return h->stage ? h->N : h->N - 1; }

// Launches that take the handle from iteration N to N + n.
int64_t launches_for(const fd_handle* h, int64_t n) {
    return n > 0 ? h->N + n - 1 - launched_through(h) : 0;
}

// Two iterations in one launch (fused_kernel): b.src holds S_n, b.dst
// receives S_n+2; side_in P_n, side_out P_n+2.  Stage A treats rows beyond
// a wall as ghost rows, which holds when the walls are the edge rows (every
// population that would stream in across y = 0 / H-1 is bounced there).
// With the walls of fingering.py:573 on rows 1 and H-2, rows 0 and H-1 pull
// across the edge through np.roll's wrap: those domains run one launch per
// iteration.
// The domain's geometry allows the two-iteration kernel: H >= 8, and y
// periodic or walled on the edge rows.
bool geom_fusable(const fd_params& p) {
    const bool edge_walls = p.y_bc != FD_Y_WALLS || (p.wall_lo == 0 && p.wall_hi == p.H - 1);
    return p.H >= 8 && edge_walls;
}

bool fusable(const fd_handle* h) { return h->fuse >= 2 && geom_fusable(h->p); }

// Workgroups across x and chunks of the two-iteration kernel on a domain of
// width W and height H with launch shape wx x rows.
int fused_strip_groups(int W, int wx) { return ((W + kFuseCols - 1) / kFuseCols + wx - 1) / wx; }

int launch_fused(fd_handle* h, const KBufs& b, hipStream_t stream = nullptr, int wx = 0, int rows = 0) {
    if (wx <= 0) wx = h->fwx;
    if (rows <= 0) rows = h->frows;
    KParams k = h->k;
    k.rows = rows;
    k.nsb = fused_strip_groups(h->p.W, wx);
    k.sb0 = 0;
    const int nchunks = (h->p.H + rows - 1) / rows;
    const dim3 grid(k.nsb * nchunks), block(64 * wx);
    if (!stream) stream = h->stream;
    switch (wx) {
        case 1: hipLaunchKernelGGL((fused_kernel<1>), grid, block, 0, stream, k, b); break;
        case 4: hipLaunchKernelGGL((fused_kernel<4>), grid, block, 0, stream, k, b); break;
        default: hipLaunchKernelGGL((fused_kernel<2>), grid, block, 0, stream, k, b); break;
    }
    FD_HIP(hipGetLastError());
    return FD_OK;
}

// Waves of the three-iteration kernel the device holds at once (one-wave
// workgroups): the quantum of its launch rounds.
int fused3_slots(int dev, const hipDeviceProp_t& prop, int* slots) {
    int per_cu = 0;
    FD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&fused3_kernel<1>), 64,
                                                        0));
    (void)dev;
    *slots = std::max(1, per_cu) * std::max(1, prop.multiProcessorCount);
    return FD_OK;
}

// One launch of the three-iteration kernel (one or two waves per
// workgroup).
int launch_fused3(fd_handle* h, const KBufs& b, hipStream_t stream = nullptr) {
    const int wx = h->f3wx >= 2 ? 2 : 1;
    KParams k = h->k;
    k.rows = h->f3rows;
    const int nstrips = (h->p.W + kFuse3Cols - 1) / kFuse3Cols;
    k.nsb = (nstrips + wx - 1) / wx;
    k.sb0 = 0;
    const int nchunks = (h->p.H + h->f3rows - 1) / h->f3rows;
    const dim3 grid(k.nsb * nchunks), block(64 * wx);
    if (!stream) stream = h->stream;
    if (wx == 2)
        hipLaunchKernelGGL((fused3_kernel<2>), grid, block, 0, stream, k, b);
    else
        hipLaunchKernelGGL((fused3_kernel<1>), grid, block, 0, stream, k, b);
    FD_HIP(hipGetLastError());
    return FD_OK;
}

// One launch of `it` iterations (1, 2 or 3) from the buffers b.
int launch_iterations(fd_handle* h, int it, const KBufs& b) {
    return it == 3 ? launch_fused3(h, b) : (it == 2 ? launch_fused(h, b) : launch_step(h, 0, b));
}

// Buffers of one launch from population parity c and side parity sc.
KBufs step_bufs(const fd_handle* h, int c, int sc) {
    KBufs b{};
    b.src = h->pop[c];
    b.dst = h->pop[1 - c];
    b.flags = h->flags;
    b.side_in = h->side[sc];
    b.side_out = h->side[1 - sc];
    b.uprof = h->uprof;
    return b;
}

// The captured pair of launches of `per` iterations each, from parity 0:
// (re)captured when the launch configuration changed since the last one.
int step_graph(fd_handle* h, int per) {
    const int key[5] = {per, per == 3 ? h->f3wx : h->fwx, per == 3 ? h->f3rows : h->frows, h->wx, h->rows};
    if (h->gexec && std::memcmp(key, h->gkey, sizeof key) == 0) return FD_OK;
    if (h->gexec) {
        (void)hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
    }
    hipGraph_t g = nullptr;
    FD_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = FD_OK;
    for (int p = 0; p < 2 && !rc; ++p) rc = launch_iterations(h, per, step_bufs(h, p, p));
    const hipError_t ec = hipStreamEndCapture(h->stream, &g);
    if (rc) {
        if (ec == hipSuccess && g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return fail(FD_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        h->gexec = nullptr;
        return fail(FD_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    std::memcpy(h->gkey, key, sizeof key);
    return FD_OK;
}

// Number of FP64 partial-sum slots (one per wave) a moments launch writes.
int moments_grid_blocks(const fd_handle* h) { return h->wx * step_grid(h, nullptr); }

int ensure_partials(fd_handle* h, int n) {
    if (n <= h->npart) return FD_OK;
    if (h->partials) FD_HIP(hipFree(h->partials));
    h->partials = nullptr;
    h->npart = 0;  // a failed allocation below must not leave a stale capacity behind
    FD_HIP(hipMalloc(&h->partials, sizeof(double) * 4 * (size_t)n));
    h->npart = n;
    return FD_OK;
}

// Fields (and optionally post-stream populations into `pops`) of the current
// iteration, plus the FP64 sums in h->mass.  With populations staged by
// fd_set_populations the post-stream populations already sit in
// pop[1 - cur]: `pops` may then only be that buffer (or NULL).
int compute_fields(fd_handle* h, float* pops, double* scratch = nullptr, int scratch_slots = 0) {
    if (!h->inited) return fail(FD_E_STATE, "fd_init has not been called");
    KBufs b{};
    b.flags = h->flags;
    b.uprof = h->uprof;
    b.fields = h->fields;
    b.dst = pops;
    int nblk;
    // FP64 partial sums: the handle's buffer (grown on demand), or a
    // caller-provided fixed scratch area
    double* part = nullptr;
    auto take_partials = [&](int n) -> int {
        if (scratch) {
            if (n > scratch_slots)
                return fail(FD_E_ARG, "this launch shape needs " + std::to_string(n) +
                                          " partial-sum slots; the fixed scratch holds " +
                                          std::to_string(scratch_slots));
            part = scratch;
            return FD_OK;
        }
        int rc = ensure_partials(h, n);
        part = h->partials;
        return rc;
    };
    const dim3 ngrid((h->p.W + 63) / 64, (h->p.H + 3) / 4);
    if (h->stage == kStageInit) {
        nblk = ngrid.x * ngrid.y;
        int rc = take_partials(nblk);
        if (rc) return rc;
        b.partials = part;
        if (!pops) b.dst = h->pop[1 - h->cur];
        // after fd_init_fields the staged velocity (fields planes 2, 3) is
        // part of the N = 0 state; each node reads its own value before
        // writing the same one back
        const float* ux0 = h->init_u ? h->fields + 2 * h->Q : nullptr;
        const float* uy0 = h->init_u ? h->fields + 3 * h->Q : nullptr;
        hipLaunchKernelGGL((init_kernel<1>), ngrid, dim3(64, 4), 0, h->stream, h->k, b, h->psi0, h->rho0,
                           h->p.init_mode == FD_INIT_WALLED ? 1 : 0, ux0, uy0);
        FD_HIP(hipGetLastError());
    } else if (h->stage == kStagePops) {
        if (pops && pops != h->pop[1 - h->cur]) return fail(FD_E_STATE, "internal: staged populations elsewhere");
        nblk = ngrid.x * ngrid.y;
        int rc = take_partials(nblk);
        if (rc) return rc;
        b.partials = part;
        hipLaunchKernelGGL((restage_kernel<0>), ngrid, dim3(64, 4), 0, h->stream, h->k, b, h->pop[1 - h->cur]);
        FD_HIP(hipGetLastError());
    } else {
        nblk = moments_grid_blocks(h);
        int rc = take_partials(nblk);
        if (rc) return rc;
        b.partials = part;
        b.src = h->pop[h->cur];
        b.side_in = h->side[h->sc];
        rc = launch_step(h, 1, b);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(256), 0, h->stream, part, nblk, h->mass);
    FD_HIP(hipGetLastError());
    return FD_OK;
}

// Fields of columns [x0, x1) only (the moments pass over the strips that
// hold them; no FP64 sums).  At N = 0 the full initial fields are used.
int compute_fields_window(fd_handle* h, int x0, int x1) {
    if (!h->inited) return fail(FD_E_STATE, "fd_init has not been called");
    if (h->stage) return compute_fields(h, nullptr);
    const int sb_lo = (x0 / 64) / h->wx, sb_hi = ((x1 - 1) / 64) / h->wx + 1;
    int rc = ensure_partials(h, h->wx * step_grid(h, nullptr, sb_lo, sb_hi));
    if (rc) return rc;
    KBufs b{};
    b.flags = h->flags;
    b.uprof = h->uprof;
    b.fields = h->fields;
    b.partials = h->partials;
    b.src = h->pop[h->cur];
    b.side_in = h->side[h->sc];
    return launch_step(h, 1, b, sb_lo, sb_hi);
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
    if (h->ens_refs > 0) {
        // an ensemble's device table still points at this handle's buffers
        g_err = "fd_destroy: the handle belongs to an ensemble; destroy the ensemble first (handle kept)";
        return;
    }
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
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->evs) (void)hipEventDestroy(h->evs);
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
    std::vector<uint16_t> fl;
    h->inert = pack_flags(*p, solid, links, h->Wp, fl);
    fill_kparams(h);
    // Launch shapes: rows per wave chosen so that a launch has enough waves
    // to fill the device (about 8192 for one iteration per launch, 12288 for
    // two; measured on MI355X: 4 x 16 and 1 x 8 on 4096 x 2048, 1 row per
    // wave on 400 x 100).  fd_set_tile / fd_set_fusion override them.
    {
        const int64_t strips64 = (p->W + 63) / 64, strips58 = (p->W + kFuseCols - 1) / kFuseCols;
        h->wx = strips64 >= 16 ? 4 : 1;
        h->rows = (int)std::min<int64_t>(16, std::max<int64_t>(1, (strips64 * p->H + 4096) / 8192));
        h->fwx = 1;
        h->frows = (int)std::min<int64_t>(8, std::max<int64_t>(1, (strips58 * p->H + 6144) / 12288));
        int slots = 0;
        if ((rc = fused3_slots(dev, prop, &slots))) {
            delete h;
            return rc;
        }
        h->f3rows = kFuse3Rows;
        // three iterations per launch on domains that fill the device at
        // least twice over with waves of 16 rows (measured on 4096 x 2048:
        // 0.1455 ms per iteration at 1 x 16 against 0.175 with two)
        if (geom_fusable(*p) &&
            (int64_t)((p->W + kFuse3Cols - 1) / kFuse3Cols) * ((p->H + h->f3rows - 1) / h->f3rows) >= 2 * (int64_t)slots)
            h->fuse = 3;
    }

    const size_t Q = h->Q, H = p->H;
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
    FD_HIP_H(hipEventCreateWithFlags(&h->evs, hipEventDisableTiming));
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
    FD_HIP_H(hipMalloc(&h->mass, sizeof(double) * 4));
    FD_HIP_H(hipMemcpyAsync(h->flags, fl.data(), sizeof(uint16_t) * Q, hipMemcpyHostToDevice, h->stream));
    FD_HIP_H(hipMemcpyAsync(h->uprof, up.data(), sizeof(float) * H, hipMemcpyHostToDevice, h->stream));
    FD_HIP_H(hipStreamSynchronize(h->stream));
#undef FD_HIP_H
    *out = h;
    return FD_OK;
}

namespace {
int init_impl(fd_handle* h, const double* psi0, const double* rho0, const double* ux0, const double* uy0) {
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
    const float* dux = nullptr;
    const float* duy = nullptr;
    std::vector<float> vx, vy;
    if (ux0) {
        // the field buffer's velocity planes stage the given velocity
        vx.assign(h->Q, 0.f);
        vy.assign(h->Q, 0.f);
        for (size_t y = 0; y < H; ++y)
            for (size_t x = 0; x < W; ++x) {
                vx[y * Wp + x] = (float)ux0[y * W + x];
                vy[y * Wp + x] = (float)uy0[y * W + x];
            }
        FD_HIP(hipMemcpyAsync(h->fields + 2 * h->Q, vx.data(), sizeof(float) * h->Q, hipMemcpyHostToDevice, h->stream));
        FD_HIP(hipMemcpyAsync(h->fields + 3 * h->Q, vy.data(), sizeof(float) * h->Q, hipMemcpyHostToDevice, h->stream));
        dux = h->fields + 2 * h->Q;
        duy = h->fields + 3 * h->Q;
    }
    hipLaunchKernelGGL((init_kernel<0>), grid, dim3(64, 4), 0, h->stream, h->k, b, h->psi0, h->rho0,
                       h->p.init_mode == FD_INIT_WALLED ? 1 : 0, dux, duy);
    FD_HIP(hipGetLastError());
    FD_HIP(hipStreamSynchronize(h->stream));  // host staging vectors go out of scope
    h->cur = 0;
    h->sc = 0;
    h->N = 0;
    h->inited = true;
    h->stage = kStageInit;
    h->init_u = ux0 != nullptr;
    h->have_timing = false;
    return FD_OK;
}
}  // namespace

int fd_init(fd_handle* h, const double* psi0, const double* rho0) {
    if (!h || !psi0) return fail(FD_E_ARG, "handle/psi0 is NULL");
    return init_impl(h, psi0, rho0, nullptr, nullptr);
}

int fd_init_fields(fd_handle* h, const double* psi, const double* rho, const double* ux, const double* uy) {
    if (!h || !psi || !rho || !ux || !uy) return fail(FD_E_ARG, "handle or a field is NULL");
    return init_impl(h, psi, rho, ux, uy);
}

int fd_step(fd_handle* h, int32_t nsteps) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    if (nsteps < 0) return fail(FD_E_ARG, "nsteps < 0");
    if (!h->inited) return fail(FD_E_STATE, "fd_init has not been called");
    FD_HIP(hipSetDevice(h->device));
    // The buffers hold the post-collision state of iteration
    // launched_through(h); reaching iteration N' needs launches up to N'-1.
    const int64_t have = launched_through(h);
    const int64_t nl = launches_for(h, nsteps);
    int64_t i = 0;  // iterations issued
    auto failed = [&](int rc) {
        // the launches already issued stand: the buffers hold the
        // post-collision state of iteration have + i
        if (i > 0) {
            h->N = have + i + 1;
            h->stage = 0;
        }
        h->have_timing = false;
        return rc;
    };
    auto issue = [&](int it) {
        int rc = launch_iterations(h, it, step_bufs(h, h->cur, h->sc));
        if (rc) return rc;
        h->cur ^= 1;
        h->sc ^= 1;
        i += it;
        return FD_OK;
    };
    if (nl > 0) FD_HIP(hipEventRecord(h->ev0, h->stream));
    // groups of iterations go to the two- or three-iteration kernel; long
    // loops replay a captured pair of launches (buffer parity 0 -> 1 -> 0)
    const int per = fusable(h) ? h->fuse : 1;
    int rc = FD_OK;
    if (h->use_graph && h->cur == h->sc && nl >= 4 * per) {
        if (h->cur != 0 && (rc = issue(per))) return failed(rc);
        if ((rc = step_graph(h, per))) return failed(rc);
        for (; nl - i >= 2 * per; i += 2 * per) {
            const hipError_t e = hipGraphLaunch(h->gexec, h->stream);
            if (e != hipSuccess) return failed(fail(FD_E_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e)));
        }
    }
    while (i < nl)
        if ((rc = issue((int)std::min<int64_t>(per, nl - i)))) return failed(rc);
    if (nl > 0) {
        FD_HIP(hipEventRecord(h->ev1, h->stream));
        h->timed_launches = (int)nl;
        h->have_timing = true;
    }
    h->N += nsteps;
    if (nsteps > 0) h->stage = 0;
    return FD_OK;
}

int fd_use_graph(fd_handle* h, int32_t on) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    h->use_graph = on != 0;
    return FD_OK;
}

int64_t fd_steps_done(const fd_handle* h) { return h ? h->N : -1; }

int fd_solid_inert(const fd_handle* h) { return h ? (h->inert ? 1 : 0) : -1; }

int fd_copy_probe(int64_t nbytes, int32_t reps, double* ms_per_copy) {
    if (nbytes < ((int64_t)32 << 20) || reps < 1 || !ms_per_copy) return fail(FD_E_ARG, "copy probe needs >= 32 MiB");
    const int64_t n = nbytes / 16;
    // the best of three launch shapes (few long-lived waves to one lane per vector)
    const int grids[3] = {1024, 4096, (int)std::min<int64_t>((n + 255) / 256, 1 << 20)};
    float4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FD_OK;
    double best = -1.0;
#define FD_HIP_P(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            rc = fail(FD_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
            goto done;                                                                     \
        }                                                                                  \
    } while (0)
    FD_HIP_P(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    FD_HIP_P(hipEventCreate(&e0));
    FD_HIP_P(hipEventCreate(&e1));
    FD_HIP_P(hipMalloc(&a, n * 16));
    FD_HIP_P(hipMalloc(&b, n * 16));
    FD_HIP_P(hipMemsetAsync(a, 0, n * 16, st));
    for (int gi = 0; gi < 3; ++gi) {
        float ms = 0.f;
        hipLaunchKernelGGL(fd_copy_kernel, dim3(grids[gi]), dim3(256), 0, st, a, b, n);  // warm-up
        FD_HIP_P(hipGetLastError());
        FD_HIP_P(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(fd_copy_kernel, dim3(grids[gi]), dim3(256), 0, st, a, b, n);
        FD_HIP_P(hipGetLastError());
        FD_HIP_P(hipEventRecord(e1, st));
        FD_HIP_P(hipEventSynchronize(e1));
        FD_HIP_P(hipEventElapsedTime(&ms, e0, e1));
        if (best < 0.0 || (double)ms / reps < best) best = (double)ms / reps;
    }
    *ms_per_copy = best;
#undef FD_HIP_P
done:
    if (st) (void)hipStreamSynchronize(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

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

int fd_set_fusion(fd_handle* h, int32_t depth, int32_t waves_x, int32_t rows) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    if (depth < 1 || depth > 3) return fail(FD_E_ARG, "depth must be 1, 2 or 3");
    if (waves_x > 0 && waves_x != 1 && waves_x != 2 && waves_x != 4)
        return fail(FD_E_ARG, "waves_x must be 1, 2 or 4 (or <= 0 to keep)");
    h->fuse = depth;
    if (depth == 3) {
        if (waves_x > 0) h->f3wx = waves_x;
        if (rows > 0) h->f3rows = rows;
    } else {
        if (waves_x > 0) h->fwx = waves_x;
        if (rows > 0) h->frows = rows;
    }
    return FD_OK;
}

int fd_get_fusion(const fd_handle* h, int32_t* depth, int32_t* waves_x, int32_t* rows) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    const int d = fusable(h) ? h->fuse : 1;
    if (depth) *depth = d;
    if (waves_x) *waves_x = d == 3 ? h->f3wx : (d == 2 ? h->fwx : h->wx);
    if (rows) *rows = d == 3 ? h->f3rows : (d == 2 ? h->frows : h->rows);
    return FD_OK;
}

int fd_set_tile(fd_handle* h, int32_t waves_x, int32_t rows) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    if (waves_x != 1 && waves_x != 2 && waves_x != 4 && waves_x != 8)
        return fail(FD_E_ARG, "waves_x must be 1, 2, 4 or 8");
    if (rows < 1) return fail(FD_E_ARG, "rows must be >= 1");
    h->wx = waves_x;
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

int fd_get_fields_window(fd_handle* h, int32_t x0, int32_t x1, float* psi, float* rho, float* ux, float* uy) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    if (x0 < 0 || x1 > h->p.W || x0 >= x1) return fail(FD_E_ARG, "need 0 <= x0 < x1 <= W");
    FD_HIP(hipSetDevice(h->device));
    int rc = compute_fields_window(h, x0, x1);
    if (rc) return rc;
    const size_t w = (size_t)(x1 - x0);
    float* outs[4] = {psi, rho, ux, uy};
    for (int i = 0; i < 4; ++i)
        if (outs[i])
            FD_HIP(hipMemcpy2DAsync(outs[i], sizeof(float) * w, h->fields + i * h->Q + x0, sizeof(float) * h->Wp,
                                    sizeof(float) * w, h->p.H, hipMemcpyDeviceToHost, h->stream));
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

}  // extern "C"

// Fixed FP64 partial-sum scratch of fd_phase_mass_noalloc: kScratchSlots
// slots of 4 doubles (32 B), 4 MiB.  A device global: every GPU the library
// runs on holds its own copy, at its own address, until the process ends.
// In host code `scratch` names a host-side stand-in, never the device
// memory: kernels and copies get the address from scratch_of().
namespace {
constexpr int kScratchSlots = 131072;
}
static __device__ __align__(256) double scratch[4 * kScratchSlots];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of `scratch` on device dev (the current device),
// looked up once per device.
int scratch_of(int dev, double** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(FD_E_ARG, "device index out of range for the fixed scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(FD_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return FD_OK;
}
}  // namespace

extern "C" {

int fd_phase_mass_noalloc(fd_handle* h, double out[3]) {
    if (!h || !out) return fail(FD_E_ARG, "handle/out is NULL");
    FD_HIP(hipSetDevice(h->device));
    double* sp = nullptr;
    int rc = scratch_of(h->device, &sp);
    if (rc) return rc;
    rc = compute_fields(h, nullptr, sp, kScratchSlots);
    if (rc) return rc;
    FD_HIP(hipMemcpyAsync(out, h->mass, sizeof(double) * 3, hipMemcpyDeviceToHost, h->stream));
    FD_HIP(hipStreamSynchronize(h->stream));
    return FD_OK;
}

// State blob: header (magic, H, W, N) + side[4H] + 18 planes [H][W].
// kMagicState: the planes are the post-collision state of iteration N-1 and
// the side buffer belongs to it.  kMagicStaged: the planes are the
// post-stream populations of iteration N (a state set by
// fd_set_populations and not stepped since); the side buffer is unused and
// loading it stages the populations again.
constexpr int64_t kMagicState = 0x4644424d53544154LL;
constexpr int64_t kMagicStaged = 0x4644424d53544147LL;

int64_t fd_state_bytes(const fd_handle* h) {
    if (!h) return -1;
    return 32 + (int64_t)sizeof(float) * (4 * (int64_t)h->p.H + 18 * (int64_t)h->p.H * h->p.W);
}

int fd_save_state(fd_handle* h, void* buf, int64_t nbytes) {
    if (!h || !buf) return fail(FD_E_ARG, "handle/buf is NULL");
    if (!h->inited) return fail(FD_E_STATE, "fd_init has not been called");
    if (h->stage == kStageInit)
        return fail(FD_E_STATE, "nothing to save before the first iteration: fd_init rebuilds that state");
    if (nbytes < fd_state_bytes(h)) return fail(FD_E_ARG, "buffer too small");
    FD_HIP(hipSetDevice(h->device));
    const bool staged = h->stage == kStagePops;
    int64_t hdr[4] = {staged ? kMagicStaged : kMagicState, h->p.H, h->p.W, h->N};
    std::memcpy(buf, hdr, 32);
    char* p = (char*)buf + 32;
    FD_HIP(hipMemcpyAsync(p, h->side[h->sc], sizeof(float) * 4 * h->p.H, hipMemcpyDeviceToHost, h->stream));
    p += sizeof(float) * 4 * h->p.H;
    const float* planes = staged ? h->pop[1 - h->cur] : h->pop[h->cur];
    for (int i = 0; i < 18; ++i) {
        int rc = copy_plane_out(h, (float*)p, planes + i * h->Q);
        if (rc) return rc;
        p += sizeof(float) * (size_t)h->p.H * h->p.W;
    }
    FD_HIP(hipStreamSynchronize(h->stream));
    return FD_OK;
}

}  // extern "C"

namespace {
// Upload post-stream populations of iteration N ([9][H][W] f, then g; host
// or, with kind hipMemcpyDeviceToDevice, device memory) into pop[1 - cur] and
// run the collision of iteration N from them.
int stage_populations(fd_handle* h, const float* f, const float* g, int64_t N, hipMemcpyKind kind) {
    float* st = h->pop[1 - h->cur];
    const size_t plane = (size_t)h->p.H * h->p.W;
    for (int i = 0; i < 18; ++i) {
        const float* src = i < 9 ? f + i * plane : g + (i - 9) * plane;
        FD_HIP(hipMemcpy2DAsync(st + i * h->Q, sizeof(float) * h->Wp, src, sizeof(float) * h->p.W,
                                sizeof(float) * h->p.W, h->p.H, kind, h->stream));
    }
    KBufs b{};
    b.flags = h->flags;
    b.dst = h->pop[h->cur];
    b.side_out = h->side[h->sc];
    const dim3 grid((h->p.W + 63) / 64, (h->p.H + 3) / 4);
    hipLaunchKernelGGL((restage_kernel<1>), grid, dim3(64, 4), 0, h->stream, h->k, b, st);
    FD_HIP(hipGetLastError());
    h->N = N;
    h->inited = true;
    h->stage = kStagePops;
    h->init_u = false;
    h->have_timing = false;
    return FD_OK;
}
}  // namespace

extern "C" {

int fd_set_populations(fd_handle* h, const float* f, const float* g) {
    if (!h || !f || !g) return fail(FD_E_ARG, "handle/f/g is NULL");
    FD_HIP(hipSetDevice(h->device));
    int rc = stage_populations(h, f, g, h->inited ? h->N : 0, hipMemcpyHostToDevice);
    if (rc) return rc;
    FD_HIP(hipStreamSynchronize(h->stream));  // the caller's host arrays may go away
    return FD_OK;
}

int fd_watchdog(fd_handle* h, double out[2]) {
    if (!h || !out) return fail(FD_E_ARG, "handle/out is NULL");
    FD_HIP(hipSetDevice(h->device));
    int rc = compute_fields(h, nullptr);
    if (rc) return rc;
    FD_HIP(hipMemcpyAsync(out, h->mass + 2, sizeof(double) * 2, hipMemcpyDeviceToHost, h->stream));
    FD_HIP(hipStreamSynchronize(h->stream));
    return FD_OK;
}

int fd_load_state(fd_handle* h, const void* buf, int64_t nbytes) {
    if (!h || !buf) return fail(FD_E_ARG, "handle/buf is NULL");
    if (nbytes < fd_state_bytes(h)) return fail(FD_E_ARG, "buffer too small");
    int64_t hdr[4];
    std::memcpy(hdr, buf, 32);
    const bool staged = hdr[0] == kMagicStaged;
    if ((hdr[0] != kMagicState && !staged) || hdr[1] != h->p.H || hdr[2] != h->p.W || hdr[3] < (staged ? 0 : 1))
        return fail(FD_E_ARG, "state blob does not match this handle");
    FD_HIP(hipSetDevice(h->device));
    const char* p = (const char*)buf + 32;
    if (staged) {
        const float* planes = (const float*)(p + sizeof(float) * 4 * h->p.H);
        const size_t plane = (size_t)h->p.H * h->p.W;
        int rc = stage_populations(h, planes, planes + 9 * plane, hdr[3], hipMemcpyHostToDevice);
        if (rc) return rc;
        FD_HIP(hipStreamSynchronize(h->stream));
        return FD_OK;
    }
    FD_HIP(hipMemcpyAsync(h->side[h->sc], p, sizeof(float) * 4 * h->p.H, hipMemcpyHostToDevice, h->stream));
    p += sizeof(float) * 4 * h->p.H;
    for (int i = 0; i < 18; ++i) {
        FD_HIP(hipMemcpy2DAsync(h->pop[h->cur] + i * h->Q, sizeof(float) * h->Wp, p, sizeof(float) * h->p.W,
                                sizeof(float) * h->p.W, h->p.H, hipMemcpyHostToDevice, h->stream));
        p += sizeof(float) * (size_t)h->p.H * h->p.W;
    }
    if (h->inert) {
        hipLaunchKernelGGL(zero_inert_solids, dim3((h->p.W + 63) / 64, (h->p.H + 3) / 4), dim3(64, 4), 0, h->stream,
                           h->pop[h->cur], h->flags, h->p.H, h->p.W, h->Wp, (int64_t)h->Q);
        FD_HIP(hipGetLastError());
    }
    FD_HIP(hipStreamSynchronize(h->stream));
    h->N = hdr[3];
    h->inited = true;
    h->stage = 0;
    h->init_u = false;
    h->have_timing = false;
    return FD_OK;
}

// ----------------------------------------------- device-array coupling
// Everything below is enqueued on the handle's stream and returns without
// waiting; the caller orders its own work on that stream (fd_stream).

void* fd_stream(const fd_handle* h) { return h ? (void*)h->stream : nullptr; }

}  // extern "C"

namespace {
// 0 if every non-NULL pointer is device memory of the handle's device, else
// an FD_E_ARG failure: a host pointer in a device-to-device copy would
// fault the GPU instead of failing the call.
int check_device_ptrs(const fd_handle* h, std::initializer_list<const void*> ptrs) {
    for (const void* q : ptrs) {
        if (!q) continue;
        hipPointerAttribute_t a{};
        if (hipPointerGetAttributes(&a, q) != hipSuccess) {
            (void)hipGetLastError();
            return fail(FD_E_ARG, "not a device pointer");
        }
        if (a.type != hipMemoryTypeDevice || a.device != h->device)
            return fail(FD_E_ARG, "not device memory of the handle's device");
    }
    return FD_OK;
}
}  // namespace

extern "C" {

int fd_fields_device(fd_handle* h, float* psi, float* rho, float* ux, float* uy) {
    if (!h) return fail(FD_E_ARG, "handle is NULL");
    FD_HIP(hipSetDevice(h->device));
    int rc = check_device_ptrs(h, {psi, rho, ux, uy});
    if (rc) return rc;
    rc = compute_fields(h, nullptr);
    if (rc) return rc;
    float* outs[4] = {psi, rho, ux, uy};
    for (int i = 0; i < 4; ++i)
        if (outs[i])
            FD_HIP(hipMemcpy2DAsync(outs[i], sizeof(float) * h->p.W, h->fields + i * h->Q, sizeof(float) * h->Wp,
                                    sizeof(float) * h->p.W, h->p.H, hipMemcpyDeviceToDevice, h->stream));
    return FD_OK;
}

int fd_init_fields_device(fd_handle* h, const float* psi, const float* rho, const float* ux, const float* uy) {
    if (!h || !psi || !rho || !ux || !uy) return fail(FD_E_ARG, "handle or a field is NULL");
    if (!h->inited) return fail(FD_E_STATE, "fd_init has not been called");
    FD_HIP(hipSetDevice(h->device));
    int rc = check_device_ptrs(h, {psi, rho, ux, uy});
    if (rc) return rc;
    // Padding columns x >= W of the staging planes are never read by the
    // initialisation (it covers x < W only).
    const float* src[4] = {psi, rho, ux, uy};
    float* dst[4] = {h->psi0, h->rho0, h->fields + 2 * h->Q, h->fields + 3 * h->Q};
    for (int i = 0; i < 4; ++i)
        FD_HIP(hipMemcpy2DAsync(dst[i], sizeof(float) * h->Wp, src[i], sizeof(float) * h->p.W,
                                sizeof(float) * h->p.W, h->p.H, hipMemcpyDeviceToDevice, h->stream));
    KBufs b{};
    b.flags = h->flags;
    b.dst = h->pop[0];
    b.side_out = h->side[0];
    b.uprof = h->uprof;
    const dim3 grid((h->p.W + 63) / 64, (h->p.H + 3) / 4);
    hipLaunchKernelGGL((init_kernel<0>), grid, dim3(64, 4), 0, h->stream, h->k, b, h->psi0, h->rho0,
                       h->p.init_mode == FD_INIT_WALLED ? 1 : 0, dst[2], dst[3]);
    FD_HIP(hipGetLastError());
    h->cur = 0;
    h->sc = 0;
    h->N = 0;
    h->stage = kStageInit;
    h->init_u = true;
    h->have_timing = false;
    return FD_OK;
}

// Device state blob: side[4H] then the 18 padded population planes [H][Wp].
int64_t fd_state_device_bytes(const fd_handle* h) {
    if (!h) return -1;
    return (int64_t)sizeof(float) * (4 * (int64_t)h->p.H + 18 * (int64_t)h->Q);
}

int fd_save_state_device(fd_handle* h, void* dbuf, int64_t nbytes) {
    if (!h || !dbuf) return fail(FD_E_ARG, "handle/buf is NULL");
    if (!h->inited) return fail(FD_E_STATE, "fd_init has not been called");
    if (h->stage) return fail(FD_E_STATE, "a device state blob is taken after at least one iteration since the last "
                                          "fd_init / fd_set_populations");
    if (nbytes < fd_state_device_bytes(h)) return fail(FD_E_ARG, "buffer too small");
    FD_HIP(hipSetDevice(h->device));
    int rc = check_device_ptrs(h, {dbuf});
    if (rc) return rc;
    char* p = (char*)dbuf;
    FD_HIP(hipMemcpyAsync(p, h->side[h->sc], sizeof(float) * 4 * h->p.H, hipMemcpyDeviceToDevice, h->stream));
    FD_HIP(hipMemcpyAsync(p + sizeof(float) * 4 * h->p.H, h->pop[h->cur], sizeof(float) * 18 * h->Q,
                          hipMemcpyDeviceToDevice, h->stream));
    return FD_OK;
}

int fd_load_state_device(fd_handle* h, const void* dbuf, int64_t nbytes, int64_t steps_done) {
    if (!h || !dbuf) return fail(FD_E_ARG, "handle/buf is NULL");
    if (nbytes < fd_state_device_bytes(h)) return fail(FD_E_ARG, "buffer too small");
    if (steps_done < 1) return fail(FD_E_ARG, "a device state is restored at iteration >= 1");
    FD_HIP(hipSetDevice(h->device));
    int rc = check_device_ptrs(h, {dbuf});
    if (rc) return rc;
    const char* p = (const char*)dbuf;
    FD_HIP(hipMemcpyAsync(h->side[h->sc], p, sizeof(float) * 4 * h->p.H, hipMemcpyDeviceToDevice, h->stream));
    FD_HIP(hipMemcpyAsync(h->pop[h->cur], p + sizeof(float) * 4 * h->p.H, sizeof(float) * 18 * h->Q,
                          hipMemcpyDeviceToDevice, h->stream));
    if (h->inert) {
        hipLaunchKernelGGL(zero_inert_solids, dim3((h->p.W + 63) / 64, (h->p.H + 3) / 4), dim3(64, 4), 0, h->stream,
                           h->pop[h->cur], h->flags, h->p.H, h->p.W, h->Wp, (int64_t)h->Q);
        FD_HIP(hipGetLastError());
    }
    h->N = steps_done;
    h->inited = true;
    h->stage = 0;
    h->init_u = false;
    h->have_timing = false;
    return FD_OK;
}


// ------------------------------------------------------------------ ensemble
}  // extern "C"

struct fd_ensemble {
    std::vector<fd_handle*> hs;
    std::vector<int> cur0;           // handle parities when the table was built
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evd = nullptr;
    EnsDomain* table = nullptr;      // single-iteration launch
    EnsDomain* ftable = nullptr;     // two- or three-iteration launch (fdepth)
    int nblocks = 0, fnblocks = 0, fdepth = 0;
    int wx = 4, rows = 16;           // single-iteration launch shape
    int fuse = 3, fwx = 1, frows = 8;  // iterations per launch, two-iteration launch shape
    int f3wx = 1, f3rows = 26;       // three-iteration launch shape
    bool table_ok = false;
    int64_t flips = 0;               // launches since the table was built
    bool use_graph = true;
    hipGraphExec_t exec = nullptr, fexec = nullptr;
    int timed_launches = 0;
    bool have_timing = false;
    bool counted = false;            // holds a membership count on every handle
};

namespace {

// Iterations per batched launch: e->fuse when every domain's geometry
// allows the two- and three-iteration kernels, else 1.
int ens_depth(const fd_ensemble* e) {
    if (e->fuse < 2) return 1;
    for (const fd_handle* h : e->hs)
        if (!geom_fusable(h->p)) return 1;
    return e->fuse;
}

int ens_build(fd_ensemble* e) {
    for (hipGraphExec_t* x : {&e->exec, &e->fexec})
        if (*x) {
            (void)hipGraphExecDestroy(*x);
            *x = nullptr;
        }
    std::vector<EnsDomain> t(e->hs.size()), ft;
    int64_t blk = 0, fblk = 0;
    const int depth = ens_depth(e);
    const bool fused = depth > 1;
    if (fused) ft.resize(e->hs.size());
    for (size_t i = 0; i < e->hs.size(); ++i) {
        fd_handle* h = e->hs[i];
        EnsDomain& d = t[i];
        std::memset(&d, 0, sizeof d);
        d.k = h->k;
        d.k.rows = e->rows;
        d.k.nsb = ((h->p.W + 63) / 64 + e->wx - 1) / e->wx;
        d.k.sb0 = 0;
        d.flags = h->flags;
        d.uprof = h->uprof;
        d.pop[0] = h->pop[0];
        d.pop[1] = h->pop[1];
        d.side[0] = h->side[0];
        d.side[1] = h->side[1];
        d.cur0 = h->cur;
        d.sc0 = h->sc;
        d.blk0 = (int)blk;
        e->cur0[i] = h->cur;
        blk += (int64_t)d.k.nsb * ((h->p.H + e->rows - 1) / e->rows);
        if (fused) {
            EnsDomain& fd = ft[i];
            fd = d;
            const int frows = depth == 3 ? e->f3rows : e->frows;
            fd.k.rows = frows;
            fd.k.nsb = depth == 3 ? ((h->p.W + kFuse3Cols - 1) / kFuse3Cols + e->f3wx - 1) / e->f3wx
                                  : fused_strip_groups(h->p.W, e->fwx);
            fd.blk0 = (int)fblk;
            fblk += (int64_t)fd.k.nsb * ((h->p.H + frows - 1) / frows);
        }
    }
    if (blk >= ((int64_t)1 << 31) || fblk >= ((int64_t)1 << 31))
        return fail(FD_E_ARG, "ensemble too large for one launch");
    FD_HIP(hipMemcpyAsync(e->table, t.data(), sizeof(EnsDomain) * t.size(), hipMemcpyHostToDevice, e->stream));
    if (fused)
        FD_HIP(hipMemcpyAsync(e->ftable, ft.data(), sizeof(EnsDomain) * ft.size(), hipMemcpyHostToDevice,
                              e->stream));
    FD_HIP(hipStreamSynchronize(e->stream));
    e->nblocks = (int)blk;
    e->fnblocks = fused ? (int)fblk : 0;
    e->fdepth = depth;
    e->flips = 0;
    e->table_ok = true;
    return FD_OK;
}

// One batched launch of `it` iterations (1, or the table's depth).
int ens_launch(fd_ensemble* e, int parity, int it = 1) {
    const int n = (int)e->hs.size();
    if (it == 3) {
        const dim3 grid(e->fnblocks), block(64 * e->f3wx);
        if (e->f3wx == 2)
            hipLaunchKernelGGL((ens_fused_kernel<2, 3>), grid, block, 0, e->stream, e->ftable, n, parity);
        else
            hipLaunchKernelGGL((ens_fused_kernel<1, 3>), grid, block, 0, e->stream, e->ftable, n, parity);
        FD_HIP(hipGetLastError());
        return FD_OK;
    }
    if (it == 2) {
        const dim3 grid(e->fnblocks), block(64 * e->fwx);
        switch (e->fwx) {
            case 1: hipLaunchKernelGGL((ens_fused_kernel<1, 2>), grid, block, 0, e->stream, e->ftable, n, parity); break;
            case 4: hipLaunchKernelGGL((ens_fused_kernel<4, 2>), grid, block, 0, e->stream, e->ftable, n, parity); break;
            default: hipLaunchKernelGGL((ens_fused_kernel<2, 2>), grid, block, 0, e->stream, e->ftable, n, parity); break;
        }
        FD_HIP(hipGetLastError());
        return FD_OK;
    }
    const dim3 grid(e->nblocks), block(64 * e->wx);
    switch (e->wx) {
        case 1: hipLaunchKernelGGL((ens_kernel<1>), grid, block, 0, e->stream, e->table, n, parity); break;
        case 2: hipLaunchKernelGGL((ens_kernel<2>), grid, block, 0, e->stream, e->table, n, parity); break;
        case 8: hipLaunchKernelGGL((ens_kernel<8>), grid, block, 0, e->stream, e->table, n, parity); break;
        default: hipLaunchKernelGGL((ens_kernel<4>), grid, block, 0, e->stream, e->table, n, parity); break;
    }
    FD_HIP(hipGetLastError());
    return FD_OK;
}

// Two launches (parity 0, then 1) captured once and replayed: the buffer
// sets alternate, so a pair returns every pointer to where it started.
int ens_graph(fd_ensemble* e, int per) {
    hipGraphExec_t& exec = per > 1 ? e->fexec : e->exec;
    if (exec) return FD_OK;
    hipGraph_t g = nullptr;
    FD_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = ens_launch(e, 0, per);
    if (!rc) rc = ens_launch(e, 1, per);
    const hipError_t ec = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (ec == hipSuccess && g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return fail(FD_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        exec = nullptr;
        return fail(FD_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    return FD_OK;
}

int ens_step(fd_ensemble* e, int32_t nsteps, bool sequential) {
    if (!e) return fail(FD_E_ARG, "ensemble is NULL");
    if (nsteps < 0) return fail(FD_E_ARG, "nsteps < 0");
    const int64_t N = e->hs[0]->N;
    const int64_t have = launched_through(e->hs[0]);
    for (fd_handle* h : e->hs) {
        if (!h->inited) return fail(FD_E_STATE, "fd_init has not been called on every handle");
        if (h->N != N || launched_through(h) != have)
            return fail(FD_E_STATE, "the handles of an ensemble must be at the same iteration and stage");
    }
    FD_HIP(hipSetDevice(e->device));
    const int64_t nl = launches_for(e->hs[0], nsteps);
    if (nl > 0) {
        int rc = FD_OK;
        bool consistent = e->table_ok;
        for (size_t i = 0; consistent && i < e->hs.size(); ++i)
            consistent = e->hs[i]->cur == (e->cur0[i] ^ (int)(e->flips & 1));
        if (!consistent && (rc = ens_build(e))) return rc;
        // the handles' own queued work comes first
        for (fd_handle* h : e->hs) {
            FD_HIP(hipEventRecord(h->evs, h->stream));
            FD_HIP(hipStreamWaitEvent(e->stream, h->evs, 0));
        }
        FD_HIP(hipEventRecord(e->ev0, e->stream));
        // iterations per launch: groups go to the two- or three-iteration
        // kernel, a shorter remainder to the single-iteration one
        const int per = e->fdepth > 1 ? e->fdepth : 1;
        if (sequential) {
            // the baseline: one launch per domain per launch step, one stream
            for (int64_t i = 0; i < nl;) {
                const int it = nl - i >= per ? per : 1;
                for (fd_handle* h : e->hs) {
                    const KBufs b = step_bufs(h, h->cur, h->sc);
                    if (it == 3) {
                        const int wx3 = h->f3wx, rows3 = h->f3rows;
                        h->f3wx = e->f3wx;
                        h->f3rows = e->f3rows;
                        rc = launch_fused3(h, b, e->stream);
                        h->f3wx = wx3;
                        h->f3rows = rows3;
                    } else if (it == 2) {
                        rc = launch_fused(h, b, e->stream, e->fwx, e->frows);
                    } else {
                        const int wx = h->wx, rows = h->rows;
                        h->wx = e->wx;
                        h->rows = e->rows;
                        rc = launch_step(h, 0, b, 0, -1, e->stream);
                        h->wx = wx;
                        h->rows = rows;
                    }
                    if (rc) return rc;
                    h->cur ^= 1;
                    h->sc ^= 1;
                }
                i += it;
                ++e->flips;
            }
        } else {
            // every launch, of one iteration or more, flips the buffer parity
            int64_t i = 0, launches = 0;
            auto one = [&](int it) {
                const int r = ens_launch(e, (int)(e->flips & 1), it);
                if (!r) {
                    i += it;
                    ++e->flips;
                    ++launches;
                }
                return r;
            };
            if ((e->flips & 1) != 0 && (rc = one(nl >= per ? per : 1))) return rc;
            if (e->use_graph && nl - i >= 2 * per) {
                if ((rc = ens_graph(e, per))) return rc;
                for (; nl - i >= 2 * per; i += 2 * per, e->flips += 2, launches += 2)
                    FD_HIP(hipGraphLaunch(per > 1 ? e->fexec : e->exec, e->stream));
            }
            while (i < nl)
                if ((rc = one(nl - i >= per ? per : 1))) return rc;
            for (fd_handle* h : e->hs) {
                h->cur ^= (int)(launches & 1);
                h->sc ^= (int)(launches & 1);
            }
        }
        FD_HIP(hipEventRecord(e->ev1, e->stream));
        // later work on a handle's own stream waits for the ensemble
        FD_HIP(hipEventRecord(e->evd, e->stream));
        for (fd_handle* h : e->hs) FD_HIP(hipStreamWaitEvent(h->stream, e->evd, 0));
        e->timed_launches = (int)nl;
        e->have_timing = true;
    }
    for (fd_handle* h : e->hs) {
        h->N += nsteps;
        if (nsteps > 0) h->stage = 0;
        h->have_timing = false;
    }
    return FD_OK;
}

}  // namespace

extern "C" {

int fd_ens_create(fd_handle* const* hs, int32_t n, fd_ensemble** out) {
    if (!out) return fail(FD_E_ARG, "out is NULL");
    *out = nullptr;
    if (!hs || n < 1) return fail(FD_E_ARG, "need at least one handle");
    std::vector<fd_handle*> v(hs, hs + n);
    for (int i = 0; i < n; ++i) {
        if (!v[i]) return fail(FD_E_ARG, "handle is NULL");
        if (v[i]->device != v[0]->device) return fail(FD_E_ARG, "the handles of an ensemble must share a device");
        for (int j = 0; j < i; ++j)
            if (v[j] == v[i]) return fail(FD_E_ARG, "a handle appears twice in the ensemble");
    }
    fd_ensemble* e = new (std::nothrow) fd_ensemble();
    if (!e) return fail(FD_E_ARG, "out of host memory");
    e->hs = v;
    e->cur0.assign(n, 0);
    e->device = v[0]->device;
#define FD_HIP_E(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            fd_ens_destroy(e);                                                             \
            return fail(FD_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
        }                                                                                  \
    } while (0)
    FD_HIP_E(hipSetDevice(e->device));
    FD_HIP_E(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    FD_HIP_E(hipEventCreate(&e->ev0));
    FD_HIP_E(hipEventCreate(&e->ev1));
    FD_HIP_E(hipEventCreateWithFlags(&e->evd, hipEventDisableTiming));
    FD_HIP_E(hipMalloc(&e->table, sizeof(EnsDomain) * (size_t)n));
    FD_HIP_E(hipMalloc(&e->ftable, sizeof(EnsDomain) * (size_t)n));
#undef FD_HIP_E
    for (fd_handle* h : e->hs) ++h->ens_refs;
    e->counted = true;
    *out = e;
    return FD_OK;
}

int fd_ens_step(fd_ensemble* e, int32_t nsteps) { return ens_step(e, nsteps, false); }

int fd_ens_step_sequential(fd_ensemble* e, int32_t nsteps) { return ens_step(e, nsteps, true); }

int fd_ens_set_tile(fd_ensemble* e, int32_t waves_x, int32_t rows) {
    if (!e) return fail(FD_E_ARG, "ensemble is NULL");
    if (waves_x != 1 && waves_x != 2 && waves_x != 4 && waves_x != 8)
        return fail(FD_E_ARG, "waves_x must be 1, 2, 4 or 8");
    if (rows < 1) return fail(FD_E_ARG, "rows must be >= 1");
    e->wx = waves_x;
    e->rows = rows;
    e->table_ok = false;
    return FD_OK;
}

int fd_ens_set_fusion(fd_ensemble* e, int32_t depth, int32_t waves_x, int32_t rows) {
    if (!e) return fail(FD_E_ARG, "ensemble is NULL");
    if (depth < 1 || depth > 3) return fail(FD_E_ARG, "depth must be 1, 2 or 3");
    if (waves_x > 0 && waves_x != 1 && waves_x != 2 && waves_x != 4) return fail(FD_E_ARG, "waves_x must be 1, 2 or 4");
    if (depth == 3 && waves_x > 2) return fail(FD_E_ARG, "three iterations per launch take 1 or 2 waves per workgroup");
    e->fuse = depth;
    if (depth == 3) {
        if (waves_x > 0) e->f3wx = waves_x;
        if (rows > 0) e->f3rows = rows;
    } else {
        if (waves_x > 0) e->fwx = waves_x;
        if (rows > 0) e->frows = rows;
    }
    e->table_ok = false;
    return FD_OK;
}

int fd_ens_use_graph(fd_ensemble* e, int32_t on) {
    if (!e) return fail(FD_E_ARG, "ensemble is NULL");
    e->use_graph = on != 0;
    return FD_OK;
}

int fd_ens_synchronize(fd_ensemble* e) {
    if (!e) return fail(FD_E_ARG, "ensemble is NULL");
    FD_HIP(hipSetDevice(e->device));
    FD_HIP(hipStreamSynchronize(e->stream));
    return FD_OK;
}

double fd_ens_last_step_ms(const fd_ensemble* e) {
    if (!e || !e->have_timing || e->timed_launches <= 0) return -1.0;
    if (hipEventSynchronize(e->ev1) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e->ev0, e->ev1) != hipSuccess) return -1.0;
    return (double)ms / e->timed_launches;
}

void fd_ens_destroy(fd_ensemble* e) {
    if (!e) return;
    if (e->counted)
        for (fd_handle* h : e->hs) --h->ens_refs;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->fexec) (void)hipGraphExecDestroy(e->fexec);
    if (e->table) (void)hipFree(e->table);
    if (e->ftable) (void)hipFree(e->ftable);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->evd) (void)hipEventDestroy(e->evd);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

}  // extern "C"
