// imexlbm.hip -- context, device memory, launch sequencing and the C ABI.
//
// The reference's step is Collision() (in place) then Streaming() (ghost fill,
// pull into ft, swap), src/lbm.cpp:131-207.  Here the context holds either
// the post-streaming state S (as the reference does) or a post-collision
// state P whose pull-stream is still owed.  From P one fused launch does
// "pull-stream, then collide" and yields the next P, so each step reads and
// writes the populations exactly once.  Anything that observes S (the
// conserved sums, a download) gathers it through the pull offsets, which is
// bit-identical to materialising it because streaming is a copy.
//
// Where the boxes allow it, pairs of fused steps run as one pass (k_step2:
// the intermediate step lives in LDS), bit-identical to two launches.
// imex_step hands the pairs out directly; Streaming() leaves at most one
// fused step owed until the next call that needs the state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/imexlbm.h"
#include "kernels.hpp"

using namespace imex;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPC(call)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(IMEX_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

#define TRY(call)                      \
    do {                               \
        int r_ = (call);               \
        if (r_ != IMEX_OK) return r_;  \
    } while (0)

constexpr double TGV_PI = 3.1415926;  // the literal the reference uses, src/lbm.cpp:3

int storage_bytes(int prec)
{
    switch (prec) {
    case IMEX_F64: return 8;
    case IMEX_F32: return 4;
    case IMEX_F16: return 2;
    case IMEX_BF16: return 2;
    default: return 0;
    }
}

long long round_up(long long v, long long a) { return (v + a - 1) / a * a; }

struct Instance {
    imex_domain dom;
    double omega;
    int box_begin, box_end;
    int tile_begin, tile_end;
    int tab[3];
};

// LBM::setup_Cartesian's factorisation (src/mpi_view_transfer.cpp:55-91): the
// first strict minimum of the spread of the int local extents, x outer.
// whole_x keeps x uncut (rx = 1) and factors y and z by the same rule.
void cartesian(int n, const int gl[3], int r[3], bool whole_x = false)
{
    double min_e = 1e16;
    r[0] = r[1] = r[2] = 1;
    for (int i = 1; i <= (whole_x ? 1 : n); ++i) {
        if (n % i) continue;
        for (int j = 1; j <= n / i; ++j) {
            if ((n / i) % j) continue;
            const int k = n / i / j;
            const double a = gl[0] / i, b = gl[1] / j, c = gl[2] / k;
            const double m = (a + b + c) / 3.0;
            const double e = (a - m) * (a - m) + (b - m) * (b - m) + (c - m) * (c - m);
            if (e < min_e) {
                min_e = e;
                r[0] = i; r[1] = j; r[2] = k;
            }
        }
    }
}

// setup_Local's block distribution (src/mpi_view_transfer.cpp:165-188).
int block_len(int gl, int r, int p) { return (p - gl % r >= 0) ? gl / r : gl / r + 1; }
int block_lo(int gl, int r, int p) { return p * (gl / r) + std::min(p, gl % r); }

// Groups of the per-phase profiler, as the reference's LBM_PROFILE build
// reports them (src/profile.cpp:64-97): the first six carry its group names,
// byte model (collision and stream 2 x cells x 27 x sizeof(storage), pack and
// unpack twice the exchange-buffer bytes, computeMacroscopic the populations
// plus four doubles per cell) and flop model (594 per collided cell, 189 per
// cell of moments).  On one GPU the pack buffer is read by the unpack
// directly, so "mpi" records nothing.  The last three have no counterpart
// there: the fused pull-stream + collide launch (one or two lattice updates;
// its bytes are those the launch must move, every population read once and
// written once), the direct ghost fill and the conserved sums.
enum Phase { PH_COLLISION = 0, PH_PACK, PH_MPI, PH_UNPACK, PH_STREAM, PH_MACRO, PH_FUSED, PH_DIRECT,
             PH_CONSERVED, PH_N };
const char *const kPhaseNames[PH_N] = {"collision", "pack (enq+fence)", "mpi (post+wait)",
                                       "unpack (enq+fence)", "stream (enq+fence)", "computeMacroscopic",
                                       "collision+stream (fused)", "ghost fill (direct)", "conserved"};
constexpr double FLOP_COLLISION = 594.0;   // src/profile.cpp:38, src/main.cpp:49
constexpr double FLOP_MOMENTS = 189.0;     // src/profile.cpp:39
struct PhaseStat {
    double ms = 0, mn = 0, mx = 0, bytes = 0, flops = 0;
    long long calls = 0;
};

}  // namespace

struct imex_ctx {
    imex_config cfg{};
    int prec = IMEX_F64;
    int ssize = 8;
    bool tiled = false;
    bool whole_x = false;        // tiled: tiles keep x whole, kernels wrap x
    int device = 0;
    int ncu = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};

    std::vector<Instance> inst;
    std::vector<BoxDesc> boxes;
    std::vector<int> box_coord;  // tile coordinates, 3 per box (tiled mode)
    std::vector<WorkTile> tiles;
    std::vector<HaloSeg> segs;
    std::vector<int2> pack_blocks, unpack_blocks, direct_blocks;
    std::vector<int2> direct_blocks_nox;   // the direct fill without the x faces
    // the fill before a pass in side_mode: x-crossing channels into the side buffers
    std::vector<int2> unpack_blocks_pass, direct_blocks_pass, direct_blocks_nox_pass;
    int2 *d_unpack_pass = nullptr, *d_direct_pass = nullptr, *d_direct_nox_pass = nullptr;
    int rgrid[3] = {1, 1, 1};

    BoxDesc *d_boxes = nullptr;
    WorkTile *d_tiles = nullptr;
    size_t d_tiles_cap = 0;
    HaloSeg *d_segs = nullptr;
    int2 *d_pack = nullptr, *d_unpack = nullptr, *d_direct = nullptr, *d_direct_nox = nullptr;
    double *d_tabs = nullptr;
    double *d_partial = nullptr;
    double *d_out = nullptr;
    double *h_out = nullptr;
    double *d_stage = nullptr;
    size_t stage_cap = 0;
    void *pool[3] = {nullptr, nullptr, nullptr};   // [2]: the autotuner's temporary pool
    long long pool_elems = 0;
    void *xbuf = nullptr;
    long long xbuf_elems = 0;

    int cur = 0;
    bool initialized = false;
    bool state_S = true;       // pool[cur] holds S (true) or P with a pull owed
    bool pending = false;      // a Collision() not yet applied
    bool ghosts_valid = false; // tiled: ghost layers of pool[cur] are current
    bool side_mode = false;    // tiled two-step with x side buffers (add_side_buffers)
    bool pghosts_valid = false; // side_mode: the ghosts the pass reads (side buffers and
                                // the y/z ghosts) of pool[cur] are current
    bool xghost_ok = false;    // tiled: the x ghost columns of pool[cur] were written by
                               // the two-step pass that produced it (BoxDesc::xnb)

    int block = 256, cpt = 1, persistent = 0, nt = 0, dir_pad = 0, halo_depth = 1, xcds = 8;
    int contiguous = 0;
    int io16 = 0;
    int xcd_run = 0;
    // two steps per pass: on/off, column tile, z-planes per sweep, tile order
    int two_step = 0, ts_tile = 0, ts_zseg = 1, ts_xcds = 1, ts_pf = 0, ts_buf = 0;
    bool halo_all = false;       // tiled two-step: ghosts carry all 27 directions
    bool halo_direct = true;     // tiled: one direct copy instead of pack + unpack
    long long direct_bytes = 0;  // bytes the direct fill moves
    long long direct_bytes_nox = 0;
    bool xdirect = false;        // the tiled two-step pass writes its x neighbours' ghosts
    PassBox *d_pass = nullptr;   // per-box workgroup table of the two-step pass
    int pass_key[2] = {-1, -1};  // tile and z-sweep the table was built for
    long long pass_grid = 0;
    int deferred = 0;            // fused steps owed by Streaming() calls, < 2
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    bool gflip[2] = {false, false};   // replaying graph[p] leaves the state in pool[1-p]

    long long cells = 0;
    long long device_bytes = 0;
    long long halo_bytes = 0;    // bytes one pack + unpack moves
    long long pack_bytes = 0;    // bytes the pack launch moves (2 x the buffer)

    bool profile = false;        // time every phase with events (adds a sync)
    bool capturing = false;      // inside a hipGraph capture: no timing
    PhaseStat ph[PH_N];
    hipEvent_t pev[2] = {nullptr, nullptr};
};

namespace {

// Times one phase with a pair of events and a synchronisation when the
// profiler is on; a no-op otherwise.
struct PhaseScope {
    imex_ctx *c;
    int p;
    double bytes, flops;
    bool on;
    PhaseScope(imex_ctx *c_, int p_, double b, double fl = 0.0)
        : c(c_), p(p_), bytes(b), flops(fl), on(c_->profile && !c_->capturing)
    {
        if (on) (void)hipEventRecord(c->pev[0], c->stream);
    }
    ~PhaseScope()
    {
        if (!on) return;
        (void)hipEventRecord(c->pev[1], c->stream);
        float ms = 0.f;
        if (hipEventSynchronize(c->pev[1]) != hipSuccess ||
            hipEventElapsedTime(&ms, c->pev[0], c->pev[1]) != hipSuccess)
            return;
        PhaseStat &st = c->ph[p];
        st.mn = st.calls ? std::min(st.mn, (double)ms) : ms;
        st.mx = st.calls ? std::max(st.mx, (double)ms) : ms;
        st.ms += ms;
        st.bytes += bytes;
        st.flops += flops;
        ++st.calls;
    }
};

}  // namespace

namespace {

void *pool_ptr(imex_ctx *c, int which) { return c->pool[which]; }

// --------------------------------------------------------------------------
// Geometry
// --------------------------------------------------------------------------
int layout_boxes(imex_ctx *c)
{
    const long long align = 128 / c->ssize;   // elements per 128 bytes
    c->boxes.clear();
    c->box_coord.clear();
    long long start = 0;
    c->cells = 0;
    for (size_t in = 0; in < c->inst.size(); ++in) {
        Instance &I = c->inst[in];
        const int gl[3] = {I.dom.nx, I.dom.ny, I.dom.nz};
        int r[3] = {1, 1, 1};
        if (c->tiled) {
            cartesian(c->cfg.ntiles, gl, r, c->whole_x);
            std::copy(r, r + 3, c->rgrid);
        }
        I.box_begin = (int)c->boxes.size();
        const int nb = r[0] * r[1] * r[2];
        for (int me = 0; me < nb; ++me) {
            // rank order me = px + rx*pz + rx*rz*py (src/mpi_view_transfer.cpp:96-98)
            const int p[3] = {me % r[0], me / r[0] / r[2], (me / r[0]) % r[2]};
            BoxDesc b{};
            b.xnb[0] = b.xnb[1] = -1;
            b.xside[0] = b.xside[1] = -1;
            int n[3];
            for (int a = 0; a < 3; ++a) {
                n[a] = block_len(gl[a], r[a], p[a]);
                b.lo[a] = block_lo(gl[a], r[a], p[a]);
                b.gl[a] = gl[a];
                b.tab[a] = I.tab[a];
                if (n[a] < 1) return fail(IMEX_EINVAL, "more tiles than cells along an axis");
                if (c->tiled && n[a] < c->halo_depth && !(a == 0 && c->whole_x))   // whole_x: no x ghosts
                    return fail(IMEX_EINVAL, "tile thinner than the halo depth");
            }
            b.nx = n[0]; b.ny = n[1]; b.nz = n[2];
            // f[z][ii][y][x]: 27 direction sub-planes per z-plane
            long long rows, zplanes, org_xy, zorg;
            if (c->tiled && c->whole_x) {
                // ghost rows and planes only: x wraps in the kernels
                b.sy = (int)round_up(n[0], align);
                rows = n[1] + 6;
                zplanes = n[2] + 6;
                org_xy = 3LL * b.sy;
                b.wrap_x = 1;
            } else if (c->tiled) {
                const long long xoff = std::max<long long>(align, 3);
                b.sy = (int)round_up(xoff + n[0] + 3, align);
                rows = n[1] + 6;
                zplanes = n[2] + 6;
                org_xy = xoff + 3LL * b.sy;
            } else {
                b.sy = n[0];
                rows = n[1];
                zplanes = n[2];
                org_xy = 0;
            }
            b.dstride = round_up((long long)b.sy * rows, align) + std::max(c->dir_pad, 0);
            const long long sz = 27LL * b.dstride;
            // one z-plane (all directions) must fit a 2 GiB buffer window, and
            // in-plane cell offsets must fit 32-bit indexing
            if (sz * c->ssize >= (1LL << 31) || (long long)n[0] * n[1] * (long long)n[0] >= (1LL << 32))
                return fail(IMEX_ERANGE, "z-plane too large for 32-bit buffer addressing");
            b.sz = (int)sz;
            zorg = c->tiled ? 3 * sz : 0;
            const long long vol = sz * zplanes;
            b.base = start;
            b.zorg = zorg;
            b.org_xy = (int)org_xy;
            b.org = start + zorg + org_xy;
            b.mag = ((1ULL << 32) + (unsigned long long)n[0] - 1) / (unsigned long long)n[0];
            b.inst = (int)in;
            b.omega = I.omega;
            b.omega_f = (float)I.omega;
            b.rho0 = I.dom.rho0;
            b.u0 = I.dom.u0;
            start += round_up(vol, align * 2);
            c->boxes.push_back(b);
            c->box_coord.insert(c->box_coord.end(), p, p + 3);
            c->cells += (long long)n[0] * n[1] * n[2];
        }
        I.box_end = (int)c->boxes.size();
    }
    c->pool_elems = start;
    return IMEX_OK;
}

// The x ghost columns of the two-step pass in side buffers (side_index):
// tiled domains split along x whose ghost tables were built for two steps
// per pass.  The pass reads its x ghosts there, and stores its first and
// last two columns there for its x neighbours; the ghost fill before a pass
// writes the x-crossing channels there.  The boxes' own ghost columns stay
// for the one-step kernels and the reads of the state, which a full fill
// serves.  Each pool holds the side buffers of the state it carries.
void add_side_buffers(imex_ctx *c)
{
    c->side_mode = c->tiled && c->halo_all && !c->whole_x;
    if (!c->side_mode) return;
    const long long align = 128 / c->ssize;
    for (BoxDesc &b : c->boxes)
        for (int sd = 0; sd < 2; ++sd) {
            b.xside[sd] = c->pool_elems;
            c->pool_elems += round_up(side_elems(b.ny, b.nz), 2 * align);
        }
}

void build_tiles(imex_ctx *c)
{
    c->tiles.clear();
    const int chunk = c->block * c->cpt;
    for (Instance &I : c->inst) {
        I.tile_begin = (int)c->tiles.size();
        for (int bi = I.box_begin; bi < I.box_end; ++bi) {
            const BoxDesc &b = c->boxes[bi];
            const int plane = b.nx * b.ny;
            for (int z = 0; z < b.nz; ++z)
                for (int p0 = 0; p0 < plane; p0 += chunk)
                    c->tiles.push_back({bi, z, p0, std::min(chunk, plane - p0)});
        }
        I.tile_end = (int)c->tiles.size();
    }
}

int upload_tiles(imex_ctx *c)
{
    const size_t need = c->tiles.size() * sizeof(WorkTile);
    if (need > c->d_tiles_cap) {
        if (c->d_tiles) HIPC(hipFree(c->d_tiles));
        c->d_tiles = nullptr;
        HIPC(hipMalloc(&c->d_tiles, need));
        c->d_tiles_cap = need;
    }
    HIPC(hipMemcpy(c->d_tiles, c->tiles.data(), need, hipMemcpyHostToDevice));
    return IMEX_OK;
}

// Directions that cross from a box into its neighbour at offset ch: every
// non-zero component of e equals that of ch.
unsigned channel_mask(const int ch[3])
{
    unsigned m = 0;
    for (int ii = 0; ii < Q; ++ii) {
        const int e[3] = {EX[ii], EY[ii], EZ[ii]};
        bool ok = true;
        for (int a = 0; a < 3; ++a)
            if (ch[a] != 0 && e[a] != ch[a]) ok = false;
        if (ok) m |= 1u << ii;
    }
    return m;
}

// Directions of a ghost cell that a two-step pass reads, for the cell at
// ghost layer kk[a] (1 or 2; 0: both, a slab two layers deep) on each axis a
// the channel ch crosses: the pass pulls from layer 1 and the ring one cell
// outside the box, so e (oriented along ch) must satisfy e_a >= kk[a] - 1:
// the crossing and the in-layer directions at layer 1, the crossing ones only
// at layer 2 (a subset, so a two-layer slab carries layer 1's set).
unsigned layer_mask(const int ch[3], const int kk[3])
{
    unsigned m = 0;
    for (int ii = 0; ii < Q; ++ii) {
        const int e[3] = {EX[ii], EY[ii], EZ[ii]};
        bool ok = true;
        for (int a = 0; a < 3; ++a)
            if (ch[a] != 0 && e[a] * ch[a] < std::max(kk[a], 1) - 1) ok = false;
        if (ok) m |= 1u << ii;
    }
    return m;
}

// Halo plan for the tiled mode: for every tile and each of its 26 neighbour
// channels (src/mpi_view_transfer.cpp:260-449), one pack segment (boundary
// slab -> exchange buffer) and one unpack segment (exchange buffer -> the
// neighbour's ghost slab on the opposite side).  One step per pass carries
// the directions that cross the channel (9 per face, 3 per edge, 1 per
// corner); two steps per pass fill two layers, each with what it needs.
void build_halo(imex_ctx *c)
{
    c->segs.clear();
    c->pack_blocks.clear();
    c->unpack_blocks.clear();
    c->direct_blocks.clear();
    c->direct_blocks_nox.clear();
    c->unpack_blocks_pass.clear();
    c->direct_blocks_pass.clear();
    c->direct_blocks_nox_pass.clear();
    std::vector<HaloSeg> direct;
    std::vector<char> xface;   // direct segment of an x face (channel (+-1, 0, 0))
    // target of each unpack / direct segment: box, x of the channel, first ghost row and plane
    std::vector<std::array<int, 4>> utarget, dtarget;
    std::vector<std::array<int, 5>> direct_key;   // box, channel (x, y, z), layer combination
    std::map<std::array<int, 5>, int> direct_at;
    const int d = c->halo_depth;
    long long xoff = 0;
    std::vector<HaloSeg> unpack;
    for (const Instance &I : c->inst) {
        const int nb = I.box_end - I.box_begin;
        for (int bi = I.box_begin; bi < I.box_end; ++bi) {
            const BoxDesc &b = c->boxes[bi];
            const int *p = &c->box_coord[3 * bi];
            for (int cx = -1; cx <= 1; ++cx)
                for (int cy = -1; cy <= 1; ++cy)
                    for (int cz = -1; cz <= 1; ++cz) {
                        if (!cx && !cy && !cz) continue;
                        if (cx && c->whole_x) continue;   // x wraps in the kernels
                        const int ch[3] = {cx, cy, cz};
                        int q[3];
                        for (int a = 0; a < 3; ++a) q[a] = (p[a] + ch[a] + c->rgrid[a]) % c->rgrid[a];
                        const int me = q[0] + c->rgrid[0] * q[2] + c->rgrid[0] * c->rgrid[2] * q[1];
                        if (me >= nb) continue;
                        const BoxDesc &t = c->boxes[I.box_begin + me];
                        const int n[3] = {b.nx, b.ny, b.nz};
                        const int nt[3] = {t.nx, t.ny, t.nz};
                        // one step per pass: every filled layer at once, with the
                        // crossing directions.  Two steps per pass: layers 1 and 2
                        // of the y and z ghost axes apart, each with the directions
                        // read there; along x both layers in one slab (they share
                        // cache lines, as x is the contiguous axis)
                        int ga[3], ng = 0;
                        for (int a = c->halo_all ? 1 : 0; a < 3; ++a)
                            if (ch[a]) ga[ng++] = a;
                        const int combos = c->halo_all ? (1 << ng) : 1;
                        for (int cb = 0; cb < combos; ++cb) {
                            int kk[3] = {0, 0, 0};   // ghost layer per axis (0: layers 1..d)
                            for (int j = 0; j < ng; ++j) kk[ga[j]] = c->halo_all ? 1 + ((cb >> j) & 1) : 0;
                            // (x, or every axis for one step per pass, stays 0: layers 1..d)
                            int s0[3], d0[3], ext[3];
                            for (int a = 0; a < 3; ++a) {
                                if (ch[a] == 0) { s0[a] = 0; d0[a] = 0; ext[a] = n[a]; }
                                else if (kk[a] == 0 && ch[a] > 0) { s0[a] = n[a] - d; d0[a] = -d; ext[a] = d; }
                                else if (kk[a] == 0) { s0[a] = 0; d0[a] = nt[a]; ext[a] = d; }
                                else if (ch[a] > 0) { s0[a] = n[a] - kk[a]; d0[a] = -kk[a]; ext[a] = 1; }
                                else { s0[a] = kk[a] - 1; d0[a] = nt[a] + kk[a] - 1; ext[a] = 1; }
                            }
                            const unsigned mask = c->halo_all ? layer_mask(ch, kk) : channel_mask(ch);
                            const int ndir = __builtin_popcount(mask);
                            HaloSeg ps{};
                            ps.ex = ext[0]; ps.ey = ext[1]; ps.ez = ext[2];
                            ps.ncell = ext[0] * ext[1] * ext[2];
                            ps.mask = mask;
                            ps.pack = 1;
                            ps.partner = -1;
                            ps.s_off = b.org + s0[0] + (long long)s0[1] * b.sy + (long long)s0[2] * b.sz;
                            ps.s_ds = b.dstride;
                            ps.s_sy = b.sy; ps.s_sz = b.sz;
                            ps.d_off = xoff;
                            ps.d_ds = ps.ncell;
                            ps.d_sy = ext[0]; ps.d_sz = ext[0] * ext[1];
                            HaloSeg us = ps;
                            us.pack = 0;
                            us.s_off = xoff;
                            us.s_ds = ps.ncell;
                            us.s_sy = ext[0]; us.s_sz = ext[0] * ext[1];
                            us.d_off = t.org + d0[0] + (long long)d0[1] * t.sy + (long long)d0[2] * t.sz;
                            us.d_ds = t.dstride;
                            us.d_sy = t.sy; us.d_sz = t.sz;
                            xoff += round_up((long long)ndir * ps.ncell, 64);
                            HaloSeg ds = ps;   // boundary slab of b -> ghost slab of t
                            ds.pack = 2;
                            ds.d_off = us.d_off;
                            ds.d_ds = us.d_ds;
                            ds.d_sy = us.d_sy; ds.d_sz = us.d_sz;
                            c->segs.push_back(ps);
                            unpack.push_back(us);
                            utarget.push_back({I.box_begin + me, cx, d0[1], d0[2]});
                            dtarget.push_back({I.box_begin + me, cx, d0[1], d0[2]});
                            direct_at[{bi, cx, cy, cz, cb}] = (int)direct.size();
                            direct_key.push_back({I.box_begin + me, -cx, -cy, -cz, cb});
                            direct.push_back(ds);
                            xface.push_back(cx != 0 && cy == 0 && cz == 0);
                        }
                    }
        }
    }
    const int npack = (int)c->segs.size();
    for (int s = 0; s < npack; ++s)
        for (int c0 = 0; c0 < c->segs[s].ncell; c0 += 256) c->pack_blocks.push_back({s, c0});
    for (int s = 0; s < (int)unpack.size(); ++s)
        for (int c0 = 0; c0 < unpack[s].ncell; c0 += 256) c->unpack_blocks.push_back({npack + s, c0});
    c->segs.insert(c->segs.end(), unpack.begin(), unpack.end());
    const int nsu = (int)c->segs.size();
    // pair each direct segment with its reverse: (b -> t across ch) with
    // (t -> b across -ch), which has the same layer combination and extents
    for (int s = 0; s < (int)direct.size(); ++s) {
        const auto it = direct_at.find(direct_key[s]);
        const int q = it == direct_at.end() ? -1 : it->second;
        const bool same = q >= 0 && direct[q].ex == direct[s].ex && direct[q].ey == direct[s].ey &&
                          direct[q].ez == direct[s].ez;
        if (same && q > s) direct[s].partner = nsu + q;
        if (same && q < s) continue;               // copied with its partner
        for (int c0 = 0; c0 < direct[s].ncell; c0 += 256) {
            c->direct_blocks.push_back({nsu + s, c0});
            if (!xface[s]) c->direct_blocks_nox.push_back({nsu + s, c0});
        }
    }
    c->segs.insert(c->segs.end(), direct.begin(), direct.end());
    // With side buffers, the fill before a pass writes the channels that
    // cross x into them: a copy of each such unpack and direct segment whose
    // destination is the side buffer (two columns; rows and planes as in the
    // box, ghosts included), the others shared with the full fill.
    if (c->side_mode) {
        const int nseg = (int)c->segs.size();
        std::vector<int> side_copy(nseg, -1);
        for (int s = npack; s < nseg; ++s) {
            const std::array<int, 4> &tg = s < nsu ? utarget[s - npack] : dtarget[s - nsu];
            if (tg[1] == 0) continue;
            const BoxDesc &t = c->boxes[tg[0]];
            HaloSeg sg = c->segs[s];
            // crossing +x fills the lower side of the target, -x the upper
            sg.d_off = side_index(t, tg[1] > 0 ? 0 : 1, 0, tg[2], tg[3], 0);
            sg.d_ds = 2LL * (t.ny + 6);
            sg.d_sy = 2;
            sg.d_sz = 2 * Q * (t.ny + 6);
            side_copy[s] = (int)c->segs.size();
            c->segs.push_back(sg);
        }
        for (int s = 0; s < nseg; ++s)
            if (side_copy[s] >= 0 && c->segs[s].partner >= 0)
                c->segs[side_copy[s]].partner = side_copy[c->segs[s].partner];
        auto remap = [&](const std::vector<int2> &from, std::vector<int2> &to) {
            for (const int2 &bk : from) to.push_back({side_copy[bk.x] >= 0 ? side_copy[bk.x] : bk.x, bk.y});
        };
        remap(c->unpack_blocks, c->unpack_blocks_pass);
        remap(c->direct_blocks, c->direct_blocks_pass);
        remap(c->direct_blocks_nox, c->direct_blocks_nox_pass);
    }
    c->xbuf_elems = xoff;
    c->halo_bytes = 0;
    c->pack_bytes = 0;
    c->direct_bytes = 0;
    c->direct_bytes_nox = 0;
    // each segment reads and writes its cells (a side-buffer copy as many as its original)
    for (int s = 0; s < nsu + (int)direct.size(); ++s) {
        const HaloSeg &sg = c->segs[s];
        const long long b = 2LL * __builtin_popcount(sg.mask) * sg.ncell * c->ssize;
        if (sg.pack == 2) {
            c->direct_bytes += b;
            if (!xface[s - nsu]) c->direct_bytes_nox += b;
        } else {
            c->halo_bytes += b;
            if (sg.pack == 1) c->pack_bytes += b;
        }
    }
    // Two steps per pass with the direct fill: the pass writes the two x ghost
    // columns of its x neighbours itself (every direction, from the values it
    // stores), and the fill before the next pass leaves the x faces out.  The
    // x neighbours of a box share its y and z extents (one Cartesian row).
    for (BoxDesc &b : c->boxes) b.xnb[0] = b.xnb[1] = -1;
    c->xdirect = c->halo_direct && c->halo_all && c->halo_depth == 2 && !c->whole_x && c->inst.size() == 1;
    if (c->xdirect) {
        const Instance &I = c->inst[0];
        const int nb = I.box_end - I.box_begin;
        for (int bi = I.box_begin; bi < I.box_end && c->xdirect; ++bi) {
            const int *p = &c->box_coord[3 * bi];
            for (int side = 0; side < 2; ++side) {
                const int qx = (p[0] + (side ? 1 : -1) + c->rgrid[0]) % c->rgrid[0];
                const int me = qx + c->rgrid[0] * p[2] + c->rgrid[0] * c->rgrid[2] * p[1];
                const BoxDesc &b = c->boxes[bi];
                if (me >= nb) { c->xdirect = false; break; }
                const BoxDesc &t = c->boxes[I.box_begin + me];
                // the pass writes into the neighbour's side buffer of that side
                if (t.ny != b.ny || t.nz != b.nz || t.nx < 2 || b.nx < 2 || t.xside[1 - side] < 0) {
                    c->xdirect = false;
                    break;
                }
                c->boxes[bi].xnb[side] = I.box_begin + me;
            }
        }
        if (!c->xdirect)
            for (BoxDesc &b : c->boxes) b.xnb[0] = b.xnb[1] = -1;
    }
}

// --------------------------------------------------------------------------
// Launchers
// --------------------------------------------------------------------------
template <class S>
int launch_step_t(imex_ctx *c, int op, int src, int dst, int t0, int n)
{
    const S *sp = (const S *)pool_ptr(c, src);
    S *dp = (S *)pool_ptr(c, dst);
    const WorkTile *tl = c->d_tiles + t0;
    const int grid = c->persistent > 0 ? std::min(c->persistent, n) : n;
    const dim3 g(grid), b(c->block);
    const bool nt = c->nt != 0;
    // two cells per thread (tiles of 2*block cells or more) use the paired kernel
    const bool pair = c->cpt >= 2 && op == OP_FUSED;
    const bool g16 = c->io16 == 1;
#define LAUNCH(W, O, N, P, G) hipLaunchKernelGGL((k_step<S, W, O, N, P, G>), g, b, 0, c->stream, sp, dp, c->d_boxes, tl, n, c->xcds, c->xcd_run)
#define FUSED_G(W, G)                                                                                  \
    do {                                                                                               \
        if (pair) { if (nt) LAUNCH(W, OP_FUSED, true, true, G); else LAUNCH(W, OP_FUSED, false, true, G); } \
        else { if (nt) LAUNCH(W, OP_FUSED, true, false, G); else LAUNCH(W, OP_FUSED, false, false, G); }   \
    } while (0)
#define FUSED(W)                                                 \
    do {                                                         \
        if (g16) FUSED_G(W, true); else FUSED_G(W, false);       \
    } while (0)
    if (!c->tiled) {
        if (op == OP_FUSED) FUSED(true);
        else if (op == OP_COLLIDE) LAUNCH(true, OP_COLLIDE, false, false, false);
        else LAUNCH(true, OP_STREAM, false, false, false);
    } else {
        if (op == OP_FUSED) FUSED(false);
        else if (op == OP_COLLIDE) LAUNCH(false, OP_COLLIDE, false, false, false);
        else LAUNCH(false, OP_STREAM, false, false, false);
    }
#undef FUSED
#undef FUSED_G
#undef LAUNCH
    HIPC(hipGetLastError());
    return IMEX_OK;
}

int launch_step(imex_ctx *c, int op, int src, int dst)
{
    PhaseScope ps(c, op == OP_FUSED ? PH_FUSED : (op == OP_COLLIDE ? PH_COLLISION : PH_STREAM),
                  2.0 * Q * c->ssize * (double)c->cells, op == OP_STREAM ? 0.0 : FLOP_COLLISION * (double)c->cells);
    auto one = [&](int t0, int n) -> int {
        if (n <= 0) return IMEX_OK;
        switch (c->prec) {
        case IMEX_F64: return launch_step_t<double>(c, op, src, dst, t0, n);
        case IMEX_F32: return launch_step_t<float>(c, op, src, dst, t0, n);
        case IMEX_F16: return launch_step_t<_Float16>(c, op, src, dst, t0, n);
        default: return launch_step_t<bf16_t>(c, op, src, dst, t0, n);
        }
    };
    if (c->cfg.sequential && c->inst.size() > 1) {
        for (const Instance &I : c->inst) TRY(one(I.tile_begin, I.tile_end - I.tile_begin));
        return IMEX_OK;
    }
    return one(0, (int)c->tiles.size());
}

// Two-step pass: column tile {TX, TY} of each tile variant (3 is the 32x8
// tile with 384 threads: one ring cell per thread; 4 the 64x16 tile with
// 1024 threads, 2-byte formats only: its ring takes 128 KB of LDS).
void ts_dims(int tile, int &tx, int &ty)
{
    tx = (tile == 0 || tile == 4) ? 64 : 32;
    ty = (tile == 1 || tile == 4) ? 16 : 8;
}

// Periodic boxes (one domain or a batch, at least half a column tile across
// and one column tile deep in y; the last column may be ragged, and a box
// narrower than the tile wraps its ring twice) or the tiles of one tiled
// domain (divided
// exactly by the column tile, at least two planes deep), in one-launch mode;
// FP64 builds only the 32x8 tiles (its LDS ring needs 147 KB).
bool ts_allowed(const imex_ctx *c, int tile)
{
    if (tile < 0 || tile > 4 || (c->tiled && c->inst.size() != 1)) return false;
    if (c->inst.size() > 1 && c->cfg.sequential) return false;
    if (c->ssize == 8 && (tile < 2 || tile == 4)) return false;
    if (c->ssize == 4 && tile == 4) return false;
    int tx, ty;
    ts_dims(tile, tx, ty);
    for (const BoxDesc &b : c->boxes) {
        if (c->tiled ? (b.nx % tx || b.ny % ty || b.nz < 2) : (2 * b.nx < tx || b.ny < ty)) return false;
    }
    return true;
}

// Prefetch mode of the two-step pass: 0 none, 1 every ring cell's
// populations in registers, 2 the partial prefetch (2-byte formats, tiles
// whose every thread owns a ring cell: 0, 1, 2 and 4; elsewhere 2 means 1).
// A negative request picks 2 for 2-byte formats and 1 for the others, except
// for BF16 in the 64x16 tile over ghost cells, whose 1024 threads leave 128
// registers each: there the prefetch spills, so it runs without.  (FP16 fits:
// its collision over ghost-bounded boxes reads the halves without widening
// them first.)
int ts_pf_for(const imex_ctx *c, int tile, int want)
{
    const bool partial_ok = c->ssize == 2 && tile != 3;
    if (want < 0) {
        if (c->prec == IMEX_BF16 && tile == 4 && c->tiled) return 0;
        return partial_ok ? 2 : 1;
    }
    if (want >= 2) return partial_ok ? 2 : 1;
    return want;
}

// Cells the pass's column tiles cover over the cells of the boxes (1 when
// the tile divides every box).
double ts_cover(const imex_ctx *c, int tile)
{
    int tx, ty;
    ts_dims(tile, tx, ty);
    double cov = 0, cells = 0;
    for (const BoxDesc &b : c->boxes) {
        cov += (double)((b.nx + tx - 1) / tx * tx) * ((b.ny + ty - 1) / ty * ty) * b.nz;
        cells += (double)b.nx * b.ny * b.nz;
    }
    return cells > 0 ? cov / cells : 1e30;
}

// Per-box workgroup table of the two-step pass for the current tile and
// z-sweep length; rebuilt (synchronously, outside any capture) when they
// change.
int ts_prepare(imex_ctx *c)
{
    if (c->pass_key[0] == c->ts_tile && c->pass_key[1] == c->ts_zseg) return IMEX_OK;
    if (c->capturing) return fail(IMEX_ESTATE, "two-step table changed under capture");
    int tx, ty;
    ts_dims(c->ts_tile, tx, ty);
    std::vector<PassBox> t;
    long long wg = 0;
    for (int bi = 0; bi < (int)c->boxes.size(); ++bi) {
        const BoxDesc &b = c->boxes[bi];
        const int zseg = std::max(1, std::min(c->ts_zseg, b.nz));
        const int tiles_x = (b.nx + tx - 1) / tx, ntile = tiles_x * ((b.ny + ty - 1) / ty);
        t.push_back({(int)wg, bi, tiles_x, ntile});
        wg += (long long)ntile * ((b.nz + zseg - 1) / zseg);
    }
    if (wg > 0x7fffffffLL) return fail(IMEX_ERANGE, "two-step grid too large");
    HIPC(hipMemcpy(c->d_pass, t.data(), t.size() * sizeof(PassBox), hipMemcpyHostToDevice));
    c->pass_grid = wg;
    c->pass_key[0] = c->ts_tile;
    c->pass_key[1] = c->ts_zseg;
    return IMEX_OK;
}

// Workgroups of a two-step pass with z-sweeps of zseg planes.
long long ts_grid(const imex_ctx *c, int tile, int zseg)
{
    int tx, ty;
    ts_dims(tile, tx, ty);
    long long g = 0;
    for (const BoxDesc &b : c->boxes)
        g += (long long)((b.nx + tx - 1) / tx) * ((b.ny + ty - 1) / ty) * ((b.nz + zseg - 1) / zseg);
    return g;
}

// Planes per z-sweep: more sweeps give more workgroups, and with the tile
// order by XCD each XCD a z-range of its own, but every sweep recomputes two
// intermediate planes.  Autotuned optima at 256^3 and 512^3: the 64x16 tile
// (one 128 KB workgroup per CU) sweeps as far as one workgroup per CU
// allows; otherwise near nz/8 (at least 32) for 2-byte formats and nz/4 (at
// least 64) for the others, whose passes carry more work per plane, halved
// (not below 16 planes) until the pass has two workgroups per CU.  Returns 0
// when the pass would not have one workgroup per CU even then: such a small
// domain runs faster one step per pass.
int ts_default_zseg(const imex_ctx *c, int tile)
{
    int nz = 1;
    for (const BoxDesc &b : c->boxes) nz = std::max(nz, b.nz);
    const long long want = (tile == 4 ? 1LL : 2LL) * c->ncu;
    int z = tile == 4 ? nz : std::min(nz, c->ssize == 2 ? std::max(32, nz / 8) : std::max(64, nz / 4));
    while (z > 16 && ts_grid(c, tile, z) < want) z = std::max(16, z / 2);
    return ts_grid(c, tile, z) >= c->ncu ? z : 0;
}

// Planes per sweep tried for a tile (by the default choice and the
// autotuner): the rule above, and nz, nz/2, ..., nz/16 down to 16 planes.
std::vector<int> ts_zseg_candidates(const imex_ctx *c, int tile)
{
    int nz = 1;
    for (const BoxDesc &b : c->boxes) nz = std::max(nz, b.nz);
    std::vector<int> zs;
    if (const int z = ts_default_zseg(c, tile)) zs.push_back(z);
    for (int div : {1, 2, 4, 8, 16})
        if (nz / div >= std::min(16, nz) && std::find(zs.begin(), zs.end(), nz / div) == zs.end())
            zs.push_back(nz / div);
    return zs;
}

// Modelled duration of a two-step pass: waves of workgroups x planes of the
// longest sweep (two of them recomputed) x the plane time of the tile.  The
// plane times (microseconds per plane for one workgroup, every CU busy) and
// the workgroups that fit a CU (LDS) were measured at 512^3 on MI355X.  The
// wave count prices the idle tail of a pass whose workgroups do not fill its
// last wave, the main reason one sweep length or tile beats another at sizes
// other than powers of two.  Only the default choice uses it; the autotuner
// measures.
double ts_model(const imex_ctx *c, int tile, int zseg)
{
    static const double us2[5] = {7.23, 8.53, 9.08, 10.5, 6.37};    // 2-byte formats, tiles 0..4
    static const double us4[4] = {5.98, 6.29, 6.56, 7.12};          // FP32
    static const double us8[4] = {0.0, 0.0, 6.66, 6.15};            // FP64 (tiles 2, 3)
    static const int wpc2[5] = {2, 2, 4, 4, 1}, wpc4[4] = {1, 1, 2, 2};
    const double us = c->ssize == 2 ? us2[tile] : (c->ssize == 4 ? us4[tile] : us8[tile]);
    const int wpc = c->ssize == 2 ? wpc2[tile] : (c->ssize == 4 ? wpc4[tile] : 1);
    int tx, ty;
    ts_dims(tile, tx, ty);
    long long g = 0;
    int zl = 0;
    for (const BoxDesc &b : c->boxes) {
        const int z = std::max(1, std::min(zseg, b.nz));
        zl = std::max(zl, z);
        g += (long long)((b.nx + tx - 1) / tx) * ((b.ny + ty - 1) / ty) * ((b.nz + z - 1) / z);
    }
    const long long slots = (long long)c->ncu * wpc;
    return (double)((g + slots - 1) / slots) * (zl + 2) * us;
}

template <class S, int TX, int TY, int NTH>
int launch_ts(imex_ctx *c, int src, int dst)
{
    TRY(ts_prepare(c));
    const dim3 g((unsigned)c->pass_grid), b(NTH);
    const S *sp = (const S *)pool_ptr(c, src);
    S *dp = (S *)pool_ptr(c, dst);
    const int npb = (int)c->boxes.size(), zseg = std::max(1, c->ts_zseg);
#define TS(W, P, G) hipLaunchKernelGGL((k_step2<S, W, TX, TY, NTH, true, P, G>), g, b, 0, c->stream, sp, dp, \
                                       c->d_boxes, c->d_pass, npb, c->ts_xcds, zseg)
    // partial prefetch where every thread owns a ring cell (2-byte formats)
    constexpr bool PARTIAL = sizeof(S) == 2 && NTH <= (TX + 2) * (TY + 2);
    if constexpr (PARTIAL) {
        if (c->ts_pf == 2) {
#define TSP(W, G) hipLaunchKernelGGL((k_step2<S, W, TX, TY, NTH, true, false, G, PARTIAL_PD>), g, b, 0, c->stream, \
                                     sp, dp, c->d_boxes, c->d_pass, npb, c->ts_xcds, zseg)
            if (c->tiled) TSP(false, true);
            else if (c->ts_buf) TSP(true, false);
            else TSP(true, true);
#undef TSP
            HIPC(hipGetLastError());
            return IMEX_OK;
        }
    }
    if (c->tiled) {
        if (c->ts_pf) TS(false, true, true); else TS(false, false, true);
    } else if (c->ts_buf) {
        if (c->ts_pf) TS(true, true, false); else TS(true, false, false);
    } else {
        if (c->ts_pf) TS(true, true, true); else TS(true, false, true);
    }
#undef TS
    HIPC(hipGetLastError());
    return IMEX_OK;
}

template <class S>
int launch_step2_t(imex_ctx *c, int src, int dst)
{
    if constexpr (sizeof(S) == 8) {
        if (c->ts_tile == 3) return launch_ts<S, 32, 8, 384>(c, src, dst);
        return launch_ts<S, 32, 8, 256>(c, src, dst);
    } else {
        switch (c->ts_tile) {
        case 0: return launch_ts<S, 64, 8, 512>(c, src, dst);
        case 1: return launch_ts<S, 32, 16, 512>(c, src, dst);
        case 3: return launch_ts<S, 32, 8, 384>(c, src, dst);
        case 4:
            if constexpr (sizeof(S) == 2) return launch_ts<S, 64, 16, 1024>(c, src, dst);
            return fail(IMEX_EINVAL, "two-step tile 4 is built for 2-byte formats only");
        default: return launch_ts<S, 32, 8, 256>(c, src, dst);
        }
    }
}

// Two fused steps pool[src] -> pool[dst] in one pass (the caller checked
// ts_allowed for c->ts_tile).
int launch_step2(imex_ctx *c, int src, int dst)
{
    PhaseScope ps(c, PH_FUSED, 2.0 * Q * c->ssize * (double)c->cells, 2.0 * FLOP_COLLISION * (double)c->cells);
    switch (c->prec) {
    case IMEX_F64: return launch_step2_t<double>(c, src, dst);
    case IMEX_F32: return launch_step2_t<float>(c, src, dst);
    case IMEX_F16: return launch_step2_t<_Float16>(c, src, dst);
    default: return launch_step2_t<bf16_t>(c, src, dst);
    }
}

// The ghost fill of pool[which]: the full one (every ghost of every box), or
// with side buffers and `pass`, the one the two-step pass reads (the side
// buffers and the y and z ghosts).
template <class S>
int launch_halo_t(imex_ctx *c, int which, bool pass)
{
    S *f = (S *)pool_ptr(c, which);
    S *x = (S *)c->xbuf;
    const bool side = pass && c->side_mode;
    if (c->halo_direct) {
        // x faces left out when the pass that produced the pool wrote them
        // (with side buffers it wrote those, which only this fill serves)
        const bool nox = c->xghost_ok && !c->capturing && (side || !c->side_mode);
        const std::vector<int2> &bl = side ? (nox ? c->direct_blocks_nox_pass : c->direct_blocks_pass)
                                           : (nox ? c->direct_blocks_nox : c->direct_blocks);
        int2 *dbl = side ? (nox ? c->d_direct_nox_pass : c->d_direct_pass) : (nox ? c->d_direct_nox : c->d_direct);
        const int nd = (int)bl.size();
        PhaseScope ps(c, PH_DIRECT, (double)(nox ? c->direct_bytes_nox : c->direct_bytes));
        if (nd) hipLaunchKernelGGL((k_halo<S>), dim3(nd), dim3(256), 0, c->stream, f, f, c->d_segs, dbl, nd);
        HIPC(hipGetLastError());
        return IMEX_OK;
    }
    const int np = (int)c->pack_blocks.size(), nu = (int)c->unpack_blocks.size();
    {
        PhaseScope ps(c, PH_PACK, (double)c->pack_bytes);
        if (np) hipLaunchKernelGGL((k_halo<S>), dim3(np), dim3(256), 0, c->stream, f, x, c->d_segs, c->d_pack, np);
        HIPC(hipGetLastError());
    }
    PhaseScope ps(c, PH_UNPACK, (double)(c->halo_bytes - c->pack_bytes));
    if (nu)
        hipLaunchKernelGGL((k_halo<S>), dim3(nu), dim3(256), 0, c->stream, x, f, c->d_segs,
                           side ? c->d_unpack_pass : c->d_unpack, nu);
    HIPC(hipGetLastError());
    return IMEX_OK;
}

// Fill the ghosts of pool[cur] unless they are current: those of the
// two-step pass (`pass`) or all of them.  Without side buffers the two are
// the same fill.
int ensure_ghosts(imex_ctx *c, bool pass = false)
{
    if (!c->tiled) return IMEX_OK;
    const bool side = pass && c->side_mode;
    if (side ? c->pghosts_valid : c->ghosts_valid) return IMEX_OK;
    switch (c->prec) {
    case IMEX_F64: TRY(launch_halo_t<double>(c, c->cur, side)); break;
    case IMEX_F32: TRY(launch_halo_t<float>(c, c->cur, side)); break;
    case IMEX_F16: TRY(launch_halo_t<_Float16>(c, c->cur, side)); break;
    default: TRY(launch_halo_t<bf16_t>(c, c->cur, side)); break;
    }
    (side ? c->pghosts_valid : c->ghosts_valid) = true;
    return IMEX_OK;
}

int do_collide(imex_ctx *c)
{
    TRY(launch_step(c, OP_COLLIDE, c->cur, c->cur));
    c->ghosts_valid = c->pghosts_valid = false;
    c->xghost_ok = false;
    return IMEX_OK;
}

int do_fused(imex_ctx *c)
{
    TRY(ensure_ghosts(c));
    TRY(launch_step(c, OP_FUSED, c->cur, 1 - c->cur));
    c->cur = 1 - c->cur;
    c->ghosts_valid = c->pghosts_valid = false;
    c->xghost_ok = false;
    return IMEX_OK;
}

int do_fused2(imex_ctx *c)
{
    TRY(ensure_ghosts(c, true));
    TRY(launch_step2(c, c->cur, 1 - c->cur));
    c->cur = 1 - c->cur;
    c->ghosts_valid = c->pghosts_valid = false;
    c->xghost_ok = c->xdirect && !c->capturing;
    return IMEX_OK;
}

// Two fused steps: one pass when the two-step mode is on.
int do_pair(imex_ctx *c)
{
    if (c->two_step) return do_fused2(c);
    TRY(do_fused(c));
    return do_fused(c);
}

// Launch a fused step that a Streaming() call left owed.
int flush_deferred(imex_ctx *c)
{
    if (c->deferred > 0) {
        c->deferred = 0;
        TRY(do_fused(c));
    }
    return IMEX_OK;
}

int do_stream(imex_ctx *c)
{
    TRY(ensure_ghosts(c));
    TRY(launch_step(c, OP_STREAM, c->cur, 1 - c->cur));
    c->cur = 1 - c->cur;
    c->ghosts_valid = c->pghosts_valid = false;
    c->xghost_ok = false;
    return IMEX_OK;
}

// Apply a deferred Collision() whose Streaming() has not come: the logical
// state becomes the collided one, stored as S.
int flush_pending(imex_ctx *c)
{
    TRY(flush_deferred(c));
    if (!c->pending) return IMEX_OK;
    if (c->state_S) TRY(do_collide(c));
    else TRY(do_fused(c));
    c->state_S = true;
    c->pending = false;
    return IMEX_OK;
}

int one_step(imex_ctx *c)
{
    TRY(flush_pending(c));
    if (c->state_S) TRY(do_collide(c));
    else TRY(do_fused(c));
    c->state_S = false;
    return IMEX_OK;
}

void drop_graphs(imex_ctx *c)
{
    for (int k = 0; k < 2; ++k) {
        if (c->gexec[k]) (void)hipGraphExecDestroy(c->gexec[k]);
        if (c->graph[k]) (void)hipGraphDestroy(c->graph[k]);
        c->gexec[k] = nullptr;
        c->graph[k] = nullptr;
    }
}

// Two fused steps starting from pool[parity], captured once and replayed.
int build_graph(imex_ctx *c, int parity)
{
    const int save_cur = c->cur;
    const bool save_gv = c->ghosts_valid, save_pgv = c->pghosts_valid;
    c->cur = parity;
    c->ghosts_valid = c->pghosts_valid = false;
    c->xghost_ok = false;
    if (c->two_step) TRY(ts_prepare(c));
    HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    const int r = do_pair(c);
    c->capturing = false;
    c->gflip[parity] = c->cur != parity;   // one two-step pass flips the pools, two steps do not
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->cur = save_cur;
    c->ghosts_valid = save_gv;
    c->pghosts_valid = save_pgv;
    if (r != IMEX_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return fail(IMEX_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    c->graph[parity] = g;
    HIPC(hipGraphInstantiate(&c->gexec[parity], g, nullptr, nullptr, 0));
    return IMEX_OK;
}

int run_steps(imex_ctx *c, int nsteps)
{
    if (nsteps > 0) TRY(flush_deferred(c));
    while (nsteps > 0) {
        if (nsteps >= 2 && !c->state_S && !c->pending) {
            if (c->cfg.use_graph && !c->profile) {
                if (!c->gexec[c->cur]) TRY(build_graph(c, c->cur));
                HIPC(hipGraphLaunch(c->gexec[c->cur], c->stream));
                if (c->gflip[c->cur]) c->cur = 1 - c->cur;
                c->ghosts_valid = c->pghosts_valid = false;
                c->xghost_ok = false;
            } else if (c->two_step) {
                TRY(do_fused2(c));
            } else {
                TRY(do_pair(c));
            }
            nsteps -= 2;
            continue;
        }
        TRY(one_step(c));
        --nsteps;
    }
    return IMEX_OK;
}

template <class S>
int launch_conserved_t(imex_ctx *c, int t0, int n, bool gather)
{
    const S *sp = (const S *)pool_ptr(c, c->cur);
    const int grid = std::min(n, 1024);
    const WorkTile *tl = c->d_tiles + t0;
    if (!c->tiled) {
        if (gather) hipLaunchKernelGGL((k_conserved<S, true, true>), dim3(grid), dim3(RED_BLOCK), 0, c->stream, sp, c->d_boxes, tl, n, c->cpt, c->d_partial);
        else hipLaunchKernelGGL((k_conserved<S, true, false>), dim3(grid), dim3(RED_BLOCK), 0, c->stream, sp, c->d_boxes, tl, n, c->cpt, c->d_partial);
    } else {
        if (gather) hipLaunchKernelGGL((k_conserved<S, false, true>), dim3(grid), dim3(RED_BLOCK), 0, c->stream, sp, c->d_boxes, tl, n, c->cpt, c->d_partial);
        else hipLaunchKernelGGL((k_conserved<S, false, false>), dim3(grid), dim3(RED_BLOCK), 0, c->stream, sp, c->d_boxes, tl, n, c->cpt, c->d_partial);
    }
    hipLaunchKernelGGL(k_conserved_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, c->d_partial, grid, c->d_out);
    HIPC(hipGetLastError());
    return IMEX_OK;
}

template <class S>
int launch_export_t(imex_ctx *c, int box, bool gather)
{
    const S *sp = (const S *)pool_ptr(c, c->cur);
    const BoxDesc &b = c->boxes[box];
    const long long ncell = (long long)b.nx * b.ny * b.nz;
    const int grid = (int)std::min<long long>((ncell + 255) / 256, 65536);
    if (!c->tiled) {
        if (gather) hipLaunchKernelGGL((k_export<S, true, true>), dim3(grid), dim3(256), 0, c->stream, sp, c->d_boxes, box, c->d_stage);
        else hipLaunchKernelGGL((k_export<S, true, false>), dim3(grid), dim3(256), 0, c->stream, sp, c->d_boxes, box, c->d_stage);
    } else {
        if (gather) hipLaunchKernelGGL((k_export<S, false, true>), dim3(grid), dim3(256), 0, c->stream, sp, c->d_boxes, box, c->d_stage);
        else hipLaunchKernelGGL((k_export<S, false, false>), dim3(grid), dim3(256), 0, c->stream, sp, c->d_boxes, box, c->d_stage);
    }
    HIPC(hipGetLastError());
    return IMEX_OK;
}

template <class S>
int launch_macro_t(imex_ctx *c, int box, bool gather, double *out)
{
    const S *sp = (const S *)pool_ptr(c, c->cur);
    const BoxDesc &b = c->boxes[box];
    const long long ncell = (long long)b.nx * b.ny * b.nz;
    const int grid = (int)std::min<long long>((ncell + 255) / 256, 65536);
    if (!c->tiled) {
        if (gather) hipLaunchKernelGGL((k_macro<S, true, true>), dim3(grid), dim3(256), 0, c->stream, sp, c->d_boxes, box, out);
        else hipLaunchKernelGGL((k_macro<S, true, false>), dim3(grid), dim3(256), 0, c->stream, sp, c->d_boxes, box, out);
    } else {
        if (gather) hipLaunchKernelGGL((k_macro<S, false, true>), dim3(grid), dim3(256), 0, c->stream, sp, c->d_boxes, box, out);
        else hipLaunchKernelGGL((k_macro<S, false, false>), dim3(grid), dim3(256), 0, c->stream, sp, c->d_boxes, box, out);
    }
    HIPC(hipGetLastError());
    return IMEX_OK;
}

template <class S>
int launch_import_t(imex_ctx *c, int box)
{
    S *dp = (S *)pool_ptr(c, c->cur);
    const BoxDesc &b = c->boxes[box];
    const long long ncell = (long long)b.nx * b.ny * b.nz;
    const int grid = (int)std::min<long long>((ncell + 255) / 256, 65536);
    hipLaunchKernelGGL((k_import<S>), dim3(grid), dim3(256), 0, c->stream, dp, c->d_boxes, box, c->d_stage);
    HIPC(hipGetLastError());
    return IMEX_OK;
}

template <class S>
int launch_init_t(imex_ctx *c)
{
    S *dp = (S *)pool_ptr(c, c->cur);
    const int n = (int)c->tiles.size();
    hipLaunchKernelGGL((k_init<S>), dim3(std::min(n, 65536)), dim3(c->block), 0, c->stream, dp, c->d_boxes, c->d_tiles, n, c->cpt, c->d_tabs);
    HIPC(hipGetLastError());
    return IMEX_OK;
}

#define DISPATCH(fn, ...)                                                   \
    (c->prec == IMEX_F64 ? fn<double>(__VA_ARGS__)                          \
     : c->prec == IMEX_F32 ? fn<float>(__VA_ARGS__)                         \
     : c->prec == IMEX_F16 ? fn<_Float16>(__VA_ARGS__)                      \
                           : fn<bf16_t>(__VA_ARGS__))

int ensure_stage(imex_ctx *c, size_t doubles)
{
    if (doubles <= c->stage_cap) return IMEX_OK;
    if (c->d_stage) HIPC(hipFree(c->d_stage));
    c->d_stage = nullptr;
    c->stage_cap = 0;
    if (hipMalloc(&c->d_stage, doubles * sizeof(double)) != hipSuccess)
        return fail(IMEX_ENOMEM, "staging buffer allocation failed");
    c->stage_cap = doubles;
    return IMEX_OK;
}

int check_ctx(imex_ctx *c)
{
    if (!c) return fail(IMEX_EINVAL, "null context");
    HIPC(hipSetDevice(c->device));
    return IMEX_OK;
}

int check_domain(imex_ctx *c, int domain)
{
    if (domain < 0 || domain >= (int)c->inst.size()) return fail(IMEX_EINVAL, "domain index out of range");
    if (!c->initialized) return fail(IMEX_ESTATE, "imex_initialize has not been called");
    return IMEX_OK;
}

// --------------------------------------------------------------------------
// Scratch of the allocation-free conserved sums
// --------------------------------------------------------------------------
constexpr int kNoallocTile = 512;   // cells per work tile: a default context's 256 threads x 2 cells
constexpr int kScratchDoubles = IMEX_NOALLOC_PARTIALS * 10 + 8;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
double *g_scratch_at[kMaxDevices] = {};   // device address of `scratch` per device ordinal

}  // namespace

// Per-workgroup partials (5 sums and their compensations) and the 5 results
// of imex_conserved_noalloc.  Each GPU the library runs on holds its own copy
// until the process ends.  In host code this name is not a device address:
// only the address hipGetSymbolAddress returns for the current device is.
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
int scratch_address(double **out)
{
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(IMEX_ENODEV, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_at[dev]) {
        void *p = nullptr;
        HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(IMEX_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_at[dev] = (double *)p;
    }
    *out = g_scratch_at[dev];
    return IMEX_OK;
}

template <class S>
int launch_conserved_box_t(const void *f, const BoxDesc &b, int ntiles, int grid, double *scr, hipStream_t st)
{
    hipLaunchKernelGGL((k_conserved_box<S>), dim3(grid), dim3(RED_BLOCK), 0, st, (const S *)f, b, kNoallocTile,
                       ntiles, scr);
    hipLaunchKernelGGL(k_conserved_final, dim3(1), dim3(RED_BLOCK), 0, st, scr, grid,
                       scr + 10 * IMEX_NOALLOC_PARTIALS);
    HIPC(hipGetLastError());
    return IMEX_OK;
}

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================
extern "C" {

void imex_config_default(imex_config *cfg)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->precision = IMEX_F64;
    cfg->mode = IMEX_WRAP;
    cfg->ntiles = 8;
    cfg->halo_depth = 1;
    cfg->dir_pad = -1;
    cfg->nontemporal = -1;
    cfg->xcd_run = -1;
    cfg->io16 = -1;
    cfg->two_step = -1;
    cfg->two_step_tile = -1;
    cfg->two_step_prefetch = -1;
    cfg->two_step_io = -1;
    cfg->halo_direct = -1;
}

const char *imex_strerror(int s)
{
    switch (s) {
    case IMEX_OK: return "ok";
    case IMEX_EINVAL: return "invalid argument";
    case IMEX_ENOMEM: return "out of memory";
    case IMEX_EHIP: return "HIP runtime error";
    case IMEX_ENODEV: return "no GPU";
    case IMEX_ESTATE: return "call out of order";
    case IMEX_ERANGE: return "size out of range";
    default: return "unknown status";
    }
}

const char *imex_last_error(void) { return g_err.c_str(); }

int imex_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void imex_destroy(imex_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (void *p : {(void *)c->d_boxes, (void *)c->d_tiles, (void *)c->d_segs, (void *)c->d_pack,
                    (void *)c->d_unpack, (void *)c->d_tabs, (void *)c->d_partial, (void *)c->d_out,
                    (void *)c->d_stage, (void *)c->d_pass, (void *)c->d_direct, (void *)c->d_direct_nox,
                    (void *)c->d_unpack_pass, (void *)c->d_direct_pass, (void *)c->d_direct_nox_pass, c->pool[0],
                    c->pool[1], c->xbuf})
        if (p) (void)hipFree(p);
    if (c->h_out) (void)hipHostFree(c->h_out);
    for (hipEvent_t e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->pev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int imex_create(imex_ctx **out, const imex_config *cfg_in, const imex_domain *doms, int ndomains)
{
    if (!out || !doms || ndomains < 1) return fail(IMEX_EINVAL, "null output, null domains or ndomains < 1");
    *out = nullptr;
    imex_config cfg;
    imex_config_default(&cfg);
    if (cfg_in) cfg = *cfg_in;
    if (storage_bytes(cfg.precision) == 0) return fail(IMEX_EINVAL, "unknown precision");
    if (cfg.mode != IMEX_WRAP && cfg.mode != IMEX_TILED) return fail(IMEX_EINVAL, "unknown mode");
    if (cfg.mode == IMEX_TILED && ndomains != 1) return fail(IMEX_EINVAL, "tiled mode takes one domain");
    if (cfg.ntiles <= 0) cfg.ntiles = 8;
    if (cfg.halo_depth <= 0) cfg.halo_depth = 1;
    if (cfg.halo_depth > 3) return fail(IMEX_EINVAL, "halo_depth must be 1..3");
    if (cfg.whole_x != 0 && cfg.whole_x != 1) return fail(IMEX_EINVAL, "whole_x must be 0 or 1");
    if (cfg.block != 0 && (cfg.block % 64 != 0 || cfg.block > 512)) return fail(IMEX_EINVAL, "block must be a multiple of 64, at most 512");
    if (cfg.cells_per_thread < 0 || cfg.cells_per_thread > 16) return fail(IMEX_EINVAL, "cells_per_thread must be 0..16");
    for (int d = 0; d < ndomains; ++d) {
        const imex_domain &D = doms[d];
        if (D.nx < 1 || D.ny < 1 || D.nz < 1) return fail(IMEX_EINVAL, "domain extents must be positive");
        if (!(D.tau0 + 0.5 > 0.0) || !std::isfinite(D.tau0)) return fail(IMEX_EINVAL, "tau0 + 0.5 must be positive");
        if (!std::isfinite(D.rho0) || !std::isfinite(D.u0)) return fail(IMEX_EINVAL, "rho0 and u0 must be finite");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(IMEX_ENODEV, "no HIP device");
    if (cfg.device < 0 || cfg.device >= ndev) return fail(IMEX_EINVAL, "device ordinal out of range");

    imex_ctx *c = new imex_ctx();
    c->cfg = cfg;
    c->device = cfg.device;
    c->prec = cfg.precision;
    c->ssize = storage_bytes(cfg.precision);
    c->tiled = cfg.mode == IMEX_TILED;
    c->whole_x = c->tiled && cfg.whole_x != 0;
    c->halo_depth = cfg.halo_depth;
    c->halo_direct = cfg.halo_direct != 0;
    c->block = cfg.block ? cfg.block : 256;
    c->cpt = cfg.cells_per_thread ? cfg.cells_per_thread : 2;
    c->persistent = std::max(cfg.persistent, 0);
    c->nt = cfg.nontemporal >= 0 ? (cfg.nontemporal ? 1 : 0) : 1;
    // Direction planes of a power-of-two size (256^3, 512^3, ...) start a
    // multiple of a large power of two apart, so one row's lines of all 27
    // directions share L2 sets: the two-step pass of 8- and 4-byte formats
    // then fetched 1.27-1.65x its reads (the one-cell ring re-read from HBM)
    // and 1.06-1.21x with 32 elements between planes.  2-byte formats stay
    // unpadded (1.02-1.04x already; 32 elements would break the planes'
    // 128-byte alignment), and tiled boxes have no power-of-two planes.
    c->dir_pad = cfg.dir_pad >= 0 ? cfg.dir_pad : (c->ssize >= 4 && !c->tiled ? 32 : 0);
    // Launch defaults per storage width, from same-allocation A/B timings at
    // 512^3: 8- and 4-byte formats stream fastest when each XCD owns one
    // contiguous range of tiles (4-byte: runs of 64 tiles), 2-byte formats in
    // plain dispatch order with global 16-bit loads and stores.
    c->xcds = cfg.xcd_groups > 0 ? cfg.xcd_groups : (c->ssize == 2 ? 1 : 8);
    c->xcd_run = cfg.xcd_run >= 0 ? cfg.xcd_run : (c->ssize == 4 ? 64 : 0);
    c->io16 = cfg.io16 >= 0 ? (cfg.io16 ? 1 : 0) : (c->ssize == 2 ? 1 : 0);
    auto bail = [&](int r) { imex_destroy(c); return r; };

    if (hipSetDevice(c->device) != hipSuccess) return bail(fail(IMEX_EHIP, "hipSetDevice failed"));
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && ncu > 0) c->ncu = ncu;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(IMEX_EHIP, "stream creation failed"));
    for (hipEvent_t &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(IMEX_EHIP, "event creation failed"));
    for (hipEvent_t &e : c->pev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(IMEX_EHIP, "event creation failed"));
    c->profile = cfg.profile != 0;

    // instances and their init tables
    int tab = 0;
    for (int d = 0; d < ndomains; ++d) {
        Instance I{};
        I.dom = doms[d];
        I.omega = 1.0 / (I.dom.tau0 + 0.5);   // src/lbm.cpp:133
        const int gl[3] = {I.dom.nx, I.dom.ny, I.dom.nz};
        for (int a = 0; a < 3; ++a) {
            I.tab[a] = tab;
            tab += 3 * gl[a];
        }
        c->inst.push_back(I);
    }
    int r = layout_boxes(c);
    if (r != IMEX_OK) return bail(r);
    build_tiles(c);
    {
        // default shape: the tile and sweep length of least modelled time among
        // those with a workgroup for nine CUs in ten (any allowed one when two
        // steps per pass are requested); without one, the tile covering the
        // fewest cells is reported.  Sweep lengths: the candidates the
        // autotuner tries and the balanced ones, ceil(nz/k), down to 16
        // planes.  Those of the tiles other than 64x16 are capped (FP64 nz/4
        // and FP32 nz/8, at least 64 planes; 2-byte formats nz/8, at least
        // 32): longer ones measured slower at 512^3 than the model says.
        c->ts_tile = cfg.two_step_tile >= 0 ? cfg.two_step_tile : (c->ssize == 8 ? 3 : (c->ssize == 2 ? 4 : 0));
        if (cfg.two_step_tile < 0) {
            double best = 1e30;
            for (int t : {4, 0, 1, 3, 2})
                if (ts_allowed(c, t) && ts_cover(c, t) < best - 1e-9) { best = ts_cover(c, t); c->ts_tile = t; }
        }
        int zdef = 0;
        for (int pass = 0; pass < 2 && zdef == 0; ++pass) {
            if (pass == 1 && cfg.two_step != 1) break;
            double best = 1e300;
            for (int t : {4, 0, 1, 3, 2}) {
                if ((cfg.two_step_tile >= 0 && t != cfg.two_step_tile) || !ts_allowed(c, t)) continue;
                int nz = 1;
                for (const BoxDesc &b : c->boxes) nz = std::max(nz, b.nz);
                std::vector<int> zs = cfg.two_step_zseg > 0 ? std::vector<int>{cfg.two_step_zseg}
                                                            : ts_zseg_candidates(c, t);
                if (cfg.two_step_zseg <= 0)
                    for (int k = 1; (nz + k - 1) / k >= std::min(16, nz); ++k) {
                        const int z = (nz + k - 1) / k;
                        if (std::find(zs.begin(), zs.end(), z) == zs.end()) zs.push_back(z);
                        if (z == 1) break;
                    }
                const int zcap = t == 4 ? nz
                                        : std::min(nz, c->ssize == 2 ? std::max(32, nz / 8)
                                                                     : std::max(64, nz / (c->ssize == 4 ? 8 : 4)));
                for (int z : zs) {
                    if ((pass == 0 && 10 * ts_grid(c, t, z) < 9LL * c->ncu) || (cfg.two_step_zseg <= 0 && z > zcap))
                        continue;
                    const double m = ts_model(c, t, z);
                    if (m < best - 1e-9) { best = m; c->ts_tile = t; zdef = z; }
                }
            }
        }
        const bool ok = ts_allowed(c, c->ts_tile);
        if (cfg.two_step == 1 && !ok)
            return bail(fail(IMEX_EINVAL, "two_step needs periodic boxes at least half a column tile wide and one "
                                          "tile deep in y, or tiled boxes whose nx and ny the column tile divides, "
                                          "at least 2 planes deep (batches: one launch for all; FP64: 32x8 tiles "
                                          "only)"));
        c->two_step = ok && (cfg.two_step == 1 || (cfg.two_step < 0 && zdef > 0)) ? 1 : 0;
        if (c->two_step && c->tiled) {
            // the second step pulls through the first step's ring: two ghost
            // layers, with every direction of the first one
            c->halo_all = true;
            c->halo_depth = 2;
        }
        c->ts_zseg = cfg.two_step_zseg > 0 ? cfg.two_step_zseg : (zdef > 0 ? zdef : 16);
        c->ts_xcds = cfg.two_step_xcds > 0 ? cfg.two_step_xcds : 8;
        c->ts_pf = ts_pf_for(c, c->ts_tile, cfg.two_step_prefetch);
        c->ts_buf = cfg.two_step_io == 0 && !c->tiled ? 1 : 0;
    }
    add_side_buffers(c);
    if (c->tiled) build_halo(c);

    const size_t pool_bytes = (size_t)c->pool_elems * c->ssize;
    c->contiguous = cfg.contiguous ? 1 : 0;
    for (int k = 0; k < 2; ++k) {
        if (c->contiguous &&
            hipExtMallocWithFlags(&c->pool[k], pool_bytes, hipDeviceMallocContiguous) != hipSuccess) {
            (void)hipGetLastError();
            c->pool[k] = nullptr;
            c->contiguous = 0;
        }
        if (!c->pool[k] && hipMalloc(&c->pool[k], pool_bytes) != hipSuccess)
            return bail(fail(IMEX_ENOMEM, "population pool allocation failed (" + std::to_string(pool_bytes) + " bytes)"));
    }
    c->device_bytes = 2LL * (long long)pool_bytes;
    if (c->tiled) {
        const size_t xb = std::max<long long>(c->xbuf_elems, 1) * (size_t)c->ssize;
        if (hipMalloc(&c->xbuf, xb) != hipSuccess) return bail(fail(IMEX_ENOMEM, "exchange buffer allocation failed"));
        c->device_bytes += (long long)xb;
        const size_t sb = c->segs.size() * sizeof(HaloSeg);
        if (hipMalloc(&c->d_segs, sb) != hipSuccess ||
            hipMalloc(&c->d_pack, std::max<size_t>(c->pack_blocks.size(), 1) * sizeof(int2)) != hipSuccess ||
            hipMalloc(&c->d_unpack, std::max<size_t>(c->unpack_blocks.size(), 1) * sizeof(int2)) != hipSuccess ||
            hipMalloc(&c->d_direct, std::max<size_t>(c->direct_blocks.size(), 1) * sizeof(int2)) != hipSuccess ||
            hipMalloc(&c->d_direct_nox, std::max<size_t>(c->direct_blocks_nox.size(), 1) * sizeof(int2)) != hipSuccess)
            return bail(fail(IMEX_ENOMEM, "halo table allocation failed"));
        if (hipMemcpy(c->d_segs, c->segs.data(), sb, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->d_pack, c->pack_blocks.data(), c->pack_blocks.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->d_unpack, c->unpack_blocks.data(), c->unpack_blocks.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->d_direct, c->direct_blocks.data(), c->direct_blocks.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->d_direct_nox, c->direct_blocks_nox.data(), c->direct_blocks_nox.size() * sizeof(int2),
                      hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(IMEX_EHIP, "halo table upload failed"));
        if (c->side_mode) {
            auto upload = [](const std::vector<int2> &v, int2 **d) {
                return hipMalloc(d, std::max<size_t>(v.size(), 1) * sizeof(int2)) == hipSuccess &&
                       hipMemcpy(*d, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice) == hipSuccess;
            };
            if (!upload(c->unpack_blocks_pass, &c->d_unpack_pass) || !upload(c->direct_blocks_pass, &c->d_direct_pass) ||
                !upload(c->direct_blocks_nox_pass, &c->d_direct_nox_pass))
                return bail(fail(IMEX_EHIP, "halo table upload failed"));
        }
    }
    const size_t bb = c->boxes.size() * sizeof(BoxDesc);
    if (hipMalloc(&c->d_boxes, bb) != hipSuccess || hipMalloc(&c->d_tabs, std::max(tab, 1) * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->d_partial, 1024 * 10 * sizeof(double)) != hipSuccess || hipMalloc(&c->d_out, 8 * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&c->h_out, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return bail(fail(IMEX_ENOMEM, "descriptor allocation failed"));
    if (hipMemcpy(c->d_boxes, c->boxes.data(), bb, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(IMEX_EHIP, "descriptor upload failed"));
    if (hipMalloc(&c->d_pass, c->boxes.size() * sizeof(PassBox)) != hipSuccess)
        return bail(fail(IMEX_ENOMEM, "descriptor allocation failed"));
    if (c->two_step && (r = ts_prepare(c)) != IMEX_OK) return bail(r);
    r = upload_tiles(c);
    if (r != IMEX_OK) return bail(r);

    // Per-axis tables of the analytic field, evaluated with the C library
    // exactly as init_macro does (src/lbm.cpp:81-93).
    std::vector<double> h(std::max(tab, 1));
    for (const Instance &I : c->inst) {
        const int gl[3] = {I.dom.nx, I.dom.ny, I.dom.nz};
        for (int a = 0; a < 3; ++a)
            for (int x = 0; x < gl[a]; ++x) {
                const double xx = (double)x / (double)gl[a] * 2.0 * TGV_PI;
                h[I.tab[a] + x] = std::sin(xx);
                h[I.tab[a] + gl[a] + x] = std::cos(xx);
                h[I.tab[a] + 2 * gl[a] + x] = std::cos(2.0 * xx);
            }
    }
    if (hipMemcpy(c->d_tabs, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(IMEX_EHIP, "table upload failed"));
    *out = c;
    return IMEX_OK;
}

int imex_initialize(imex_ctx *c)
{
    TRY(check_ctx(c));
    drop_graphs(c);
    c->cur = 0;
    TRY(DISPATCH(launch_init_t, c));
    c->initialized = true;
    c->state_S = true;
    c->pending = false;
    c->deferred = 0;
    c->ghosts_valid = c->pghosts_valid = false;
    c->xghost_ok = false;
    HIPC(hipStreamSynchronize(c->stream));
    return IMEX_OK;
}

int imex_collision(imex_ctx *c)
{
    TRY(check_ctx(c));
    if (!c->initialized) return fail(IMEX_ESTATE, "imex_initialize has not been called");
    if (c->pending) TRY(flush_pending(c));
    c->pending = true;
    return IMEX_OK;
}

int imex_streaming(imex_ctx *c)
{
    TRY(check_ctx(c));
    if (!c->initialized) return fail(IMEX_ESTATE, "imex_initialize has not been called");
    if (c->pending) {
        c->pending = false;
        if (c->state_S) {
            TRY(do_collide(c));
        } else if (c->two_step) {
            // fused steps go out in pairs, one pass each; an odd one left
            // over is launched by the next call that needs the state
            if (++c->deferred == 2) {
                c->deferred = 0;
                TRY(do_fused2(c));
            }
        } else {
            TRY(do_fused(c));
        }
    } else if (!c->state_S) {
        TRY(flush_deferred(c));
        TRY(do_stream(c));
    }
    c->state_S = false;
    return IMEX_OK;
}

int imex_step(imex_ctx *c, int nsteps)
{
    TRY(check_ctx(c));
    if (!c->initialized) return fail(IMEX_ESTATE, "imex_initialize has not been called");
    if (nsteps < 0) return fail(IMEX_EINVAL, "nsteps < 0");
    return run_steps(c, nsteps);
}

int imex_conserved(imex_ctx *c, int domain, double out[5])
{
    TRY(check_ctx(c));
    TRY(check_domain(c, domain));
    if (!out) return fail(IMEX_EINVAL, "null output");
    TRY(flush_pending(c));
    const bool gather = !c->state_S;
    if (gather) TRY(ensure_ghosts(c));
    const Instance &I = c->inst[domain];
    PhaseScope ps(c, PH_CONSERVED, (double)Q * c->ssize * I.dom.nx * (double)I.dom.ny * I.dom.nz,
                  FLOP_MOMENTS * I.dom.nx * (double)I.dom.ny * I.dom.nz);
    TRY(DISPATCH(launch_conserved_t, c, I.tile_begin, I.tile_end - I.tile_begin, gather));
    HIPC(hipMemcpyAsync(c->h_out, c->d_out, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    std::memcpy(out, c->h_out, 5 * sizeof(double));
    return IMEX_OK;
}

int imex_conserved_noalloc(const void *f, int precision, int nx, int ny, int nz, long long dstride, double rho0,
                           void *stream, double out[5])
{
    if (!f || !out) return fail(IMEX_EINVAL, "null argument");
    const int ss = storage_bytes(precision);
    if (!ss) return fail(IMEX_EINVAL, "unknown precision");
    if (nx < 1 || ny < 1 || nz < 1) return fail(IMEX_EINVAL, "extents must be positive");
    const long long plane = (long long)nx * ny;
    if (dstride == 0) dstride = plane;
    if (dstride < plane) return fail(IMEX_EINVAL, "direction stride below nx*ny");
    // the limits of the step kernels' addressing: one z-plane (27 direction
    // planes) within 2 GiB, in-plane cell indices within 32 bits
    if (27LL * dstride * ss >= (1LL << 31) || plane * nx >= (1LL << 32))
        return fail(IMEX_ERANGE, "z-plane too large for 32-bit addressing (27 * dstride * sizeof must stay "
                                 "below 2 GiB and nx * ny * nx below 2^32)");
    const long long ntiles = (plane + kNoallocTile - 1) / kNoallocTile * nz;
    if (ntiles > INT_MAX) return fail(IMEX_ERANGE, "more than 2^31 - 1 work tiles");
    double *scr = nullptr;
    TRY(scratch_address(&scr));
    BoxDesc b{};
    b.org = 0;
    b.dstride = dstride;
    b.mag = ((1ULL << 32) + (unsigned long long)nx - 1) / (unsigned long long)nx;
    b.nx = nx; b.ny = ny; b.nz = nz;
    b.sy = nx;
    b.sz = (int)(27LL * dstride);
    b.rho0 = rho0;
    b.xnb[0] = b.xnb[1] = -1;
    // one partial per workgroup: the grid never exceeds the scratch's rows
    const int grid = (int)std::min<long long>(ntiles, IMEX_NOALLOC_PARTIALS);
    hipStream_t st = (hipStream_t)stream;
    switch (precision) {
    case IMEX_F64: TRY(launch_conserved_box_t<double>(f, b, (int)ntiles, grid, scr, st)); break;
    case IMEX_F32: TRY(launch_conserved_box_t<float>(f, b, (int)ntiles, grid, scr, st)); break;
    case IMEX_F16: TRY(launch_conserved_box_t<_Float16>(f, b, (int)ntiles, grid, scr, st)); break;
    default: TRY(launch_conserved_box_t<bf16_t>(f, b, (int)ntiles, grid, scr, st)); break;
    }
    HIPC(hipMemcpyAsync(out, scr + 10 * IMEX_NOALLOC_PARTIALS, 5 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    return IMEX_OK;
}

int imex_get_f(imex_ctx *c, int domain, double *host)
{
    TRY(check_ctx(c));
    TRY(check_domain(c, domain));
    if (!host) return fail(IMEX_EINVAL, "null output");
    TRY(flush_pending(c));
    const bool gather = !c->state_S;
    if (gather) TRY(ensure_ghosts(c));
    const Instance &I = c->inst[domain];
    const long long NX = I.dom.nx, NY = I.dom.ny, NZ = I.dom.nz;
    const long long gcell = NX * NY * NZ;
    std::vector<double> tmp;
    for (int bi = I.box_begin; bi < I.box_end; ++bi) {
        const BoxDesc &b = c->boxes[bi];
        const long long ncell = (long long)b.nx * b.ny * b.nz;
        TRY(ensure_stage(c, 27 * (size_t)ncell));
        TRY(DISPATCH(launch_export_t, c, bi, gather));
        tmp.resize(27 * (size_t)ncell);
        HIPC(hipMemcpyAsync(tmp.data(), c->d_stage, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        for (int ii = 0; ii < Q; ++ii)
            for (int k = 0; k < b.nz; ++k)
                for (int j = 0; j < b.ny; ++j)
                    for (int i = 0; i < b.nx; ++i) {
                        const long long X = b.lo[0] + i, Y = b.lo[1] + j, Z = b.lo[2] + k;
                        host[ii * gcell + (X * NY + Y) * NZ + Z] =
                            tmp[(size_t)ii * ncell + ((size_t)k * b.ny + j) * b.nx + i];
                    }
    }
    return IMEX_OK;
}

int imex_get_macro(imex_ctx *c, int domain, double *p, double *u, double *v, double *w)
{
    TRY(check_ctx(c));
    TRY(check_domain(c, domain));
    if (!p || !u || !v || !w) return fail(IMEX_EINVAL, "null output");
    TRY(flush_pending(c));
    const bool gather = !c->state_S;
    if (gather) TRY(ensure_ghosts(c));
    const Instance &I = c->inst[domain];
    const long long NY = I.dom.ny, NZ = I.dom.nz;
    double *outs[4] = {p, u, v, w};
    // every box's fields into one staging area, so that the macroscopic
    // kernels of one call are timed as one phase
    std::vector<size_t> at;
    size_t total = 0;
    for (int bi = I.box_begin; bi < I.box_end; ++bi) {
        const BoxDesc &b = c->boxes[bi];
        at.push_back(total);
        total += 4 * (size_t)b.nx * b.ny * b.nz;
    }
    TRY(ensure_stage(c, total));
    {
        const double ncell = (double)I.dom.nx * I.dom.ny * I.dom.nz;
        PhaseScope ps(c, PH_MACRO, (Q * c->ssize + 32.0) * ncell, FLOP_MOMENTS * ncell);
        for (int bi = I.box_begin; bi < I.box_end; ++bi)
            TRY(DISPATCH(launch_macro_t, c, bi, gather, c->d_stage + at[bi - I.box_begin]));
    }
    std::vector<double> tmp;
    for (int bi = I.box_begin; bi < I.box_end; ++bi) {
        const BoxDesc &b = c->boxes[bi];
        const long long ncell = (long long)b.nx * b.ny * b.nz;
        tmp.resize(4 * (size_t)ncell);
        HIPC(hipMemcpyAsync(tmp.data(), c->d_stage + at[bi - I.box_begin], tmp.size() * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        for (int m = 0; m < 4; ++m)
            for (int k = 0; k < b.nz; ++k)
                for (int j = 0; j < b.ny; ++j)
                    for (int i = 0; i < b.nx; ++i) {
                        const long long X = b.lo[0] + i, Y = b.lo[1] + j, Z = b.lo[2] + k;
                        outs[m][(X * NY + Y) * NZ + Z] =
                            tmp[(size_t)m * ncell + ((size_t)k * b.ny + j) * b.nx + i];
                    }
    }
    return IMEX_OK;
}

int imex_set_f(imex_ctx *c, int domain, const double *host)
{
    TRY(check_ctx(c));
    TRY(check_domain(c, domain));
    if (!host) return fail(IMEX_EINVAL, "null input");
    TRY(flush_pending(c));
    if (!c->state_S) {
        // materialise S for every instance before overwriting one of them
        TRY(do_stream(c));
        c->state_S = true;
    }
    drop_graphs(c);
    const Instance &I = c->inst[domain];
    const long long NY = I.dom.ny, NZ = I.dom.nz;
    const long long gcell = (long long)I.dom.nx * NY * NZ;
    std::vector<double> tmp;
    for (int bi = I.box_begin; bi < I.box_end; ++bi) {
        const BoxDesc &b = c->boxes[bi];
        const long long ncell = (long long)b.nx * b.ny * b.nz;
        TRY(ensure_stage(c, 27 * (size_t)ncell));
        tmp.resize(27 * (size_t)ncell);
        for (int ii = 0; ii < Q; ++ii)
            for (int k = 0; k < b.nz; ++k)
                for (int j = 0; j < b.ny; ++j)
                    for (int i = 0; i < b.nx; ++i) {
                        const long long X = b.lo[0] + i, Y = b.lo[1] + j, Z = b.lo[2] + k;
                        tmp[(size_t)ii * ncell + ((size_t)k * b.ny + j) * b.nx + i] =
                            host[ii * gcell + (X * NY + Y) * NZ + Z];
                    }
        HIPC(hipMemcpyAsync(c->d_stage, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        TRY(DISPATCH(launch_import_t, c, bi));
        HIPC(hipStreamSynchronize(c->stream));
    }
    c->ghosts_valid = c->pghosts_valid = false;
    c->xghost_ok = false;
    return IMEX_OK;
}

int imex_synchronize(imex_ctx *c)
{
    TRY(check_ctx(c));
    TRY(flush_pending(c));
    HIPC(hipStreamSynchronize(c->stream));
    return IMEX_OK;
}

int imex_timed_steps(imex_ctx *c, int nsteps, double *wall_ms, double *kernel_ms)
{
    TRY(check_ctx(c));
    if (!c->initialized) return fail(IMEX_ESTATE, "imex_initialize has not been called");
    if (nsteps < 0) return fail(IMEX_EINVAL, "nsteps < 0");
    HIPC(hipStreamSynchronize(c->stream));
    const auto t0 = std::chrono::steady_clock::now();
    HIPC(hipEventRecord(c->ev[0], c->stream));
    TRY(run_steps(c, nsteps));
    HIPC(hipEventRecord(c->ev[1], c->stream));
    HIPC(hipEventSynchronize(c->ev[1]));
    const auto t1 = std::chrono::steady_clock::now();
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    if (wall_ms) *wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (kernel_ms) *kernel_ms = (double)ms;
    return IMEX_OK;
}

static int autotune_impl(imex_ctx *c, int reps, double *best_ms)
{
    if (reps < 1) reps = 3;
    // IMEXLBM_AUTOTUNE_FAIL_AT=k stops the tuning with an error after its
    // k-th timed candidate (counting one-step candidates first), so that the
    // failure path can be tested from the host
    const char *fail_env = std::getenv("IMEXLBM_AUTOTUNE_FAIL_AT");
    const long fail_at = fail_env && *fail_env ? std::atol(fail_env) : -1;
    long timed = 0;
    auto stop_here = [&]() { return fail_at >= 0 && ++timed > fail_at; };
    TRY(flush_pending(c));
    TRY(ensure_ghosts(c));
    TRY(ensure_ghosts(c, true));
    HIPC(hipStreamSynchronize(c->stream));
    // Candidates: workgroup size, cells per thread (2 = paired kernel), tile
    // order over the XCDs (dispatch order, one range per XCD, runs of 4 or 64
    // tiles) and the plane access path.  Streaming stores and one tile per
    // workgroup are kept: the other settings never won in sweeps.
    struct Cand { int block, cpt, xcds, run, io16; };
    std::vector<Cand> cands;
    for (int blk : {256, 512})
        for (int cpt : {1, 2})
            for (int io : {0, 1})
                for (int order = 0; order < 4; ++order) {
                    const int x = order == 0 ? 1 : 8;
                    const int run = order <= 1 ? 0 : (order == 2 ? 4 : 64);
                    cands.push_back({blk, cpt, x, run, io});
                }
    c->persistent = 0;
    c->nt = 1;
    auto apply = [&](const Cand &k) -> int {
        c->block = k.block; c->cpt = k.cpt; c->xcds = k.xcds; c->xcd_run = k.run; c->io16 = k.io16;
        build_tiles(c);
        return upload_tiles(c);
    };
    // Timed launches never write the live pool, so the state is untouched.
    // Launches from the live pool into the other one, repeated, would keep a
    // pool of up to ~250 MB in the 256 MB Infinity Cache, which a run does not
    // (it alternates two pools): the timings would be optimistic at those
    // sizes.  Up to 512 MB a temporary third pool, a copy of the live one,
    // is allocated, and the launches alternate between it and the other pool.
    struct TmpPool {
        imex_ctx *c;
        void *p = nullptr;
        ~TmpPool()
        {
            if (!p) return;
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(p);
            c->pool[2] = nullptr;
        }
    } tmp{c};
    {
        const size_t pool_bytes = (size_t)c->pool_elems * c->ssize;
        if (pool_bytes <= (512ull << 20)) {
            if (hipMalloc(&tmp.p, pool_bytes) == hipSuccess &&
                hipMemcpyAsync(tmp.p, c->pool[c->cur], pool_bytes, hipMemcpyDeviceToDevice, c->stream) == hipSuccess)
                c->pool[2] = tmp.p;
            else
                (void)hipGetLastError();
        }
    }
    // source and destination pool of timed launch r (-1: the warm-up one)
    auto route = [&](int r, int &s, int &d) {
        if (!c->pool[2]) { s = c->cur; d = 1 - c->cur; return; }
        const bool a = r < 0 || (r % 2) == 1;
        s = a ? 2 : 1 - c->cur;
        d = a ? 1 - c->cur : 2;
    };
    auto time_it = [&](int n, double &ms_out) -> int {
        int s, d;
        route(-1, s, d);
        TRY(launch_step(c, OP_FUSED, s, d));
        HIPC(hipEventRecord(c->ev[0], c->stream));
        for (int r = 0; r < n; ++r) {
            route(r, s, d);
            TRY(launch_step(c, OP_FUSED, s, d));
        }
        HIPC(hipEventRecord(c->ev[1], c->stream));
        HIPC(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        ms_out = ms / n;
        return IMEX_OK;
    };
    std::vector<std::pair<double, int>> t(cands.size());
    for (size_t k = 0; k < cands.size(); ++k) {
        TRY(apply(cands[k]));
        TRY(time_it(reps, t[k].first));
        if (stop_here()) return fail(IMEX_ESTATE, "autotune stopped by IMEXLBM_AUTOTUNE_FAIL_AT");
        t[k].second = (int)k;
        if (std::getenv("IMEXLBM_AUTOTUNE_LOG"))
            std::fprintf(stderr, "autotune block=%d cpt=%d xcds=%d run=%d io16=%d: %.4f ms\n", cands[k].block,
                         cands[k].cpt, cands[k].xcds, cands[k].run, cands[k].io16, t[k].first);
    }
    // confirmation round: the four fastest again over longer runs, which
    // alone decide (a few launches misorder close shapes)
    std::sort(t.begin(), t.end());
    const size_t top = std::min<size_t>(4, t.size());
    for (size_t k = 0; k < top; ++k) {
        TRY(apply(cands[t[k].second]));
        TRY(time_it(std::max(16, 4 * reps), t[k].first));
        if (std::getenv("IMEXLBM_AUTOTUNE_LOG"))
            std::fprintf(stderr, "autotune confirm block=%d cpt=%d xcds=%d run=%d io16=%d: %.4f ms\n",
                         cands[t[k].second].block, cands[t[k].second].cpt, cands[t[k].second].xcds,
                         cands[t[k].second].run, cands[t[k].second].io16, t[k].first);
    }
    std::sort(t.begin(), t.begin() + top);
    TRY(apply(cands[t[0].second]));
    double best = t[0].first;
    // the chosen shape as a persistent launch: that many workgroups stride
    // over the work-tile table (a batch of uneven instances leaves a ragged
    // last wave of workgroups that a persistent grid does not)
    {
        const int keep_p = c->persistent;
        int best_p = 0;
        for (int per_cu : {2, 4, 8}) {
            c->persistent = per_cu * c->ncu;
            if (c->persistent >= (int)c->tiles.size()) continue;
            double ms = 0;
            TRY(time_it(std::max(16, 4 * reps), ms));
            if (std::getenv("IMEXLBM_AUTOTUNE_LOG"))
                std::fprintf(stderr, "autotune persistent=%d: %.4f ms\n", c->persistent, ms);
            if (ms < best) { best = ms; best_p = c->persistent; }
        }
        c->persistent = best_p ? best_p : keep_p;
    }

    // Two steps per pass: column tile x z-planes per sweep x tile order,
    // timed per pass (two steps) against two single-step launches.
    // (a tiled context only when its ghost tables were built for two steps per
    // pass: two layers, every direction of the first)
    if (c->cfg.two_step != 0 && (!c->tiled || c->halo_all)) {
        struct Cand2 { int tile, zseg, xcds, pf, buf; };
        std::vector<Cand2> c2;
        const int want_pf = c->cfg.two_step_prefetch;
        const int buf_max = (!c->tiled && c->cfg.two_step_io != 1) ? 1 : 0;
        const int buf_min = (!c->tiled && c->cfg.two_step_io == 0) ? 1 : 0;
        for (int tile : {0, 1, 2, 3, 4}) {
            if (!ts_allowed(c, tile)) continue;
            std::vector<int> zs = ts_zseg_candidates(c, tile);
            // the shape in use (the modelled default or an earlier pick) is a candidate too
            if (tile == c->ts_tile && std::find(zs.begin(), zs.end(), c->ts_zseg) == zs.end()) zs.push_back(c->ts_zseg);
            // prefetch modes: the requested one, or all that the tile builds
            std::vector<int> pfs;
            if (want_pf >= 0) pfs.push_back(ts_pf_for(c, tile, want_pf));
            else for (int pf = 0; pf <= ts_pf_for(c, tile, -1); ++pf) pfs.push_back(pf);
            for (int z : zs)
                for (int x : {1, 8})
                    for (int pf : pfs)
                        for (int buf = buf_min; buf <= buf_max; ++buf) c2.push_back({tile, z, x, pf, buf});
        }
        const int keep[6] = {c->two_step, c->ts_tile, c->ts_zseg, c->ts_xcds, c->ts_pf, c->ts_buf};
        // time candidate k over n passes (after one untimed pass), per step
        auto time2 = [&](size_t k, int n, double &per_step) -> int {
            c->ts_tile = c2[k].tile; c->ts_zseg = c2[k].zseg; c->ts_xcds = c2[k].xcds; c->ts_pf = c2[k].pf;
            c->ts_buf = c2[k].buf;
            int s, d;
            route(-1, s, d);
            TRY(launch_step2(c, s, d));
            HIPC(hipEventRecord(c->ev[0], c->stream));
            for (int r = 0; r < n; ++r) {
                route(r, s, d);
                TRY(launch_step2(c, s, d));
            }
            HIPC(hipEventRecord(c->ev[1], c->stream));
            HIPC(hipEventSynchronize(c->ev[1]));
            float ms = 0.f;
            HIPC(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            per_step = ms / n / 2.0;
            if (std::getenv("IMEXLBM_AUTOTUNE_LOG"))
                std::fprintf(stderr, "autotune two-step tile=%d zseg=%d xcds=%d pf=%d buf=%d passes=%d: %.4f ms per step\n",
                             c2[k].tile, c2[k].zseg, c2[k].xcds, c2[k].pf, c2[k].buf, n, per_step);
            return IMEX_OK;
        };
        std::vector<std::pair<double, int>> t2(c2.size());
        for (size_t k = 0; k < c2.size(); ++k) {
            TRY(time2(k, reps, t2[k].first));
            if (stop_here()) return fail(IMEX_ESTATE, "autotune stopped by IMEXLBM_AUTOTUNE_FAIL_AT");
            t2[k].second = (int)k;
        }
        // confirmation: the four fastest again over longer runs, which alone
        // decide.  A few passes misorder close candidates: at 512^3 FP16 the
        // buffer-descriptor path measured 2% faster over 3 passes and runs 4%
        // slower over 40.
        std::sort(t2.begin(), t2.end());
        const size_t top2 = std::min<size_t>(4, t2.size());
        for (size_t k = 0; k < top2; ++k) TRY(time2(t2[k].second, std::max(16, 4 * reps), t2[k].first));
        std::sort(t2.begin(), t2.begin() + top2);
        const double best2 = t2.empty() ? 1e300 : t2[0].first;
        const int pick = t2.empty() ? -1 : t2[0].second;
        if (pick >= 0 && (best2 < best || c->cfg.two_step == 1)) {
            // faster, or requested: keep it with its fastest shape
            c->two_step = 1;
            c->ts_tile = c2[pick].tile; c->ts_zseg = c2[pick].zseg; c->ts_xcds = c2[pick].xcds;
            c->ts_pf = c2[pick].pf;
            c->ts_buf = c2[pick].buf;
            best = std::min(best, best2);
        } else {
            c->two_step = pick >= 0 ? 0 : keep[0];
            c->ts_tile = keep[1]; c->ts_zseg = keep[2]; c->ts_xcds = keep[3]; c->ts_pf = keep[4];
            c->ts_buf = keep[5];
        }
    }
    drop_graphs(c);
    if (best_ms) *best_ms = best;
    return IMEX_OK;
}

int imex_autotune(imex_ctx *c, int reps, double *best_ms)
{
    TRY(check_ctx(c));
    if (!c->initialized) return fail(IMEX_ESTATE, "imex_initialize has not been called");
    // the launch shape in use, restored if tuning fails part way: its work
    // tiles and two-step table are rebuilt for it, and captured graphs (which
    // may point at tables a candidate replaced) are dropped
    const int keep[13] = {c->block, c->cpt, c->xcds, c->xcd_run, c->io16, c->persistent, c->nt,
                          c->two_step, c->ts_tile, c->ts_zseg, c->ts_xcds, c->ts_pf, c->ts_buf};
    const int r = autotune_impl(c, reps, best_ms);
    if (r == IMEX_OK) return IMEX_OK;
    const std::string why = g_err;
    c->block = keep[0]; c->cpt = keep[1]; c->xcds = keep[2]; c->xcd_run = keep[3]; c->io16 = keep[4];
    c->persistent = keep[5]; c->nt = keep[6]; c->two_step = keep[7]; c->ts_tile = keep[8];
    c->ts_zseg = keep[9]; c->ts_xcds = keep[10]; c->ts_pf = keep[11]; c->ts_buf = keep[12];
    drop_graphs(c);
    (void)hipStreamSynchronize(c->stream);
    build_tiles(c);
    const int u = upload_tiles(c);
    c->pass_key[0] = c->pass_key[1] = -1;   // the two-step table is rebuilt on its next use
    if (u != IMEX_OK) return u;
    return fail(r, why + " (launch shape restored)");
}

int imex_copy_probe(int device, long long bytes, int reps, double *gbs)
{
    if (bytes < (1 << 20) || reps < 1 || !gbs) return fail(IMEX_EINVAL, "copy probe needs >= 1 MiB, reps >= 1");
    HIPC(hipSetDevice(device));
    const long long n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return fail(IMEX_ENOMEM, "copy probe allocation failed");
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return fail(IMEX_ENOMEM, "copy probe allocation failed");
    }
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int r = IMEX_OK;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        r = fail(IMEX_EHIP, "copy probe stream/event creation failed");
    // best of four variants: one or four 16-byte loads per lane (one 4 KiB or
    // 16 KiB run per workgroup), plain or streaming stores
    double best = 0.0;
    for (int v = 0; v < 4 && r == IMEX_OK; ++v) {
        const int u = v < 2 ? 1 : 4;
        const long long grid = (n + 256LL * u - 1) / (256LL * u);
        auto launch = [&]() {
            const u32x4 *sa = (const u32x4 *)a;
            u32x4 *db = (u32x4 *)b;
            switch (v) {
            case 0: hipLaunchKernelGGL((k_copy<false, 1>), dim3(grid), dim3(256), 0, s, sa, db, n); break;
            case 1: hipLaunchKernelGGL((k_copy<true, 1>), dim3(grid), dim3(256), 0, s, sa, db, n); break;
            case 2: hipLaunchKernelGGL((k_copy<false, 4>), dim3(grid), dim3(256), 0, s, sa, db, n); break;
            default: hipLaunchKernelGGL((k_copy<true, 4>), dim3(grid), dim3(256), 0, s, sa, db, n); break;
            }
        };
        launch();
        (void)hipEventRecord(e0, s);
        for (int k = 0; k < reps; ++k) launch();
        (void)hipEventRecord(e1, s);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            r = fail(IMEX_EHIP, "copy probe timing failed");
        else
            best = std::max(best, 2.0 * (double)n * 16.0 * reps / (ms * 1e-3) / 1e9);
    }
    if (r == IMEX_OK) *gbs = best;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return r;
}

const char *imex_profile_phase_name(int phase)
{
    return (phase >= 0 && phase < PH_N) ? kPhaseNames[phase] : nullptr;
}

int imex_profile_reset(imex_ctx *c)
{
    if (!c) return fail(IMEX_EINVAL, "null context");
    for (PhaseStat &p : c->ph) p = PhaseStat();
    return IMEX_OK;
}

int imex_profile_read(imex_ctx *c, int phase, double out[5], long long *calls)
{
    if (!c || !out || !calls) return fail(IMEX_EINVAL, "null argument");
    if (phase < 0 || phase >= PH_N) return fail(IMEX_EINVAL, "phase out of range");
    const PhaseStat &p = c->ph[phase];
    out[0] = p.ms;
    out[1] = p.mn;
    out[2] = p.mx;
    out[3] = p.bytes;
    out[4] = p.flops;
    *calls = p.calls;
    return IMEX_OK;
}

int imex_query(imex_ctx *c, imex_info *info)
{
    if (!c || !info) return fail(IMEX_EINVAL, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->cells = c->cells;
    info->storage_bytes = c->ssize;
    info->bytes_per_update = 2LL * Q * c->ssize;
    info->device_bytes = c->device_bytes;
    info->ndomains = (int)c->inst.size();
    info->ntiles = (int)c->boxes.size();
    info->work_tiles = (int)c->tiles.size();
    info->block = c->block;
    info->cells_per_thread = c->cpt;
    info->persistent = c->persistent;
    info->nontemporal = c->nt;
    info->dir_pad = c->dir_pad;
    info->use_graph = c->cfg.use_graph;
    info->xcd_groups = c->xcds;
    info->contiguous = c->contiguous;
    info->io16 = c->io16;
    info->xcd_run = c->xcd_run;
    info->two_step = c->two_step;
    info->two_step_tile = c->ts_tile;
    info->two_step_zseg = c->ts_zseg;
    info->two_step_xcds = c->ts_xcds;
    info->two_step_prefetch = c->ts_pf;
    info->two_step_io = c->ts_buf ? 0 : 1;
    info->halo_direct = c->tiled ? (c->halo_direct ? 1 : 0) : 0;
    info->whole_x = c->whole_x ? 1 : 0;
    info->halo_send_bytes = c->tiled ? c->pack_bytes / 2 : 0;
    info->halo_depth = c->tiled ? c->halo_depth : 0;
    info->grid = c->tiled ? 1 : 0;
    info->rx = c->rgrid[0]; info->ry = c->rgrid[1]; info->rz = c->rgrid[2];
    return IMEX_OK;
}

}  // extern "C"
