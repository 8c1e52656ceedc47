// kernels.hpp -- gfx950 kernels of the IMEXLBM hot path.
//
// Data layout in HBM: structure of arrays with x stride-1, grouped by
// z-plane: f[z][ii][y][x].  Each z-plane holds the 27 direction sub-planes
// back to back (dstride elements apart, sz = 27*dstride between z-planes), so
// everything one z-plane of cells touches lies in one window that a single
// buffer descriptor can address with 32-bit offsets.  A "box" is one stored
// lattice block: a whole periodic domain (no ghost cells, periodic wrap in the
// index math) or one tile of a tiled domain (three ghost layers, rows padded
// so interior rows start on a 128-byte boundary).  Two pools A and B with
// identical layout hold the ping-pong state.
//
// Work is described by a table of work tiles, each a run of up to
// block*cells_per_thread consecutive cells of one z-plane of one box.  Every
// step kernel walks that table, so a single periodic domain, the tiles of a
// tiled domain and a batch of independent instances of uneven size all run as
// one launch per step.  Wavefronts own 64 consecutive x cells, so each of the
// 27 loads and 27 stores of a wavefront is one contiguous 64-element run.
#pragma once
#include <type_traits>
#include <utility>

#include "d3q27.hpp"

namespace imex {

struct BoxDesc {
    long long org;      // element offset of interior cell (0,0,0), direction 0
    long long dstride;  // elements between direction planes
    unsigned long long mag;  // ceil(2^32 / nx): row = (p * mag) >> 32
    int nx, ny, nz;     // interior extents
    int sy, sz;         // element strides of y rows and z planes
    int inst;           // index of the instance (logical domain)
    int lo[3];          // offset of this box inside its instance
    int gl[3];          // instance extents
    int tab[3];         // offsets of the instance's per-axis init tables
    double omega;       // 1/(tau0+0.5)
    double rho0, u0;
    float omega_f;
    int org_xy;         // in-plane offset of interior cell (0,0) from the row-0 start
    long long base;     // element offset of the box allocation (ghosts included)
    long long zorg;     // offset of interior z-plane 0 from base (3*sz with ghosts)
    int wrap_x;         // box with ghosts in y and z only: x wraps in the kernels
    int xnb[2];         // tiled two-step pass: boxes whose two x ghost columns this
                        // box's first / last two columns fill (the lower / upper
                        // x neighbour), written by the pass itself into their
                        // side buffers; -1: none
    long long xside[2]; // tiled two-step pass: element offsets of this box's
                        // lower / upper x ghost columns in their side buffers
                        // (side_index below); -1: the ghosts live in the rows
};

// Side buffer of the two x ghost columns on one side of a tiled box, read by
// the two-step pass: [z][direction][y][column], ghost planes and rows
// included (z and y from -3), so that the ghosts of a whole plane and
// direction share a few 128-byte lines instead of one line per row.  Column
// 0 and 1 are x = -2, -1 on the lower side and x = nx, nx + 1 on the upper.
__host__ __device__ inline long long side_offset(int ny, int col, int y, int z, int ii)
{
    return (((long long)(z + 3) * Q + ii) * (ny + 6) + (y + 3)) * 2 + col;
}

__host__ __device__ inline long long side_index(const BoxDesc &d, int side, int col, int y, int z, int ii)
{
    return d.xside[side] + side_offset(d.ny, col, y, z, ii);
}

__host__ __device__ inline long long side_elems(int ny, int nz) { return (long long)(nz + 6) * Q * (ny + 6) * 2; }

struct WorkTile { int box, z, p0, n; };

// One halo segment: a box of cells copied for a subset of directions, either
// from a tile into the exchange buffer (pack) or back into ghost cells
// (unpack).  On the tile side the direction plane is ii; in the buffer the
// selected directions are numbered 0..ndir-1.
struct HaloSeg {
    long long s_off, d_off;     // element offset of the first cell, dir plane 0
    long long s_ds, d_ds;       // direction strides
    int ex, ey, ez, ncell;      // extents
    int s_sy, s_sz, d_sy, d_sz; // row / plane strides
    unsigned mask;              // directions carried (bit ii)
    int pack;                   // 1: tile -> buffer, 0: buffer -> tile,
                                // 2: tile boundary -> neighbour's ghosts
    int partner;                // direct fill: the reverse segment, copied by
                                // the same thread for the same cell (-1: none)
};

enum { OP_COLLIDE = 0, OP_FUSED = 1, OP_STREAM = 2 };

// Offsets, relative to a cell, of the cells at x-1, x+1, y-1, y+1, z-1, z+1.
struct Nbr { int xm, xp, ym, yp; long long zm, zp; };

template <bool WRAP>
__device__ inline Nbr neighbours(const BoxDesc &d, int i, int j, int k)
{
    Nbr n;
    if constexpr (WRAP) {
        n.xm = (i == 0) ? d.nx - 1 : -1;
        n.xp = (i == d.nx - 1) ? 1 - d.nx : 1;
        n.ym = (j == 0) ? (d.ny - 1) * d.sy : -d.sy;
        n.yp = (j == d.ny - 1) ? (1 - d.ny) * d.sy : d.sy;
        n.zm = (k == 0) ? (long long)(d.nz - 1) * d.sz : -(long long)d.sz;
        n.zp = (k == d.nz - 1) ? (long long)(1 - d.nz) * d.sz : (long long)d.sz;
    } else {
        n.xm = (d.wrap_x && i == 0) ? d.nx - 1 : -1;
        n.xp = (d.wrap_x && i == d.nx - 1) ? 1 - d.nx : 1;
        n.ym = -d.sy; n.yp = d.sy;
        n.zm = -d.sz; n.zp = d.sz;
    }
    return n;
}

// Pull source of direction ii: the cell at x - e_ii.
template <int ii>
__device__ inline long long pull_delta(const Nbr &n)
{
    long long o = 0;
    if constexpr (EX[ii] == 1) o += n.xm; else if constexpr (EX[ii] == -1) o += n.xp;
    if constexpr (EY[ii] == 1) o += n.ym; else if constexpr (EY[ii] == -1) o += n.yp;
    if constexpr (EZ[ii] == 1) o += n.zm; else if constexpr (EZ[ii] == -1) o += n.zp;
    return o;
}

template <class S, bool GATHER, bool WRAP, int... I>
__device__ inline void load_cell(const S *__restrict__ base, long long ds, int o,
                                 const Nbr &n, typename Store<S>::real (&f)[Q],
                                 std::integer_sequence<int, I...>)
{
    ((f[I] = Store<S>::ld(base + (long long)I * ds + o + (GATHER ? pull_delta<I>(n) : 0))), ...);
}

template <class S, bool NT, int... I>
__device__ inline void store_cell(S *__restrict__ base, long long ds, int o,
                                  const typename Store<S>::real (&f)[Q],
                                  std::integer_sequence<int, I...>)
{
    (Store<S>::template st<NT>(base + (long long)I * ds + o, f[I]), ...);
}

// Buffer-descriptor access used by the step kernel.  A descriptor's base is
// the start of one z-plane of a box; the per-lane offset is the in-plane cell
// (x, y) in bytes and the scalar offset selects the direction sub-plane, so a
// load or store needs no per-lane 64-bit address arithmetic.
typedef __amdgpu_buffer_rsrc_t rsrc_t;

__device__ inline rsrc_t plane_rsrc(const void *p)
{
    // stride 0 (raw), 2 GiB window, gfx9 default dword3
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, 0x7fffffff, 0x00020000);
}

template <class S> struct BufIO;

template <> struct BufIO<double> {
    __device__ static double ld(rsrc_t r, int vo, int so)
    {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    }
    template <bool NT> __device__ static void st(double v, rsrc_t r, int vo, int so)
    {
        using B = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B, v), r, vo, so, NT ? 2 : 0);
    }
};

template <> struct BufIO<float> {
    __device__ static float ld(rsrc_t r, int vo, int so)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
    }
    template <bool NT> __device__ static void st(float v, rsrc_t r, int vo, int so)
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo, so, NT ? 2 : 0);
    }
};

template <> struct BufIO<_Float16> {
    __device__ static float ld(rsrc_t r, int vo, int so)
    {
        return (float)__builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0));
    }
    template <bool NT> __device__ static void st(float v, rsrc_t r, int vo, int so)
    {
        const _Float16 h = (_Float16)v;
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, h), r, vo, so, NT ? 2 : 0);
    }
};

template <> struct BufIO<bf16_t> {
    __device__ static float ld(rsrc_t r, int vo, int so)
    {
        return bf16_to_float(__builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0));
    }
    template <bool NT> __device__ static void st(float v, rsrc_t r, int vo, int so)
    {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (__bf16)v), r, vo, so,
                                              NT ? 2 : 0);
    }
};

// In-plane pull offset (elements) of direction ii: the cell at (x, y) - e.
template <int ii>
__device__ inline int pull_xy(const Nbr &n)
{
    int o = 0;
    if constexpr (EX[ii] == 1) o += n.xm; else if constexpr (EX[ii] == -1) o += n.xp;
    if constexpr (EY[ii] == 1) o += n.ym; else if constexpr (EY[ii] == -1) o += n.yp;
    return o;
}

// The one of collide32.  For FP16 storage in kernels over ghost-bounded
// boxes (!WRAP) it is a 1.0 the compiler cannot fold (an SGPR set by
// s_mov_b32): the collision then reads widened halves through v_fma_mix_f32,
// which drops the 27 widening conversions and about 25 registers per thread.
// Those kernels are register-bound (the 64x16 two-step tile spilled without
// it) and run 10-25% faster at 256^3; the periodic kernels ran about 3%
// slower with it at 512^3 and keep the conversions.  Other formats: the
// literal.
template <class S, bool WRAP> __device__ inline float collide_one()
{
    if constexpr (std::is_same<S, _Float16>::value && !WRAP) {
        float r;
        asm("s_mov_b32 %0, 1.0" : "=s"(r));
        return r;
    } else {
        return 1.0f;
    }
}

template <class S> __device__ inline float widen16(unsigned short b);
template <> __device__ inline float widen16<_Float16>(unsigned short b)
{
    return (float)__builtin_bit_cast(_Float16, b);
}
template <> __device__ inline float widen16<bf16_t>(unsigned short b) { return bf16_to_float(b); }

template <class S> __device__ inline unsigned short narrow16(float v);
template <> __device__ inline unsigned short narrow16<_Float16>(float v)
{
    return __builtin_bit_cast(unsigned short, (_Float16)v);        // round to nearest even
}
template <> __device__ inline unsigned short narrow16<bf16_t>(float v)
{
    return __builtin_bit_cast(unsigned short, (__bf16)v);          // round to nearest even
}

// The planes one work tile touches: the 27 direction sub-planes of source
// z-planes z-1, z, z+1 and of destination z-plane z.  Offsets are bytes: the
// lane's in-plane cell (vo) plus the direction sub-plane (so, wavefront
// uniform).  4- and 8-byte formats use buffer descriptors, which take the
// direction offset as a scalar operand.  Any format can instead use global
// loads and stores with a uniform z-plane base and one 32-bit lane offset
// (the io16 launch knob, "global IO"): which of the two streams faster is
// measured per format, not assumed.
template <class S, bool GLOBAL> struct Planes;

template <class S> struct Planes<S, false> {
    using R = typename Store<S>::real;
    rsrc_t m, c, p, w;
    __device__ Planes(const S *zm, const S *z0, const S *zp, S *zw)
        : m(plane_rsrc(zm)), c(plane_rsrc(z0)), p(plane_rsrc(zp)), w(plane_rsrc(zw)) {}
    // source plane of a direction with z velocity EZV: z - EZV
    template <int EZV> __device__ R ld(int vo, int so) const
    {
        return BufIO<S>::ld(EZV == 0 ? c : (EZV == 1 ? m : p), vo, so);
    }
    template <bool NT> __device__ void st(R v, int vo, int so) const { BufIO<S>::template st<NT>(v, w, vo, so); }
    template <int EZV> __device__ unsigned short raw(int vo, int so) const
    {
        return __builtin_amdgcn_raw_buffer_load_b16(EZV == 0 ? c : (EZV == 1 ? m : p), vo, so, 0);
    }
};

template <class S> struct Planes<S, true> {
    using R = typename Store<S>::real;
    const char *m, *c, *p;
    char *w;
    __device__ Planes(const S *zm, const S *z0, const S *zp, S *zw)
        : m(reinterpret_cast<const char *>(zm)), c(reinterpret_cast<const char *>(z0)),
          p(reinterpret_cast<const char *>(zp)), w(reinterpret_cast<char *>(zw)) {}
    template <int EZV> __device__ const char *src() const { return EZV == 0 ? c : (EZV == 1 ? m : p); }
    template <int EZV> __device__ unsigned short raw(int vo, int so) const
    {
        return *reinterpret_cast<const unsigned short *>(src<EZV>() + (unsigned long long)(unsigned)(so + vo));
    }
    template <int EZV> __device__ R ld(int vo, int so) const
    {
        const char *q = src<EZV>() + (unsigned long long)(unsigned)(so + vo);
        if constexpr (sizeof(S) == 2) return widen16<S>(*reinterpret_cast<const unsigned short *>(q));
        else return *reinterpret_cast<const S *>(q);
    }
    template <bool NT> __device__ void st(R v, int vo, int so) const
    {
        char *q = w + (unsigned long long)(unsigned)(so + vo);
        if constexpr (sizeof(S) == 2) {
            const unsigned short b = narrow16<S>(v);
            if constexpr (NT) __builtin_nontemporal_store(b, reinterpret_cast<unsigned short *>(q));
            else *reinterpret_cast<unsigned short *>(q) = b;
        } else {
            if constexpr (NT) __builtin_nontemporal_store((S)v, reinterpret_cast<S *>(q));
            else *reinterpret_cast<S *>(q) = (S)v;
        }
    }
};

template <class S, bool GATHER, class PL, int... I>
__device__ inline void load_cell_io(const PL &P, int dsb, int xy, const Nbr &n,
                                    typename Store<S>::real (&f)[Q], std::integer_sequence<int, I...>)
{
    ((f[I] = P.template ld<GATHER ? EZ[I] : 0>((xy + (GATHER ? pull_xy<I>(n) : 0)) * (int)sizeof(S), I * dsb)),
     ...);
}

template <class S, bool NT, class PL, int... I>
__device__ inline void store_cell_io(const PL &P, int dsb, int xy,
                                     const typename Store<S>::real (&f)[Q], std::integer_sequence<int, I...>)
{
    (P.template st<NT>(f[I], xy * (int)sizeof(S), I * dsb), ...);
}

// Two cells of a 2-byte format loaded as raw halves packed two to a register
// (low half: first cell, high half: second cell), widened only when needed.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

template <class S, bool GATHER, class PL, int... I>
__device__ inline void load_pair_raw16(const PL &P, int dsb, int xy1, const Nbr &n1, int xy2,
                                       const Nbr &n2, u16x2 (&raw)[Q], std::integer_sequence<int, I...>)
{
    ((raw[I].x = P.template raw<GATHER ? EZ[I] : 0>((xy1 + (GATHER ? pull_xy<I>(n1) : 0)) * 2, I * dsb),
      raw[I].y = P.template raw<GATHER ? EZ[I] : 0>((xy2 + (GATHER ? pull_xy<I>(n2) : 0)) * 2, I * dsb)),
     ...);
}

// ---------------------------------------------------------------------------
// Step kernel.  OP_COLLIDE: collide in place (src == dst).  OP_FUSED: pull
// from src, collide, store to dst.  OP_STREAM: pull only (a pure copy).
// ---------------------------------------------------------------------------
// Workgroups are dealt to the XCDs round-robin (workgroup w runs on XCD
// w % X), and each XCD has its own L2.  xcd_slot renumbers workgroups so that
// each XCD takes runs of `run` consecutive work tiles: the cache line an
// x-shifted direction straddles between two tiles of a run is then read from
// one L2 instead of being fetched twice, while all XCDs still work on nearby
// addresses.  run = 0 gives every XCD one contiguous range; the tail that does
// not fill whole runs keeps dispatch order.
__device__ inline int xcd_slot(int w, int g, int x, int run)
{
    if (x <= 1) return w;
    if (run <= 0) {
        const int q = g / x, rem = g % x, k = w % x;
        return k * q + min(k, rem) + w / x;
    }
    const int full = g / (x * run) * (x * run);
    if (w >= full) return w;
    const int idx = w / x;
    return ((idx / run) * x + w % x) * run + idx % run;
}

// PAIR: each thread takes two cells of its tile, blockDim apart, and issues
// the loads of both before computing either, which doubles the bytes each
// wavefront keeps in flight (used for the 2-byte formats, whose 27 loads per
// cell are otherwise too few to cover HBM latency).
template <class S, bool WRAP, int OP, bool NT, bool PAIR = false, bool G16 = false>
__global__ void __launch_bounds__(512)
k_step(const S *__restrict__ src, S *__restrict__ dst, const BoxDesc *__restrict__ boxes,
       const WorkTile *__restrict__ tiles, int ntiles, int xcds, int xcd_run)
{
    using R = typename Store<S>::real;
    constexpr bool GATHER = (OP != OP_COLLIDE);
    const int r = xcd_slot((int)blockIdx.x, (int)gridDim.x, xcds, xcd_run);
    // one tile per workgroup, or (persistent grid) a contiguous run of tiles
    const int per = ((int)gridDim.x == ntiles) ? 1 : (ntiles + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t_end = min(ntiles, (r + 1) * per);
#pragma clang loop unroll(disable)
    for (int t = r * per; t < t_end; ++t) {
        const WorkTile wt = tiles[t];
        const BoxDesc &d = boxes[wt.box];
        const int z = wt.z;
        int zm = z - 1, zp = z + 1;
        if constexpr (WRAP) {
            if (zm < 0) zm = d.nz - 1;
            if (zp == d.nz) zp = 0;
        }
        const S *sbox = src + d.base + d.zorg;
        const Planes<S, G16> P(sbox + (long long)zm * d.sz, sbox + (long long)z * d.sz, sbox + (long long)zp * d.sz,
                          dst + d.base + d.zorg + (long long)z * d.sz);
        const int dsb = (int)(d.dstride * (long long)sizeof(S));
        if constexpr (PAIR) {
            const int nb = (int)blockDim.x;
            for (int q = (int)threadIdx.x; q < wt.n; q += 2 * nb) {
                // the second cell repeats the first when the tile runs out; it
                // is then stored twice with identical bits by the same lane
                const int q2 = (q + nb < wt.n) ? q + nb : q;
                const int p1 = wt.p0 + q, p2 = wt.p0 + q2;
                const int j1 = (int)(((unsigned long long)(unsigned)p1 * d.mag) >> 32);
                const int j2 = (int)(((unsigned long long)(unsigned)p2 * d.mag) >> 32);
                const int i1 = p1 - j1 * d.nx, i2 = p2 - j2 * d.nx;
                const int xy1 = d.org_xy + i1 + j1 * d.sy, xy2 = d.org_xy + i2 + j2 * d.sy;
                const Nbr n1 = neighbours<WRAP>(d, i1, j1, z), n2 = neighbours<WRAP>(d, i2, j2, z);
                R f1[Q], f2[Q];
                if constexpr (sizeof(S) == 2) {
                    u16x2 raw[Q];
                    load_pair_raw16<S, GATHER>(P, dsb, xy1, n1, xy2, n2, raw, std::make_integer_sequence<int, Q>{});
                    // keep both cells' loads ahead of any arithmetic or store
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int ii = 0; ii < Q; ++ii) {
                        f1[ii] = widen16<S>(raw[ii].x);
                        f2[ii] = widen16<S>(raw[ii].y);
                    }
                } else {
                    load_cell_io<S, GATHER>(P, dsb, xy1, n1, f1, std::make_integer_sequence<int, Q>{});
                    load_cell_io<S, GATHER>(P, dsb, xy2, n2, f2, std::make_integer_sequence<int, Q>{});
                    __builtin_amdgcn_sched_barrier(0);
                }
                if constexpr (OP != OP_STREAM) {
                    if constexpr (Store<S>::shifted) {
                        collide32(f1, d.omega_f, collide_one<S, WRAP>());
                        collide32(f2, d.omega_f, collide_one<S, WRAP>());
                    }
                    else { collide64(f1, d.omega); collide64(f2, d.omega); }
                }
                store_cell_io<S, NT>(P, dsb, xy1, f1, std::make_integer_sequence<int, Q>{});
                store_cell_io<S, NT>(P, dsb, xy2, f2, std::make_integer_sequence<int, Q>{});
            }
            continue;
        }
        for (int q = (int)threadIdx.x; q < wt.n; q += (int)blockDim.x) {
            const int p = wt.p0 + q;
            const int j = (int)(((unsigned long long)(unsigned)p * d.mag) >> 32);
            const int i = p - j * d.nx;
            const int xy = d.org_xy + i + j * d.sy;
            const Nbr n = neighbours<WRAP>(d, i, j, z);
            R f[Q];
            load_cell_io<S, GATHER>(P, dsb, xy, n, f, std::make_integer_sequence<int, Q>{});
            if constexpr (OP != OP_STREAM) {
                if constexpr (Store<S>::shifted) collide32(f, d.omega_f, collide_one<S, WRAP>());
                else collide64(f, d.omega);
            }
            store_cell_io<S, NT>(P, dsb, xy, f, std::make_integer_sequence<int, Q>{});
        }
    }
}

// ---------------------------------------------------------------------------
// Two steps in one pass (temporal blocking) for one periodic box.
//
// A workgroup owns a column of TX x TY cells and sweeps z over [z0, z1).  For
// every z-plane p it first computes the intermediate step A(p) = C(pull(P)) on
// the column plus a one-cell ring (the cells the second step pulls from),
// reading P from HBM, and keeps A(p) in LDS; then the output plane
// B(k) = C(pull(A)) of the plane below, whose pull reads A(k-1), A(k), A(k+1)
// from LDS, goes to HBM.  Each population is read from HBM once and written
// once per two steps; the ring cells (a fraction (TX+2)(TY+2)/(TX*TY) - 1 of
// the column) are collided twice.  A(p) is rounded to the storage format on
// its way into LDS, exactly as a stored step would be, so the result is
// bit-identical to two separate steps.
//
// LDS holds A by direction group: B(k) pulls the 9 directions with e_z = -1
// from A(k+1), the 9 with e_z = 0 from A(k) and the 9 with e_z = +1 from
// A(k-1).  So the e_z = -1 group needs one plane slot, e_z = 0 two and e_z = +1
// three: 54 direction planes of (TX+2)(TY+2) cells instead of 81.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int ez_group(int ii) { return EZ[ii] + 1; }   // 0: -1, 1: 0, 2: +1
__host__ __device__ constexpr int ez_index(int ii)
{
    int k = 0;
    for (int j = 0; j < ii; ++j)
        if (EZ[j] == EZ[ii]) ++k;
    return k;
}

template <class S> struct LdsWord { using type = S; };
template <> struct LdsWord<_Float16> { using type = unsigned short; };
template <> struct LdsWord<bf16_t> { using type = unsigned short; };

template <class S> __device__ inline typename LdsWord<S>::type to_word(typename Store<S>::real v)
{
    if constexpr (sizeof(S) == 2) return narrow16<S>(v);
    else return (S)v;
}
template <class S> __device__ inline typename Store<S>::real from_word(typename LdsWord<S>::type v)
{
    if constexpr (sizeof(S) == 2) return widen16<S>(v);
    else return v;
}

template <class S, int RC, int... I>
__device__ inline void lds_store_cell(typename LdsWord<S>::type *gm, typename LdsWord<S>::type *g0,
                                      typename LdsWord<S>::type *gp, int q,
                                      const typename Store<S>::real (&f)[Q], std::integer_sequence<int, I...>)
{
    (((ez_group(I) == 0 ? gm : (ez_group(I) == 1 ? g0 : gp))[ez_index(I) * RC + q] = to_word<S>(f[I])), ...);
}

// pull of ring cell r: direction ii comes from r - (ex, ey) in its group's slot
template <class S, int RC, int RX, int... I>
__device__ inline void lds_pull_cell(const typename LdsWord<S>::type *gm, const typename LdsWord<S>::type *g0,
                                     const typename LdsWord<S>::type *gp, int r, typename Store<S>::real (&f)[Q],
                                     std::integer_sequence<int, I...>)
{
    ((f[I] = from_word<S>((ez_group(I) == 0 ? gm : (ez_group(I) == 1 ? g0 : gp))[ez_index(I) * RC + r -
                                                                                EY[I] * RX - EX[I]])),
     ...);
}

// PD > 0 (2-byte formats): partial prefetch.  Only the first PD populations
// of each thread's first ring cell of plane p + 2 are loaded ahead, held as
// stored (a 2-byte value widened at once would make the wave wait for its
// load right there).  That fits the registers that a full prefetch of a
// 2-byte pass exceeds, and still takes PD/27 of the first step's load burst
// off the time between the barriers.
//
// PF: prefetch.  The populations of plane p + 2 are loaded into registers
// before the barrier and the second step of plane p, and collided in the next
// iteration, so the HBM latency of the first step hides behind the second
// (barriers wait for LDS traffic only).  Costs 27 registers per ring cell a
// thread owns; used where the workgroup's LDS share leaves few waves per CU.
// Workgroups of a two-step pass per box: the first one, the box and its
// column grid (tiles_x columns across, ntile per z-segment).
struct PassBox { int wg0, box, tiles_x, ntile; };

// WRAP: periodic box, ring and z-planes wrap, and the last column tile may
// run past nx or ny (TX <= 2 nx, TY <= ny): its ring wraps onto real cells,
// twice in x for a box narrower than the tile (the same cell then appears
// twice in the ring, collided identically), and its output cells beyond the
// box are not stored.  Otherwise the tile divides
// the box, and the ring and the planes z0 - 1 and z1 are ghost cells, filled
// two layers deep (with the directions each layer is read for) before the
// pass; a box that keeps x whole (wrap_x) wraps the ring in x instead.
// Pulled populations PD..26 of a cell (the first PD come from a prefetch).
template <class S, int PD, class PL, int... I>
__device__ inline void load_cell_tail(const PL &P, int dsb, int xy, const Nbr &n,
                                      typename Store<S>::real (&f)[Q], std::integer_sequence<int, I...>)
{
    ((I >= PD ? (void)(f[I] = P.template ld<EZ[I]>((xy + pull_xy<I>(n)) * (int)sizeof(S), I * dsb)) : (void)0), ...);
}

// Pulled populations 0..PD-1 of a cell of a 2-byte format, as stored.
template <int PD, class PL, int... I>
__device__ inline void load_cell_head_raw(const PL &P, int dsb, int xy, const Nbr &n, unsigned short (&h)[PD],
                                          std::integer_sequence<int, I...>)
{
    ((h[I] = P.template raw<EZ[I]>((xy + pull_xy<I>(n)) * 2, I * dsb)), ...);
}

// Pulled population ii of a ring cell (x, y) of plane p in the outer two
// columns of a box with side buffers: from the box when its source x - e_x
// is inside, else from the side buffer of that side.  The address is
// selected and loaded once: a branch per population would make each load
// wait for the ones before it.  (Global-pointer planes only.)
template <class S, int ii, class PL>
__device__ inline void load_edge_one(const PL &P, int dsb, int xy, const Nbr &n, const S *pool, const BoxDesc &d,
                                     int x, int y, int p, typename Store<S>::real &v)
{
    const int xs = x - EX[ii];
    const char *box = P.template src<EZ[ii]>() + (unsigned long long)(unsigned)(ii * dsb + (xy + pull_xy<ii>(n)) * (int)sizeof(S));
    const char *sb = reinterpret_cast<const char *>(
        pool + side_index(d, xs < 0 ? 0 : 1, xs < 0 ? xs + 2 : xs - d.nx, y - EY[ii], p - EZ[ii], ii));
    const char *a = (xs >= 0 && xs < d.nx) ? box : sb;
    if constexpr (sizeof(S) == 2) v = widen16<S>(*reinterpret_cast<const unsigned short *>(a));
    else v = *reinterpret_cast<const S *>(a);
}

template <class S, class PL, int... I>
__device__ inline void load_cell_edge(const PL &P, int dsb, int xy, const Nbr &n, const S *pool, const BoxDesc &d,
                                      int x, int y, int p, typename Store<S>::real (&f)[Q],
                                      std::integer_sequence<int, I...>)
{
    (load_edge_one<S, I>(P, dsb, xy, n, pool, d, x, y, p, f[I]), ...);
}

// One cell's populations into a box's side buffer, from wp (the cell's
// direction-0 element; the directions are ds elements apart): the tiled
// two-step pass storing its x neighbours' ghosts.  Ordinary stores: a wave
// writes two lanes per row here, and the rows of many waves share a line,
// which the L2 gathers before writing it back (streaming stores would send
// each few bytes to memory on their own).
template <class S>
__device__ inline void x_ghost_store_side(S *wp, long long ds, const typename Store<S>::real (&f)[Q])
{
#pragma unroll
    for (int ii = 0; ii < Q; ++ii) Store<S>::template st<false>(wp + ii * ds, f[ii]);
}

// Directions of the partial prefetch of 2-byte formats (PD): the first 11
// populations of each thread's first ring cell.  More made no further gain at
// 512^3; every thread owns a ring cell in the tiles that use it.
constexpr int PARTIAL_PD = 11;

template <class S, bool WRAP, int TX, int TY, int NTH, bool NT, bool PF, bool GIO = true, int PD = 0>
__global__ void __launch_bounds__(NTH)
k_step2(const S *__restrict__ src, S *__restrict__ dst, const BoxDesc *__restrict__ boxes,
        const PassBox *__restrict__ pb, int npb, int xcds, int zseg)
{
    using R = typename Store<S>::real;
    using Wd = typename LdsWord<S>::type;
    constexpr int RX = TX + 2, RC = RX * (TY + 2), IC = TX * TY;
    __shared__ Wd lds[54 * RC];
    constexpr auto seq = std::make_integer_sequence<int, Q>{};
    const int w = xcd_slot((int)blockIdx.x, (int)gridDim.x, xcds, 0);
    int bi = 0;
    while (bi + 1 < npb && pb[bi + 1].wg0 <= w) ++bi;
    const PassBox B = pb[bi];
    const BoxDesc &d = boxes[B.box];
    const int tiles_x = B.tiles_x, ntile = B.ntile;
    const int tile = (w - B.wg0) % ntile, seg = (w - B.wg0) / ntile;
    const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
    const int z0 = seg * zseg, z1 = min(d.nz, z0 + zseg);
    const S *sbox = src + d.base + d.zorg;
    S *dbox = dst + d.base + d.zorg;
    const int dsb = (int)(d.dstride * (long long)sizeof(S));

    // ring cell q of plane p: its source planes, in-plane offset and neighbours
    auto ring_at = [&](int p, int q, int &xy, Nbr &n) {
        int z = p, zm = p - 1, zp = p + 1;
        int x = x0 - 1 + q % RX, y = y0 - 1 + q / RX;
        if constexpr (WRAP) {
            z = p < 0 ? p + d.nz : (p >= d.nz ? p - d.nz : p);
            zm = z == 0 ? d.nz - 1 : z - 1;
            zp = z == d.nz - 1 ? 0 : z + 1;
            if (x < 0) x += d.nx; else if (x >= d.nx) x -= d.nx;
            if (d.nx < TX && x >= d.nx) x -= d.nx;   // a box narrower than the tile
            if (y < 0) y += d.ny; else if (y >= d.ny) y -= d.ny;
        } else if (d.wrap_x) {
            if (x < 0) x += d.nx; else if (x >= d.nx) x -= d.nx;
        }
        n = neighbours<WRAP>(d, x, y, z);
        xy = d.org_xy + x + y * d.sy;
        return Planes<S, GIO>(sbox + (long long)zm * d.sz, sbox + (long long)z * d.sz,
                              sbox + (long long)zp * d.sz, dbox);
    };
    // Boxes with x side buffers: in a tile at the box's lower (upper) x edge
    // the ring cells of the first (last) two columns pull some populations
    // from the side buffers.  The ring of such a tile is walked by slots: the
    // other cells first, row by row, then from the next wave boundary the
    // edge-column cells, with idle slots between.  A wave then takes the
    // side-buffer path with all its lanes or none, and never runs both.
    const bool side = !WRAP && d.xside[0] >= 0;
    const int elo = side && x0 == 0, ehi = side && x0 + TX >= d.nx - 1;
    const bool edge_tile = elo || ehi;
    const int nce = 2 * (elo + ehi);             // edge columns of the ring
    const int nec = (RX - nce) * (TY + 2);       // ring cells clear of them
    const int nep = (nec + 63) & ~63;            // slot of the first edge-column cell
    constexpr int RS = WRAP ? RC : std::max((RX * (TY + 2) - 4 * (TY + 2) + 63) / 64 * 64 + 4 * (TY + 2),
                                            ((RX - 2) * (TY + 2) + 63) / 64 * 64 + 2 * (TY + 2));   // ring slots
    constexpr int NC = (RS + NTH - 1) / NTH;     // ring slots per thread
    static_assert(NC == (RC + NTH - 1) / NTH, "the edge-column slots add no slot per thread");
    // ring cell of slot s, -1 for an idle slot
    auto ring_q = [&](int s) -> int {
        if (!edge_tile) return s < RC ? s : -1;
        if (s < nec) {
            const int r = s / (RX - nce);
            return r * RX + 2 * elo + (s - r * (RX - nce));
        }
        const int e = s - nep;
        const int r = nce == 4 ? e >> 2 : e >> 1, k = nce == 4 ? e & 3 : e & 1;
        if (e < 0 || r >= TY + 2) return -1;
        return r * RX + (elo && k < 2 ? k : RX - 2 + k - 2 * elo);
    };
    // loads the pulled populations of ring cell q of plane p into f; the
    // edge-column slots take those whose source lies beyond the box in x
    // from the side buffers
    auto load_ring = [&](int p, int s, int q, R (&f)[Q]) {
        int xy;
        Nbr n;
        const auto P = ring_at(p, q, xy, n);
        if constexpr (WRAP || !GIO) {
            load_cell_io<S, true>(P, dsb, xy, n, f, seq);
        } else {
            if (!edge_tile || s < nep) load_cell_io<S, true>(P, dsb, xy, n, f, seq);
            else load_cell_edge<S>(P, dsb, xy, n, src, d, x0 - 1 + q % RX, y0 - 1 + q / RX, p, f, seq);
        }
    };
    // collide ring cell q of the intermediate plane p and keep it in LDS
    auto put_ring = [&](int p, int q, R (&f)[Q]) {
        Wd *gm = lds, *g0 = lds + (9 + 9 * ((p + 4) & 1)) * RC, *gp = lds + (27 + 9 * ((p + 3) % 3)) * RC;
        if constexpr (Store<S>::shifted) collide32(f, d.omega_f, collide_one<S, WRAP>());
        else collide64(f, d.omega);
        lds_store_cell<S, RC>(gm, g0, gp, q, f, seq);
    };
    // A(p) for p in [z0 - 1, z1]: plane p mod nz of the intermediate step
    auto first = [&](int p) {
        for (int s = (int)threadIdx.x; s < RS; s += NTH) {
            const int q = ring_q(s);
            if (q < 0) continue;
            R f[Q];
            load_ring(p, s, q, f);
            put_ring(p, q, f);
        }
    };
    // The side buffers this box's pass fills: the upper one of its lower x
    // neighbour and the lower one of its upper x neighbour (the host sets
    // xnb only where the neighbour has side buffers; x neighbours share ny).
    // Read once here: inside the sweep these dependent scalar loads stalled
    // every wave of every plane (4-9% of the tiled pass).
    static_assert(NTH >= IC, "the second step gives each thread one output cell at most");
    const int otx = (int)threadIdx.x < IC ? (int)threadIdx.x % TX : -1, oty = (int)threadIdx.x / TX;
    long long side_lo = -1, side_hi = -1;
    if constexpr (!WRAP) {
        if (d.xnb[0] >= 0) side_lo = boxes[d.xnb[0]].xside[1];
        if (d.xnb[1] >= 0) side_hi = boxes[d.xnb[1]].xside[0];
    }
    // B(k) from A(k-1), A(k), A(k+1)
    auto second = [&](int k) {
        const Wd *gm = lds, *g0 = lds + (9 + 9 * ((k + 4) & 1)) * RC, *gp = lds + (27 + 9 * ((k + 2) % 3)) * RC;
        S *wp = dbox + (long long)k * d.sz;
        const Planes<S, GIO> P(wp, wp, wp, wp);
        if (otx >= 0) {
            const int tx = otx, ty = oty;
            if (WRAP && (x0 + tx >= d.nx || y0 + ty >= d.ny)) return;   // past a ragged edge
            R f[Q];
            lds_pull_cell<S, RC, RX>(gm, g0, gp, (ty + 1) * RX + tx + 1, f, seq);
            if constexpr (Store<S>::shifted) collide32(f, d.omega_f, collide_one<S, WRAP>());
            else collide64(f, d.omega);
            store_cell_io<S, NT>(P, dsb, d.org_xy + (x0 + tx) + (y0 + ty) * d.sy, f, seq);
            if constexpr (!WRAP) {
                // the first and last two columns are also the x neighbours'
                // two ghost columns: stored there too, every direction
                const int gx = x0 + tx, gy = y0 + ty;
                const long long ds = 2LL * (d.ny + 6);
                if (gx < 2 && side_lo >= 0) x_ghost_store_side<S>(dst + side_lo + side_offset(d.ny, gx, gy, k, 0), ds, f);
                if (gx >= d.nx - 2 && side_hi >= 0)
                    x_ghost_store_side<S>(dst + side_hi + side_offset(d.ny, gx - d.nx + 2, gy, k, 0), ds, f);
            }
        }
    };
    first(z0 - 1);
    first(z0);
    if constexpr (PD > 0) {
        static_assert(sizeof(S) == 2 && NTH <= RC && !PF, "partial prefetch: 2-byte formats, every thread owns a ring cell");
        // each thread's first ring cell is clear of the outer columns
        static_assert(WRAP || (RX - 4) * (TY + 2) >= NTH, "partial prefetch: the first ring cells take no side-buffer loads");
        const int q1 = ring_q((int)threadIdx.x);
        unsigned short pre[PD];
        auto fetch_head = [&](int p) {
            int xy;
            Nbr n;
            const auto P = ring_at(p, q1, xy, n);
            load_cell_head_raw<PD>(P, dsb, xy, n, pre, std::make_integer_sequence<int, PD>{});
        };
        fetch_head(z0 + 1);
#pragma clang loop unroll(disable)
        for (int k = z0; k < z1; ++k) {
            {
                int xy;
                Nbr n;
                const auto P = ring_at(k + 1, q1, xy, n);
                R f[Q];
                load_cell_tail<S, PD>(P, dsb, xy, n, f, seq);
#pragma unroll
                for (int ii = 0; ii < PD; ++ii) f[ii] = widen16<S>(pre[ii]);
                put_ring(k + 1, q1, f);
            }
            for (int s = (int)threadIdx.x + NTH; s < RS; s += NTH) {
                const int q = ring_q(s);
                if (q < 0) continue;
                R f[Q];
                load_ring(k + 1, s, q, f);
                put_ring(k + 1, q, f);
            }
            if (k + 1 < z1) fetch_head(k + 2);
            __syncthreads();
            second(k);
            __syncthreads();
        }
    } else if constexpr (PF) {
        R pf[NC][Q];
        int qs[NC];   // ring cell of each slot of this thread (-1: idle)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            qs[c] = ring_q((int)threadIdx.x + c * NTH);
            if (qs[c] >= 0) load_ring(z0 + 1, (int)threadIdx.x + c * NTH, qs[c], pf[c]);
        }
#pragma clang loop unroll(disable)
        for (int k = z0; k < z1; ++k) {
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (qs[c] >= 0) put_ring(k + 1, qs[c], pf[c]);
            if (k + 1 < z1) {
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    if (qs[c] >= 0) load_ring(k + 2, (int)threadIdx.x + c * NTH, qs[c], pf[c]);
            }
            __syncthreads();
            second(k);
            __syncthreads();
        }
    } else {
#pragma clang loop unroll(disable)
        for (int k = z0; k < z1; ++k) {
            first(k + 1);
            __syncthreads();
            second(k);
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Initial state (src/lbm.cpp:78-108) from per-axis tables [sin | cos | cos2x]
// evaluated on the host, in FP64, stored as f (FP64) or g = f - w*rho0.
// ---------------------------------------------------------------------------
template <class S>
__global__ void __launch_bounds__(512)
k_init(S *__restrict__ dst, const BoxDesc *__restrict__ boxes, const WorkTile *__restrict__ tiles,
       int ntiles, int /*cells_per_thread: sets the tile size only*/, const double *__restrict__ tabs)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const WorkTile wt = tiles[t];
        const BoxDesc &d = boxes[wt.box];
        S *__restrict__ dp = dst + d.org + (long long)wt.z * d.sz;
        for (int q = (int)threadIdx.x; q < wt.n; q += (int)blockDim.x) {
            const int p = wt.p0 + q;
            const int j = (int)(((unsigned long long)(unsigned)p * d.mag) >> 32);
            const int i = p - j * d.nx;
            const int X = d.lo[0] + i, Y = d.lo[1] + j, Z = d.lo[2] + wt.z;
            const double *tx = tabs + d.tab[0], *ty = tabs + d.tab[1], *tz = tabs + d.tab[2];
            const double sx = tx[X], cx = tx[d.gl[0] + X], c2x = tx[2 * d.gl[0] + X];
            const double sy = ty[Y], cy = ty[d.gl[1] + Y], c2y = ty[2 * d.gl[1] + Y];
            const double cz = tz[d.gl[2] + Z], c2z = tz[2 * d.gl[2] + Z];
            const double u0 = d.u0, rho0 = d.rho0;
            const double cs2 = 1.0 / 3.0;
            const double u = u0 * sx * cy * cz;
            const double v = -u0 * cx * sy * cz;
            const double pr = rho0 * cs2 + rho0 * u0 * u0 / 16.0 * (c2x + c2y) * (c2z + 2.0);
            double f[Q];
            equilibrium64(pr, u, v, f);
            const int o = i + j * d.sy;
#pragma unroll
            for (int ii = 0; ii < Q; ++ii) {
                if constexpr (Store<S>::shifted) {
                    const double g = f[ii] - weight(ii) * rho0;
                    Store<S>::template st<false>(dp + (long long)ii * d.dstride + o, (float)g);
                } else {
                    Store<S>::template st<false>(dp + (long long)ii * d.dstride + o, f[ii]);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Compensated FP64 reduction of {mass, px, py, pz, ke} (src/lbm.cpp:271-294)
// over a range of work tiles.  Each thread keeps Neumaier sums; workgroups
// merge them in a fixed tree, so the result does not depend on timing.
// GATHER reads S_n = stream(P_{n-1}) through the pull offsets.
// ---------------------------------------------------------------------------
struct Acc { double s, c; };

__device__ inline void acc_add(Acc &a, double v)
{
    const double t = a.s + v;
    if (fabs(a.s) >= fabs(v)) a.c += (a.s - t) + v; else a.c += (v - t) + a.s;
    a.s = t;
}
__device__ inline void acc_merge(Acc &a, const Acc &b)
{
    acc_add(a, b.s);
    a.c += b.c;
}

constexpr int RED_BLOCK = 256;

// One work tile's cells into the thread's five Neumaier sums.
template <class S, bool WRAP, bool GATHER>
__device__ __forceinline__ void conserved_tile(const S *__restrict__ src, const BoxDesc &d, const WorkTile &wt,
                                               Acc acc[5])
{
    using R = typename Store<S>::real;
    const S *__restrict__ sp = src + d.org + (long long)wt.z * d.sz;
    for (int q = (int)threadIdx.x; q < wt.n; q += (int)blockDim.x) {
        const int p = wt.p0 + q;
        const int j = (int)(((unsigned long long)(unsigned)p * d.mag) >> 32);
        const int i = p - j * d.nx;
        const Nbr n = neighbours<WRAP>(d, i, j, wt.z);
        R g[Q];
        load_cell<S, GATHER, WRAP>(sp, d.dstride, i + j * d.sy, n, g, std::make_integer_sequence<int, Q>{});
        double f[Q];
#pragma unroll
        for (int ii = 0; ii < Q; ++ii) {
            if constexpr (Store<S>::shifted) f[ii] = (double)g[ii] + weight(ii) * d.rho0;
            else f[ii] = (double)g[ii];
        }
        const Moments64 m = moments64(f);
        acc_add(acc[0], m.pl);
        acc_add(acc[1], m.ul);
        acc_add(acc[2], m.vl);
        acc_add(acc[3], m.wl);
        acc_add(acc[4], 0.5 * (m.ul * m.ul + m.vl * m.vl + m.wl * m.wl));
    }
}

// The workgroup's threads merged in a fixed tree; one partial (5 sums and
// their compensations) per workgroup.
__device__ __forceinline__ void conserved_partial(const Acc acc[5], double *__restrict__ partial)
{
    __shared__ double sh[2 * 5][RED_BLOCK];
    for (int k = 0; k < 5; ++k) {
        sh[2 * k][threadIdx.x] = acc[k].s;
        sh[2 * k + 1][threadIdx.x] = acc[k].c;
    }
    __syncthreads();
    for (int w = RED_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 5; ++k) {
                Acc a = {sh[2 * k][threadIdx.x], sh[2 * k + 1][threadIdx.x]};
                const Acc b = {sh[2 * k][threadIdx.x + w], sh[2 * k + 1][threadIdx.x + w]};
                acc_merge(a, b);
                sh[2 * k][threadIdx.x] = a.s;
                sh[2 * k + 1][threadIdx.x] = a.c;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 10) partial[blockIdx.x * 10 + threadIdx.x] = sh[threadIdx.x][0];
}

template <class S, bool WRAP, bool GATHER>
__global__ void __launch_bounds__(RED_BLOCK)
k_conserved(const S *__restrict__ src, const BoxDesc *__restrict__ boxes,
            const WorkTile *__restrict__ tiles, int ntiles, int /*cells_per_thread: sets the tile size only*/, double *__restrict__ partial)
{
    Acc acc[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const WorkTile wt = tiles[t];
        conserved_tile<S, WRAP, GATHER>(src, boxes[wt.box], wt, acc);
    }
    conserved_partial(acc, partial);
}

// The same sums over one periodic box of post-streaming populations whose
// descriptor comes by value, with the work tiles of a single-domain context
// (runs of `chunk` cells of one z-plane, in plane order) computed instead of
// read from a table: the same cells in the same order as k_conserved over
// that context's table, so the same bits.
template <class S>
__global__ void __launch_bounds__(RED_BLOCK)
k_conserved_box(const S *__restrict__ src, const BoxDesc box, int chunk, int ntiles, double *__restrict__ partial)
{
    Acc acc[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const int plane = box.nx * box.ny;
    const int per_plane = (plane + chunk - 1) / chunk;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        WorkTile wt;
        wt.box = 0;
        wt.z = t / per_plane;
        wt.p0 = (t - wt.z * per_plane) * chunk;
        wt.n = min(chunk, plane - wt.p0);
        conserved_tile<S, true, false>(src, box, wt, acc);
    }
    conserved_partial(acc, partial);
}

// Second stage: one workgroup folds the per-workgroup partials, in order.
__global__ void __launch_bounds__(RED_BLOCK)
k_conserved_final(const double *__restrict__ partial, int nparts, double *__restrict__ out)
{
    Acc acc[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int b = threadIdx.x; b < nparts; b += RED_BLOCK)
        for (int k = 0; k < 5; ++k) {
            const Acc v = {partial[b * 10 + 2 * k], partial[b * 10 + 2 * k + 1]};
            acc_merge(acc[k], v);
        }
    __shared__ double sh[10][RED_BLOCK];
    for (int k = 0; k < 5; ++k) {
        sh[2 * k][threadIdx.x] = acc[k].s;
        sh[2 * k + 1][threadIdx.x] = acc[k].c;
    }
    __syncthreads();
    for (int w = RED_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 5; ++k) {
                Acc a = {sh[2 * k][threadIdx.x], sh[2 * k + 1][threadIdx.x]};
                const Acc b = {sh[2 * k][threadIdx.x + w], sh[2 * k + 1][threadIdx.x + w]};
                acc_merge(a, b);
                sh[2 * k][threadIdx.x] = a.s;
                sh[2 * k + 1][threadIdx.x] = a.c;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) out[threadIdx.x] = sh[2 * threadIdx.x][0] + sh[2 * threadIdx.x + 1][0];
}

// ---------------------------------------------------------------------------
// Export / import of one box as doubles f in [ii][z][y][x] order (no ghosts).
// Export with GATHER materialises S_n from the post-collision state.
// ---------------------------------------------------------------------------
template <class S, bool WRAP, bool GATHER>
__global__ void __launch_bounds__(256)
k_export(const S *__restrict__ src, const BoxDesc *__restrict__ boxes, int box,
         double *__restrict__ out)
{
    using R = typename Store<S>::real;
    const BoxDesc &d = boxes[box];
    const long long ncell = (long long)d.nx * d.ny * d.nz;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % d.nx);
        const int j = (int)((c / d.nx) % d.ny);
        const int k = (int)(c / ((long long)d.nx * d.ny));
        const Nbr n = neighbours<WRAP>(d, i, j, k);
        R g[Q];
        load_cell<S, GATHER, WRAP>(src + d.org + (long long)k * d.sz, d.dstride, i + j * d.sy, n, g,
                                   std::make_integer_sequence<int, Q>{});
#pragma unroll
        for (int ii = 0; ii < Q; ++ii) {
            double f = (double)g[ii];
            if constexpr (Store<S>::shifted) f += weight(ii) * d.rho0;
            out[ii * ncell + c] = f;
        }
    }
}

// Macroscopic fields of S_n (ComputeMacroscopic, src/lbm.cpp:218-236):
// p = pl*(1/3), u = ul, v = vl, w = wl from the FP64 moments of the cell, as
// four planes [field][z][y][x] of doubles.
template <class S, bool WRAP, bool GATHER>
__global__ void __launch_bounds__(256)
k_macro(const S *__restrict__ src, const BoxDesc *__restrict__ boxes, int box,
        double *__restrict__ out)
{
    using R = typename Store<S>::real;
    const BoxDesc &d = boxes[box];
    const long long ncell = (long long)d.nx * d.ny * d.nz;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % d.nx);
        const int j = (int)((c / d.nx) % d.ny);
        const int k = (int)(c / ((long long)d.nx * d.ny));
        const Nbr n = neighbours<WRAP>(d, i, j, k);
        R g[Q];
        load_cell<S, GATHER, WRAP>(src + d.org + (long long)k * d.sz, d.dstride, i + j * d.sy, n, g,
                                   std::make_integer_sequence<int, Q>{});
        double f[Q];
#pragma unroll
        for (int ii = 0; ii < Q; ++ii) {
            f[ii] = (double)g[ii];
            if constexpr (Store<S>::shifted) f[ii] += weight(ii) * d.rho0;
        }
        const Moments64 m = moments64(f);
        out[c] = m.pl * (1.0 / 3.0);
        out[ncell + c] = m.ul;
        out[2 * ncell + c] = m.vl;
        out[3 * ncell + c] = m.wl;
    }
}

template <class S>
__global__ void __launch_bounds__(256)
k_import(S *__restrict__ dst, const BoxDesc *__restrict__ boxes, int box,
         const double *__restrict__ in)
{
    const BoxDesc &d = boxes[box];
    const long long ncell = (long long)d.nx * d.ny * d.nz;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % d.nx);
        const int j = (int)((c / d.nx) % d.ny);
        const int k = (int)(c / ((long long)d.nx * d.ny));
        S *p = dst + d.org + (long long)k * d.sz + i + (long long)j * d.sy;
        for (int ii = 0; ii < Q; ++ii) {
            const double f = in[ii * ncell + c];
            if constexpr (Store<S>::shifted)
                Store<S>::template st<false>(p + ii * d.dstride, (float)(f - weight(ii) * d.rho0));
            else
                Store<S>::template st<false>(p + ii * d.dstride, f);
        }
    }
}

// ---------------------------------------------------------------------------
// Halo pack / unpack over every segment of every tile in one launch, or the
// direct fill (pack == 2: a tile's boundary slab straight into the
// neighbour's ghost slab of the same pool; the slabs read are interior cells,
// those written ghost cells, so no cell is both).  Each workgroup copies up
// to blockDim cells of one segment for all of its directions, loading every
// direction before storing any.  Copies are bit-exact moves of the stored
// format.
//
// A direct segment (tile b -> the ghosts of its neighbour t across channel
// ch) and its partner (t -> the ghosts of b across -ch) have the same extents
// and read and write the same cache lines: b's boundary cells and b's ghost
// cells on that side lie in the same rows, and so do t's.  Along x, the
// contiguous axis, a row holds only a few bytes of either, so one thread
// copies the same cell of both segments: every line is read and then written
// while it is still in the cache, instead of being fetched once per segment
// and written back partially.
// ---------------------------------------------------------------------------
struct HaloCell {
    long long so, dofs;
};

__device__ inline HaloCell halo_cell(const HaloSeg &sg, int c)
{
    const int x = c % sg.ex;
    const int y = (c / sg.ex) % sg.ey;
    const int z = c / (sg.ex * sg.ey);
    return {sg.s_off + x + (long long)y * sg.s_sy + (long long)z * sg.s_sz,
            sg.d_off + x + (long long)y * sg.d_sy + (long long)z * sg.d_sz};
}

// s_waitcnt operand (gfx9 encoding) that waits for every vector memory
// access: vmcnt 0, expcnt and lgkmcnt at their maxima (not waited for).
constexpr int VMCNT0 = (7 << 4) | (15 << 8);

// The bits of one stored value.  A 2-byte value is held in a 32-bit register
// of its own: packing two into one would make each load wait for the last.
template <class S> struct HaloWord { using type = S; };
template <> struct HaloWord<_Float16> { using type = unsigned; };
template <> struct HaloWord<bf16_t> { using type = unsigned; };

template <class S, class W = typename HaloWord<S>::type>
__device__ inline void halo_load(const S *__restrict__ src, const HaloSeg &sg, const HaloCell &hc, W (&v)[Q])
{
    int slot = 0;
#pragma unroll
    for (int ii = 0; ii < Q; ++ii)
        if (sg.mask & (1u << ii)) {
            const S *p = src + hc.so + (long long)(sg.pack ? ii : slot) * sg.s_ds;
            if constexpr (sizeof(S) == 2) v[ii] = *reinterpret_cast<const unsigned short *>(p);
            else v[ii] = *p;
            ++slot;
        }
}

template <class S, class W = typename HaloWord<S>::type>
__device__ inline void halo_store(S *__restrict__ dst, const HaloSeg &sg, const HaloCell &hc, const W (&v)[Q])
{
    int slot = 0;
#pragma unroll
    for (int ii = 0; ii < Q; ++ii)
        if (sg.mask & (1u << ii)) {
            S *p = dst + hc.dofs + (long long)(sg.pack == 1 ? slot : ii) * sg.d_ds;
            if constexpr (sizeof(S) == 2) *reinterpret_cast<unsigned short *>(p) = (unsigned short)v[ii];
            else *p = v[ii];
            ++slot;
        }
}

template <class S>
__global__ void __launch_bounds__(256)
k_halo(const S *__restrict__ src, S *__restrict__ dst, const HaloSeg *__restrict__ segs,
       const int2 *__restrict__ blocks, int nblocks)
{
    using W = typename HaloWord<S>::type;
    // one workgroup per block entry (the grid is nblocks)
    if ((int)blockIdx.x >= nblocks) return;
    const int2 bk = blocks[blockIdx.x];
    const HaloSeg sg = segs[bk.x];
    const int c = bk.y + (int)threadIdx.x;
    if (c >= sg.ncell) return;
    const HaloCell hc = halo_cell(sg, c);
    W v[Q];
    halo_load(src, sg, hc, v);
    if (sg.partner >= 0) {
        const HaloSeg sp = segs[sg.partner];
        const HaloCell hp = halo_cell(sp, c);
        W w[Q];
        halo_load(src, sp, hp, w);
        // every load has landed: the stores then need no wait of their own
        // (the compiler cannot count the loads taken across the per-direction
        // branches and would wait for all outstanding accesses before each)
        __builtin_amdgcn_s_waitcnt(VMCNT0);
        halo_store(dst, sg, hc, v);
        halo_store(dst, sp, hp, w);
    } else {
        __builtin_amdgcn_s_waitcnt(VMCNT0);
        halo_store(dst, sg, hc, v);
    }
}

// ---------------------------------------------------------------------------
// Plain device copy, 16 bytes per lane, U loads in flight per lane (each
// workgroup copies one run of U * 4 KiB): the measured HBM ceiling that the
// step kernels are compared with.
// ---------------------------------------------------------------------------
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <bool NT, int U>
__global__ void __launch_bounds__(256)
k_copy(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, long long n)
{
    const long long base = (long long)blockIdx.x * (256 * U) + threadIdx.x;
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (base + u * 256 < n) v[u] = src[base + u * 256];
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (base + u * 256 < n) {
            if constexpr (NT) __builtin_nontemporal_store(v[u], dst + base + u * 256);
            else dst[base + u * 256] = v[u];
        }
}

}  // namespace imex
