/*
 * imexlbm.h -- C ABI of the MI355X-native IMEXLBM hot path.
 *
 * The library owns every device buffer; the caller owns the context.  All
 * calls on one context are stream-ordered on the context's own HIP stream and
 * must come from one host thread.  Only imex_conserved, imex_get_f,
 * imex_synchronize and imex_timed_steps block the host.  Every function
 * returns IMEX_OK (0) or a negative status; imex_strerror names it and
 * imex_last_error gives the detail of the most recent failure.
 *
 * Lattice state semantics follow the reference exactly: after n calls of
 * Collision + Streaming the populations are the post-streaming state S_n
 * (src/main.cpp:31-61, src/lbm.cpp:131-207).  Internally the library keeps a
 * post-collision state and fuses the pull-stream of one step with the
 * collision of the next; whatever reads S_n (imex_conserved, imex_get_f)
 * gathers it on the fly, which is bit-identical because streaming is a copy.
 */
#ifndef IMEXLBM_H
#define IMEXLBM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMEX_Q 27

typedef struct imex_ctx imex_ctx;

/* Storage precision of the populations.  IMEX_F64 stores f and computes in
 * FP64 in the reference's operation order.  The others store the shifted
 * populations g_i = f_i - w_i*rho0 and compute the collision in FP32. */
enum imex_precision { IMEX_F64 = 0, IMEX_F32 = 1, IMEX_F16 = 2, IMEX_BF16 = 3 };

/* Periodic ghost fill.  IMEX_WRAP: periodic index wrap inside the stream
 * kernel, no ghost cells.  IMEX_TILED: the domain is split into subdomains
 * with three ghost layers each, filled every step by one device pack and one
 * device unpack launch over all 26 neighbour channels of all subdomains. */
enum imex_mode { IMEX_WRAP = 0, IMEX_TILED = 1 };

enum imex_status {
    IMEX_OK = 0,
    IMEX_EINVAL = -1,   /* bad argument                                  */
    IMEX_ENOMEM = -2,   /* device or host allocation failed              */
    IMEX_EHIP = -3,     /* a HIP runtime call failed                     */
    IMEX_ENODEV = -4,   /* no usable GPU                                 */
    IMEX_ESTATE = -5,   /* call out of order (e.g. before initialize)    */
    IMEX_ERANGE = -6    /* size beyond what the kernels index            */
};

/* One Taylor-Green instance: the numbers the reference's System derives from
 * input.in (src/System.cpp:22-28) and hands to LBM::LBM (src/lbm.hpp:179). */
typedef struct imex_domain {
    int nx, ny, nz;     /* global interior cells (glx, gly, glz)          */
    double tau0;        /* relaxation parameter; rate is 1/(tau0+0.5)     */
    double rho0;        /* reference density                              */
    double u0;          /* Taylor-Green velocity amplitude                */
} imex_domain;

typedef struct imex_config {
    int precision;        /* enum imex_precision                          */
    int mode;             /* enum imex_mode                               */
    int ntiles;           /* IMEX_TILED: subdomain count, factored as the
                             reference's setup_Cartesian; 0 means 8       */
    int halo_depth;       /* IMEX_TILED: ghost layers filled, 1..3; 0 = 1 */
    int block;            /* threads per workgroup, multiple of 64; 0 = auto */
    int cells_per_thread; /* 0 = 2; 2 or more loads two cells per thread
                             before computing either                      */
    int persistent;       /* 0: one workgroup per work tile; >0: that many
                             workgroups stride over the work-tile table   */
    int nontemporal;      /* 1: streaming (non-temporal) stores, 0: plain
                             stores, -1 = default (streaming)             */
    int dir_pad;          /* extra elements between direction planes; -1 =
                             auto (32 for periodic FP64/FP32 contexts,
                             which keeps one row's 27 direction lines out
                             of a single L2 set; else 0)                 */
    int use_graph;        /* 1: imex_step replays a captured hipGraph     */
    int sequential;       /* several domains: one launch per domain per
                             step instead of one launch for all           */
    int device;           /* HIP device ordinal                           */
    int xcd_groups;       /* workgroups are renumbered so each of this many
                             XCDs takes its own work tiles (see xcd_run);
                             1 = dispatch order; 0 = auto: 8 (MI355X) for
                             8- and 4-byte formats, 1 for 2-byte ones     */
    int contiguous;       /* 1: ask for physically contiguous population
                             pools (falls back to a plain allocation)     */
    int io16;             /* 0 = buffer-descriptor loads and stores, 1 =
                             global loads and stores (any format);
                             -1 = auto (1 for 2-byte formats, else 0)     */
    int profile;          /* 1: time every phase with HIP events (one
                             synchronisation per phase, graphs off)       */
    int xcd_run;          /* with xcd_groups > 1: each XCD takes runs of
                             this many consecutive work tiles; 0 = one
                             contiguous range per XCD; -1 = auto (64 for
                             4-byte formats, else 0)                      */
    int two_step;         /* temporal blocking: 1 = advance pairs of steps
                             in one pass (each workgroup sweeps z over a
                             column of cells and keeps the intermediate
                             step in LDS), 0 = one pass per step, -1 =
                             auto (where the domain allows it and the
                             modelled pass time wins: periodic boxes at
                             least half a column tile across and one
                             tile deep in y, or tiled boxes that the
                             column tile divides)                        */
    int two_step_tile;    /* column tile: 0 = 64x8 cells (512 threads),
                             1 = 32x16 (512), 2 = 32x8 (256), 3 = 32x8
                             (384), 4 = 64x16 (1024, 2-byte formats only);
                             FP64 builds only 2 and 3.  -1 = auto: with
                             two_step_zseg, the shape of least modelled
                             pass time (workgroup waves x planes per
                             sweep x the tile's measured plane time)     */
    int two_step_zseg;    /* z-planes per workgroup sweep; 0 = auto       */
    int two_step_xcds;    /* tile order of the two-step pass over the
                             XCDs, as xcd_groups; 0 = auto (8: each XCD
                             sweeps its own z-range)                     */
    int two_step_prefetch; /* load the next plane's populations into
                             registers ahead of the barrier: 1 = every
                             ring cell's, 2 = the first 11 of each
                             thread's first ring cell (2-byte formats,
                             tiles 0, 1, 2 and 4; elsewhere as 1), 0 =
                             off, -1 = auto (2 for 2-byte formats, but 0
                             for BF16 in tile 4 over ghost cells, where
                             it spills; 1 for the others; the autotuner
                             tries each)                                 */
    int two_step_io;      /* periodic boxes: the pass's loads and stores
                             through 1 = global addresses, 0 = buffer
                             descriptors (direction offset in a scalar
                             operand); -1 = auto (1)                     */
    int halo_direct;      /* IMEX_TILED: 1 = each tile's boundary slabs are
                             copied straight into its neighbours' ghosts
                             (one launch), 0 = packed into an exchange
                             buffer and unpacked from it (two launches:
                             the shape of a cross-GPU exchange); -1 = auto
                             (1)                                         */
    int whole_x;          /* IMEX_TILED: 1 = tiles cut y and z only (the
                             setup_Cartesian factorisation with rx = 1)
                             and the kernels wrap x themselves, so no x
                             ghosts are kept or filled; 0 = the reference's
                             factorisation over x, y and z (default)     */
} imex_config;

typedef struct imex_info {
    long long cells;          /* interior cells over all domains          */
    long long bytes_per_update; /* algorithmic bytes per lattice update:
                                 2 * 27 * sizeof(storage)                  */
    long long device_bytes;   /* bytes the context holds on the device    */
    int storage_bytes;        /* sizeof one stored population             */
    int ndomains;
    int ntiles;               /* subdomains (IMEX_TILED) or 1 per domain  */
    int work_tiles;           /* entries in the work-tile table           */
    int block, cells_per_thread, persistent, nontemporal, dir_pad, use_graph;
    int grid;                 /* tile grid rx, ry, rz as setup_Cartesian  */
    int rx, ry, rz;
    int xcd_groups;
    int contiguous;           /* 1 if the pools were allocated contiguous */
    int io16;
    int xcd_run;
    int two_step;             /* 1 if pairs of steps run as one pass      */
    int two_step_tile, two_step_zseg, two_step_xcds, two_step_prefetch, two_step_io;
    int halo_direct;          /* IMEX_TILED: 1 direct ghost fill, 0 pack/unpack */
    int halo_depth;           /* IMEX_TILED: ghost layers filled (2 when two
                                 steps run per pass)                          */
    int whole_x;              /* IMEX_TILED: 1 if tiles keep x whole and the
                                 kernels wrap it                              */
    long long halo_send_bytes; /* IMEX_TILED: bytes of the exchange buffer
                                 one pack fills (the 26 send buffers of the
                                 reference, src/profile.hpp:90-93); 0 in
                                 IMEX_WRAP                                   */
} imex_info;

/* Defaults for every field of imex_config: FP64, wrap, 256-thread
 * workgroups, two cells per thread, streaming stores, and the per-format
 * automatic tile order and 16-bit access path. */
void imex_config_default(imex_config *cfg);

/* LBM::LBM + the allocation half of LBM::Initialize (src/lbm.hpp:179-183,
 * src/lbm.cpp:38-57).  ndomains > 1 creates a batch of independent periodic
 * instances advanced by one launch per step (IMEX_WRAP only). */
int imex_create(imex_ctx **out, const imex_config *cfg,
                const imex_domain *domains, int ndomains);

/* LBM::Initialize (src/lbm.cpp:38-111): analytic Taylor-Green field and the
 * FP64 pressure-based equilibrium, computed on the device from per-axis
 * sin/cos tables that the host evaluates with the C library. */
int imex_initialize(imex_ctx *ctx);

/* LBM::Collision (src/lbm.cpp:131-166).  Deferred: it is applied, fused with
 * the next pull-stream, by imex_streaming or by whatever reads the state. */
int imex_collision(imex_ctx *ctx);

/* LBM::Streaming (src/lbm.cpp:168-207) including passf's ghost fill
 * (src/mpi_view_transfer.cpp:451-456). */
int imex_streaming(imex_ctx *ctx);

/* nsteps x (LBM::Collision; LBM::Streaming), the driver's loop body
 * (src/main.cpp:31-61).  Asynchronous. */
int imex_step(imex_ctx *ctx, int nsteps);

/* LBM::Conserved (src/lbm.cpp:266-306) for one domain: out = {mass, px, py,
 * pz, ke} of S_n, summed in compensated FP64.  Blocks. */
int imex_conserved(imex_ctx *ctx, int domain, double out[5]);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc): LBM::Conserved
 * (src/lbm.cpp:266-306) over post-streaming populations the caller holds in
 * device memory, with no context and no device allocation.  The partial sums
 * live in a fixed array inside the library (IMEX_NOALLOC_PARTIALS workgroup
 * partials of 80 bytes plus the result, 82 KB per GPU, held until the
 * process ends); every call on one GPU uses that GPU's copy, so calls that
 * overlap on one GPU (two host threads, or two streams) race.
 *
 * f: one periodic domain of nx x ny x nz cells in the library's layout,
 * f[z][ii][y][x] with x fastest, `dstride` elements between direction
 * planes (0 means nx*ny) and 27*dstride between z-planes; FP64 holds f,
 * FP32/FP16/BF16 hold the shifted g = f - w_ii*rho0 (enum imex_precision).
 * The sums are those of imex_conserved (compensated FP64, {mass, px, py, pz,
 * ke}), accumulated in the order a default single-domain context uses (256
 * threads and 2 cells per thread: runs of 512 cells of one z-plane), so they
 * equal its result bit for bit.  Runs on `stream` (NULL: the null stream),
 * then blocks until out holds the sums.  IMEX_ERANGE when one z-plane of the
 * layout exceeds the kernels' 32-bit addressing (27*dstride*sizeof >= 2 GiB,
 * or nx*ny*nx >= 2^32). */
#define IMEX_NOALLOC_PARTIALS 1024
int imex_conserved_noalloc(const void *f, int precision, int nx, int ny, int nz, long long dstride,
                           double rho0, void *stream, double out[5]);

/* Copy S_n of one domain to the host as doubles in the reference's
 * (ii, i, j, k) order, host[((ii*nx + i)*ny + j)*nz + k].  Blocks. */
int imex_get_f(imex_ctx *ctx, int domain, double *host);

/* LBM::ComputeMacroscopic (src/lbm.cpp:213-240) for one domain, copied to
 * the host: p = (Σf)/3 and (u, v, w) = Σf e of S_n, each nx*ny*nz doubles in
 * the reference's (i, j, k) order.  Blocks. */
int imex_get_macro(imex_ctx *ctx, int domain, double *p, double *u, double *v,
                   double *w);

/* Replace one domain's state by S given in the same order (tests use it to
 * drive collision and streaming from arbitrary inputs). */
int imex_set_f(imex_ctx *ctx, int domain, const double *host);

/* Wait for all work queued on the context, including a Collision() whose
 * Streaming() has not come yet (the counterpart of Kokkos::fence()). */
int imex_synchronize(imex_ctx *ctx);

/* Run nsteps steps between two synchronisations and report the wall time
 * and the summed duration of the step kernels, measured with HIP events on
 * the context's stream.  kernel_ms covers the fused step launches only. */
int imex_timed_steps(imex_ctx *ctx, int nsteps, double *wall_ms,
                     double *kernel_ms);

/* Time candidate launch shapes (workgroup size, cells per thread, tile order
 * over the XCDs, plane access path) on the live arrays without changing the
 * state, confirm the four fastest with a second timing, and keep the best.
 * Streaming stores and one work tile per workgroup are fixed.  Blocks.
 * best_ms (may be NULL) receives its step time. */
int imex_autotune(imex_ctx *ctx, int reps, double *best_ms);

int imex_query(imex_ctx *ctx, imex_info *info);

/* Per-phase profiler, the counterpart of the reference's LBM_PROFILE build
 * (src/profile.hpp:42-113, src/profile.cpp:64-204).  With cfg.profile set,
 * every launch is timed with HIP events and charged to a group.  Groups 0-5
 * carry the reference's group names and byte and flop models
 * (src/profile.cpp:64-97): "collision" (in-place collide), "pack (enq+fence)",
 * "mpi (post+wait)" (nothing on one GPU: the unpack reads the pack buffer),
 * "unpack (enq+fence)", "stream (enq+fence)" (bare pull-stream) and
 * "computeMacroscopic".  Groups 6-8 have no counterpart there: the fused
 * pull-stream + collide launch ("collision+stream (fused)", one or two
 * lattice updates per launch), the direct ghost fill and the conserved sums.
 * imex_profile_read gives out = {total ms, fastest ms, slowest ms, bytes,
 * flops} summed over the calls, and the call count; bytes are algorithmic:
 * 2 x cells x 27 x sizeof(storage) for a step launch (each population read
 * and written once), twice the exchange-buffer bytes for pack and unpack,
 * 27 x sizeof + 32 per cell for the macroscopic fields; flops are 594 per
 * collided cell and 189 per cell of moments.  imex_profile_phase_name
 * returns NULL past the last group. */
const char *imex_profile_phase_name(int phase);
int imex_profile_reset(imex_ctx *ctx);
int imex_profile_read(imex_ctx *ctx, int phase, double out[5], long long *calls);

/* Measured HBM ceiling: a plain device-to-device copy of `bytes` (16 bytes
 * per lane; one or four loads per lane, plain or streaming stores: the best
 * of the four), timed over `reps` launches; *gbs receives read plus write
 * bytes per second / 1e9.  No counterpart in the reference, whose
 * phase profiler assumes a fixed 1600 GB/s peak instead
 * (src/profile.hpp:67-73); the step kernels are reported against this. */
int imex_copy_probe(int device, long long bytes, int reps, double *gbs);
void imex_destroy(imex_ctx *ctx);

int imex_device_count(void);
const char *imex_strerror(int status);
const char *imex_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
