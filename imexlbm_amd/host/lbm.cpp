// lbm.cpp -- the solver object over the C ABI.  Every failure from the
// library is fatal with a message, as MPI and Kokkos errors are fatal in the
// reference.
#include "lbm.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

// Peak device memory bandwidth per GPU for the profiler's "% BW" column;
// the reference assumes 1600 GB/s (src/profile.hpp:67-73), MI355X's
// published HBM3E figure is 8 TB/s.  Override with -DLBM_PROF_PEAK_GBPS=...
#ifndef LBM_PROF_PEAK_GBPS
#define LBM_PROF_PEAK_GBPS 8000.0
#endif

namespace {

void check(int status, const char *what)
{
    if (status != IMEX_OK)
    {
        std::fprintf(stderr, "FATAL: %s: %s (%s)\n", what, imex_strerror(status), imex_last_error());
        std::abort();
    }
}

int env_int(const char *name, int dflt)
{
    const char *v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}

}  // namespace

LBM::LBM(MPI_Comm comm_, int sx, int sy, int sz, double &tau, double &rho0_, double &u0_)
    : comm(comm_), glx(sx), gly(sy), glz(sz), rho0(rho0_), tau0(tau), u0(u0_)
{
    cs2 = 1.0 / 3.0;
    mu = 0.0;
    imex_config_default(&cfg);
    if (const char *p = std::getenv("IMEXLBM_PRECISION"))
    {
        const std::string s(p);
        if (s == "fp64") cfg.precision = IMEX_F64;
        else if (s == "fp32") cfg.precision = IMEX_F32;
        else if (s == "fp16") cfg.precision = IMEX_F16;
        else if (s == "bf16") cfg.precision = IMEX_BF16;
        else { std::fprintf(stderr, "FATAL: IMEXLBM_PRECISION=%s\n", p); std::abort(); }
    }
    if (const char *m = std::getenv("IMEXLBM_MODE"))
    {
        const std::string s(m);
        if (s == "wrap") cfg.mode = IMEX_WRAP;
        else if (s == "tiled") cfg.mode = IMEX_TILED;
        else { std::fprintf(stderr, "FATAL: IMEXLBM_MODE=%s\n", m); std::abort(); }
    }
    cfg.ntiles = env_int("IMEXLBM_TILES", 8);
    cfg.halo_depth = env_int("IMEXLBM_HALO", 1);
#ifdef LBM_PROFILE
    cfg.profile = env_int("IMEXLBM_PROFILE", 1);
#else
    cfg.profile = env_int("IMEXLBM_PROFILE", 0);
#endif
    cfg.two_step = env_int("IMEXLBM_TWO_STEP", -1);
    cfg.whole_x = env_int("IMEXLBM_WHOLE_X", 0);
    cfg.halo_direct = env_int("IMEXLBM_HALO_DIRECT", -1);

    // one rank owns the whole domain (src/mpi_view_transfer.cpp:163-195 at rx=ry=rz=1)
    l_l[0] = glx; l_l[1] = gly; l_l[2] = glz;
    lx = glx + 2 * LBM_GHOST; ly = gly + 2 * LBM_GHOST; lz = glz + 2 * LBM_GHOST;
    for (int d = 0; d < 3; ++d) { l_s[d] = LBM_GHOST; l_e[d] = LBM_GHOST + l_l[d]; }
    x_hi = glx - 1; y_hi = gly - 1; z_hi = glz - 1;
}

LBM::~LBM()
{
    imex_destroy(ctx);
}

void LBM::Initialize()
{
    // setup_MPI (src/mpi_view_transfer.cpp:41-51): this process's rank and
    // the communicator's size.  One GPU process owns the whole domain here.
    MPI_Comm_rank(comm, &me);
    MPI_Comm_size(comm, &nranks);
    if (nranks != 1)
    {
        if (me == 0)
            std::fprintf(stderr,
                         "FATAL: this build advances the whole domain on one GPU per process; "
                         "run it with one MPI rank (got %d)\n", nranks);
        std::fflush(stderr);
        MPI_Abort(comm, 1);
    }
    imex_domain d;
    d.nx = glx; d.ny = gly; d.nz = glz;
    d.tau0 = tau0; d.rho0 = rho0; d.u0 = u0;
    check(imex_create(&ctx, &cfg, &d, 1), "imex_create");
    imex_info info;
    check(imex_query(ctx, &info), "imex_query");
    if (me == 0 && cfg.mode == IMEX_TILED)
        std::printf("rx=%d,ry=%d,rz=%d\n", info.rx, info.ry, info.rz);
    check(imex_initialize(ctx), "imex_initialize");
    if (env_int("IMEXLBM_AUTOTUNE", 0)) check(imex_autotune(ctx, 3, nullptr), "imex_autotune");
}

void LBM::Collision() { check(imex_collision(ctx), "imex_collision"); }

void LBM::Streaming() { check(imex_streaming(ctx), "imex_streaming"); }

void LBM::Conserved(double &mass, double &mom_x, double &mom_y, double &mom_z, double &ke)
{
    double out[5];
    check(imex_conserved(ctx, 0, out), "imex_conserved");
    // the reference's global reduction of the five sums (src/lbm.cpp:297-299)
    MPI_Allreduce(MPI_IN_PLACE, out, 5, MPI_DOUBLE, MPI_SUM, comm);
    mass = out[0]; mom_x = out[1]; mom_y = out[2]; mom_z = out[3]; ke = out[4];
}

void LBM::fence() { check(imex_synchronize(ctx), "imex_synchronize"); }

void LBM::download_f(double *host) { check(imex_get_f(ctx, 0, host), "imex_get_f"); }

void LBM::ProfileStepBegin()
{
    if (cfg.profile) step_t0 = MPI_Wtime();
}

void LBM::ProfileStepEnd()
{
    if (!cfg.profile) return;
    // every launch of a profiled context completes before its call returns
    const double dt = MPI_Wtime() - step_t0;
    step_min = step_calls ? std::min(step_min, dt) : dt;
    step_max = step_calls ? std::max(step_max, dt) : dt;
    step_acc += dt;
    ++step_calls;
}

void LBM::ProfileReset()
{
    if (!cfg.profile) return;
    check(imex_profile_reset(ctx), "imex_profile_reset");
    step_acc = step_min = step_max = 0.0;
    step_calls = 0;
}

// The reference's report (src/profile.cpp:101-204) for one GPU: the phase
// table (per-call mean, fastest and slowest call, spread and share of the
// step) and the group table, the only place rates are given: GFLOP/s from
// 594 flop per collided cell and 189 per cell of moments, GB/s from the
// algorithmic bytes of each group, and "% BW" against LBM_PROF_PEAK_GBPS.
void LBM::ProfileReport(int nsteps)
{
    if (!cfg.profile) return;
    if (nsteps <= 0)
    {
        if (me == 0) std::printf("\n[profile] nothing to report (nsteps <= 0)\n");
        return;
    }
    imex_info info;
    check(imex_query(ctx, &info), "imex_query");
    if (me != 0) return;

    const double step = step_calls ? 1e3 * step_acc / step_calls : 0.0;
    std::printf("\n=== per-phase profile: %d ranks, %d timed steps ===\n", nranks, nsteps);
    std::printf("mean ms = per-call mean; spread%% = (slowest - fastest) / mean\n");
    std::printf("imbal   = max/min across ranks (1.00 with one rank)\n");
    std::printf("device groups are timed with HIP events, one synchronisation per launch\n\n");
    std::printf("%-36s %6s %9s %8s %8s %7s %6s %7s\n", "phase", "calls", "mean ms", "min ms", "max ms",
                "spread%", "imbal", "% step");
    std::printf("%-36s %6s %9s %8s %8s %7s %6s %7s\n", "------------------------------------", "------",
                "---------", "--------", "--------", "-------", "------", "-------");
    std::printf("%-36s %6ld %9.4f %8.4f %8.4f %7.1f %6.2f %7.1f\n", "STEP TOTAL", step_calls, step,
                1e3 * step_min, 1e3 * step_max, step > 0 ? 100.0 * 1e3 * (step_max - step_min) / step : 0.0,
                1.0, 100.0);
    double attributed = 0.0;   // per step
    for (int p = 0; imex_profile_phase_name(p); ++p)
    {
        double v[5];
        long long n;
        check(imex_profile_read(ctx, p, v, &n), "imex_profile_read");
        if (n == 0) continue;
        const double mean = v[0] / n, per_step = v[0] / nsteps;
        attributed += per_step;
        std::printf("  %-34s %6lld %9.4f %8.4f %8.4f %7.1f %6.2f %7.1f\n", imex_profile_phase_name(p), n, mean,
                    v[1], v[2], mean > 0 ? 100.0 * (v[2] - v[1]) / mean : 0.0, 1.0,
                    step > 0 ? 100.0 * per_step / step : 0.0);
    }
    std::printf("%-36s %6s %9.4f %8s %8s %7s %6s %7.1f\n", "  (unattributed)", "", step - attributed, "", "",
                "", "", step > 0 ? 100.0 * (step - attributed) / step : 0.0);

    std::printf("\n%-26s %9s %7s %10s %10s %7s\n", "group", "ms/step", "% step", "GFLOP/s", "GB/s", "% BW");
    std::printf("%-26s %9s %7s %10s %10s %7s\n", "--------------------------", "---------", "-------",
                "----------", "----------", "-------");
    for (int p = 0; imex_profile_phase_name(p); ++p)
    {
        double v[5];
        long long n;
        check(imex_profile_read(ctx, p, v, &n), "imex_profile_read");
        if (n == 0) continue;
        const double secs = 1e-3 * v[0];
        char gf[16], gb[16], bw[16];
        if (v[4] > 0 && secs > 0) std::snprintf(gf, sizeof gf, "%10.1f", v[4] / secs / 1e9);
        else std::snprintf(gf, sizeof gf, "%10s", "-");
        if (v[3] > 0 && secs > 0)
        {
            const double gbs = v[3] / secs / 1e9;
            std::snprintf(gb, sizeof gb, "%10.1f", gbs);
            std::snprintf(bw, sizeof bw, "%7.1f", 100.0 * gbs / LBM_PROF_PEAK_GBPS);
        }
        else
        {
            std::snprintf(gb, sizeof gb, "%10s", "-");
            std::snprintf(bw, sizeof bw, "%7s", "-");
        }
        std::printf("%-26s %9.4f %7.1f %s %s %s\n", imex_profile_phase_name(p), v[0] / nsteps,
                    step > 0 ? 100.0 * v[0] / nsteps / step : 0.0, gf, gb, bw);
    }
    std::printf("\nworkload/rank: %lld interior cells, halo %.2f MB send + %.2f MB recv\n", info.cells,
                info.halo_send_bytes / 1048576.0, info.halo_send_bytes / 1048576.0);
    std::printf("bytes are algorithmic: a fused launch reads and writes every population once\n");
    std::printf("  for one or two lattice updates; %% BW is against %.0f GB/s.\n", (double)LBM_PROF_PEAK_GBPS);
    std::printf("PROFILE run (one synchronisation per launch) -- take throughput from SCALING.\n\n");
}
