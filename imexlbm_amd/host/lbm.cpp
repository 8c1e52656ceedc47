// lbm.cpp -- the solver object over the C ABI.  Every failure from the
// library is fatal with a message, as MPI and Kokkos errors are fatal in the
// reference.
#include "lbm.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

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

LBM::LBM(lbm_comm comm_, int sx, int sy, int sz, double &tau, double &rho0_, double &u0_)
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
    cfg.profile = env_int("IMEXLBM_PROFILE", 0);
    cfg.two_step = env_int("IMEXLBM_TWO_STEP", -1);
    cfg.whole_x = env_int("IMEXLBM_WHOLE_X", 0);

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
    mass = out[0]; mom_x = out[1]; mom_y = out[2]; mom_z = out[3]; ke = out[4];
}

void LBM::fence() { check(imex_synchronize(ctx), "imex_synchronize"); }

void LBM::download_f(double *host) { check(imex_get_f(ctx, 0, host), "imex_get_f"); }

void LBM::ProfileReset()
{
    if (cfg.profile) check(imex_profile_reset(ctx), "imex_profile_reset");
}

// Table in the spirit of src/profile.cpp:99-204: per phase the mean time per
// timed step, the fastest and slowest call, the share of the step, and the
// algorithmic bandwidth against a copy ceiling measured on this device
// (where the reference assumes 1600 GB/s).
void LBM::ProfileReport(int nsteps)
{
    if (!cfg.profile || nsteps <= 0) return;
    double peak = 0.0;
    if (imex_copy_probe(cfg.device, 4LL << 30, 5, &peak) != IMEX_OK) peak = 0.0;
    std::printf("\n=== per-phase profile: 1 GPU, %d timed steps ===\n", nsteps);
    std::printf("%-34s %6s %10s %9s %9s %7s %9s %7s\n", "phase", "calls", "ms/step", "min ms", "max ms",
                "step%", "GB/s", "%copy");
    double step_ms = 0.0;
    for (int p = 0; imex_profile_phase_name(p); ++p) {
        double v[4];
        long long n;
        check(imex_profile_read(ctx, p, v, &n), "imex_profile_read");
        step_ms += v[0];
    }
    for (int p = 0; imex_profile_phase_name(p); ++p) {
        double v[4];
        long long n;
        check(imex_profile_read(ctx, p, v, &n), "imex_profile_read");
        if (n == 0) continue;
        const double gbs = v[0] > 0 ? v[3] / (v[0] * 1e-3) / 1e9 : 0.0;
        std::printf("%-34s %6lld %10.4f %9.4f %9.4f %7.1f %9.1f %7.1f\n", imex_profile_phase_name(p), n,
                    v[0] / nsteps, v[1], v[2], step_ms > 0 ? 100.0 * v[0] / step_ms : 0.0, gbs,
                    peak > 0 ? 100.0 * gbs / peak : 0.0);
    }
    std::printf("copy ceiling measured on this device: %.1f GB/s\n", peak);
    std::printf("profile run (one synchronisation per phase): take throughput from the MLUPS lines\n\n");
}
