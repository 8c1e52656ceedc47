// lbm.hpp -- drop-in for the reference's solver object (src/lbm.hpp:113-211).
//
// Same constructor arguments, the same step methods (Initialize, Collision,
// Streaming, Conserved) with the same semantics, and the same public
// geometry fields, so a driver written against the reference compiles
// against this header.  The populations live on the GPU behind the C ABI of
// include/imexlbm.h; there is no Kokkos and no per-population View.
//
// One process drives one GPU.  The communicator argument is kept for source
// compatibility; only the one-rank world exists (me = 0, nranks = 1), and
// the 26-neighbour periodic exchange runs on the device (in-kernel wrap, or
// device pack/unpack between subdomain tiles when IMEXLBM_MODE=tiled).
//
// Storage precision, halo mode and launch shape are chosen from environment
// variables read by the constructor, because the reference's nine-number
// deck has no place for them:
//   IMEXLBM_PRECISION = fp64 (default) | fp32 | fp16 | bf16
//   IMEXLBM_MODE      = wrap (default) | tiled
//   IMEXLBM_TILES     = subdomain count in tiled mode (default 8)
//   IMEXLBM_HALO      = ghost layers filled per exchange, 1..3 (default 1)
//   IMEXLBM_AUTOTUNE  = 1 to time launch shapes after Initialize
//   IMEXLBM_OUTPUT    = 1 to write output<n>.plt at every report
//   IMEXLBM_PROFILE   = 1 for the per-phase profile table at the end
#ifndef IMEXLBM_LBM_HPP
#define IMEXLBM_LBM_HPP

#include <vector>

#include "imexlbm.h"

// Lattice constants of the reference (src/lbm.hpp:22-24), as constants
// rather than macros so that no later identifier is clobbered.
constexpr int LBM_Q = 27;
constexpr int LBM_DIM = 3;
constexpr int LBM_GHOST = 3;

typedef int lbm_comm;             // the one-rank world
constexpr lbm_comm LBM_COMM_WORLD = 0;

struct LBM
{
    lbm_comm comm;
    int nranks = 1;

    int rx = 1, ry = 1, rz = 1;
    int me = 0;
    int px = 0, py = 0, pz = 0;

    int glx, gly, glz;
    // include ghost nodes
    int lx, ly, lz;
    // local start, local end, local length
    int l_s[3], l_e[3], l_l[3];
    int x_lo = 0, x_hi = 0, y_lo = 0, y_hi = 0, z_lo = 0, z_hi = 0;
    double rho0, mu, cs2, tau0, u0;

    imex_config cfg;        // device-side options
    imex_ctx *ctx = nullptr;

    LBM(lbm_comm comm_, int sx, int sy, int sz, double &tau, double &rho0, double &u0);
    ~LBM();
    LBM(const LBM &) = delete;
    LBM &operator=(const LBM &) = delete;

    void Initialize();
    void Collision();
    void Streaming();
    void Conserved(double &mass, double &mom_x, double &mom_y, double &mom_z, double &ke);

    // Field output (src/lbm.cpp:213-240, 308-636): p, u, v, w of the current
    // state on the host, a Tecplot binary "output<n>.plt" and the ASCII
    // "output<n><me>.dat" of the rest population.
    void ComputeMacroscopic();
    void MPIoutput(int n);
    void Output(int n);
    std::vector<double> p_h, u_h, v_h, w_h;   // (i, j, k), k fastest

    // Per-phase profile (the reference's LBM_PROF_RESET / LBM_PROF_REPORT).
    void ProfileReset();
    void ProfileReport(int nsteps);

    // Device synchronisation, the counterpart of Kokkos::fence().
    void fence();
    // Copy S_n to the host in the reference's (ii, i, j, k) order.
    void download_f(double *host);
};

#endif
