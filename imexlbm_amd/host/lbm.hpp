// lbm.hpp -- drop-in for the reference's solver object (src/lbm.hpp:113-211).
//
// Same constructor (an MPI communicator, the global extents and tau, rho0,
// u0 by reference, src/lbm.hpp:179-183), the same step methods (Initialize,
// Collision, Streaming, Conserved) with the same semantics, the same public
// geometry fields and the lattice names q, dim and ghost, so a driver written
// against the reference compiles against this header.  The populations live
// on the GPU behind the C ABI of include/imexlbm.h; there is no Kokkos and no
// per-population View.
//
// One process drives one GPU and owns the whole domain: Initialize() reads
// the communicator's rank and size (src/mpi_view_transfer.cpp:41-51) and
// stops with a message on more than one rank.  The 26-neighbour periodic
// exchange runs on the device (in-kernel wrap, or device pack/unpack between
// subdomain tiles when IMEXLBM_MODE=tiled).  Conserved() is collective as in
// the reference: the device sums go through MPI_Allreduce on the
// communicator.  Builds without an MPI library use the one-rank mpi.h under
// mpi_stub/.
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
//   IMEXLBM_PROFILE   = 1 for the per-phase profile tables at the end (also
//                       on in a build with -DLBM_PROFILE)
#ifndef IMEXLBM_LBM_HPP
#define IMEXLBM_LBM_HPP

#include <mpi.h>

#include <vector>

#include "imexlbm.h"

// The lattice constants as typed constants, for code that must not see the
// macros below.
constexpr int LBM_Q = 27;
constexpr int LBM_DIM = 3;
constexpr int LBM_GHOST = 3;

struct LBM
{
    MPI_Comm comm;
    int nranks = 1;

    int rx = 1, ry = 1, rz = 1;
    // rank
    int me = 0;
    // axis for each rank
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

    LBM(MPI_Comm comm_, int sx, int sy, int sz, double &tau, double &rho0, double &u0);
    ~LBM();
    LBM(const LBM &) = delete;
    LBM &operator=(const LBM &) = delete;

    void Initialize();
    void Collision();
    void Streaming();
    // Collective over comm, blocking: {mass, momentum, KE} of the current
    // (post-streaming) state, src/lbm.cpp:266-306.
    void Conserved(double &mass, double &mom_x, double &mom_y, double &mom_z, double &ke);

    // Field output (src/lbm.cpp:213-240, 308-636): p, u, v, w of the current
    // state on the host, a Tecplot binary "output<n>.plt" and the ASCII
    // "output<n><me>.dat" of the rest population.
    void ComputeMacroscopic();
    void MPIoutput(int n);
    void Output(int n);
    std::vector<double> p_h, u_h, v_h, w_h;   // (i, j, k), k fastest

    // Per-phase profile (the reference's LBM_PROF_BEGIN/END(P_STEP),
    // LBM_PROF_RESET and LBM_PROF_REPORT, src/profile.hpp:96-113): the step
    // is timed on the host around Collision + Streaming, the device groups
    // with HIP events inside the library.
    bool profiling() const { return cfg.profile != 0; }
    void ProfileStepBegin();
    void ProfileStepEnd();
    void ProfileReset();
    void ProfileReport(int nsteps);
    double step_t0 = 0.0, step_acc = 0.0, step_min = 0.0, step_max = 0.0;
    long step_calls = 0;

    // Device synchronisation, the counterpart of Kokkos::fence(): every step
    // the calls so far have issued, including a Collision() whose Streaming()
    // has not come, has executed when it returns.
    void fence();
    // Copy S_n to the host in the reference's (ii, i, j, k) order.
    void download_f(double *host);
};

// The reference's lattice macros (src/lbm.hpp:22-24), defined after every
// header this one includes, as the reference defines them after its own.
// As there, code that includes this header cannot use q, dim or ghost as
// identifiers.
#define q 27
#define dim 3
#define ghost 3

#endif
