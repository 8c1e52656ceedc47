// main.cpp -- the driver loop of the reference (src/main.cpp:6-110) on one
// GPU: MPI start-up, the deck, the same time loop and reporting cadence, and
// the same printf formats, which downstream tools parse
// (tools/parse_scaling.sh reads the MLUPS line, tests/regression.sh the cons
// line).  MLUPS counts interior cells times steps over wall time and
// excludes the Conserved reduction.
//
// Like the reference, a default build times the collision of every step
// between two fences and prints "collision | ... ms | ... GFLOPS" with 594
// flop per cell; -DLBM_NO_PERSTEP_TIMING (the scaling build) leaves that out.
// In this library the collision runs fused with the pull-stream of the
// previous step, so the timed region is that fused launch: the work of one
// step.  The per-step fences also keep steps from pairing into two-step
// passes, so take throughput from the scaling build.
#include <mpi.h>

#include <cstdio>
#include <cstdlib>

#include "System.hpp"
#include "lbm.hpp"

int main(int argc, char *argv[])
{
    MPI_Init(&argc, &argv);
    {
        int rank = 0;
        MPI_Comm_rank(MPI_COMM_WORLD, &rank);

        System s1;
        if (rank == 0) s1.Monitor();
        MPI_Barrier(MPI_COMM_WORLD);
        LBM l1(MPI_COMM_WORLD, s1.sx, s1.sy, s1.sz, s1.tau, s1.rho0, s1.u0);

        l1.Initialize();

        l1.fence();
        MPI_Barrier(MPI_COMM_WORLD);
        const double start = MPI_Wtime();
        double t_last = start;
        const bool write_fields = std::getenv("IMEXLBM_OUTPUT") && std::getenv("IMEXLBM_OUTPUT")[0] == '1';

        for (int it = 1; it <= s1.Time; it++)
        {
            l1.ProfileStepBegin();
#ifndef LBM_NO_PERSTEP_TIMING
            l1.fence();
            const double t0 = MPI_Wtime();
            l1.Collision();
            l1.fence();
            const double dt_coll = MPI_Wtime() - t0;
            if (l1.me == 0)
            {
                const double gflops = 594.0 * s1.sx * s1.sy * s1.sz / dt_coll / 1.0e9;
                std::printf("collision | %8.4f ms | %10.2f GFLOPS\n", dt_coll * 1e3, gflops);
            }
#else
            l1.Collision();
#endif
            l1.Streaming();
            l1.ProfileStepEnd();

            // step 1 carries the one-time first collision and start-up costs
            // (src/main.cpp:66-68)
            if (it == 1) l1.ProfileReset();

            if (it % s1.inter == 0)
            {
                l1.fence();
                MPI_Barrier(MPI_COMM_WORLD);
                const double now = MPI_Wtime();
                if (l1.me == 0)
                {
                    const double dt_int = now - t_last;
                    const double mlups = (double)s1.sx * s1.sy * s1.sz * s1.inter / dt_int / 1.0e6;
                    std::printf("step %6d | interval %8.4f s | total %8.4f s | %10.2f MLUPS\n",
                                it, dt_int, now - start, mlups);
                }

                // collective; printed without an "MLUPS" token, which the
                // scaling parser selects on
                double mass, px, py, pz, ke;
                l1.Conserved(mass, px, py, pz, ke);
                if (l1.me == 0)
                {
                    std::printf("cons  step %6d | mass %.15e | px %.15e | py %.15e | pz %.15e | ke %.15e\n",
                                it, mass, px, py, pz, ke);
                    std::fflush(stdout);
                }
                // the reference leaves this call commented out (src/main.cpp:99)
                if (write_fields) l1.MPIoutput(it / s1.inter);
                // the diagnostics are not charged to the next interval
                t_last = MPI_Wtime();
            }
        }

        // step 1 was discarded above
        l1.ProfileReport(s1.Time - 1);
    }
    MPI_Finalize();
    return 0;
}
