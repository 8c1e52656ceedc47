// main.cpp -- the driver loop of the reference (src/main.cpp:6-110) on one
// GPU: the same time loop, the same reporting cadence and the same two
// printf formats, which downstream tools parse (tools/parse_scaling.sh reads
// the MLUPS line, tests/regression.sh the cons line).  MLUPS counts interior
// cells times steps over wall time and excludes the Conserved reduction.
// Timing follows the reference's scaling build (no per-step fences).
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "System.hpp"
#include "lbm.hpp"

static double wtime()
{
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

int main()
{
    {
        System s1;
        s1.Monitor();
        LBM l1(LBM_COMM_WORLD, s1.sx, s1.sy, s1.sz, s1.tau, s1.rho0, s1.u0);

        l1.Initialize();

        l1.fence();
        const double start = wtime();
        double t_last = start;

        for (int it = 1; it <= s1.Time; it++)
        {
            l1.Collision();
            l1.Streaming();
            // step 1 carries the one-time first collision (src/main.cpp:66-68)
            if (it == 1) l1.ProfileReset();

            if (it % s1.inter == 0)
            {
                l1.fence();
                const double now = wtime();
                if (l1.me == 0)
                {
                    const double dt_int = now - t_last;
                    const double mlups = (double)s1.sx * s1.sy * s1.sz * s1.inter / dt_int / 1.0e6;
                    std::printf("step %6d | interval %8.4f s | total %8.4f s | %10.2f MLUPS\n",
                                it, dt_int, now - start, mlups);
                }

                double mass, px, py, pz, ke;
                l1.Conserved(mass, px, py, pz, ke);
                if (l1.me == 0)
                {
                    std::printf("cons  step %6d | mass %.15e | px %.15e | py %.15e | pz %.15e | ke %.15e\n",
                                it, mass, px, py, pz, ke);
                    std::fflush(stdout);
                }
                // the reference leaves this call commented out (src/main.cpp:99)
                if (std::getenv("IMEXLBM_OUTPUT") && std::getenv("IMEXLBM_OUTPUT")[0] == '1')
                    l1.MPIoutput(it / s1.inter);
                t_last = wtime();
            }
        }
        l1.ProfileReport(s1.Time - 1);
    }
    return 0;
}
