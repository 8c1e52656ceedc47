// A driver written against the reference's solver interface
// (src/lbm.hpp:22-24,113-210, src/System.hpp:7-20) in the way its main
// program calls it (src/main.cpp:6-107): MPI start-up, the deck, the solver
// constructed on MPI_COMM_WORLD, Initialize, Collision/Streaming steps and
// the collective Conserved.  The tests compile it against the drop-in
// headers, with the one-rank mpi.h and with an installed MPI.
#include <mpi.h>

#include <cstdio>

#include "System.hpp"
#include "lbm.hpp"

static_assert(q == 27, "q is the number of lattice directions");
static_assert(dim == 3, "dim is the number of space dimensions");
static_assert(ghost == 3, "ghost is the number of ghost layers");

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    {
        int rank = -1, size = 0;
        MPI_Comm_rank(MPI_COMM_WORLD, &rank);
        MPI_Comm_size(MPI_COMM_WORLD, &size);
        System s1;
        if (rank == 0) s1.Monitor();
        MPI_Barrier(MPI_COMM_WORLD);

        LBM l1(MPI_COMM_WORLD, s1.sx, s1.sy, s1.sz, s1.tau, s1.rho0, s1.u0);
        // geometry fields a driver may read before Initialize
        std::printf("driver: ranks %d lx %d ly %d lz %d l_s %d l_e %d cs2 %.17g\n", size, l1.lx, l1.ly,
                    l1.lz, l1.l_s[0], l1.l_e[0], l1.cs2);
        std::fflush(stdout);

        l1.Initialize();
        for (int it = 1; it <= s1.Time; ++it)
        {
            l1.Collision();
            l1.Streaming();
        }
        double mass, px, py, pz, ke;
        l1.Conserved(mass, px, py, pz, ke);
        if (l1.me == 0)
            std::printf("cons  step %6d | mass %.15e | px %.15e | py %.15e | pz %.15e | ke %.15e\n", s1.Time,
                        mass, px, py, pz, ke);
    }
    MPI_Finalize();
    return 0;
}
