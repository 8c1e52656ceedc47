// System.cpp -- reads ./input.in: "rho0 R Re / u0 Time inter / nx ny nz",
// trailing text ignored, and derives Ma, miu and tau with the reference's
// expressions and the literal 3.1415926 (src/System.cpp:11-44).  A missing
// deck aborts, as in the reference.
#include "System.hpp"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>

System::System()
{
    cs2 = 1.0 / 3.0;
    cs = std::sqrt(cs2);

    std::ifstream input("input.in");
    if (!input)
    {
        std::cerr << "FATAL: cannot open input.in" << std::endl;
        std::abort();
    }
    input >> rho0 >> R >> Re;
    input >> u0 >> Time >> inter;
    input >> sx >> sy >> sz;
    if (!input)
    {
        std::cerr << "FATAL: input.in must hold nine numbers: rho0 R Re / u0 Time inter / nx ny nz"
                  << std::endl;
        std::abort();
    }

    Ma = u0 / cs;
    miu = rho0 * u0 * sx / 2.0 / 3.1415926 / Re;
    tau = u0 * sx / Re / cs2 / 2.0 / 3.1415926;
}

void System::Monitor()
{
    std::cout << "============================" << std::endl
              << "3D Taylor-Green Vortex" << std::endl
              << "Re    = " << Re << std::endl
              << "Ma    = " << Ma << std::endl
              << "rho   = " << rho0 << std::endl
              << "miu   = " << miu << std::endl
              << "tau   = " << tau << std::endl
              << "Time  = " << Time << std::endl
              << "inter = " << inter << std::endl
              << "nx,ny,nz = " << sx << " " << sy << " " << sz << std::endl
              << "============================" << std::endl;
}
