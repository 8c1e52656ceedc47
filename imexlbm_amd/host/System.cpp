// The deck reader and the derived quantities below follow src/System.cpp
// and src/System.hpp of IMEXLBM, distributed under this licence:
//
// BSD 3-Clause License
//
// Copyright (c) 2025, Chunheng Zhao
// All rights reserved.
//
// Redistribution and use in source and binary forms, with or without
// modification, are permitted provided that the following conditions are met:
//
// 1. Redistributions of source code must retain the above copyright notice, this
//    list of conditions and the following disclaimer.
//
// 2. Redistributions in binary form must reproduce the above copyright notice,
//    this list of conditions and the following disclaimer in the documentation
//    and/or other materials provided with the distribution.
//
// 3. Neither the name of the copyright holder nor the names of its
//    contributors may be used to endorse or promote products derived from
//    this software without specific prior written permission.
//
// THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS IS"
// AND ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT LIMITED TO, THE
// IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR PURPOSE ARE
// DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT HOLDER OR CONTRIBUTORS BE LIABLE
// FOR ANY DIRECT, INDIRECT, INCIDENTAL, SPECIAL, EXEMPLARY, OR CONSEQUENTIAL
// DAMAGES (INCLUDING, BUT NOT LIMITED TO, PROCUREMENT OF SUBSTITUTE GOODS OR
// SERVICES; LOSS OF USE, DATA, OR PROFITS; OR BUSINESS INTERRUPTION) HOWEVER
// CAUSED AND ON ANY THEORY OF LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY,
// OR TORT (INCLUDING NEGLIGENCE OR OTHERWISE) ARISING OUT OF THE USE OF THIS
// SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
//
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
