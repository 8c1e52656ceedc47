// System.hpp -- the run deck, as the reference's System class
// (src/System.hpp:7-20): nine numbers read from ./input.in.
#ifndef IMEXLBM_SYSTEM_HPP
#define IMEXLBM_SYSTEM_HPP

class System
{
public:
    System();
    void Monitor();

    int sx, sy, sz;
    int Time, inter;
    double miu, u0, R, rho0, Ma, tau;

    int Re;
    double cs2, cs;
};
#endif
