// output.cpp -- macroscopic fields and field files of the solver object.
//
// ComputeMacroscopic (src/lbm.cpp:213-240) evaluates p = (Σf)/3 and u = Σf e
// of the current state on the device and keeps them on the host.
// MPIoutput (src/lbm.cpp:308-604) writes them as a Tecplot binary file in the
// reference's exact layout: "#!TDV112", a 332-byte header (7 variables
// x y z u v w p, one ordered zone "ZONE 001" of glx x gly x glz, solution time
// n, min/max of every variable), then the 7 variables as FP64 arrays in
// Fortran order (x fastest), coordinates (i)/(glx-1).  One process writes
// the whole file, which is what the reference's collective write produces.
// Output (src/lbm.cpp:606-636) writes the rest population as ASCII.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "lbm.hpp"

namespace {

struct Writer {
    std::FILE *f;
    void i32(int32_t v) { std::fwrite(&v, 4, 1, f); }
    void f32(float v) { std::fwrite(&v, 4, 1, f); }
    void f64(double v) { std::fwrite(&v, 8, 1, f); }
    void name(const char *s)
    {
        for (; *s; ++s) i32((int32_t)(unsigned char)*s);
        i32(0);
    }
};

void die(const char *what)
{
    std::fprintf(stderr, "FATAL: %s\n", what);
    std::abort();
}

}  // namespace

void LBM::ComputeMacroscopic()
{
    const size_t n = (size_t)glx * gly * glz;
    p_h.resize(n);
    u_h.resize(n);
    v_h.resize(n);
    w_h.resize(n);
    const int s = imex_get_macro(ctx, 0, p_h.data(), u_h.data(), v_h.data(), w_h.data());
    if (s != IMEX_OK) die(imex_last_error());
}

void LBM::MPIoutput(int n)
{
    // The reference writes whatever ComputeMacroscopic last left in its
    // fields; here they are refreshed from the current state first.
    ComputeMacroscopic();
    double mn[4], mx[4];
    const std::vector<double> *fld[4] = {&u_h, &v_h, &w_h, &p_h};
    for (int m = 0; m < 4; ++m) {
        const auto r = std::minmax_element(fld[m]->begin(), fld[m]->end());
        mn[m] = *r.first;
        mx[m] = *r.second;
    }
    const std::string path = "output" + std::to_string(n) + ".plt";
    std::FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) die("cannot open the Tecplot output file");
    Writer w{fp};
    std::fwrite("#!TDV112", 1, 8, fp);
    w.i32(1);             // byte order
    w.i32(0);             // file type: full
    w.i32(0);             // title: empty string
    w.i32(7);             // number of variables
    for (const char *v : {"x", "y", "z", "u", "v", "w", "p"}) w.name(v);
    w.f32(299.0f);        // zone marker
    w.name("ZONE 001");
    w.i32(-1);            // strand id
    w.f64((double)n);     // solution time
    w.i32(0);             // not used
    w.i32(-1);            // zone colour
    w.i32(0);             // zone type: ordered
    w.i32(0);             // no variable location list
    w.i32(0);             // no raw face neighbours
    w.i32(0);             // no miscellaneous face connections
    w.i32(glx);
    w.i32(gly);
    w.i32(glz);
    w.i32(0);             // no auxiliary data
    w.f32(357.0f);        // end of header
    w.f32(299.0f);        // data section
    for (int m = 0; m < 7; ++m) w.i32(2);   // every variable FP64
    w.i32(0);             // no passive variables
    w.i32(0);             // no variable sharing
    w.i32(-1);            // no connectivity sharing
    for (int m = 0; m < 3; ++m) { w.f64(0.0); w.f64(1.0); }
    for (int m = 0; m < 4; ++m) { w.f64(mn[m]); w.f64(mx[m]); }
    if (std::ftell(fp) != 332) die("Tecplot header size");
    // 7 arrays, x fastest; fields are held (i, j, k) with k fastest
    std::vector<double> buf((size_t)glx);
    for (int var = 0; var < 7; ++var)
        for (int k = 0; k < glz; ++k)
            for (int j = 0; j < gly; ++j) {
                for (int i = 0; i < glx; ++i) {
                    double val;
                    switch (var) {
                    case 0: val = (double)(x_lo + i) / (glx - 1); break;
                    case 1: val = (double)(y_lo + j) / (gly - 1); break;
                    case 2: val = (double)(z_lo + k) / (glz - 1); break;
                    default: val = (*fld[var - 3])[((size_t)i * gly + j) * glz + k]; break;
                    }
                    buf[i] = val;
                }
                std::fwrite(buf.data(), sizeof(double), buf.size(), fp);
            }
    if (std::fclose(fp) != 0) die("writing the Tecplot output file failed");
}

void LBM::Output(int n)
{
    std::vector<double> f((size_t)LBM_Q * glx * gly * glz);
    download_f(f.data());
    std::ofstream out("output" + std::to_string(n) + std::to_string(me) + ".dat");
    out << std::setiosflags(std::ios::fixed);
    out << "variables=x,y,z,f" << std::endl;
    out << "zone I=" << glx << ",J=" << gly << ",K=" << glz << std::endl;
    for (int k = 0; k < glz; ++k)
        for (int j = 0; j < gly; ++j)
            for (int i = 0; i < glx; ++i)
                out << std::setprecision(8) << std::setiosflags(std::ios::left) << x_lo + i << " " << y_lo + j
                    << " " << z_lo + k << " " << f[((size_t)i * gly + j) * glz + k] << std::endl;
    if (me == 0) {
        std::printf("\nThe result %d is writen\n\n============================\n", n);
    }
}
