// d3q27.hpp -- D3Q27 lattice, storage formats and the BGK collision for gfx950.
//
// Direction order, weights and velocities are the reference's
// (src/lbm.cpp:9-35): 0 rest, 1-6 faces, 7-18 edges, 19-26 corners, with each
// direction followed by its opposite.  That pairing is what both collision
// forms below exploit: e_b = -e_a makes e_b.u = -(e_a.u) exactly in IEEE
// arithmetic, so one pair shares every product.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace imex {

constexpr int Q = 27;

// Velocities e[ii] = (EX, EY, EZ), src/lbm.cpp:20-30.
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0,
                       1, -1, 1, -1, 1, -1, 1, -1};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1,
                       1, -1, -1, 1, 1, -1, -1, 1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1,
                       1, -1, 1, -1, -1, 1, -1, 1};

// Weights as the same correctly rounded quotients, src/lbm.cpp:10-17.
constexpr double W_REST = 8.0 / 27.0;
constexpr double W_FACE = 2.0 / 27.0;
constexpr double W_EDGE = 1.0 / 54.0;
constexpr double W_CORNER = 1.0 / 216.0;

__host__ __device__ constexpr double weight(int ii)
{
    return ii == 0 ? W_REST : (ii <= 6 ? W_FACE : (ii <= 18 ? W_EDGE : W_CORNER));
}

// ---------------------------------------------------------------------------
// Storage formats.  FP64 stores f itself.  The reduced formats store the
// shifted population g = f - w*rho0 and compute in FP32.
// ---------------------------------------------------------------------------
struct bf16_t { uint16_t bits; };

template <class S> struct Store;

template <> struct Store<double> {
    using real = double;
    static constexpr bool shifted = false;
    __device__ static double ld(const double *p) { return *p; }
    template <bool NT> __device__ static void st(double *p, double v)
    {
        if constexpr (NT) __builtin_nontemporal_store(v, p); else *p = v;
    }
};

template <> struct Store<float> {
    using real = float;
    static constexpr bool shifted = true;
    __device__ static float ld(const float *p) { return *p; }
    template <bool NT> __device__ static void st(float *p, float v)
    {
        if constexpr (NT) __builtin_nontemporal_store(v, p); else *p = v;
    }
};

template <> struct Store<_Float16> {
    using real = float;
    static constexpr bool shifted = true;
    __device__ static float ld(const _Float16 *p) { return (float)*p; }
    template <bool NT> __device__ static void st(_Float16 *p, float v)
    {
        const _Float16 h = (_Float16)v;   // round to nearest even
        if constexpr (NT) __builtin_nontemporal_store(h, p); else *p = h;
    }
};

// bfloat16 as raw bits: widening is a shift.  Device code narrows with the
// hardware conversion (round to nearest even); the host form below is the same
// rounding, for host-side use.
__host__ __device__ inline float bf16_to_float(uint16_t b)
{
    union { uint32_t u; float f; } c;
    c.u = (uint32_t)b << 16;
    return c.f;
}
__host__ __device__ inline uint16_t float_to_bf16(float f)
{
    union { uint32_t u; float f; } c;
    c.f = f;
    if ((c.u & 0x7f800000u) == 0x7f800000u && (c.u & 0x007fffffu))
        return (uint16_t)((c.u >> 16) | 0x40u);
    const uint32_t r = c.u + 0x7fffu + ((c.u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

template <> struct Store<bf16_t> {
    using real = float;
    static constexpr bool shifted = true;
    __device__ static float ld(const bf16_t *p)
    {
        return bf16_to_float(*reinterpret_cast<const uint16_t *>(p));
    }
    template <bool NT> __device__ static void st(bf16_t *p, float v)
    {
        const uint16_t b = __builtin_bit_cast(uint16_t, (__bf16)v);   // round to nearest even
        uint16_t *q = reinterpret_cast<uint16_t *>(p);
        if constexpr (NT) __builtin_nontemporal_store(b, q); else *q = b;
    }
};

// ---------------------------------------------------------------------------
// FP64 collision in the reference's operation order (src/lbm.cpp:140-164),
// compiled without contraction.  Moments skip the e = 0 terms: adding
// f*0 = +-0 leaves a sum unchanged except for the sign of an exact zero, which
// no later non-zero result can see.  Sums run in direction order 0..26.
// ---------------------------------------------------------------------------
struct Moments64 { double pl, ul, vl, wl; };

__device__ inline Moments64 moments64(const double (&f)[Q])
{
    double pl = 0.0, ul = 0.0, vl = 0.0, wl = 0.0;
#pragma unroll
    for (int ii = 0; ii < Q; ++ii) {
        pl += f[ii];
        if (EX[ii] == 1) ul += f[ii]; else if (EX[ii] == -1) ul -= f[ii];
        if (EY[ii] == 1) vl += f[ii]; else if (EY[ii] == -1) vl -= f[ii];
        if (EZ[ii] == 1) wl += f[ii]; else if (EZ[ii] == -1) wl -= f[ii];
    }
    return {pl, ul, vl, wl};
}

// feq for the pair (a, a+1) with e_a.u = s and e_{a+1}.u = -s, as the
// reference writes it: w*(3*pp + 3*edu + 4.5*edu*edu - 1.5*udu), left to right.
__device__ inline void feq_pair64(double p3, double y, double s, double w,
                                  double &fa, double &fb)
{
    const double s3 = 3.0 * s;
    const double x = (4.5 * s) * s;
    fa = w * (((p3 + s3) + x) - y);
    fb = w * (((p3 - s3) + x) - y);
}

// The 13 pair projections e_a.u, each with the reference's association
// ((e0*u + e1*v) + e2*w).
__device__ inline void pair_proj(double u, double v, double w, double (&s)[13])
{
    s[0] = u;            s[1] = v;            s[2] = w;
    s[3] = u + v;        s[4] = u - v;        s[5] = u + w;
    s[6] = u - w;        s[7] = v + w;        s[8] = v - w;
    s[9] = (u + v) + w;  s[10] = (u - v) + w; s[11] = (u + v) - w;
    s[12] = (u - v) - w;
}

__device__ inline void collide64(double (&f)[Q], double inv_tau)
{
    const Moments64 m = moments64(f);
    const double pp = m.pl * (1.0 / 3.0);
    const double udu = m.ul * m.ul + m.vl * m.vl + m.wl * m.wl;
    const double p3 = 3.0 * pp;
    const double y = 1.5 * udu;
    double s[13];
    pair_proj(m.ul, m.vl, m.wl, s);
    {
        const double feq = W_REST * (p3 - y);
        f[0] = f[0] - (f[0] - feq) * inv_tau;
    }
#pragma unroll
    for (int pr = 0; pr < 13; ++pr) {
        const int a = 1 + 2 * pr, b = a + 1;
        double fa, fb;
        feq_pair64(p3, y, s[pr], weight(a), fa, fb);
        f[a] = f[a] - (f[a] - fa) * inv_tau;
        f[b] = f[b] - (f[b] - fb) * inv_tau;
    }
}

// Initial equilibrium (src/lbm.cpp:101-106): the same expression with the
// analytic pressure p in place of pp and w = 0.
__device__ inline void equilibrium64(double p, double u, double v, double (&f)[Q])
{
    const double w = 0.0;
    const double udu = u * u + v * v + w * w;
    const double p3 = 3.0 * p;
    const double y = 1.5 * udu;
    double s[13];
    pair_proj(u, v, w, s);
    f[0] = W_REST * (p3 - y);
#pragma unroll
    for (int pr = 0; pr < 13; ++pr) {
        const int a = 1 + 2 * pr;
        feq_pair64(p3, y, s[pr], weight(a), f[a], f[a + 1]);
    }
}

// ---------------------------------------------------------------------------
// FP32 collision on shifted populations g = f - w*rho0.  With Σw = 1 and
// Σw e = 0: Σf = Σg + rho0 and Σf e = Σg e, so the shifted equilibrium is
// geq = w*(Σg + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u) and rho0 drops out.  The update
// is g' = g + ω (geq - g), relaxing the small non-equilibrium part.
//
// Momentum: the e.u term of geq carries all of the equilibrium's momentum,
// Σ_i 3 w_i (e_i.u) e_i = u because 2*(3w_face + 4*3w_edge + 4*3w_corner) = 1.
// With float-rounded weights that sum is 1 + δ (δ ~ 3e-8) and the relaxed
// momentum gains the same factor every step, a drift of order 1e-4 in the
// kinetic energy after a few thousand steps.  The 3w constants below are
// float values chosen so the identity holds exactly in real arithmetic;
// what remains is unbiased per-operation rounding.
// ---------------------------------------------------------------------------
constexpr float W3_FACE = 0x1.c71c72p-3f;    // 3 * 2/27  = 2/9
constexpr float W3_EDGE = 0x1.c71c72p-5f;    // 3 * 1/54  = 1/18
constexpr float W3_CORNER = 0x1.c71c7p-7f;   // 3 * 1/216 = 1/72
static_assert(2.0 * ((double)W3_FACE + 4.0 * (double)W3_EDGE + 4.0 * (double)W3_CORNER) == 1.0,
              "momentum weights must sum exactly");

__host__ __device__ constexpr float weight3f(int ii)
{
    return ii <= 6 ? W3_FACE : (ii <= 18 ? W3_EDGE : W3_CORNER);
}

// one: 1.0.  Each add of two populations, and each subtraction of one, is
// written as a fused multiply-add by one, which rounds exactly like the add.
// With the literal 1.0 the compiler emits the plain add.  Passed in a register
// it cannot see through (the FP16 kernels over ghost-bounded boxes), a widened
// FP16 operand is read by v_fma_mix_f32 in place, and the 27 widening
// conversions per cell go.
__device__ inline void collide32(float (&g)[Q], float omega, float one = 1.0f)
{
    auto add = [one](float a, float b) { return __builtin_fmaf(a, one, b); };
    auto sub = [one](float a, float b) { return __builtin_fmaf(b, -one, a); };   // a - b
    const float xp = add(g[25], (add(g[1], g[7]) + add(g[9], g[11])) + (add(g[13], g[19]) + add(g[21], g[23])));
    const float xm = add(g[26], (add(g[2], g[8]) + add(g[10], g[12])) + (add(g[14], g[20]) + add(g[22], g[24])));
    const float x0 = add(g[18], (add(g[0], g[3]) + add(g[4], g[5])) + (add(g[6], g[15]) + add(g[16], g[17])));
    const float yp = add(g[26], (add(g[3], g[7]) + add(g[10], g[15])) + (add(g[17], g[19]) + add(g[22], g[23])));
    const float ym = add(g[25], (add(g[4], g[8]) + add(g[9], g[16])) + (add(g[18], g[20]) + add(g[21], g[24])));
    const float zp = add(g[26], (add(g[5], g[11]) + add(g[14], g[15])) + (add(g[18], g[19]) + add(g[21], g[24])));
    const float zm = add(g[25], (add(g[6], g[12]) + add(g[13], g[16])) + (add(g[17], g[20]) + add(g[22], g[23])));
    const float sig = (xp + xm) + x0;
    const float u = xp - xm, v = yp - ym, w = zp - zm;
    const float udu = __builtin_fmaf(u, u, __builtin_fmaf(v, v, w * w));
    const float A = __builtin_fmaf(-1.5f, udu, sig);
    g[0] = __builtin_fmaf(omega, sub((float)W_REST * A, g[0]), g[0]);
    const float s[13] = {u, v, w, u + v, u - v, u + w, u - w, v + w, v - w,
                         (u + v) + w, (u - v) + w, (u + v) - w, (u - v) - w};
    // w*A once per weight group; the symmetric part of the pair is then
    // w*A + (4.5 w s) s, one multiply and one fused multiply-add
    const float wA[3] = {(float)W_FACE * A, (float)W_EDGE * A, (float)W_CORNER * A};
#pragma unroll
    for (int pr = 0; pr < 13; ++pr) {
        const int a = 1 + 2 * pr, b = a + 1;
        const int grp = a <= 6 ? 0 : (a <= 18 ? 1 : 2);
        const float B = __builtin_fmaf(s[pr] * (float)(4.5 * weight(a)), s[pr], wA[grp]);
        const float C = s[pr] * weight3f(a);
        g[a] = __builtin_fmaf(omega, sub(B + C, g[a]), g[a]);
        g[b] = __builtin_fmaf(omega, sub(B - C, g[b]), g[b]);
    }
}

}  // namespace imex
