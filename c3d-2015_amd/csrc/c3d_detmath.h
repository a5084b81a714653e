// c3d_detmath.h -- exp and log built from + - * / only, so that every IEEE binary64
// implementation (x86-64 host, gfx950 device) returns the same bits for the same input.
//
// exp: Cody & Waite (1980) argument reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split
//      into a 42-bit head (k*LN2_HI exact for |k| < 2^11) and a tail; e^r by its degree-13
//      Taylor polynomial in Horner form (truncation < 5e-18 relative); scaling by 2^k with
//      ldexp (exact, or correctly rounded into the subnormal range).  Max error ~1 ulp.
// log: x = 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1),
//      |s| <= 0.1716, by the odd series to s^23; e ln2 with the same split.  ~1 ulp.
// Used where the closure's control flow depends on an exponential or a logarithm, so the
// device and a CPU evaluation of the same method make the same decisions.  Compile with
// FP contraction off.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define C3D_HD __host__ __device__
#else
#define C3D_HD
#endif

namespace c3d {

#pragma clang fp contract(off)

C3D_HD inline double det_exp(double x)
{
    const double kInvLn2 = 0x1.71547652b82fep+0;
    const double kLn2Hi = 0x1.62e42fefa3800p-1;
    const double kLn2Lo = 0x1.ef35793c76730p-45;
    if (x != x)
        return x + x;
    if (x > 710.0)
        return HUGE_VAL;
    if (x < -746.0)
        return 0.0;
    const double k = rint(x * kInvLn2);
    const double r = (x - k * kLn2Hi) - k * kLn2Lo;
    double p = 0x1.6124613a86d09p-33;  // 1/13!
    p = p * r + 0x1.1eed8eff8d898p-29;  // 1/12!
    p = p * r + 0x1.ae64567f544e4p-26;  // 1/11!
    p = p * r + 0x1.27e4fb7789f5cp-22;  // 1/10!
    p = p * r + 0x1.71de3a556c734p-19;  // 1/9!
    p = p * r + 0x1.a01a01a01a01ap-16;  // 1/8!
    p = p * r + 0x1.a01a01a01a01ap-13;  // 1/7!
    p = p * r + 0x1.6c16c16c16c17p-10;  // 1/6!
    p = p * r + 0x1.1111111111111p-7;   // 1/5!
    p = p * r + 0x1.5555555555555p-5;   // 1/4!
    p = p * r + 0x1.5555555555555p-3;   // 1/3!
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)k);
}

C3D_HD inline double det_log(double x)
{
    const double kLn2Hi = 0x1.62e42fefa3800p-1;
    const double kLn2Lo = 0x1.ef35793c76730p-45;
    if (!(x > 0.0))
        return (x == 0.0) ? -HUGE_VAL : (x - x) / (x - x);
    if (x == HUGE_VAL)
        return x;
    int e = 0;
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < 0x1.6a09e667f3bcdp-1) {  // sqrt(1/2)
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    const double lm = 2.0 * s * p;
    const double de = (double)e;
    return de * kLn2Hi + (lm + de * kLn2Lo);
}

}  // namespace c3d
