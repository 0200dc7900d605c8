// FP64 sine, cosine and arctangent of the axisymmetric mover (the DCOS, DSIN and DATAN of
// MoveOne.f90:174-176, :221-225, :248-262). They are written out here, in a fixed sequence of
// IEEE additions, multiplications and divisions (built without contraction), so that the
// device and a CPU evaluation of the same sequence give the same bits; a math library's
// routines are not bit-identical across the two. Accuracy is about 1 ulp of the exact value.
//
//   sincos: x = k pi/2 + r, k = rint(2x/pi), r by a three-part Cody-Waite split of pi/2 (the
//           first two parts carry 33 bits, so k * part is exact for |k| < 2^20); sin r and
//           cos r by their Taylor series through r^15 / r^16 (truncation below 2^-56 for
//           |r| <= pi/4); the quadrant k mod 4 selects and signs the pair.
//   atan:   |t| > 1 -> pi/2 - atan(1/|t|); |t| > tan(pi/8) -> pi/4 + atan((|t|-1)/(|t|+1));
//           then the series u - u^3/3 + ... - u^39/39 + u^41/41 for |u| <= tan(pi/8)
//           (truncation below 2^-55 relative); pi/2 and pi/4 as two-part constants.
// Constants: correctly rounded doubles of pi/2 (split), 2/pi and the series coefficients
// (-1)^n/(2n+1)!, (-1)^n/(2n)!, (-1)^n/(2n+1).
#pragma once
#include <math.h>

#ifndef PIC_HD
#define PIC_HD __host__ __device__ inline
#endif

#define PIC_TRIG_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define PIC_TRIG_PIO2_1 0x1.921fb54400000p+0
#define PIC_TRIG_PIO2_2 0x1.0b4611a600000p-34
#define PIC_TRIG_PIO2_3 0x1.3198a2e037073p-69
#define PIC_TRIG_PIO2_HI 0x1.921fb54442d18p+0
#define PIC_TRIG_PIO2_LO 0x1.1a62633145c07p-54
#define PIC_TRIG_PIO4_HI 0x1.921fb54442d18p-1
#define PIC_TRIG_PIO4_LO 0x1.1a62633145c07p-55
#define PIC_TRIG_TAN_PI8 0x1.a827999fcef32p-2

PIC_HD void pic_sincos(double x, double* s, double* c) {
    const double kd = rint(x * PIC_TRIG_TWO_OVER_PI);
    const double r = ((x - kd * PIC_TRIG_PIO2_1) - kd * PIC_TRIG_PIO2_2) - kd * PIC_TRIG_PIO2_3;
    const double z = r * r;
    const double ps = -0x1.5555555555555p-3 + z * (0x1.1111111111111p-7 + z * (-0x1.a01a01a01a01ap-13 +
                      z * (0x1.71de3a556c734p-19 + z * (-0x1.ae64567f544e4p-26 + z * (0x1.6124613a86d09p-33 +
                      z * -0x1.ae7f3e733b81fp-41)))));
    const double sr = r + r * z * ps;
    const double pc = 0x1.5555555555555p-5 + z * (-0x1.6c16c16c16c17p-10 + z * (0x1.a01a01a01a01ap-16 +
                      z * (-0x1.27e4fb7789f5cp-22 + z * (0x1.1eed8eff8d898p-29 + z * (-0x1.93974a8c07c9dp-37 +
                      z * 0x1.ae7f3e733b81fp-45)))));
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    const double cr = w + (((1.0 - w) - hz) + z * z * pc);
    switch ((int)kd & 3) {
        case 0: *s = sr; *c = cr; break;
        case 1: *s = cr; *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}

PIC_HD double pic_atan(double t) {
    if (!(t == t)) return t;
    double a = fabs(t);
    const bool inv = a > 1.0;
    if (inv) a = 1.0 / a;
    const bool red = a > PIC_TRIG_TAN_PI8;
    const double u = red ? (a - 1.0) / (a + 1.0) : a;
    const double z = u * u;
    double q = 0x1.8f9c18f9c18fap-6;
    q = -0x1.a41a41a41a41ap-6 + z * q;
    q = 0x1.bacf914c1bad0p-6 + z * q;
    q = -0x1.d41d41d41d41dp-6 + z * q;
    q = 0x1.f07c1f07c1f08p-6 + z * q;
    q = -0x1.0842108421084p-5 + z * q;
    q = 0x1.1a7b9611a7b96p-5 + z * q;
    q = -0x1.2f684bda12f68p-5 + z * q;
    q = 0x1.47ae147ae147bp-5 + z * q;
    q = -0x1.642c8590b2164p-5 + z * q;
    q = 0x1.8618618618618p-5 + z * q;
    q = -0x1.af286bca1af28p-5 + z * q;
    q = 0x1.e1e1e1e1e1e1ep-5 + z * q;
    q = -0x1.1111111111111p-4 + z * q;
    q = 0x1.3b13b13b13b14p-4 + z * q;
    q = -0x1.745d1745d1746p-4 + z * q;
    q = 0x1.c71c71c71c71cp-4 + z * q;
    q = -0x1.2492492492492p-3 + z * q;
    q = 0x1.999999999999ap-3 + z * q;
    q = -0x1.5555555555555p-2 + z * q;
    double p = u + u * z * q;
    if (red) p = PIC_TRIG_PIO4_HI + (p + PIC_TRIG_PIO4_LO);
    if (inv) p = PIC_TRIG_PIO2_HI - (p - PIC_TRIG_PIO2_LO);
    return t < 0.0 ? -p : p;
}
