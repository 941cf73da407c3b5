// Bessel functions of the first kind J_0, J_1, J_2 in float64 (j012, jn) and
// for float32 arguments and results (j012_f32), host and device.
//
// Methods (Abramowitz & Stegun, Handbook of Mathematical Functions, 1964):
//   |x| <  14: ascending power series, A&S 9.1.10
//       J_n(x) = (x/2)^n sum_k (-x^2/4)^k / (k! (n+k)!),
//     as a degree-31 polynomial in q = -x^2/4 (Horner); the largest term at
//     x = 14 is ~2.7e4, so ~4 of the 16 digits are lost to cancellation --
//     well beyond float32 needs (max |error| 5.4e-12 against scipy.special.jv);
//   |x| >= 14: Hankel asymptotic expansion, A&S 9.2.5 with 9.2.9/9.2.10
//       J_n(x) = sqrt(2/(pi x)) (P cos chi - Q sin chi),  chi = x - (n/2+1/4) pi,
//     P and Q as polynomials in 1/x^2 of 12 terms each (max |error| 3e-14 for
//     x >= 14); the three orders share one sincos, since chi_1 = chi_0 - pi/2
//     and chi_2 = chi_0 - pi.
// The coefficients (exact rationals rounded once to float64) are printed by
// gen_bessel_coefs.py.  Polynomials have fixed degree, so there are no
// data-dependent loops or divisions.  J_n(-x) = (-1)^n J_n(x).
// j012_f32 keeps the float64 series below 14 and evaluates the expansion in
// float32 above it, with 6 terms of P and Q.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace bpmbessel {

// Hankel asymptotic expansion coefficients of P_n and Q_n in 1/x^2 (A&S
// 9.2.9/9.2.10 for n = 0, 1, 2), exact rationals rounded once to float64.
constexpr double P0[12] = {
    1.0, -0.0703125, 0.112152099609375,
    -0.5725014209747314, 6.074042001273483, -110.01714026924674,
    3038.090510922384, -118838.42625678325, 6252951.493434797,
    -425939216.5047669, 36468400807.06556, -3833534661393.9443};
constexpr double Q0[12] = {
    -0.125, 0.0732421875, -0.22710800170898438,
    1.7277275025844574, -24.380529699556064, 551.3358961220206,
    -18257.755474293175, 832859.3040162893, -50069589.531988926,
    3836255180.2304335, -364901081884.98334, 42189715702840.97};
constexpr double P1[12] = {
    1.0, 0.1171875, -0.144195556640625,
    0.6765925884246826, -6.883914268109947, 121.59789187653587,
    -3302.2722944808525, 127641.2726461746, -6656367.718817688,
    450278600.3050393, -38338575207.427895, 4011838599133.1978};
constexpr double Q1[12] = {
    0.375, -0.1025390625, 0.2775764465332031,
    -1.993531733751297, 27.248827311268542, -603.8440767050702,
    19718.37591223663, -890297.8767070678, 53104110.10968523,
    -4043620325.107754, 382701134659.8606, -44064814178522.79};
constexpr double P2[12] = {
    1.0, -0.8203125, 0.317230224609375,
    -1.1276543140411377, 10.061105468776077, -164.51479489178382,
    4245.778664332524, -158275.1780812565, 8033547.2468489325,
    -532147436.72413737, 44555641457.28107, -4598936930713.666};
constexpr double Q2[12] = {
    1.875, 0.3076171875, -0.5154991149902344,
    3.0809126794338226, -38.14835823577596, 794.5316798750923,
    -24862.30006325488, 1088141.8493086386, -63382324.9696243,
    4736812380.840512, -441578232299.8392, 50213392901107.36};

template <int K>
__host__ __device__ inline double horner(const double (&c)[K], double q) {
  double r = c[K - 1];
#pragma unroll
  for (int k = K - 2; k >= 0; --k) r = fma(r, q, c[k]);
  return r;
}

// J_0, J_1, J_2 at |x| (signs applied by the callers)
__host__ __device__ inline void j012_abs(double ax, double& j0, double& j1, double& j2) {
  if (ax < 14.0) {
    constexpr double S0[32] = {
        1.0, 1.0, 0.25,
        0.027777777777777776, 0.001736111111111111, 6.944444444444444e-05,
        1.9290123456790124e-06, 3.936759889140842e-08, 6.151187326782565e-10,
        7.594058428126624e-12, 7.594058428126623e-14, 6.276081345559193e-16,
        4.358389823304995e-18, 2.5789288895295828e-20, 1.3157800456783586e-22,
        5.8479113141260385e-25, 2.2843403570804838e-27, 7.904291893012054e-30,
        2.4395962632753253e-32, 6.757884385804225e-35, 1.6894710964510564e-37,
        3.8310002187098785e-40, 7.915289708078262e-43, 1.4962740468957016e-45,
        2.5976979980828152e-48, 4.156316796932504e-51, 6.14839762859838e-54,
        8.434015951438106e-57, 1.0757673407446564e-59, 1.2791526049282477e-62,
        1.4212806721424974e-65, 1.4789601166935458e-68};
    constexpr double S1[32] = {
        1.0, 0.5, 0.08333333333333333,
        0.006944444444444444, 0.00034722222222222224, 1.1574074074074073e-05,
        2.755731922398589e-07, 4.920949861426052e-09, 6.834652585313961e-11,
        7.594058428126623e-13, 6.903689480115112e-15, 5.230067787965994e-17,
        3.352607556388458e-19, 1.842092063949702e-21, 8.771866971189057e-24,
        3.654944571328774e-26, 1.3437296218120491e-28, 4.391273273895586e-31,
        1.2839980333028028e-33, 3.3789421929021126e-36, 8.045100459290744e-39,
        1.7413637357772174e-41, 3.4414303078601135e-44, 6.234475195398757e-47,
        1.0390791992331261e-49, 1.5985833834355786e-52, 2.2771843068882885e-55,
        3.0121485540850376e-58, 3.7095425542919185e-61, 4.263842016427493e-64,
        4.5847763617499917e-67, 4.6217503646673306e-70};
    constexpr double S2[32] = {
        0.5, 0.16666666666666666, 0.020833333333333332,
        0.001388888888888889, 5.787037037037037e-05, 1.6534391534391535e-06,
        3.444664902998236e-08, 5.467722068251169e-10, 6.834652585313961e-12,
        6.903689480115112e-14, 5.753074566762594e-16, 4.023129067666149e-18,
        2.3947196831346127e-20, 1.228061375966468e-22, 5.482416856993161e-25,
        2.1499673948992786e-27, 7.465164565622496e-30, 2.311196459945045e-32,
        6.419990166514014e-35, 1.609020091858149e-37, 3.6568638451321564e-40,
        7.57114667729225e-43, 1.433929294941714e-45, 2.4937900781595025e-48,
        3.9964584585889467e-51, 5.92067919790955e-54, 8.132801096029602e-57,
        1.038671915201737e-59, 1.2365141847639728e-62, 1.3754329085249975e-65,
        1.4327426130468724e-68, 1.4005304135355547e-71};
    const double h = 0.5 * ax, q = -h * h;
    j0 = horner(S0, q);
    j1 = h * horner(S1, q);
    j2 = h * h * horner(S2, q);
  } else {
    const double y = 1.0 / ax, y2 = y * y;
    const double amp = sqrt(2.0 / (M_PI * ax));
    double s, c;
    sincos(ax - 0.25 * M_PI, &s, &c);  // chi_0; chi_1 = chi_0 - pi/2, chi_2 = chi_0 - pi
    const double p0 = horner(P0, y2), q0 = y * horner(Q0, y2);
    const double p1 = horner(P1, y2), q1 = y * horner(Q1, y2);
    const double p2 = horner(P2, y2), q2 = y * horner(Q2, y2);
    j0 = amp * (p0 * c - q0 * s);
    j1 = amp * (p1 * s + q1 * c);     // cos chi_1 = sin chi_0, sin chi_1 = -cos chi_0
    j2 = -amp * (p2 * c - q2 * s);    // cos chi_2 = -cos chi_0, sin chi_2 = -sin chi_0
  }
}

// The first K coefficients of c in float32 Horner form.
template <int K, int N>
__host__ __device__ inline float horner_f32(const double (&c)[N], float q) {
  static_assert(K <= N, "not enough coefficients");
  float r = (float)c[K - 1];
#pragma unroll
  for (int k = K - 2; k >= 0; --k) r = fmaf(r, q, (float)c[k]);
  return r;
}

// J_0, J_1, J_2 at ax >= 14 in float32: the same Hankel expansion with the
// first 6 terms of P and Q (the omitted ones are below 2e-9 of the leading
// term at x = 14), and the phase chi_0 = x - pi/4 taken from sincosf of the
// exact argument rotated by pi/4 (sin chi_0 = (sin x - cos x)/sqrt 2,
// cos chi_0 = (cos x + sin x)/sqrt 2), so no rounding of x - pi/4 enters.
// Within a few float32 ulp of the float64 evaluation.
__host__ __device__ inline void j012_asym_f32(float ax, float& j0, float& j1, float& j2) {
  const float y = 1.f / ax, y2 = y * y;
  const float amp = sqrtf(0.636619772f * y);  // sqrt(2 / (pi x))
  float sx, cx;
  sincosf(ax, &sx, &cx);
  const float r2 = 0.707106781f;
  const float s = (sx - cx) * r2, c = (cx + sx) * r2;
  const float p0 = horner_f32<6>(P0, y2), q0 = y * horner_f32<6>(Q0, y2);
  const float p1 = horner_f32<6>(P1, y2), q1 = y * horner_f32<6>(Q1, y2);
  const float p2 = horner_f32<6>(P2, y2), q2 = y * horner_f32<6>(Q2, y2);
  j0 = amp * (p0 * c - q0 * s);
  j1 = amp * (p1 * s + q1 * c);
  j2 = -amp * (p2 * c - q2 * s);
}

// The float64 series branch of j012_f32, kept out of line: a kernel that
// loops over many arguments would otherwise hold the 96 series
// coefficients in registers across its loop.
struct J012f {
  float j0, j1, j2;
};
__host__ __device__ __noinline__ inline J012f j012_series_f32(float ax) {
  double d0, d1, d2;
  j012_abs((double)ax, d0, d1, d2);
  return J012f{(float)d0, (float)d1, (float)d2};
}

// J_0, J_1, J_2 of a float32 argument as float32: the float64 power series
// below 14 (where float32 would lose digits to cancellation), the float32
// expansion above.  OUTLINE calls the series out of line (same values).
template <bool OUTLINE = false>
__host__ __device__ inline void j012_f32(float x, float& j0, float& j1, float& j2) {
  const float ax = fabsf(x);
  if (ax < 14.f) {
    if (OUTLINE) {
      const J012f r = j012_series_f32(ax);
      j0 = r.j0;
      j1 = r.j1;
      j2 = r.j2;
    } else {
      double d0, d1, d2;
      j012_abs((double)ax, d0, d1, d2);
      j0 = (float)d0;
      j1 = (float)d1;
      j2 = (float)d2;
    }
  } else {
    j012_asym_f32(ax, j0, j1, j2);
  }
  if (x < 0.f) j1 = -j1;
}

// J_0(x), J_1(x), J_2(x)
__host__ __device__ inline void j012(double x, double& j0, double& j1, double& j2) {
  j012_abs(fabs(x), j0, j1, j2);
  if (x < 0.0) j1 = -j1;
}

// J_n(x) for n in {0, 1, 2}
__host__ __device__ inline double jn(int n, double x) {
  double j0, j1, j2;
  j012(x, j0, j1, j2);
  return n == 0 ? j0 : (n == 1 ? j1 : j2);
}

}  // namespace bpmbessel
