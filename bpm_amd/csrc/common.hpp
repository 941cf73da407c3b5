// Shared device/host helpers for the gfx950 BPM and Debye-Wolf kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#define BPMH_DEV __device__ __forceinline__
#define BPMH_HD __host__ __device__ __forceinline__

// Complex helpers on float2 (x = real, y = imag).
BPMH_HD float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
BPMH_HD float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
BPMH_HD float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a*b with the component formula (a.x b.x - a.y b.y, a.x b.y + a.y b.x)
BPMH_HD float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a*conj(b)
BPMH_HD float2 c_mulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV>
BPMH_HD float2 c_rot90(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// IEEE float32 arithmetic with FMA contraction disabled, used where a formula
// of the reference has to be reproduced operation by operation.
BPMH_DEV float f_mul(float a, float b) { return __fmul_rn(a, b); }
BPMH_DEV float f_add(float a, float b) { return __fadd_rn(a, b); }
BPMH_DEV float f_sub(float a, float b) { return __fsub_rn(a, b); }
BPMH_DEV float f_div(float a, float b) { return __fdiv_rn(a, b); }

// cos/sin of a float32 angle, evaluated in float64 and rounded to float32:
// bit-reproducible against a float64 libm for all but ~2^-28 of inputs.
BPMH_DEV float2 sincos_f32_via_f64(float ph) {
  double s, c;
  sincos((double)ph, &s, &c);
  return make_float2((float)c, (float)s);
}

// The reference's M_PI in bpm_3d_kernels.cl:3 (a float literal).
#define BPMH_PI_F 3.14159265358979f
