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

// IEEE float32 arithmetic that is never fused into an FMA with a neighbouring
// operation, used where a formula of the reference has to be reproduced
// operation by operation (HIP's __fmul_rn / __fadd_rn are plain operators
// that the compiler contracts; the pragma marks these operations as not
// contractible, so each rounds on its own in every kernel that inlines
// them).
BPMH_DEV float f_mul(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
BPMH_DEV float f_add(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
BPMH_DEV float f_sub(float a, float b) {
#pragma clang fp contract(off)
  return a - b;
}
BPMH_DEV float f_div(float a, float b) {
#pragma clang fp contract(off)
  return a / b;
}
BPMH_DEV double d_mul(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
BPMH_DEV double d_add(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}
BPMH_DEV double d_sub(double a, double b) {
#pragma clang fp contract(off)
  return a - b;
}

// cos/sin of a float32 angle, evaluated in float64 and rounded to float32:
// bit-reproducible against a float64 libm for all but ~2^-28 of inputs.
BPMH_DEV float2 sincos_f32_via_f64(float ph) {
  double s, c;
  sincos((double)ph, &s, &c);
  return make_float2((float)c, (float)s);
}

// sin and cos of a float32 angle.  For |x| <= pi/4 (every phase screen of a
// physical index contrast: |k0 dz dn| is far below that) the Taylor series of
// sin to x^9 and cos to x^10, in Horner form with FMA; the first omitted terms
// are below 2.5e-9 relative, so the result is within about 1 ulp.  Larger
// arguments go to the range-reduced library sincosf.
BPMH_DEV void sincos_fast(float x, float* s, float* c) {
  if (fabsf(x) <= 0.78539816f) {
    const float x2 = x * x;
    float ps = fmaf(x2, 2.75573192e-6f, -1.98412698e-4f);  // 1/9!, -1/7!
    ps = fmaf(x2, ps, 8.33333333e-3f);                      // 1/5!
    ps = fmaf(x2, ps, -1.66666667e-1f);                     // -1/3!
    *s = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, -2.75573192e-7f, 2.48015873e-5f);   // -1/10!, 1/8!
    pc = fmaf(x2, pc, -1.38888889e-3f);                     // -1/6!
    pc = fmaf(x2, pc, 4.16666667e-2f);                      // 1/4!
    pc = fmaf(x2, pc, -0.5f);
    *c = fmaf(x2, pc, 1.f);
  } else {
    sincosf(x, s, c);
  }
}

// Streaming (nontemporal) global accesses.  Used where measurement favours
// them: the row pass's plane loads and stores (the column tiles that read and
// write the plane next run on other XCDs, so the lines are not reused from
// this XCD's L2) and the write-once per-plane output volumes.  The column
// pass keeps ordinary accesses: neighbouring tiles share 128-B lines in L2.
typedef float bpmh_f2v __attribute__((ext_vector_type(2)));
BPMH_DEV float2 ld_stream(const float2* q) {
  const bpmh_f2v w = __builtin_nontemporal_load((const bpmh_f2v*)q);
  return make_float2(w.x, w.y);
}
BPMH_DEV float ld_stream(const float* q) { return __builtin_nontemporal_load(q); }
BPMH_DEV void st_stream(float2* q, float2 v) {
  const bpmh_f2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, (bpmh_f2v*)q);
}
BPMH_DEV void st_stream(float* q, float v) { __builtin_nontemporal_store(v, q); }

// The reference's M_PI in bpm_3d_kernels.cl:3 (a float literal).
#define BPMH_PI_F 3.14159265358979f
