// Vectorial Debye-Wolf focus field on gfx950.
//
// For a lateral point (x, y) and axial z the reference kernels
// (bpm/psf/kernels/psf_debye.cl:24-362) sum, over theta samples t of each
// aperture annulus, the Richards-Wolf integrands
//   I0 += w(t) (cos t + 1) J0(kr sin t) e^{i kz cos t}
//   I1 += w(t)  sin t     J1(kr sin t) e^{i kz cos t}
//   I2 += w(t) (cos t - 1) J2(kr sin t) e^{i kz cos t}
// and combine them into Ex, Ey, Ez and I.  The phase factor depends on
// (z, t) only and the real amplitudes on (x, y, t) only, so for the whole
// grid
//   I_n[z][col] = sum_s P[z][s] A_n[s][col],   P = e^{i kz cos t}
// is a product of a (Nz x S') complex and an (S' x NxNy) real matrix.  The
// grid path therefore tabulates A_n (Bessel functions in float64, once per
// sample and column) and P (one sincos per sample and z, instead of one per
// sample and voxel), then runs a register-tiled product whose epilogue
// combines the three sums into the outputs.  Every output sums its samples
// in index order with one fused multiply-add each, as the per-voxel form.
#include "common.hpp"
#include "bessel.hpp"
#include "bpm_hip.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

int bpmh_set_error(int code, const char* msg);

#define DW_MAX_ALPHAS 16

struct DwArgs {
  int nx, ny, nz;
  float x1, x2, y1, y2, z1, z2;
  float lam;
  float alphas[DW_MAX_ALPHAS];
  int n_alphas;
  int S;          // INT_STEPS; S+1 samples per annulus
  float ex0, ey0;
  int variant;    // BPMH_DW_*
  float exp_prefac;
  int i_mode;
  float2 *ex, *ey, *ez;
  float* I;
  const float *px, *py, *pz;  // points mode
  float2* itab;               // integral table (3, nz, ncols): the product's output when set
};

// coordinate along one axis: a1 + i*(a2-a1)/(n-1.f) (psf_debye.cl:46-48)
BPMH_DEV float axis_coord(int i, int n, float a1, float a2) {
  if (n == 1) return a1;
  return f_add(a1, f_div(f_mul((float)i, f_sub(a2, a1)), (float)n - 1.f));
}

// 2*M_PI/lam*v with the double M_PI in scope of psf_debye.cl, rounded to float
BPMH_DEV float k_times(float lam, float v) {
  return (float)((2.0 * 3.141592653589793 / (double)lam) * (double)v);
}

// Per-sample weights of the three variants (psf_debye.cl:66-78, :160-168, :320-328).
struct Sample {
  float co, a0, a1, a2;
};

// The parts of a sample that do not depend on the lateral point: cos t,
// sin t and the quadrature weight w.
struct SampleParams {
  float co, si, w;
};

BPMH_DEV SampleParams sample_params(const DwArgs& a, int i_n, int i_t) {
  const float al1 = a.alphas[2 * i_n], al2 = a.alphas[2 * i_n + 1];
  const float dt = f_div(f_sub(al2, al1), (float)a.S - 1.f);
  const float t = f_add(al1, f_mul((float)i_t, dt));
  double sd, cd;
  sincos((double)t, &sd, &cd);
  const float co = (float)cd, si = (float)sd;
  const float sq = __fsqrt_rn(co);
  float w;
  if (a.variant == BPMH_DW_INDEX) {
    w = (i_t == 0 || i_t == a.S) ? 0.5f : 1.f;
    w = f_mul(w, f_mul(dt, sq));
  } else {
    w = (t == al1 || t == al2) ? 0.5f : 1.f;
    float rhs = f_mul(f_mul(dt, sq), si);
    if (a.variant == BPMH_DW_GAUSS)
      rhs = f_mul(rhs, (float)exp((double)f_mul(f_mul(t, t), a.exp_prefac)));
    w = f_mul(w, rhs);
  }
  return SampleParams{co, si, w};
}

// amplitudes w (cos t + 1) J0, w sin t J1, w (cos t - 1) J2 at kr sin t
BPMH_DEV Sample sample_amps(SampleParams p, float kr) {
  float j0, j1, j2;
  bpmbessel::j012_f32(f_mul(kr, p.si), j0, j1, j2);
  Sample s;
  s.co = p.co;
  s.a0 = f_mul(f_mul(p.w, f_add(p.co, 1.f)), j0);
  s.a1 = f_mul(f_mul(p.w, p.si), j1);
  s.a2 = f_mul(f_mul(p.w, f_sub(p.co, 1.f)), j2);
  return s;
}

BPMH_DEV Sample make_sample(const DwArgs& a, int i_n, int i_t, float kr) {
  return sample_amps(sample_params(a, i_n, i_t), kr);
}

// Ex, Ey, Ez, I from the three integrals (psf_debye.cl:88-109), every
// product and sum rounded on its own (no contraction), so each kernel that
// inlines this gives the same bits.
struct DwOut {
  float2 ex, ey, ez;
  float I;
};

// cos 2phi, sin 2phi, cos phi, sin phi of a lateral point (float64
// evaluation, rounded), computed once per point and shared by its planes.
struct DwTrig {
  float c2, s2, cp, sp;
};

BPMH_DEV DwTrig dw_trig(float phi) {
  double s2d, c2d, s1d, c1d;
  const float phi2 = f_mul(2.f, phi);
  sincos((double)phi2, &s2d, &c2d);
  sincos((double)phi, &s1d, &c1d);
  return DwTrig{(float)c2d, (float)s2d, (float)c1d, (float)s1d};
}

BPMH_DEV DwOut dw_fields(const DwArgs& a, const DwTrig& tr, float2 I0, float2 I1, float2 I2) {
  const float c2 = tr.c2, s2 = tr.s2, cp = tr.cp, sp = tr.sp;
  const float ey0s2 = f_mul(a.ey0, s2), ex0s2 = f_mul(a.ex0, s2);
  DwOut o;
  o.ex = make_float2(f_add(f_mul(a.ex0, f_add(I0.x, f_mul(I2.x, c2))), f_mul(ey0s2, I2.x)),
                     f_add(f_mul(a.ex0, f_add(I0.y, f_mul(I2.y, c2))), f_mul(ey0s2, I2.y)));
  o.ey = make_float2(f_add(f_mul(a.ey0, f_sub(I0.x, f_mul(I2.x, c2))), f_mul(ex0s2, I2.x)),
                     f_add(f_mul(a.ey0, f_sub(I0.y, f_mul(I2.y, c2))), f_mul(ex0s2, I2.y)));
  const float m = f_add(f_mul(a.ex0, cp), f_mul(a.ey0, sp));
  o.ez = make_float2(f_mul(f_mul(2.f, I1.y), m), f_mul(f_mul(-2.f, I1.x), m));
  const float vx = f_add(f_mul(o.ex.x, o.ex.x), f_mul(o.ex.y, o.ex.y));
  const float vy = f_add(f_mul(o.ey.x, o.ey.x), f_mul(o.ey.y, o.ey.y));
  const float vz = f_add(f_mul(o.ez.x, o.ez.x), f_mul(o.ez.y, o.ez.y));
  o.I = a.i_mode ? vx : f_add(f_add(vx, vy), vz);
  return o;
}

BPMH_DEV void combine(const DwArgs& a, const DwTrig& tr, float2 I0, float2 I1, float2 I2,
                      size_t o) {
  const DwOut v = dw_fields(a, tr, I0, I1, I2);
  if (a.ex) a.ex[o] = v.ex;
  if (a.ey) a.ey[o] = v.ey;
  if (a.ez) a.ez[o] = v.ez;
  if (a.I) a.I[o] = v.I;
}

// ----------------------------------------------------------------------------
// grid path: tables, tiled product, spread
// ----------------------------------------------------------------------------
// A[n][s][col] = the three real amplitudes of sample s at lateral point
// col = j*nx + i, or at column pair col (below); P[s][z] = e^{i kz cos t_s}
// with the float32 argument kz*cos t.  Both are zero-padded to tp samples,
// cp columns and zp planes (A: row stride cp, plane stride tp*cp; P: row
// stride zp).

// Column pair p -> (i, j), i <= j, in the order (0,0), (0,1), (1,1), (0,2), ...
BPMH_DEV void pair_ij(int p, int& i, int& j) {
  j = (int)((sqrtf(8.f * (float)p + 1.f) - 1.f) * 0.5f);
  while (j * (j + 1) / 2 > p) --j;
  while ((j + 1) * (j + 2) / 2 <= p) ++j;
  i = p - j * (j + 1) / 2;
}

// The table indexed by column pairs, for identical x and y axes (nx == ny,
// x1 == y1, x2 == y2, as for the octant of a centred grid): kr(i, j) and
// kr(j, i) are then bit-identical (the same coordinates, and x^2 + y^2
// rounds the same in either order), so column p holds the amplitudes of
// both lateral points (i, j) and (j, i) of pair p; npairs = nx (nx + 1) / 2
// columns, zero-padded to cp.

// All tables of one grid call in a single launch, block row r = samples
// SB*r .. SB*r+SB-1: blocks x < nbx fill the amplitudes A_n[s][col] (column
// pairs when PAIRS), block nbx the phases P[s][z] = e^{i kz cos t_s}, and
// block nbx + 1 the azimuth terms of the lateral points (when trig is set;
// points r*256 .., strided by the rows).  The SB samples' parameters (cos t,
// sin t, w) are evaluated once per block by SB threads in parallel, each
// thread's lateral point (kr) once for its SB samples, and the float64
// Bessel series is called out of line (inlined in the sample loop, its 96
// coefficients were kept in registers: 212 instead of 88).  The launch is
// bound by its vector instruction issue.  Every value equals the
// one-kernel-per-table form bit for bit.
template <bool PAIRS, int SB>
__global__ void __launch_bounds__(256) k_dw_tables(DwArgs a, int total, int ncols, int tp, int cp,
                                                   int zp, float* __restrict__ A,
                                                   float2* __restrict__ P,
                                                   float4* __restrict__ trig) {
  const int s0 = (int)blockIdx.y * SB;
  const int nbx = (cp + 255) / 256;
  const int bx = blockIdx.x;
  if (bx == nbx + 1) {
    if (!trig) return;
    for (int k = (int)blockIdx.y * 256 + (int)threadIdx.x; k < a.nx * a.ny;
         k += (int)gridDim.y * 256) {
      const int i = k % a.nx, j = k / a.nx;
      const float phi = (float)atan2((double)axis_coord(j, a.ny, a.y1, a.y2),
                                     (double)axis_coord(i, a.nx, a.x1, a.x2));
      const DwTrig tr = dw_trig(phi);
      trig[k] = make_float4(tr.c2, tr.s2, tr.cp, tr.sp);
    }
    return;
  }
  // per sample: cos t, sin t, and the amplitude factors w (cos t + 1),
  // w sin t, w (cos t - 1) of sample_amps (the same rounded products)
  __shared__ float sp[SB][5];
  if (threadIdx.x < SB) {
    const int s = s0 + (int)threadIdx.x;
    float* q = sp[threadIdx.x];
    if (s < total) {
      const int per_ann = a.S + 1;
      const SampleParams p = sample_params(a, s / per_ann, s % per_ann);
      q[0] = p.co;
      q[1] = p.si;
      q[2] = f_mul(p.w, f_add(p.co, 1.f));
      q[3] = f_mul(p.w, p.si);
      q[4] = f_mul(p.w, f_sub(p.co, 1.f));
    } else {
      q[0] = q[1] = q[2] = q[3] = q[4] = 0.f;
    }
  }
  __syncthreads();
  if (bx == nbx) {
    for (int q = 0; q < SB; ++q) {
      const int s = s0 + q;
      for (int z = threadIdx.x; z < zp; z += blockDim.x) {
        float2 v = make_float2(0.f, 0.f);
        if (z < a.nz && s < total) {
          const float kz = k_times(a.lam, axis_coord(z, a.nz, a.z1, a.z2));
          sincosf(kz * sp[q][0], &v.y, &v.x);
        }
        P[(size_t)s * zp + z] = v;
      }
    }
    return;
  }
  const int col = bx * 256 + (int)threadIdx.x;
  if (col >= cp) return;
  const size_t plane = (size_t)tp * cp;
  float kr = 0.f;
  if (col < ncols) {
    int i, j;
    if (PAIRS) pair_ij(col, i, j);
    else i = col % a.nx, j = col / a.nx;
    const float x = axis_coord(i, a.nx, a.x1, a.x2);
    const float y = axis_coord(j, a.ny, a.y1, a.y2);
    kr = k_times(a.lam, __fsqrt_rn(f_add(f_mul(x, x), f_mul(y, y))));
  }
#pragma unroll 1
  for (int q = 0; q < SB; ++q) {
    const int s = s0 + q;
    const size_t o = (size_t)s * cp + col;
    if (s >= total || col >= ncols) {
      A[o] = A[plane + o] = A[2 * plane + o] = 0.f;
      continue;
    }
    float j0, j1, j2;
    bpmbessel::j012_f32<true>(f_mul(kr, sp[q][1]), j0, j1, j2);
    A[o] = f_mul(sp[q][2], j0);
    A[plane + o] = f_mul(sp[q][3], j1);
    A[2 * plane + o] = f_mul(sp[q][4], j2);
  }
}

// MFMA form of the product (v_mfma_f32_16x16x4f32; wave64 operand layout:
// lane l holds A[l%16][l/16] of the 16x4 left operand, B[l/16][l%16] of the
// 4x16 right operand, and D[4(l/16)+j][l%16], j < 4, of the 16x16 result).
// Left operand: Re or Im of P for 16 planes x 4 samples; right operand: A_n
// for 4 samples x 16 lateral points.  Each wavefront owns ZT tiles of 16
// planes (the kernel launches ZT = 1) x DM_CT tiles of 16 table columns (6
// accumulators per tile pair: Re, Im of I0, I1, I2) and reads its operands
// straight from the tables into registers, the next 16 samples in flight
// while the current ones are multiplied; no LDS and no barriers, so
// wavefronts run independently.  The wavefronts of a work-group share
// their table columns (their right operands hit in L1).  One column tile
// per wavefront: 74 registers, so the 256^3 octant's 4752 wavefronts are
// all resident at 4-5 per SIMD (89 us); two column tiles need 126
// registers and give 2376 wavefronts at 2-3 per SIMD, the busiest SIMDs
// then holding 1.3x the mean work (107 us).  (Also measured at 256^3: two
// plane tiles per wavefront 6 % slower; sample-contiguous tables with one
// 16-byte operand load per row and batch 6 % slower.)
constexpr int DM_CT = 1;  // 16-column tiles per wavefront
typedef float dm_v4f __attribute__((ext_vector_type(4)));
typedef unsigned dw_u2v __attribute__((ext_vector_type(2)));

// ZT 16-plane tiles (consecutive, from zt0) x DM_CT 16-column tiles per
// wavefront; each loaded operand feeds ZT (left) or DM_CT (right) MFMAs.
template <int ZT, bool PAIRS, bool ITAB>
BPMH_DEV void dw_mfma_tiles(const DwArgs& a, int ncols, int zp, int tp, int cp, int zt0, int c0, int l,
                            const float2* __restrict__ P, const float* __restrict__ A) {
  const int lr = l & 15, lk = l >> 4;
  const int plane = tp * cp;
  // The tables are zero-padded to whole tiles (tp samples, zp planes, cp
  // columns), so every load below is in range.  They are read with buffer
  // loads: one 32-bit per-lane byte offset per table, and everything that
  // varies with the sample, the order or the column tile in the scalar
  // offset (the host keeps both tables below 2 GiB).  Against per-lane
  // 64-bit addresses this removes most of the loop's VALU work and ~40
  // registers; the product ran 113 -> 77 us on the 256^3 octant in an
  // isolated test.
  const __amdgpu_buffer_rsrc_t rsA =
      __builtin_amdgcn_make_buffer_rsrc((void*)A, (short)0, 3 * plane * (int)sizeof(float), 0x00020000);
  const __amdgpu_buffer_rsrc_t rsP =
      __builtin_amdgcn_make_buffer_rsrc((void*)P, (short)0, tp * zp * (int)sizeof(float2), 0x00020000);
  const int voffA = (lk * cp + c0 + lr) * (int)sizeof(float);
  const int voffP = (lk * zp + zt0 * 16 + lr) * (int)sizeof(float2);
  dm_v4f acc[ZT][DM_CT][6];
#pragma unroll
  for (int r = 0; r < ZT; ++r)
#pragma unroll
    for (int t = 0; t < DM_CT; ++t)
#pragma unroll
      for (int i = 0; i < 6; ++i) acc[r][t][i] = dm_v4f{0.f, 0.f, 0.f, 0.f};

  // two operand buffers used in turn (no register copies between chunks)
  float2 p[2][ZT][4];
  float b[2][DM_CT][3][4];
  auto load = [&](int buf, int s0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int s = s0 + 4 * u;
#pragma unroll
      for (int r = 0; r < ZT; ++r) {
        const dw_u2v pv = __builtin_amdgcn_raw_buffer_load_b64(
            rsP, voffP, (s * zp + 16 * r) * (int)sizeof(float2), 0);
        p[buf][r][u] = make_float2(__uint_as_float(pv.x), __uint_as_float(pv.y));
      }
#pragma unroll
      for (int t = 0; t < DM_CT; ++t)
#pragma unroll
        for (int n = 0; n < 3; ++n)
          b[buf][t][n][u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
              rsA, voffA, (n * plane + s * cp + 16 * t) * (int)sizeof(float), 0));
    }
  };
  auto mma = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int r = 0; r < ZT; ++r)
#pragma unroll
        for (int t = 0; t < DM_CT; ++t)
#pragma unroll
          for (int n = 0; n < 3; ++n) {
            acc[r][t][2 * n] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                p[buf][r][u].x, b[buf][t][n][u], acc[r][t][2 * n], 0, 0, 0);
            acc[r][t][2 * n + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                p[buf][r][u].y, b[buf][t][n][u], acc[r][t][2 * n + 1], 0, 0, 0);
          }
  };
  // chunks of 16 samples in order: buffer 0 holds chunk s0, buffer 1 chunk
  // s0 + 16, each loaded one chunk ahead of its products
  load(0, 0);
  int s0 = 0;
  for (; s0 + 32 <= tp; s0 += 32) {
    load(1, s0 + 16);
    mma(0);
    if (s0 + 32 < tp) load(0, s0 + 32);
    mma(1);
  }
  if (s0 < tp) mma(0);

  if constexpr (ITAB) {
    // the three integrals of each (plane, table column), 24 contiguous bytes,
    // columns consecutive across the lanes of a tile; the fields are formed
    // where they are spread
#pragma unroll
    for (int t = 0; t < DM_CT; ++t) {
      const int c = c0 + 16 * t + lr;
      if (c >= ncols) continue;
#pragma unroll
      for (int r = 0; r < ZT; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int z = (zt0 + r) * 16 + 4 * lk + q;
          if (z >= a.nz) continue;
          float2* o = a.itab + ((size_t)z * ncols + c) * 3;
#pragma unroll
          for (int n = 0; n < 3; ++n) o[n] = make_float2(acc[r][t][2 * n][q], acc[r][t][2 * n + 1][q]);
        }
    }
    return;
  }
  // table column c is lateral point c, or (PAIRS) the points (i, j) and
  // (j, i) of pair c, each combined with its own azimuth
  const size_t out_plane = (size_t)a.nx * a.ny;
#pragma unroll
  for (int t = 0; t < DM_CT; ++t) {
    const int c = c0 + 16 * t + lr;
    if (c >= ncols) continue;
    int i, j;
    if (PAIRS) pair_ij(c, i, j);
    else i = c % a.nx, j = c / a.nx;
    const bool second = PAIRS && i != j;
    const float phi1 = (float)atan2((double)axis_coord(j, a.ny, a.y1, a.y2),
                                    (double)axis_coord(i, a.nx, a.x1, a.x2));
    const float phi2 = second ? (float)atan2((double)axis_coord(i, a.ny, a.y1, a.y2),
                                             (double)axis_coord(j, a.nx, a.x1, a.x2))
                              : 0.f;
    const size_t o1 = (size_t)j * a.nx + i, o2 = (size_t)i * a.nx + j;
    const DwTrig tr1 = dw_trig(phi1);
    const DwTrig tr2 = second ? dw_trig(phi2) : tr1;
#pragma unroll
    for (int r = 0; r < ZT; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int z = (zt0 + r) * 16 + 4 * lk + q;
        if (z >= a.nz) continue;
        const float2 I0 = make_float2(acc[r][t][0][q], acc[r][t][1][q]);
        const float2 I1 = make_float2(acc[r][t][2][q], acc[r][t][3][q]);
        const float2 I2 = make_float2(acc[r][t][4][q], acc[r][t][5][q]);
        combine(a, tr1, I0, I1, I2, (size_t)z * out_plane + o1);
        if (second) combine(a, tr2, I0, I1, I2, (size_t)z * out_plane + o2);
      }
  }
}

// Wavefronts per product work-group for nzt plane tiles: the W in 4, 3, 2
// that leaves the fewest idle wavefronts (larger W on ties), or nzt itself
// when it is below 4.
static int dw_waves_per_group(int nzt) {
  if (nzt <= 4) return nzt < 1 ? 1 : nzt;
  int best = 4, idle = (4 - nzt % 4) % 4;
  for (int w = 3; w >= 2; --w) {
    const int i = (w - nzt % w) % w;
    if (i < idle) best = w, idle = i;
  }
  return best;
}

// Wavefront w of a work-group of W wavefronts (W = blockDim.x / 64) owns
// the 16-plane tile W*g + w of its plane group g and the 16*DM_CT table
// columns of its column block.  W is chosen by the host so that W divides
// the number of plane tiles when it can: every wavefront then has work and
// all work-groups of a 256^3 call are resident at once (four-wavefront
// groups over nine plane tiles had left groups for a second round).  (Two
// plane tiles per wavefront reuse each loaded column operand twice but
// need more registers, fewer wavefronts per SIMD, and leave a tail of
// unequal wavefronts: 0.364 vs 0.341 ms per 256^3 call when measured.)
template <bool PAIRS, bool ITAB>
__global__ void __launch_bounds__(256, 2) k_dw_mfma(DwArgs a, int ncols, int nzt, int tp, int cp,
                                                const float2* __restrict__ P,
                                                const float* __restrict__ A) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wpg = (int)(blockDim.x >> 6);
  // one-dimensional grid dealt round-robin over the 8 XCDs: XCD x takes the
  // column blocks x*per .. x*per+per-1, the plane groups of a block one after
  // another, so the work-groups reading a block's table columns share an L2
  const int ncb = cp / (16 * DM_CT), ngz = (nzt + wpg - 1) / wpg;
  const int per = (ncb + 7) / 8;
  const int id = (int)blockIdx.x;
  const int k = id >> 3;
  const int cb = (id & 7) * per + k / ngz;
  const int zt0 = (k % ngz) * wpg + w;
  if (cb >= ncb || zt0 >= nzt) return;  // whole wavefront; the kernel has no barriers
  const int c0 = cb * (16 * DM_CT);
  dw_mfma_tiles<1, PAIRS, ITAB>(a, ncols, nzt * 16, tp, cp, zt0, c0, l, P, A);
}


__global__ void __launch_bounds__(64) k_debye_points(DwArgs a, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float x = a.px[p], y = a.py[p], z = a.pz[p];
  const float kr = k_times(a.lam, __fsqrt_rn(f_add(f_mul(x, x), f_mul(y, y))));
  const float kz = k_times(a.lam, z);
  const float phi = (float)atan2((double)y, (double)x);
  float2 I0 = make_float2(0.f, 0.f), I1 = I0, I2 = I0;
  for (int i_n = 0; i_n < a.n_alphas / 2; ++i_n) {
    for (int i_t = 0; i_t <= a.S; ++i_t) {
      const Sample sm = make_sample(a, i_n, i_t, kr);
      float sn, cs;
      sincosf(kz * sm.co, &sn, &cs);
      I0.x += sm.a0 * cs; I0.y += sm.a0 * sn;
      I1.x += sm.a1 * cs; I1.y += sm.a1 * sn;
      I2.x += sm.a2 * cs; I2.y += sm.a2 * sn;
    }
  }
  combine(a, dw_trig(phi), I0, I1, I2, (size_t)p);
}

// Full-grid planes of one launch of the spread: blockIdx.z < n_hi gives
// fz = hi0 + blockIdx.z, the rest fz = lo0 - (blockIdx.z - n_hi) (the two
// images of a range of octant planes; hi0 = 0, n_hi = nz0 covers every
// plane).
struct PlaneMap {
  int hi0, n_hi, lo0;
};
BPMH_DEV int plane_of(PlaneMap m) {
  const int b = blockIdx.z;
  return b < m.n_hi ? m.hi0 + b : m.lo0 - (b - m.n_hi);
}

// Full-grid spread of the octant (_focus_fields_debye.py:116-143).
__global__ void k_debye_mirror(int nx0, int ny0, int nz0, const float* uo, const float2* exo,
                               const float2* eyo, const float2* ezo, float* u, float2* ex,
                               float2* ey, float2* ez, PlaneMap pm) {
  const int fx = blockIdx.x * blockDim.x + threadIdx.x;
  const int fy = blockIdx.y;
  const int fz = plane_of(pm);
  if (fx >= nx0) return;
  const int nx = nx0 / 2 + 1, ny = ny0 / 2 + 1;
  const int ox = abs(fx - nx0 / 2), oy = abs(fy - ny0 / 2), oz = abs(fz - nz0 / 2);
  const size_t so = ((size_t)oz * ny + oy) * nx + ox;
  const size_t sf = ((size_t)fz * ny0 + fy) * nx0 + fx;
  const bool conj = fz <= nz0 / 2;
  if (u) u[sf] = uo[so];
  if (ex) { float2 v = exo[so]; if (conj) v.y = -v.y; ex[sf] = v; }
  if (ey) { float2 v = eyo[so]; if (conj) v.y = -v.y; ey[sf] = v; }
  if (ez) { float2 v = ezo[so]; if (conj) v.y = -v.y; ez[sf] = v; }
}

// The same spread, two x-neighbours per lane with 16-byte (8-byte for I)
// streaming stores: the outputs are written once and not read back here.
// Needs even nx0 (the caller checks it) and 16-byte aligned Ex/Ey/Ez, 8-byte
// aligned I.
typedef float dw_f4v __attribute__((ext_vector_type(4)));
typedef float dw_f2v __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(128) k_debye_mirror2(int nx0, int ny0, int nz0, const float* uo,
                                                       const float2* exo, const float2* eyo,
                                                       const float2* ezo, float* u, float2* ex,
                                                       float2* ey, float2* ez, PlaneMap pm) {
  const int fx = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
  const int fy = blockIdx.y;
  const int fz = plane_of(pm);
  if (fx >= nx0) return;
  const int nx = nx0 / 2 + 1, ny = ny0 / 2 + 1;
  const int ox0 = abs(fx - nx0 / 2), ox1 = abs(fx + 1 - nx0 / 2);
  const int oy = abs(fy - ny0 / 2), oz = abs(fz - nz0 / 2);
  const size_t row = ((size_t)oz * ny + oy) * nx;
  const size_t sf = ((size_t)fz * ny0 + fy) * nx0 + fx;
  const bool conj = fz <= nz0 / 2;
  if (u) __builtin_nontemporal_store(dw_f2v{uo[row + ox0], uo[row + ox1]}, (dw_f2v*)(u + sf));
  const float2* src[3] = {exo, eyo, ezo};
  float2* dst[3] = {ex, ey, ez};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!dst[q]) continue;
    float2 a = src[q][row + ox0], b = src[q][row + ox1];
    if (conj) { a.y = -a.y; b.y = -b.y; }
    __builtin_nontemporal_store(dw_f4v{a.x, a.y, b.x, b.y}, (dw_f4v*)(dst[q] + sf));
  }
}

// The spread from the pair-indexed integral table (identical octant axes).
// A work-group takes `rows` octant rows oy0.. (all ox) of octant plane
// oz = blockIdx.z: it first gathers their integrals into LDS -- pair
// (min, max) of point (ox, oy), rows i <= j read along ox and the
// transposed points along oy, so both are runs of consecutive pairs -- then
// forms each point's fields once and writes them to the full-grid rows
// fy = ny0/2 +- oy of both images of the plane, fz = nz0/2 + oz (for
// oz < nz0/2) and fz = nz0/2 - oz (for oz >= 1), two x-neighbours per lane
// with 16-byte streaming stores.  Each voxel gets the fields of its octant
// point (ox, oy, oz) formed with that point's azimuth terms -- the values
// the octant would hold -- conjugated for fz <= nz0/2.
__global__ void __launch_bounds__(512) k_debye_spread_pairs(DwArgs a, int nx0, int ny0, int nz0,
                                                            int rows,
                                                            const float2* __restrict__ itab,
                                                            const float4* __restrict__ trig,
                                                            float* u, float2* ex, float2* ey,
                                                            float2* ez) {
  extern __shared__ float2 S[];  // [rows][nxo][3]
  const int nxo = nx0 / 2 + 1;
  const int npairs = nxo * (nxo + 1) / 2;
  const int oy0 = blockIdx.x * rows;
  const int oz = blockIdx.z;
  const float2* it = itab + (size_t)oz * npairs * 3;
  const int nr = min(rows, nxo - oy0);
  // points with ox <= oy: pair tri(oy) + ox, consecutive along ox
  for (int e = threadIdx.x; e < nr * nxo; e += blockDim.x) {
    const int r = e / nxo, ox = e - r * nxo, oy = oy0 + r;
    if (ox > oy) continue;
    const float2* q = it + ((size_t)oy * (oy + 1) / 2 + ox) * 3;
    float2* d = S + ((size_t)r * nxo + ox) * 3;
    d[0] = q[0];
    d[1] = q[1];
    d[2] = q[2];
  }
  // points with ox > oy: pair tri(ox) + oy, consecutive along oy
  for (int e = threadIdx.x; e < nr * nxo; e += blockDim.x) {
    const int ox = e / nr, r = e - ox * nr, oy = oy0 + r;
    if (ox <= oy) continue;
    const float2* q = it + ((size_t)ox * (ox + 1) / 2 + oy) * 3;
    float2* d = S + ((size_t)r * nxo + ox) * 3;
    d[0] = q[0];
    d[1] = q[1];
    d[2] = q[2];
  }
  __syncthreads();
  const int hx = nx0 / 2, hy = ny0 / 2, hz = nz0 / 2;
  // the two images: fz = hz + oz (conjugated only for oz = 0) and hz - oz
  const bool has_hi = oz < hz, has_lo = oz >= 1;
  const bool conj_hi = oz == 0;
  const size_t plane_hi = (size_t)(hz + oz) * ny0, plane_lo = (size_t)(hz - oz) * ny0;
  const int lanes_per_row = nx0 / 2;  // two x-neighbours per lane
  // full-grid rows of this block: fy = hy + oy (oy < hy) and hy - oy (oy >= 1)
  const int nrows = 2 * nr;
  for (int e = threadIdx.x; e < nrows * lanes_per_row; e += blockDim.x) {
    const int k = e / lanes_per_row;
    const int fx = 2 * (e - k * lanes_per_row);
    const int r = k >> 1, oy = oy0 + r;
    const int fy = (k & 1) ? hy - oy : hy + oy;
    if ((k & 1) ? (oy < 1) : (oy >= hy)) continue;
    DwOut o[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int ox = abs(fx + h - hx);
      const float4 tv = trig[(size_t)oy * nxo + ox];
      const DwTrig tr{tv.x, tv.y, tv.z, tv.w};
      const float2* v = S + ((size_t)r * nxo + ox) * 3;
      o[h] = dw_fields(a, tr, v[0], v[1], v[2]);
    }
    const dw_f2v vu{o[0].I, o[1].I};
    const dw_f4v vx{o[0].ex.x, o[0].ex.y, o[1].ex.x, o[1].ex.y};
    const dw_f4v vy{o[0].ey.x, o[0].ey.y, o[1].ey.x, o[1].ey.y};
    const dw_f4v vz{o[0].ez.x, o[0].ez.y, o[1].ez.x, o[1].ez.y};
    const dw_f4v cj{1.f, -1.f, 1.f, -1.f};
    if (has_hi) {
      const size_t sf = (plane_hi + fy) * nx0 + fx;
      const dw_f4v m = conj_hi ? cj : dw_f4v{1.f, 1.f, 1.f, 1.f};
      if (u) __builtin_nontemporal_store(vu, (dw_f2v*)(u + sf));
      if (ex) __builtin_nontemporal_store(vx * m, (dw_f4v*)(ex + sf));
      if (ey) __builtin_nontemporal_store(vy * m, (dw_f4v*)(ey + sf));
      if (ez) __builtin_nontemporal_store(vz * m, (dw_f4v*)(ez + sf));
    }
    if (has_lo) {
      const size_t sf = (plane_lo + fy) * nx0 + fx;
      if (u) __builtin_nontemporal_store(vu, (dw_f2v*)(u + sf));
      if (ex) __builtin_nontemporal_store(vx * cj, (dw_f4v*)(ex + sf));
      if (ey) __builtin_nontemporal_store(vy * cj, (dw_f4v*)(ey + sf));
      if (ez) __builtin_nontemporal_store(vz * cj, (dw_f4v*)(ez + sf));
    }
  }
}

// the float32 evaluation that the Debye sample tables use
__global__ void k_bessel_fill(int n, const float* x, float* out, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float j0, j1, j2;
  bpmbessel::j012_f32(x[i], j0, j1, j2);
  out[i] = n == 0 ? j0 : (n == 1 ? j1 : j2);
}

static int fill_args(DwArgs& a, float lam, const float* alphas, int n_alphas, int int_steps,
                     float ex0, float ey0, int variant, float sig, int i_mode) {
  memset(&a, 0, sizeof a);
  if (!alphas) return bpmh_set_error(BPMH_ERR_INVALID, "null alphas");
  if (n_alphas < 2 || n_alphas % 2 || n_alphas > DW_MAX_ALPHAS)
    return bpmh_set_error(BPMH_ERR_INVALID, "alphas must be an even count in [2, 16]");
  if (int_steps < 2) return bpmh_set_error(BPMH_ERR_INVALID, "int_steps must be >= 2");
  if (variant < BPMH_DW_INDEX || variant > BPMH_DW_GAUSS)
    return bpmh_set_error(BPMH_ERR_INVALID, "unknown Debye variant");
  for (int i = 0; i < n_alphas; ++i) a.alphas[i] = alphas[i];
  a.n_alphas = n_alphas;
  a.lam = lam;
  a.S = int_steps;
  a.ex0 = ex0;
  a.ey0 = ey0;
  a.variant = variant;
  a.i_mode = i_mode;
  if (variant == BPMH_DW_GAUSS) {
    // exp_prefac = -.5f/alpha_max/alpha_max/sig/sig  (psf_debye.cl:307-308)
    const float am = alphas[n_alphas - 1];
    volatile float e = -0.5f;
    e = e / am;
    e = e / am;
    e = e / sig;
    e = e / sig;
    a.exp_prefac = e;
  }
  return BPMH_OK;
}

// Scratch of the grid path (tables, integral table or octant): one
// grow-only device buffer per device, reused in stream order.  A call that
// runs on another stream than the buffer's last user first makes its stream
// wait for that use (an event), so calls on several streams stay ordered
// without host synchronisation; growing waits for the last use on the host
// and reallocates (hipMalloc), which happens only when a call needs more than
// any call before it on this device.  A stream-ordered pool allocation per
// call (hipMallocFromPoolAsync) was measured here to stall the host until
// the previous call had finished, leaving the GPU idle between calls.
struct DwScratch {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipEvent_t last_use = nullptr;
  hipStream_t last_stream = nullptr;
  bool used = false;
};

static std::mutex g_dw_mu;
static DwScratch g_dw_scr[64];

// Locks g_dw_mu (released by dw_scratch_release) and returns a buffer of at
// least `bytes` that stream st may use after its pending work.
static int dw_scratch_acquire(size_t bytes, hipStream_t st, void** out, int* dev_out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
  if (dev < 0 || dev >= 64) return bpmh_set_error(BPMH_ERR_HIP, "device index out of range");
  g_dw_mu.lock();
  DwScratch& d = g_dw_scr[dev];
  if (!d.last_use) {
    e = hipEventCreateWithFlags(&d.last_use, hipEventDisableTiming);
    if (e != hipSuccess) {
      g_dw_mu.unlock();
      return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
    }
  }
  if (d.bytes < bytes) {
    if (d.used) (void)hipEventSynchronize(d.last_use);
    if (d.ptr) (void)hipFree(d.ptr);
    d.ptr = nullptr;
    d.bytes = 0;
    e = hipMalloc(&d.ptr, bytes);
    if (e != hipSuccess) {
      d.ptr = nullptr;
      g_dw_mu.unlock();
      return bpmh_set_error(BPMH_ERR_NOMEM, "device allocation failed for the Debye tables");
    }
    d.bytes = bytes;
    d.used = false;
  } else if (d.used && d.last_stream != st) {
    e = hipStreamWaitEvent(st, d.last_use, 0);
    if (e != hipSuccess) {
      g_dw_mu.unlock();
      return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
    }
  }
  *out = d.ptr;
  *dev_out = dev;
  return BPMH_OK;
}

// Marks the buffer used by the work just enqueued on st and unlocks.
static void dw_scratch_release(int dev, hipStream_t st) {
  DwScratch& d = g_dw_scr[dev];
  if (hipEventRecord(d.last_use, st) == hipSuccess) {
    d.last_stream = st;
    d.used = true;
  }
  g_dw_mu.unlock();
}

// Optional HIP-event timing of the grid path's kernels (this thread's calls):
// kinds 0 tables (sample parameters, amplitudes, phases, azimuth terms: one
// launch), 1 product, 2 full-grid spread; 3 and 4 stay zero.
struct DwTiming {
  bool on = false;
  std::vector<hipEvent_t> ev;  // pairs
  std::vector<int> kind;
  size_t used = 0;
};
static thread_local DwTiming g_dwt;

extern "C" int bpmh_debye_set_timing(int enable) {
  g_dwt.on = enable != 0;
  g_dwt.used = 0;
  return BPMH_OK;
}

extern "C" int bpmh_debye_kernel_times(float* ms, int* counts) {
  if (!ms || !counts) return bpmh_set_error(BPMH_ERR_INVALID, "null argument");
  for (int k = 0; k < 5; ++k) {
    ms[k] = 0.f;
    counts[k] = 0;
  }
  for (size_t i = 0; i < g_dwt.used; ++i) {
    float t = 0.f;
    hipError_t e = hipEventSynchronize(g_dwt.ev[2 * i + 1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&t, g_dwt.ev[2 * i], g_dwt.ev[2 * i + 1]);
    if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
    ms[g_dwt.kind[i]] += t;
    counts[g_dwt.kind[i]] += 1;
  }
  g_dwt.used = 0;
  return BPMH_OK;
}

// Run `launch` on st, between an event pair when timing is on.
template <class F>
static int dw_timed(int kind, hipStream_t st, F&& launch) {
  if (!g_dwt.on) {
    launch();
    return BPMH_OK;
  }
  while (g_dwt.ev.size() < 2 * (g_dwt.used + 1)) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, "hipEventCreate failed");
    g_dwt.ev.push_back(e);
  }
  if (g_dwt.kind.size() < g_dwt.used + 1) g_dwt.kind.resize(g_dwt.used + 1);
  const size_t i = g_dwt.used++;
  g_dwt.kind[i] = kind;
  (void)hipEventRecord(g_dwt.ev[2 * i], st);
  launch();
  (void)hipEventRecord(g_dwt.ev[2 * i + 1], st);
  return BPMH_OK;
}

// Full-grid outputs for debye_grid_run: the grid in `a` is the octant of an
// (nz0, ny0, nx0) grid; the product writes the octant into pool scratch and
// one spread kernel then writes these outputs.
struct DwFull {
  int nx0, ny0, nz0;
  float* u;
  float2 *ex, *ey, *ez;
};

// Launch the spread of octant planes [oz0, oz1) (both their full-grid
// images) on stream st.
static void launch_mirror(int nx0, int ny0, int nz0, const float* uo, const float2* exo,
                          const float2* eyo, const float2* ezo, float* u, float2* ex, float2* ey,
                          float2* ez, int oz0, int oz1, hipStream_t st) {
  const int hz = nz0 / 2;
  PlaneMap pm;
  pm.hi0 = hz + oz0;                                   // images hz + o, o < hz
  pm.n_hi = (oz1 < hz ? oz1 : hz) - oz0;
  if (pm.n_hi < 0) pm.n_hi = 0;
  const int lo_begin = oz0 > 1 ? oz0 : 1;              // images hz - o, o >= 1
  const int n_lo = oz1 - lo_begin > 0 ? oz1 - lo_begin : 0;
  pm.lo0 = hz - lo_begin;
  const int planes = pm.n_hi + n_lo;
  if (planes == 0) return;
  const bool aligned = ((uintptr_t)u % 8 == 0) && ((uintptr_t)ex % 16 == 0) &&
                       ((uintptr_t)ey % 16 == 0) && ((uintptr_t)ez % 16 == 0);
  if (aligned) {
    hipLaunchKernelGGL(k_debye_mirror2, dim3((nx0 / 2 + 127) / 128, ny0, planes), dim3(128), 0, st,
                       nx0, ny0, nz0, uo, exo, eyo, ezo, u, ex, ey, ez, pm);
  } else {
    hipLaunchKernelGGL(k_debye_mirror, dim3((nx0 + 127) / 128, ny0, planes), dim3(128), 0, st,
                       nx0, ny0, nz0, uo, exo, eyo, ezo, u, ex, ey, ez, pm);
  }
}

// The spread of every octant plane (nz0/2 + 1 of them, each to its two
// full-grid images) from the pair-indexed integral table (16-byte aligned
// Ex/Ey/Ez, 8-byte aligned I: the caller checks).
static void launch_spread_pairs(const DwArgs& a, int nx0, int ny0, int nz0, const float2* itab,
                                const float4* trig, float* u, float2* ex, float2* ey, float2* ez,
                                hipStream_t st) {
  const int planes = nz0 / 2 + 1;
  const int nxo = nx0 / 2 + 1;
  // two octant rows (four full-grid rows of each image) per 512-thread
  // work-group, one when two do not fit in 48 KiB of LDS (the caller keeps
  // nx0 <= 2048, so one row fits): at 256^3 this measured 84 us against 92
  // to 99 us for 1, 3, 4 or 8 rows, 128- or 256-thread groups, and for
  // groups that write their image rows in memory order
  int rows = (int)((48 * 1024) / ((size_t)nxo * 3 * sizeof(float2)));
  rows = rows < 1 ? 1 : (rows > 2 ? 2 : rows);
  const size_t lds = (size_t)rows * nxo * 3 * sizeof(float2);
  hipLaunchKernelGGL(k_debye_spread_pairs, dim3((nxo + rows - 1) / rows, 1, planes), dim3(512),
                     lds, st, a, nx0, ny0, nz0, rows, itab, trig, u, ex, ey, ez);
}

// samples per block row of the tables launch (divides 16)
constexpr int DW_SB = 4;

// Tables and product for the grid in `a` (octant or whole grid) on stream
// st; with `full`, the octant is then spread to the full grid.
static int debye_grid_run(DwArgs& a, int int_steps, int n_alphas, hipStream_t st,
                          const DwFull* full = nullptr) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int ncols = nx * ny;
  // samples index the y extent of the tables grid (at most 65535)
  const long long total_ll = ((long long)int_steps + 1) * (n_alphas / 2);
  if (total_ll > 65000 || nz > (1 << 23) ||
      (long long)ncols * (total_ll + 64) > (1LL << 31) ||
      (long long)(nz + 16) * (total_ll + 64) > (1LL << 31))
    return bpmh_set_error(BPMH_ERR_INVALID, "grid x samples too large");
  const int total = (int)total_ll;
  // stream-ordered scratch, zero-padded to whole product tiles: A (3 x tp x
  // cp), co (tp), P (tp x zp), per-sample parameters (3 x total), and for
  // the full-grid outputs the octant itself (28 B per octant voxel)
  // identical x and y axes: the table and the product run over column pairs
  const bool pairs = nx == ny && a.x1 == a.y1 && a.x2 == a.y2;
  const int ntab = pairs ? nx * (nx + 1) / 2 : ncols;  // table columns
  const int tp = (total + 15) / 16 * 16;
  const int nzt = (nz + 15) / 16, zp = nzt * 16;
  const int cp = (ntab + 16 * DM_CT - 1) / (16 * DM_CT) * (16 * DM_CT);
  const size_t nA = (size_t)3 * tp * cp, nP = (size_t)tp * zp;
  // the product's one-dimensional grid: column blocks x plane groups, padded
  // to a multiple of 8 XCDs
  const int wpg = dw_waves_per_group(nzt);
  const long long mblocks = 8LL * ((cp / (16 * DM_CT) + 7) / 8) * ((nzt + wpg - 1) / wpg);
  if (mblocks > (1LL << 31) - 1 || nA * sizeof(float) > (size_t)INT32_MAX ||
      nP * sizeof(float2) > (size_t)INT32_MAX)
    return bpmh_set_error(BPMH_ERR_INVALID, "grid x samples too large");
  // full grid from the pair table: the product stores the three integrals
  // per (plane, pair) and the spread forms the fields (needs the 16-byte
  // aligned outputs of the vector spread); otherwise the octant's fields
  // are stored and mirrored
  const bool itab_path = full && pairs && full->nx0 <= 2048 && ((uintptr_t)full->u % 8 == 0) &&
                         ((uintptr_t)full->ex % 16 == 0) && ((uintptr_t)full->ey % 16 == 0) &&
                         ((uintptr_t)full->ez % 16 == 0);
  const size_t noct = full ? (size_t)nx * ny * nz : 0;
  const size_t out_bytes = !full ? 0
                           : itab_path ? (size_t)3 * nz * ntab * sizeof(float2) +
                                             (size_t)nx * ny * sizeof(float4)
                                       : noct * 28;
  const size_t tab_bytes = nA * sizeof(float) + nP * sizeof(float2) + 256;
  void* scr = nullptr;
  int sdev = 0;
  const int arc = dw_scratch_acquire(tab_bytes + out_bytes + 64, st, &scr, &sdev);
  if (arc) return arc;
  float* A = (float*)scr;
  float2* P = (float2*)(((uintptr_t)(A + nA) + 15) & ~(uintptr_t)15);
  float4* trig = nullptr;
  if (itab_path) {
    // integral table (3, nz, npairs), then the azimuth terms (16-byte aligned)
    a.itab = (float2*)(((uintptr_t)scr + tab_bytes + 15) & ~(uintptr_t)15);
    trig = (float4*)(a.itab + (size_t)3 * nz * ntab);
  } else if (full) {
    // octant: Ex, Ey, Ez (16-byte aligned), then I
    float2* oex = (float2*)(((uintptr_t)scr + tab_bytes + 15) & ~(uintptr_t)15);
    a.ex = oex;
    a.ey = oex + noct;
    a.ez = oex + 2 * noct;
    a.I = (float*)(oex + 3 * noct);
  }
  int rc = dw_timed(0, st, [&] {
    const dim3 tgrid((cp + 255) / 256 + 2, tp / DW_SB);  // tp: a multiple of 16
    if (pairs)
      hipLaunchKernelGGL((k_dw_tables<true, DW_SB>), tgrid, dim3(256), 0, st, a, total, ntab, tp, cp,
                         zp, A, P, trig);
    else
      hipLaunchKernelGGL((k_dw_tables<false, DW_SB>), tgrid, dim3(256), 0, st, a, total, ncols, tp,
                         cp, zp, A, P, trig);
  });
  const dim3 mgrid((unsigned)mblocks);
  if (!rc) rc = dw_timed(1, st, [&] {
    if (itab_path)
      hipLaunchKernelGGL((k_dw_mfma<true, true>), mgrid, dim3(64 * wpg), 0, st, a, ntab, nzt, tp, cp,
                         (const float2*)P, (const float*)A);
    else if (pairs)
      hipLaunchKernelGGL((k_dw_mfma<true, false>), mgrid, dim3(64 * wpg), 0, st, a, ntab, nzt, tp,
                         cp, (const float2*)P, (const float*)A);
    else
      hipLaunchKernelGGL((k_dw_mfma<false, false>), mgrid, dim3(64 * wpg), 0, st, a, ncols, nzt, tp,
                         cp, (const float2*)P, (const float*)A);
  });
  if (!rc && itab_path)
    rc = dw_timed(2, st, [&] {
      launch_spread_pairs(a, full->nx0, full->ny0, full->nz0, a.itab, trig, full->u, full->ex,
                          full->ey, full->ez, st);
    });
  else if (!rc && full)
    rc = dw_timed(2, st, [&] {
      launch_mirror(full->nx0, full->ny0, full->nz0, a.I, a.ex, a.ey, a.ez, full->u, full->ex,
                    full->ey, full->ez, 0, nz, st);
    });
  const hipError_t e = hipGetLastError();
  dw_scratch_release(sdev, st);
  if (rc) return rc;
  if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
  return BPMH_OK;
}

extern "C" int bpmh_debye_wolf(int nx, int ny, int nz, float x1, float x2, float y1, float y2,
                               float z1, float z2, float lam, const float* alphas, int n_alphas,
                               int int_steps, float ex0, float ey0, int variant, float sig,
                               int i_mode, float* ex, float* ey, float* ez, float* I,
                               void* stream) {
  if (nx < 1 || ny < 1 || nz < 1) return bpmh_set_error(BPMH_ERR_INVALID, "bad grid");
  if ((long long)nx * ny > (1LL << 31) - 1) return bpmh_set_error(BPMH_ERR_INVALID, "grid too large");
  DwArgs a;
  int rc = fill_args(a, lam, alphas, n_alphas, int_steps, ex0, ey0, variant, sig, i_mode);
  if (rc) return rc;
  a.nx = nx; a.ny = ny; a.nz = nz;
  a.x1 = x1; a.x2 = x2; a.y1 = y1; a.y2 = y2; a.z1 = z1; a.z2 = z2;
  a.ex = (float2*)ex; a.ey = (float2*)ey; a.ez = (float2*)ez; a.I = I;
  return debye_grid_run(a, int_steps, n_alphas, (hipStream_t)stream);
}

extern "C" int bpmh_focus_field_debye(int nx0, int ny0, int nz0, double dx, double dy, double dz,
                                      float lam, const float* alphas, int n_alphas,
                                      int int_steps, float* ex, float* ey, float* ez, float* I,
                                      void* stream) {
  if (nx0 < 2 || ny0 < 2 || nz0 < 2 || (nx0 | ny0 | nz0) & 1)
    return bpmh_set_error(BPMH_ERR_UNSUPPORTED, "odd shapes not supported yet");
  if (ny0 > 65535 || nz0 > 65535 || (long long)nx0 * ny0 > (1LL << 31) - 1)
    return bpmh_set_error(BPMH_ERR_INVALID, "grid too large");
  DwArgs a;
  int rc = fill_args(a, lam, alphas, n_alphas, int_steps, 1.f, 0.f, BPMH_DW_INDEX, 0.f, 0);
  if (rc) return rc;
  // the octant (nz0/2+1, ny0/2+1, nx0/2+1) at x = 0 .. (nx0/2)*dx, the
  // extents of _focus_fields_debye.py:102-109 formed in float64 and passed
  // as float32 kernel arguments
  a.nx = nx0 / 2 + 1; a.ny = ny0 / 2 + 1; a.nz = nz0 / 2 + 1;
  a.x1 = 0.f; a.x2 = (float)(dx * (a.nx - 1.));
  a.y1 = 0.f; a.y2 = (float)(dy * (a.ny - 1.));
  a.z1 = 0.f; a.z2 = (float)(dz * (a.nz - 1.));
  DwFull full;
  full.nx0 = nx0; full.ny0 = ny0; full.nz0 = nz0;
  full.u = I; full.ex = (float2*)ex; full.ey = (float2*)ey; full.ez = (float2*)ez;
  return debye_grid_run(a, int_steps, n_alphas, (hipStream_t)stream, &full);
}

extern "C" int bpmh_debye_wolf_points(int n, const float* x, const float* y, const float* z,
                                      float lam, const float* alphas, int n_alphas,
                                      int int_steps, float ex0, float ey0, int variant,
                                      float sig, int i_mode, float* ex, float* ey, float* ez,
                                      float* I, void* stream) {
  if (n < 0) return bpmh_set_error(BPMH_ERR_INVALID, "negative point count");
  if (n > 0 && (!x || !y || !z)) return bpmh_set_error(BPMH_ERR_INVALID, "null coordinates");
  DwArgs a;
  int rc = fill_args(a, lam, alphas, n_alphas, int_steps, ex0, ey0, variant, sig, i_mode);
  if (rc) return rc;
  if (n == 0) return BPMH_OK;
  a.px = x; a.py = y; a.pz = z;
  a.ex = (float2*)ex; a.ey = (float2*)ey; a.ez = (float2*)ez; a.I = I;
  hipLaunchKernelGGL(k_debye_points, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, a, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
  return BPMH_OK;
}

extern "C" int bpmh_debye_mirror(int nx0, int ny0, int nz0, const float* u_oct,
                                 const float* ex_oct, const float* ey_oct, const float* ez_oct,
                                 float* u, float* ex, float* ey, float* ez, void* stream) {
  if (nx0 < 2 || ny0 < 2 || nz0 < 2 || (nx0 | ny0 | nz0) & 1)
    return bpmh_set_error(BPMH_ERR_UNSUPPORTED, "odd shapes not supported yet");
  if ((u && !u_oct) || (ex && !ex_oct) || (ey && !ey_oct) || (ez && !ez_oct))
    return bpmh_set_error(BPMH_ERR_INVALID, "output requested without its octant");
  if (ny0 > 65535 || nz0 > 65535) return bpmh_set_error(BPMH_ERR_INVALID, "grid too large");
  const bool aligned = ((uintptr_t)u % 8 == 0) && ((uintptr_t)ex % 16 == 0) &&
                       ((uintptr_t)ey % 16 == 0) && ((uintptr_t)ez % 16 == 0);
  if (aligned) {
    dim3 grid((nx0 / 2 + 127) / 128, ny0, nz0);
    hipLaunchKernelGGL(k_debye_mirror2, grid, dim3(128), 0, (hipStream_t)stream, nx0, ny0, nz0,
                       u_oct, (const float2*)ex_oct, (const float2*)ey_oct, (const float2*)ez_oct,
                       u, (float2*)ex, (float2*)ey, (float2*)ez, PlaneMap{0, nz0, 0});
  } else {
    dim3 grid((nx0 + 127) / 128, ny0, nz0);
    hipLaunchKernelGGL(k_debye_mirror, grid, dim3(128), 0, (hipStream_t)stream, nx0, ny0, nz0,
                       u_oct, (const float2*)ex_oct, (const float2*)ey_oct, (const float2*)ez_oct,
                       u, (float2*)ex, (float2*)ey, (float2*)ez, PlaneMap{0, nz0, 0});
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
  return BPMH_OK;
}

extern "C" int bpmh_bessel_fill(int n, const float* x, float* out, int count, void* stream) {
  if (n < 0 || n > 2) return bpmh_set_error(BPMH_ERR_UNSUPPORTED, "only J0, J1, J2 are provided");
  if (count < 0) return bpmh_set_error(BPMH_ERR_INVALID, "negative count");
  if (count == 0) return BPMH_OK;
  if (!x || !out) return bpmh_set_error(BPMH_ERR_INVALID, "null array");
  hipLaunchKernelGGL(k_bessel_fill, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     n, x, out, count);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
  return BPMH_OK;
}

extern "C" int bpmh_bessel_host(int n, const double* x, double* out, int count) {
  if (n < 0 || n > 2) return bpmh_set_error(BPMH_ERR_UNSUPPORTED, "only J0, J1, J2 are provided");
  if (count < 0 || (count > 0 && (!x || !out))) return bpmh_set_error(BPMH_ERR_INVALID, "bad arrays");
  for (int i = 0; i < count; ++i) out[i] = bpmbessel::jn(n, x[i]);
  return BPMH_OK;
}

// ----------------------------------------------------------------------------
// cylindrical lens (bpm/psf/kernels/psf_cylindrical.cl:74-185): per (y, z),
//   E = sum_t w(t) sqrt(cos t) e^{i(ky sin t + kz cos t)} (1, cos t, sin t)
// over S+1 samples t = -alpha + i_t dt, dt = 2 alpha/(S-1); end weights by
// index (psf_cylindrical) or by float equality with -+alpha (_plane).
// ----------------------------------------------------------------------------
__global__ void k_cylindrical(int ny, int nz, float y1, float y2, float z1, float z2, float lam,
                              float alpha, int S, int plane_variant, float2* ex, float2* ey,
                              float2* ez, float* I) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= ny) return;
  const float y = axis_coord(i, ny, y1, y2);
  const float z = axis_coord(j, nz, z1, z2);
  const float kf = f_div(f_mul(2.f, 3.141592653589793f), lam);
  const float ky = f_mul(kf, y), kz = f_mul(kf, z);
  const float dt = f_div(f_mul(2.f, alpha), (float)S - 1.f);
  float2 sx = make_float2(0.f, 0.f), sy = sx, sz = sx;
  for (int it = 0; it <= S; ++it) {
    const float t = f_add(-alpha, f_mul((float)it, dt));
    double sd, cd;
    sincos((double)t, &sd, &cd);
    const float co = (float)cd, si = (float)sd;
    float w;
    if (plane_variant) w = (t == -alpha || t == alpha) ? .5f : 1.f;
    else w = (it == 0 || it == S) ? .5f : 1.f;
    w = f_mul(w, f_mul(dt, __fsqrt_rn(co)));
    float sn, cs;
    sincosf(f_add(f_mul(ky, si), f_mul(kz, co)), &sn, &cs);
    sx.x += w * cs; sx.y += w * sn;
    const float wc = f_mul(w, co), ws = f_mul(w, si);
    sy.x += wc * cs; sy.y += wc * sn;
    sz.x += ws * cs; sz.y += ws * sn;
  }
  const size_t o = (size_t)j * ny + i;
  if (ex) ex[o] = sx;
  if (ey) ey[o] = sy;
  if (ez) ez[o] = sz;
  if (I) I[o] = sx.x * sx.x + sx.y * sx.y + sy.x * sy.x + sy.y * sy.y + sz.x * sz.x + sz.y * sz.y;
}

extern "C" int bpmh_cylindrical(int ny, int nz, float y1, float y2, float z1, float z2, float lam,
                                float alpha, int int_steps, int plane_variant, float* ex,
                                float* ey, float* ez, float* I, void* stream) {
  if (ny < 1 || nz < 1 || nz > 65535) return bpmh_set_error(BPMH_ERR_INVALID, "bad grid");
  if (int_steps < 2) return bpmh_set_error(BPMH_ERR_INVALID, "int_steps must be >= 2");
  hipLaunchKernelGGL(k_cylindrical, dim3((ny + 63) / 64, nz), dim3(64), 0, (hipStream_t)stream, ny,
                     nz, y1, y2, z1, z2, lam, alpha, int_steps, plane_variant, (float2*)ex,
                     (float2*)ey, (float2*)ez, I);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
  return BPMH_OK;
}
