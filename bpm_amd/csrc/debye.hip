// Vectorial Debye-Wolf focus field on gfx950.
//
// For a lateral point (x, y) and axial z the reference kernels
// (bpm/psf/kernels/psf_debye.cl:24-362) sum, over theta samples t of each
// aperture annulus, the Richards-Wolf integrands
//   I0 += w(t) (cos t + 1) J0(kr sin t) e^{i kz cos t}
//   I1 += w(t)  sin t     J1(kr sin t) e^{i kz cos t}
//   I2 += w(t) (cos t - 1) J2(kr sin t) e^{i kz cos t}
// and combine them into Ex, Ey, Ez and I.  Everything except the phase
// e^{i kz cos t} is independent of z, so a work-group takes one (x, y)
// column: it first tabulates, per theta sample, cos t and the three real
// amplitudes (Bessel functions evaluated once per column in float64) in LDS,
// then each lane accumulates the phase sums for its z values.  The column's
// 3x(S+1) Bessel evaluations are thus shared by all Nz voxels instead of
// being recomputed per voxel.  Samples are processed in LDS chunks of
// CHUNK so any number of integration steps fits.
#include "common.hpp"
#include "bessel.hpp"
#include "bpm_hip.h"

#include <cmath>
#include <cstring>

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

BPMH_DEV Sample make_sample(const DwArgs& a, int i_n, int i_t, float kr) {
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
  const double ks = (double)f_mul(kr, si);
  Sample s;
  s.co = co;
  s.a0 = f_mul(f_mul(w, f_add(co, 1.f)), (float)bpmbessel::jn(0, ks));
  s.a1 = f_mul(f_mul(w, si), (float)bpmbessel::jn(1, ks));
  s.a2 = f_mul(f_mul(w, f_sub(co, 1.f)), (float)bpmbessel::jn(2, ks));
  return s;
}

// Ex, Ey, Ez, I from the three integrals (psf_debye.cl:88-109)
BPMH_DEV void combine(const DwArgs& a, float phi, float2 I0, float2 I1, float2 I2, size_t o) {
  double s2d, c2d, s1d, c1d;
  const float phi2 = f_mul(2.f, phi);
  sincos((double)phi2, &s2d, &c2d);
  sincos((double)phi, &s1d, &c1d);
  const float c2 = (float)c2d, s2 = (float)s2d, cp = (float)c1d, sp = (float)s1d;
  const float2 ex = make_float2(a.ex0 * (I0.x + I2.x * c2) + (a.ey0 * s2) * I2.x,
                                a.ex0 * (I0.y + I2.y * c2) + (a.ey0 * s2) * I2.y);
  const float2 ey = make_float2(a.ey0 * (I0.x - I2.x * c2) + (a.ex0 * s2) * I2.x,
                                a.ey0 * (I0.y - I2.y * c2) + (a.ex0 * s2) * I2.y);
  const float m = a.ex0 * cp + a.ey0 * sp;
  const float2 ez = make_float2((2.f * I1.y) * m, (-2.f * I1.x) * m);
  if (a.ex) a.ex[o] = ex;
  if (a.ey) a.ey[o] = ey;
  if (a.ez) a.ez[o] = ez;
  if (a.I) {
    const float vx = ex.x * ex.x + ex.y * ex.y;
    const float vy = ey.x * ey.x + ey.y * ey.y;
    const float vz = ez.x * ez.x + ez.y * ez.y;
    a.I[o] = a.i_mode ? vx : vx + vy + vz;
  }
}

constexpr int DW_THREADS = 64;   // one wavefront per (x, y) column
constexpr int DW_ZPT = 4;        // z values per lane per sweep
constexpr int DW_CHUNK = 512;    // theta samples per LDS chunk

__global__ void __launch_bounds__(DW_THREADS) k_debye_grid(DwArgs a) {
  __shared__ float s_co[DW_CHUNK], s_a0[DW_CHUNK], s_a1[DW_CHUNK], s_a2[DW_CHUNK];
  const int col = blockIdx.x;
  const int i = col % a.nx;
  const int j = col / a.nx;
  const float x = axis_coord(i, a.nx, a.x1, a.x2);
  const float y = axis_coord(j, a.ny, a.y1, a.y2);
  const float kr = k_times(a.lam, __fsqrt_rn(f_add(f_mul(x, x), f_mul(y, y))));
  const float phi = (float)atan2((double)y, (double)x);
  const int per_ann = a.S + 1;
  const int total = per_ann * (a.n_alphas / 2);
  const size_t plane = (size_t)a.nx * a.ny;

  for (int z0 = 0; z0 < a.nz; z0 += DW_THREADS * DW_ZPT) {
    float kz[DW_ZPT];
    float2 I0[DW_ZPT], I1[DW_ZPT], I2[DW_ZPT];
#pragma unroll
    for (int q = 0; q < DW_ZPT; ++q) {
      const int k = z0 + threadIdx.x + q * DW_THREADS;
      kz[q] = k < a.nz ? k_times(a.lam, axis_coord(k, a.nz, a.z1, a.z2)) : 0.f;
      I0[q] = I1[q] = I2[q] = make_float2(0.f, 0.f);
    }
    const int nq = min(DW_ZPT, (a.nz - z0 + DW_THREADS - 1) / DW_THREADS);
    for (int c0 = 0; c0 < total; c0 += DW_CHUNK) {
      const int cn = min(DW_CHUNK, total - c0);
      __syncthreads();
      for (int s = threadIdx.x; s < cn; s += DW_THREADS) {
        const int g = c0 + s;
        const Sample sm = make_sample(a, g / per_ann, g % per_ann, kr);
        s_co[s] = sm.co;
        s_a0[s] = sm.a0;
        s_a1[s] = sm.a1;
        s_a2[s] = sm.a2;
      }
      __syncthreads();
      for (int s = 0; s < cn; ++s) {
        const float co = s_co[s], r0 = s_a0[s], r1 = s_a1[s], r2 = s_a2[s];
#pragma unroll
        for (int q = 0; q < DW_ZPT; ++q) {
          if (q < nq) {
            float sn, cs;
            sincosf(kz[q] * co, &sn, &cs);
            I0[q].x += r0 * cs; I0[q].y += r0 * sn;
            I1[q].x += r1 * cs; I1[q].y += r1 * sn;
            I2[q].x += r2 * cs; I2[q].y += r2 * sn;
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < DW_ZPT; ++q) {
      const int k = z0 + threadIdx.x + q * DW_THREADS;
      if (k < a.nz) combine(a, phi, I0[q], I1[q], I2[q], (size_t)k * plane + col);
    }
  }
}

__global__ void k_debye_points(DwArgs a, int n) {
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
  combine(a, phi, I0, I1, I2, (size_t)p);
}

// Full-grid spread of the octant (_focus_fields_debye.py:116-143).
__global__ void k_debye_mirror(int nx0, int ny0, int nz0, const float* uo, const float2* exo,
                               const float2* eyo, const float2* ezo, float* u, float2* ex,
                               float2* ey, float2* ez) {
  const int fx = blockIdx.x * blockDim.x + threadIdx.x;
  const int fy = blockIdx.y;
  const int fz = blockIdx.z;
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

__global__ void k_bessel_fill(int n, const float* x, float* out, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (float)bpmbessel::jn(n, (double)x[i]);
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
  hipLaunchKernelGGL(k_debye_grid, dim3(nx * ny), dim3(DW_THREADS), 0, (hipStream_t)stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
  return BPMH_OK;
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
  dim3 grid((nx0 + 127) / 128, ny0, nz0);
  hipLaunchKernelGGL(k_debye_mirror, grid, dim3(128), 0, (hipStream_t)stream, nx0, ny0, nz0,
                     u_oct, (const float2*)ex_oct, (const float2*)ey_oct, (const float2*)ez_oct,
                     u, (float2*)ex, (float2*)ey, (float2*)ez);
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
